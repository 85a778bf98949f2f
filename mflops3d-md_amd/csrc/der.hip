// Compact finite-difference derivatives on gfx950.
//
// One lane owns one grid line and runs the reference's per-line algorithm:
// the explicit stencil right-hand side (class_derivatives.f90:262-295) fused
// with the forward recurrence of the packed quasi-pentadiagonal LU solve
// (pentares, class_derivatives.f90:425-466), then the backward recurrence.
// All lanes of a wavefront walk their lines in lockstep, so the coefficient
// loads are wave-uniform (scalar loads) and, for y- and z-lines whose
// neighbouring lanes are neighbouring i, every field access is a coalesced
// 512-byte wavefront access.  Field values are prefetched CH positions ahead
// into registers so that a wave keeps CH loads in flight per lane; the
// forward-sweep intermediate goes through the output array and comes back
// from L2/MALL in the backward sweep.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <vector>

#include "internal.h"
#include "mflops/mf.h"

namespace mf {

struct LineTab {      // device view of one packed scheme (dl LU + dr)
  const double *dr;   // (n,7) row-major: dr[7*i + m]
  const double *p, *f, *d, *e, *c;  // (n) each
  double alpha, beta;
  int n;
};

// y[i*s] <- scheme applied to x[i*s], i = 0..n-1 (n = T.n >= 8)
template <int CH>
__device__ __forceinline__ void solve_line(const double *__restrict__ x, double *__restrict__ y,
                                           long s, const LineTab &T) {
  const int n = T.n;
  const double *__restrict__ dr = T.dr;
  // ---- rows 0 and 1: 7-point one-sided stencils on f[0..6]
  double w[7];
#pragma unroll
  for (int m = 0; m < 7; ++m) w[m] = x[m * s];
  double r0 = 0.0, r1 = 0.0;
#pragma unroll
  for (int m = 0; m < 7; ++m) {
    r0 = fma(dr[m], w[m], r0);
    r1 = fma(dr[7 + m], w[m], r1);
  }
  double ym2 = r0;
  double ym1 = r1 - T.f[1] * ym2;
  double ym3 = 0.0;
  y[0] = ym2;
  y[s] = ym1;
  // window a[0..4] = f[i-2..i+2] for the interior rows
  double a0 = w[0], a1 = w[1], a2 = w[2], a3 = w[3], a4 = w[4];
  // lookahead values f[i+2] for rows i in [i0, i0+CH): buf[r] = f[i0+2+r]
  double buf[CH], nbuf[CH];
  const int last = n - 3;  // last interior row
  int i0 = 2;
  // f[4] is a4 already; buf holds f[5..] for rows 3..
#pragma unroll
  for (int r = 0; r < CH; ++r) {
    int q = i0 + 3 + r;  // f index for row i0+1+r
    buf[r] = (q < n) ? x[(long)q * s] : 0.0;
  }
  // row 2 first (window already complete)
  {
    const double *c = dr + 7 * 2;
    double rr = c[0] * a0 + c[1] * a1 + c[2] * a2 + c[3] * a3 + c[4] * a4;
    double yi = rr - T.p[2] * ym2 - T.f[2] * ym1;
    y[2 * s] = yi;
    ym3 = ym2; ym2 = ym1; ym1 = yi;
  }
  for (i0 = 3; i0 <= last; i0 += CH) {
#pragma unroll
    for (int r = 0; r < CH; ++r) {
      int q = i0 + CH + 2 + r;
      nbuf[r] = (q < n) ? x[(long)q * s] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < CH; ++r) {
      const int i = i0 + r;
      if (i <= last) {
        a0 = a1; a1 = a2; a2 = a3; a3 = a4; a4 = buf[r];
        const double *c = dr + 7 * i;
        double rr = c[0] * a0 + c[1] * a1 + c[2] * a2 + c[3] * a3 + c[4] * a4;
        double yi = rr - T.p[i] * ym2 - T.f[i] * ym1;
        y[(long)i * s] = yi;
        ym3 = ym2; ym2 = ym1; ym1 = yi;
      }
    }
#pragma unroll
    for (int r = 0; r < CH; ++r) buf[r] = nbuf[r];
  }
  // ---- rows n-2 and n-1: 7-point stencils on f[n-7..n-1]
  {
    double v[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) v[m] = x[(long)(n - 7 + m) * s];
    const double *c = dr + 7 * (n - 2);
    double rr = 0.0;
#pragma unroll
    for (int m = 0; m < 7; ++m) rr = fma(c[m], v[m], rr);
    double yi = rr - T.p[n - 2] * ym2 - T.f[n - 2] * ym1;
    y[(long)(n - 2) * s] = yi;
    ym3 = ym2; ym2 = ym1; ym1 = yi;
    c = dr + 7 * (n - 1);
    rr = 0.0;
#pragma unroll
    for (int m = 0; m < 7; ++m) rr = fma(c[m], v[m], rr);
    yi = rr - T.p[n - 1] * ym2 - T.f[n - 1] * ym1 - T.beta * ym3;
    // ---- backward sweep
    double xp1 = yi / T.c[n - 1];                                    // x[n-1]
    y[(long)(n - 1) * s] = xp1;
    double xc = (ym1 - T.d[n - 2] * xp1) / T.c[n - 2];               // x[n-2]
    y[(long)(n - 2) * s] = xc;
    double xp2 = xp1;
    xp1 = xc;
    // rows n-3 .. 1, prefetching y downward
    double yb[CH], nyb[CH];
#pragma unroll
    for (int r = 0; r < CH; ++r) {
      int q = n - 3 - r;
      yb[r] = (q >= 1) ? y[(long)q * s] : 0.0;
    }
    double x1 = 0.0, x2 = 0.0, x3 = 0.0;  // x[1], x[2], x[3] for row 0
    for (int j0 = n - 3; j0 >= 1; j0 -= CH) {
#pragma unroll
      for (int r = 0; r < CH; ++r) {
        int q = j0 - CH - r;
        nyb[r] = (q >= 1) ? y[(long)q * s] : 0.0;
      }
#pragma unroll
      for (int r = 0; r < CH; ++r) {
        const int i = j0 - r;
        if (i >= 1) {
          double xi = (yb[r] - T.e[i] * xp2 - T.d[i] * xp1) / T.c[i];
          y[(long)i * s] = xi;
          xp2 = xp1;
          xp1 = xi;
          if (i == 3) x3 = xi;
          if (i == 2) x2 = xi;
          if (i == 1) x1 = xi;
        }
      }
#pragma unroll
      for (int r = 0; r < CH; ++r) yb[r] = nyb[r];
    }
    const double y0 = y[0];
    y[0] = (y0 - T.e[0] * x2 - T.d[0] * x1 - T.alpha * x3) / T.c[0];
  }
}

constexpr int kCH = 16;

// 3-D driver (class_field.f90:629-978).  mode 0 = plain (every line full
// scheme), 1 = solver (interior lines full, one-boundary-plane lines on the
// sub-line scheme with zero ends, two-boundary-plane lines zero).
template <int AXIS>
__global__ void __launch_bounds__(256) der3d_kernel(const double *__restrict__ x,
                                                    double *__restrict__ out, int nx, int ny,
                                                    int nz, LineTab full, LineTab sub, int mode) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int a, b, na, nb;
  long base, s;
  if (AXIS == 0) {
    na = ny; nb = nz;
    if (t >= (long)ny * nz) return;
    a = (int)(t % ny); b = (int)(t / ny);
    base = (long)nx * (a + (long)ny * b); s = 1;
  } else if (AXIS == 1) {
    na = nx; nb = nz;
    if (t >= (long)nx * nz) return;
    a = (int)(t % nx); b = (int)(t / nx);
    base = a + (long)nx * ny * b; s = nx;
  } else {
    na = nx; nb = ny;
    if (t >= (long)nx * ny) return;
    a = (int)(t % nx); b = (int)(t / nx);
    base = a + (long)nx * b; s = (long)nx * ny;
  }
  int kind = 0;  // 0 full, 1 sub, 2 zero
  if (mode) {
    int nbd = (a == 0 || a == na - 1) + (b == 0 || b == nb - 1);
    kind = nbd;
  }
  const int n = full.n;
  if (kind == 0) {
    solve_line<kCH>(x + base, out + base, s, full);
  } else if (kind == 1) {
    out[base] = 0.0;
    out[base + (long)(n - 1) * s] = 0.0;
    solve_line<kCH>(x + base + s, out + base + s, s, sub);
  } else {
    for (int i = 0; i < n; ++i) out[base + (long)i * s] = 0.0;
  }
}

// batched lines: line l at x + l*ld, contiguous points (der/dder/der_s/dder_s)
__global__ void __launch_bounds__(256) der_lines_kernel(const double *__restrict__ x,
                                                        double *__restrict__ out, int nlines,
                                                        long ld, LineTab T, int offset) {
  const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nlines) return;
  solve_line<kCH>(x + l * ld + offset, out + l * ld + offset, 1, T);
}

}  // namespace mf

// ---------------------------------------------------------------------------
// Handle and C ABI
// ---------------------------------------------------------------------------
struct mf_der {
  int n = 0, so = 6, solver = 0;
  double *dbuf = nullptr;      // device tables
  mf::LineTab tab[2][2];       // [order-1][sub]
  std::vector<double> host[2][2][2];  // [order-1][sub][0: dl_lu(n,5), 1: dr(n,7)]
};

bool mf::der_host_tables(const mf_der *dc, int order, int sub, const double **dl_lu,
                         const double **dr, int *n) {
  if (!dc || order < 1 || order > 2 || (sub && !dc->solver)) return false;
  const auto &t = dc->host[order - 1][sub ? 1 : 0];
  *dl_lu = t[0].data();
  *dr = t[1].data();
  *n = sub ? dc->n - 2 : dc->n;
  return true;
}

namespace {
// pack one scheme: dr row-major (7n) | p | f | d | e | c
void pack_scheme(int n, const std::vector<double> &dl, const std::vector<double> &dr,
                 std::vector<double> &dst) {
  dst.assign((size_t)12 * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int m = 0; m < 7; ++m) dst[(size_t)7 * i + m] = dr[i + (size_t)n * m];
  double *p = &dst[(size_t)7 * n];
  for (int i = 0; i < n; ++i) {
    p[i] = dl[i];
    p[n + i] = dl[n + i];
    p[2 * n + i] = dl[3 * (size_t)n + i];
    p[3 * n + i] = dl[4 * (size_t)n + i];
    p[4 * n + i] = dl[2 * (size_t)n + i];
  }
}
#define HIPCHK(call)                                                             \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess)                                                        \
      return mf::fail(MF_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)
}  // namespace

extern "C" int mf_der_coeffs_1d(const double *grid, int n, int der, int so, double *dl_lu,
                                double *dr) {
  if (!grid || !dl_lu || !dr || n < 8 || (der != 1 && der != 2) ||
      (so != 2 && so != 4 && so != 6 && so != 8))
    return mf::fail(MF_ERR_ARG, "mf_der_coeffs_1d: invalid argument");
  if (!mf::der_coeffs_global(grid, n, der, so, dl_lu, dr))
    return mf::fail(MF_ERR_SETUP, "mf_der_coeffs_1d: singular Taylor system");
  mf::pentalu(n, dl_lu);
  return MF_OK;
}

extern "C" int mf_der_create(const double *grid, int n, int solver_mode, int so, mf_der **out) {
  if (!grid || !out) return mf::fail(MF_ERR_ARG, "mf_der_create: null pointer");
  const int nmin = solver_mode ? 10 : 8;
  if (n < nmin) return mf::fail(MF_ERR_ARG, "mf_der_create: line too short for the 7-point closures");
  if (so != 2 && so != 4 && so != 6 && so != 8)
    return mf::fail(MF_ERR_ARG, "mf_der_create: so must be 2, 4, 6 or 8");
  mf_der *h = new mf_der();
  h->n = n;
  h->so = so;
  h->solver = solver_mode ? 1 : 0;
  std::vector<double> packed[2][2];
  for (int o = 0; o < 2; ++o)
    for (int sub = 0; sub < 1 + h->solver; ++sub) {
      const int m = sub ? n - 2 : n;
      auto &dl = h->host[o][sub][0];
      auto &dr = h->host[o][sub][1];
      dl.assign((size_t)5 * m, 0.0);
      dr.assign((size_t)7 * m, 0.0);
      if (!mf::der_coeffs_global(grid + sub, m, o + 1, so, dl.data(), dr.data())) {
        delete h;
        return mf::fail(MF_ERR_SETUP, "mf_der_create: singular Taylor system");
      }
      mf::pentalu(m, dl.data());
      pack_scheme(m, dl, dr, packed[o][sub]);
    }
  size_t total = 0;
  for (int o = 0; o < 2; ++o)
    for (int sub = 0; sub < 2; ++sub) total += packed[o][sub].size();
  hipError_t e = hipMalloc(&h->dbuf, total * sizeof(double));
  if (e != hipSuccess) {
    delete h;
    return mf::fail(MF_ERR_DEVICE, std::string("mf_der_create: hipMalloc: ") + hipGetErrorString(e));
  }
  size_t off = 0;
  for (int o = 0; o < 2; ++o)
    for (int sub = 0; sub < 2; ++sub) {
      auto &pk = packed[o][sub];
      mf::LineTab &T = h->tab[o][sub];
      std::memset(&T, 0, sizeof(T));
      if (pk.empty()) continue;
      const int m = sub ? n - 2 : n;
      e = hipMemcpy(h->dbuf + off, pk.data(), pk.size() * sizeof(double), hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        (void)hipFree(h->dbuf);
        delete h;
        return mf::fail(MF_ERR_DEVICE, std::string("mf_der_create: hipMemcpy: ") + hipGetErrorString(e));
      }
      double *b = h->dbuf + off;
      T.dr = b;
      T.p = b + 7 * (size_t)m;
      T.f = T.p + m;
      T.d = T.f + m;
      T.e = T.d + m;
      T.c = T.e + m;
      T.alpha = h->host[o][sub][0][0];                       // dl(1,1)
      T.beta = h->host[o][sub][0][4 * (size_t)m + m - 1];    // dl(m,5) after pentalu
      T.n = m;
      off += pk.size();
    }
  *out = h;
  return MF_OK;
}

extern "C" int mf_der_destroy(mf_der *dc) {
  if (!dc) return MF_OK;
  if (dc->dbuf) (void)hipFree(dc->dbuf);
  delete dc;
  return MF_OK;
}

extern "C" int mf_der_lines(const mf_der *dc, int order, int sub, const double *x, double *out,
                            int nlines, long long ld, void *stream) {
  if (!dc || !x || !out || (order != 1 && order != 2) || nlines < 0 || ld < dc->n)
    return mf::fail(MF_ERR_ARG, "mf_der_lines: invalid argument");
  if (sub && !dc->solver)
    return mf::fail(MF_ERR_ARG, "mf_der_lines: sub-line scheme needs a solver-mode handle");
  if (nlines == 0) return MF_OK;
  const mf::LineTab &T = dc->tab[order - 1][sub ? 1 : 0];
  const int blocks = (nlines + 255) / 256;
  hipLaunchKernelGGL(mf::der_lines_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, x,
                     out, nlines, (long)ld, T, sub ? 1 : 0);
  HIPCHK(hipGetLastError());
  return MF_OK;
}

extern "C" int mf_der3d(const mf_der *dc, int axis, int order, const double *x, double *out,
                        int nx, int ny, int nz, void *stream) {
  if (!dc || !x || !out || axis < 0 || axis > 2 || (order != 1 && order != 2) || nx < 1 ||
      ny < 1 || nz < 1)
    return mf::fail(MF_ERR_ARG, "mf_der3d: invalid argument");
  const int dims[3] = {nx, ny, nz};
  if (dims[axis] != dc->n)
    return mf::fail(MF_ERR_ARG, "mf_der3d: field size along the axis differs from the handle's grid");
  if (x == out) return mf::fail(MF_ERR_ARG, "mf_der3d: in-place derivative is not supported");
  const long lines = (long)nx * ny * nz / dims[axis];
  const int blocks = (int)((lines + 255) / 256);
  const mf::LineTab &F = dc->tab[order - 1][0];
  const mf::LineTab &S = dc->tab[order - 1][1];
  const int mode = dc->solver;
  hipStream_t st = (hipStream_t)stream;
  if (axis == 0)
    hipLaunchKernelGGL(mf::der3d_kernel<0>, dim3(blocks), dim3(256), 0, st, x, out, nx, ny, nz, F, S, mode);
  else if (axis == 1)
    hipLaunchKernelGGL(mf::der3d_kernel<1>, dim3(blocks), dim3(256), 0, st, x, out, nx, ny, nz, F, S, mode);
  else
    hipLaunchKernelGGL(mf::der3d_kernel<2>, dim3(blocks), dim3(256), 0, st, x, out, nx, ny, nz, F, S, mode);
  HIPCHK(hipGetLastError());
  return MF_OK;
}

extern "C" int mf_der3d_host(const mf_der *dc, int axis, int order, const double *x, double *out,
                             int nx, int ny, int nz) {
  if (!x || !out || nx < 1 || ny < 1 || nz < 1) return mf::fail(MF_ERR_ARG, "mf_der3d_host: invalid argument");
  const size_t bytes = (size_t)nx * ny * nz * sizeof(double);
  double *dx = nullptr, *dy = nullptr;
  HIPCHK(hipMalloc(&dx, bytes));
  hipError_t e = hipMalloc(&dy, bytes);
  if (e != hipSuccess) {
    (void)hipFree(dx);
    return mf::fail(MF_ERR_DEVICE, std::string("mf_der3d_host: hipMalloc: ") + hipGetErrorString(e));
  }
  int rc = MF_OK;
  e = hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    rc = mf_der3d(dc, axis, order, dx, dy, nx, ny, nz, nullptr);
    if (rc == MF_OK) e = hipMemcpy(out, dy, bytes, hipMemcpyDeviceToHost);
  }
  (void)hipFree(dx);
  (void)hipFree(dy);
  if (rc != MF_OK) return rc;
  if (e != hipSuccess)
    return mf::fail(MF_ERR_DEVICE, std::string("mf_der3d_host: ") + hipGetErrorString(e));
  return MF_OK;
}
