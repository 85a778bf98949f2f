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

#include "der_seg.h"
#include "internal.h"
#include "stage.h"
#include "mflops/mf.h"

namespace mf {

struct LineTab {      // device view of one packed scheme (dl LU + dr)
  const double *dr;   // (n,7) row-major: dr[7*i + m]
  const double *p, *f, *d, *e, *c;  // (n) each
  const double *rc;   // 1/c
  double alpha, beta;
  int n;
};

// y[i*ys] <- scheme applied to x[i*xs], i = 0..n-1 (n = T.n >= 8).
// The interior rows run in chunks of CH with the next chunk's inputs loaded
// ahead; every load and store is unconditional (addresses clamped into the
// line) and every branch is wave-uniform, so the compiler can wait for
// exactly the loads it needs.  A surplus lane of a block passes a scratch
// line as y (ys = 1): it walks the recurrences and stores only there.
// The coefficient rows come from CP-typed pointers: the block's copy of the
// scheme in LDS (coef_lds), whose wave-uniform reads are broadcasts that the
// compiler can issue ahead of the recurrence.
struct Coef {
  const double *dr, *p, *f, *d, *e, *rc;
  double alpha, beta;
  int n;
};

// Copy the packed scheme (dr 7n | p | f | d | e | c | 1/c) into LDS.
__device__ __forceinline__ Coef coef_lds(const LineTab &T, double *lds) {
  const int n = T.n;
  for (int q = threadIdx.x; q < 13 * n; q += blockDim.x) lds[q] = T.dr[q];
  __syncthreads();
  Coef c;
  c.dr = lds;
  c.p = lds + 7 * n;
  c.f = c.p + n;
  c.d = c.f + n;
  c.e = c.d + n;
  c.rc = c.e + 2 * n;
  c.alpha = T.alpha;
  c.beta = T.beta;
  c.n = n;
  return c;
}

template <int CH>
__device__ __forceinline__ void solve_line(const double *__restrict__ x, long xs,
                                           double *__restrict__ y, long ys, const Coef &T) {
  const int n = T.n;
  typedef const double *cptr;
  const cptr dr = T.dr, Tp = T.p, Tf = T.f, Td = T.d, Te = T.e, Trc = T.rc;
  auto X = [&](int q) { return x[(long)q * xs]; };
  // ---- rows 0 and 1: 7-point one-sided stencils on f[0..6]
  double w[7];
#pragma unroll
  for (int m = 0; m < 7; ++m) w[m] = X(m);
  double r0 = 0.0, r1 = 0.0;
#pragma unroll
  for (int m = 0; m < 7; ++m) {
    r0 = fma(dr[m], w[m], r0);
    r1 = fma(dr[7 + m], w[m], r1);
  }
  double ym2 = r0;
  double ym1 = r1 - Tf[1] * ym2;
  double ym3 = 0.0;
  y[0] = ym2;
  y[ys] = ym1;
  // window a[0..4] = f[i-2..i+2] for the interior rows 2..n-3
  double a0 = w[0], a1 = w[1], a2 = w[2], a3 = w[3], a4 = w[4];
  auto row = [&](int i, double fnext) {
    a0 = a1; a1 = a2; a2 = a3; a3 = a4; a4 = fnext;
    const cptr c = dr + 7 * i;
    const double rr = c[0] * a0 + c[1] * a1 + c[2] * a2 + c[3] * a3 + c[4] * a4;
    const double yi = rr - Tp[i] * ym2 - Tf[i] * ym1;
    y[(long)i * ys] = yi;
    ym3 = ym2; ym2 = ym1; ym1 = yi;
  };
  // row 2: its window f[0..4] is complete already
  {
    const cptr c = dr + 7 * 2;
    const double rr = c[0] * a0 + c[1] * a1 + c[2] * a2 + c[3] * a3 + c[4] * a4;
    const double yi = rr - Tp[2] * ym2 - Tf[2] * ym1;
    y[2 * ys] = yi;
    ym3 = ym2; ym2 = ym1; ym1 = yi;
  }
  const int last = n - 3;  // last interior row
  int i0 = 3;
  double buf[CH];
#pragma unroll
  for (int r = 0; r < CH; ++r) buf[r] = X(min(i0 + 2 + r, n - 1));  // f[i+2], rows i0..
  for (; i0 + CH - 1 <= last; i0 += CH) {
    double nbuf[CH];
#pragma unroll
    for (int r = 0; r < CH; ++r) nbuf[r] = X(min(i0 + CH + 2 + r, n - 1));
#pragma unroll
    for (int r = 0; r < CH; ++r) row(i0 + r, buf[r]);
#pragma unroll
    for (int r = 0; r < CH; ++r) buf[r] = nbuf[r];
  }
  // remaining interior rows (fewer than CH): buf holds their inputs
#pragma unroll
  for (int r = 0; r < CH - 1; ++r)
    if (i0 + r <= last) row(i0 + r, buf[r]);
  // ---- rows n-2 and n-1: 7-point stencils on f[n-7..n-1]
  double v[7];
#pragma unroll
  for (int m = 0; m < 7; ++m) v[m] = X(n - 7 + m);
  {
    cptr c = dr + 7 * (n - 2);
    double rr = 0.0;
#pragma unroll
    for (int m = 0; m < 7; ++m) rr = fma(c[m], v[m], rr);
    double yi = rr - Tp[n - 2] * ym2 - Tf[n - 2] * ym1;
    y[(long)(n - 2) * ys] = yi;
    ym3 = ym2; ym2 = ym1; ym1 = yi;
    c = dr + 7 * (n - 1);
    rr = 0.0;
#pragma unroll
    for (int m = 0; m < 7; ++m) rr = fma(c[m], v[m], rr);
    yi = rr - Tp[n - 1] * ym2 - Tf[n - 1] * ym1 - T.beta * ym3;
    // ---- backward sweep
    double xp1 = yi * Trc[n - 1];                                    // x[n-1]
    y[(long)(n - 1) * ys] = xp1;
    double xc = (ym1 - Td[n - 2] * xp1) * Trc[n - 2];               // x[n-2]
    y[(long)(n - 2) * ys] = xc;
    double xp2 = xp1;
    xp1 = xc;
    double x1 = 0.0, x2 = 0.0, x3 = 0.0;  // x[1], x[2], x[3] for row 0
    auto brow = [&](int i, double yv) {
      const double xi = (yv - Te[i] * xp2 - Td[i] * xp1) * Trc[i];
      y[(long)i * ys] = xi;
      xp2 = xp1;
      xp1 = xi;
    };
    // rows n-3 .. 1 in chunks, prefetching y downward
    int j0 = n - 3;
    double yb[CH];
#pragma unroll
    for (int r = 0; r < CH; ++r) yb[r] = y[(long)max(j0 - r, 1) * ys];
    for (; j0 - CH + 1 >= 4; j0 -= CH) {  // full chunks that stay above row 3
      double nyb[CH];
#pragma unroll
      for (int r = 0; r < CH; ++r) nyb[r] = y[(long)max(j0 - CH - r, 1) * ys];
#pragma unroll
      for (int r = 0; r < CH; ++r) brow(j0 - r, yb[r]);
#pragma unroll
      for (int r = 0; r < CH; ++r) yb[r] = nyb[r];
    }
    // remaining rows j0 .. 1 (fewer than CH + 3), keeping x[3], x[2], x[1]
#pragma unroll
    for (int r = 0; r < CH + 2; ++r) {
      const int i = j0 - r;
      if (i >= 1) {
        const double yv = r < CH ? yb[r] : y[(long)i * ys];
        brow(i, yv);
        if (i == 3) x3 = xp1;
        if (i == 2) x2 = xp1;
        if (i == 1) x1 = xp1;
      }
    }
    const double y0 = y[0];
    y[0] = (y0 - Te[0] * x2 - Td[0] * x1 - T.alpha * x3) * Trc[0];
  }
}

constexpr int kCH = 16;
constexpr int kMaxLine = 630;  // 13 coefficient doubles per point in 64 KiB of LDS

// 3-D driver (class_field.f90:629-978).  mode 0 = plain (every line full
// scheme), 1 = solver (interior lines full, one-boundary-plane lines on the
// sub-line scheme with zero ends, two-boundary-plane lines zero).  Blocks
// [0, b_full) take full-scheme lines, [b_full, b_full + b_sub) the sub-line
// ones and the last block the zero lines, so every line solve runs in
// wave-uniform control flow (the coefficients then come through the scalar
// cache); surplus lanes of a block walk its last line and store into the
// handle's scratch line.
// Up to kMaxBatch fields of one shape go in one launch, field = blockIdx.y.
constexpr int kMaxBatch = 8;
struct FieldSet {
  const double *x[kMaxBatch];
  double *out[kMaxBatch];
};

template <int AXIS>
__global__ void __launch_bounds__(256) der3d_kernel(FieldSet fs, int nx, int ny, int nz,
                                                    LineTab full, LineTab sub, int mode,
                                                    int b_full, int b_sub,
                                                    double *__restrict__ scratch) {
  const double *__restrict__ x = fs.x[blockIdx.y];
  double *__restrict__ out = fs.out[blockIdx.y];
  int na, nb;
  long s;
  if (AXIS == 0) { na = ny; nb = nz; s = 1; }
  else if (AXIS == 1) { na = nx; nb = nz; s = nx; }
  else { na = nx; nb = ny; s = (long)nx * ny; }
  auto base_of = [&](int a, int b) -> long {
    if (AXIS == 0) return (long)nx * (a + (long)ny * b);
    if (AXIS == 1) return a + (long)nx * ny * b;
    return a + (long)nx * b;
  };
  extern __shared__ double coef[];
  const int n = full.n;
  const int blk = blockIdx.x;
  const long tid = threadIdx.x;
  if (!mode) {
    const Coef C = coef_lds(full, coef);
    const long cnt = (long)na * nb;
    long t = (long)blk * 256 + tid;
    const bool act = t < cnt;
    t = act ? t : cnt - 1;
    const long base = base_of((int)(t % na), (int)(t / na));
    solve_line<kCH>(x + base, s, act ? out + base : scratch, act ? s : 1, C);
  } else if (blk < b_full) {
    const Coef C = coef_lds(full, coef);
    const long cnt = (long)(na - 2) * (nb - 2);
    long t = (long)blk * 256 + tid;
    const bool act = t < cnt;
    t = act ? t : cnt - 1;
    const long base = base_of(1 + (int)(t % (na - 2)), 1 + (int)(t / (na - 2)));
    solve_line<kCH>(x + base, s, act ? out + base : scratch, act ? s : 1, C);
  } else if (blk < b_full + b_sub) {
    const Coef C = coef_lds(sub, coef);
    const long cnt = 2L * (na - 2) + 2L * (nb - 2);
    long t = (long)(blk - b_full) * 256 + tid;
    const bool act = t < cnt;
    t = act ? t : cnt - 1;
    int a, b;
    if (t < 2L * (na - 2)) {
      a = 1 + (int)(t % (na - 2));
      b = (t / (na - 2)) ? nb - 1 : 0;
    } else {
      t -= 2L * (na - 2);
      b = 1 + (int)(t % (nb - 2));
      a = (t / (nb - 2)) ? na - 1 : 0;
    }
    const long base = base_of(a, b);
    if (act) {
      out[base] = 0.0;
      out[base + (long)(n - 1) * s] = 0.0;
    }
    solve_line<kCH>(x + base + s, s, act ? out + base + s : scratch, act ? s : 1, C);
  } else {
    for (long t = tid; t < 4L * n; t += 256) {
      const int line = (int)(t / n), i = (int)(t % n);
      const long base = base_of((line & 1) ? na - 1 : 0, (line & 2) ? nb - 1 : 0);
      out[base + (long)i * s] = 0.0;
    }
  }
}

// 3-D driver on segmented line solves (seg_solve): blocks of 64 lines, one
// wavefront per segment.  Plain mode (mode 0): every line on the full scheme,
// blocks [0, b_full).  Solver mode: blocks [0, b_full) take the interior lines
// (full scheme), the next nb4.x .. nb4.w blocks the four families of
// one-boundary-plane lines (sub-line scheme on points 1..n-2), and the last
// block writes the zeros (two-boundary-plane lines, sub-line end points).
template <int AXIS>
__global__ void __launch_bounds__(64 * SEG_KMAX) der3d_seg_kernel(FieldSet fs, int nx, int ny,
                                                                  int nz, SegTab full,
                                                                  SegTab sub, int mode,
                                                                  int b_full, int4 nb4) {
  extern __shared__ double lds[];
  const double *__restrict__ x = fs.x[blockIdx.y];
  double *__restrict__ out = fs.out[blockIdx.y];
  int na, nb;
  long s, sa, sb;
  if (AXIS == 0) { na = ny; nb = nz; s = 1; sa = nx; sb = (long)nx * ny; }
  else if (AXIS == 1) { na = nx; nb = nz; s = nx; sa = 1; sb = (long)nx * ny; }
  else { na = nx; nb = ny; s = (long)nx * ny; sa = 1; sb = nx; }
  const int n = full.n;
  int blk = blockIdx.x;
  SegLines L;
  L.sa = sa; L.sb = sb; L.s = s;
  // one call site of seg_solve: pick the scheme, the line family and the
  // field offset of this block, then solve
  SegTab T = full;
  long off = 0;
  bool solve = true;
  if (blk < b_full) {
    if (mode) { L.a0 = 1; L.na = na - 2; L.b0 = 1; L.nb = nb - 2; }
    else { L.a0 = 0; L.na = na; L.b0 = 0; L.nb = nb; }
  } else {
    blk -= b_full;
    if (blk < nb4.x + nb4.y) {
      const int f = blk >= nb4.x;
      blk -= f ? nb4.x : 0;
      L.a0 = 1; L.na = na - 2; L.b0 = f ? nb - 1 : 0; L.nb = 1;
    } else if (blk < nb4.x + nb4.y + nb4.z + nb4.w) {
      blk -= nb4.x + nb4.y;
      const int f = blk >= nb4.z;
      blk -= f ? nb4.z : 0;
      L.a0 = f ? na - 1 : 0; L.na = 1; L.b0 = 1; L.nb = nb - 2;
    } else {
      solve = false;
    }
    T = sub;
    off = s;  // the sub-line scheme acts on points 1..n-2
  }
  if (solve) {
    L.nlines = L.na * L.nb;
    seg_solve(x + off, out + off, T, L, blk, lds);
    return;
  }
  // zeros: the four two-boundary-plane lines and both ends of the sub lines
  const int nt = blockDim.x;
  for (long t = threadIdx.x; t < 4L * n; t += nt) {
    const int line = (int)(t / n), i = (int)(t % n);
    out[((line & 1) ? na - 1 : 0) * sa + (long)((line & 2) ? nb - 1 : 0) * sb + (long)i * s] = 0.0;
  }
  const long ns = 2L * (na - 2) + 2L * (nb - 2);
  for (long t = threadIdx.x; t < ns; t += nt) {
    int a, b;
    if (t < 2L * (na - 2)) { a = 1 + (int)(t % (na - 2)); b = (t / (na - 2)) ? nb - 1 : 0; }
    else { const long r = t - 2L * (na - 2); b = 1 + (int)(r % (nb - 2)); a = (r / (nb - 2)) ? na - 1 : 0; }
    const long base = a * sa + b * sb;
    out[base] = 0.0;
    out[base + (long)(n - 1) * s] = 0.0;
  }
}

// batched lines: line l at x + l*ld, contiguous points (der/dder/der_s/dder_s);
// surplus lanes walk the last line into the scratch line (as der3d_kernel)
__global__ void __launch_bounds__(256) der_lines_kernel(const double *__restrict__ x,
                                                        double *__restrict__ out, int nlines,
                                                        long ld, LineTab T, int offset,
                                                        double *__restrict__ scratch) {
  extern __shared__ double coef[];
  const Coef C = coef_lds(T, coef);
  long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = l < nlines;
  l = act ? l : nlines - 1;
  solve_line<kCH>(x + l * ld + offset, 1, act ? out + l * ld + offset : scratch, 1, C);
}

}  // namespace mf

// ---------------------------------------------------------------------------
// Handle and C ABI
// ---------------------------------------------------------------------------
struct mf_der {
  int n = 0, so = 6, solver = 0;
  mutable mf::HostStage stage;  // device staging of the host-array entry points
  double *dbuf = nullptr;      // device tables, then one scratch line of n
  double *scratch = nullptr;   // surplus lanes' output line
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
// pack one scheme: dr row-major (7n) | p | f | d | e | c | 1/c, then the
// segmented solve's tables (seg_tables): rows (SEG_ROW n) | ab (2n) | cd (2n)
constexpr int kPackPerPoint = 13 + mf::SEG_ROW + 4;
void pack_scheme(int n, const std::vector<double> &dl, const std::vector<double> &dr,
                 std::vector<double> &dst) {
  dst.assign((size_t)kPackPerPoint * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int m = 0; m < 7; ++m) dst[(size_t)7 * i + m] = dr[i + (size_t)n * m];
  double *p = &dst[(size_t)7 * n];
  for (int i = 0; i < n; ++i) {
    p[i] = dl[i];
    p[n + i] = dl[n + i];
    p[2 * n + i] = dl[3 * (size_t)n + i];
    p[3 * n + i] = dl[4 * (size_t)n + i];
    p[4 * n + i] = dl[2 * (size_t)n + i];
    p[5 * n + i] = 1.0 / dl[2 * (size_t)n + i];
  }
  // alpha = dl(1,1) and beta = dl(n,5) of the factored matrix
  double *rows = p + 6 * n;
  mf::seg_tables(n, dr.data(), p, p + n, p + 2 * n, p + 3 * n, p + 5 * n, dl[0],
                 dl[4 * (size_t)n + n - 1], rows, rows + mf::SEG_ROW * n,
                 rows + (mf::SEG_ROW + 2) * n);
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
  if (n > mf::kMaxLine)
    return mf::fail(MF_ERR_ARG, "mf_der_create: lines longer than " + std::to_string(mf::kMaxLine) +
                                    " points exceed the per-block coefficient copy (64 KiB LDS)");
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
  hipError_t e = hipMalloc(&h->dbuf, (total + n) * sizeof(double));
  if (e == hipSuccess) h->scratch = h->dbuf + total;
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
      T.rc = T.c + m;
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
  hipLaunchKernelGGL(mf::der_lines_kernel, dim3(blocks), dim3(256), 13 * T.n * sizeof(double),
                     (hipStream_t)stream, x, out, nlines, (long)ld, T, sub ? 1 : 0, dc->scratch);
  HIPCHK(hipGetLastError());
  return MF_OK;
}

extern "C" int mf_der3d_batch(const mf_der *dc, int axis, int order, int nfields,
                              const double *const *x, double *const *out, int nx, int ny, int nz,
                              void *stream) {
  if (!dc || !x || !out || axis < 0 || axis > 2 || (order != 1 && order != 2) || nx < 1 ||
      ny < 1 || nz < 1 || nfields < 1 || nfields > mf::kMaxBatch)
    return mf::fail(MF_ERR_ARG, "mf_der3d: invalid argument");
  const int dims[3] = {nx, ny, nz};
  if (dims[axis] != dc->n)
    return mf::fail(MF_ERR_ARG, "mf_der3d: field size along the axis differs from the handle's grid");
  mf::FieldSet fs;
  std::memset(&fs, 0, sizeof(fs));
  for (int f = 0; f < nfields; ++f) {
    if (!x[f] || !out[f]) return mf::fail(MF_ERR_ARG, "mf_der3d: null field pointer");
    for (int g = 0; g < nfields; ++g)
      if (x[g] == out[f]) return mf::fail(MF_ERR_ARG, "mf_der3d: an output aliases an input");
    fs.x[f] = x[f];
    fs.out[f] = out[f];
  }
  const int na = axis == 0 ? ny : nx, nb = axis == 2 ? ny : nz;
  const mf::LineTab &F = dc->tab[order - 1][0];
  const mf::LineTab &S = dc->tab[order - 1][1];
  const int mode = dc->solver;
  int blocks, b_full = 0, b_sub = 0;
  if (!mode) {
    blocks = (int)(((long)na * nb + 255) / 256);
  } else {
    if (na < 3 || nb < 3) return mf::fail(MF_ERR_ARG, "mf_der3d: solver mode needs 3+ points across the lines");
    b_full = (int)(((long)(na - 2) * (nb - 2) + 255) / 256);
    b_sub = (int)((2L * (na - 2) + 2L * (nb - 2) + 255) / 256);
    blocks = b_full + b_sub + 1;
  }
  hipStream_t st = (hipStream_t)stream;
  if (mf::seg_count(F.n) <= mf::SEG_KMAX) {
    // segmented solves: 64 lines per block, one wavefront per segment; the
    // row tables follow the packed 1/c
    auto seg_of = [](const mf::LineTab &t) {
      if (!t.rc) return mf::SegTab{nullptr, nullptr, nullptr, 0.0, 0.0, 0};
      const double *rows = t.rc + t.n;
      return mf::SegTab{rows, rows + mf::SEG_ROW * t.n, rows + (mf::SEG_ROW + 2) * t.n,
                        t.alpha, t.beta, t.n};
    };
    const mf::SegTab SF = seg_of(F), SS = seg_of(S);
    int4 nb4 = make_int4(0, 0, 0, 0);
    int sb_full, sblocks;
    if (!mode) {
      sb_full = (int)(((long)na * nb + 63) / 64);
      sblocks = sb_full;
    } else {
      sb_full = (int)(((long)(na - 2) * (nb - 2) + 63) / 64);
      nb4.x = nb4.y = (na - 2 + 63) / 64;
      nb4.z = nb4.w = (nb - 2 + 63) / 64;
      sblocks = sb_full + nb4.x + nb4.y + nb4.z + nb4.w + 1;
    }
    const size_t lds = mf::seg_lds_doubles(F.n) * sizeof(double);
    const dim3 grid(sblocks, nfields), block(64 * mf::seg_count(F.n));
    if (axis == 0)
      hipLaunchKernelGGL(mf::der3d_seg_kernel<0>, grid, block, lds, st, fs, nx, ny, nz, SF, SS, mode, sb_full, nb4);
    else if (axis == 1)
      hipLaunchKernelGGL(mf::der3d_seg_kernel<1>, grid, block, lds, st, fs, nx, ny, nz, SF, SS, mode, sb_full, nb4);
    else
      hipLaunchKernelGGL(mf::der3d_seg_kernel<2>, grid, block, lds, st, fs, nx, ny, nz, SF, SS, mode, sb_full, nb4);
    HIPCHK(hipGetLastError());
    return MF_OK;
  }
  const size_t lds = 13 * (size_t)F.n * sizeof(double);  // the larger (full) scheme
  const dim3 grid(blocks, nfields);
  if (axis == 0)
    hipLaunchKernelGGL(mf::der3d_kernel<0>, grid, dim3(256), lds, st, fs, nx, ny, nz, F, S, mode, b_full, b_sub, dc->scratch);
  else if (axis == 1)
    hipLaunchKernelGGL(mf::der3d_kernel<1>, grid, dim3(256), lds, st, fs, nx, ny, nz, F, S, mode, b_full, b_sub, dc->scratch);
  else
    hipLaunchKernelGGL(mf::der3d_kernel<2>, grid, dim3(256), lds, st, fs, nx, ny, nz, F, S, mode, b_full, b_sub, dc->scratch);
  HIPCHK(hipGetLastError());
  return MF_OK;
}

extern "C" int mf_der3d(const mf_der *dc, int axis, int order, const double *x, double *out,
                        int nx, int ny, int nz, void *stream) {
  if (x && x == out) return mf::fail(MF_ERR_ARG, "mf_der3d: in-place derivative is not supported");
  return mf_der3d_batch(dc, axis, order, 1, &x, &out, nx, ny, nz, stream);
}

extern "C" int mf_der3d_host(const mf_der *dc, int axis, int order, const double *x, double *out,
                             int nx, int ny, int nz) {
  if (!dc || !x || !out || nx < 1 || ny < 1 || nz < 1)
    return mf::fail(MF_ERR_ARG, "mf_der3d_host: invalid argument");
  const size_t bytes = (size_t)nx * ny * nz * sizeof(double);
  std::lock_guard<std::mutex> lock(dc->stage.mu);
  hipError_t e = dc->stage.reserve(2 * bytes);
  if (e != hipSuccess)
    return mf::fail(MF_ERR_DEVICE, std::string("mf_der3d_host: staging: ") + hipGetErrorString(e));
  double *dx = (double *)dc->stage.dev, *dy = dx + (size_t)nx * ny * nz;
  int rc = MF_OK;
  e = hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    rc = mf_der3d(dc, axis, order, dx, dy, nx, ny, nz, nullptr);
    if (rc == MF_OK) e = hipMemcpy(out, dy, bytes, hipMemcpyDeviceToHost);
  }
  if (rc != MF_OK) return rc;
  if (e != hipSuccess)
    return mf::fail(MF_ERR_DEVICE, std::string("mf_der3d_host: ") + hipGetErrorString(e));
  return MF_OK;
}

extern "C" int mf_der_lines_host(const mf_der *dc, int order, int sub, const double *x, double *out,
                                 int nlines, long long ld) {
  if (!dc || !x || !out || nlines < 0 || ld < dc->n)
    return mf::fail(MF_ERR_ARG, "mf_der_lines_host: invalid argument");
  if (nlines == 0) return MF_OK;
  const size_t bytes = (size_t)nlines * (size_t)ld * sizeof(double);
  std::lock_guard<std::mutex> lock(dc->stage.mu);
  hipError_t e = dc->stage.reserve(2 * bytes);
  if (e != hipSuccess)
    return mf::fail(MF_ERR_DEVICE, std::string("mf_der_lines_host: staging: ") + hipGetErrorString(e));
  double *dx = (double *)dc->stage.dev, *dy = dx + (size_t)nlines * ld;
  e = hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice);
  // the sub-line scheme leaves the end points of out untouched
  if (e == hipSuccess && sub) e = hipMemcpy(dy, out, bytes, hipMemcpyHostToDevice);
  int rc = MF_OK;
  if (e == hipSuccess) {
    rc = mf_der_lines(dc, order, sub, dx, dy, nlines, ld, nullptr);
    if (rc == MF_OK) e = hipMemcpy(out, dy, bytes, hipMemcpyDeviceToHost);
  }
  if (rc != MF_OK) return rc;
  if (e != hipSuccess)
    return mf::fail(MF_ERR_DEVICE, std::string("mf_der_lines_host: ") + hipGetErrorString(e));
  return MF_OK;
}
