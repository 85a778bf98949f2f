// Prediction-projection Navier-Stokes time step on the device
// (navier3d.f90:48-286, class_navier_3D.f90), single domain, no mapping:
// skew-symmetric (nlt=2, the default) or advective (nlt=1) nonlinear term,
// BDF order tou and pressure extrapolation order top-1 (1..3 each),
// projection type pt = 1 (the default: non-incremental rotational correction
// with the tangential boundary-gradient correction), 2 (incremental), 3
// (velocity correction with the rot-rot pressure boundary term) or 4
// (incremental velocity correction), and up to nsubite sub-iterations per
// step, stopped by the reference's convergence test.
//
// All time levels stay in HBM.  The nonlinear term of a time level is
// computed once, when the level is first needed, and cached (the reference
// recomputes it for all nt=4 levels every step and then uses at most three,
// class_navier_3D.f90:1193-1220).  The derivatives of phi taken in the
// projection are kept per level and reused by the next step's boundary
// gradient (:729-790), and div(u*) is taken from the pressure right-hand side
// (div(fac1 u*) / fac1) instead of three more derivative passes (:1895-1899).
// The constant-flow-rate forcing (:1124-1164) updates its uniform force on
// the device, so a step never waits for the host unless a mixed-precision
// velocity solve reads back its residual norms.
#include <hip/hip_runtime.h>

#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"
#include "mflops/mf.h"

namespace mf {
namespace ns {

constexpr int NT = 4;

struct Dims {
  int nx, ny, nz;
};

__device__ __forceinline__ bool on_edge(long q, const Dims &d) {
  const int i = (int)(q % d.nx);
  const long r = q / d.nx;
  const int j = (int)(r % d.ny), k = (int)(r / d.ny);
  const int nb = (i == 0 || i == d.nx - 1) + (j == 0 || j == d.ny - 1) + (k == 0 || k == d.nz - 1);
  return nb >= 2;
}

// Products q u_j of the velocity components: uu, uv, uw, vv, vw, ww
__global__ void __launch_bounds__(256) prod6_kernel(const double *__restrict__ u,
                                                    const double *__restrict__ v,
                                                    const double *__restrict__ w,
                                                    double *__restrict__ uu, double *__restrict__ uv,
                                                    double *__restrict__ uw, double *__restrict__ vv,
                                                    double *__restrict__ vw, double *__restrict__ ww,
                                                    long N) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  const double a = u[t], b = v[t], c = w[t];
  uu[t] = a * a; uv[t] = a * b; uw[t] = a * c;
  vv[t] = b * b; vw[t] = b * c; ww[t] = c * c;
}

// The first presolve's velocity right-hand sides, formed by nl_kernel from
// the nonlinear term it has just computed (pt = 1, no LES; u, v, w are q^n):
// rhs_kernel's arithmetic, in its order.  out[0] null: not formed.
struct NlRhs {
  double *out[3];
  const double *n1[3], *n2[3];     // N at the older extrapolation levels
  const double *q1[3], *q2[3];     // q^n-1, q^n-2
  const double *force[3];          // or null
  const double *gflow[3];          // uniform force (device scalar) or null
  double c0, c1, c2, f2, f3, f4, re;
  Dims d;
};

struct NlArgs {
  const double *u, *v, *w;
  const double *g[3][3];   // g[axis][comp] = d q_comp / d x_axis
  const double *e[3][3];   // e[axis][comp] = d (q_comp u_axis) / d x_axis (nlt=2)
  double *out[3];
  long N;
  int nlt;
  NlRhs r;
};

// N_c = (u q_x + v q_y + w q_z)/2 + (q u)_x/2 + (q v)_y/2 + (q w)_z/2 (nlt=2),
// or the advective part alone (nlt=1), for c = u, v, w
__global__ void __launch_bounds__(256) nl_kernel(NlArgs a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.N) return;
  const double u = a.u[t], v = a.v[t], w = a.w[t];
  const bool rhs = a.r.out[0] != nullptr;
  const bool edge = rhs && on_edge(t, a.r.d);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double adv = u * a.g[0][c][t] + v * a.g[1][c][t] + w * a.g[2][c][t];
    double nc;
    if (a.nlt == 1) {
      nc = adv;
    } else {
      nc = adv * 0.5;
      nc = nc + a.e[0][c][t] * 0.5;
      nc = nc + a.e[1][c][t] * 0.5;
      nc = nc + a.e[2][c][t] * 0.5;
    }
    a.out[c][t] = nc;
    if (rhs) {
      const NlRhs &h = a.r;
      double ne = h.c0 * nc;
      if (h.c1 != 0.0) ne += h.c1 * h.n1[c][t];
      if (h.c2 != 0.0) ne += h.c2 * h.n2[c][t];
      if (edge) ne = 0.0;
      double f = h.force[c] ? h.force[c][t] : 0.0;
      if (h.gflow[c]) f += h.gflow[c][0];
      double rr = f - ne;
      if (edge) rr = 0.0;
      const double q0 = c == 0 ? u : (c == 1 ? v : w);
      double g = h.f2 * q0 + h.f3 * h.q1[c][t];
      if (h.f4 != 0.0) g += h.f4 * h.q2[c][t];
      h.out[c][t] = h.re * (g - rr);
    }
  }
}

// Partial sums of W * extrapol(u) per block (256 threads); extrapolation
// coefficients c0..c2 on levels a0..a2 (unused levels have c = 0).
__global__ void __launch_bounds__(256) wsum_kernel(const double *__restrict__ W,
                                                   const double *__restrict__ a0,
                                                   const double *__restrict__ a1,
                                                   const double *__restrict__ a2, double c0,
                                                   double c1, double c2, long N,
                                                   double *__restrict__ part, Dims d) {
  double s = 0.0;
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < N; t += (long)gridDim.x * 256) {
    double e = c0 * a0[t];
    if (c1 != 0.0) e += c1 * a1[t];
    if (c2 != 0.0) e += c2 * a2[t];
    if (on_edge(t, d)) e = 0.0;
    s += W[t] * e;
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// flow-rate state s = {g, debit, initialised, g0 (<= 0: 3 debit)}:
// q = sum(part)/vol; first call: debit = q (1 if q <= 1e-10), g = g0;
// then g -= (q - debit) fac1.
// One block of 1024 threads; the partials are summed in a fixed tree order
// (deterministic).
__global__ void __launch_bounds__(1024) flow_kernel(const double *__restrict__ part, int np,
                                                    double vol, double fac1,
                                                    double *__restrict__ s) {
  __shared__ double sh[1024];
  double v = 0.0;
  for (int b = threadIdx.x; b < np; b += 1024) v += part[b];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double q = sh[0] / vol;
  if (s[2] == 0.0) {
    s[1] = q > 1e-10 ? q : 1.0;
    s[0] = s[3] > 0.0 ? s[3] : 3.0 * s[1];
    s[2] = 1.0;
  }
  s[0] -= (q - s[1]) * fac1;
}

struct RhsArgs {
  const double *q0, *q1, *q2;     // q^n, q^n-1, q^n-2
  const double *n0, *n1, *n2;     // extrapolated levels of N (unused ones have c = 0)
  double f2, f3, f4;              // fac(2..4)
  double c0, c1, c2;              // extrapolation coefficients
  const double *force;            // or null
  const double *gflow;            // uniform force (device scalar) or null
  const double *les;              // -les_f: sum_j d_j (nu (du_ij + du_ji)), or null
  const double *rhs_in;           // a stored rhs1 to use instead of computing it, or null
  double *rhs_out;                // where to keep the computed rhs1, or null
  const double *dp;               // pressure-gradient component, or null
  double re;
  double *out;                    // velocity right-hand side, or null
  long N;
  Dims d;
};

// rhs1 = zero_edges(f - zero_edges(extrapol N))   (navier_presolve_u/v/w)
// g = Re (fac2 q^n + fac3 q^n-1 + fac4 q^n-2 + dp - rhs1)   (navier_solve_u/v/w)
__global__ void __launch_bounds__(256) rhs_kernel(RhsArgs a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.N) return;
  double rhs;
  if (a.rhs_in) {
    rhs = a.rhs_in[t];
  } else {
    const bool edge = on_edge(t, a.d);
    double ne = a.c0 * a.n0[t];
    if (a.c1 != 0.0) ne += a.c1 * a.n1[t];
    if (a.c2 != 0.0) ne += a.c2 * a.n2[t];
    if (edge) ne = 0.0;
    double f = a.force ? a.force[t] : 0.0;
    if (a.gflow) f += a.gflow[0];
    if (a.les) f += a.les[t];
    rhs = f - ne;
    if (edge) rhs = 0.0;
    if (a.rhs_out) a.rhs_out[t] = rhs;
  }
  if (!a.out) return;
  double g = a.f2 * a.q0[t] + a.f3 * a.q1[t];
  if (a.f4 != 0.0) g += a.f4 * a.q2[t];
  if (a.dp) g += a.dp[t];
  a.out[t] = a.re * (g - rhs);
}

// out = zero_edges(c0 a0 + c1 a1 + c2 a2) (navier_extrapol; a level with
// c = 0 is not read)
__global__ void __launch_bounds__(256) extrap_kernel(const double *__restrict__ a0,
                                                     const double *__restrict__ a1,
                                                     const double *__restrict__ a2, double c0,
                                                     double c1, double c2,
                                                     double *__restrict__ out, long N, Dims d) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  double e = 0.0;
  if (c0 != 0.0) e = c0 * a0[t];
  if (c1 != 0.0) e += c1 * a1[t];
  if (c2 != 0.0) e += c2 * a2[t];
  out[t] = on_edge(t, d) ? 0.0 : e;
}

// Right-hand sides rhs_p of the pressure equation (navier_phi_rhs,
// class_navier_3D.f90:1961-2007), zero on the edges, for the three components:
//   mode 0 (pt <= 2):              fac1 q(it1)
//   mode 3 (pt = 3):               rhs1 - fac2 q^n - fac3 q^n-1 - fac4 q^n-2
//   mode 4 (pt = 4, subite = 1):   mode 3 - rhs2 + fac1 q^n + fac2 q^n-1 + fac3 q^n-2 + fac4 q(it1)
//   mode 5 (pt = 4, subite > 1):   rhs1 - rhs2 + fac1 q(it1)
struct RpArgs {
  const double *qi[3], *q0[3], *q1[3], *q2[3];  // levels it(1), it(nt), it(nt-1), it(nt-2)
  const double *r1[3], *r2[3];
  double f1, f2, f3, f4;
  int mode;
  double *out[3];
  long N;
  Dims d;
};

__global__ void __launch_bounds__(256) rp_kernel(RpArgs a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.N) return;
  const bool e = on_edge(t, a.d);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double r;
    if (a.mode == 0) {
      r = a.f1 * a.qi[c][t];
    } else if (a.mode == 5) {
      r = a.r1[c][t] - a.r2[c][t];
      r = r + a.f1 * a.qi[c][t];
    } else {
      const double x0 = a.q0[c][t], x1 = a.q1[c][t], x2 = a.q2[c][t];
      r = a.r1[c][t] - a.f2 * x0;
      r = r - a.f3 * x1;
      r = r - a.f4 * x2;
      if (a.mode == 4) {
        r = r - a.r2[c][t];
        r = r + a.f1 * x0;
        r = r + a.f2 * x1;
        r = r + a.f3 * x2;
        r = r + a.f4 * a.qi[c][t];
      }
    }
    a.out[c][t] = e ? 0.0 : r;
  }
}

// add_boundary_rotrot (class_navier_3D.f90:793-846) on the two faces normal
// to axis `axis`: bcphi -= (d_a(d_b e_b + d_c e_c) - dd_b e_a - dd_c e_a)/Re,
// with t2 = d_a(...), db = dd_b e_a, dc = dd_c e_a given as fields
__global__ void __launch_bounds__(256) rotrot_kernel(int axis, const double *__restrict__ t2,
                                                     const double *__restrict__ db,
                                                     const double *__restrict__ dc, double re,
                                                     double *__restrict__ bcphi, Dims d) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const long fx = (long)(ny - 2) * (nz - 2), fy = (long)(nx - 2) * (nz - 2), fz = (long)(nx - 2) * (ny - 2);
  const long n = axis == 0 ? fx : (axis == 1 ? fy : fz);
  if (t >= 2 * n) return;
  const int side = (int)(t / n);
  const long r = t % n;
  long i, j, k, off;
  if (axis == 0) {
    j = 1 + r % (ny - 2); k = 1 + r / (ny - 2); i = side ? nx - 1 : 0; off = t;
  } else if (axis == 1) {
    i = 1 + r % (nx - 2); k = 1 + r / (nx - 2); j = side ? ny - 1 : 0; off = 2 * fx + t;
  } else {
    i = 1 + r % (nx - 2); j = 1 + r / (nx - 2); k = side ? nz - 1 : 0; off = 2 * (fx + fy) + t;
  }
  const long q = i + (long)nx * (j + (long)ny * k);
  bcphi[off] -= (t2[q] - db[q] - dc[q]) / re;
}

// Partial sums of the sub-iteration test (testconv, navier3d.f90:366-412) per
// block: W |q|^2, W |q - q_sub|^2 and W (p - p_sub)
__global__ void __launch_bounds__(256) conv1_kernel(const double *__restrict__ W,
                                                    const double *__restrict__ u,
                                                    const double *__restrict__ v,
                                                    const double *__restrict__ w,
                                                    const double *__restrict__ su,
                                                    const double *__restrict__ sv,
                                                    const double *__restrict__ sw,
                                                    const double *__restrict__ p,
                                                    const double *__restrict__ sp, long N,
                                                    double *__restrict__ part) {
  double a = 0.0, b = 0.0, c = 0.0;
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < N; t += (long)gridDim.x * 256) {
    const double wt = W[t];
    const double x = u[t], y = v[t], z = w[t];
    const double dx = su[t] - x, dy = sv[t] - y, dz = sw[t] - z;
    a += wt * (x * x + y * y + z * z);
    b += wt * (dx * dx + dy * dy + dz * dz);
    c += wt * (p[t] - sp[t]);
  }
  __shared__ double sh[3][256];
  sh[0][threadIdx.x] = a;
  sh[1][threadIdx.x] = b;
  sh[2][threadIdx.x] = c;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int m = 0; m < 3; ++m) sh[m][threadIdx.x] += sh[m][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[threadIdx.x * gridDim.x + blockIdx.x] = sh[threadIdx.x][0];
}

// second pass: with m = sum W (p - p_sub) / sum W, partial sums of
// W (p_sub + m)^2 and W (p - p_sub - m)^2
__global__ void __launch_bounds__(256) conv2_kernel(const double *__restrict__ W,
                                                    const double *__restrict__ p,
                                                    const double *__restrict__ sp,
                                                    const double *__restrict__ sums, double vol,
                                                    long N, double *__restrict__ part) {
  const double m = sums[2] / vol;
  double a = 0.0, b = 0.0;
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < N; t += (long)gridDim.x * 256) {
    const double wt = W[t];
    const double s = sp[t] + m, e = p[t] - s;
    a += wt * s * s;
    b += wt * e * e;
  }
  __shared__ double sh[2][256];
  sh[0][threadIdx.x] = a;
  sh[1][threadIdx.x] = b;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int k = 0; k < 2; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 2) part[threadIdx.x * gridDim.x + blockIdx.x] = sh[threadIdx.x][0];
}

// out[q] = sum of part[q np .. q np + np - 1] in a fixed tree order, q < nq
// (one block of 1024 threads per quantity)
__global__ void __launch_bounds__(1024) psum_kernel(const double *__restrict__ part, int np,
                                                    double *__restrict__ out) {
  __shared__ double sh[1024];
  const double *pp = part + (long)blockIdx.x * np;
  double v = 0.0;
  for (int b = threadIdx.x; b < np; b += 1024) v += pp[b];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

struct BcArgs {
  // packed per component: bcx (ny-2)(nz-2)2 | bcy (nx-2)(nz-2)2 | bcz (nx-2)(ny-2)2
  const double *bc0[3];
  double *bcw[3];
  const double *dp[3][3];  // [level slot][axis] derivatives of phi (slot unused: null)
  double c[3];             // extrapolation coefficients of the slots
  double inv_fac1;
  Dims d;
};

// bc(it1) = bc0 + tangential grad(extrapol_p phi)/fac1 (add_boundary_gradient)
__global__ void __launch_bounds__(256) bc_kernel(BcArgs a) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = a.d.nx, ny = a.d.ny, nz = a.d.nz;
  const long fx = (long)(ny - 2) * (nz - 2), fy = (long)(nx - 2) * (nz - 2), fz = (long)(nx - 2) * (ny - 2);
  const long per = 2 * (fx + fy + fz);
  if (t >= 3 * per) return;
  const int comp = (int)(t / per);
  long r = t % per;
  const long off = r;
  int face;  // 0 x, 1 y, 2 z
  long i, j, k;
  int side;
  if (r < 2 * fx) {
    face = 0; side = (int)(r / fx); r %= fx;
    j = 1 + r % (ny - 2); k = 1 + r / (ny - 2); i = side ? nx - 1 : 0;
  } else if ((r -= 2 * fx) < 2 * fy) {
    face = 1; side = (int)(r / fy); r %= fy;
    i = 1 + r % (nx - 2); k = 1 + r / (nx - 2); j = side ? ny - 1 : 0;
  } else {
    r -= 2 * fy;
    face = 2; side = (int)(r / fz); r %= fz;
    i = 1 + r % (nx - 2); j = 1 + r / (nx - 2); k = side ? nz - 1 : 0;
  }
  double v = a.bc0[comp][off];
  // tangential component: the derivative along the component's own axis,
  // taken on faces normal to another axis
  if (comp != face) {
    const long q = i + (long)nx * (j + (long)ny * k);
    double g = 0.0;
    for (int s = 0; s < 3; ++s)
      if (a.dp[s][comp]) g += a.c[s] * a.dp[s][comp][q];
    v += g * a.inv_fac1;
  }
  a.bcw[comp][off] = v;
}

__global__ void __launch_bounds__(256) sum3_kernel(const double *__restrict__ a,
                                                   const double *__restrict__ b,
                                                   const double *__restrict__ c,
                                                   double *__restrict__ out, long N) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  out[t] = a[t] + b[t] + c[t];
}

__global__ void __launch_bounds__(256) sum2_kernel(const double *__restrict__ a,
                                                   const double *__restrict__ b,
                                                   double *__restrict__ out, long N) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  out[t] = a[t] + b[t];
}

// bcphi faces: rhs_p at the face minus fac1 times the normal velocity bc
// (navier_bc_pressure, class_navier_3D.f90:368-384)
__global__ void __launch_bounds__(256) bcphi_kernel(const double *__restrict__ ru,
                                                    const double *__restrict__ rv,
                                                    const double *__restrict__ rw,
                                                    const double *__restrict__ bu,
                                                    const double *__restrict__ bv,
                                                    const double *__restrict__ bw, double fac1,
                                                    double *__restrict__ out, Dims d) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const long fx = (long)(ny - 2) * (nz - 2), fy = (long)(nx - 2) * (nz - 2), fz = (long)(nx - 2) * (ny - 2);
  const long off = t;
  if (t < 2 * fx) {
    const int side = (int)(t / fx);
    const long r = t % fx;
    const long j = 1 + r % (ny - 2), k = 1 + r / (ny - 2), i = side ? nx - 1 : 0;
    out[off] = ru[i + (long)nx * (j + (long)ny * k)] - fac1 * bu[off];
  } else if ((t -= 2 * fx) < 2 * fy) {  // t: index within the y faces
    const int side = (int)(t / fy);
    const long r = t % fy;
    const long i = 1 + r % (nx - 2), k = 1 + r / (nx - 2), j = side ? ny - 1 : 0;
    out[off] = rv[i + (long)nx * (j + (long)ny * k)] - fac1 * bv[t];
  } else if ((t -= 2 * fy) < 2 * fz) {  // t: index within the z faces
    const int side = (int)(t / fz);
    const long r = t % fz;
    const long i = 1 + r % (nx - 2), j = 1 + r / (nx - 2), k = side ? nz - 1 : 0;
    out[off] = rw[i + (long)nx * (j + (long)ny * k)] - fac1 * bw[t];
  }
}

__global__ void __launch_bounds__(256) zero_edges_kernel(double *__restrict__ f, Dims d) {
  // one thread per edge point: 4 x-lines, 4 y-lines, 4 z-lines
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  long i, j, k;
  if (t < 4L * nx) {
    i = t % nx; const int e = (int)(t / nx); j = (e & 1) ? ny - 1 : 0; k = (e & 2) ? nz - 1 : 0;
  } else if ((t -= 4L * nx) < 4L * ny) {
    j = t % ny; const int e = (int)(t / ny); i = (e & 1) ? nx - 1 : 0; k = (e & 2) ? nz - 1 : 0;
  } else if ((t -= 4L * ny) < 4L * nz) {
    k = t % nz; const int e = (int)(t / nz); i = (e & 1) ? nx - 1 : 0; j = (e & 2) ? ny - 1 : 0;
  } else {
    return;
  }
  f[i + (long)nx * (j + (long)ny * k)] = 0.0;
}

// p = zero_edges(phi - fphi/(fac1 Re)); q_c = zero_edges((rp_c - dphi_c)/fac1)
// aux_i = nu (du_ij + du_ji), i = 0..2, for one derivative direction j
// (f(nav,var), class_navier_3D.f90:1090-1099)
__global__ void __launch_bounds__(256) les_aux_kernel(const double *__restrict__ nu,
                                                      const double *__restrict__ a0,
                                                      const double *__restrict__ a1,
                                                      const double *__restrict__ a2,
                                                      const double *__restrict__ b0,
                                                      const double *__restrict__ b1,
                                                      const double *__restrict__ b2,
                                                      double *__restrict__ o0,
                                                      double *__restrict__ o1,
                                                      double *__restrict__ o2, long N) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  const double n = nu[t];
  o0[t] = n * (a0[t] + b0[t]);
  o1[t] = n * (a1[t] + b1[t]);
  o2[t] = n * (a2[t] + b2[t]);
}

// acc_i (=|+=) d_i, i = 0..2
__global__ void __launch_bounds__(256) les_acc_kernel(const double *__restrict__ d0,
                                                      const double *__restrict__ d1,
                                                      const double *__restrict__ d2,
                                                      double *__restrict__ c0,
                                                      double *__restrict__ c1,
                                                      double *__restrict__ c2, int first, long N) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  if (first) {
    c0[t] = d0[t]; c1[t] = d1[t]; c2[t] = d2[t];
  } else {
    c0[t] += d0[t]; c1[t] += d1[t]; c2[t] += d2[t];
  }
}

struct ProjArgs {
  const double *phi;
  const double *pe;        // extrapol_p(p) (pt = 2, 4) or null
  const double *dv;        // divergence term or null
  int dv_fac1;             // 1: dv is div(fac1 u*) (pt <= 2), 0: div(extrapol_p(q)) (pt = 4)
  const double *r[3];      // rhs_p
  const double *g[3];      // grad phi
  double fac1, re;
  double *p, *q[3];
  long N;
  Dims d;
};

// navier_projection (:1847-1959): p = zero_edges(phi [+ pe] [- div/Re]);
// q_c = zero_edges((rhs_pc - d_c phi)/fac1)
__global__ void __launch_bounds__(256) proj_kernel(ProjArgs a) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.N) return;
  const bool e = on_edge(t, a.d);
  double pv = a.phi[t];
  if (a.pe) pv = pv + a.pe[t];
  if (a.dv) pv = pv - (a.dv_fac1 ? (a.dv[t] / a.fac1) / a.re : a.dv[t] / a.re);
  a.p[t] = e ? 0.0 : pv;
#pragma unroll
  for (int c = 0; c < 3; ++c) a.q[c][t] = e ? 0.0 : (a.r[c][t] - a.g[c][t]) / a.fac1;
}

__global__ void __launch_bounds__(256) weights_kernel(const double *__restrict__ gx,
                                                      const double *__restrict__ gy,
                                                      const double *__restrict__ gz,
                                                      double *__restrict__ W, long N, Dims d) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  if (on_edge(t, d)) { W[t] = 0.0; return; }
  const int i = (int)(t % d.nx);
  const long r = t / d.nx;
  const int j = (int)(r % d.ny), k = (int)(r / d.ny);
  auto w1 = [](const double *g, int n, int a) {
    if (a == 0) return (g[1] - g[0]) * 0.5;
    if (a == n - 1) return (g[n - 1] - g[n - 2]) * 0.5;
    return (g[a + 1] - g[a - 1]) * 0.5;
  };
  W[t] = w1(gx, d.nx, i) * w1(gy, d.ny, j) * w1(gz, d.nz, k);
}

}  // namespace ns
}  // namespace mf

using namespace mf::ns;

struct mf_ns {
  int nx = 0, ny = 0, nz = 0;
  long N = 0, nbc = 0;  // points; packed bc doubles per component
  double re = 1.0, dt = 1.0, fac[4] = {0, 0, 0, 0};
  int tou = 2, top = 2, nlt = 2;
  mf_der *dc[3] = {nullptr, nullptr, nullptr};
  mf_solver3d *scu = nullptr, *scp = nullptr;
  // the v and w velocity solves on handles and streams of their own, so that
  // the three solves overlap (MF_NS_CONCURRENT=0: one handle, in turn)
  mf_solver3d *scu_vw[2] = {nullptr, nullptr};
  hipStream_t side[2] = {nullptr, nullptr};
  hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
  void *mem = nullptr;
  double *q[3][NT], *p[NT], *phi[NT], *nl[3][NT], *dphi[NT][3];
  bool nl_ok[NT], dphi_ok[NT];
  double *bc0[3], *bcw[3], *bcphi;
  double *force[3];
  bool has_force[3] = {false, false, false};
  bool flow = false;
  double flow_g0 = 0.0;
  double *W = nullptr, *fstate = nullptr, *part = nullptr;
  double vol = 0.0;
  int nparts = 0;
  double *T[24];
  int it[NT] = {0, 1, 2, 3};
  double time = 0.0;
  mf::ns::Dims d{0, 0, 0};
  std::vector<double> gx, gy, gz;
  int grad_level = -1;  // level whose velocity gradients T[0..8] hold
  // the first presolve's velocity right-hand sides formed by nl_kernel
  // (pt = 1, no LES; MF_NS_FUSED_RHS=0: by rhs_kernel)
  bool fused_rhs = true;
  // LES (navier_LES): model, constant squared, fields -les_f (3) and nu,
  // squared spacings
  int les_type = 0;
  double les_c2 = 0.0;
  void *lesmem = nullptr;
  double *lesf[3] = {nullptr, nullptr, nullptr}, *lesnu = nullptr, *d2 = nullptr;
  // learned closure: tau = tscale MLP(gscale du, lscale Delta)
  const mf_mlp *sgs = nullptr;
  double sgs_g = 1.0, sgs_l = 1.0, sgs_t = 1.0;
  int sgs_coupling = 0;  // 0: -div tau; 1: eddy viscosity max(0, -tau:S)/(2 S:S)
  void *sgsmem = nullptr;
  float *sgsX = nullptr, *sgsY = nullptr;
  // projection type and sub-iterations (navier_initialization :2133,2136):
  // rhs1/rhs2 of the presolves for pt >= 3 (rhsb[rhs_cur] is rhs1), the
  // previous sub-iterate and the partial sums of its test for nsubite > 1
  int pt = 1, nsubite = 1;
  void *schmem = nullptr;
  double *rhsb[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  int rhs_cur = 0;
  double *subq[3] = {nullptr, nullptr, nullptr}, *subp = nullptr;
  double *cpart = nullptr, *csum = nullptr;
  int subite_done = 0;
  double conv[2] = {0.0, 0.0};
};

namespace {
#define NSCHK(call)                                                                   \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return mf::fail(MF_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define RC(call)                   \
  do {                             \
    int rc_ = (call);              \
    if (rc_ != MF_OK) return rc_;  \
  } while (0)

inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

// Three independent launches launch(k, stream), k = 0, 1, 2 (one per axis or
// component): k = 0 on st, 1 and 2 on the side streams when they exist, with
// a fork and a join by events; otherwise all three on st in turn.
template <class F>
int fork3(mf_ns *s, hipStream_t st, F &&launch) {
  if (!s->side[0]) {
    for (int k = 0; k < 3; ++k) RC(launch(k, st));
    return MF_OK;
  }
  NSCHK(hipEventRecord(s->ev_fork, st));
  for (int k = 0; k < 2; ++k) NSCHK(hipStreamWaitEvent(s->side[k], s->ev_fork, 0));
  hipStream_t ss[3] = {st, s->side[0], s->side[1]};
  for (int k = 0; k < 3; ++k) RC(launch(k, ss[k]));
  for (int k = 0; k < 2; ++k) {
    NSCHK(hipEventRecord(s->ev_join[k], s->side[k]));
    NSCHK(hipStreamWaitEvent(st, s->ev_join[k], 0));
  }
  return MF_OK;
}

// extrapolation coefficients on levels it(nt), it(nt-1), it(nt-2)
void extrap_coefs(int order, double c[3]) {
  c[0] = c[1] = c[2] = 0.0;
  if (order == 1) { c[0] = 1.0; }
  else if (order == 2) { c[0] = 2.0; c[1] = -1.0; }
  else if (order == 3) { c[0] = 3.0; c[1] = -3.0; c[2] = 1.0; }
}

int der(mf_ns *s, int axis, const double *x, double *out, hipStream_t st) {
  return mf_der3d(s->dc[axis], axis, 1, x, out, s->nx, s->ny, s->nz, st);
}

int der3(mf_ns *s, int axis, const double *x0, const double *x1, const double *x2, double *o0,
         double *o1, double *o2, hipStream_t st) {
  const double *x[3] = {x0, x1, x2};
  double *o[3] = {o0, o1, o2};
  return mf_der3d_batch(s->dc[axis], axis, 1, 3, x, o, s->nx, s->ny, s->nz, st);
}

// nonlinear terms of the three components at level L (cached): the 9
// velocity gradients and, for the skew-symmetric form, the 9 product
// derivatives, one batched launch per axis (6 fields: the line recurrences
// are latency-bound, so a wider batch costs far less than a second launch)
int nonlinear(mf_ns *s, int L, hipStream_t st, const NlRhs *rhs = nullptr) {
  if (s->nl_ok[L]) return MF_OK;
  const long N = s->N;
  double *u = s->q[0][L], *v = s->q[1][L], *w = s->q[2][L];
  double **T = s->T;  // T[3a + c] gradients, T[9..14] products, T[15 + 3a + c] product derivatives
  NlArgs n;
  std::memset(&n, 0, sizeof(n));
  n.u = u; n.v = v; n.w = w;
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) n.g[a][c] = T[3 * a + c];
  if (s->nlt == 2) {
    double *uu = T[9], *uv = T[10], *uw = T[11], *vv = T[12], *vw = T[13], *ww = T[14];
    hipLaunchKernelGGL(prod6_kernel, dim3(nblk(N)), dim3(256), 0, st, (const double *)u,
                       (const double *)v, (const double *)w, uu, uv, uw, vv, vw, ww, N);
    // axis a: d_a of (u, v, w) and of (q u_a), q = u, v, w
    const double *prod[3][3] = {{uu, uv, uw}, {uv, vv, vw}, {uw, vw, ww}};
    RC(fork3(s, st, [&](int a, hipStream_t q) {
      const double *x[6] = {u, v, w, prod[a][0], prod[a][1], prod[a][2]};
      double *o[6] = {T[3 * a], T[3 * a + 1], T[3 * a + 2], T[15 + 3 * a], T[16 + 3 * a],
                      T[17 + 3 * a]};
      return mf_der3d_batch(s->dc[a], a, 1, 6, x, o, s->nx, s->ny, s->nz, q);
    }));
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) n.e[a][c] = T[15 + 3 * a + c];
  } else {
    RC(fork3(s, st, [&](int a, hipStream_t q) {
      return der3(s, a, u, v, w, T[3 * a], T[3 * a + 1], T[3 * a + 2], q);
    }));
  }
  // the fused right-hand sides overwrite the gradients T[0..2] (each point
  // after its own reads)
  s->grad_level = rhs ? -1 : L;
  for (int c = 0; c < 3; ++c) n.out[c] = s->nl[c][L];
  n.N = N;
  n.nlt = s->nlt;
  if (rhs) n.r = *rhs;
  hipLaunchKernelGGL(nl_kernel, dim3(nblk(N)), dim3(256), 0, st, n);
  NSCHK(hipGetLastError());
  s->nl_ok[L] = true;
  return MF_OK;
}

// navier_LES (:1225-1452) on level L and the LES part of f(nav,var)
// (:1083-1134): lesf_i = -les_f_i = sum_j d_j (nu (du_ij + du_ji)).  Uses
// T[0..8] (gradients), T[9..14] and keeps nu in lesnu.
int les_force(mf_ns *s, int L, hipStream_t st) {
  double **T = s->T;
  const long N = s->N;
  if (s->grad_level != L) {
    RC(fork3(s, st, [&](int a, hipStream_t q) {
      return der3(s, a, s->q[0][L], s->q[1][L], s->q[2][L], T[3 * a], T[3 * a + 1], T[3 * a + 2], q);
    }));
    s->grad_level = L;
  }
  const double *du[9];  // du[3 i + j] = d u_i / d x_j = T[3 j + i]
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) du[3 * i + j] = T[3 * j + i];
  bool first = true;
  if (s->les_type) {
    mf::les_nu_launch(du, s->d2, s->d2 + s->nx, s->d2 + s->nx + s->ny, s->nx, s->ny, s->nz,
                      s->les_type, s->les_c2, s->lesnu, st);
    for (int j = 0; j < 3; ++j) {
      hipLaunchKernelGGL(les_aux_kernel, dim3(nblk(N)), dim3(256), 0, st,
                         (const double *)s->lesnu, du[j], du[3 + j], du[6 + j], du[3 * j],
                         du[3 * j + 1], du[3 * j + 2], T[9], T[10], T[11], N);
      RC(der3(s, j, T[9], T[10], T[11], T[12], T[13], T[14], st));
      hipLaunchKernelGGL(les_acc_kernel, dim3(nblk(N)), dim3(256), 0, st, (const double *)T[12],
                         (const double *)T[13], (const double *)T[14], s->lesf[0], s->lesf[1],
                         s->lesf[2], (int)first, N);
      first = false;
    }
  }
  if (s->sgs && s->sgs_coupling == 1) {  // the learned stress as an eddy viscosity
    mf::sgs_features_launch(du, s->d2, s->d2 + s->nx, s->d2 + s->nx + s->ny, s->nx, s->ny, s->nz,
                            s->sgs_g, s->sgs_l, s->sgsX, st);
    RC(mf_mlp_forward(s->sgs, s->sgsX, s->sgsY, N, st));
    mf::sgs_nu_launch(s->sgsY, du, s->sgs_t, s->lesnu, s->nx, s->ny, s->nz, st);
    for (int j = 0; j < 3; ++j) {
      hipLaunchKernelGGL(les_aux_kernel, dim3(nblk(N)), dim3(256), 0, st,
                         (const double *)s->lesnu, du[j], du[3 + j], du[6 + j], du[3 * j],
                         du[3 * j + 1], du[3 * j + 2], T[9], T[10], T[11], N);
      RC(der3(s, j, T[9], T[10], T[11], T[12], T[13], T[14], st));
      hipLaunchKernelGGL(les_acc_kernel, dim3(nblk(N)), dim3(256), 0, st, (const double *)T[12],
                         (const double *)T[13], (const double *)T[14], s->lesf[0], s->lesf[1],
                         s->lesf[2], (int)first, N);
      first = false;
    }
  } else if (s->sgs) {  // -div tau of the learned closure
    mf::sgs_features_launch(du, s->d2, s->d2 + s->nx, s->d2 + s->nx + s->ny, s->nx, s->ny, s->nz,
                            s->sgs_g, s->sgs_l, s->sgsX, st);
    RC(mf_mlp_forward(s->sgs, s->sgsX, s->sgsY, N, st));
    for (int j = 0; j < 3; ++j) {
      mf::sgs_aux_launch(s->sgsY, j, s->sgs_t, T[9], T[10], T[11], s->nx, s->ny, s->nz, st);
      RC(der3(s, j, T[9], T[10], T[11], T[12], T[13], T[14], st));
      hipLaunchKernelGGL(les_acc_kernel, dim3(nblk(N)), dim3(256), 0, st, (const double *)T[12],
                         (const double *)T[13], (const double *)T[14], s->lesf[0], s->lesf[1],
                         s->lesf[2], (int)first, N);
      first = false;
    }
  }
  NSCHK(hipGetLastError());
  return MF_OK;
}

int dder(mf_ns *s, int axis, const double *x, double *out, hipStream_t st) {
  return mf_der3d(s->dc[axis], axis, 2, x, out, s->nx, s->ny, s->nz, st);
}

// The extrapolation navier_extrapol makes in sub-iteration `sub`: the
// order's coefficients on it(nt), it(nt-1), it(nt-2) in the first, level
// it(1) (the current iterate) with coefficient 1 in the later ones.
struct Extrap {
  int lev[3];
  double c[3];
};

Extrap extrap_of(const mf_ns *s, int order, int sub) {
  Extrap e;
  const int *it = s->it;
  if (sub > 1) {
    e.lev[0] = e.lev[1] = e.lev[2] = it[0];
    e.c[0] = 1.0; e.c[1] = e.c[2] = 0.0;
  } else {
    e.lev[0] = it[NT - 1]; e.lev[1] = it[NT - 2]; e.lev[2] = it[NT - 3];
    extrap_coefs(order, e.c);
  }
  return e;
}

void launch_extrap(mf_ns *s, double *const lv[NT], const Extrap &e, double *out, hipStream_t st) {
  hipLaunchKernelGGL(extrap_kernel, dim3(nblk(s->N)), dim3(256), 0, st, (const double *)lv[e.lev[0]],
                     (const double *)lv[e.lev[1]], (const double *)lv[e.lev[2]], e.c[0], e.c[1],
                     e.c[2], out, s->N, s->d);
}

// Velocity solves of navier_solve_u/v/w into q(it1): right-hand side from
// rhs_kernel (rhs1 computed from the cached nonlinear terms ex, or read from
// rhsb when rhs_in), the pressure gradient dp[c] when given, boundary values
// bc[c]
// rhs_done: the right-hand sides are in T[0..2] already (formed by nl_kernel).
int velocity_solves(mf_ns *s, const Extrap &ev, bool rhs_in, double *const dp[3],
                    const double *const bc[3], int precision, double tol, hipStream_t st,
                    bool rhs_done = false) {
  const int *it = s->it;
  const int n1 = it[0], L0 = it[NT - 1], L1 = it[NT - 2], L2 = it[NT - 3];
  const long sx = 2L * (s->ny - 2) * (s->nz - 2), sy = 2L * (s->nx - 2) * (s->nz - 2);
  const double sigma = -s->re * s->fac[0];
  auto rhs = [&](int c, double *out) {
    RhsArgs r;
    std::memset(&r, 0, sizeof(r));
    r.q0 = s->q[c][L0]; r.q1 = s->q[c][L1]; r.q2 = s->q[c][L2];
    r.f2 = s->fac[1]; r.f3 = s->fac[2]; r.f4 = s->tou == 3 ? s->fac[3] : 0.0;
    if (rhs_in) {
      r.rhs_in = s->rhsb[s->rhs_cur][c];
    } else {
      r.n0 = s->nl[c][ev.lev[0]]; r.n1 = s->nl[c][ev.lev[1]]; r.n2 = s->nl[c][ev.lev[2]];
      r.c0 = ev.c[0]; r.c1 = ev.c[1]; r.c2 = ev.c[2];
      r.force = s->has_force[c] ? s->force[c] : nullptr;
      r.gflow = (c == 0 && s->flow) ? s->fstate : nullptr;
      r.les = (s->les_type || s->sgs) ? s->lesf[c] : nullptr;
    }
    r.dp = dp ? dp[c] : nullptr;
    r.re = s->re;
    r.out = out;
    r.N = s->N;
    r.d = s->d;
    hipLaunchKernelGGL(rhs_kernel, dim3(nblk(s->N)), dim3(256), 0, st, r);
  };
  if (!s->scu_vw[0]) {
    for (int c = 0; c < 3; ++c) {
      double *g = rhs_done ? s->T[c] : s->T[0];
      if (!rhs_done) rhs(c, g);
      NSCHK(hipGetLastError());
      RC(mf_solver3d_solve(s->scu, g, sigma, s->q[c][n1], bc[c], bc[c] + sx,
                           bc[c] + sx + sy, precision, tol, 0, nullptr, st));
    }
  } else {
    // right-hand sides into T[0..2] on st; u's solve on st, v's and w's on
    // the side streams after them; the host waits for all three (mixed
    // precision checks each residual), and st waits for the side streams
    if (!rhs_done)
      for (int c = 0; c < 3; ++c) rhs(c, s->T[c]);
    NSCHK(hipGetLastError());
    NSCHK(hipEventRecord(s->ev_fork, st));
    for (int k = 0; k < 2; ++k) NSCHK(hipStreamWaitEvent(s->side[k], s->ev_fork, 0));
    mf_solver3d *hs[3] = {s->scu, s->scu_vw[0], s->scu_vw[1]};
    hipStream_t ss[3] = {st, s->side[0], s->side[1]};
    for (int c = 0; c < 3; ++c)
      RC(mf::solver3d_solve_begin(hs[c], s->T[c], sigma, s->q[c][n1], bc[c], bc[c] + sx,
                                  bc[c] + sx + sy, precision, tol, 0, ss[c]));
    for (int c = 0; c < 3; ++c) RC(mf::solver3d_solve_end(hs[c], nullptr));
    for (int k = 0; k < 2; ++k) {
      NSCHK(hipEventRecord(s->ev_join[k], s->side[k]));
      NSCHK(hipStreamWaitEvent(st, s->ev_join[k], 0));
    }
  }
  s->nl_ok[n1] = false;
  return MF_OK;
}

// testconv (navier3d.f90:366-412) on level it(1) against the previous
// sub-iterate: relative changes of |q| and of p (mean difference removed),
// norme2 = sqrt(sum W x^2); true when both are below 1e-9.  Synchronises.
int testconv(mf_ns *s, bool *done, hipStream_t st) {
  const int n1 = s->it[0];
  const int np = s->nparts;
  hipLaunchKernelGGL(conv1_kernel, dim3(np), dim3(256), 0, st, (const double *)s->W,
                     (const double *)s->q[0][n1], (const double *)s->q[1][n1],
                     (const double *)s->q[2][n1], (const double *)s->subq[0],
                     (const double *)s->subq[1], (const double *)s->subq[2],
                     (const double *)s->p[n1], (const double *)s->subp, s->N, s->cpart);
  hipLaunchKernelGGL(psum_kernel, dim3(3), dim3(1024), 0, st, (const double *)s->cpart, np, s->csum);
  hipLaunchKernelGGL(conv2_kernel, dim3(np), dim3(256), 0, st, (const double *)s->W,
                     (const double *)s->p[n1], (const double *)s->subp, (const double *)s->csum,
                     s->vol, s->N, s->cpart);
  hipLaunchKernelGGL(psum_kernel, dim3(2), dim3(1024), 0, st, (const double *)s->cpart, np,
                     s->csum + 3);
  NSCHK(hipGetLastError());
  double h[5];
  NSCHK(hipMemcpyAsync(h, s->csum, sizeof(h), hipMemcpyDeviceToHost, st));
  NSCHK(hipStreamSynchronize(st));
  const double e1 = std::sqrt(h[1]) / std::sqrt(h[0]);
  const double e2 = std::sqrt(h[4]) / std::sqrt(h[3]);
  s->conv[0] = e1;
  s->conv[1] = e2;
  *done = e1 < 1e-9 && e2 < 1e-9;
  return MF_OK;
}

// One pass of the sub-iteration loop (navier3d.f90:73-147, 254-255) on level
// it(1); *done when testconv ends the loop.
int substep(mf_ns *s, int sub, int precision, double tol, bool *done, hipStream_t st) {
  const int *it = s->it;
  const int n1 = it[0], L0 = it[NT - 1], L1 = it[NT - 2], L2 = it[NT - 3];
  const long N = s->N;
  const int pt = s->pt;
  const double fac1 = s->fac[0];
  double **T = s->T;
  const Extrap ev = extrap_of(s, s->tou, sub), ep = extrap_of(s, s->top - 1, sub);
  // The flow-rate force and the velocity boundary values depend on neither
  // the nonlinear terms nor the LES force: with the side streams they run on
  // side stream 1 beside the product kernel of the nonlinear terms (the
  // derivatives those fork to side stream 1 queue after them), and st joins
  // them before the presolve.
  hipStream_t sp = st;
  if (s->side[1]) {
    sp = s->side[1];
    NSCHK(hipEventRecord(s->ev_fork, st));
    NSCHK(hipStreamWaitEvent(sp, s->ev_fork, 0));
  }
  // constant-flow-rate force from extrapol_v(u), updated at every presolve
  if (s->flow) {
    hipLaunchKernelGGL(wsum_kernel, dim3(s->nparts), dim3(256), 0, sp, (const double *)s->W,
                       (const double *)s->q[0][ev.lev[0]], (const double *)s->q[0][ev.lev[1]],
                       (const double *)s->q[0][ev.lev[2]], ev.c[0], ev.c[1], ev.c[2], N, s->part,
                       s->d);
    hipLaunchKernelGGL(flow_kernel, dim3(1), dim3(1024), 0, sp, (const double *)s->part, s->nparts,
                       s->vol, fac1, s->fstate);
  }
  // velocity boundary values: with the tangential gradient of extrapol_p(phi)
  // for pt = 1 (add_boundary_gradient), as given otherwise
  const double *bc[3] = {s->bc0[0], s->bc0[1], s->bc0[2]};
  if (pt == 1) {
    BcArgs b;
    std::memset(&b, 0, sizeof(b));
    for (int c = 0; c < 3; ++c) { b.bc0[c] = s->bc0[c]; b.bcw[c] = s->bcw[c]; bc[c] = s->bcw[c]; }
    for (int m = 0; m < 3; ++m) {
      b.c[m] = ep.c[m];
      if (ep.c[m] != 0.0 && s->dphi_ok[ep.lev[m]])
        for (int a = 0; a < 3; ++a) b.dp[m][a] = s->dphi[ep.lev[m]][a];
    }
    b.inv_fac1 = 1.0 / fac1;
    b.d = s->d;
    hipLaunchKernelGGL(bc_kernel, dim3(nblk(3 * s->nbc)), dim3(256), 0, sp, b);
  }
  NSCHK(hipGetLastError());
  // nonlinear terms of the levels the extrapolation reads, the older ones
  // first: for pt = 1 without LES the newest one's kernel also forms the
  // velocity right-hand sides (into T[0..2]), which read the older ones
  const bool fuse_rhs = s->fused_rhs && pt == 1 && sub == 1 && !s->les_type && !s->sgs &&
                        ev.c[0] != 0.0 && ev.lev[0] == L0 && !s->nl_ok[ev.lev[0]];
  for (int m = 2; m >= 1; --m)
    if (ev.c[m] != 0.0) RC(nonlinear(s, ev.lev[m], st));
  if (ev.c[0] != 0.0) {
    NlRhs h;
    std::memset(&h, 0, sizeof(h));
    if (fuse_rhs) {
      for (int c = 0; c < 3; ++c) {
        h.out[c] = T[c];
        h.n1[c] = s->nl[c][ev.lev[1]];
        h.n2[c] = s->nl[c][ev.lev[2]];
        h.q1[c] = s->q[c][L1];
        h.q2[c] = s->q[c][L2];
        h.force[c] = s->has_force[c] ? s->force[c] : nullptr;
        h.gflow[c] = (c == 0 && s->flow) ? s->fstate : nullptr;
      }
      h.c0 = ev.c[0]; h.c1 = ev.c[1]; h.c2 = ev.c[2];
      h.f2 = s->fac[1]; h.f3 = s->fac[2]; h.f4 = s->tou == 3 ? s->fac[3] : 0.0;
      h.re = s->re;
      h.d = s->d;
    }
    RC(nonlinear(s, ev.lev[0], st, fuse_rhs ? &h : nullptr));
  }
  if (sp != st) {
    NSCHK(hipEventRecord(s->ev_join[1], sp));
    NSCHK(hipStreamWaitEvent(st, s->ev_join[1], 0));
  }
  // eddy viscosity and -div(tau) from the latest level (navier3d.f90:68); the
  // level does not change between sub-iterations, so neither does the force
  if (sub == 1 && (s->les_type || s->sgs)) RC(les_force(s, L0, st));
  s->grad_level = -1;  // T is scratch from here on
  // presolve: rhs2 = rhs1, rhs1 = f - extrapol_v(N), kept for pt >= 3
  if (pt >= 3) {
    s->rhs_cur ^= 1;
    for (int c = 0; c < 3; ++c) {
      RhsArgs r;
      std::memset(&r, 0, sizeof(r));
      r.n0 = s->nl[c][ev.lev[0]]; r.n1 = s->nl[c][ev.lev[1]]; r.n2 = s->nl[c][ev.lev[2]];
      r.c0 = ev.c[0]; r.c1 = ev.c[1]; r.c2 = ev.c[2];
      r.force = s->has_force[c] ? s->force[c] : nullptr;
      r.gflow = (c == 0 && s->flow) ? s->fstate : nullptr;
      r.les = (s->les_type || s->sgs) ? s->lesf[c] : nullptr;
      r.rhs_out = s->rhsb[s->rhs_cur][c];
      r.N = N;
      r.d = s->d;
      hipLaunchKernelGGL(rhs_kernel, dim3(nblk(N)), dim3(256), 0, st, r);
    }
    NSCHK(hipGetLastError());
  }
  const long sx = 2L * (s->ny - 2) * (s->nz - 2), sy = 2L * (s->nx - 2) * (s->nz - 2);
  // predicted velocity (pt <= 2), with d_c extrapol_p(p) for pt = 2
  if (pt <= 2) {
    double *dp[3] = {T[21], T[22], T[23]};
    if (pt == 2) {
      launch_extrap(s, s->p, ep, T[20], st);
      RC(fork3(s, st, [&](int a, hipStream_t q) { return der(s, a, T[20], dp[a], q); }));
    }
    RC(velocity_solves(s, ev, false, pt == 2 ? dp : nullptr, bc, precision, tol, st, fuse_rhs));
  }
  // pressure: rhs_p, fphi = div rhs_p, Neumann values, solve
  {
    RpArgs r;
    std::memset(&r, 0, sizeof(r));
    for (int c = 0; c < 3; ++c) {
      r.qi[c] = s->q[c][n1]; r.q0[c] = s->q[c][L0]; r.q1[c] = s->q[c][L1]; r.q2[c] = s->q[c][L2];
      if (pt >= 3) { r.r1[c] = s->rhsb[s->rhs_cur][c]; r.r2[c] = s->rhsb[s->rhs_cur ^ 1][c]; }
      r.out[c] = T[c];
    }
    r.f1 = s->fac[0]; r.f2 = s->fac[1]; r.f3 = s->fac[2]; r.f4 = s->fac[3];
    r.mode = pt <= 2 ? 0 : (pt == 3 ? 3 : (sub == 1 ? 4 : 5));
    r.N = N;
    r.d = s->d;
    hipLaunchKernelGGL(rp_kernel, dim3(nblk(N)), dim3(256), 0, st, r);
  }
  RC(fork3(s, st, [&](int a, hipStream_t q) { return der(s, a, T[a], T[3 + a], q); }));
  hipLaunchKernelGGL(sum3_kernel, dim3(nblk(N)), dim3(256), 0, st, (const double *)T[3],
                     (const double *)T[4], (const double *)T[5], T[6], N);
  hipLaunchKernelGGL(bcphi_kernel, dim3(nblk(s->nbc)), dim3(256), 0, st, (const double *)T[0],
                     (const double *)T[1], (const double *)T[2], bc[0], bc[1] + sx, bc[2] + sx + sy,
                     fac1, s->bcphi, s->d);
  NSCHK(hipGetLastError());
  if (pt == 3) {  // add_boundary_rotrot on extrapol_p of the velocity
    double *e[3] = {T[7], T[8], T[9]}, *de[3] = {T[10], T[11], T[12]};
    for (int c = 0; c < 3; ++c) launch_extrap(s, s->q[c], ep, e[c], st);
    for (int c = 0; c < 3; ++c) RC(der(s, c, e[c], de[c], st));  // d_c e_c
    for (int a = 0; a < 3; ++a) {
      const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
      hipLaunchKernelGGL(sum2_kernel, dim3(nblk(N)), dim3(256), 0, st, (const double *)de[b],
                         (const double *)de[c], T[13], N);
      RC(der(s, a, T[13], T[14], st));
      RC(dder(s, b, e[a], T[15], st));
      RC(dder(s, c, e[a], T[16], st));
      const long nf = a == 0 ? sx : (a == 1 ? sy : 2L * (s->nx - 2) * (s->ny - 2));
      hipLaunchKernelGGL(rotrot_kernel, dim3(nblk(nf)), dim3(256), 0, st, a, (const double *)T[14],
                         (const double *)T[15], (const double *)T[16], s->re, s->bcphi, s->d);
    }
    NSCHK(hipGetLastError());
  }
  RC(mf_solver3d_solve(s->scp, T[6], 0.0, s->phi[n1], s->bcphi, s->bcphi + sx, s->bcphi + sx + sy,
                       MF_FP64, 0.0, 0, nullptr, st));
  // projection
  const long ne = 4L * (s->nx + s->ny + s->nz);
  hipLaunchKernelGGL(zero_edges_kernel, dim3(nblk(ne)), dim3(256), 0, st, s->phi[n1], s->d);
  RC(fork3(s, st, [&](int a, hipStream_t q) { return der(s, a, s->phi[n1], s->dphi[n1][a], q); }));
  s->dphi_ok[n1] = true;
  {
    ProjArgs pj;
    std::memset(&pj, 0, sizeof(pj));
    pj.phi = s->phi[n1];
    if (pt == 2 || pt == 4) {
      launch_extrap(s, s->p, ep, T[7], st);
      pj.pe = T[7];
    }
    if (pt <= 2) {
      pj.dv = T[6];  // div(fac1 u*)
      pj.dv_fac1 = 1;
    } else if (pt == 4) {  // div(extrapol_p(q))
      for (int c = 0; c < 3; ++c) {
        launch_extrap(s, s->q[c], ep, T[8 + c], st);
        RC(der(s, c, T[8 + c], T[11 + c], st));
      }
      hipLaunchKernelGGL(sum3_kernel, dim3(nblk(N)), dim3(256), 0, st, (const double *)T[11],
                         (const double *)T[12], (const double *)T[13], T[14], N);
      pj.dv = T[14];
    }
    for (int c = 0; c < 3; ++c) { pj.r[c] = T[c]; pj.g[c] = s->dphi[n1][c]; pj.q[c] = s->q[c][n1]; }
    pj.fac1 = fac1;
    pj.re = s->re;
    pj.p = s->p[n1];
    pj.N = N;
    pj.d = s->d;
    hipLaunchKernelGGL(proj_kernel, dim3(nblk(N)), dim3(256), 0, st, pj);
    NSCHK(hipGetLastError());
  }
  s->nl_ok[n1] = false;
  // velocity correction (pt >= 3) with the gradient of the new pressure
  if (pt >= 3) {
    double *dp[3] = {T[21], T[22], T[23]};
    RC(fork3(s, st, [&](int a, hipStream_t q) { return der(s, a, s->p[n1], dp[a], q); }));
    RC(velocity_solves(s, ev, true, dp, bc, precision, tol, st));
  }
  *done = false;
  if (sub > 1) RC(testconv(s, done, st));
  if (s->nsubite > 1) {
    const size_t b = N * sizeof(double);
    for (int c = 0; c < 3; ++c)
      NSCHK(hipMemcpyAsync(s->subq[c], s->q[c][n1], b, hipMemcpyDeviceToDevice, st));
    NSCHK(hipMemcpyAsync(s->subp, s->p[n1], b, hipMemcpyDeviceToDevice, st));
  }
  return MF_OK;
}

int one_step(mf_ns *s, int precision, double tol, hipStream_t st) {
  s->time += s->dt;
  s->subite_done = 0;
  for (int sub = 1; sub <= s->nsubite; ++sub) {
    bool done = false;
    RC(substep(s, sub, precision, tol, &done, st));
    s->subite_done = sub;
    if (done) break;
  }
  // rotate the time levels
  const int first = s->it[0];
  for (int m = 0; m < NT - 1; ++m) s->it[m] = s->it[m + 1];
  s->it[NT - 1] = first;
  return MF_OK;
}
}  // namespace

extern "C" int mf_ns_destroy(mf_ns *s) {
  if (!s) return MF_OK;
  for (int a = 0; a < 3; ++a) mf_der_destroy(s->dc[a]);
  mf_solver3d_destroy(s->scu);
  mf_solver3d_destroy(s->scp);
  for (int k = 0; k < 2; ++k) {
    mf_solver3d_destroy(s->scu_vw[k]);
    if (s->side[k]) (void)hipStreamDestroy(s->side[k]);
    if (s->ev_join[k]) (void)hipEventDestroy(s->ev_join[k]);
  }
  if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
  if (s->mem) (void)hipFree(s->mem);
  if (s->lesmem) (void)hipFree(s->lesmem);
  if (s->sgsmem) (void)hipFree(s->sgsmem);
  if (s->schmem) (void)hipFree(s->schmem);
  delete s;
  return MF_OK;
}

extern "C" int mf_ns_create(const double *gx, int nx, const double *gy, int ny, const double *gz,
                            int nz, double re, double dt, int tou, int top, int nlt, int so,
                            mf_ns **out) {
  if (!gx || !gy || !gz || !out) return mf::fail(MF_ERR_ARG, "mf_ns_create: null pointer");
  if (tou < 1 || tou > 3 || top < 1 || top > 3 || (nlt != 1 && nlt != 2))
    return mf::fail(MF_ERR_ARG, "mf_ns_create: tou and top must be 1..3, nlt 1 or 2");
  if (!(re > 0.0) || !(dt > 0.0)) return mf::fail(MF_ERR_ARG, "mf_ns_create: re and dt must be positive");
  mf_ns *s = new mf_ns();
  s->nx = nx; s->ny = ny; s->nz = nz;
  s->gx.assign(gx, gx + nx);
  s->gy.assign(gy, gy + ny);
  s->gz.assign(gz, gz + nz);
  s->d = Dims{nx, ny, nz};
  s->N = (long)nx * ny * nz;
  s->nbc = 2L * ((long)(ny - 2) * (nz - 2) + (long)(nx - 2) * (nz - 2) + (long)(nx - 2) * (ny - 2));
  s->re = re; s->dt = dt; s->tou = tou; s->top = top; s->nlt = nlt;
  // BDF factors (navier_initialization, class_navier_3D.f90:2145-2160)
  if (tou == 1) { s->fac[0] = 1.0 / dt; s->fac[1] = -1.0 / dt; }
  else if (tou == 2) { s->fac[0] = 1.5 / dt; s->fac[1] = -2.0 / dt; s->fac[2] = 0.5 / dt; }
  else {
    s->fac[0] = 11.0 / (6.0 * dt); s->fac[1] = -18.0 / (6.0 * dt);
    s->fac[2] = 9.0 / (6.0 * dt); s->fac[3] = -2.0 / (6.0 * dt);
  }
  const int bu[6] = {1, 1, 1, 1, 1, 1}, bp[6] = {2, 2, 2, 2, 2, 2};
  int rc = mf_der_create(gx, nx, 1, so, &s->dc[0]);
  if (rc == MF_OK) rc = mf_der_create(gy, ny, 1, so, &s->dc[1]);
  if (rc == MF_OK) rc = mf_der_create(gz, nz, 1, so, &s->dc[2]);
  if (rc == MF_OK) rc = mf_solver3d_create(gx, nx, gy, ny, gz, nz, bu, so, &s->scu);
  if (rc == MF_OK) rc = mf_solver3d_create(gx, nx, gy, ny, gz, nz, bp, so, &s->scp);
  bool conc = true;
  if (const char *env = std::getenv("MF_NS_CONCURRENT")) conc = std::atoi(env) != 0;
  if (const char *env = std::getenv("MF_NS_FUSED_RHS")) s->fused_rhs = std::atoi(env) != 0;
  for (int k = 0; conc && k < 2 && rc == MF_OK; ++k)
    rc = mf_solver3d_create(gx, nx, gy, ny, gz, nz, bu, so, &s->scu_vw[k]);
  if (conc && rc == MF_OK) {
    // default priority, non-blocking: no implicit ordering with the null stream
    hipError_t he = hipSuccess;
    for (int k = 0; k < 2 && he == hipSuccess; ++k) {
      he = hipStreamCreateWithFlags(&s->side[k], hipStreamNonBlocking);
      if (he == hipSuccess) he = hipEventCreateWithFlags(&s->ev_join[k], hipEventDisableTiming);
    }
    if (he == hipSuccess) he = hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming);
    if (he != hipSuccess) {
      mf_ns_destroy(s);
      return mf::fail(MF_ERR_DEVICE, std::string("mf_ns_create: streams: ") + hipGetErrorString(he));
    }
  }
  if (rc != MF_OK) {
    const std::string msg = mf_last_error();
    mf_ns_destroy(s);
    return mf::fail(rc, "mf_ns_create: " + msg);
  }
  // one allocation: fields, then bc arrays, weights, scalars
  const long N = s->N;
  const long Na = (N + 31) / 32 * 32, Ba = (s->nbc + 31) / 32 * 32;
  const int nfields = 3 * NT + NT + NT + 3 * NT + 3 * NT + 3 + 24 + 1;
  s->nparts = 1024;
  const size_t total = (size_t)nfields * Na + 7 * Ba + 64 + s->nparts;
  hipError_t e = hipMalloc(&s->mem, total * sizeof(double));
  if (e != hipSuccess) {
    mf_ns_destroy(s);
    return mf::fail(MF_ERR_DEVICE, std::string("mf_ns_create: hipMalloc: ") + hipGetErrorString(e));
  }
  double *m = (double *)s->mem;
  auto take = [&](long cnt) { double *r = m; m += cnt; return r; };
  for (int c = 0; c < 3; ++c) for (int L = 0; L < NT; ++L) s->q[c][L] = take(Na);
  for (int L = 0; L < NT; ++L) s->p[L] = take(Na);
  for (int L = 0; L < NT; ++L) s->phi[L] = take(Na);
  for (int c = 0; c < 3; ++c) for (int L = 0; L < NT; ++L) s->nl[c][L] = take(Na);
  for (int L = 0; L < NT; ++L) for (int a = 0; a < 3; ++a) s->dphi[L][a] = take(Na);
  for (int c = 0; c < 3; ++c) s->force[c] = take(Na);
  for (int k = 0; k < 24; ++k) s->T[k] = take(Na);
  s->W = take(Na);
  for (int c = 0; c < 3; ++c) s->bc0[c] = take(Ba);
  for (int c = 0; c < 3; ++c) s->bcw[c] = take(Ba);
  s->bcphi = take(Ba);
  s->fstate = take(64);
  s->part = take(s->nparts);
  e = hipMemset(s->mem, 0, total * sizeof(double));
  double *dg = nullptr;
  if (e == hipSuccess) e = hipMalloc(&dg, (size_t)(nx + ny + nz) * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(dg, gx, nx * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dg + nx, gy, ny * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dg + nx + ny, gz, nz * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(weights_kernel, dim3(nblk(N)), dim3(256), 0, 0, (const double *)dg,
                       (const double *)(dg + nx), (const double *)(dg + nx + ny), s->W, N, s->d);
    e = hipDeviceSynchronize();
  }
  if (dg) (void)hipFree(dg);
  if (e != hipSuccess) {
    mf_ns_destroy(s);
    return mf::fail(MF_ERR_DEVICE, std::string("mf_ns_create: set-up: ") + hipGetErrorString(e));
  }
  // volume of the weights (host, same rule)
  auto w1 = [](const double *g, int n, int a) {
    if (a == 0) return (g[1] - g[0]) * 0.5;
    if (a == n - 1) return (g[n - 1] - g[n - 2]) * 0.5;
    return (g[a + 1] - g[a - 1]) * 0.5;
  };
  double vol = 0.0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int nb = (i == 0 || i == nx - 1) + (j == 0 || j == ny - 1) + (k == 0 || k == nz - 1);
        if (nb < 2) vol += w1(gx, nx, i) * w1(gy, ny, j) * w1(gz, nz, k);
      }
  s->vol = vol;
  for (int L = 0; L < NT; ++L) s->nl_ok[L] = s->dphi_ok[L] = false;
  *out = s;
  return MF_OK;
}

extern "C" int mf_ns_set_bc(mf_ns *s, int c, const double *bcx, const double *bcy,
                            const double *bcz, void *stream) {
  if (!s || c < 0 || c > 2 || !bcx || !bcy || !bcz) return mf::fail(MF_ERR_ARG, "mf_ns_set_bc: invalid argument");
  hipStream_t st = (hipStream_t)stream;
  const long sx = 2L * (s->ny - 2) * (s->nz - 2), sy = 2L * (s->nx - 2) * (s->nz - 2),
             sz = 2L * (s->nx - 2) * (s->ny - 2);
  NSCHK(hipMemcpyAsync(s->bc0[c], bcx, sx * sizeof(double), hipMemcpyDeviceToDevice, st));
  NSCHK(hipMemcpyAsync(s->bc0[c] + sx, bcy, sy * sizeof(double), hipMemcpyDeviceToDevice, st));
  NSCHK(hipMemcpyAsync(s->bc0[c] + sx + sy, bcz, sz * sizeof(double), hipMemcpyDeviceToDevice, st));
  return MF_OK;
}

extern "C" int mf_ns_set_forcing(mf_ns *s, int c, const double *f, void *stream) {
  if (!s || c < 0 || c > 2) return mf::fail(MF_ERR_ARG, "mf_ns_set_forcing: invalid argument");
  s->has_force[c] = f != nullptr;
  if (f)
    NSCHK(hipMemcpyAsync(s->force[c], f, s->N * sizeof(double), hipMemcpyDeviceToDevice,
                         (hipStream_t)stream));
  return MF_OK;
}

extern "C" int mf_ns_set_flow_rate(mf_ns *s, int enable, double g0, void *stream) {
  if (!s) return mf::fail(MF_ERR_ARG, "mf_ns_set_flow_rate: null handle");
  s->flow = enable != 0;
  const double st4[4] = {0.0, 0.0, 0.0, g0 > 0.0 ? g0 : 0.0};
  NSCHK(hipMemcpyAsync(s->fstate, st4, sizeof(st4), hipMemcpyHostToDevice, (hipStream_t)stream));
  NSCHK(hipStreamSynchronize((hipStream_t)stream));
  return MF_OK;
}

namespace {
// force accumulators, nu and the squared spacings (first use of LES or SGS)
int les_alloc(mf_ns *s) {
  if (!s->lesmem) {
    const long Na = (s->N + 31) / 32 * 32;
    const int ng = s->nx + s->ny + s->nz;
    NSCHK(hipMalloc(&s->lesmem, (4 * Na + ng) * sizeof(double)));
    double *m = (double *)s->lesmem;
    for (int c3 = 0; c3 < 3; ++c3) s->lesf[c3] = m + c3 * Na;
    s->lesnu = m + 3 * Na;
    s->d2 = m + 4 * Na;
    // squared spacings of the grids (navier_LES :1416-1446)
    std::vector<double> h2(ng);
    mf::les_spacing2(s->gx.data(), s->nx, h2.data());
    mf::les_spacing2(s->gy.data(), s->ny, h2.data() + s->nx);
    mf::les_spacing2(s->gz.data(), s->nz, h2.data() + s->nx + s->ny);
    NSCHK(hipMemcpy(s->d2, h2.data(), ng * sizeof(double), hipMemcpyHostToDevice));
    NSCHK(hipMemset(s->lesnu, 0, Na * sizeof(double)));
  }
  return MF_OK;
}
}  // namespace

extern "C" int mf_ns_set_les(mf_ns *s, int les_type, double c, void *stream) {
  if (!s || les_type < 0 || les_type > 3 || !(c >= 0.0))
    return mf::fail(MF_ERR_ARG, "mf_ns_set_les: les_type must be 0..3 and c >= 0");
  if (les_type) RC(les_alloc(s));
  s->les_type = les_type;
  s->les_c2 = c * c;  // navier_initialization: les_c = cmd%les_c**2 (:2138)
  (void)stream;
  return MF_OK;
}

extern "C" int mf_ns_set_scheme(mf_ns *s, int pt, int nsubite, void *stream) {
  if (!s || pt < 1 || pt > 4 || nsubite < 1)
    return mf::fail(MF_ERR_ARG, "mf_ns_set_scheme: pt must be 1..4 and nsubite >= 1");
  hipStream_t st = (hipStream_t)stream;
  NSCHK(hipStreamSynchronize(st));
  if (s->schmem) {
    NSCHK(hipFree(s->schmem));
    s->schmem = nullptr;
  }
  for (int k = 0; k < 2; ++k)
    for (int c = 0; c < 3; ++c) s->rhsb[k][c] = nullptr;
  for (int c = 0; c < 3; ++c) s->subq[c] = nullptr;
  s->subp = s->cpart = s->csum = nullptr;
  const long Na = (s->N + 31) / 32 * 32;
  const long nrhs = pt >= 3 ? 6 : 0, nsub = nsubite > 1 ? 4 : 0;
  const size_t total = (size_t)(nrhs + nsub) * Na + (nsub ? 3L * s->nparts + 8 : 0);
  if (total) {
    NSCHK(hipMalloc(&s->schmem, total * sizeof(double)));
    NSCHK(hipMemsetAsync(s->schmem, 0, total * sizeof(double), st));
    double *m = (double *)s->schmem;
    for (int k = 0; k < 2 && nrhs; ++k)
      for (int c = 0; c < 3; ++c, m += Na) s->rhsb[k][c] = m;
    if (nsub) {
      for (int c = 0; c < 3; ++c, m += Na) s->subq[c] = m;
      s->subp = m; m += Na;
      s->cpart = m; m += 3L * s->nparts;
      s->csum = m;
    }
    NSCHK(hipStreamSynchronize(st));
  }
  s->pt = pt;
  s->nsubite = nsubite;
  s->rhs_cur = 0;
  s->subite_done = 0;
  s->conv[0] = s->conv[1] = 0.0;
  return MF_OK;
}

extern "C" int mf_ns_subiterations(const mf_ns *s, int *count, double *dvel, double *dp) {
  if (!s) return mf::fail(MF_ERR_ARG, "mf_ns_subiterations: null handle");
  if (count) *count = s->subite_done;
  if (dvel) *dvel = s->conv[0];
  if (dp) *dp = s->conv[1];
  return MF_OK;
}

extern "C" int mf_ns_set_sgs(mf_ns *s, const mf_mlp *mlp, double gscale, double lscale,
                             double tscale, void *stream) {
  if (!s) return mf::fail(MF_ERR_ARG, "mf_ns_set_sgs: null handle");
  if (mlp && !(tscale > 0.0)) return mf::fail(MF_ERR_ARG, "mf_ns_set_sgs: tscale must be positive");
  if (mlp) {
    RC(les_alloc(s));
    if (!s->sgsmem) {
      NSCHK(hipMalloc(&s->sgsmem, (size_t)s->N * 16 * sizeof(float)));
      s->sgsX = (float *)s->sgsmem;
      s->sgsY = s->sgsX + (size_t)s->N * 10;
    }
  }
  s->sgs = mlp;
  s->sgs_g = gscale;
  s->sgs_l = lscale;
  s->sgs_t = tscale;
  (void)stream;
  return MF_OK;
}

namespace {
// rhs1 = rhs2 = 0, as after navier_initialization (field_init zeroes them,
// class_navier_3D.f90:2240-2244; the restart files do not hold them)
int clear_presolve(mf_ns *s, hipStream_t st) {
  for (int k = 0; k < 2; ++k)
    for (int c = 0; c < 3; ++c)
      if (s->rhsb[k][c]) NSCHK(hipMemsetAsync(s->rhsb[k][c], 0, s->N * sizeof(double), st));
  s->rhs_cur = 0;
  return MF_OK;
}
}  // namespace

extern "C" int mf_ns_set_sgs_coupling(mf_ns *s, int coupling) {
  if (!s || coupling < 0 || coupling > 1)
    return mf::fail(MF_ERR_ARG, "mf_ns_set_sgs_coupling: coupling must be 0 or 1");
  s->sgs_coupling = coupling;
  return MF_OK;
}

// ---------------------------------------------------------------------------
// Restart files (restart_write / restart_read, class_navier_3D.f90:186-340)
// ---------------------------------------------------------------------------
namespace {
const char *const kDims[3] = {"resolution_x", "resolution_y", "resolution_z"};

// tn(nav) (:169-183): floor(time) and the first 8 decimals, 8 digits each
std::string time_name(double t) {
  char b[64];
  const double f = std::floor(t);
  std::snprintf(b, sizeof(b), "%08lld.%08lld", (long long)f, (long long)((t - f) * 1e8));
  return b;
}
}  // namespace

extern "C" int mf_ns_restart_write(const mf_ns *s, const char *dirname, char *dir_out,
                                   int dir_out_len, void *stream) {
  if (!s) return mf::fail(MF_ERR_ARG, "mf_ns_restart_write: null handle");
  const std::string dir = (dirname && *dirname) ? std::string(dirname) : "restart_" + time_name(s->time);
  if (::mkdir(dir.c_str(), 0755) != 0 && errno != EEXIST)
    return mf::fail(MF_ERR_ARG, "mf_ns_restart_write: cannot create " + dir);
  NSCHK(hipStreamSynchronize((hipStream_t)stream));
  // param.dat: time (real 8) then it(1:nt) (integer 4), unformatted stream
  {
    FILE *fp = std::fopen((dir + "/param.dat").c_str(), "wb");
    if (!fp) return mf::fail(MF_ERR_ARG, "mf_ns_restart_write: cannot write " + dir + "/param.dat");
    int32_t it1[NT];
    for (int m = 0; m < NT; ++m) it1[m] = s->it[m] + 1;
    bool ok = std::fwrite(&s->time, sizeof(double), 1, fp) == 1 &&
              std::fwrite(it1, sizeof(int32_t), NT, fp) == (size_t)NT;
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) return mf::fail(MF_ERR_ARG, "mf_ns_restart_write: write error on param.dat");
  }
  std::vector<double> h(s->N), zero(s->N, 0.0);
  std::string err;
  auto put = [&](const std::string &file, const char *name, const double *src, bool dbl) -> int {
    const double *host = zero.data();
    if (src) {
      hipError_t e = hipMemcpy(h.data(), src, s->N * sizeof(double), hipMemcpyDeviceToHost);
      if (e != hipSuccess) return mf::fail(MF_ERR_DEVICE, std::string("mf_ns_restart_write: ") + hipGetErrorString(e));
      host = h.data();
    }
    const int rc = mf::nc_write_field((dir + "/" + file + ".nc").c_str(), name, kDims, s->nx, s->ny,
                                      s->nz, host, dbl, err);
    return rc == MF_OK ? MF_OK : mf::fail(rc, "mf_ns_restart_write: " + err);
  };
  // every time level in the order of it(:), in FP64; the right-hand sides
  // fu, fv, fw, fp, fphi are rebuilt from zero before each use in the
  // reference (:1615, :1810), so their files are written as zero fields
  for (int i = 1; i <= NT; ++i) {
    const int L = s->it[i - 1];
    const std::string n = std::to_string(i);
    RC(put("vel_u_" + n, "U", s->q[0][L], true));
    RC(put("vel_v_" + n, "V", s->q[1][L], true));
    RC(put("vel_w_" + n, "W", s->q[2][L], true));
    RC(put("vel_p_" + n, "P", s->p[L], true));
    RC(put("vel_phi_" + n, "P", s->phi[L], true));
    RC(put("rhs_u_" + n, "RHS_U", nullptr, true));
    RC(put("rhs_v_" + n, "RHS_V", nullptr, true));
    RC(put("rhs_w_" + n, "RHS_W", nullptr, true));
    RC(put("rhs_p_" + n, "RHS_P", nullptr, true));
  }
  RC(put("rhs_phi_" + std::to_string(NT), "RHS_PHI", nullptr, true));
  // the latest level in single precision (the reference writes w's data to
  // vel_p, :256; here vel_p holds p)
  const int Ln = s->it[NT - 1];
  RC(put("vel_u", "U", s->q[0][Ln], false));
  RC(put("vel_v", "V", s->q[1][Ln], false));
  RC(put("vel_w", "W", s->q[2][Ln], false));
  RC(put("vel_p", "P", s->p[Ln], false));
  if (dir_out && dir_out_len > 0) {
    std::strncpy(dir_out, dir.c_str(), dir_out_len - 1);
    dir_out[dir_out_len - 1] = 0;
  }
  return MF_OK;
}

extern "C" int mf_ns_restart_read(mf_ns *s, const char *dirname, void *stream) {
  if (!s) return mf::fail(MF_ERR_ARG, "mf_ns_restart_read: null handle");
  const std::string dir = (dirname && *dirname) ? std::string(dirname) : "restart";
  hipStream_t st = (hipStream_t)stream;
  double time = 0.0;
  int32_t it1[NT];
  {
    FILE *fp = std::fopen((dir + "/param.dat").c_str(), "rb");
    if (!fp) return mf::fail(MF_ERR_ARG, "mf_ns_restart_read: no " + dir + "/param.dat");
    const bool ok = std::fread(&time, sizeof(double), 1, fp) == 1 &&
                    std::fread(it1, sizeof(int32_t), NT, fp) == (size_t)NT;
    std::fclose(fp);
    if (!ok) return mf::fail(MF_ERR_ARG, "mf_ns_restart_read: short param.dat");
  }
  int seen = 0;
  for (int m = 0; m < NT; ++m) {
    if (it1[m] < 1 || it1[m] > NT || (seen & (1 << it1[m])))
      return mf::fail(MF_ERR_ARG, "mf_ns_restart_read: param.dat levels are not a permutation of 1..4");
    seen |= 1 << it1[m];
  }
  std::vector<double> h(s->N);
  std::string err;
  auto get = [&](const std::string &file, const char *name, double *dst) -> int {
    int dims[3] = {s->nx, s->ny, s->nz};
    const int rc = mf::nc_read_field((dir + "/" + file + ".nc").c_str(), name, dims, nullptr,
                                     nullptr, h.data(), err);
    if (rc != MF_OK) return mf::fail(rc, "mf_ns_restart_read: " + err);
    hipError_t e = hipMemcpy(dst, h.data(), s->N * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) return mf::fail(MF_ERR_DEVICE, std::string("mf_ns_restart_read: ") + hipGetErrorString(e));
    return MF_OK;
  };
  NSCHK(hipStreamSynchronize(st));
  for (int i = 1; i <= NT; ++i) {
    const int L = it1[i - 1] - 1;
    const std::string n = std::to_string(i);
    RC(get("vel_u_" + n, "U", s->q[0][L]));
    RC(get("vel_v_" + n, "V", s->q[1][L]));
    RC(get("vel_w_" + n, "W", s->q[2][L]));
    RC(get("vel_p_" + n, "P", s->p[L]));
    RC(get("vel_phi_" + n, "P", s->phi[L]));
  }
  for (int m = 0; m < NT; ++m) s->it[m] = it1[m] - 1;
  s->time = time;
  // derived state: nonlinear terms recomputed when needed, the boundary
  // gradient's phi derivatives now; the flow-rate force re-initialises from
  // the restarted field at the next step (its state lives in save variables
  // of f(nav,var), :1136-1147, not in the restart files)
  for (int L = 0; L < NT; ++L) {
    s->nl_ok[L] = false;
    for (int a = 0; a < 3; ++a) RC(der(s, a, s->phi[L], s->dphi[L][a], st));
    s->dphi_ok[L] = true;
  }
  s->grad_level = -1;
  RC(clear_presolve(s, st));
  const double zero = 0.0;
  NSCHK(hipMemcpyAsync(s->fstate + 2, &zero, sizeof(double), hipMemcpyHostToDevice, st));
  NSCHK(hipStreamSynchronize(st));
  return MF_OK;
}

extern "C" int mf_ns_les_nu(const mf_ns *s, double *nu, void *stream) {
  if (!s || !nu) return mf::fail(MF_ERR_ARG, "mf_ns_les_nu: invalid argument");
  if (!s->lesmem) return mf::fail(MF_ERR_ARG, "mf_ns_les_nu: LES was never enabled");
  NSCHK(hipMemcpyAsync(nu, s->lesnu, s->N * sizeof(double), hipMemcpyDeviceToDevice,
                       (hipStream_t)stream));
  return MF_OK;
}

extern "C" int mf_ns_set_state(mf_ns *s, const double *u, const double *v, const double *w,
                               void *stream) {
  if (!s || !u || !v || !w) return mf::fail(MF_ERR_ARG, "mf_ns_set_state: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const double *src[3] = {u, v, w};
  const long ne = 4L * (s->nx + s->ny + s->nz);
  for (int c = 0; c < 3; ++c)
    for (int L = 0; L < NT; ++L) {
      NSCHK(hipMemcpyAsync(s->q[c][L], src[c], s->N * sizeof(double), hipMemcpyDeviceToDevice, st));
      hipLaunchKernelGGL(zero_edges_kernel, dim3(nblk(ne)), dim3(256), 0, st, s->q[c][L], s->d);
      s->nl_ok[L] = false;
    }
  for (int L = 0; L < NT; ++L) {
    NSCHK(hipMemsetAsync(s->p[L], 0, s->N * sizeof(double), st));
    NSCHK(hipMemsetAsync(s->phi[L], 0, s->N * sizeof(double), st));
    s->dphi_ok[L] = false;
  }
  for (int m = 0; m < NT; ++m) s->it[m] = m;
  s->time = 0.0;
  s->grad_level = -1;
  RC(clear_presolve(s, st));
  NSCHK(hipGetLastError());
  return MF_OK;
}

extern "C" int mf_ns_step(mf_ns *s, int nsteps, int precision, double tol, void *stream) {
  if (!s || nsteps < 0) return mf::fail(MF_ERR_ARG, "mf_ns_step: invalid argument");
  for (int k = 0; k < nsteps; ++k) RC(one_step(s, precision, tol, (hipStream_t)stream));
  return MF_OK;
}

extern "C" int mf_ns_get(const mf_ns *s, double *u, double *v, double *w, double *p, double *phi,
                         void *stream) {
  if (!s) return mf::fail(MF_ERR_ARG, "mf_ns_get: null handle");
  hipStream_t st = (hipStream_t)stream;
  const int L = s->it[NT - 1];
  const size_t b = s->N * sizeof(double);
  double *dst[5] = {u, v, w, p, phi};
  const double *src[5] = {s->q[0][L], s->q[1][L], s->q[2][L], s->p[L], s->phi[L]};
  for (int k = 0; k < 5; ++k)
    if (dst[k]) NSCHK(hipMemcpyAsync(dst[k], src[k], b, hipMemcpyDeviceToDevice, st));
  return MF_OK;
}

extern "C" int mf_ns_info(const mf_ns *s, double *time, double *flow_force, void *stream) {
  if (!s) return mf::fail(MF_ERR_ARG, "mf_ns_info: null handle");
  if (time) *time = s->time;
  if (flow_force) {
    NSCHK(hipStreamSynchronize((hipStream_t)stream));
    NSCHK(hipMemcpy(flow_force, s->fstate, sizeof(double), hipMemcpyDeviceToHost));
  }
  return MF_OK;
}
