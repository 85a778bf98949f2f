// Spline filter, LES eddy viscosities and the learned SGS closure on gfx950.
//
// * Filter (LSQUnivariateSpline, src/class_filter/class_filter.f90:130-174):
//   the least-squares spline of degree k with curfit's iopt = -1 knots fitted
//   to a line and evaluated at the output points is linear in the data; per
//   axis it is a matrix F (n_out x n_in), built on the host (B-spline bases by
//   de Boor's recurrence, normal equations by Cholesky in 80-bit arithmetic)
//   and applied to a 3-D field as three FP64 MFMA GEMMs (x, y, z), the shapes
//   of the solver's eigenvector transforms.
// * Eddy viscosities (navier_LES, src/class_navier/class_navier_3D.f90:
//   1225-1449; les_sigma :1453-1497): one pointwise pass over the nine
//   velocity-gradient fields.
// * Learned closure: an MLP 10 -> 64 -> 64 -> 6 (tanh) from the filtered
//   velocity gradients and the filter width to the six SGS stresses.  Weights
//   are bf16 (rounded from FP32 master weights after every optimiser step),
//   arithmetic FP32.  One block of 64 threads handles 64 samples: forward in
//   registers with the activations kept in LDS, backward into per-block
//   gradient partials that a second kernel sums in a fixed order
//   (deterministic), then Adam.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "gemm.h"
#include "internal.h"
#include "mflops/mf.h"

namespace mf {
namespace les {

constexpr int I0 = 10, H = 64, O = 6;
constexpr int NW1 = H * I0, NW2 = H * H, NW3 = O * H;
constexpr int OB1 = NW1, OW2 = OB1 + H, OB2 = OW2 + NW2, OW3 = OB2 + H, OB3 = OW3 + NW3;
constexpr int NPARAM = OB3 + O;  // 5254
constexpr int SPB = 64;          // samples per block

__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

__device__ __forceinline__ unsigned short f2bf(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

// bf16 copies of the three weight matrices (W1 | W2 | W3)
__global__ void __launch_bounds__(256) round_kernel(const float *__restrict__ p,
                                                    unsigned short *__restrict__ wb) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t < NW1) wb[t] = f2bf(p[t]);
  else if (t < NW1 + NW2) wb[t] = f2bf(p[OW2 + (t - NW1)]);
  else if (t < NW1 + NW2 + NW3) wb[t] = f2bf(p[OW3 + (t - NW1 - NW2)]);
}

struct Net {
  float w1[NW1], w2[NW2], w3[NW3], b1[H], b2[H], b3[O];
};

__device__ void load_net(Net &n, const unsigned short *__restrict__ wb, const float *__restrict__ p) {
  for (int t = threadIdx.x; t < NW1; t += blockDim.x) n.w1[t] = bf2f(wb[t]);
  for (int t = threadIdx.x; t < NW2; t += blockDim.x) n.w2[t] = bf2f(wb[NW1 + t]);
  for (int t = threadIdx.x; t < NW3; t += blockDim.x) n.w3[t] = bf2f(wb[NW1 + NW2 + t]);
  for (int t = threadIdx.x; t < H; t += blockDim.x) { n.b1[t] = p[OB1 + t]; n.b2[t] = p[OB2 + t]; }
  for (int t = threadIdx.x; t < O; t += blockDim.x) n.b3[t] = p[OB3 + t];
}

// Y[s][0..5] for samples s < n; X row-major (n x 10)
__global__ void __launch_bounds__(64) mlp_forward_kernel(const float *__restrict__ X,
                                                         float *__restrict__ Y, long n,
                                                         const unsigned short *__restrict__ wb,
                                                         const float *__restrict__ p) {
  __shared__ Net net;
  __shared__ float A[SPB][H + 1];  // h1, one row per sample
  load_net(net, wb, p);
  __syncthreads();
  const int tid = threadIdx.x;
  const long s = (long)blockIdx.x * SPB + tid;
  if (s >= n) return;
  float x[I0];
#pragma unroll
  for (int j = 0; j < I0; ++j) x[j] = X[s * I0 + j];
  for (int i = 0; i < H; ++i) {
    float a = net.b1[i];
#pragma unroll
    for (int j = 0; j < I0; ++j) a = fmaf(net.w1[i * I0 + j], x[j], a);
    A[tid][i] = tanhf(a);
  }
  float y[O];
#pragma unroll
  for (int o = 0; o < O; ++o) y[o] = net.b3[o];
  for (int i = 0; i < H; ++i) {
    float a = net.b2[i];
    for (int j = 0; j < H; ++j) a = fmaf(net.w2[i * H + j], A[tid][j], a);
    const float h2 = tanhf(a);
    #pragma unroll
    for (int o = 0; o < O; ++o) y[o] = fmaf(net.w3[o * H + i], h2, y[o]);
  }
  for (int o = 0; o < O; ++o) Y[s * O + o] = y[o];
}

// Gradient of sum_s sum_c (Y - T)^2 / n over the block's samples:
// part[blk * (NPARAM + 1) + q] (q = NPARAM: the block's loss term).
__global__ void __launch_bounds__(64) mlp_grad_kernel(const float *__restrict__ X,
                                                      const float *__restrict__ T, long n,
                                                      const unsigned short *__restrict__ wb,
                                                      const float *__restrict__ p,
                                                      float *__restrict__ part) {
  __shared__ Net net;
  __shared__ float A[SPB][H + 1], B[SPB][H + 1];  // h1 -> dh1, h2 -> dh2
  __shared__ float Xs[SPB][I0], dY[SPB][O];
  load_net(net, wb, p);
  const int tid = threadIdx.x;
  const long s = (long)blockIdx.x * SPB + tid;
  const bool on = s < n;
  for (int j = 0; j < I0; ++j) Xs[tid][j] = on ? X[s * I0 + j] : 0.0f;
  __syncthreads();
  // forward
  for (int i = 0; i < H; ++i) {
    float a = net.b1[i];
    for (int j = 0; j < I0; ++j) a = fmaf(net.w1[i * I0 + j], Xs[tid][j], a);
    A[tid][i] = tanhf(a);
  }
  float y[O];
  for (int o = 0; o < O; ++o) y[o] = net.b3[o];
  for (int i = 0; i < H; ++i) {
    float a = net.b2[i];
    for (int j = 0; j < H; ++j) a = fmaf(net.w2[i * H + j], A[tid][j], a);
    const float h2 = tanhf(a);
    B[tid][i] = h2;
    #pragma unroll
    for (int o = 0; o < O; ++o) y[o] = fmaf(net.w3[o * H + i], h2, y[o]);
  }
  const float inv_n = 1.0f / (float)n;
  float loss = 0.0f;
  for (int o = 0; o < O; ++o) {
    const float d = on ? y[o] - T[s * O + o] : 0.0f;
    loss = fmaf(d, d, loss);
    dY[tid][o] = 2.0f * d * inv_n;
  }
  __syncthreads();
  float *g = part + (long)blockIdx.x * (NPARAM + 1);
  // W3 and b3: thread tid owns column j = tid of the six rows
  for (int o = 0; o < O; ++o) {
    float a = 0.0f;
    for (int q = 0; q < SPB; ++q) a = fmaf(dY[q][o], B[q][tid], a);
    g[OW3 + o * H + tid] = a;
  }
  if (tid < O) {
    float a = 0.0f;
    for (int q = 0; q < SPB; ++q) a += dY[q][tid];
    g[OB3 + tid] = a;
  }
  __syncthreads();
  // dh2 = (W3^T dY) (1 - h2^2), in place of h2 (own row)
  for (int i = 0; i < H; ++i) {
    float a = 0.0f;
    for (int o = 0; o < O; ++o) a = fmaf(net.w3[o * H + i], dY[tid][o], a);
    const float h2 = B[tid][i];
    B[tid][i] = a * (1.0f - h2 * h2);
  }
  __syncthreads();
  // W2 (row i = tid) and b2
  {
    float acc[H];
    for (int j = 0; j < H; ++j) acc[j] = 0.0f;
    float gb = 0.0f;
    for (int q = 0; q < SPB; ++q) {
      const float d = B[q][tid];
      gb += d;
      for (int j = 0; j < H; ++j) acc[j] = fmaf(d, A[q][j], acc[j]);
    }
    for (int j = 0; j < H; ++j) g[OW2 + tid * H + j] = acc[j];
    g[OB2 + tid] = gb;
  }
  __syncthreads();
  // dh1 = (W2^T dh2) (1 - h1^2), in place of h1 (own row)
  for (int j = 0; j < H; ++j) {
    float a = 0.0f;
    for (int i = 0; i < H; ++i) a = fmaf(net.w2[i * H + j], B[tid][i], a);
    const float h1 = A[tid][j];
    A[tid][j] = a * (1.0f - h1 * h1);
  }
  __syncthreads();
  // W1 (row i = tid) and b1
  {
    float acc[I0];
    for (int j = 0; j < I0; ++j) acc[j] = 0.0f;
    float gb = 0.0f;
    for (int q = 0; q < SPB; ++q) {
      const float d = A[q][tid];
      gb += d;
      for (int j = 0; j < I0; ++j) acc[j] = fmaf(d, Xs[q][j], acc[j]);
    }
    for (int j = 0; j < I0; ++j) g[tid * I0 + j] = acc[j];
    g[OB1 + tid] = gb;
  }
  // the block's loss (fixed-order tree over its samples)
  __shared__ float ls[SPB];
  ls[tid] = loss * inv_n;
  __syncthreads();
  for (int o = SPB / 2; o > 0; o >>= 1) {
    if (tid < o) ls[tid] += ls[tid + o];
    __syncthreads();
  }
  if (tid == 0) g[NPARAM] = ls[0];
}

// grad[q] = sum_b part[b (NPARAM + 1) + q], b ascending (FP64 sum)
__global__ void __launch_bounds__(256) mlp_reduce_kernel(const float *__restrict__ part, int nb,
                                                         float *__restrict__ grad) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q > NPARAM) return;
  double a = 0.0;
  for (int b = 0; b < nb; ++b) a += part[(long)b * (NPARAM + 1) + q];
  grad[q] = (float)a;
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) with bias correction at step t
__global__ void __launch_bounds__(256) adam_kernel(float *__restrict__ p, float *__restrict__ m,
                                                   float *__restrict__ v,
                                                   const float *__restrict__ g, float lr, float c1,
                                                   float c2) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= NPARAM) return;
  const float gq = g[q];
  const float mq = 0.9f * m[q] + 0.1f * gq;
  const float vq = 0.999f * v[q] + 0.001f * gq * gq;
  m[q] = mq;
  v[q] = vq;
  p[q] -= lr * (mq * c1) / (sqrtf(vq * c2) + 1e-8f);
}

// ---- eddy viscosity (one thread per point) ----
struct NuArgs {
  const double *du[9];     // du[3 i + j] = d u_i / d x_j
  const double *d2x, *d2y, *d2z;  // squared spacings
  int nx, ny, nz, type;
  double c;
  double *nu;
};

__device__ void sigmas(const double g2[3][3], double &s1, double &s2, double &s3) {
  double g[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) g[i][j] = g2[i][0] * g2[j][0] + g2[i][1] * g2[j][1] + g2[i][2] * g2[j][2];
  const double i1 = g[0][0] + g[1][1] + g[2][2];
  const double i2 = g[0][0] * g[1][1] + g[0][0] * g[2][2] + g[1][1] * g[2][2] - g[0][1] * g[1][0] -
                    g[0][2] * g[2][0] - g[1][2] * g[2][1];
  const double i3 = g[0][0] * g[1][1] * g[2][2] + g[1][0] * g[2][1] * g[0][2] +
                    g[2][0] * g[0][1] * g[1][2] - g[0][2] * g[1][1] * g[2][0] -
                    g[1][2] * g[2][1] * g[0][0] - g[2][2] * g[0][1] * g[1][0];
  const double a1 = i1 * i1 / 9.0 - i2 / 3.0;
  const double a2 = i1 * i1 * i1 / 27.0 - i1 * i2 / 6.0 + i3 / 2.0;
  double a3 = 0.0;
  if (!(a1 < 1e-10)) a3 = acos(fmin(1.0, fmax(-1.0, a2 / pow(a1, 1.5)))) / 3.0;
  const double sa = sqrt(fmax(a1, 0.0));
  s1 = i1 / 3.0 + 2.0 * sa * cos(a3);
  s2 = i1 / 3.0 - 2.0 * sa * cos(a3 + M_PI / 3.0);
  s3 = i1 - s1 - s2;
  s1 = sqrt(fmax(s1, 0.0));
  s2 = sqrt(fmax(s2, 0.0));
  s3 = sqrt(fmax(s3, 0.0));
}

__global__ void __launch_bounds__(256) les_nu_kernel(NuArgs a) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  const long N = (long)a.nx * a.ny * a.nz;
  if (t >= N) return;
  const int i = (int)(t % a.nx), j = (int)((t / a.nx) % a.ny), k = (int)(t / ((long)a.nx * a.ny));
  const int nb = (i == 0 || i == a.nx - 1) + (j == 0 || j == a.ny - 1) + (k == 0 || k == a.nz - 1);
  if (nb >= 2) {  // box edges (field_zero_edges)
    a.nu[t] = 0.0;
    return;
  }
  double d[3][3];
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q) d[p][q] = a.du[3 * p + q][t];
  double nu = 0.0;
  if (a.type == 1 || a.type == 2) {
    double ss = 0.0;
    for (int p = 0; p < 3; ++p) {
      ss += d[p][p] * d[p][p];
      for (int q = p + 1; q < 3; ++q) ss += 0.5 * (d[p][q] + d[q][p]) * (d[p][q] + d[q][p]);
    }
    if (a.type == 1) {
      nu = a.c * sqrt(2.0 * ss);
    } else {
      double tr = 0.0;
      for (int m = 0; m < 3; ++m)
        for (int q = 0; q < 3; ++q) tr += d[m][q] * d[q][m];
      double sd = 0.0;
      for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) {
          double v = 0.0;
          for (int m = 0; m < 3; ++m) v += 0.5 * (d[p][m] * d[m][q] + d[q][m] * d[m][p]);
          if (p == q) v -= tr / 3.0;
          sd += v * v;
        }
      nu = (sd < 1e-10) ? 0.0 : a.c * pow(sd, 1.5) / (pow(ss, 2.5) + pow(sd, 1.25));
    }
  } else if (a.type == 3) {
    double s1, s2, s3;
    sigmas(d, s1, s2, s3);
    nu = (s1 < 1e-10) ? 0.0 : a.c * s3 * (s1 - s2) * (s2 - s3) / (s1 * s1);
  }
  a.nu[t] = nu * (a.d2x[i] + a.d2y[j] + a.d2z[k]) / 3.0;
}

inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

// ---- learned-closure data path ----
__global__ void __launch_bounds__(256) prod_kernel(const double *__restrict__ a,
                                                   const double *__restrict__ b,
                                                   double *__restrict__ o, long N) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t < N) o[t] = a[t] * b[t];
}

// T[t][col] = (F(u_i u_j) - F(u_i) F(u_j)) / tscale
__global__ void __launch_bounds__(256) tau_kernel(const double *__restrict__ uu,
                                                  const double *__restrict__ ui,
                                                  const double *__restrict__ uj,
                                                  float *__restrict__ T, int col, double inv, long N) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t < N) T[t * O + col] = (float)((uu[t] - ui[t] * uj[t]) * inv);
}

struct FeatArgs {
  const double *du[9];  // du[3 i + j] = d u_i / d x_j
  const double *d2x, *d2y, *d2z;
  int nx, ny, nz;
  double gscale, lscale;
  float *X;
};

// X[t] = (gscale du_ij (row-major ij), lscale sqrt((dx^2+dy^2+dz^2)/3));
// gradients zero on the box edges (field_zero_edges)
__global__ void __launch_bounds__(256) features_kernel(FeatArgs a) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  const long N = (long)a.nx * a.ny * a.nz;
  if (t >= N) return;
  const int i = (int)(t % a.nx), j = (int)((t / a.nx) % a.ny), k = (int)(t / ((long)a.nx * a.ny));
  const int nb = (i == 0 || i == a.nx - 1) + (j == 0 || j == a.ny - 1) + (k == 0 || k == a.nz - 1);
  float *x = a.X + t * I0;
#pragma unroll
  for (int q = 0; q < 9; ++q) x[q] = nb >= 2 ? 0.0f : (float)(a.gscale * a.du[q][t]);
  x[9] = (float)(a.lscale * sqrt((a.d2x[i] + a.d2y[j] + a.d2z[k]) / 3.0));
}

// column of tau_ij in the six outputs (xx, yy, zz, xy, xz, yz)
__device__ __forceinline__ int pair_col(int i, int j) {
  return i == j ? i : (i + j == 1 ? 3 : (i + j == 2 ? 4 : 5));
}

// o_i = -tscale Y[t][pair(i, j)] for one direction j, zero on the box edges
struct SgsNuArgs {
  const double *du[9];
  double tscale;
  double *nu;
  int nx, ny, nz;
};

__global__ void __launch_bounds__(256) sgs_aux_kernel(const float *__restrict__ Y, int jd,
                                                      double tscale, double *__restrict__ o0,
                                                      double *__restrict__ o1,
                                                      double *__restrict__ o2, int nx, int ny,
                                                      int nz) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  const long N = (long)nx * ny * nz;
  if (t >= N) return;
  const int i = (int)(t % nx), j = (int)((t / nx) % ny), k = (int)(t / ((long)nx * ny));
  const int nb = (i == 0 || i == nx - 1) + (j == 0 || j == ny - 1) + (k == 0 || k == nz - 1);
  const float *y = Y + t * O;
  const double s = nb >= 2 ? 0.0 : -tscale;
  o0[t] = s * (double)y[pair_col(0, jd)];
  o1[t] = s * (double)y[pair_col(1, jd)];
  o2[t] = s * (double)y[pair_col(2, jd)];
}

// Eddy viscosity of the learned stress projected on the strain rate:
// nu = max(0, -(tau : S)) / (2 S : S), tau = tscale Y (xx yy zz xy xz yz),
// S_ij = (du_ij + du_ji)/2; 0 on the box edges and where S : S = 0.
__global__ void __launch_bounds__(256) sgs_nu_kernel(const float *__restrict__ Y, SgsNuArgs a) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  const long N = (long)a.nx * a.ny * a.nz;
  if (t >= N) return;
  const int i = (int)(t % a.nx), j = (int)((t / a.nx) % a.ny), k = (int)(t / ((long)a.nx * a.ny));
  const int nb = (i == 0 || i == a.nx - 1) + (j == 0 || j == a.ny - 1) + (k == 0 || k == a.nz - 1);
  double g[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) g[q] = a.du[q][t];
  const double sxx = g[0], syy = g[4], szz = g[8];
  const double sxy = 0.5 * (g[1] + g[3]), sxz = 0.5 * (g[2] + g[6]), syz = 0.5 * (g[5] + g[7]);
  const float *y = Y + t * O;
  const double ts = a.tscale;
  const double tS = ts * ((double)y[0] * sxx + (double)y[1] * syy + (double)y[2] * szz +
                          2.0 * ((double)y[3] * sxy + (double)y[4] * sxz + (double)y[5] * syz));
  const double SS = sxx * sxx + syy * syy + szz * szz + 2.0 * (sxy * sxy + sxz * sxz + syz * syz);
  a.nu[t] = (nb >= 2 || !(SS > 0.0)) ? 0.0 : fmax(0.0, -tS) / (2.0 * SS);
}

// Eddy-viscosity stress targets of feature rows: tau = -2 c^2 Delta^2 |S| S
// (xx yy zz xy xz yz), S from the nine gradients (row-major ij), Delta the
// tenth column, |S| = sqrt(2 S:S): the Smagorinsky stress of constant c
__global__ void __launch_bounds__(256) eddy_targets_kernel(const float *__restrict__ X, long n,
                                                           double c2, float *__restrict__ T) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const float *x = X + t * I0;
  const double sxx = x[0], syy = x[4], szz = x[8];
  const double sxy = 0.5 * ((double)x[1] + (double)x[3]), sxz = 0.5 * ((double)x[2] + (double)x[6]),
               syz = 0.5 * ((double)x[5] + (double)x[7]);
  const double SS = sxx * sxx + syy * syy + szz * szz + 2.0 * (sxy * sxy + sxz * sxz + syz * syz);
  const double d = x[9];
  const double f = -2.0 * c2 * d * d * sqrt(2.0 * SS);
  float *o = T + t * O;
  o[0] = (float)(f * sxx); o[1] = (float)(f * syy); o[2] = (float)(f * szz);
  o[3] = (float)(f * sxy); o[4] = (float)(f * sxz); o[5] = (float)(f * syz);
}

}  // namespace les
}  // namespace mf

extern "C" int mf_sgs_eddy_targets(const float *X, long long n, double c, float *T, void *stream) {
  if (!X || !T || n < 0 || !(c >= 0.0)) return mf::fail(MF_ERR_ARG, "mf_sgs_eddy_targets: invalid argument");
  if (n == 0) return MF_OK;
  hipLaunchKernelGGL(mf::les::eddy_targets_kernel, dim3(mf::les::nblk(n)), dim3(256), 0,
                     (hipStream_t)stream, X, (long)n, c * c, T);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return mf::fail(MF_ERR_DEVICE, std::string("mf_sgs_eddy_targets: ") + hipGetErrorString(e));
  return MF_OK;
}

void mf::sgs_nu_launch(const float *Y, const double *const du[9], double tscale, double *nu, int nx,
                       int ny, int nz, void *stream) {
  mf::les::SgsNuArgs a;
  for (int q = 0; q < 9; ++q) a.du[q] = du[q];
  a.tscale = tscale;
  a.nu = nu;
  a.nx = nx; a.ny = ny; a.nz = nz;
  hipLaunchKernelGGL(mf::les::sgs_nu_kernel, dim3(mf::les::nblk((long)nx * ny * nz)), dim3(256), 0,
                     (hipStream_t)stream, Y, a);
}

void mf::sgs_features_launch(const double *const du[9], const double *d2x, const double *d2y,
                             const double *d2z, int nx, int ny, int nz, double gscale,
                             double lscale, float *X, void *stream) {
  mf::les::FeatArgs a;
  for (int q = 0; q < 9; ++q) a.du[q] = du[q];
  a.d2x = d2x; a.d2y = d2y; a.d2z = d2z;
  a.nx = nx; a.ny = ny; a.nz = nz; a.gscale = gscale; a.lscale = lscale; a.X = X;
  hipLaunchKernelGGL(mf::les::features_kernel, dim3(mf::les::nblk((long)nx * ny * nz)), dim3(256),
                     0, (hipStream_t)stream, a);
}

void mf::sgs_aux_launch(const float *Y, int j, double tscale, double *o0, double *o1, double *o2,
                        int nx, int ny, int nz, void *stream) {
  hipLaunchKernelGGL(mf::les::sgs_aux_kernel, dim3(mf::les::nblk((long)nx * ny * nz)), dim3(256),
                     0, (hipStream_t)stream, Y, j, tscale, o0, o1, o2, nx, ny, nz);
}

// ---------------------------------------------------------------------------
// Filter handle
// ---------------------------------------------------------------------------
struct mf_filter {
  int nin = 0, nout = 0;
  std::vector<double> F;   // nout x nin, column-major
  double *dF = nullptr;    // F (nout x nin) column-major
  double *dFt = nullptr;   // F^T (nin x nout) column-major
};

namespace {
#define LESCHK(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return mf::fail(MF_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// B-splines of degree k on knots t at x (deriv 0), or their first
// derivatives (deriv 1): the k+1 nonzero values in N and the index of the
// first.  de Boor's recurrence to degree k - deriv; the right end point
// belongs to the last nonempty span.
int bspline(const std::vector<long double> &t, int k, long double x, int deriv, long double *N) {
  const int nc = (int)t.size() - k - 1;
  int last = nc - 1;
  while (t[last] >= t[last + 1]) --last;
  int l = (int)(std::upper_bound(t.begin(), t.end(), x) - t.begin()) - 1;
  l = std::min(std::max(l, k), last);
  const int deg = k - deriv;
  long double M[8], left[8], right[8];
  M[0] = 1.0L;
  for (int j = 1; j <= deg; ++j) {
    left[j] = x - t[l + 1 - j];
    right[j] = t[l + j] - x;
    long double saved = 0.0L;
    for (int q = 0; q < j; ++q) {
      const long double tmp = M[q] / (right[q + 1] + left[j - q]);
      M[q] = saved + right[q + 1] * tmp;
      saved = left[j - q] * tmp;
    }
    M[j] = saved;
  }
  if (deriv == 0) {
    for (int a = 0; a <= k; ++a) N[a] = M[a];
  } else {  // N'_{i,k} = k (N_{i,k-1} / (t_{i+k} - t_i) - N_{i+1,k-1} / (t_{i+k+1} - t_{i+1}))
    for (int a = 0; a <= k; ++a) {
      const int i = l - k + a;
      const long double lo = a >= 1 ? M[a - 1] : 0.0L, hi = a <= k - 1 ? M[a] : 0.0L;
      const long double d1 = t[i + k] - t[i], d2 = t[i + k + 1] - t[i + 1];
      N[a] = k * ((d1 > 0 ? lo / d1 : 0.0L) - (d2 > 0 ? hi / d2 : 0.0L));
    }
  }
  return l - k;
}
}  // namespace

extern "C" int mf_filter_create(const double *xi, const double *wi, int nin, const double *xo,
                                int nout, const double *xn, int nn, int k, int deriv,
                                mf_filter **out) {
  if (!xi || !xo || !xn || !out) return mf::fail(MF_ERR_ARG, "mf_filter_create: null pointer");
  if (k < 1 || k > 5 || nn < 2 || nin < nn + k - 1 || nout < 1 || nout > 256 || deriv < 0 ||
      deriv > 1)
    return mf::fail(MF_ERR_ARG, "mf_filter_create: need 1 <= k <= 5, nn >= 2, nin >= nn+k-1, "
                                "1 <= nout <= 256, deriv 0|1");
  for (int r = 1; r < nin; ++r)
    if (!(xi[r] > xi[r - 1])) return mf::fail(MF_ERR_ARG, "mf_filter_create: xi must increase");
  for (int r = 1; r < nn; ++r)
    if (!(xn[r] > xn[r - 1])) return mf::fail(MF_ERR_ARG, "mf_filter_create: xn must increase");
  if (wi)
    for (int r = 0; r < nin; ++r)
      if (!(wi[r] > 0)) return mf::fail(MF_ERR_ARG, "mf_filter_create: weights must be positive");
  // the end nodes are the input end points (the knots below use xi's); two
  // grids built independently may differ there by rounding
  {
    const double span = std::fabs(xi[nin - 1]) + std::fabs(xi[0]) + (xi[nin - 1] - xi[0]);
    const double tol = 16.0 * 2.220446049250313e-16 * span;
    if (std::fabs(xn[0] - xi[0]) > tol || std::fabs(xn[nn - 1] - xi[nin - 1]) > tol)
      return mf::fail(MF_ERR_ARG, "mf_filter_create: the nodes must start and end at the input end points");
  }
  // knots (class_filter.f90:149-154)
  std::vector<long double> t;
  for (int q = 0; q <= k; ++q) t.push_back(xi[0]);
  for (int q = 1; q < nn - 1; ++q) t.push_back(xn[q]);
  for (int q = 0; q <= k; ++q) t.push_back(xi[nin - 1]);
  const int nc = nn + k - 1;
  // normal equations G = B^T W^2 B (nc x nc) and B^T W^2 (nc x nin)
  std::vector<long double> G((size_t)nc * nc, 0.0L), Bt((size_t)nc * nin, 0.0L), N(k + 1);
  for (int r = 0; r < nin; ++r) {
    const int f = bspline(t, k, xi[r], 0, N.data());
    const long double w2 = wi ? (long double)wi[r] * wi[r] : 1.0L;
    for (int a = 0; a <= k; ++a) {
      Bt[(size_t)(f + a) * nin + r] = w2 * N[a];
      for (int b = 0; b <= k; ++b) G[(size_t)(f + a) * nc + f + b] += w2 * N[a] * N[b];
    }
  }
  // Cholesky G = L L^T
  for (int j = 0; j < nc; ++j) {
    const long double g0 = G[(size_t)j * nc + j];
    long double d = g0;
    for (int q = 0; q < j; ++q) d -= G[(size_t)j * nc + q] * G[(size_t)j * nc + q];
    if (!(d > 1e-18L * g0))  // zero column, or a pivot lost to rounding
      return mf::fail(MF_ERR_SETUP, "mf_filter_create: a spline has no data points (singular fit)");
    d = std::sqrt(d);
    G[(size_t)j * nc + j] = d;
    for (int i = j + 1; i < nc; ++i) {
      long double v = G[(size_t)i * nc + j];
      for (int q = 0; q < j; ++q) v -= G[(size_t)i * nc + q] * G[(size_t)j * nc + q];
      G[(size_t)i * nc + j] = v / d;
    }
  }
  // C = G^-1 B^T (nc x nin): forward and back substitution per column
  for (int r = 0; r < nin; ++r) {
    for (int i = 0; i < nc; ++i) {
      long double v = Bt[(size_t)i * nin + r];
      for (int q = 0; q < i; ++q) v -= G[(size_t)i * nc + q] * Bt[(size_t)q * nin + r];
      Bt[(size_t)i * nin + r] = v / G[(size_t)i * nc + i];
    }
    for (int i = nc - 1; i >= 0; --i) {
      long double v = Bt[(size_t)i * nin + r];
      for (int q = i + 1; q < nc; ++q) v -= G[(size_t)q * nc + i] * Bt[(size_t)q * nin + r];
      Bt[(size_t)i * nin + r] = v / G[(size_t)i * nc + i];
    }
  }
  mf_filter *h = new mf_filter();
  h->nin = nin;
  h->nout = nout;
  h->F.assign((size_t)nout * nin, 0.0);
  for (int o = 0; o < nout; ++o) {
    const int f = bspline(t, k, xo[o], deriv, N.data());
    for (int r = 0; r < nin; ++r) {
      long double v = 0.0L;
      for (int a = 0; a <= k; ++a) v += N[a] * Bt[(size_t)(f + a) * nin + r];
      h->F[o + (size_t)nout * r] = (double)v;
    }
  }
  std::vector<double> Ft((size_t)nin * nout);
  for (int o = 0; o < nout; ++o)
    for (int r = 0; r < nin; ++r) Ft[r + (size_t)nin * o] = h->F[o + (size_t)nout * r];
  hipError_t e = hipMalloc(&h->dF, h->F.size() * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&h->dFt, Ft.size() * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(h->dF, h->F.data(), h->F.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(h->dFt, Ft.data(), Ft.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (h->dF) (void)hipFree(h->dF);
    if (h->dFt) (void)hipFree(h->dFt);
    delete h;
    return mf::fail(MF_ERR_DEVICE, std::string("mf_filter_create: ") + hipGetErrorString(e));
  }
  *out = h;
  return MF_OK;
}

extern "C" int mf_filter_destroy(mf_filter *h) {
  if (!h) return MF_OK;
  if (h->dF) (void)hipFree(h->dF);
  if (h->dFt) (void)hipFree(h->dFt);
  delete h;
  return MF_OK;
}

extern "C" int mf_filter_matrix(const mf_filter *h, double *F) {
  if (!h || !F) return mf::fail(MF_ERR_ARG, "mf_filter_matrix: null pointer");
  std::memcpy(F, h->F.data(), h->F.size() * sizeof(double));
  return MF_OK;
}

namespace {
template <bool WLEFT>
void launch_fgemm(const double *A, long lda, long sA, const double *B, long ldb, double *C, long ldc,
                  long sC, int M, int N, int K, int batch, hipStream_t st) {
  mf::EpiArgs ep;
  std::memset(&ep, 0, sizeof(ep));
  const int d = WLEFT ? M : N;
  const int t = d <= 64 ? 64 : (d <= 128 ? 128 : 256);
  auto go = [&](auto bm, auto bn) {
    constexpr int BM = decltype(bm)::value, BN = decltype(bn)::value;
    dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN, batch);
    hipLaunchKernelGGL((mf::gemm_mode_kernel<double, double, double, double, BM, BN, mf::EPI_STORE>),
                       grid, dim3(256), 0, st, A, lda, sA, B, ldb, C, ldc, sC, M, N, K, ep);
  };
  using I64 = std::integral_constant<int, 64>;
  using I128 = std::integral_constant<int, 128>;
  using I256 = std::integral_constant<int, 256>;
  if (WLEFT) {
    if (t == 64) go(I64(), I256());
    else if (t == 128) go(I128(), I128());
    else go(I256(), I64());
  } else {
    if (t == 64) go(I256(), I64());
    else if (t == 128) go(I128(), I128());
    else go(I64(), I256());
  }
}
}  // namespace

extern "C" int mf_filter3d(const mf_filter *fx, const mf_filter *fy, const mf_filter *fz,
                           const double *in, double *out, double *work, void *stream) {
  if (!fx || !fy || !fz || !in || !out || !work) return mf::fail(MF_ERR_ARG, "mf_filter3d: null pointer");
  const int nxi = fx->nin, nyi = fy->nin, nzi = fz->nin;
  const int nxo = fx->nout, nyo = fy->nout, nzo = fz->nout;
  hipStream_t st = (hipStream_t)stream;
  double *w1 = work, *w2 = work + (size_t)nxo * nyi * nzi;
  // x: w1 (nxo, nyi nzi) = Fx in
  launch_fgemm<true>(fx->dF, nxo, 0, in, nxi, w1, nxo, 0, nxo, nyi * nzi, nxi, 1, st);
  // y: w2_k (nxo, nyo) = w1_k Fy^T, batched over z
  launch_fgemm<false>(w1, nxo, (long)nxo * nyi, fy->dFt, nyi, w2, nxo, (long)nxo * nyo, nxo, nyo, nyi,
                      nzi, st);
  // z: out (nxo nyo, nzo) = w2 Fz^T
  launch_fgemm<false>(w2, (long)nxo * nyo, 0, fz->dFt, nzi, out, (long)nxo * nyo, 0, nxo * nyo, nzo,
                      nzi, 1, st);
  LESCHK(hipGetLastError());
  return MF_OK;
}

// ---- mf_filter3d without allocation or workspace argument -----------------
// The two intermediate fields of mf_filter3d live in one file-scope device
// array (each GPU holds its own copy, for the life of the process).  Host
// code never passes the array's name to a kernel: it is a host-side
// stand-in there.  Its device address comes from hipGetSymbolAddress, once
// per device, with that device current, on the first call made there.
namespace {
static __device__ __align__(256) double scratch[MF_FILTER_SCRATCH_DOUBLES];

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
double *scratch_ptr[kMaxDevices] = {};

int scratch_on_current_device(double **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return mf::fail(MF_ERR_DEVICE, std::string("mf_filter3d_noalloc: hipGetDevice: ") +
                                       hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return mf::fail(MF_ERR_DEVICE, "mf_filter3d_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return mf::fail(MF_ERR_DEVICE, std::string("mf_filter3d_noalloc: hipGetSymbolAddress: ") +
                                         hipGetErrorString(e));
    scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = scratch_ptr[dev];
  return MF_OK;
}
}  // namespace

extern "C" int mf_filter3d_noalloc(const mf_filter *fx, const mf_filter *fy, const mf_filter *fz,
                                   const double *in, double *out, void *stream) {
  if (!fx || !fy || !fz || !in || !out)
    return mf::fail(MF_ERR_ARG, "mf_filter3d_noalloc: null pointer");
  const size_t need = (size_t)fx->nout * fy->nin * fz->nin + (size_t)fx->nout * fy->nout * fz->nin;
  if (need > (size_t)MF_FILTER_SCRATCH_DOUBLES)
    return mf::fail(MF_ERR_ARG, "mf_filter3d_noalloc: nxo*nyi*nzi + nxo*nyo*nzi exceeds "
                                "MF_FILTER_SCRATCH_DOUBLES");
  double *ws = nullptr;
  const int rc = scratch_on_current_device(&ws);
  if (rc != MF_OK) return rc;
  // every element of both intermediates is written before it is read: no
  // clearing, so the previous call's contents do not matter
  return mf_filter3d(fx, fy, fz, in, out, ws, stream);
}

void mf::les_spacing2(const double *g, int n, double *out) {
  for (int i = 0; i < n; ++i) {
    const double d = (i == 0) ? g[1] - g[0] : (i == n - 1 ? g[i] - g[i - 1] : 0.5 * (g[i + 1] - g[i - 1]));
    out[i] = d * d;
  }
}

void mf::les_nu_launch(const double *const du[9], const double *d2x, const double *d2y,
                       const double *d2z, int nx, int ny, int nz, int type, double c, double *nu,
                       void *stream) {
  mf::les::NuArgs a;
  for (int q = 0; q < 9; ++q) a.du[q] = du[q];
  a.d2x = d2x; a.d2y = d2y; a.d2z = d2z;
  a.nx = nx; a.ny = ny; a.nz = nz; a.type = type; a.c = c; a.nu = nu;
  const long N = (long)nx * ny * nz;
  hipLaunchKernelGGL(mf::les::les_nu_kernel, dim3(mf::les::nblk(N)), dim3(256), 0,
                     (hipStream_t)stream, a);
}

extern "C" int mf_les_nu(const double *const *du, const double *gx, int nx, const double *gy, int ny,
                         const double *gz, int nz, int les_type, double c, double *nu, void *stream) {
  if (!du || !gx || !gy || !gz || !nu) return mf::fail(MF_ERR_ARG, "mf_les_nu: null pointer");
  if (les_type < 1 || les_type > 3) return mf::fail(MF_ERR_ARG, "mf_les_nu: les_type must be 1, 2 or 3");
  if (nx < 2 || ny < 2 || nz < 2) return mf::fail(MF_ERR_ARG, "mf_les_nu: sizes must be >= 2");
  for (int q = 0; q < 9; ++q)
    if (!du[q]) return mf::fail(MF_ERR_ARG, "mf_les_nu: null gradient field");
  // squared spacings (navier_LES :1416-1446)
  std::vector<double> d2((size_t)nx + ny + nz);
  mf::les_spacing2(gx, nx, d2.data());
  mf::les_spacing2(gy, ny, d2.data() + nx);
  mf::les_spacing2(gz, nz, d2.data() + nx + ny);
  hipStream_t st = (hipStream_t)stream;
  double *dd = nullptr;
  LESCHK(hipMallocAsync(&dd, d2.size() * sizeof(double), st));
  LESCHK(hipMemcpyAsync(dd, d2.data(), d2.size() * sizeof(double), hipMemcpyHostToDevice, st));
  mf::les_nu_launch(du, dd, dd + nx, dd + nx + ny, nx, ny, nz, les_type, c, nu, st);
  LESCHK(hipGetLastError());
  LESCHK(hipFreeAsync(dd, st));
  LESCHK(hipStreamSynchronize(st));  // d2 goes out of scope
  return MF_OK;
}

// ---------------------------------------------------------------------------
// MLP handle
// ---------------------------------------------------------------------------
struct mf_mlp {
  float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;  // NPARAM (+1 for g)
  unsigned short *wb = nullptr;
  float *part = nullptr;
  long part_blocks = 0;
  long step = 0;
};

extern "C" int mf_mlp_create(const float *params, mf_mlp **out) {
  using namespace mf::les;
  if (!params || !out) return mf::fail(MF_ERR_ARG, "mf_mlp_create: null pointer");
  mf_mlp *h = new mf_mlp();
  hipError_t e = hipMalloc(&h->p, 4 * (size_t)(NPARAM + 1) * sizeof(float));
  if (e == hipSuccess) {
    h->m = h->p + (NPARAM + 1);
    h->v = h->m + (NPARAM + 1);
    h->g = h->v + (NPARAM + 1);
    e = hipMemset(h->p, 0, 4 * (size_t)(NPARAM + 1) * sizeof(float));
  }
  if (e == hipSuccess) e = hipMalloc(&h->wb, (NW1 + NW2 + NW3) * sizeof(unsigned short));
  if (e == hipSuccess) e = hipMemcpy(h->p, params, NPARAM * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (h->p) (void)hipFree(h->p);
    if (h->wb) (void)hipFree(h->wb);
    delete h;
    return mf::fail(MF_ERR_DEVICE, std::string("mf_mlp_create: ") + hipGetErrorString(e));
  }
  hipLaunchKernelGGL(round_kernel, dim3(nblk(NW1 + NW2 + NW3)), dim3(256), 0, nullptr,
                     (const float *)h->p, h->wb);
  if (hipDeviceSynchronize() != hipSuccess) {
    mf_mlp *hh = h;
    (void)hipFree(hh->p);
    (void)hipFree(hh->wb);
    delete hh;
    return mf::fail(MF_ERR_DEVICE, "mf_mlp_create: rounding kernel");
  }
  *out = h;
  return MF_OK;
}

extern "C" int mf_mlp_destroy(mf_mlp *h) {
  if (!h) return MF_OK;
  if (h->p) (void)hipFree(h->p);
  if (h->wb) (void)hipFree(h->wb);
  if (h->part) (void)hipFree(h->part);
  delete h;
  return MF_OK;
}

extern "C" int mf_mlp_params(const mf_mlp *h, float *params) {
  if (!h || !params) return mf::fail(MF_ERR_ARG, "mf_mlp_params: null pointer");
  LESCHK(hipMemcpy(params, h->p, mf::les::NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  return MF_OK;
}

extern "C" int mf_mlp_forward(const mf_mlp *h, const float *X, float *Y, long long n, void *stream) {
  using namespace mf::les;
  if (!h || !X || !Y || n < 0) return mf::fail(MF_ERR_ARG, "mf_mlp_forward: invalid argument");
  if (n == 0) return MF_OK;
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)((n + SPB - 1) / SPB)), dim3(SPB), 0,
                     (hipStream_t)stream, X, Y, (long)n, (const unsigned short *)h->wb, (const float *)h->p);
  LESCHK(hipGetLastError());
  return MF_OK;
}

namespace {
int mlp_grad(mf_mlp *h, const float *X, const float *T, long n, hipStream_t st) {
  using namespace mf::les;
  const long nb = (n + SPB - 1) / SPB;
  if (nb > h->part_blocks) {
    if (h->part) {
      LESCHK(hipStreamSynchronize(st));
      (void)hipFree(h->part);
      h->part = nullptr;
    }
    LESCHK(hipMalloc(&h->part, (size_t)nb * (NPARAM + 1) * sizeof(float)));
    h->part_blocks = nb;
  }
  hipLaunchKernelGGL(mlp_grad_kernel, dim3((unsigned)nb), dim3(SPB), 0, st, X, T, n,
                     (const unsigned short *)h->wb, (const float *)h->p, h->part);
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3(nblk(NPARAM + 1)), dim3(256), 0, st,
                     (const float *)h->part, (int)nb, h->g);
  LESCHK(hipGetLastError());
  return MF_OK;
}
}  // namespace

extern "C" int mf_mlp_loss_grad(mf_mlp *h, const float *X, const float *T, long long n,
                                float *loss, float *grad, void *stream) {
  using namespace mf::les;
  if (!h || !X || !T || n < 1) return mf::fail(MF_ERR_ARG, "mf_mlp_loss_grad: invalid argument");
  hipStream_t st = (hipStream_t)stream;
  int rc = mlp_grad(h, X, T, (long)n, st);
  if (rc != MF_OK) return rc;
  if (grad) LESCHK(hipMemcpyAsync(grad, h->g, NPARAM * sizeof(float), hipMemcpyDeviceToDevice, st));
  if (loss) {
    LESCHK(hipMemcpyAsync(loss, h->g + NPARAM, sizeof(float), hipMemcpyDeviceToHost, st));
    LESCHK(hipStreamSynchronize(st));
  }
  return MF_OK;
}

extern "C" int mf_mlp_adam_step(mf_mlp *h, const float *X, const float *T, long long n, float lr,
                                float *loss, void *stream) {
  using namespace mf::les;
  if (!h || !X || !T || n < 1 || !(lr > 0.0f)) return mf::fail(MF_ERR_ARG, "mf_mlp_adam_step: invalid argument");
  hipStream_t st = (hipStream_t)stream;
  int rc = mlp_grad(h, X, T, (long)n, st);
  if (rc != MF_OK) return rc;
  h->step += 1;
  const float c1 = 1.0f / (1.0f - std::pow(0.9f, (float)h->step));
  const float c2 = 1.0f / (1.0f - std::pow(0.999f, (float)h->step));
  hipLaunchKernelGGL(adam_kernel, dim3(nblk(NPARAM)), dim3(256), 0, st, h->p, h->m, h->v,
                     (const float *)h->g, lr, c1, c2);
  hipLaunchKernelGGL(round_kernel, dim3(nblk(NW1 + NW2 + NW3)), dim3(256), 0, st, (const float *)h->p,
                     h->wb);
  LESCHK(hipGetLastError());
  if (loss) {
    LESCHK(hipMemcpyAsync(loss, h->g + NPARAM, sizeof(float), hipMemcpyDeviceToHost, st));
    LESCHK(hipStreamSynchronize(st));
  }
  return MF_OK;
}

// ---------------------------------------------------------------------------
// Learned-closure data path
// ---------------------------------------------------------------------------
extern "C" int mf_sgs_targets(const mf_filter *fx, const mf_filter *fy, const mf_filter *fz,
                              const double *u, const double *v, const double *w, double *ubar,
                              double *vbar, double *wbar, float *tau, double tscale, double *work,
                              void *stream) {
  using namespace mf::les;
  if (!fx || !fy || !fz || !u || !v || !w || !ubar || !vbar || !wbar || !tau || !work)
    return mf::fail(MF_ERR_ARG, "mf_sgs_targets: null pointer");
  if (!(tscale > 0.0)) return mf::fail(MF_ERR_ARG, "mf_sgs_targets: tscale must be positive");
  hipStream_t st = (hipStream_t)stream;
  const long Nf = (long)fx->nin * fy->nin * fz->nin;
  const long Nc = (long)fx->nout * fy->nout * fz->nout;
  double *prod = work, *fb = work + Nf;
  double *fw = fb + Nc;  // filter work
  const double *q[3] = {u, v, w};
  double *qb[3] = {ubar, vbar, wbar};
  for (int c = 0; c < 3; ++c) {
    int rc = mf_filter3d(fx, fy, fz, q[c], qb[c], fw, st);
    if (rc != MF_OK) return rc;
  }
  const int pi[6] = {0, 1, 2, 0, 0, 1}, pj[6] = {0, 1, 2, 1, 2, 2};
  for (int col = 0; col < 6; ++col) {
    hipLaunchKernelGGL(prod_kernel, dim3(nblk(Nf)), dim3(256), 0, st, q[pi[col]], q[pj[col]], prod, Nf);
    int rc = mf_filter3d(fx, fy, fz, prod, fb, fw, st);
    if (rc != MF_OK) return rc;
    hipLaunchKernelGGL(tau_kernel, dim3(nblk(Nc)), dim3(256), 0, st, (const double *)fb,
                       (const double *)qb[pi[col]], (const double *)qb[pj[col]], tau, col,
                       1.0 / tscale, Nc);
  }
  LESCHK(hipGetLastError());
  return MF_OK;
}

extern "C" int mf_sgs_features(const mf_der *dx, const mf_der *dy, const mf_der *dz,
                               const double *u, const double *v, const double *w, int nx, int ny,
                               int nz, const double *gx, const double *gy, const double *gz,
                               double gscale, double lscale, float *X, double *work,
                               void *stream) {
  if (!dx || !dy || !dz || !u || !v || !w || !gx || !gy || !gz || !X || !work)
    return mf::fail(MF_ERR_ARG, "mf_sgs_features: null pointer");
  if (nx < 2 || ny < 2 || nz < 2) return mf::fail(MF_ERR_ARG, "mf_sgs_features: sizes must be >= 2");
  hipStream_t st = (hipStream_t)stream;
  const long N = (long)nx * ny * nz;
  double *g[9];  // g[3 a + c] = d q_c / d x_a
  for (int q = 0; q < 9; ++q) g[q] = work + q * N;
  const double *in[3] = {u, v, w};
  const mf_der *dc[3] = {dx, dy, dz};
  for (int a = 0; a < 3; ++a) {
    int rc = mf_der3d_batch(dc[a], a, 1, 3, in, g + 3 * a, nx, ny, nz, st);
    if (rc != MF_OK) return rc;
  }
  const double *du[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) du[3 * i + j] = g[3 * j + i];
  std::vector<double> d2((size_t)nx + ny + nz);
  mf::les_spacing2(gx, nx, d2.data());
  mf::les_spacing2(gy, ny, d2.data() + nx);
  mf::les_spacing2(gz, nz, d2.data() + nx + ny);
  double *dd = nullptr;
  LESCHK(hipMallocAsync(&dd, d2.size() * sizeof(double), st));
  LESCHK(hipMemcpyAsync(dd, d2.data(), d2.size() * sizeof(double), hipMemcpyHostToDevice, st));
  mf::sgs_features_launch(du, dd, dd + nx, dd + nx + ny, nx, ny, nz, gscale, lscale, X, st);
  LESCHK(hipGetLastError());
  LESCHK(hipFreeAsync(dd, st));
  LESCHK(hipStreamSynchronize(st));  // d2 goes out of scope
  return MF_OK;
}
