// FP64 residual of the multiplied-through operator in one z-marching pass.
//
//   r' = P g_f - B u,   P = Mx My Mz,
//   B u = Rz (Mx My u) + Mz (Rx My u + Mx Ry u + sigma Mx My u)
// with M = lhs and R = rhs of each direction.  Every row of these banded
// operators fits a 5-wide window [c0, c0+4] with i-4 <= c0 <= i (lhs is
// tri- or pentadiagonal, rhs pentadiagonal, one-sided boundary rows reach 4
// points); the windows are centred (c0 = i-2) away from the boundaries.  So
//   r'_k = sum_p Mz[k,p] G_p - sum_p Rz[k,p] A_p - FB_k,
//   A_p = Mx My u_p,   G_p = Mx My (g - sigma u)_p - (Rx My + Mx Ry) u_p
// are 2-D stencils on plane p.  The boundary fold g_f = g - sum_faces
// inv(lhs) (boundary columns of rhs) bc becomes, after multiplying by P, a
// term FB on the two rows next to each face only:
//   P (f_x1 (x) bcx1) = e_x1 (x) (My Mz bcx1),  e_x1 = lhs_x f_x1 = the
// boundary column of rhs_x (rows 1, 2), and likewise for the other faces.
// The face products (My Mz bcx1 ...) are small 2-D arrays computed once per
// solve by face_fold_kernel.
//
// A block owns a 56 x 16 tile of (i,j) and a chunk of RZ_CZ output planes and
// marches over the planes the chunk's z-windows reach (RZ_CZ + 4 inside the
// domain).  Each plane of u and g is staged with a 4-point halo in x and y,
// 64 x 24 points: one wavefront spans the staged width, so wave w computes
// the y-stencils of output rows 4w..4w+3 (rows uniform per wave: inside the
// domain the four rows reuse one 8-row register segment) and then the
// x-stencils of the same rows, lane = column.  Each plane's A and G are added
// into the chunk's rows in registers; weights come from per-block LDS tables
// and planes are fetched two ahead of their use.  HBM traffic: u and g read
// once (halo re-reads hit L2), r' written once in FP32 (the inner solve's
// input).
#pragma once
#include <hip/hip_runtime.h>

namespace mf {

struct BandRows {      // one banded 1-D operator: row i acts on c0[i]..c0[i]+4
  const int *c0;
  const double *w;     // (n,8), 5 used
};

constexpr int BAND_W = 5;  // window width of every banded row
constexpr int RZ_H = 4;    // halo: the farthest reach of a row
constexpr int RZ_NW = 4, RZ_RPW = 4;  // wavefronts per block, tile rows per wavefront
constexpr int RZ_TX = 64 - 2 * RZ_H, RZ_TY = RZ_NW * RZ_RPW, RZ_CZ = 8;

struct ResArgs {
  int n1, n2, n3;
  const double *u, *g;   // interior origin of the full arrays: f[i + ldj*j + ldk*k]
  long ldj, ldk;
  double sigma;
  BandRows Mx, Rx, My, Ry, Mz, Rz;
  const double *ex, *ey, *ez;  // e_a1 (n_a) then e_a2 (n_a): nonzero on 2 rows each
  const double *fbx, *fby, *fbz;  // face products (n2,n3,2), (n1,n3,2), (n1,n2,2)
  float *out;            // r', interior-contiguous
  double *part;          // per-block partial sums of r'^2
  unsigned *amax;        // NULL or: max |r'| (float bits, atomic max per block)
  unsigned *zero;        // NULL or: nzero scalars block 0 sets to zero
  int nzero;
  const double *zq;      // residual_zs_kernel: Mz, Rz rows on planes b(k)..b(k)+4, e_z (n3,12)
  int cz;                // residual_zs_kernel: output planes per block
  double *out64;         // residual_zs_kernel<false>: NULL or b' in FP64 (interior-contiguous)
  const double *bp;      // residual_zs_kernel<true>: b' in FP64, read instead of g
};

// RES = false: right-hand side b' = P g_f (u is not read);
// RES = true:  residual r' = P g_f - B u.
// block (64, RZ_NW), grid: one block per tile (ceil(n1/RZ_TX) ceil(n2/RZ_TY)
// ceil(n3/RZ_CZ) in all, any shape); consecutive block ids go round-robin
// to the 8 XCDs, so the tile order is remapped to give each XCD a contiguous
// run of neighbouring tiles (their halos then hit the XCD's own L2).
template <bool RES>
__global__ void __launch_bounds__(64 * RZ_NW) residual_kernel(ResArgs a) {
  constexpr int NT = 64 * RZ_NW;
  constexpr int SW = RZ_TX + 2 * RZ_H, SH = RZ_TY + 2 * RZ_H;  // staged tile 64 x 24
  constexpr int SP = SW + 1;                                   // row pitch
  constexpr int NS = SW * SH / NT;                             // staged points per thread
  constexpr int SEG = RZ_RPW + BAND_W - 1;                     // centred y-window rows
  constexpr int NP = RZ_CZ + 2 * RZ_H;                         // most planes a chunk reaches
  static_assert(SW == 64 && NS * NT == SW * SH, "one wavefront per staged row");
  __shared__ double U[SH][SP], H[SH][SP];
  __shared__ double T1[RZ_TY][SP], T2[RZ_TY][SP], T3[RZ_TY][SP];
  __shared__ double WY[RZ_TY][2][BAND_W];       // My, Ry rows of the tile
  __shared__ int CY[RZ_TY][2];                  // their window starts, staged rows
  __shared__ double2 WZ[NP][RZ_CZ];             // (Mz, Rz) weight of plane s in row kk
  const int lane = threadIdx.x, tid = lane + 64 * threadIdx.y;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);  // wave index (uniform)
  const int n1 = a.n1, n2 = a.n2, n3 = a.n3;
  const int nbx = (n1 + RZ_TX - 1) / RZ_TX, nby = (n2 + RZ_TY - 1) / RZ_TY;
  int i0, j0, z0;
  {
    const int nb = gridDim.x, bid = blockIdx.x;
    const int q = nb / 8, r = nb % 8, x = bid % 8, l = bid / 8;
    const int t = (x < r) ? x * (q + 1) + l : r * (q + 1) + (x - r) * q + l;
    i0 = (t % nbx) * RZ_TX;
    j0 = ((t / nbx) % nby) * RZ_TY;
    z0 = (t / (nbx * nby)) * RZ_CZ;
  }
  // planes reached by the chunk's rows (window starts are non-decreasing)
  const int zl = min(z0 + RZ_CZ, n3) - 1;
  int p_lo = a.Mz.c0[z0], p_hi = a.Mz.c0[zl] + BAND_W - 1;
  if (RES) {
    p_lo = min(p_lo, a.Rz.c0[z0]);
    p_hi = max(p_hi, a.Rz.c0[zl] + BAND_W - 1);
  }
  // per-block weight tables
  if (tid < RZ_TY * 2 * BAND_W) {
    const int r = tid / (2 * BAND_W), q = (tid / BAND_W) % 2, m = tid % BAND_W;
    const int j = min(j0 + r, n2 - 1);
    const BandRows &B = q == 0 ? a.My : a.Ry;
    WY[r][q][m] = (q == 0 || RES) ? B.w[8 * j + m] : 0.0;
    if (m == 0) CY[r][q] = B.c0[j] - (j0 - RZ_H);
  }
  for (int e = tid; e < NP * RZ_CZ; e += NT) {
    const int sp = e / RZ_CZ, kk = e % RZ_CZ;
    const int p = p_lo + sp, k = min(z0 + kk, n3 - 1);
    const int dm = p - a.Mz.c0[k], dr = p - a.Rz.c0[k];
    double2 v;
    v.x = (dm >= 0 && dm < BAND_W) ? a.Mz.w[8 * k + dm] : 0.0;
    v.y = (RES && dr >= 0 && dr < BAND_W) ? a.Rz.w[8 * k + dr] : 0.0;
    WZ[sp][kk] = v;
  }
  // staging: thread stages column `lane` of rows w, w+4, ... (clamped; out of range -> 0)
  const int si = i0 - RZ_H + lane;
  const bool iok = si >= 0 && si < n1;
  const int sic = min(max(si, 0), n1 - 1);
  long soff[NS];
  bool sok[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int sj = j0 - RZ_H + w + RZ_NW * s;
    sok[s] = iok && sj >= 0 && sj < n2;
    soff[s] = sic + a.ldj * min(max(sj, 0), n2 - 1);
  }
  // two register sets: planes are fetched two ahead of their use
  double ru[2][NS], rg[2][NS];
  auto fetch = [&](int p, int b) {
    const long po = a.ldk * p;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      rg[b][s] = a.g[soff[s] + po];
      if (RES) ru[b][s] = a.u[soff[s] + po];
    }
  };
  fetch(p_lo, 0);
  if (p_lo + 1 <= p_hi) fetch(p_lo + 1, 1);
  // x: lane = staged column; outputs for lanes RZ_H .. RZ_H+RZ_TX-1
  const int i = i0 - RZ_H + lane;
  const bool xout = lane >= RZ_H && lane < RZ_H + RZ_TX && i < n1;
  const int ic = min(max(i, 0), n1 - 1);
  // (halo lanes compute nothing kept; their windows are clamped into the tile)
  const int cmx = min(max(a.Mx.c0[ic] - (i0 - RZ_H), 0), SW - BAND_W);
  const int crx = min(max(a.Rx.c0[ic] - (i0 - RZ_H), 0), SW - BAND_W);
  double wmx[BAND_W], wrx[BAND_W];
#pragma unroll
  for (int m = 0; m < BAND_W; ++m) {
    wmx[m] = a.Mx.w[8 * ic + m];
    wrx[m] = RES ? a.Rx.w[8 * ic + m] : 0.0;
  }
  double acc[RZ_CZ][RZ_RPW];
#pragma unroll
  for (int kk = 0; kk < RZ_CZ; ++kk)
#pragma unroll
    for (int d = 0; d < RZ_RPW; ++d) acc[kk][d] = 0.0;
  __syncthreads();
  // y rows of this wave: output rows R w .. R w + R-1 (R = RZ_RPW); centred
  // windows use staged rows R w + d + 2 .. R w + d + 6, one SEG-row segment
  const int jr0 = j0 + RZ_RPW * w;
  bool yfast = jr0 + RZ_RPW - 1 < n2;
#pragma unroll
  for (int d = 0; d < RZ_RPW; ++d) {
    const int r = RZ_RPW * w + d;
    yfast = yfast && CY[r][0] == r + 2 && (!RES || CY[r][1] == r + 2);
  }

  auto plane = [&](int p, int b) {
    // 1. stage plane p: U = u, H = g - sigma u
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const double uv = (RES && sok[s]) ? ru[b][s] : 0.0;
      const double gv = sok[s] ? rg[b][s] : 0.0;
      U[w + RZ_NW * s][lane] = uv;
      H[w + RZ_NW * s][lane] = RES ? fma(-a.sigma, uv, gv) : gv;
    }
    __syncthreads();
    if (p + 2 <= p_hi) fetch(p + 2, b);
    // 2. y-stencils of the wave's rows, column = lane
    if (yfast) {
      double su[SEG], sh[SEG];
#pragma unroll
      for (int t = 0; t < SEG; ++t) {
        sh[t] = H[RZ_RPW * w + 2 + t][lane];
        if (RES) su[t] = U[RZ_RPW * w + 2 + t][lane];
      }
#pragma unroll
      for (int d = 0; d < RZ_RPW; ++d) {
        const double *wm = WY[RZ_RPW * w + d][0], *wr = WY[RZ_RPW * w + d][1];
        double t1 = 0.0, t2 = 0.0, t3 = 0.0;
#pragma unroll
        for (int m = 0; m < BAND_W; ++m) {
          t3 = fma(wm[m], sh[d + m], t3);
          if (RES) {
            t1 = fma(wm[m], su[d + m], t1);
            t2 = fma(wr[m], su[d + m], t2);
          }
        }
        T3[RZ_RPW * w + d][lane] = t3;
        if (RES) {
          T1[RZ_RPW * w + d][lane] = t1;
          T2[RZ_RPW * w + d][lane] = t2;
        }
      }
    } else {
#pragma unroll
      for (int d = 0; d < RZ_RPW; ++d) {
        const int r = RZ_RPW * w + d;
        const int cm = CY[r][0], cr = CY[r][1];
        double t1 = 0.0, t2 = 0.0, t3 = 0.0;
#pragma unroll
        for (int m = 0; m < BAND_W; ++m) {
          t3 = fma(WY[r][0][m], H[cm + m][lane], t3);
          if (RES) {
            t1 = fma(WY[r][0][m], U[cm + m][lane], t1);
            t2 = fma(WY[r][1][m], U[cr + m][lane], t2);
          }
        }
        T3[r][lane] = t3;
        if (RES) {
          T1[r][lane] = t1;
          T2[r][lane] = t2;
        }
      }
    }
    __syncthreads();
    // 3. x-stencils of the wave's rows at column `lane`, then into the chunk's rows
    const int sp = p - p_lo;
#pragma unroll
    for (int d = 0; d < RZ_RPW; ++d) {
      const int r = RZ_RPW * w + d;
      double A = 0.0, G = 0.0;
#pragma unroll
      for (int m = 0; m < BAND_W; ++m) {
        G = fma(wmx[m], T3[r][cmx + m], G);
        if (RES) {
          A = fma(wmx[m], T1[r][cmx + m], A);
          G = fma(-wrx[m], T1[r][crx + m], G);
          G = fma(-wmx[m], T2[r][cmx + m], G);
        }
      }
#pragma unroll
      for (int kk = 0; kk < RZ_CZ; ++kk) {
        const double2 wz = WZ[sp][kk];
        acc[kk][d] = fma(wz.x, G, acc[kk][d]);
        if (RES) acc[kk][d] = fma(-wz.y, A, acc[kk][d]);
      }
    }
    // (the next plane's first barrier orders these T reads before its writes)
  };
  for (int p = p_lo; p <= p_hi; p += 2) {
    plane(p, 0);
    if (p + 1 <= p_hi) plane(p + 1, 1);
  }
  // 4. boundary fold, store r' (FP32), norm partials
  double contrib = 0.0;
  float vmax = 0.0f;
  if (xout) {
    const double ex1 = a.ex[i], ex2 = a.ex[n1 + i];
    const long sx = (long)n2 * n3, sy = (long)n1 * n3, sz = (long)n1 * n2;
#pragma unroll
    for (int d = 0; d < RZ_RPW; ++d) {
      const int j = jr0 + d;
      if (j >= n2) continue;
      const double ey1 = a.ey[j], ey2 = a.ey[n2 + j];
#pragma unroll
      for (int kk = 0; kk < RZ_CZ; ++kk) {
        const int k = z0 + kk;
        if (k >= n3) continue;
        double v = acc[kk][d];
        if (ex1 != 0.0) v = fma(-ex1, a.fbx[j + (long)n2 * k], v);
        if (ex2 != 0.0) v = fma(-ex2, a.fbx[sx + j + (long)n2 * k], v);
        if (ey1 != 0.0) v = fma(-ey1, a.fby[i + (long)n1 * k], v);
        if (ey2 != 0.0) v = fma(-ey2, a.fby[sy + i + (long)n1 * k], v);
        const double ez1 = a.ez[k], ez2 = a.ez[n3 + k];
        if (ez1 != 0.0) v = fma(-ez1, a.fbz[i + (long)n1 * j], v);
        if (ez2 != 0.0) v = fma(-ez2, a.fbz[sz + i + (long)n1 * j], v);
        a.out[i + (long)n1 * (j + (long)n2 * k)] = (float)v;
        contrib = fma(v, v, contrib);
        vmax = fmaxf(vmax, fabsf((float)v));
      }
    }
  }
  __shared__ double wsum[RZ_NW];
  __shared__ float wmax[RZ_NW];
  for (int off = 32; off > 0; off >>= 1) contrib += __shfl_down(contrib, off, 64);
  if (a.amax)
    for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off, 64));
  if (lane == 0) {
    wsum[w] = contrib;
    wmax[w] = vmax;
  }
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int q = 0; q < RZ_NW; ++q) t += wsum[q];
    a.part[blockIdx.x] = t;
    if (a.amax) {
      float b = wmax[0];
      for (int q = 1; q < RZ_NW; ++q) b = fmaxf(b, wmax[q]);
      atomicMax(a.amax, __float_as_uint(b));
    }
    if (a.zero && blockIdx.x == 0)
      for (int q = 0; q < a.nzero; ++q) a.zero[q] = 0u;
  }
}

// Face products of the boundary fold: for the x faces F(j,k,s) =
// sum My[j,j'] Mz[k,k'] bcx(j',k',s), and likewise (Mx, Mz) for y and
// (Mx, My) for z.  One thread per face point; grid covers
// 2 (n2 n3 + n1 n3 + n1 n2) points.
struct FaceFoldArgs {
  int n1, n2, n3;
  BandRows Mx, My, Mz;
  const double *bcx, *bcy, *bcz;
  double *fbx, *fby, *fbz;
  unsigned *zero;  // NULL or a scalar thread 0 sets to zero
};

__global__ void __launch_bounds__(256) face_fold_kernel(FaceFoldArgs f) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f.zero && t == 0) *f.zero = 0u;
  const long sx = (long)f.n2 * f.n3, sy = (long)f.n1 * f.n3, sz = (long)f.n1 * f.n2;
  const double *src;
  double *dst;
  BandRows A, B;
  int la;   // leading size of the face array
  long fs;  // points per face
  if (t < 2 * sx) { src = f.bcx; dst = f.fbx; A = f.My; B = f.Mz; la = f.n2; fs = sx; }
  else if ((t -= 2 * sx) < 2 * sy) { src = f.bcy; dst = f.fby; A = f.Mx; B = f.Mz; la = f.n1; fs = sy; }
  else if ((t -= 2 * sy) < 2 * sz) { src = f.bcz; dst = f.fbz; A = f.Mx; B = f.My; la = f.n1; fs = sz; }
  else return;
  const int side = (int)(t / fs);
  const long r = t % fs;
  const int ia = (int)(r % la), ib = (int)(r / la);
  const double *s0 = src + side * fs;
  const int ca = A.c0[ia], cb = B.c0[ib];
  double v = 0.0;
  for (int mb = 0; mb < BAND_W; ++mb) {
    double row = 0.0;
    for (int ma = 0; ma < BAND_W; ++ma)
      row = fma(A.w[8 * ia + ma], s0[(ca + ma) + (long)la * (cb + mb)], row);
    v = fma(B.w[8 * ib + mb], row, v);
  }
  dst[side * fs + r] = v;
}

}  // namespace mf
