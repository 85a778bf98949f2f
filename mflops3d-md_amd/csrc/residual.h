// FP64 residual of the multiplied-through operator in two passes.
//
//   B u = Rz (Mx My u) + Mz (Rx My u + Mx Ry u + sigma Mx My u)
//
// Pass 1 (plane_kernel), per z-plane: a = Mx My u and c = (Rx My + Mx Ry +
// sigma Mx My) u from one LDS-staged 64 x 16 tile of the plane with a 6-point
// halo (the widest one-sided boundary row reaches 6 points inward).
// Pass 2 (zpass_kernel): r' = b' - Rz a - Mz c along z, coalesced over i,
// with per-block partial sums of |r'|^2.
// The right-hand side b' = Mz (Mx My g_f) uses the same passes; plane_kernel
// then folds the boundary values into g on the fly (g_f never stored).
#pragma once
#include <hip/hip_runtime.h>

namespace mf {

struct BandRows {      // one banded 1-D operator: row i acts on c0[i]..c0[i]+6
  const int *c0;
  const double *w;     // (n,8)
};

constexpr int PT_X = 64, PT_Y = 16, PT_H = 6;

struct PlaneArgs {
  int n1, n2, n3;
  const double *u;     // input (strided): u[i + ldj*j + ldk*k]
  long ldj, ldk;
  double *a, *c;       // outputs, interior-contiguous
  BandRows Mx, Rx, My, Ry;
  double sigma;
  // fold on the fly (RHS mode): g_f = g - fold terms
  const double *fold, *bcx, *bcy, *bcz;
};

// MODE 0: residual pass 1 on u; MODE 1: RHS pass 1 on g (a only, folded)
// block (64,4), grid (ceil(n1/64), ceil(n2/16), n3)
template <int MODE>
__global__ void __launch_bounds__(256) plane_kernel(PlaneArgs p) {
  __shared__ double U[PT_Y + 2 * PT_H][PT_X + 2 * PT_H + 1];
  __shared__ double T1[PT_Y][PT_X + 2 * PT_H + 1];
  __shared__ double T2[PT_Y][PT_X + 2 * PT_H + 1];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = blockIdx.x * PT_X, j0 = blockIdx.y * PT_Y, k = blockIdx.z;
  const int n1 = p.n1, n2 = p.n2;
  const int tid = tx + 64 * ty;
  // 1. plane tile with halo
  constexpr int UW = PT_X + 2 * PT_H, UH = PT_Y + 2 * PT_H;
  for (int e = tid; e < UW * UH; e += 256) {
    const int li = e % UW, lj = e / UW;
    const int i = i0 - PT_H + li, j = j0 - PT_H + lj;
    double v = 0.0;
    if (i >= 0 && i < n1 && j >= 0 && j < n2) {
      v = p.u[i + p.ldj * j + p.ldk * k];
      if (MODE == 1) {
        const int n3 = p.n3;
        const double *fx1 = p.fold, *fx2 = fx1 + n1, *fy1 = fx2 + n1, *fy2 = fy1 + n2;
        const double *fz1 = fy2 + n2, *fz2 = fz1 + n3;
        const long sx = (long)n2 * n3, sy = (long)n1 * n3, sz = (long)n1 * n2;
        v -= fx1[i] * p.bcx[j + (long)n2 * k];
        v -= fx2[i] * p.bcx[sx + j + (long)n2 * k];
        v -= fy1[j] * p.bcy[i + (long)n1 * k];
        v -= fy2[j] * p.bcy[sy + i + (long)n1 * k];
        v -= fz1[k] * p.bcz[i + (long)n1 * j];
        v -= fz2[k] * p.bcz[sz + i + (long)n1 * j];
      }
    }
    U[lj][li] = v;
  }
  __syncthreads();
  // 2. y-stencils on the tile's rows, all halo columns
  for (int e = tid; e < UW * PT_Y; e += 256) {
    const int li = e % UW, lr = e / UW;
    const int j = j0 + lr;
    double t1 = 0.0, t2 = 0.0;
    if (j < n2) {
      const int cm = p.My.c0[j] - (j0 - PT_H), cr = p.Ry.c0[j] - (j0 - PT_H);
      const double *wm = p.My.w + 8 * j, *wr = p.Ry.w + 8 * j;
#pragma unroll
      for (int m = 0; m < 7; ++m) {
        t1 = fma(wm[m], U[cm + m][li], t1);
        if (MODE == 0) t2 = fma(wr[m], U[cr + m][li], t2);
      }
    }
    T1[lr][li] = t1;
    T2[lr][li] = t2;
  }
  __syncthreads();
  // 3. x-stencils, 4 rows per thread
  const int i = i0 + tx;
  if (i >= n1) return;
  const int cm = p.Mx.c0[i] - (i0 - PT_H), cr = p.Rx.c0[i] - (i0 - PT_H);
  const double *wm = p.Mx.w + 8 * i, *wr = p.Rx.w + 8 * i;
  for (int r = ty; r < PT_Y; r += 4) {
    const int j = j0 + r;
    if (j >= n2) break;
    double a = 0.0, c = 0.0;
#pragma unroll
    for (int m = 0; m < 7; ++m) {
      a = fma(wm[m], T1[r][cm + m], a);
      if (MODE == 0) {
        c = fma(wr[m], T1[r][cr + m], c);
        c = fma(wm[m], T2[r][cm + m], c);
      }
    }
    const long q = i + (long)n1 * (j + (long)n2 * k);
    p.a[q] = a;
    if (MODE == 0) p.c[q] = c + p.sigma * a;
  }
}

struct ZArgs {
  int n1, n2, n3;
  const double *a, *c;   // interior-contiguous
  const double *bprime;  // MODE 0
  double *out;           // r' (MODE 0) or b' (MODE 1)
  double *part;          // per-block partial sums of |out|^2
  BandRows Mz, Rz;
};

// MODE 0: out = b' - Rz a - Mz c;  MODE 1: out = Mz a.  Norm partials.
// block (64,4), grid (ceil(n1/64), ceil(n2/4), n3)
template <int MODE>
__global__ void __launch_bounds__(256) zpass_kernel(ZArgs z) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  double contrib = 0.0;
  if (i < z.n1 && j < z.n2) {
    const long plane = (long)z.n1 * z.n2;
    const long base = i + (long)z.n1 * j;
    const int cm = z.Mz.c0[k];
    const double *wm = z.Mz.w + 8 * k;
    double v = 0.0;
    if (MODE == 0) {
      const int cr = z.Rz.c0[k];
      const double *wr = z.Rz.w + 8 * k;
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < 7; ++m) {
        s = fma(wr[m], z.a[base + plane * (cr + m)], s);
        s = fma(wm[m], z.c[base + plane * (cm + m)], s);
      }
      v = z.bprime[base + plane * k] - s;
    } else {
#pragma unroll
      for (int m = 0; m < 7; ++m) v = fma(wm[m], z.a[base + plane * (cm + m)], v);
    }
    z.out[base + plane * k] = v;
    contrib = v * v;
  }
  __shared__ double wsum[4];
  for (int off = 32; off > 0; off >>= 1) contrib += __shfl_down(contrib, off, 64);
  if (threadIdx.x == 0) wsum[threadIdx.y] = contrib;
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0)
    z.part[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] =
        (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
}

}  // namespace mf
