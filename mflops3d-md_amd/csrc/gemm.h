// MFMA tensor-mode products for the diagonalisation solver (gfx950).
//
// Every eigenvector transform of solve_3d (dgx/dgy/dgz,
// class_solver_3d.f90:307-378) is a column-major GEMM C = A B:
//   x-transform:  C(n1, n2 n3) = W(n1,n1) . F(n1, n2 n3)          (W left)
//   y-transform:  C_k(n1, n2)  = F_k(n1,n2) . Wt(n2,n2), k = 1..n3 (W right, batched)
//   z-transform:  C(n1 n2, n3) = F(n1 n2, n3) . Wt(n3,n3)         (W right)
// In all three A is M-contiguous and B is K-contiguous, so one kernel serves
// them.  The block tile spans the whole eigenvector dimension (256 >= n), so
// each field element is read from HBM exactly once per transform and the
// eigenvector matrix streams from L2.
//
// MFMA operand order is swapped (D = B^T-fragment x A-fragment) so that the
// 16 lanes holding one D row are 16 consecutive m: stores of the
// column-major C are 128-byte contiguous segments.
#pragma once
#include <hip/hip_runtime.h>

namespace mf {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
// Operand layout (both types): lane l supplies A[l%16][l/16] and B[l/16][l%16].
// Output layout differs: result r of lane l is D[row(l,r)][l%16] with
// row = l/16 + 4r for the f64 instruction and 4(l/16) + r for the f32 one
// (measured on gfx950).
template <>
struct Mfma<double> {
  typedef d4_t acc_t;
  static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
  typedef f4_t acc_t;
  static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

enum EpiMode { EPI_STORE = 0, EPI_DIVIDE = 1, EPI_SCATTER = 2, EPI_SCATTER_ADD = 3 };

struct EpiArgs {
  // EPI_DIVIDE: m = i1 + n1*i2, n = i3 -> value / (lx[i1]+ly[i2]+lz[i3]+sigma),
  // 0 where max(|sigma|,|lx|,|ly|,|lz|) < 1e-11 (class_solver_3d.f90:202-225)
  const double *lx, *ly, *lz;
  double sigma;
  int n1, n2;
  // EPI_SCATTER(_ADD): m = i, n = j + n2*k -> u[(i+1) + nx((j+1) + ny(k+1))]
  double *u;
  int nx, ny;
};

// Block tile BM x BN on 4 waves.  A side of 64/128/256 is split into 64-wide
// wave tiles (4 x 4 MFMA blocks of 16); a side of 192 into 4 wave tiles of 48
// (3 MFMA blocks), for eigenvector dimensions of 129..192 points.
template <typename T, typename TA, typename TB, typename TC, int BM, int BN, int EPI>
__global__ void __launch_bounds__(256)
    gemm_mode_kernel(const TA *__restrict__ A, long lda, long sA, const TB *__restrict__ B,
                     long ldb, TC *__restrict__ C, long ldc, long sC, int M, int N, int K,
                     EpiArgs ep) {
  constexpr int BK = 16;
  constexpr int NA = BM * BK / 256;
  constexpr int NB = BN * BK / 256;
  // wave grid WM x WN, each wave MI x NI MFMA blocks of 16 x 16
  constexpr int MI = BM == 192 ? 3 : 4, NI = BN == 192 ? 3 : 4;
  constexpr int WM = BM / (16 * MI), WN = BN / (16 * NI);
  static_assert(WM * WN == 4 && WM * 16 * MI == BM && WN * 16 * NI == BN, "4 wave tiles");
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ T As[BK][BM];
  __shared__ T Bs[BN][BK + 1];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const long m0 = (long)blockIdx.x * BM;
  const long n0 = (long)blockIdx.y * BN;
  A += (long)blockIdx.z * sA;
  if (EPI == EPI_STORE || EPI == EPI_DIVIDE) C += (long)blockIdx.z * sC;
  // four wave tiles arranged WM x WN
  const int wm = (wave % WM) * 16 * MI;
  const int wn = (wave / WM) * 16 * NI;

  acc_t acc[MI][NI];
#pragma unroll
  for (int a = 0; a < MI; ++a)
#pragma unroll
    for (int b = 0; b < NI; ++b) acc[a][b] = (acc_t){0, 0, 0, 0};

  T ra[NA], rb[NB];
  auto load = [&](int kt) {
#pragma unroll
    for (int q = 0; q < NA; ++q) {
      const int e = tid + 256 * q;
      const long m = m0 + (e % BM);
      const int k = kt + e / BM;
      ra[q] = (m < M && k < K) ? (T)A[m + lda * k] : (T)0;
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int e = tid + 256 * q;
      const int k = kt + e % BK;
      const long n = n0 + e / BK;
      rb[q] = (n < N && k < K) ? (T)B[k + ldb * n] : (T)0;
    }
  };
  load(0);
  for (int kt = 0; kt < K; kt += BK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NA; ++q) {
      const int e = tid + 256 * q;
      As[e / BM][e % BM] = ra[q];
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int e = tid + 256 * q;
      Bs[e / BK][e % BK] = rb[q];
    }
    __syncthreads();
    if (kt + BK < K) load(kt + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T af[MI], bf[NI];
#pragma unroll
      for (int i = 0; i < MI; ++i) af[i] = As[kk + (lane >> 4)][wm + i * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < NI; ++i) bf[i] = Bs[wn + i * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = Mfma<T>::op(bf[ni], af[mi], acc[mi][ni]);
    }
  }
  // epilogue: lane holds C(m0+wm+16mi+lane%16, n0+wn+16ni+row(lane,r)).
  // Index splits (m -> i1,i2; n -> j,k) take one 32-bit division per lane and
  // then advance with wrap-around.
  int i1 = 0, i2 = 0;
  if (EPI == EPI_DIVIDE) {
    const int mb = (int)(m0 + wm + (lane & 15));
    i1 = mb % ep.n1;
    i2 = mb / ep.n1;
  }
#pragma unroll
  for (int mi = 0; mi < MI; ++mi) {
    const long m = m0 + wm + mi * 16 + (lane & 15);
    if (mi > 0 && EPI == EPI_DIVIDE) {
      i1 += 16;
      while (i1 >= ep.n1) { i1 -= ep.n1; ++i2; }
    }
    if (m >= M) continue;
    double lxy = 0.0, big_xy = 0.0;
    if (EPI == EPI_DIVIDE) {
      lxy = ep.lx[i1] + ep.ly[i2];
      big_xy = fmax(fabs(ep.lx[i1]), fabs(ep.ly[i2]));
    }
    long ubase = 0;
    int jb = 0, kb = 0;
    if (EPI == EPI_SCATTER || EPI == EPI_SCATTER_ADD) {
      ubase = m + 1;
      const int nb = (int)(n0 + wn);
      jb = nb % ep.n2;
      kb = nb / ep.n2;
    }
#pragma unroll
    for (int ni = 0; ni < NI; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int dn = ni * 16 + Mfma<T>::row(lane, r);
        const long n = n0 + wn + dn;
        if (n >= N) continue;
        const double v = (double)acc[mi][ni][r];
        if (EPI == EPI_STORE) {
          C[m + ldc * n] = (TC)v;
        } else if (EPI == EPI_DIVIDE) {
          const double lzn = ep.lz[n];
          const double big = fmax(fmax(fabs(ep.sigma), big_xy), fabs(lzn));
          const double den = lxy + lzn + ep.sigma;
          if (sizeof(TC) == 8)
            C[m + ldc * n] = (TC)((big < 1e-11) ? 0.0 : v / den);
          else  // FP32 inner solve: FP32 reciprocal
            C[m + ldc * n] = (TC)((big < 1e-11) ? 0.0f : (float)v * __builtin_amdgcn_rcpf((float)den));
        } else {
          int j = jb + dn, k = kb;
          while (j >= ep.n2) { j -= ep.n2; ++k; }
          double *p = ep.u + ubase + (long)ep.nx * ((j + 1) + (long)ep.ny * (k + 1));
          if (EPI == EPI_SCATTER)
            *p = v;
          else
            *p += v;
        }
      }
  }
}

}  // namespace mf
