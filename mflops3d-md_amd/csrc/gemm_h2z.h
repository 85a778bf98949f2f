// The z-forward transform, the eigenvalue divide and the z-backward transform
// of one fp16x2 inner solve in one kernel.
//
// The z transforms act along the last index: the field is (n1 n2) rows of n3
// values, F[r + n12 k].  A block owns 64 rows and, when n3 <= 256, every
// output z of both products, so the divided intermediate never leaves the
// block: it is split into two fp16 parts in LDS (scaled by the block's own
// largest magnitude, a power of two) and is the field operand of the
// backward product.  Against separate launches this saves one write and one
// read of the FP32 field and one launch.
//
//   C1 = (F W1^T) / (lx + ly + lz + sigma)   (first phase, the divide of
//                                              gemm_h2_kernel's EPI_DIVIDE)
//   C  = C1 W2^T                              (second phase, stored in FP32)
//
// Splitting, products, accumulator layout and scales are gemm_h2_kernel's.
#pragma once
#include <hip/hip_runtime.h>

#include "gemm_h2.h"

namespace mf {

constexpr int H2Z_BM = 64, H2Z_BN = 256;  // rows per block; outputs per block (all of z)

// W1split, W2split: fp16 [2][KP/8][WROWS][8] (scaled by wscale1, wscale2),
// WROWS == H2Z_BN, KP a multiple of 64 <= 256 (N <= KP).  grid: ceil(M / 64)
// blocks of 256 threads.
// fmax_in: the field's largest magnitude (float bits); fmax_out: receives
// the largest stored magnitude.
__global__ void __launch_bounds__(256, 2)
    gemm_h2z_kernel(const float *__restrict__ F, long ldf, const _Float16 *__restrict__ W1split,
                    const _Float16 *__restrict__ W2split, int KP, int WROWS,
                    float *__restrict__ C, long ldc, int M, int N, EpiArgs ep,
                    const unsigned *__restrict__ fmax_in, float wscale1, float wscale2,
                    unsigned *__restrict__ fmax_out) {
  constexpr int NT = 256, BM = H2Z_BM, BN = H2Z_BN;
  constexpr int BK = 32, PAD = 8;
  constexpr int KC = BN + PAD;            // row pitch of the intermediate (fp16 elements)
  constexpr int NG = BM * (BK / 8) / NT;  // groups of 8 field values per thread (1)
  static_assert(NG == 1, "tile");
  // phase 1 stages field k-steps (Fs), phase 2 holds the intermediate (Cs)
  constexpr int FS_BYTES = 2 * 2 * BM * (BK + PAD) * 2, CS_BYTES = 2 * BM * KC * 2;
  __shared__ __attribute__((aligned(16))) char smem[FS_BYTES > CS_BYTES ? FS_BYTES : CS_BYTES];
  __shared__ float bmax[NT / 64];
  auto Fs = reinterpret_cast<_Float16(*)[2][BM][BK + PAD]>(smem);  // [buf][part][row][k]
  auto Cs = reinterpret_cast<_Float16(*)[BM][KC]>(smem);           // [part][row][k]
  const int K = N;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long m0 = (long)blockIdx.x * BM;
  const int wn = wave * 64;  // this wave's first output z
  const float fs = h2_scale(__uint_as_float(*fmax_in));
  f4_t acc[4][4];
  auto zero_acc = [&] {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = (f4_t){0, 0, 0, 0};
  };
  zero_acc();
  // eigenvector fragments: lane l holds row wn + 16 q + l%16, k = 32 t + 8 (l/16) .. +7
  const unsigned wlane = ((unsigned)(lane >> 4) * WROWS + wn + (lane & 15)) * 16u;
  const unsigned wpart = (unsigned)(KP / 8) * WROWS * 16u, wstep = (unsigned)WROWS * 64u;
  auto loadW = [&](const _Float16 *Wsplit, int t, f16x8_t (&wf)[2][4]) {
    const char *Wb = reinterpret_cast<const char *>(Wsplit);
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        wf[p][q] = *reinterpret_cast<const f16x8_t *>(
            Wb + (wlane + p * wpart + (unsigned)t * wstep + q * 256u));
  };
  // three products per 16x16 tile, field fragments ff (D columns, rows m),
  // eigenvector fragments wf (D rows, outputs n)
  auto mfma3 = [&](const f16x8_t (&ff)[2][4], const f16x8_t (&wf)[2][4]) {
#pragma unroll
    for (int ni = 0; ni < 4; ++ni)
#pragma unroll
      for (int mi = 0; mi < 4; ++mi) {
        f4_t c = acc[mi][ni];
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(wf[1][ni], ff[0][mi], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(wf[0][ni], ff[1][mi], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(wf[0][ni], ff[0][mi], c, 0, 0, 0);
        acc[mi][ni] = c;
      }
  };
  const int ko = 8 * (lane >> 4);

  // ---- phase 1: C1 = F W1^T, field staged through LDS by k-steps
  // field element (row r, k): F[r + ldf k], lanes along r; columns k >= K
  // are read at K-1 and replaced by zero when split
  const int frow = tid % BM, fkq = (tid / BM) * 8;
  const long fr = m0 + frow < M ? m0 + frow : M - 1;
  const float *Fr = F + fr;
  float freg[8];
  auto loadF = [&](int t) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k0 = t * BK + fkq + q;
      freg[q] = Fr[ldf * (k0 < K ? k0 : K - 1)];
    }
  };
  auto storeF = [&](int t, int buf) {
    f16x8_t v0, v1;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const float x = (t * BK + fkq + q < K) ? freg[q] * fs : 0.0f;
      _Float16 a, b;
      split2h(x, a, b);
      v0[q] = a;
      v1[q] = b;
    }
    *reinterpret_cast<f16x8_t *>(&Fs[buf][0][frow][fkq]) = v0;
    *reinterpret_cast<f16x8_t *>(&Fs[buf][1][frow][fkq]) = v1;
  };
  auto stepF = [&](int buf, const f16x8_t (&wf)[2][4]) {
    f16x8_t ff[2][4];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        ff[p][i] = *reinterpret_cast<const f16x8_t *>(&Fs[buf][p][16 * i + (lane & 15)][ko]);
    mfma3(ff, wf);
  };
  const int nk2 = KP / BK;  // even: KP is a multiple of 64
  f16x8_t wA[2][4], wB[2][4];
  loadF(0);
  loadW(W1split, 0, wA);
  for (int t = 0; t < nk2; t += 2) {
    storeF(t, 0);
    loadF(t + 1);
    __syncthreads();
    loadW(W1split, t + 1, wB);
    stepF(0, wA);
    storeF(t + 1, 1);
    loadF(min(t + 2, nk2 - 1));
    __syncthreads();
    loadW(W1split, min(t + 2, nk2 - 1), wA);
    stepF(1, wB);
  }
  // divide (unscaled by 1 / (fs wscale1)), block maximum
  {
    const float unscale = 1.0f / (fs * wscale1);
    float lzf[4][4];
    bool nullz[4][4];
#pragma unroll
    for (int ni = 0; ni < 4; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = wn + 4 * (lane >> 4) + ni * 16 + r;
        const double lz = ep.lz[n < N ? n : N - 1];
        lzf[ni][r] = (float)lz;
        nullz[ni][r] = fabs(lz) < 1e-11;
      }
    float vmax = 0.0f;
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
      const long m = m0 + mi * 16 + (lane & 15);
      const long mc = m < M ? m : M - 1;
      const int i1 = (int)(mc % ep.n1), i2 = (int)(mc / ep.n1);
      const float lxy = (float)(ep.lx[i1] + ep.ly[i2] + ep.sigma);
      const bool null_xy = fmax(fmax(fabs(ep.lx[i1]), fabs(ep.ly[i2])), fabs(ep.sigma)) < 1e-11;
#pragma unroll
      for (int ni = 0; ni < 4; ++ni)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = wn + 4 * (lane >> 4) + ni * 16 + r;
          const float inv = __builtin_amdgcn_rcpf(lxy + lzf[ni][r]);
          const float v = acc[mi][ni][r] * unscale;
          const float o = (m >= M || n >= N || (null_xy && nullz[ni][r])) ? 0.0f : v * inv;
          acc[mi][ni][r] = o;
          vmax = fmaxf(vmax, fabsf(o));
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off, 64));
    if (lane == 0) bmax[wave] = vmax;
  }
  __syncthreads();  // bmax complete; every wave is done with Fs
  float b = bmax[0];
#pragma unroll
  for (int q = 1; q < NT / 64; ++q) b = fmaxf(b, bmax[q]);
  const float fs2 = h2_scale(b);
  // the intermediate, split, as the field operand: Cs[part][row m][k = n]
  // (a lane's four r values are consecutive n: one 8-byte write per part)
  typedef _Float16 f16x4_t __attribute__((ext_vector_type(4)));
#pragma unroll
  for (int mi = 0; mi < 4; ++mi)
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
      const int ml = mi * 16 + (lane & 15), n = wn + 4 * (lane >> 4) + ni * 16;
      f16x4_t p0, p1;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        _Float16 a0, a1;
        split2h(acc[mi][ni][r] * fs2, a0, a1);
        p0[r] = a0;
        p1[r] = a1;
      }
      *reinterpret_cast<f16x4_t *>(&Cs[0][ml][n]) = p0;
      *reinterpret_cast<f16x4_t *>(&Cs[1][ml][n]) = p1;
    }
  __syncthreads();

  // ---- phase 2: C = C1 W2^T, field operand from LDS, no barriers
  zero_acc();
  loadW(W2split, 0, wA);
  for (int t = 0; t < nk2; t += 2) {
    f16x8_t ff[2][4];
    loadW(W2split, t + 1, wB);
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        ff[p][i] = *reinterpret_cast<const f16x8_t *>(&Cs[p][16 * i + (lane & 15)][t * BK + ko]);
    mfma3(ff, wA);
    loadW(W2split, min(t + 2, nk2 - 1), wA);
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        ff[p][i] = *reinterpret_cast<const f16x8_t *>(&Cs[p][16 * i + (lane & 15)][(t + 1) * BK + ko]);
    mfma3(ff, wB);
  }
  // ---- store (unscaled by 1 / (fs2 wscale2)), largest magnitude
  const float unscale2 = 1.0f / (fs2 * wscale2);
  float vmax = 0.0f;
#pragma unroll
  for (int mi = 0; mi < 4; ++mi) {
    const long m = m0 + mi * 16 + (lane & 15);
    if (m >= M) continue;
#pragma unroll
    for (int ni = 0; ni < 4; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = wn + 4 * (lane >> 4) + ni * 16 + r;
        if (n >= N) continue;
        const float v = acc[mi][ni][r] * unscale2;
        C[m + ldc * n] = v;
        vmax = fmaxf(vmax, fabsf(v));
      }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off, 64));
  __syncthreads();  // the reads of bmax above are done
  if (lane == 0) bmax[wave] = vmax;
  __syncthreads();
  if (tid == 0) {
    float c = bmax[0];
#pragma unroll
    for (int q = 1; q < NT / 64; ++q) c = fmaxf(c, bmax[q]);
    atomicMax(fmax_out, __float_as_uint(c));
  }
}

}  // namespace mf
