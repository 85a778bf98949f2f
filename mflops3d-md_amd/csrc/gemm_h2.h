// Eigenvector transforms with near-FP32 accuracy on the FP16 matrix cores.
//
// Operands are scaled by powers of two into FP16's range and split into two
// FP16 parts, x s = x0 + x1 with x0 = fp16(x s), x1 = fp16(x s - x0): 22
// significant bits.  A product keeps the three terms a0b0 + a0b1 + a1b0,
// each a v_mfma_f32_16x16x32_f16 accumulated in FP32 (products of FP16 values
// are exact in FP32); the dropped a1b1 is below 2^-22 |a||b|.  Half the
// matrix-core work of the three-part bf16 form (three products instead of
// six) and two thirds of its operand traffic, for two more bits of error per
// transform, which the FP64-residual refinement absorbs.
//
// Scales: the eigenvector matrix carries a fixed power of two (its largest
// entry lands in [2^13, 2^14)); the field operand is scaled from the largest
// magnitude of the field, a device scalar (float bits, written with atomic
// max by the kernel that produced the field), so that its largest entry lands
// in [2^13, 2^14): no overflow, and entries far below the largest keep an
// absolute error of ~2^-24 of it.  The epilogue multiplies by the exact
// inverse of both scales and, when asked, folds the largest magnitude of
// what it stores into the next field's scalar.
//
// Layout, tiles, wave arrangement, operand swap and epilogues are those of
// gemm_bf3_kernel.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "gemm.h"

namespace mf {

typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));

// 2^(13 - floor(log2 m)) for a finite m > 0, else 1: m s lands in [2^13, 2^14)
__device__ __forceinline__ float h2_scale(float m) {
  const unsigned b = __float_as_uint(m);
  const int e = (int)((b >> 23) & 0xffu);  // biased exponent
  if (!(m > 0.0f) || e == 0xff || e == 0) return 1.0f;
  const int se = 127 + 13 - (e - 127);
  if (se <= 0 || se >= 255) return 1.0f;
  return __uint_as_float((unsigned)se << 23);
}

__device__ __forceinline__ void split2h(float xs, _Float16 &p0, _Float16 &p1) {
  p0 = (_Float16)xs;
  p1 = (_Float16)(xs - (float)p0);
}

// Epilogue of the fp16x2 transforms (as gemm_bf3_kernel's): acc[mi][ni][r]
// holds D(m, n), m = m0 + wm + 16 mi + lane % 16, n = n0 + wn + 4 (lane / 16)
// + 16 ni + r, scaled by 1 / unscale (an exact power of two).  EPI_STORE and
// EPI_DIVIDE store FP32 (divide: by lx + ly + lz + sigma, 0 at the null
// mode) and fold the block's largest stored magnitude into fmax_out;
// EPI_SCATTER / EPI_SCATTER_ADD write or add into the interior of u (FP64).
// amax receives the largest |acc| stored (before unscale).
//
// Addressing: one 64-bit base per lane, then offsets that are wave-uniform
// per (mi, ni, r) (scalar arithmetic); bound checks only in waves whose 64 x
// 64 tile crosses M or N (a wave-uniform test picks the checked body).
template <typename TC, int BM, int BN, int EPI>
__device__ __forceinline__ void h2_store(f4_t (&acc)[4][4], float unscale, long m0, long n0,
                                         int wm, int wn, int lane, TC *__restrict__ C, long ldc,
                                         int M, int N, const EpiArgs &ep, float &amax) {
  const int lm = lane & 15, ln = 4 * (lane >> 4);
  const long mw = m0 + wm, nw = n0 + wn;  // wave-uniform
  const bool full = mw + 64 <= M && nw + 64 <= N;
  if (EPI == EPI_STORE || EPI == EPI_DIVIDE) {
    TC *pl = C + (mw + lm) + ldc * (nw + ln);
    float lzf[4][4];
    bool nullz[4][4];
    if (EPI == EPI_DIVIDE) {
#pragma unroll
      for (int ni = 0; ni < 4; ++ni)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const long n = nw + ln + ni * 16 + r;
          const double lz = ep.lz[n < N ? n : N - 1];
          lzf[ni][r] = (float)lz;
          nullz[ni][r] = fabs(lz) < 1e-11;
        }
    }
    auto body = [&](auto checked) {
#pragma unroll
      for (int mi = 0; mi < 4; ++mi) {
        const long m = mw + 16 * mi + lm;
        if (decltype(checked)::value && m >= M) continue;
        float lxy = 0.0f;
        bool null_xy = false;
        if (EPI == EPI_DIVIDE) {
          const int i1 = (int)(m % ep.n1), i2 = (int)(m / ep.n1);
          lxy = (float)(ep.lx[i1] + ep.ly[i2] + ep.sigma);
          null_xy = fmax(fmax(fabs(ep.lx[i1]), fabs(ep.ly[i2])), fabs(ep.sigma)) < 1e-11;
        }
#pragma unroll
        for (int ni = 0; ni < 4; ++ni)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            if (decltype(checked)::value && nw + ln + 16 * ni + r >= N) continue;
            float v = acc[mi][ni][r];
            if (EPI == EPI_DIVIDE)
              v = (null_xy && nullz[ni][r]) ? 0.0f : v * __builtin_amdgcn_rcpf(lxy + lzf[ni][r]);
            amax = fmaxf(amax, fabsf(v));
            pl[16 * mi + ldc * (16 * ni + r)] = (TC)(v * unscale);
          }
      }
    };
    if (full) body(std::false_type{});
    else body(std::true_type{});
  } else {
    // u[(i+1) + nx((j+1) + ny(k+1))] for m = i, n = j + n2 k.  With nb = n0 +
    // wn + 4 (lane / 16) = jb + n2 kb and dn = 16 ni + r < 64 <= n2, n = nb +
    // dn wraps at most once: the offset is nx dn past the lane's base, plus 2
    // nx (the two boundary rows of the next plane) when jb + dn >= n2.
    const long nb = nw + ln;
    const int jb = (int)(nb % ep.n2), kb = (int)(nb / ep.n2);
    const long nx = ep.nx, nxy = (long)ep.nx * ep.ny;
    double *ul = ep.u + (mw + lm + 1) + nx * (jb + 1) + nxy * (kb + 1);
    auto body = [&](auto checked) {
#pragma unroll
      for (int ni = 0; ni < 4; ++ni)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int dn = 16 * ni + r;
          if (decltype(checked)::value && nb + dn >= N) continue;
          long off;
          if (ep.n2 >= 64) {
            off = nx * dn + (jb + dn >= ep.n2 ? 2 * nx : 0);
          } else {
            const int q = (jb + dn) / ep.n2;  // planes crossed
            off = nx * dn + 2 * nx * q;
          }
#pragma unroll
          for (int mi = 0; mi < 4; ++mi) {
            if (decltype(checked)::value && mw + 16 * mi + lm >= M) continue;
            double *pu = ul + 16 * mi + off;
            const double v = (double)(acc[mi][ni][r] * unscale);
            if (EPI == EPI_SCATTER_ADD) *pu = *pu + v;
            else *pu = v;
          }
        }
    };
    if (full) body(std::false_type{});
    else body(std::true_type{});
  }
}

// h2_store, then the block's largest stored magnitude into fmax_out (one
// atomic per block).  Every thread of the block must call it.
template <typename TC, int BM, int BN, int EPI>
__device__ __forceinline__ void h2_epilogue(f4_t (&acc)[4][4], float unscale, long m0, long n0,
                                            int wm, int wn, int lane, int wave, int tid,
                                            TC *__restrict__ C, long ldc, int M, int N,
                                            const EpiArgs &ep, unsigned *__restrict__ fmax_out) {
  constexpr int NT = 256;
  float amax = 0.0f;  // largest |acc| stored (before unscale)
  h2_store<TC, BM, BN, EPI>(acc, unscale, m0, n0, wm, wn, lane, C, ldc, M, N, ep, amax);
  if ((EPI == EPI_STORE || EPI == EPI_DIVIDE) && fmax_out) {
    // block maximum, then one atomic per block
    __shared__ float bmax[NT / 64];
    float vmax = amax * unscale;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off, 64));
    if (lane == 0) bmax[wave] = vmax;
    __syncthreads();
    if (tid == 0) {
      float b = bmax[0];
#pragma unroll
      for (int q = 1; q < NT / 64; ++q) b = fmaxf(b, bmax[q]);
      atomicMax(fmax_out, __float_as_uint(b));
    }
  }
}

// Wsplit: fp16 [2][KP/8][WROWS][8] of the scaled matrix (wscale applied).
// fmax_in: the field's largest magnitude (float bits); fmax_out: NULL or the
// scalar receiving the largest stored magnitude (EPI_STORE / EPI_DIVIDE).
// zero_slot: NULL or a scalar that block 0 sets to zero (the caller's next
// producer starts from it).
template <typename TF, typename TC, int BM, int BN, bool WLEFT, int EPI>
__global__ void __launch_bounds__(256, 2)
    gemm_h2_kernel(const TF *__restrict__ F, long ldf, long sF, const _Float16 *__restrict__ Wsplit,
                   int KP, int WROWS, TC *__restrict__ C, long ldc, long sC, int M, int N, int K,
                   EpiArgs ep, const unsigned *__restrict__ fmax_in, float wscale,
                   unsigned *__restrict__ fmax_out, unsigned *__restrict__ zero_slot) {
  constexpr int NT = 256;
  constexpr int BK = 32;
  constexpr int PAD = 8;                  // fp16 elements of row padding (16 B)
  constexpr int RF = WLEFT ? BN : BM;     // field rows per tile
  constexpr int NG = RF * (BK / 8) / NT;  // groups of 8 field values per thread
  static_assert(BM * BN == NT * 64 && NG * NT == RF * (BK / 8), "tile");
  __shared__ __attribute__((aligned(16))) _Float16 Fs[2][2][RF][BK + PAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const long m0 = (long)blockIdx.x * BM, n0 = (long)blockIdx.y * BN;
  F += (long)blockIdx.z * sF;
  if (EPI == EPI_STORE || EPI == EPI_DIVIDE) C += (long)blockIdx.z * sC;
  if (zero_slot && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && tid == 0) *zero_slot = 0u;
  const float fs = h2_scale(__uint_as_float(*fmax_in));
  constexpr int WM = BM / 64;
  const int wm = (wave % WM) * 64, wn = (wave / WM) * 64;
  const int wr0 = (int)(WLEFT ? m0 + wm : n0 + wn);  // first eigenvector row of this wave
  const long f0 = WLEFT ? n0 : m0;                   // first field row of the tile
  const long FR = WLEFT ? N : M;                     // field rows
  f4_t acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = (f4_t){0, 0, 0, 0};

  // eigenvector fragments (byte offsets): part p, 16-row tile q of k-step t:
  // lane l holds row wr0 + 16 q + l%16, k = 32 t + 8 (l/16) .. +7
  const char *Wb = reinterpret_cast<const char *>(Wsplit);
  const unsigned wlane = ((unsigned)(lane >> 4) * WROWS + wr0 + (lane & 15)) * 16u;
  const unsigned wpart = (unsigned)(KP / 8) * WROWS * 16u, wstep = (unsigned)WROWS * 64u;
  auto loadW = [&](int t, f16x8_t (&wf)[2][4]) {
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        wf[p][q] = *reinterpret_cast<const f16x8_t *>(
            Wb + (wlane + p * wpart + (unsigned)t * wstep + q * 256u));
  };
  // field element (row r, column k) of the tile: WLEFT -> F[k + ldf r]
  // (k contiguous, 4 threads per row); else F[r + ldf k] (lanes along r).
  // Columns k >= K are read at K-1 and replaced by zero when split.
  auto frow = [&](int g) { return WLEFT ? (tid + NT * g) / (BK / 8) : (tid + NT * g) % RF; };
  auto fkq = [&](int g) { return WLEFT ? ((tid + NT * g) % (BK / 8)) * 8 : ((tid + NT * g) / RF) * 8; };
  const char *Fb = reinterpret_cast<const char *>(F);
  const unsigned ldfu = (unsigned)ldf;
  unsigned frb[NG];
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    const long r0 = f0 + frow(g);
    const unsigned r = (unsigned)(r0 < FR ? r0 : FR - 1);
    frb[g] = WLEFT ? ldfu * r : r;
  }
  TF freg[NG][8];
  auto loadF = [&](int t) {
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int k0 = t * BK + fkq(g) + q;
        const unsigned k = (unsigned)(k0 < K ? k0 : K - 1);
        const unsigned e = WLEFT ? frb[g] + k : frb[g] + ldfu * k;
        freg[g][q] = *reinterpret_cast<const TF *>(Fb + e * (unsigned)sizeof(TF));
      }
  };
  auto storeF = [&](int t, int buf) {
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      f16x8_t v0, v1;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const float x = (t * BK + fkq(g) + q < K) ? (float)freg[g][q] * fs : 0.0f;
        _Float16 a, b;
        split2h(x, a, b);
        v0[q] = a; v1[q] = b;
      }
      const int r = frow(g), kq = fkq(g);
      *reinterpret_cast<f16x8_t *>(&Fs[buf][0][r][kq]) = v0;
      *reinterpret_cast<f16x8_t *>(&Fs[buf][1][r][kq]) = v1;
    }
  };
  // ---- three products per 16x16 tile; lane l holds k = 8(l/16)..+7.
  // D columns (m) come from the A operand, D rows (n) from the B operand.
  const int ko = 8 * (lane >> 4);
  const int fro = WLEFT ? wn : wm;  // this wave's first field row in the tile
  auto step = [&](int buf, const f16x8_t (&wf)[2][4]) {
    f16x8_t ff[2][4];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        ff[p][i] = *reinterpret_cast<const f16x8_t *>(&Fs[buf][p][fro + 16 * i + (lane & 15)][ko]);
#pragma unroll
    for (int ni = 0; ni < 4; ++ni)
#pragma unroll
      for (int mi = 0; mi < 4; ++mi) {
        // parts sit 4 apart in both fragment arrays
        const f16x8_t *fa = WLEFT ? &wf[0][mi] : &ff[0][mi];
        const f16x8_t *fb = WLEFT ? &ff[0][ni] : &wf[0][ni];
        f4_t c = acc[mi][ni];
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(fb[4], fa[0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(fb[0], fa[4], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(fb[0], fa[0], c, 0, 0, 0);
        acc[mi][ni] = c;
      }
  };
  // An even number of k-steps (KP is a multiple of 64; the steps past K
  // multiply zeros); the prefetch after the last step re-reads it.
  const int nk2 = ((K + 2 * BK - 1) / (2 * BK)) * 2;
  f16x8_t wA[2][4], wB[2][4];
  loadF(0);
  loadW(0, wA);
  for (int t = 0; t < nk2; t += 2) {
    storeF(t, 0);
    loadF(t + 1);
    __syncthreads();
    loadW(t + 1, wB);
    step(0, wA);
    storeF(t + 1, 1);
    loadF(min(t + 2, nk2 - 1));
    __syncthreads();
    loadW(min(t + 2, nk2 - 1), wA);
    step(1, wB);
  }
  h2_epilogue<TC, BM, BN, EPI>(acc, 1.0f / (fs * wscale), m0, n0, wm, wn, lane, wave, tid, C, ldc,
                               M, N, ep, fmax_out);
}

}  // namespace mf
