// Eigenvector transforms with FP32-class accuracy on the BF16 matrix cores.
//
// Every operand value x (FP32) is split into three bf16 parts
//   x = x0 + x1 + x2,  x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1)
// (24 significant bits in all), and a product keeps the six terms whose
// part orders sum to <= 2:  a0b0 + (a0b1 + a1b0) + (a0b2 + a1b1 + a2b0),
// each a v_mfma_f32_16x16x32_bf16 accumulated in FP32.  The dropped terms are
// below 2^-24 |a||b|, the same order as FP32 rounding, at 6/16 of the cost of
// the FP32 MFMA path on gfx950 (bf16 16x16x32 takes 16 cycles, f32 16x16x4
// takes 32 for 1/8 of the work).
//
// The eigenvector operand is split once on the host (k padded to a multiple
// of 64); the field operand is split while it is staged to LDS.  Tiles, wave
// layout, operand swap (lanes contiguous in m for the column-major stores)
// and epilogues are those of gemm_mode_kernel.
#pragma once
#include <hip/hip_runtime.h>

#include "gemm.h"

namespace mf {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

__device__ __forceinline__ void split3(float x, __bf16 &p0, __bf16 &p1, __bf16 &p2) {
  p0 = (__bf16)x;
  const float r1 = x - (float)p0;
  p1 = (__bf16)r1;
  const float r2 = r1 - (float)p1;
  p2 = (__bf16)r2;
}

// WLEFT: A is the eigenvector matrix (x-transform), B the field; otherwise
// A is the field and B the eigenvector matrix (y/z transforms).
// Wsplit: bf16 [3][KP/8][WROWS][8] (part, 8-column group of k, row, k within
// the group), zero beyond the matrix, KP a multiple of 64 and WROWS a
// multiple of the tile side along the eigenvector dimension, so its loads
// need no guards.  A wave's eigenvector fragments (its 64 rows, three parts)
// are loaded straight from global memory (L2) into registers in MFMA operand
// layout, one k-step ahead; only the field operand goes through LDS, split
// into its three parts while staged, in two buffers so that one barrier per
// k-step suffices.  Every load is unconditional (field addresses clamped
// into the array, 32-bit offsets from the kernel-argument base), so the
// compiler waits for exactly the loads each instruction needs.
// The block has BM*BN/4096 wavefronts, each computing a 64 x 64 tile.
template <typename TF, typename TC, int BM, int BN, bool WLEFT, int EPI>
__global__ void __launch_bounds__(256, 2)
    gemm_bf3_kernel(const TF *__restrict__ F, long ldf, long sF, const __bf16 *__restrict__ Wsplit,
                    int KP, int WROWS, TC *__restrict__ C, long ldc, long sC, int M, int N, int K,
                    EpiArgs ep) {
  constexpr int NT = 256;
  constexpr int BK = 32;
  constexpr int PAD = 8;                  // bf16 elements of row padding (16 B)
  constexpr int RF = WLEFT ? BN : BM;     // field rows per tile
  constexpr int NG = RF * (BK / 8) / NT;  // groups of 8 field values per thread
  static_assert(BM * BN == NT * 64 && NG * NT == RF * (BK / 8), "tile");
  __shared__ __attribute__((aligned(16))) __bf16 Fs[2][3][RF][BK + PAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long m0 = (long)blockIdx.x * BM, n0 = (long)blockIdx.y * BN;
  F += (long)blockIdx.z * sF;
  if (EPI == EPI_STORE || EPI == EPI_DIVIDE) C += (long)blockIdx.z * sC;
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
  auto loadW = [&](int t, bf16x8_t (&wf)[3][4]) {
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        wf[p][q] = *reinterpret_cast<const bf16x8_t *>(
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
      bf16x8_t v0, v1, v2;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const float x = (t * BK + fkq(g) + q < K) ? (float)freg[g][q] : 0.0f;
        __bf16 a, b, c;
        split3(x, a, b, c);
        v0[q] = a; v1[q] = b; v2[q] = c;
      }
      const int r = frow(g), kq = fkq(g);
      *reinterpret_cast<bf16x8_t *>(&Fs[buf][0][r][kq]) = v0;
      *reinterpret_cast<bf16x8_t *>(&Fs[buf][1][r][kq]) = v1;
      *reinterpret_cast<bf16x8_t *>(&Fs[buf][2][r][kq]) = v2;
    }
  };
  // ---- six products per 16x16 tile; lane l holds k = 8(l/16)..+7.
  // D columns (m) come from the A operand, D rows (n) from the B operand.
  const int ko = 8 * (lane >> 4);
  const int fro = WLEFT ? wn : wm;  // this wave's first field row in the tile
  auto step = [&](int buf, const bf16x8_t (&wf)[3][4]) {
    bf16x8_t ff[3][4];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        ff[p][i] = *reinterpret_cast<const bf16x8_t *>(&Fs[buf][p][fro + 16 * i + (lane & 15)][ko]);
#pragma unroll
    for (int ni = 0; ni < 4; ++ni)
#pragma unroll
      for (int mi = 0; mi < 4; ++mi) {
        // parts sit 4 apart in both fragment arrays
        const bf16x8_t *fa = WLEFT ? &wf[0][mi] : &ff[0][mi];
        const bf16x8_t *fb = WLEFT ? &ff[0][ni] : &wf[0][ni];
        f4_t c = acc[mi][ni];
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[8], fa[0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[4], fa[4], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[0], fa[8], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[4], fa[0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[0], fa[4], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[0], fa[0], c, 0, 0, 0);
        acc[mi][ni] = c;
      }
  };
  // An even number of k-steps (KP is a multiple of 64; the steps past K
  // multiply zeros); the prefetch after the last step re-reads it.
  const int nk2 = ((K + 2 * BK - 1) / (2 * BK)) * 2;
  bf16x8_t wA[3][4], wB[3][4];
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
  // ---- epilogue (as gemm_mode_kernel; f32 output layout).  The divide uses
  // the FP32 reciprocal of an FP32 eigenvalue sum (lx+ly+sigma rounded once
  // per row, plus lz): the result is an FP32 value of an approximate inner
  // solve.  The null-mode test stays on the FP64 values.  Index splits
  // advance from one 32-bit division per lane.
  int i1 = 0, i2 = 0;
  if (EPI == EPI_DIVIDE) {
    const int mb = (int)(m0 + wm + (lane & 15));
    i1 = mb % ep.n1;
    i2 = mb / ep.n1;
  }
  int jb = 0, kb = 0;
  if (EPI == EPI_SCATTER || EPI == EPI_SCATTER_ADD) {
    const int nb = (int)(n0 + wn + 4 * (lane >> 4));
    jb = nb % ep.n2;
    kb = nb / ep.n2;
  }
  // this lane's 16 columns of lz (FP32) and their null-mode flags
  float lzf[4][4];
  bool nullz[4][4];
  if (EPI == EPI_DIVIDE) {
#pragma unroll
    for (int ni = 0; ni < 4; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long n = n0 + wn + 4 * (lane >> 4) + ni * 16 + r;
        const double lz = ep.lz[n < N ? n : N - 1];
        lzf[ni][r] = (float)lz;
        nullz[ni][r] = fabs(lz) < 1e-11;
      }
  }
  auto uidx = [&](long m, int dn) {
    int j = jb + dn, k = kb;
    while (j >= ep.n2) { j -= ep.n2; ++k; }
    return (m + 1) + (long)ep.nx * ((j + 1) + (long)ep.ny * (k + 1));
  };
#pragma unroll
  for (int mi = 0; mi < 4; ++mi) {
    const long m = m0 + wm + mi * 16 + (lane & 15);
    if (mi > 0 && EPI == EPI_DIVIDE) {
      i1 += 16;
      while (i1 >= ep.n1) { i1 -= ep.n1; ++i2; }
    }
    if (m >= M) continue;
    float lxy = 0.0f;
    bool null_xy = false;
    if (EPI == EPI_DIVIDE) {
      lxy = (float)(ep.lx[i1] + ep.ly[i2] + ep.sigma);
      null_xy = fmax(fmax(fabs(ep.lx[i1]), fabs(ep.ly[i2])), fabs(ep.sigma)) < 1e-11;
    }
    if (EPI == EPI_SCATTER_ADD) {
      // all loads of this row first, then the stores
      double old[4][4];
#pragma unroll
      for (int ni = 0; ni < 4; ++ni)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int dn = ni * 16 + r;
          const bool ok = n0 + wn + 4 * (lane >> 4) + dn < N;
          old[ni][r] = ok ? ep.u[uidx(m, dn)] : 0.0;
        }
#pragma unroll
      for (int ni = 0; ni < 4; ++ni)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int dn = ni * 16 + r;
          if (n0 + wn + 4 * (lane >> 4) + dn >= N) continue;
          ep.u[uidx(m, dn)] = old[ni][r] + (double)acc[mi][ni][r];
        }
      continue;
    }
#pragma unroll
    for (int ni = 0; ni < 4; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int dn = ni * 16 + r;
        const long n = n0 + wn + 4 * (lane >> 4) + dn;
        if (n >= N) continue;
        const float v = acc[mi][ni][r];
        if (EPI == EPI_STORE) {
          C[m + ldc * n] = (TC)v;
        } else if (EPI == EPI_DIVIDE) {
          const float inv = __builtin_amdgcn_rcpf(lxy + lzf[ni][r]);
          C[m + ldc * n] = (TC)((null_xy && nullz[ni][r]) ? 0.0f : v * inv);
        } else {
          ep.u[uidx(m, dn)] = (double)v;
        }
      }
  }
}

}  // namespace mf
