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
// The eigenvector operand is split once on the host and stored as
// [part][row][k] (k contiguous, padded to a multiple of 32); the field
// operand is split while it is staged to LDS.  Tiles, wave layout, operand
// swap (lanes contiguous in m for the column-major stores) and epilogues are
// those of gemm_mode_kernel.
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
// the group), zero beyond the matrix, with WROWS a multiple of the tile side
// along the eigenvector dimension, so its tile loads need no guards; lanes
// read consecutive rows (coalesced) and write them to LDS rows 80 bytes apart
// (bank-conflict free).  Loads are branch-free (clamped addresses, k >= K selected to 0)
// and the next k-step's tile is fetched into registers while the current one
// is multiplied.
// The block has BM*BN/4096 wavefronts, each computing a 64 x 64 tile.
template <typename TF, typename TC, int BM, int BN, bool WLEFT, int EPI>
__global__ void __launch_bounds__(BM * BN / 64)
    gemm_bf3_kernel(const TF *__restrict__ F, long ldf, long sF, const __bf16 *__restrict__ Wsplit,
                    int KP, int WROWS, TC *__restrict__ C, long ldc, long sC, int M, int N, int K,
                    EpiArgs ep) {
  constexpr int NT = BM * BN / 64;    // threads
  constexpr int BK = 32;
  constexpr int PAD = 8;  // bf16 elements of row padding (16 B)
  constexpr int RW = WLEFT ? BM : BN;  // eigenvector rows per tile
  constexpr int RF = WLEFT ? BN : BM;  // field rows (m or n) per tile
  constexpr int NW = 3 * RW * (BK / 8) / NT;  // 16-byte eigenvector chunks per thread
  constexpr int NG = RF * (BK / 8) / NT;      // groups of 8 field values per thread
  static_assert(NW * NT == 3 * RW * (BK / 8) && NG * NT == RF * (BK / 8), "tile");
  __shared__ __attribute__((aligned(16))) __bf16 Ws[3][RW][BK + PAD];
  __shared__ __attribute__((aligned(16))) __bf16 Fs[3][RF][BK + PAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long m0 = (long)blockIdx.x * BM, n0 = (long)blockIdx.y * BN;
  F += (long)blockIdx.z * sF;
  if (EPI == EPI_STORE || EPI == EPI_DIVIDE) C += (long)blockIdx.z * sC;
  constexpr int WM = BM / 64;
  const int wm = (wave % WM) * 64, wn = (wave / WM) * 64;
  const long w0 = WLEFT ? m0 : n0;   // first eigenvector row of the tile
  const long f0 = WLEFT ? n0 : m0;   // first field row of the tile
  const long FR = WLEFT ? N : M;     // field rows
  f4_t acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = (f4_t){0, 0, 0, 0};

  bf16x8_t wreg[NW];
  TF freg[NG][8];
  // field element (row r, column k) of the tile: WLEFT -> F[k + ldf r]
  // (k contiguous, 4 threads per row); else F[r + ldf k] (lanes along r)
  auto frow = [&](int g) { return WLEFT ? (tid + NT * g) / (BK / 8) : (tid + NT * g) % RF; };
  auto fkq = [&](int g) { return WLEFT ? ((tid + NT * g) % (BK / 8)) * 8 : ((tid + NT * g) / RF) * 8; };
  auto load = [&](int kt) {
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      const int e = tid + NT * i;
      const int r = e % RW, c = (e / RW) % (BK / 8), p = e / ((BK / 8) * RW);
      wreg[i] = *reinterpret_cast<const bf16x8_t *>(
          Wsplit + (((size_t)p * (KP / 8) + kt / 8 + c) * WROWS + w0 + r) * 8);
    }
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      const long r0 = f0 + frow(g);
      const long r = r0 < FR ? r0 : FR - 1;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int k0 = kt + fkq(g) + q;
        const long k = k0 < K ? k0 : K - 1;
        freg[g][q] = WLEFT ? F[k + ldf * r] : F[r + ldf * k];
      }
    }
  };
  load(0);
  for (int kt = 0; kt < K; kt += BK) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      const int e = tid + NT * i;
      const int r = e % RW, c = (e / RW) % (BK / 8), p = e / ((BK / 8) * RW);
      *reinterpret_cast<bf16x8_t *>(&Ws[p][r][8 * c]) = wreg[i];
    }
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      bf16x8_t v0, v1, v2;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const float x = (kt + fkq(g) + q < K) ? (float)freg[g][q] : 0.0f;
        __bf16 a, b, c;
        split3(x, a, b, c);
        v0[q] = a; v1[q] = b; v2[q] = c;
      }
      const int r = frow(g), kq = fkq(g);
      *reinterpret_cast<bf16x8_t *>(&Fs[0][r][kq]) = v0;
      *reinterpret_cast<bf16x8_t *>(&Fs[1][r][kq]) = v1;
      *reinterpret_cast<bf16x8_t *>(&Fs[2][r][kq]) = v2;
    }
    __syncthreads();
    if (kt + BK < K) load(kt + BK);
    // ---- six products per 16x16 tile; lane l holds k = 8(l/16)..+7
    // D columns (m) come from the A tile, D rows (n) from the B tile
    const int ko = 8 * (lane >> 4);
    auto afrag = [&](int p, int row) {
      if constexpr (WLEFT) return *reinterpret_cast<const bf16x8_t *>(&Ws[p][row][ko]);
      else return *reinterpret_cast<const bf16x8_t *>(&Fs[p][row][ko]);
    };
    auto bfrag = [&](int p, int row) {
      if constexpr (WLEFT) return *reinterpret_cast<const bf16x8_t *>(&Fs[p][row][ko]);
      else return *reinterpret_cast<const bf16x8_t *>(&Ws[p][row][ko]);
    };
    bf16x8_t fa[3][4];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int i = 0; i < 4; ++i) fa[p][i] = afrag(p, wm + 16 * i + (lane & 15));
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) {
      bf16x8_t fb[3];
#pragma unroll
      for (int p = 0; p < 3; ++p) fb[p] = bfrag(p, wn + 16 * ni + (lane & 15));
#pragma unroll
      for (int mi = 0; mi < 4; ++mi) {
        f4_t c = acc[mi][ni];
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[2], fa[0][mi], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[1], fa[1][mi], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[0], fa[2][mi], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[1], fa[0][mi], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[0], fa[1][mi], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[0], fa[0][mi], c, 0, 0, 0);
        acc[mi][ni] = c;
      }
    }
  }
  // ---- epilogue (as gemm_mode_kernel; f32 output layout).  The divide uses
  // an FP32 reciprocal of the FP64 eigenvalue sum: the result is an FP32
  // value of an approximate inner solve.  Index splits advance from one
  // 32-bit division per lane.
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
    double lxy = 0.0, big_xy = 0.0;
    if (EPI == EPI_DIVIDE) {
      lxy = ep.lx[i1] + ep.ly[i2] + ep.sigma;
      big_xy = fmax(fmax(fabs(ep.lx[i1]), fabs(ep.ly[i2])), fabs(ep.sigma));
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
          const double lzn = ep.lz[n];
          const double big = fmax(big_xy, fabs(lzn));
          const float inv = __builtin_amdgcn_rcpf((float)(lxy + lzn));
          C[m + ldc * n] = (TC)((big < 1e-11) ? 0.0f : v * inv);
        } else {
          ep.u[uidx(m, dn)] = (double)v;
        }
      }
  }
}

}  // namespace mf
