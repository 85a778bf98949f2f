// fp16x2 eigenvector transform with the eigenvector matrix resident in
// registers (eigenvector dimension 129..256).
//
// Measured on the earlier kernels (gemm_h2_kernel and variants): removing
// the MFMAs, the field loads or the stores each saves only a few of ~40 us
// per launch; the k-loop waits on the eigenvector fragments it re-reads from
// L2 for every 64-column tile, whose latency under load exceeds the MFMA
// work one prefetch step covers.  Here a persistent block of 8 waves (one
// per CU) keeps the whole split eigenvector matrix in registers: wave w
// holds rows 32 w .. 32 w + 31 for every k (two fp16 parts, KP/32 k-steps:
// at most 128 VGPRs), loaded once per launch.  The block walks the 64-row
// field tiles blockIdx.x, blockIdx.x + gridDim.x, ...:
//   - the next tile is copied into an FP32 staging buffer with direct
//     global-to-LDS loads, issued before the current tile's k-loop;
//   - the k-loop reads only LDS (the current tile, already split into its
//     two fp16 parts) and registers;
//   - after the k-loop each wave waits for its staging loads and for the
//     previous tile's stores (both issued a k-loop earlier), the block
//     splits the staged tile, and only then are the finished tile's outputs
//     stored, so that they complete during the next k-loop.
// Tile orientation as gemm_h2_kernel: WLEFT C(M, N) = W(M, K) F(K, N) with
// F[k + ldf n], field rows = columns n; !WLEFT C(M, N) = F(M, K) W^T with
// F[m + ldf k] (+ sF per batch), field rows = rows m.  The wave tile is 32
// eigenvector rows x 64 field rows.
#pragma once
#include <hip/hip_runtime.h>

#include "gemm_h2.h"

namespace mf {

constexpr int H2R_RF = 64;    // field rows per tile
constexpr int H2R_KP = 256;   // largest k extent
constexpr int H2R_PX = 260;   // staging row pitch (floats), WLEFT [row][k]
constexpr int H2R_PY = 66;    // staging row pitch (floats), !WLEFT [k][row]
constexpr int H2R_SP = 264;   // split tile row pitch (halves), [part][row][k]
constexpr int H2R_STAGE = H2R_RF * H2R_PX > H2R_KP * H2R_PY ? H2R_RF * H2R_PX : H2R_KP * H2R_PY;

// Wsplit: fp16 [2][KP/8][256][8] (scaled by wscale), zero in k >= K; K <=
// KP <= 256, KP a multiple of 64; the eigenvector side (WLEFT: M, else N)
// <= 256.  Field rows FR = WLEFT ? N : M per batch, batch entries sF apart
// (outputs sC apart).  grid: min(tiles, CUs) blocks of 512 threads, tiles =
// ceil(FR / 64) * batch.
template <typename TC, bool WLEFT, int EPI>
__global__ void __launch_bounds__(512, 1)
    gemm_h2r_kernel(const float *__restrict__ F, long ldf, long sF,
                    const _Float16 *__restrict__ Wsplit, int KP, TC *__restrict__ C, long ldc,
                    long sC, int M, int N, int K, int batch, EpiArgs ep,
                    const unsigned *__restrict__ fmax_in, float wscale,
                    unsigned *__restrict__ fmax_out, unsigned *__restrict__ zero_slot) {
  constexpr int BK = 32, WROWS = 256;
  __shared__ __attribute__((aligned(16))) float stage[H2R_STAGE];
  __shared__ __attribute__((aligned(16))) _Float16 S[2][H2R_RF][H2R_SP];
  __shared__ float bmax[8];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (zero_slot && blockIdx.x == 0 && tid == 0) *zero_slot = 0u;
  const float fs = h2_scale(__uint_as_float(*fmax_in));
  const float unscale = 1.0f / (fs * wscale);
  const long FR = WLEFT ? N : M;
  const int ftiles = (int)((FR + H2R_RF - 1) / H2R_RF);
  const int ntiles = ftiles * batch;
  const int nk = KP / BK;

  // ---- this wave's eigenvector rows 32 wave + 16 j + l%16, every k-step
  f16x8_t wf[H2R_KP / BK][2][2];
  {
    const char *Wb = reinterpret_cast<const char *>(Wsplit);
    const unsigned row = 32u * wave + (lane & 15);
#pragma unroll
    for (int t = 0; t < H2R_KP / BK; ++t) {
      if (t >= nk) break;
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          wf[t][p][j] = *reinterpret_cast<const f16x8_t *>(
              Wb + (((unsigned)p * (KP / 8) + 4u * t + (lane >> 4)) * WROWS + row + 16u * j) * 16u);
    }
  }

  // ---- staging: tile -> stage (this wave's share, direct to LDS)
  auto tile_pos = [&](int tile, long &f0, long &z) {
    z = tile / ftiles;
    f0 = (long)(tile - z * ftiles) * H2R_RF;
  };
  auto stage_tile = [&](int tile) {
    long f0, z;
    tile_pos(tile, f0, z);
    const float *Fz = F + z * sF;
    if (WLEFT) {
      // rows wave, wave + 8, ...: KP/64 chunks of 64 k each
      for (int r = wave; r < H2R_RF; r += 8) {
        const long fr = f0 + r < FR ? f0 + r : FR - 1;
        const float *rowp = Fz + fr * ldf;
        for (int c = 0; c < KP / 64; ++c) {
          const unsigned k0 = 64u * c + lane;
          __builtin_amdgcn_global_load_lds(rowp + (k0 < (unsigned)K ? k0 : (unsigned)K - 1u),
                                           stage + r * H2R_PX + 64 * c, 4, 0, 0);
        }
      }
    } else {
      // k rows wave, wave + 8, ...: the 64 field rows of one k
      const unsigned fr = (unsigned)(f0 + lane < FR ? lane : FR - 1 - f0);
      const float *F0 = Fz + f0;
      for (int k0 = wave; k0 < KP; k0 += 8) {
        const long k = k0 < K ? k0 : K - 1;
        __builtin_amdgcn_global_load_lds(F0 + ldf * k + fr, stage + k0 * H2R_PY, 4, 0, 0);
      }
    }
  };
  // ---- split: stage (FP32) -> S[part][row][k] (x fs), cooperative.
  // Columns k >= K hold clamped copies of column K-1 (finite); the
  // eigenvector split is zero there.
  auto split_tile = [&] {
    const int row = WLEFT ? tid >> 3 : lane;
    const int kq0 = WLEFT ? (tid & 7) : wave;
#pragma unroll
    for (int j = 0; j < H2R_KP / 64; ++j) {
      if (j >= KP / 64) break;
      const int k0 = 8 * (kq0 + 8 * j);
      f16x8_t p0, p1;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const float v = WLEFT ? stage[row * H2R_PX + k0 + q] : stage[(k0 + q) * H2R_PY + row];
        _Float16 a, b;
        split2h(v * fs, a, b);
        p0[q] = a;
        p1[q] = b;
      }
      *reinterpret_cast<f16x8_t *>(&S[0][row][k0]) = p0;
      *reinterpret_cast<f16x8_t *>(&S[1][row][k0]) = p1;
    }
  };

  // ---- products on the split tile: acc[mi][ni] over MI x NI 16 x 16 blocks
  constexpr int MI = WLEFT ? 2 : 4, NI = WLEFT ? 4 : 2;
  const int ko = 8 * (lane >> 4), rl = lane & 15;
  float amax = 0.0f;
  auto compute = [&](f4_t (&acc)[MI][NI]) {
#pragma unroll
    for (int a = 0; a < MI; ++a)
#pragma unroll
      for (int b = 0; b < NI; ++b) acc[a][b] = (f4_t){0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < H2R_KP / BK; ++t) {
      if (t >= nk) break;
      f16x8_t ff[2][4];
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          ff[p][i] = *reinterpret_cast<const f16x8_t *>(&S[p][16 * i + rl][t * BK + ko]);
#pragma unroll
      for (int ni = 0; ni < NI; ++ni)
#pragma unroll
        for (int mi = 0; mi < MI; ++mi) {
          // D columns (m) from the second operand, rows (n) from the first
          const f16x8_t &a0 = WLEFT ? wf[t][0][mi] : ff[0][mi];
          const f16x8_t &a1 = WLEFT ? wf[t][1][mi] : ff[1][mi];
          const f16x8_t &b0 = WLEFT ? ff[0][ni] : wf[t][0][ni];
          const f16x8_t &b1 = WLEFT ? ff[1][ni] : wf[t][1][ni];
          f4_t c = acc[mi][ni];
          c = __builtin_amdgcn_mfma_f32_16x16x32_f16(b1, a0, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_f16(b0, a1, c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_f16(b0, a0, c, 0, 0, 0);
          acc[mi][ni] = c;
        }
    }
  };
  auto store = [&](int tile, f4_t (&acc)[MI][NI]) {
    // outputs: m = mw + 16 mi + l%16, n = nw + 16 ni + 4 (l/16) + r
    long f0, z;
    tile_pos(tile, f0, z);
    const long mw = WLEFT ? 32L * wave : f0, nw = WLEFT ? f0 : 32L * wave;
    const int lm = lane & 15, ln = 4 * (lane >> 4);
    const bool full = mw + 16 * MI <= M && nw + 16 * NI <= N;
    if (EPI == EPI_STORE) {
      TC *pl = C + z * sC + (mw + lm) + ldc * (nw + ln);
      auto body = [&](auto checked) {
#pragma unroll
        for (int mi = 0; mi < MI; ++mi) {
          if (decltype(checked)::value && mw + 16 * mi + lm >= M) continue;
#pragma unroll
          for (int ni = 0; ni < NI; ++ni)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              if (decltype(checked)::value && nw + ln + 16 * ni + r >= N) continue;
              const float v = acc[mi][ni][r];
              amax = fmaxf(amax, fabsf(v));
              pl[16 * mi + ldc * (16 * ni + r)] = (TC)(v * unscale);
            }
        }
      };
      if (full) body(std::false_type{});
      else body(std::true_type{});
    } else {
      // u[(i+1) + nx((j+1) + ny(k+1))], m = i, n = j + n2 k (WLEFT only);
      // n = nb + dn with dn < 64 <= n2 wraps at most once (see h2_store)
      const long nb = nw + ln;
      const int jb = (int)(nb % ep.n2), kb = (int)(nb / ep.n2);
      const long nx = ep.nx, nxy = (long)ep.nx * ep.ny;
      double *ul = ep.u + (mw + lm + 1) + nx * (jb + 1) + nxy * (kb + 1);
      auto body = [&](auto checked) {
#pragma unroll
        for (int ni = 0; ni < NI; ++ni)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int dn = 16 * ni + r;
            if (decltype(checked)::value && nb + dn >= N) continue;
            long off;
            if (ep.n2 >= 64) {
              off = nx * dn + (jb + dn >= ep.n2 ? 2 * nx : 0);
            } else {
              const int q = (jb + dn) / ep.n2;
              off = nx * dn + 2 * nx * q;
            }
#pragma unroll
            for (int mi = 0; mi < MI; ++mi) {
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
  };

  // ---- pipeline: per tile [stage next] [products] [wait: the stage loads
  // and the previous tile's stores, both issued a k-loop ago] [split next]
  // [this tile's stores, completing during the next k-loop]
  if ((int)blockIdx.x < ntiles) stage_tile(blockIdx.x);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  split_tile();
  __syncthreads();
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int next = tile + gridDim.x;
    const bool more = next < ntiles;  // block-uniform
    if (more) stage_tile(next);
    f4_t acc[MI][NI];
    compute(acc);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();  // every wave is done with S and every stage load has landed
    if (more) split_tile();
    __syncthreads();
    store(tile, acc);
  }
  if (EPI == EPI_STORE && fmax_out) {
    float vmax = amax * unscale;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off, 64));
    if (lane == 0) bmax[wave] = vmax;
    __syncthreads();
    if (tid == 0) {
      float b = bmax[0];
#pragma unroll
      for (int q = 1; q < 8; ++q) b = fmaxf(b, bmax[q]);
      atomicMax(fmax_out, __float_as_uint(b));
    }
  }
}

}  // namespace mf
