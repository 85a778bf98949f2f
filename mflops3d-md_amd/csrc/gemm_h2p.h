// fp16x2 transform pairs on whole z-planes: the x and y transforms of one
// plane in one block, the intermediate kept in registers.
//
// The forward pair U = (Wx R) Wy^T and the backward pair u (+)= Vx (V Vy^T)
// act on one z-plane at a time (n1, n2 <= 256).  A block of 8 waves (one per
// CU, persistent over planes z = blockIdx.x, blockIdx.x + gridDim.x, ...)
// forms the first product with each wave owning 32 rows of its result: the
// wave's 32 x 256 slab sits in 128 accumulator VGPRs, which is every value
// the second product needs for those rows, so the second product reads its
// left operand from registers (split into two fp16 parts in place, with a
// per-wave power-of-two scale) and only the second matrix streams through
// LDS.  Against two separate transforms this saves one FP32 field write and
// read per pair (8 B/point of HBM traffic) and one launch.
//
// A 16x16x32 MFMA accumulator holds, in lane l, rows 4 (l/16) + r (r < 4)
// of its 16-block: the two 16-blocks of one 32-wide k-step are, seen as an
// operand, the k order perm(8 g + q) = q < 4 ? 4 g + q : 16 + 4 g + q - 4
// (g = l/16).  The matrix of the second product is stored with its k index
// in that order (h2p_perm), so register accumulators feed the MFMAs as they
// are.
//
// Phase 1 (first product), per 32-wide k-slice: the slice of the field is
// copied to an FP32 staging buffer with direct global-to-LDS loads (issued
// one slice ahead), split cooperatively into two fp16 parts, and multiplied
// by the wave's own rows of the first matrix (fragments from global memory,
// one slice ahead).  Phase 2 (second product), per 32-row chunk of the
// second matrix: the chunk (already split, k permuted) is copied to LDS one
// chunk ahead, the wave's 32 x 32 outputs are formed and stored while the
// next chunk's products run.
#pragma once
#include <hip/hip_runtime.h>

#include "gemm_h2.h"

namespace mf {

constexpr int H2P_N = 256;               // largest n1, n2 (8 waves x 32 rows)
constexpr int H2P_SP = 40;               // split slice row pitch (halves): 80-byte rows
constexpr int H2P_STAGE = 32 * H2P_N;    // FP32 values per staging buffer (32 KB)
// LDS: staging buffer 0, then the split slice S[2][256][H2P_SP] (phase 1);
// staging buffer 1 (phase 2) overlays S.  72 KB: two blocks fit one CU.
constexpr int H2P_LDS = H2P_STAGE + 2 * H2P_N * H2P_SP / 2;
// backward add: each wave's 32 x 32 old values of u (FP64) per x chunk, rows
// of 32 doubles at a pitch of 72 floats (the four row groups a wave reads at
// once fall on disjoint banks)
constexpr int H2P_UP = 72;

// contraction index held at position qq (0..31) of a permuted 32-wide k-step
__host__ __device__ constexpr int h2p_perm(int qq) {
  return (qq & 7) < 4 ? 4 * (qq >> 3) + (qq & 7) : 16 + 4 * (qq >> 3) + (qq & 7) - 4;
}

// Copy chunk c (rows 32 c .. 32 c + 31, every k) of a split matrix
// [2][kp/8][256][8] (halves) into dst [2][kp/8][32][8]: kp/2 wave loads of
// 256 bytes, wave w of NW taking w, w + NW, ...
template <int NW>
__device__ __forceinline__ void h2p_stage_chunk(const _Float16 *__restrict__ W, int kp, int c,
                                                float *dst, int wave, int lane) {
  const float *Wf = reinterpret_cast<const float *>(W);
  const int n = kp / 2;  // 2 parts x kp/8 groups x 2 half-rows of 256 bytes
  for (int i = wave; i < n; i += NW) {
    const int pk = i >> 1, h = i & 1;
    __builtin_amdgcn_global_load_lds(Wf + ((long)pk * H2P_N + 32 * c) * 4 + 64 * h + lane,
                                     dst + 64 * i, 4, 0, 0);
  }
}

// The plane and the owner index (the eigenvector rows 32 owner .. 32 owner
// + 31 of the first product) of this block's wave 0.  NW = 8: one block per
// plane, persistent over planes blockIdx.x + k gridDim.x.  NW = 4: two
// blocks per plane (each 128 rows), one plane per block; the two blocks of a
// plane are 8 apart in dispatch order, so that both run on one XCD and the
// second one's field reads hit the L2 the first one filled.
// grid: NW = 8: min(n3, CUs); NW = 4: 16 ceil(n3 / 8).
template <int NW>
__device__ __forceinline__ void h2p_block_plane(int &z0, int &zstep, int &owner0, int n3) {
  if (NW == 8) {
    z0 = blockIdx.x;
    zstep = gridDim.x;
    owner0 = 0;
  } else {
    const int b = blockIdx.x, slot = b >> 3;
    z0 = 8 * (slot >> 1) + (b & 7);
    zstep = n3;
    owner0 = NW * (slot & 1);
  }
}

// the two fp16 parts of one operand fragment: p0 at a, p1 at a + off halves
struct h2p_frag {
  f16x8_t p0, p1;
};
__device__ __forceinline__ h2p_frag h2p_read(const _Float16 *a, int off) {
  return h2p_frag{*reinterpret_cast<const f16x8_t *>(a), *reinterpret_cast<const f16x8_t *>(a + off)};
}

__device__ __forceinline__ float h2p_wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// Forward pair on planes of R[x + n1 (y + n2 z)] (FP32, largest magnitude
// in *fmax_in): U[kx + n1 (ly + n2 z)] = sum_y (sum_x Wx(kx, x) R(x, y))
// Wy(ly, y).  Wx: split [2][KPX/8][256][8] (rows kx, scaled by wsx), Wyp:
// split [2][KPY/8][256][8] (rows ly, k permuted, scaled by wsy); both zero
// outside the matrix.  The largest stored magnitude goes to *fmax_out.
// grid: h2p_block_plane, blocks of 64 NW threads; n1 <= KPX, n2 <= KPY, the
// padded k extents 192 or 256.
template <int NW, int KPX, int KPY>
__global__ void __launch_bounds__(64 * NW, 8 / NW)
    h2p_fwd_kernel(const float *__restrict__ R, int n1, int n2, int n3,
                   const _Float16 *__restrict__ Wx, const _Float16 *__restrict__ Wyp, float wsx,
                   float wsy, float *__restrict__ U,
                   const unsigned *__restrict__ fmax_in, unsigned *__restrict__ fmax_out) {
  __shared__ __attribute__((aligned(16))) float lds[H2P_LDS];
  _Float16(*S)[H2P_N][H2P_SP] = reinterpret_cast<_Float16(*)[H2P_N][H2P_SP]>(lds + H2P_STAGE);
  __shared__ float bmax[NW];
  constexpr int VPT = 128 / NW, TPR = 32 / VPT;  // split: values per thread, threads per row
  const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, rl = lane & 15;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  int z0, zstep, owner0;
  h2p_block_plane<NW>(z0, zstep, owner0, n3);
  const int own = owner0 + wave;  // this wave's rows: 32 own .. 32 own + 31
  const float fs = h2_scale(__uint_as_float(*fmax_in));
  const float u1 = 1.0f / (fs * wsx);
  constexpr int nkx = KPX / 32, nky = KPY / 32;
  const int ncy = (n2 + 31) / 32;        // ly chunks of the second
  const bool active = 32 * own < n1;     // the wave owns rows kx < n1
  const long plane = (long)n1 * n2;
  float amax = 0.0f;

  for (int z = z0; z < n3; z += zstep) {
    const float *Rz = R + z * plane;
    // raw slice t: x in [32 t, 32 t + 32), y in [0, 256) -> staging buffer 0, [y][32]
    auto stage_slice = [&](int t) {
      for (int i = wave; i < H2P_N / 2; i += NW) {
        const int y = 2 * i + (lane >> 5), x = 32 * t + (lane & 31);
        __builtin_amdgcn_global_load_lds(
            Rz + (long)(y < n2 ? y : n2 - 1) * n1 + (x < n1 ? x : n1 - 1), &lds[64 * i], 4, 0,
            0);
      }
    };
    auto split_slice = [&] {
      const int y = tid / TPR, x0 = VPT * (tid % TPR);
      f16x8_t p0[VPT / 8], p1[VPT / 8];
#pragma unroll
      for (int q = 0; q < VPT; ++q) {
        _Float16 a, b;
        split2h(lds[32 * y + x0 + q] * fs, a, b);
        p0[q >> 3][q & 7] = a;
        p1[q >> 3][q & 7] = b;
      }
#pragma unroll
      for (int h = 0; h < VPT / 8; ++h) {
        *reinterpret_cast<f16x8_t *>(&S[0][y][x0 + 8 * h]) = p0[h];
        *reinterpret_cast<f16x8_t *>(&S[1][y][x0 + 8 * h]) = p1[h];
      }
    };
    const char *Wxb = reinterpret_cast<const char *>(Wx);
    auto load_wx = [&](int t, f16x8_t (&wf)[2][2]) {
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
          wf[p][mi] = *reinterpret_cast<const f16x8_t *>(
              Wxb + (((long)p * (KPX / 8) + 4 * t + g) * H2P_N + 32 * own + 16 * mi + rl) * 16);
    };

    // ---- phase 1: T(kx, y) = sum_x Wx(kx, x) R(x, y); lane holds
    // kx = 32 own + 16 mi + l%16, y = 16 ni + 4 g + r
    f4_t T[2][16];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int ni = 0; ni < 16; ++ni) T[mi][ni] = (f4_t){0, 0, 0, 0};
    f16x8_t wc[2][2], wn[2][2];
    stage_slice(0);
    load_wx(0, wc);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    split_slice();
    __syncthreads();
    for (int t = 0; t < nkx; ++t) {
      const bool more = t + 1 < nkx;
      if (more) {
        stage_slice(t + 1);
        load_wx(t + 1, wn);
      }
      if (active) {
        // every y block up to KPY: blocks past n2 hold copies of row n2 - 1,
        // which meet zero columns of Wyp
        // (each block's fragments read one block ahead; the fence keeps
        // the compiler from hoisting every read, which spills)
        const _Float16 *Sa = &S[0][rl][8 * g];
        constexpr int SOFF = H2P_N * H2P_SP;  // part 1 - part 0, halves
        h2p_frag a = h2p_read(Sa, SOFF);
#pragma unroll
        for (int ni = 0; ni < KPY / 16; ++ni) {
          h2p_frag an = a;
          if (ni + 1 < KPY / 16) an = h2p_read(Sa + 16 * (ni + 1) * H2P_SP, SOFF);
#pragma unroll
          for (int mi = 0; mi < 2; ++mi) {
            f4_t c = T[mi][ni];
            c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.p1, wc[0][mi], c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.p0, wc[1][mi], c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.p0, wc[0][mi], c, 0, 0, 0);
            T[mi][ni] = c;
          }
          a = an;
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();  // S consumed by every wave, slice t + 1 landed
      if (more) {
        split_slice();
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
          for (int mi = 0; mi < 2; ++mi) wc[p][mi] = wn[p][mi];
      }
      __syncthreads();
    }

    // ---- the wave's slab, scaled into fp16 range and split in place
    float tm = 0.0f;
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int ni = 0; ni < 16; ++ni)
#pragma unroll
        for (int r = 0; r < 4; ++r) tm = fmaxf(tm, fabsf(T[mi][ni][r]));
    const float sT = h2_scale(h2p_wave_max(tm));
    const float u2 = 1.0f / (sT * wsy);
    f16x8_t ts[2][nky][2];  // [mi][k-step][part]
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int j = 0; j < nky; ++j)
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          _Float16 a, b;
          split2h((q < 4 ? T[mi][2 * j][q] : T[mi][2 * j + 1][q - 4]) * sT, a, b);
          ts[mi][j][0][q] = a;
          ts[mi][j][1][q] = b;
        }

    // ---- phase 2: U(kx, ly) = sum_y T(kx, y) Wy(ly, y); lane holds
    // ly = 32 c + 16 nj + 4 g + r, kx = 32 own + 16 mi + l%16
    h2p_stage_chunk<NW>(Wyp, KPY, 0, lds, wave, lane);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    float *Uz = U + z * plane;
    for (int c = 0; c < ncy; ++c) {
      if (c + 1 < ncy) h2p_stage_chunk<NW>(Wyp, KPY, c + 1, lds + H2P_STAGE * ((c + 1) & 1), wave, lane);
      const _Float16 *Wc = reinterpret_cast<const _Float16 *>(lds + H2P_STAGE * (c & 1));
      f4_t acc[2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = (f4_t){0, 0, 0, 0};
      if (active) {
        // steps s = (k-step j, 16-row block nj), fragments read one step ahead
        const _Float16 *Wa = Wc + (g * 32 + rl) * 8;
        constexpr int WOFF = (KPY / 8) * 32 * 8;  // part 1 - part 0, halves
        h2p_frag w = h2p_read(Wa, WOFF);
#pragma unroll
        for (int st = 0; st < 2 * nky; ++st) {
          const int j = st >> 1, nj = st & 1;
          h2p_frag wn = w;
          if (st + 1 < 2 * nky)
            wn = h2p_read(Wa + (4 * ((st + 1) >> 1) * 32 + 16 * ((st + 1) & 1)) * 8, WOFF);
#pragma unroll
          for (int mi = 0; mi < 2; ++mi) {
            f4_t v = acc[nj][mi];
            v = __builtin_amdgcn_mfma_f32_16x16x32_f16(w.p1, ts[mi][j][0], v, 0, 0, 0);
            v = __builtin_amdgcn_mfma_f32_16x16x32_f16(w.p0, ts[mi][j][1], v, 0, 0, 0);
            v = __builtin_amdgcn_mfma_f32_16x16x32_f16(w.p0, ts[mi][j][0], v, 0, 0, 0);
            acc[nj][mi] = v;
          }
          w = wn;
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();  // chunk c consumed by every wave, chunk c + 1 landed
      if (active) {
#pragma unroll
        for (int nj = 0; nj < 2; ++nj)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int ly = 32 * c + 16 * nj + 4 * g + r;
            if (ly >= n2) continue;
#pragma unroll
            for (int mi = 0; mi < 2; ++mi) {
              const int kx = 32 * own + 16 * mi + rl;
              if (kx >= n1) continue;
              const float v = acc[nj][mi][r] * u1 * u2;
              amax = fmaxf(amax, fabsf(v));
              Uz[kx + (long)n1 * ly] = v;
            }
          }
      }
    }
    __syncthreads();  // the staging buffers are reused by the next plane
  }
  if (fmax_out) {
    const float vmax = h2p_wave_max(amax);
    if (lane == 0) bmax[wave] = vmax;
    __syncthreads();
    if (tid == 0) {
      float b = bmax[0];
#pragma unroll
      for (int q = 1; q < NW; ++q) b = fmaxf(b, bmax[q]);
      atomicMax(fmax_out, __float_as_uint(b));
    }
  }
}

// Backward pair on planes of V[kx + n1 (ly + n2 z)] (FP32, largest magnitude
// in *fmax_in): u(x, y, z) (+)= sum_kx Vx(x, kx) (sum_ly V(kx, ly) Vy(y, ly))
// into the interior of u, u[(x+1) + nx ((y+1) + ny (z+1))] (FP64).  Vy:
// split [2][KPY/8][256][8] (rows y, scaled by wsy), Vxp: split
// [2][KPX/8][256][8] (rows x, k permuted, scaled by wsx).  zero_slot, when
// given, is cleared.  grid and limits as h2p_fwd_kernel.
template <int NW, int KPX, int KPY, bool ADD>
__global__ void __launch_bounds__(64 * NW, 8 / NW)
    h2p_bwd_kernel(const float *__restrict__ V, int n1, int n2, int n3,
                   const _Float16 *__restrict__ Vy, const _Float16 *__restrict__ Vxp, float wsy,
                   float wsx, double *__restrict__ u, int nx, int ny,
                   const unsigned *__restrict__ fmax_in, unsigned *__restrict__ zero_slot) {
  __shared__ __attribute__((aligned(16))) float lds[H2P_LDS + (ADD ? NW * 32 * H2P_UP : 0)];
  _Float16(*S)[H2P_N][H2P_SP] = reinterpret_cast<_Float16(*)[H2P_N][H2P_SP]>(lds + H2P_STAGE);
  constexpr int VPT = 128 / NW;  // split: values per thread
  const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, rl = lane & 15;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  int z0, zstep, owner0;
  h2p_block_plane<NW>(z0, zstep, owner0, n3);
  const int own = owner0 + wave;  // this wave's columns: y = 32 own .. 32 own + 31
  if (zero_slot && blockIdx.x == 0 && tid == 0) *zero_slot = 0u;
  const float fv = h2_scale(__uint_as_float(*fmax_in));
  const float u1 = 1.0f / (fv * wsy);
  constexpr int nkx = KPX / 32, nky = KPY / 32;
  const int ncx = (n1 + 31) / 32;        // x chunks of the second
  const bool active = 32 * own < n2;     // the wave owns columns y < n2
  const long plane = (long)n1 * n2;
  const long nxy = (long)nx * ny;

  for (int z = z0; z < n3; z += zstep) {
    const float *Vz = V + z * plane;
    // raw slice t: ly in [32 t, 32 t + 32), kx in [0, 256) -> staging buffer 0, [ly][256]
    auto stage_slice = [&](int t) {
      for (int i = wave; i < H2P_N / 2; i += NW) {
        const int ly = 32 * t + (i >> 2), kx = 64 * (i & 3) + lane;
        __builtin_amdgcn_global_load_lds(
            Vz + (long)(ly < n2 ? ly : n2 - 1) * n1 + (kx < n1 ? kx : n1 - 1), &lds[64 * i],
            4, 0, 0);
      }
    };
    auto split_slice = [&] {
      const int kx = tid & 255, l0 = VPT * (tid >> 8);
      f16x8_t p0[VPT / 8], p1[VPT / 8];
#pragma unroll
      for (int q = 0; q < VPT; ++q) {
        _Float16 a, b;
        split2h(lds[H2P_N * (l0 + q) + kx] * fv, a, b);
        p0[q >> 3][q & 7] = a;
        p1[q >> 3][q & 7] = b;
      }
#pragma unroll
      for (int h = 0; h < VPT / 8; ++h) {
        *reinterpret_cast<f16x8_t *>(&S[0][kx][l0 + 8 * h]) = p0[h];
        *reinterpret_cast<f16x8_t *>(&S[1][kx][l0 + 8 * h]) = p1[h];
      }
    };
    const char *Vyb = reinterpret_cast<const char *>(Vy);
    auto load_vy = [&](int t, f16x8_t (&vf)[2][2]) {
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int nj = 0; nj < 2; ++nj)
          vf[p][nj] = *reinterpret_cast<const f16x8_t *>(
              Vyb + (((long)p * (KPY / 8) + 4 * t + g) * H2P_N + 32 * own + 16 * nj + rl) * 16);
    };

    // ---- phase 1: T(kx, y) = sum_ly V(kx, ly) Vy(y, ly); lane holds
    // kx = 16 mi + 4 g + r, y = 32 own + 16 nj + l%16
    f4_t T[16][2];
#pragma unroll
    for (int mi = 0; mi < 16; ++mi)
#pragma unroll
      for (int nj = 0; nj < 2; ++nj) T[mi][nj] = (f4_t){0, 0, 0, 0};
    f16x8_t vc[2][2], vn[2][2];
    stage_slice(0);
    load_vy(0, vc);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    split_slice();
    __syncthreads();
    for (int t = 0; t < nky; ++t) {
      const bool more = t + 1 < nky;
      if (more) {
        stage_slice(t + 1);
        load_vy(t + 1, vn);
      }
      if (active) {
        // every kx block up to KPX: blocks past n1 hold copies of row n1 - 1,
        // which meet zero columns of Vxp
        const _Float16 *Sa = &S[0][rl][8 * g];
        constexpr int SOFF = H2P_N * H2P_SP;
        h2p_frag a = h2p_read(Sa, SOFF);
#pragma unroll
        for (int mi = 0; mi < KPX / 16; ++mi) {
          h2p_frag an = a;
          if (mi + 1 < KPX / 16) an = h2p_read(Sa + 16 * (mi + 1) * H2P_SP, SOFF);
#pragma unroll
          for (int nj = 0; nj < 2; ++nj) {
            f4_t c = T[mi][nj];
            c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.p1, vc[0][nj], c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.p0, vc[1][nj], c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.p0, vc[0][nj], c, 0, 0, 0);
            T[mi][nj] = c;
          }
          a = an;
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (more) {
        split_slice();
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
          for (int nj = 0; nj < 2; ++nj) vc[p][nj] = vn[p][nj];
      }
      __syncthreads();
    }

    float tm = 0.0f;
#pragma unroll
    for (int mi = 0; mi < 16; ++mi)
#pragma unroll
      for (int nj = 0; nj < 2; ++nj)
#pragma unroll
        for (int r = 0; r < 4; ++r) tm = fmaxf(tm, fabsf(T[mi][nj][r]));
    const float sT = h2_scale(h2p_wave_max(tm));
    const float u2 = 1.0f / (sT * wsx);
    f16x8_t ts[2][nkx][2];  // [nj][k-step][part]
#pragma unroll
    for (int nj = 0; nj < 2; ++nj)
#pragma unroll
      for (int j = 0; j < nkx; ++j)
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          _Float16 a, b;
          split2h((q < 4 ? T[2 * j][nj][q] : T[2 * j + 1][nj][q - 4]) * sT, a, b);
          ts[nj][j][0][q] = a;
          ts[nj][j][1][q] = b;
        }

    // ---- phase 2: u(x, y) (+)= sum_kx Vx(x, kx) T(kx, y); lane holds
    // y = 32 own + 16 nj + 4 g + r, x = 32 c + 16 xb + l%16
    h2p_stage_chunk<NW>(Vxp, KPX, 0, lds, wave, lane);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    double *uz = u + 1 + nx + nxy * (z + 1);
    for (int c = 0; c < ncx; ++c) {
      if (c + 1 < ncx) h2p_stage_chunk<NW>(Vxp, KPX, c + 1, lds + H2P_STAGE * ((c + 1) & 1), wave, lane);
      // add: this wave's old u rows y (x = 32 c .. 32 c + 31, 256 bytes each)
      // into its LDS rows, landing during the products.  Columns past n1
      // read the boundary column and the start of the next row (inside u:
      // y + 2 <= n2 + 1 = ny - 1); they are not used.
      float *uold = lds + H2P_LDS + wave * 32 * H2P_UP;
      if (ADD && active) {
        for (int yy = 0; yy < 32; ++yy) {
          const int y = 32 * own + yy;
          if (y < n2)
            __builtin_amdgcn_global_load_lds(
                reinterpret_cast<const float *>(uz + 32 * c + (long)nx * y) + lane,
                uold + yy * H2P_UP, 4, 0, 0);
        }
      }
      const _Float16 *Wc = reinterpret_cast<const _Float16 *>(lds + H2P_STAGE * (c & 1));
      f4_t acc[2][2];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = (f4_t){0, 0, 0, 0};
      if (active) {
        const _Float16 *Wa = Wc + (g * 32 + rl) * 8;
        constexpr int WOFF = (KPX / 8) * 32 * 8;
        h2p_frag w = h2p_read(Wa, WOFF);
#pragma unroll
        for (int st = 0; st < 2 * nkx; ++st) {
          const int j = st >> 1, xb = st & 1;
          h2p_frag wn = w;
          if (st + 1 < 2 * nkx)
            wn = h2p_read(Wa + (4 * ((st + 1) >> 1) * 32 + 16 * ((st + 1) & 1)) * 8, WOFF);
#pragma unroll
          for (int nj = 0; nj < 2; ++nj) {
            f4_t v = acc[nj][xb];
            v = __builtin_amdgcn_mfma_f32_16x16x32_f16(ts[nj][j][1], w.p0, v, 0, 0, 0);
            v = __builtin_amdgcn_mfma_f32_16x16x32_f16(ts[nj][j][0], w.p1, v, 0, 0, 0);
            v = __builtin_amdgcn_mfma_f32_16x16x32_f16(ts[nj][j][0], w.p0, v, 0, 0, 0);
            acc[nj][xb] = v;
          }
          w = wn;
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (active) {
#pragma unroll
        for (int nj = 0; nj < 2; ++nj)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int y = 32 * own + 16 * nj + 4 * g + r;
            if (y >= n2) continue;
#pragma unroll
            for (int xb = 0; xb < 2; ++xb) {
              const int x = 32 * c + 16 * xb + rl;
              if (x >= n1) continue;
              const double v = (double)(acc[nj][xb][r] * u1 * u2);
              if (ADD) {
                const double o = *reinterpret_cast<const double *>(
                    uold + (16 * nj + 4 * g + r) * H2P_UP + 2 * (16 * xb + rl));
                uz[x + (long)nx * y] = o + v;
              } else {
                uz[x + (long)nx * y] = v;
              }
            }
          }
      }
    }
    __syncthreads();
  }
}

}  // namespace mf
