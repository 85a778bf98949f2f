// fp16x2 transform pairs on whole z-planes: the x and y transforms of one
// plane in one block, the intermediate kept in registers.
//
// The forward pair U = (Wx R) Wy^T and the backward pair u (+)= Vx (V Vy^T)
// act on one z-plane at a time (n1, n2 <= 256).  A block of 8 waves (one per
// CU; block z takes plane z) forms the first product with each wave owning 32 rows of its result: the
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
// copied to one of two FP32 staging buffers with direct global-to-LDS loads
// (two slices in flight), split cooperatively into two fp16 parts, and
// multiplied by the wave's own rows of the first matrix (fragments from
// global memory, one slice ahead); each wave splits the part it staged, into
// one of two split slices, so a slice takes one block barrier.  Phase 2 (second product), per 32-row
// chunk of the second matrix: the chunk (already split, k permuted) is
// copied to LDS one chunk ahead, the wave's 32 x 32 outputs are formed and
// stored while the next chunk's products run.  Inside both loops each
// step's LDS fragments are read one step ahead.
//
// LDS (144 KB, one block per CU): two staging buffers of 32 KB (phase 1:
// field slices; phase 2: matrix chunks) and a third array holding two split
// slices S[2][256][40] (phase 1) and, for the backward add, each wave's old
// values of u (phase 2).  They
// are separate arrays, with the buffer of every access fixed at compile time
// (loops unrolled by buffer parity), so that the compiler's waits before an
// LDS read cover only the direct global-to-LDS loads into that array.
#pragma once
#include <hip/hip_runtime.h>

#include "gemm_h2.h"
#include "lds_sync.h"

namespace mf {

constexpr int H2P_N = 256;             // largest n1, n2 (8 waves x 32 rows)
constexpr int H2P_SP = 40;             // split slice row pitch (halves): 80-byte rows
constexpr int H2P_STAGE = 32 * H2P_N;  // FP32 values per staging buffer (32 KB)
constexpr int H2P_SLDS = 2 * H2P_N * H2P_SP / 2;  // one split slice (floats)
// backward add: each wave's 32 x 32 old values of u (FP64) per x chunk, rows
// of 32 doubles at a pitch of 72 floats (the four row groups a wave reads at
// once fall on disjoint banks)
constexpr int H2P_UP = 72;
constexpr int H2P_ULDS = 8 * 32 * H2P_UP;
static_assert(H2P_ULDS <= 2 * H2P_SLDS, "the old values of u fit the two split slices");

// contraction index held at position qq (0..31) of a permuted 32-wide k-step
__host__ __device__ constexpr int h2p_perm(int qq) {
  return (qq & 7) < 4 ? 4 * (qq >> 3) + (qq & 7) : 16 + 4 * (qq >> 3) + (qq & 7) - 4;
}

// s_waitcnt operand: at most n vector memory operations outstanding, no
// wait on the LDS/scalar and export counters (gfx9 field layout)
constexpr int h2p_vmcnt(int n) { return (n & 15) | ((n >> 4) << 14) | (7 << 4) | (15 << 8); }

// Copy chunk c (rows 32 c .. 32 c + 31, every k) of a split matrix
// [2][KP/8][256][8] (halves) into dst [2][KP/8][32][8]: KP/2 wave loads of
// 256 bytes (2 parts x KP/8 groups x 2 half-rows), wave w taking w, w + 8,
// ...: KP/16 per wave, a count fixed at compile time (the compiler's waits
// count them)
template <int KP>
__device__ __forceinline__ void h2p_stage_chunk(const _Float16 *__restrict__ W, int c, float *dst,
                                                int wave, int lane) {
  const float *Wf = reinterpret_cast<const float *>(W);
#pragma unroll
  for (int q = 0; q < KP / 16; ++q) {
    const int i = wave + 8 * q, pk = i >> 1, h = i & 1;
    __builtin_amdgcn_global_load_lds(Wf + ((long)pk * H2P_N + 32 * c) * 4 + 64 * h + lane,
                                     dst + 64 * i, 4, 0, 0);
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

// Phase 1 of both kernels: acc += the products, over NK k-slices of the
// field, of the split slice's NB 16-row blocks (fragments from S rows
// 16 i + l%16) with the wave's fragments of the first matrix.  stage(t, buf)
// issues slice t's field loads (16 wave instructions) into a staging
// buffer; split(buf, S) splits the part of a staged slice that this wave
// loaded itself into a split slice; load(t, frag) loads the wave's matrix
// fragments of slice t (4 wave instructions).  Each wave splits only what it
// staged, so a slice needs no barrier before its split, only the wave's own
// wait: its explicit vmcnt counts only the 16 field loads of the slice
// staged after it as younger (the compiler may move the fragment loads;
// waiting for them too is cheap, while counting them and finding them moved
// past the wait would leave the slice's last loads in flight).  The split
// slices alternate between two buffers, so one barrier per slice orders
// both the split before its products and the products before the next
// split into the same buffer.  The block index i goes to acc's first index
// when ROWS_FIRST, else to its second.
template <int NK, int NB, bool ROWS_FIRST, int A0, int A1, typename Stage, typename Split,
          typename Load>
__device__ __forceinline__ void h2p_phase1(f4_t (&acc)[A0][A1], float *__restrict__ stA,
                                           float *__restrict__ stB, float *__restrict__ sbuf,
                                           bool active, int rl, int g, Stage stage, Split split,
                                           Load load) {
  _Float16 *S0 = reinterpret_cast<_Float16 *>(sbuf);
  _Float16 *S1 = reinterpret_cast<_Float16 *>(sbuf + H2P_SLDS);
  f16x8_t wf[2][2][2];  // [slice parity][part][16-row block]
  load(0, wf[0]);
  stage(0, stA);
  stage(1, stB);
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(h2p_vmcnt(16));  // this wave's part of slice 0 landed
  asm volatile("" ::: "memory");
  split(stA, S0);
  lds_barrier();
  // unrolled, so that the buffers alternate without copies and the
  // compiler's own waits for the fragments count the loads issued since
#pragma unroll
  for (int t = 0; t < NK; ++t) {
    f16x8_t(&wc)[2][2] = wf[t & 1];
    if (t + 1 < NK) load(t + 1, wf[(t + 1) & 1]);
    if (t + 2 < NK) stage(t + 2, (t & 1) ? stB : stA);
    // issued here, not sunk to their first use a slice later
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    if (active) {
      // every 16-row block of S: blocks past the field's extent hold copies
      // of its last row, which meet zero columns of the second matrix
      const _Float16 *Sa = ((t & 1) ? S1 : S0) + rl * H2P_SP + 8 * g;
      constexpr int SOFF = H2P_N * H2P_SP;  // part 1 - part 0, halves
      h2p_frag a = h2p_read(Sa, SOFF);
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        h2p_frag an = a;
        if (i + 1 < NB) an = h2p_read(Sa + 16 * (i + 1) * H2P_SP, SOFF);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          f4_t *c;
          if constexpr (ROWS_FIRST) c = &acc[i][j];
          else c = &acc[j][i];
          *c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.p1, wc[0][j], *c, 0, 0, 0);
          *c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.p0, wc[1][j], *c, 0, 0, 0);
          *c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.p0, wc[0][j], *c, 0, 0, 0);
        }
        a = an;
        // one block's fragments ahead, not all of them (which spills)
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    if (t + 1 < NK) {
      // this wave's part of slice t + 1 landed: only slice t + 2's field
      // loads (if any) may remain in flight
      if (t + 2 < NK) __builtin_amdgcn_s_waitcnt(h2p_vmcnt(16));
      else __builtin_amdgcn_s_waitcnt(h2p_vmcnt(0));
      asm volatile("" ::: "memory");
      split(((t + 1) & 1) ? stB : stA, ((t + 1) & 1) ? S1 : S0);
      lds_barrier();  // slice t + 1 split; every wave is past slice t - 1's products
    }
  }
  __builtin_amdgcn_s_waitcnt(0);
  lds_barrier();  // the staging buffers are free for phase 2
}

// Forward pair on planes of R[x + n1 (y + n2 z)] (FP32, largest magnitude
// in *fmax_in): U[kx + n1 (ly + n2 z)] = sum_y (sum_x Wx(kx, x) R(x, y))
// Wy(ly, y).  Wx: split [2][KPX/8][256][8] (rows kx, scaled by wsx), Wyp:
// split [2][KPY/8][256][8] (rows ly, k permuted, scaled by wsy); both zero
// outside the matrix.  The largest stored magnitude goes to *fmax_out.
// grid: n3 blocks of 512 threads; n1 <= KPX, n2 <= KPY, the padded k extents
// 192 or 256.
template <int KPX, int KPY>
__global__ void __launch_bounds__(512, 1)
    h2p_fwd_kernel(const float *__restrict__ R, int n1, int n2, int n3,
                   const _Float16 *__restrict__ Wx, const _Float16 *__restrict__ Wyp, float wsx,
                   float wsy, float *__restrict__ U, const unsigned *__restrict__ fmax_in,
                   unsigned *__restrict__ fmax_out) {
  __shared__ __attribute__((aligned(16))) float stA[H2P_STAGE], stB[H2P_STAGE],
      sbuf[2 * H2P_SLDS];
  __shared__ float bmax[8];
  const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, rl = lane & 15;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const float fs = h2_scale(__uint_as_float(*fmax_in));
  const float u1 = 1.0f / (fs * wsx);
  constexpr int nky = KPY / 32;
  const int ncy = (n2 + 31) / 32;     // ly chunks of the second product
  const bool active = 32 * wave < n1;  // the wave owns rows kx < n1
  const long plane = (long)n1 * n2;
  const char *Wxb = reinterpret_cast<const char *>(Wx);
  float amax = 0.0f;

  {
    const int z = blockIdx.x;
    const float *Rz = R + z * plane;
    // field slice t: x in [32 t, 32 t + 32), y in [0, 256) -> buf[y][32];
    // wave w stages rows y = 2 i, 2 i + 1 of its instructions i = w + 8 q
    auto stage = [&](int t, float *buf) {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int i = wave + 8 * q, y = 2 * i + (lane >> 5), x = 32 * t + (lane & 31);
        __builtin_amdgcn_global_load_lds(
            Rz + (long)(y < n2 ? y : n2 - 1) * n1 + (x < n1 ? x : n1 - 1), buf + 64 * i, 4, 0, 0);
      }
    };
    // this wave's 32 staged rows: lane l takes 16 x of row 2 i + h, i = wave
    // + 8 (l / 4), h = (l / 2) % 2, x0 = 16 (l % 2)
    auto split = [&](const float *buf, _Float16 *Sd) {
      const int i = wave + 8 * (lane >> 2), y = 2 * i + ((lane >> 1) & 1), x0 = 16 * (lane & 1);
      f16x8_t p0[2], p1[2];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        _Float16 a, b;
        split2h(buf[32 * y + x0 + q] * fs, a, b);
        p0[q >> 3][q & 7] = a;
        p1[q >> 3][q & 7] = b;
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        *reinterpret_cast<f16x8_t *>(Sd + y * H2P_SP + x0 + 8 * h) = p0[h];
        *reinterpret_cast<f16x8_t *>(Sd + H2P_N * H2P_SP + y * H2P_SP + x0 + 8 * h) = p1[h];
      }
    };
    auto load = [&](int t, f16x8_t (&wf)[2][2]) {
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
          wf[p][mi] = *reinterpret_cast<const f16x8_t *>(
              Wxb + (((long)p * (KPX / 8) + 4 * t + g) * H2P_N + 32 * wave + 16 * mi + rl) * 16);
    };

    // ---- phase 1: T(kx, y) = sum_x Wx(kx, x) R(x, y); lane holds
    // kx = 32 wave + 16 mi + l%16, y = 16 ni + 4 g + r in T[mi][ni]
    f4_t T[2][16];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int ni = 0; ni < 16; ++ni) T[mi][ni] = (f4_t){0, 0, 0, 0};
    h2p_phase1<KPX / 32, KPY / 16, false>(T, stA, stB, sbuf, active, rl, g, stage, split, load);

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
    // ly = 32 c + 16 nj + 4 g + r, kx = 32 wave + 16 mi + l%16
    h2p_stage_chunk<KPY>(Wyp, 0, stA, wave, lane);
    __builtin_amdgcn_s_waitcnt(0);
    lds_barrier();
    float *Uz = U + z * plane;
    auto chunk = [&](int c, float *cur, float *nxt) {
      if (c + 1 < ncy) h2p_stage_chunk<KPY>(Wyp, c + 1, nxt, wave, lane);
      const _Float16 *Wc = reinterpret_cast<const _Float16 *>(cur);
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
      lds_barrier();  // chunk c consumed by every wave, chunk c + 1 landed
      if (active) {
#pragma unroll
        for (int nj = 0; nj < 2; ++nj)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int ly = 32 * c + 16 * nj + 4 * g + r;
            if (ly >= n2) continue;
#pragma unroll
            for (int mi = 0; mi < 2; ++mi) {
              const int kx = 32 * wave + 16 * mi + rl;
              if (kx >= n1) continue;
              const float v = acc[nj][mi][r] * u1 * u2;
              amax = fmaxf(amax, fabsf(v));
              Uz[kx + (long)n1 * ly] = v;
            }
          }
      }
    };
    for (int c = 0; c < ncy; c += 2) {
      chunk(c, stA, stB);
      if (c + 1 < ncy) chunk(c + 1, stB, stA);
    }
  }
  if (fmax_out) {
    const float vmax = h2p_wave_max(amax);
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

// Backward pair on planes of V[kx + n1 (ly + n2 z)] (FP32, largest magnitude
// in *fmax_in): u(x, y, z) (+)= sum_kx Vx(x, kx) (sum_ly V(kx, ly) Vy(y, ly))
// into the interior of u, u[(x+1) + nx ((y+1) + ny (z+1))] (FP64).  Vy:
// split [2][KPY/8][256][8] (rows y, scaled by wsy), Vxp: split
// [2][KPX/8][256][8] (rows x, k permuted, scaled by wsx).  zero_slot, when
// given, is cleared.  grid and limits as h2p_fwd_kernel.
template <int KPX, int KPY, bool ADD>
__global__ void __launch_bounds__(512, 1)
    h2p_bwd_kernel(const float *__restrict__ V, int n1, int n2, int n3,
                   const _Float16 *__restrict__ Vy, const _Float16 *__restrict__ Vxp, float wsy,
                   float wsx, double *__restrict__ u, int nx, int ny,
                   const unsigned *__restrict__ fmax_in, unsigned *__restrict__ zero_slot) {
  __shared__ __attribute__((aligned(16))) float stA[H2P_STAGE], stB[H2P_STAGE],
      sbuf[2 * H2P_SLDS];
  const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, rl = lane & 15;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (zero_slot && blockIdx.x == 0 && tid == 0) *zero_slot = 0u;
  const float fv = h2_scale(__uint_as_float(*fmax_in));
  const float u1 = 1.0f / (fv * wsy);
  constexpr int nkx = KPX / 32;
  const int ncx = (n1 + 31) / 32;     // x chunks of the second product
  const bool active = 32 * wave < n2;  // the wave owns columns y < n2
  const long plane = (long)n1 * n2;
  const long nxy = (long)nx * ny;
  const char *Vyb = reinterpret_cast<const char *>(Vy);

  {
    const int z = blockIdx.x;
    const float *Vz = V + z * plane;
    // field slice t: ly in [32 t, 32 t + 32), kx in [0, 256) -> buf[ly][256];
    // wave w stages the four rows ly = 32 t + 4 w .. 4 w + 3
    auto stage = [&](int t, float *buf) {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int i = 16 * wave + q, ly = 32 * t + (i >> 2), kx = 64 * (i & 3) + lane;
        __builtin_amdgcn_global_load_lds(
            Vz + (long)(ly < n2 ? ly : n2 - 1) * n1 + (kx < n1 ? kx : n1 - 1), buf + 64 * i, 4, 0,
            0);
      }
    };
    // this wave's four staged rows: lane l takes kx = l + 64 c (c < 4), the
    // four ly of each into S[part][kx][4 wave .. 4 wave + 3]
    typedef _Float16 f16x4_t __attribute__((ext_vector_type(4)));
    auto split = [&](const float *buf, _Float16 *Sd) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int kx = lane + 64 * c;
        f16x4_t p0, p1;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          _Float16 a, b;
          split2h(buf[H2P_N * (4 * wave + r) + kx] * fv, a, b);
          p0[r] = a;
          p1[r] = b;
        }
        *reinterpret_cast<f16x4_t *>(Sd + kx * H2P_SP + 4 * wave) = p0;
        *reinterpret_cast<f16x4_t *>(Sd + H2P_N * H2P_SP + kx * H2P_SP + 4 * wave) = p1;
      }
    };
    auto load = [&](int t, f16x8_t (&vf)[2][2]) {
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int nj = 0; nj < 2; ++nj)
          vf[p][nj] = *reinterpret_cast<const f16x8_t *>(
              Vyb + (((long)p * (KPY / 8) + 4 * t + g) * H2P_N + 32 * wave + 16 * nj + rl) * 16);
    };

    // ---- phase 1: T(kx, y) = sum_ly V(kx, ly) Vy(y, ly); lane holds
    // kx = 16 mi + 4 g + r, y = 32 wave + 16 nj + l%16 in T[mi][nj]
    f4_t T[16][2];
#pragma unroll
    for (int mi = 0; mi < 16; ++mi)
#pragma unroll
      for (int nj = 0; nj < 2; ++nj) T[mi][nj] = (f4_t){0, 0, 0, 0};
    h2p_phase1<KPY / 32, KPX / 16, true>(T, stA, stB, sbuf, active, rl, g, stage, split, load);

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
    // y = 32 wave + 16 nj + 4 g + r, x = 32 c + 16 xb + l%16
    h2p_stage_chunk<KPX>(Vxp, 0, stA, wave, lane);
    __builtin_amdgcn_s_waitcnt(0);
    lds_barrier();
    double *uz = u + 1 + nx + nxy * (z + 1);
    float *uold = sbuf + wave * 32 * H2P_UP;
    auto chunk = [&](int c, float *cur, float *nxt) {
      if (c + 1 < ncx) h2p_stage_chunk<KPX>(Vxp, c + 1, nxt, wave, lane);
      // add: this wave's old u rows y (x = 32 c .. 32 c + 31, 256 bytes
      // each) into its LDS rows, landing during the products.  Columns past
      // n1 read the boundary column and the start of the next row (inside
      // u: y + 2 <= n2 + 1 = ny - 1); they are not used.
      if (ADD && active) {
        for (int yy = 0; yy < 32; ++yy) {
          const int y = 32 * wave + yy;
          if (y < n2)
            __builtin_amdgcn_global_load_lds(
                reinterpret_cast<const float *>(uz + 32 * c + (long)nx * y) + lane,
                uold + yy * H2P_UP, 4, 0, 0);
        }
      }
      const _Float16 *Wc = reinterpret_cast<const _Float16 *>(cur);
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
      lds_barrier();  // chunk c consumed by every wave, chunk c + 1 landed
      if (active) {
#pragma unroll
        for (int nj = 0; nj < 2; ++nj)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int y = 32 * wave + 16 * nj + 4 * g + r;
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
    };
    for (int c = 0; c < ncx; c += 2) {
      chunk(c, stA, stB);
      if (c + 1 < ncx) chunk(c + 1, stB, stA);
    }
  }
}

// The z transforms of an inner solve with the divide between them, on
// columns: out[n + n12 z] = sum_kz Vz(z, kz) T(kz, n) with
// T(kz, n) = sum_z' Wz(kz, z') F[n + n12 z'] / (lx(i1) + ly(i2) + lz(kz) +
// sigma), n = i1 + n1 i2, and T = 0 where |lx|, |ly|, |lz| and |sigma| are
// all below 1e-11 (the null mode; class_solver_3d.f90:202-225).  A block of 8
// waves owns 256 consecutive columns n, each wave 32 of them with every kz of
// T in 128 accumulator VGPRs, so both contractions run over indices the wave
// holds: the field's values are loaded straight into operand fragments
// (registers, one z-slice ahead), and only the two z matrices, which every
// wave needs whole, stream through LDS (32-wide z-slices of Wz, then 32-row
// chunks of Vz with k permuted as in the plane pairs).  F's largest magnitude
// is in *fmax_in; the largest stored magnitude goes to *fmax_out.  Wz: split
// [2][KPZ/8][256][8] (rows kz, scaled by wsw), Vzp: split [2][KPZ/8][256][8]
// (rows z, k permuted, scaled by wsv).  grid: ceil(n12 / 256) blocks of 512
// threads; n3 <= KPZ, KPZ 192 or 256.
template <int KPZ>
__global__ void __launch_bounds__(512, 1)
    h2c_kernel(const float *__restrict__ F, long n12, int n3, const _Float16 *__restrict__ Wz,
               const _Float16 *__restrict__ Vzp, float wsw, float wsv, float *__restrict__ out,
               EpiArgs ep, const unsigned *__restrict__ fmax_in, unsigned *__restrict__ fmax_out) {
  __shared__ __attribute__((aligned(16))) float stA[H2P_STAGE], stB[H2P_STAGE];
  __shared__ __attribute__((aligned(16))) float lzs[H2P_N];
  __shared__ unsigned char lzn[H2P_N];
  __shared__ float bmax[8];
  constexpr int NK = KPZ / 32;  // z-slices of the first product, k-steps of the second
  const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, rl = lane & 15;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const long n0 = 256L * blockIdx.x + 32 * wave;  // this wave's first column
  const bool active = n0 < n12;
  const float fs = h2_scale(__uint_as_float(*fmax_in));
  const float u1 = 1.0f / (fs * wsw);
  const int ncz = (n3 + 31) / 32;
  for (int k = tid; k < H2P_N; k += 512) {
    const double l = k < n3 ? ep.lz[k] : 1.0;
    lzs[k] = (float)l;
    lzn[k] = fabs(l) < 1e-11 ? 1 : 0;
  }
  // this lane's two columns n = n0 + 16 nj + l%16 (clamped for the loads)
  long nc[2];
  float lxy[2];
  bool nxy[2];
#pragma unroll
  for (int nj = 0; nj < 2; ++nj) {
    const long n = n0 + 16 * nj + rl;
    nc[nj] = n < n12 ? n : n12 - 1;
    const int i1 = (int)(nc[nj] % ep.n1), i2 = (int)(nc[nj] / ep.n1);
    lxy[nj] = (float)(ep.lx[i1] + ep.ly[i2] + ep.sigma);
    nxy[nj] = fmax(fmax(fabs(ep.lx[i1]), fabs(ep.ly[i2])), fabs(ep.sigma)) < 1e-11;
  }
  // the field values of z-slice t this lane's fragments hold: z = 32 t + 8 g + q
  auto load_f = [&](int t, float (&fv)[2][8]) {
#pragma unroll
    for (int nj = 0; nj < 2; ++nj)
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int z = 32 * t + 8 * g + q;
        fv[nj][q] = F[nc[nj] + n12 * (z < n3 ? z : n3 - 1)];
      }
  };
  // z-slice t of Wz: [2][4][256][8] halves, 128 wave loads of 256 bytes
  auto stage_w = [&](int t, float *buf) {
    const float *Wf = reinterpret_cast<const float *>(Wz);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int i = wave + 8 * q, pg = i >> 4, h = i & 15;  // pg = 4 part + group
      __builtin_amdgcn_global_load_lds(
          Wf + ((long)((pg >> 2) * (KPZ / 8) + 4 * t + (pg & 3)) * H2P_N) * 4 + 64 * h + lane,
          buf + 64 * i, 4, 0, 0);
    }
  };

  // ---- phase 1: T(kz, n); lane holds kz = 16 mi + 4 g + r, n = n0 + 16 nj + l%16
  f4_t T[16][2];
#pragma unroll
  for (int mi = 0; mi < 16; ++mi)
#pragma unroll
    for (int nj = 0; nj < 2; ++nj) T[mi][nj] = (f4_t){0, 0, 0, 0};
  float fv[2][2][8];  // [slice parity][nj][q]
  load_f(0, fv[0]);
  stage_w(0, stA);
  __builtin_amdgcn_s_waitcnt(0);
  lds_barrier();
#pragma unroll
  for (int t = 0; t < NK; ++t) {
    const float *buf = (t & 1) ? stB : stA;
    if (t + 1 < NK) {
      stage_w(t + 1, (t & 1) ? stA : stB);
      load_f(t + 1, fv[(t + 1) & 1]);
    }
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    if (active) {
      f16x8_t ff[2][2];  // [nj][part]
#pragma unroll
      for (int nj = 0; nj < 2; ++nj)
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          _Float16 a, b;
          split2h(fv[t & 1][nj][q] * fs, a, b);
          ff[nj][0][q] = a;
          ff[nj][1][q] = b;
        }
      const _Float16 *Wa = reinterpret_cast<const _Float16 *>(buf) + (g * H2P_N + rl) * 8;
      constexpr int WOFF = 4 * H2P_N * 8;  // part 1 - part 0, halves
      h2p_frag a = h2p_read(Wa, WOFF);
#pragma unroll
      for (int mi = 0; mi < KPZ / 16; ++mi) {
        h2p_frag an = a;
        if (mi + 1 < KPZ / 16) an = h2p_read(Wa + 16 * (mi + 1) * 8, WOFF);
#pragma unroll
        for (int nj = 0; nj < 2; ++nj) {
          f4_t c = T[mi][nj];
          c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.p1, ff[nj][0], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.p0, ff[nj][1], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a.p0, ff[nj][0], c, 0, 0, 0);
          T[mi][nj] = c;
        }
        a = an;
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    // slice t + 1 (matrix in LDS from every wave, field in registers) landed
    __builtin_amdgcn_s_waitcnt(0);
    lds_barrier();
  }

  // ---- the divide, then the wave's slab scaled into fp16 range and split
  float tm = 0.0f;
#pragma unroll
  for (int mi = 0; mi < 16; ++mi) {
    const int kz0 = 16 * mi + 4 * g;
    const f4_t lz4 = *reinterpret_cast<const f4_t *>(&lzs[kz0]);
#pragma unroll
    for (int nj = 0; nj < 2; ++nj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const bool nul = nxy[nj] && lzn[kz0 + r];
        const float v = nul ? 0.0f : T[mi][nj][r] * __builtin_amdgcn_rcpf(lxy[nj] + lz4[r]);
        T[mi][nj][r] = v;
        tm = fmaxf(tm, fabsf(v));
      }
  }
  const float sT = h2_scale(h2p_wave_max(tm));
  const float u2 = 1.0f / (sT * wsv);
  f16x8_t ts[2][NK][2];  // [nj][k-step][part]
#pragma unroll
  for (int nj = 0; nj < 2; ++nj)
#pragma unroll
    for (int j = 0; j < NK; ++j)
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        _Float16 a, b;
        split2h((q < 4 ? T[2 * j][nj][q] : T[2 * j + 1][nj][q - 4]) * sT, a, b);
        ts[nj][j][0][q] = a;
        ts[nj][j][1][q] = b;
      }

  // ---- phase 2: out(z, n) = sum_kz Vz(z, kz) T(kz, n); lane holds
  // z = 32 c + 16 zb + 4 g + r, n = n0 + 16 nj + l%16
  float amax = 0.0f;
  h2p_stage_chunk<KPZ>(Vzp, 0, stA, wave, lane);
  __builtin_amdgcn_s_waitcnt(0);
  lds_barrier();
  auto chunk = [&](int c, float *cur, float *nxt) {
    if (c + 1 < ncz) h2p_stage_chunk<KPZ>(Vzp, c + 1, nxt, wave, lane);
    f4_t acc[2][2];  // [zb][nj]
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = (f4_t){0, 0, 0, 0};
    if (active) {
      const _Float16 *Wa = reinterpret_cast<const _Float16 *>(cur) + (g * 32 + rl) * 8;
      constexpr int WOFF = (KPZ / 8) * 32 * 8;
      h2p_frag w = h2p_read(Wa, WOFF);
#pragma unroll
      for (int st = 0; st < 2 * NK; ++st) {
        const int j = st >> 1, zb = st & 1;
        h2p_frag wn = w;
        if (st + 1 < 2 * NK)
          wn = h2p_read(Wa + (4 * ((st + 1) >> 1) * 32 + 16 * ((st + 1) & 1)) * 8, WOFF);
#pragma unroll
        for (int nj = 0; nj < 2; ++nj) {
          f4_t v = acc[zb][nj];
          v = __builtin_amdgcn_mfma_f32_16x16x32_f16(w.p1, ts[nj][j][0], v, 0, 0, 0);
          v = __builtin_amdgcn_mfma_f32_16x16x32_f16(w.p0, ts[nj][j][1], v, 0, 0, 0);
          v = __builtin_amdgcn_mfma_f32_16x16x32_f16(w.p0, ts[nj][j][0], v, 0, 0, 0);
          acc[zb][nj] = v;
        }
        w = wn;
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    lds_barrier();  // chunk c consumed by every wave, chunk c + 1 landed
    if (active) {
#pragma unroll
      for (int zb = 0; zb < 2; ++zb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int z = 32 * c + 16 * zb + 4 * g + r;
          if (z >= n3) continue;
#pragma unroll
          for (int nj = 0; nj < 2; ++nj) {
            const long n = n0 + 16 * nj + rl;
            if (n >= n12) continue;
            const float v = acc[zb][nj][r] * u1 * u2;
            amax = fmaxf(amax, fabsf(v));
            out[n + n12 * z] = v;
          }
        }
    }
  };
  for (int c = 0; c < ncz; c += 2) {
    chunk(c, stA, stB);
    if (c + 1 < ncz) chunk(c + 1, stB, stA);
  }
  if (fmax_out) {
    const float vmax = h2p_wave_max(amax);
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
