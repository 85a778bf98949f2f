// FP64 residual of the multiplied-through operator, z-streaming form.
//
// The same r' = P g_f - B u as residual_kernel, with the z operators
// applied first (the 1-D operators of different directions commute):
//   r' = Mx D - Rx Yu - FB,   D = My Z1 - Ry Zu,   Yu = My Zu,
//   Z1 = Mz (g - sigma u) - Rz u,   Zu = Mz u
// (RES = false: r' = Mx My Mz g - FB, the right-hand side b').
//
// A block owns a 56 x 8 tile of (i,j) and a chunk of consecutive output
// planes k.  Every thread keeps, for each of its two staged points of the
// 64 x 16 haloed tile, the five planes b(k) .. b(k)+4 of u and g - sigma u in
// registers, b(k) = clamp(k-2, 0, n3-5): the z-window of every row of Mz and
// Rz (host-checked).  Per output plane: one plane enters the queue, Z1 and Zu
// of the staged points go to LDS (double-buffered: one block barrier per
// plane), and each wave forms D and Yu of its row over the 64 staged columns
// and then the row's x-stencils (wave-local LDS rows, no block barrier).
//
// Latency: the plane that enters the queue is loaded one output plane ahead,
// and it is the plane loop's only vector-memory load apart from the face
// values of boundary rows, which travel with it (and the z-face values, read
// on the two planes next to each z face).  Everything else the loop reads
// (weights, boundary-fold rows, x-face products of the chunk) is put in LDS
// when the block starts, and each plane's r' is stored one step late, after
// the next wait: the wait at the end of a step covers only memory operations
// issued at its start (the vector-memory counter completes in order).  At
// 128 registers a second plane in flight spills, and the spill reloads wait
// on everything.  Each plane of u and g is read once per chunk plus four
// planes of warm-up.
#pragma once
#include <hip/hip_runtime.h>

#include "residual.h"

namespace mf {

constexpr int ZS_NW = 8;              // wavefronts per block = output rows per tile
constexpr int ZS_TX = 64 - 2 * RZ_H;  // output columns per tile
constexpr int ZS_TY = ZS_NW;
constexpr int ZS_CZMIN = 8, ZS_CZMAX = 64;  // output planes per block
constexpr int ZS_PT = 12;             // plane table: Mz row (5), Rz row (5), e_z1, e_z2

__device__ __forceinline__ int zs_base(int k, int n3) { return min(max(k - 2, 0), n3 - 5); }

// a wave-uniform double moved to scalar registers
__device__ __forceinline__ double zs_uniform(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readfirstlane((int)b), hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// block (64, ZS_NW); grid: ceil(n1/ZS_TX) ceil(n2/ZS_TY) ceil(n3/a.cz) blocks,
// ZS_CZMIN <= a.cz <= ZS_CZMAX.  Needs n1, n2 >= 4, n3 >= 5 and a.zq: (n3,
// ZS_PT) rows of Mz and Rz on planes b(k) .. b(k)+4, then e_z1(k), e_z2(k).
template <bool RES>
__global__ void __launch_bounds__(64 * ZS_NW, 4) residual_zs_kernel(ResArgs a) {
  constexpr int NT = 64 * ZS_NW;
  constexpr int SW = ZS_TX + 2 * RZ_H, SH = ZS_TY + 2 * RZ_H;  // staged tile 64 x 16
  constexpr int SP = SW + 1;                                   // row pitch
  constexpr int NS = SH / ZS_NW;                               // staged points per thread
  static_assert(SW == 64 && NS * ZS_NW == SH, "one wavefront per staged row");
  __shared__ double Z1[2][SH][SP], ZU[2][SH][SP];
  __shared__ double DS[ZS_TY][SP], YS[ZS_TY][SP];
  __shared__ double WX[2 * BAND_W][SW];      // Mx, Rx rows of each staged column
  __shared__ double EX[2][SW];               // e_x1, e_x2 of each staged column
  __shared__ double WY[ZS_TY][2 * BAND_W];   // My, Ry rows of the tile's rows
  __shared__ double PT[ZS_CZMAX][ZS_PT];     // the chunk's plane table
  __shared__ double FX[ZS_TY][ZS_CZMAX][2];  // x-face products of the tile's rows (x-edge tiles)
  const int lane = threadIdx.x, tid = lane + 64 * threadIdx.y;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);  // wave index = output row (uniform)
  const int n1 = a.n1, n2 = a.n2, n3 = a.n3;
  const int nbx = (n1 + ZS_TX - 1) / ZS_TX, nby = (n2 + ZS_TY - 1) / ZS_TY;
  int i0, j0, z0;
  {
    // consecutive block ids go round-robin to the 8 XCDs: give each XCD a
    // contiguous run of tiles (neighbouring halos then hit its own L2)
    const int nb = gridDim.x, bid = blockIdx.x;
    const int q = nb / 8, r = nb % 8, x = bid % 8, l = bid / 8;
    const int t = (x < r) ? x * (q + 1) + l : r * (q + 1) + (x - r) * q + l;
    i0 = (t % nbx) * ZS_TX;
    j0 = ((t / nbx) % nby) * ZS_TY;
    z0 = (t / (nbx * nby)) * a.cz;
  }
  const int z1 = min(z0 + a.cz, n3), nz = z1 - z0;
  const long sx = (long)n2 * n3, sy = (long)n1 * n3, sz = (long)n1 * n2;
  // ---- block tables
  const int i = i0 - RZ_H + lane;  // staged column (outputs: lanes RZ_H .. RZ_H+ZS_TX-1)
  const bool xout = lane >= RZ_H && lane < RZ_H + ZS_TX && i < n1;
  const int ic = min(max(i, 0), n1 - 1);
  if (w < BAND_W) {  // waves 0-4: weights, two rows each
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int m = 2 * w + h;
      WX[m][lane] = m < BAND_W ? a.Mx.w[8 * ic + m] : (RES ? a.Rx.w[8 * ic + m - BAND_W] : 0.0);
    }
  } else if (w == BAND_W) {
    EX[0][lane] = a.ex[ic];
    EX[1][lane] = a.ex[n1 + ic];
  } else if (w == BAND_W + 1 && lane < ZS_TY * 2 * BAND_W / 2) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int e = 2 * lane + h, r = e / (2 * BAND_W), c = e % (2 * BAND_W);
      const int jr = min(j0 + r, n2 - 1);
      WY[r][c] = c < BAND_W ? a.My.w[8 * jr + c] : (RES ? a.Ry.w[8 * jr + c - BAND_W] : 0.0);
    }
  }
  for (int e = tid; e < nz * ZS_PT; e += NT) {
    const int kk = e / ZS_PT, c = e % ZS_PT;
    const double v = a.zq[(long)ZS_PT * z0 + e];
    PT[kk][c] = (RES || c < BAND_W || c >= 2 * BAND_W) ? v : 0.0;
  }
  const bool xedge = i0 < 2 || i0 + ZS_TX > n1 - 2;  // the tile holds columns with e_x != 0
  if (xedge)
    for (int e = tid; e < ZS_TY * nz * 2; e += NT) {
      const int r = e / (2 * nz), kk = (e / 2) % nz, side = e % 2;
      const int jj = min(j0 + r, n2 - 1);
      FX[r][kk][side] = a.fbx[side * sx + jj + (long)n2 * (z0 + kk)];
    }
  // ---- the wave's row: y weights and window starts (staged rows), boundary fold
  const int j = j0 + w, jc = min(j, n2 - 1);
  const int cym = a.My.c0[jc] - (j0 - RZ_H);
  const int cyr = RES ? a.Ry.c0[jc] - (j0 - RZ_H) : cym;
  // e_y of the row (nonzero on rows 0, 1 and n2-2, n2-1 only: one side)
  const double ey1 = a.ey[jc], ey2 = a.ey[n2 + jc];
  const bool ybnd = j < n2 && (ey1 != 0.0 || ey2 != 0.0);
  const double eyv = ey1 != 0.0 ? ey1 : ey2;
  const double *fyp = a.fby + (ey1 != 0.0 ? 0 : sy) + ic;  // + n1 k: the y-face product at (i, k)
  // ---- staging: thread stages column `lane` of rows w and w + ZS_NW
  // (clamped addresses; points outside the domain stage zero)
  const bool iok = i >= 0 && i < n1;
  unsigned soff[NS];  // offsets inside one plane (a plane holds < 2^31 points)
  bool sok[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int sj = j0 - RZ_H + w + ZS_NW * s;
    sok[s] = iok && sj >= 0 && sj < n2;
    soff[s] = (unsigned)(ic + a.ldj * min(max(sj, 0), n2 - 1));
  }
  const int cmx = min(max(a.Mx.c0[ic] - (i0 - RZ_H), 0), SW - BAND_W);
  const int crx = min(max(a.Rx.c0[ic] - (i0 - RZ_H), 0), SW - BAND_W);
  __syncthreads();  // the tables are complete before the first plane reads them

  // ---- the z queue: planes b(k) .. b(k)+4 of u and h = g - sigma u.  The
  // plane output plane k adds, and the row's y-face value at k, are loaded
  // one output plane ahead into the slot (pu, pg, pf); a slot whose plane does
  // not enter (no shift at k) is loaded anyway, with a valid plane index, and
  // dropped.
  double qu[BAND_W][NS], qh[BAND_W][NS];
  double pu[NS], pg[NS], pf = 0.0;
  auto load_slot = [&](int p, int k, double (&du)[NS], double (&dg)[NS], double &df) {
    const double *gp = a.g + a.ldk * p, *up = RES ? a.u + a.ldk * p : nullptr;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      dg[s] = gp[soff[s]];
      du[s] = RES ? up[soff[s]] : 0.0;
    }
    if (ybnd) df = fyp[(long)n1 * k];
  };
  auto enter = [&](int m, const double (&du)[NS], const double (&dg)[NS]) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      qu[m][s] = du[s];
      qh[m][s] = RES ? fma(-a.sigma, du[s], dg[s]) : dg[s];
    }
  };
  auto shifts = [&](int k) { return zs_base(k, n3) > zs_base(k - 1, n3); };
  auto entering = [&](int k) { return zs_base(k, n3) + BAND_W - 1; };  // < n3 for any k
  {
    const int b0 = zs_base(z0, n3);
    double f0 = 0.0;
#pragma unroll
    for (int m = 0; m < BAND_W; ++m) {
      load_slot(b0 + m, z0, pu, pg, m == 0 ? pf : f0);
      enter(m, pu, pg);
    }
  }
  double contrib = 0.0;
  float vmax = 0.0f;
  // r' of plane k is stored at step k+1, after that step's wait for its
  // queue plane: a wait then never waits on a store issued just before it
  const bool sto = xout && j < n2;
  float *outp = a.out + i + (long)n1 * j;  // + n1 n2 k
  float rv = 0.0f;

  for (int k = z0; k < z1; ++k) {
    const int zb = k & 1, kk = k - z0;
    if (k > z0 && shifts(k)) {
#pragma unroll
      for (int m = 0; m < BAND_W - 1; ++m)
#pragma unroll
        for (int s = 0; s < NS; ++s) {
          qu[m][s] = qu[m + 1][s];
          qh[m][s] = qh[m + 1][s];
        }
      enter(BAND_W - 1, pu, pg);
    }
    if (k > z0 && sto) outp[sz * (k - 1)] = rv;
    const double fyk = pf;  // y-face value of this output plane (boundary rows)
    if (k + 1 < z1) load_slot(entering(k + 1), k + 1, pu, pg, pf);
    // 1. z-stencils of the staged points (plane weights: uniform, scalar registers)
    double wzm[BAND_W], wzr[BAND_W];
#pragma unroll
    for (int m = 0; m < BAND_W; ++m) {
      wzm[m] = zs_uniform(PT[kk][m]);
      wzr[m] = RES ? zs_uniform(PT[kk][BAND_W + m]) : 0.0;
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      double z1v = 0.0, zuv = 0.0;
#pragma unroll
      for (int m = 0; m < BAND_W; ++m) {
        z1v = fma(wzm[m], qh[m][s], z1v);
        if (RES) {
          z1v = fma(-wzr[m], qu[m][s], z1v);
          zuv = fma(wzm[m], qu[m][s], zuv);
        }
      }
      Z1[zb][w + ZS_NW * s][lane] = sok[s] ? z1v : 0.0;
      if (RES) ZU[zb][w + ZS_NW * s][lane] = sok[s] ? zuv : 0.0;
    }
    __syncthreads();
    // 2. y-stencil of the wave's row, column = lane
    double dv = 0.0, yv = 0.0;
#pragma unroll
    for (int m = 0; m < BAND_W; ++m) {
      const double wm = WY[w][m];
      dv = fma(wm, Z1[zb][cym + m][lane], dv);
      if (RES) {
        dv = fma(-WY[w][BAND_W + m], ZU[zb][cyr + m][lane], dv);
        yv = fma(wm, ZU[zb][cym + m][lane], yv);
      }
    }
    DS[w][lane] = dv;
    if (RES) YS[w][lane] = yv;
    // the wave's own row: order its LDS writes before its reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // 3. x-stencil, boundary fold, store
    double v = 0.0;
#pragma unroll
    for (int m = 0; m < BAND_W; ++m) {
      v = fma(WX[m][lane], DS[w][cmx + m], v);
      if (RES) v = fma(-WX[BAND_W + m][lane], YS[w][crx + m], v);
    }
    if (xout && j < n2) {
      if (xedge) {
        const double ex1 = EX[0][lane], ex2 = EX[1][lane];
        if (ex1 != 0.0) v = fma(-ex1, FX[w][kk][0], v);
        if (ex2 != 0.0) v = fma(-ex2, FX[w][kk][1], v);
      }
      if (ybnd) v = fma(-eyv, fyk, v);
      const double ez1 = PT[kk][2 * BAND_W], ez2 = PT[kk][2 * BAND_W + 1];
      if (ez1 != 0.0) v = fma(-ez1, a.fbz[i + (long)n1 * j], v);
      if (ez2 != 0.0) v = fma(-ez2, a.fbz[sz + i + (long)n1 * j], v);
      rv = (float)v;
      contrib = fma(v, v, contrib);
      vmax = fmaxf(vmax, fabsf((float)v));
    }
  }
  if (sto) outp[sz * (z1 - 1)] = rv;
  // ---- norm partials, largest magnitude, slot clearing
  __shared__ double wsum[ZS_NW];
  __shared__ float wmax[ZS_NW];
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
    for (int q = 0; q < ZS_NW; ++q) t += wsum[q];
    a.part[blockIdx.x] = t;
    if (a.amax) {
      float b = wmax[0];
      for (int q = 1; q < ZS_NW; ++q) b = fmaxf(b, wmax[q]);
      atomicMax(a.amax, __float_as_uint(b));
    }
    if (a.zero && blockIdx.x == 0)
      for (int q = 0; q < a.nzero; ++q) a.zero[q] = 0u;
  }
}

}  // namespace mf
