// FP64 right-hand side and residual of the multiplied-through operator,
// z-streaming form.
//
//   b' = P g_f = Mx My Mz g - FB                        (RES = false)
//   r' = b' - B u,  B u = Mx (My Zr + Ry Zu) + Rx My Zu  (RES = true)
//   Zu = Mz u,  Zr = Rz u + sigma Zu
// (B as in residual_kernel, its z operators applied first: the 1-D operators
// of different directions commute.)  The right-hand side kernel also writes
// b' in FP64 for the residuals of the same solve, which then read b' instead
// of g: the boundary fold FB is in b' already, and the residual's stencils
// act on u alone.
//
// A block owns a 56 x 8 tile of (i,j) and a chunk of consecutive output
// planes k.  Every thread keeps, for each of its two staged points of the
// 64 x 16 haloed tile, the five planes b(k) .. b(k)+4 of its field (g, or u)
// in registers, b(k) = clamp(k-2, 0, n3-5): the z-window of every row of Mz
// and Rz (host-checked).  Per output plane: one plane enters the queue, the
// z-stencils of the staged points go to LDS (double-buffered: one block
// barrier per plane), and each wave forms its row's y-stencils over the 64
// staged columns and then its x-stencils (from the adjacent lanes by DPP,
// or through wave-local LDS rows in the x-edge tiles; no block barrier).
//
// Latency: the plane that enters the queue (and b' of the next output plane,
// or the face values of boundary rows) is loaded one output plane ahead, and
// these are the plane loop's only vector-memory loads (plus the z-face values
// on the two planes next to each z face).  Everything else the loop reads
// (weights, boundary-fold rows, x-face products of the chunk) is put in LDS
// when the block starts, and each plane's result is stored one step late,
// after the next wait: the wait at the end of a step covers only memory
// operations issued at its start (the vector-memory counter completes in
// order).  Each plane of the field is read once per chunk plus four planes of
// warm-up.
#pragma once
#include <hip/hip_runtime.h>

#include "residual.h"

namespace mf {

constexpr int ZS_NW = 8;              // wavefronts per block = output rows per tile
constexpr int ZS_TX = 64 - 2 * RZ_H;  // output columns per tile
constexpr int ZS_TY = ZS_NW;
constexpr int ZS_CZMIN = 8, ZS_CZMAX = 48;  // output planes per block
constexpr int ZS_PT = 12;             // plane table: Mz row (5), Rz row (5), e_z1, e_z2

__device__ __forceinline__ int zs_base(int k, int n3) { return min(max(k - 2, 0), n3 - 5); }

// DPP wave shifts of a double: zs_dpp<ZS_WAVE_SHR1> gives lane l the value
// of lane l-1, zs_dpp<ZS_WAVE_SHL1> that of lane l+1 (lanes past the ends
// get 0)
constexpr int ZS_WAVE_SHL1 = 0x130, ZS_WAVE_SHR1 = 0x138;
template <int CTRL>
__device__ __forceinline__ double zs_dpp(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// a wave-uniform double moved to scalar registers
__device__ __forceinline__ double zs_uniform(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readfirstlane((int)b), hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// block (64, ZS_NW); grid: ceil(n1/ZS_TX) ceil(n2/ZS_TY) ceil(n3/a.cz) blocks,
// ZS_CZMIN <= a.cz <= ZS_CZMAX.  Needs n1, n2 >= 4, n3 >= 5 and a.zq: (n3,
// ZS_PT) rows of Mz and Rz on planes b(k) .. b(k)+4, then e_z1(k), e_z2(k).
// RES = false reads a.g (full layout) and writes a.out (FP32) and, when set,
// a.out64 (FP64), both interior-contiguous; RES = true reads a.u (full
// layout) and a.bp (b' in FP64, interior-contiguous) and writes a.out.
template <bool RES>
__global__ void __launch_bounds__(64 * ZS_NW, 6) residual_zs_kernel(ResArgs a) {
  constexpr int NT = 64 * ZS_NW;
  constexpr int SW = ZS_TX + 2 * RZ_H, SH = ZS_TY + 2 * RZ_H;  // staged tile 64 x 16
  constexpr int SP = SW + 1;                                   // row pitch
  constexpr int NS = SH / ZS_NW;                               // staged points per thread
  static_assert(SW == 64 && NS * ZS_NW == SH, "one wavefront per staged row");
  constexpr int NZ = RES ? 2 : 1;                              // z-stencil fields
  __shared__ double ZS[NZ][2][SH][SP];       // RES: Zr, Zu; else Mz g
  __shared__ double DS[NZ][ZS_TY][SP];       // RES: My Zr + Ry Zu, My Zu; else My Mz g
  __shared__ double WX[2 * BAND_W][SW];      // Mx, Rx rows of each staged column
  __shared__ double WY[ZS_TY][2 * BAND_W];   // My, Ry rows of the tile's rows
  __shared__ double PT[ZS_CZMAX][ZS_PT];     // the chunk's plane table
  // boundary-fold tables, right-hand side form only (one element in the
  // residual form, whose b' holds the fold: three blocks per CU fit the LDS)
  constexpr int EXN = RES ? 1 : 2, FXR = RES ? 1 : ZS_TY, FXK = RES ? 1 : ZS_CZMAX;
  __shared__ double EX[EXN][SW];             // e_x1, e_x2 of each staged column
  __shared__ double FX[FXR][FXK][2];         // x-face products of the tile's rows (x-edge tiles)
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
  if (w < BAND_W) {  // waves 0-4: x weights, two rows each
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int m = 2 * w + h;
      WX[m][lane] = m < BAND_W ? a.Mx.w[8 * ic + m] : (RES ? a.Rx.w[8 * ic + m - BAND_W] : 0.0);
    }
  } else if (w == BAND_W) {
    if constexpr (!RES) {
      EX[0][lane] = a.ex[ic];
      EX[1][lane] = a.ex[n1 + ic];
    }
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
  const bool xedge = !RES && (i0 < 2 || i0 + ZS_TX > n1 - 2);  // columns with e_x != 0
  if constexpr (!RES) {
    if (xedge)
      for (int e = tid; e < ZS_TY * nz * 2; e += NT) {
        const int r = e / (2 * nz), kk = (e / 2) % nz, side = e % 2;
        const int jj = min(j0 + r, n2 - 1);
        FX[r][kk][side] = a.fbx[side * sx + jj + (long)n2 * (z0 + kk)];
      }
  }
  // ---- the wave's row: window starts (staged rows), boundary fold
  const int j = j0 + w, jc = min(j, n2 - 1);
  const int cym = a.My.c0[jc] - (j0 - RZ_H);
  const int cyr = RES ? a.Ry.c0[jc] - (j0 - RZ_H) : cym;
  // e_y of the row (nonzero on rows 0, 1 and n2-2, n2-1 only: one side)
  const double ey1 = RES ? 0.0 : a.ey[jc], ey2 = RES ? 0.0 : a.ey[n2 + jc];
  const bool ybnd = !RES && j < n2 && (ey1 != 0.0 || ey2 != 0.0);
  const double eyv = ey1 != 0.0 ? ey1 : ey2;
  const double *fyp = RES ? nullptr : a.fby + (ey1 != 0.0 ? 0 : sy) + ic;  // + n1 k: y-face product
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
  const long orow = ic + (long)n1 * jc;  // interior-contiguous offset of (i, j) (+ n1 n2 k)
  // every output column's x-windows centred (all but the x-edge tiles): the
  // residual's x-stencils take their neighbours' y-stencil results from the
  // adjacent lanes (DPP), with no LDS round trip (wave-uniform; measured at
  // 256^3: 2.5% faster for the residual, 25% slower for the right-hand side,
  // which keeps the LDS rows)
  const bool lcent = !xout || (cmx == lane - 2 && crx == lane - 2);
  const bool cent = RES && __builtin_amdgcn_readfirstlane(__all(lcent)) != 0;
  __syncthreads();  // the tables are complete before the first plane reads them

  // ---- the z queue: planes b(k) .. b(k)+4 of the field.  The plane output
  // plane k adds, and the value the output row needs at k (b' for RES, the
  // y-face product of boundary rows otherwise), are loaded one output plane
  // ahead into the slot (pq, px); a slot whose plane does not enter (no
  // shift at k) is loaded anyway, with a valid plane index, and dropped.
  const double *fld = RES ? a.u : a.g;
  double q[BAND_W][NS];
  double pq[NS], px = 0.0;
  auto load_slot = [&](int p, int k, double (&dq)[NS], double &dx) {
    const double *fp = fld + a.ldk * p;
#pragma unroll
    for (int s = 0; s < NS; ++s) dq[s] = fp[soff[s]];
    if (RES) dx = a.bp[orow + sz * k];
    else if (ybnd) dx = fyp[(long)n1 * k];
  };
  auto shifts = [&](int k) { return zs_base(k, n3) > zs_base(k - 1, n3); };
  auto entering = [&](int k) { return zs_base(k, n3) + BAND_W - 1; };  // < n3 for any k
  {
    const int b0 = zs_base(z0, n3);
    double x0 = 0.0;
#pragma unroll
    for (int m = 0; m < BAND_W; ++m) {
      load_slot(b0 + m, z0, pq, m == 0 ? px : x0);
#pragma unroll
      for (int s = 0; s < NS; ++s) q[m][s] = pq[s];
    }
  }
  double contrib = 0.0;
  float vmax = 0.0f;
  // the result of plane k is stored at step k+1, after that step's wait for
  // its queue plane: a wait then never waits on a store issued just before it
  const bool sto = xout && j < n2;
  float *outp = a.out + orow;  // + n1 n2 k
  double *outp64 = RES ? nullptr : a.out64 ? a.out64 + orow : nullptr;
  double rv = 0.0;

  for (int k = z0; k < z1; ++k) {
    const int zb = k & 1, kk = k - z0;
    if (k > z0 && shifts(k)) {
#pragma unroll
      for (int m = 0; m < BAND_W - 1; ++m)
#pragma unroll
        for (int s = 0; s < NS; ++s) q[m][s] = q[m + 1][s];
#pragma unroll
      for (int s = 0; s < NS; ++s) q[BAND_W - 1][s] = pq[s];
    }
    if (k > z0 && sto) {
      outp[sz * (k - 1)] = (float)rv;
      if (outp64) outp64[sz * (k - 1)] = rv;
    }
    const double xk = px;  // b' (RES) or the y-face value of this output plane
    if (k + 1 < z1) load_slot(entering(k + 1), k + 1, pq, px);
    // 1. z-stencils of the staged points (plane weights: uniform, scalar registers)
    {
      double wzm[BAND_W], wzr[BAND_W];
#pragma unroll
      for (int m = 0; m < BAND_W; ++m) {
        wzm[m] = zs_uniform(PT[kk][m]);
        wzr[m] = RES ? zs_uniform(PT[kk][BAND_W + m]) : 0.0;
      }
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        double zm = 0.0, zr = 0.0;
#pragma unroll
        for (int m = 0; m < BAND_W; ++m) {
          zm = fma(wzm[m], q[m][s], zm);
          if (RES) zr = fma(wzr[m], q[m][s], zr);
        }
        if (RES) {
          ZS[0][zb][w + ZS_NW * s][lane] = sok[s] ? fma(a.sigma, zm, zr) : 0.0;  // Zr
          ZS[NZ - 1][zb][w + ZS_NW * s][lane] = sok[s] ? zm : 0.0;              // Zu
        } else {
          ZS[0][zb][w + ZS_NW * s][lane] = sok[s] ? zm : 0.0;
        }
      }
    }
    __syncthreads();
    // 2. y-stencils of the wave's row, column = lane
    double dv = 0.0, yv = 0.0;
#pragma unroll
    for (int m = 0; m < BAND_W; ++m) {
      const double wm = WY[w][m];
      dv = fma(wm, ZS[0][zb][cym + m][lane], dv);
      if (RES) {
        dv = fma(WY[w][BAND_W + m], ZS[NZ - 1][zb][cyr + m][lane], dv);
        yv = fma(wm, ZS[NZ - 1][zb][cym + m][lane], yv);
      }
    }
    // 3. x-stencils, boundary fold (b') or the difference (r')
    double v = 0.0;
    if (cent) {
      const double d1 = zs_dpp<ZS_WAVE_SHR1>(dv), d3 = zs_dpp<ZS_WAVE_SHL1>(dv);
      const double dd[BAND_W] = {zs_dpp<ZS_WAVE_SHR1>(d1), d1, dv, d3, zs_dpp<ZS_WAVE_SHL1>(d3)};
#pragma unroll
      for (int m = 0; m < BAND_W; ++m) v = fma(WX[m][lane], dd[m], v);
      if (RES) {
        const double y1 = zs_dpp<ZS_WAVE_SHR1>(yv), y3 = zs_dpp<ZS_WAVE_SHL1>(yv);
        const double yy[BAND_W] = {zs_dpp<ZS_WAVE_SHR1>(y1), y1, yv, y3, zs_dpp<ZS_WAVE_SHL1>(y3)};
#pragma unroll
        for (int m = 0; m < BAND_W; ++m) v = fma(WX[BAND_W + m][lane], yy[m], v);
      }
    } else {
      DS[0][w][lane] = dv;
      if (RES) DS[NZ - 1][w][lane] = yv;
      // the wave's own row: order its LDS writes before its reads
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int m = 0; m < BAND_W; ++m) {
        v = fma(WX[m][lane], DS[0][w][cmx + m], v);
        if (RES) v = fma(WX[BAND_W + m][lane], DS[NZ - 1][w][crx + m], v);
      }
    }
    if (sto) {
      if (RES) {
        v = xk - v;
      } else {
        if constexpr (!RES) {
          if (xedge) {
            const double ex1 = EX[0][lane], ex2 = EX[1][lane];
            if (ex1 != 0.0) v = fma(-ex1, FX[w][kk][0], v);
            if (ex2 != 0.0) v = fma(-ex2, FX[w][kk][1], v);
          }
        }
        if (ybnd) v = fma(-eyv, xk, v);
        const double ez1 = PT[kk][2 * BAND_W], ez2 = PT[kk][2 * BAND_W + 1];
        if (ez1 != 0.0) v = fma(-ez1, a.fbz[i + (long)n1 * j], v);
        if (ez2 != 0.0) v = fma(-ez2, a.fbz[sz + i + (long)n1 * j], v);
      }
      rv = v;
      contrib = fma(v, v, contrib);
      vmax = fmaxf(vmax, fabsf((float)v));
    }
  }
  if (sto) {
    outp[sz * (z1 - 1)] = (float)rv;
    if (outp64) outp64[sz * (z1 - 1)] = rv;
  }
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
    for (int q2 = 0; q2 < ZS_NW; ++q2) t += wsum[q2];
    a.part[blockIdx.x] = t;
    if (a.amax) {
      float b = wmax[0];
      for (int q2 = 1; q2 < ZS_NW; ++q2) b = fmaxf(b, wmax[q2]);
      atomicMax(a.amax, __float_as_uint(b));
    }
    if (a.zero && blockIdx.x == 0)
      for (int q2 = 0; q2 < a.nzero; ++q2) a.zero[q2] = 0u;
  }
}

}  // namespace mf
