// Compact-derivative line solves split into segments, one wavefront per
// segment: 64 lines per block, K segments of ~32 rows per line.
//
// The pentalu-factored solve (class_derivatives.f90:425-466) is two
// second-order linear recurrences per line:
//   forward   y_i = r_i - p_i y_{i-2} - f_i y_{i-1}   (- beta y_{n-4} on row n-1)
//   backward  x_i = (y_i - e_i x_{i+2} - d_i x_{i+1}) / c_i   (- alpha x_3 on row 0)
// On a segment [s, e) the forward values are affine in the two incoming
// values: y_i = P_i + A_i y_{s-2} + B_i y_{s-1}, with P the recurrence run from
// zero and A, B its homogeneous responses to unit incoming values, which
// depend on the coefficients only and are tabulated at set-up (likewise
// x_i = Q_i + C_i x_e + D_i x_{e+1} backward).  Each wave loads its
// segment's inputs at once, runs P in registers, publishes its two boundary
// values in LDS, chains the 2x2 maps of the segments before it to get its
// incoming values, and fixes P up; then the same backward.  The result is the
// sequential solve re-associated (rounding-level differences only), with K
// times the wavefronts of one-lane-per-line and every load of a segment in
// flight together.
#pragma once
#include <hip/hip_runtime.h>

namespace mf {

constexpr int SEG_LEN = 32;   // nominal segment length
constexpr int SEG_MAX = 38;   // longest segment (the last absorbs a short tail)
constexpr int SEG_KMAX = 10;  // segments per line (wavefronts per block)
constexpr int SEG_TAB = 17;   // coefficient doubles per point

// Segmentation of a line of n points: K segments starting at 32 j, the last
// running to n and at least 7 rows long (so that it holds the 7-point closure
// inputs and row n-4).
__host__ __device__ inline int seg_count(int n) {
  int K = (n + SEG_LEN - 1) / SEG_LEN;
  if (K > 1 && n - SEG_LEN * (K - 1) < 7) --K;
  return K;
}

// dynamic LDS doubles of a block running a scheme of n points
__host__ __device__ inline int seg_lds_doubles(int n) { return SEG_TAB * n + 4 * 64 * seg_count(n); }

// packed coefficient table of one scheme (17 n doubles):
// dr (7n, row-major) | p | f | d | e | c | 1/c | A | B | C | D
struct SegTab {
  const double *base;
  double alpha, beta;
  int n;
};

struct SegLines {
  // line l of the launch starts at (a0 + l % na) * sa + (b0 + l / na) * sb;
  // stride s between its points
  long sa, sb, s;
  int a0, na, b0, nb;
  int nlines;
};

// Solve the scheme on 64 lines per block, wave w = segment w.
// x, out: the field and result; lines per SegLines (one family of lines).
// lds: coefficients (17 n) | forward exchange [K][2][64] | backward [K][2][64].
// Every wave of the block passes the same three barriers.
__device__ __forceinline__ void seg_solve(const double *__restrict__ x, double *__restrict__ out,
                                          const SegTab &T, const SegLines &L, int lblock,
                                          double *lds) {
  const int n = T.n;
  const int K = seg_count(n);
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nt = blockDim.x;
  for (int q = threadIdx.x; q < SEG_TAB * n; q += nt) lds[q] = T.base[q];
  double *Ef = lds + SEG_TAB * n;
  double *Eb = Ef + 128 * K;
  __syncthreads();
  const double *dr = lds, *Tp = lds + 7 * n, *Tf = Tp + n, *Td = Tf + n, *Te = Td + n;
  const double *Trc = Te + 2 * n, *TA = Trc + n, *TB = TA + n, *TC = TB + n, *TD = TC + n;
  // this lane's line
  int l = lblock * 64 + lane;
  const bool act = l < L.nlines;
  l = act ? l : L.nlines - 1;
  const long base = (long)(L.a0 + l % L.na) * L.sa + (long)(L.b0 + l / L.na) * L.sb;
  const double *xl = x + base;
  double *ol = out + base;
  const long s = L.s;
  // segment rows [s0, e0)
  const int s0 = SEG_LEN * w;
  const int e0 = (w == K - 1) ? n : s0 + SEG_LEN;
  if (w < K) {
    // inputs V[k] = x[s0 - 2 + k], clamped into the line
    double V[SEG_MAX + 4];
#pragma unroll
    for (int k = 0; k < SEG_MAX + 4; ++k) {
      int q = s0 - 2 + k;
      q = q < 0 ? 0 : (q > n - 1 ? n - 1 : q);
      V[k] = xl[(long)q * s];
    }
    // forward particular solution from zero incoming values (segment 0 starts
    // at the true start)
    double Y[SEG_MAX];
    double ym1 = 0.0, ym2 = 0.0, ym3 = 0.0, yl0 = 0.0, yl1 = 0.0;
#pragma unroll
    for (int k = 0; k < SEG_MAX; ++k) {
      const int i = s0 + k;
      if (i < e0) {
        const double *c = dr + 7 * i;
        double rr = 0.0, yi;
        if (k < 2 && i <= 1) {  // 7-point head rows on x[0..6] = V[2..8] (s0 = 0)
#pragma unroll
          for (int m = 0; m < 7; ++m) rr = fma(c[m], V[2 + m], rr);
          yi = (i == 0) ? rr : rr - Tf[1] * ym1;
        } else if (k >= 4 && i >= n - 2) {  // 7-point tail rows on x[n-7..n-1]
          // x[n-7+m] = V[k-3+m] on row n-2 and V[k-4+m] on row n-1
          if (i == n - 2) {
#pragma unroll
            for (int m = 0; m < 7; ++m) rr = fma(c[m], V[k - 3 + m], rr);
          } else {
#pragma unroll
            for (int m = 0; m < 7; ++m) rr = fma(c[m], V[k - 4 + m], rr);
          }
          yi = rr - Tp[i] * ym2 - Tf[i] * ym1;
          if (i == n - 1) yi -= T.beta * ym3;
        } else {
          rr = c[0] * V[k] + c[1] * V[k + 1] + c[2] * V[k + 2] + c[3] * V[k + 3] + c[4] * V[k + 4];
          yi = rr - Tp[i] * ym2 - Tf[i] * ym1;
        }
        Y[k] = yi;
        ym3 = ym2; ym2 = ym1; ym1 = yi;
        if (i == e0 - 2) yl0 = yi;
        if (i == e0 - 1) yl1 = yi;
      }
    }
    // publish the segment's last two particular values
    Ef[(w * 2 + 0) * 64 + lane] = yl0;
    Ef[(w * 2 + 1) * 64 + lane] = yl1;
    __syncthreads();
    if (w > 0) {
      // incoming (y_{s0-2}, y_{s0-1}): segment 0 ends on its own values,
      // later segments map theirs through their homogeneous responses
      double v0 = Ef[lane], v1 = Ef[64 + lane];
      for (int j = 1; j < w; ++j) {
        const int ej = SEG_LEN * (j + 1);  // end of segment j (not the last one)
        const double p0 = Ef[(j * 2 + 0) * 64 + lane], p1 = Ef[(j * 2 + 1) * 64 + lane];
        const double n0 = p0 + TA[ej - 2] * v0 + TB[ej - 2] * v1;
        const double n1 = p1 + TA[ej - 1] * v0 + TB[ej - 1] * v1;
        v0 = n0; v1 = n1;
      }
#pragma unroll
      for (int k = 0; k < SEG_MAX; ++k) {
        const int i = s0 + k;
        if (i < e0) Y[k] = Y[k] + TA[i] * v0 + TB[i] * v1;
      }
    }
    // backward particular solution from zero incoming values (the last
    // segment ends at the true end), in place in Y
    double xp1 = 0.0, xp2 = 0.0, x1 = 0.0, x2 = 0.0, x3 = 0.0;
#pragma unroll
    for (int k = SEG_MAX - 1; k >= 0; --k) {
      const int i = s0 + k;
      if (i < e0) {
        double xi;
        if (i == n - 1) {
          xi = Y[k] * Trc[i];
        } else if (k == 0 && i == 0) {
          xi = (Y[0] - Te[0] * x2 - Td[0] * x1 - T.alpha * x3) * Trc[0];
        } else {
          xi = (Y[k] - Te[i] * xp2 - Td[i] * xp1) * Trc[i];
        }
        Y[k] = xi;
        xp2 = xp1; xp1 = xi;
        if (k == 3) x3 = xi;
        if (k == 2) x2 = xi;
        if (k == 1) x1 = xi;
      }
    }
    Eb[(w * 2 + 0) * 64 + lane] = Y[0];
    Eb[(w * 2 + 1) * 64 + lane] = Y[1];
    __syncthreads();
    if (w < K - 1) {
      // incoming (x_{e0}, x_{e0+1}) from the segments after this one
      double v0 = Eb[((K - 1) * 2 + 0) * 64 + lane], v1 = Eb[((K - 1) * 2 + 1) * 64 + lane];
      for (int j = K - 2; j > w; --j) {
        const int sj = SEG_LEN * j;
        const double q0 = Eb[(j * 2 + 0) * 64 + lane], q1 = Eb[(j * 2 + 1) * 64 + lane];
        const double n0 = q0 + TC[sj] * v0 + TD[sj] * v1;
        const double n1 = q1 + TC[sj + 1] * v0 + TD[sj + 1] * v1;
        v0 = n0; v1 = n1;
      }
#pragma unroll
      for (int k = 0; k < SEG_MAX; ++k) {
        const int i = s0 + k;
        if (i < e0) Y[k] = Y[k] + TC[i] * v0 + TD[i] * v1;
      }
    }
    if (act) {
#pragma unroll
      for (int k = 0; k < SEG_MAX; ++k) {
        const int i = s0 + k;
        if (i < e0) ol[(long)i * s] = Y[k];
      }
    }
  } else {
    // waves beyond K (blocks sized for a longer scheme) only join the barriers
    __syncthreads();
    __syncthreads();
  }
}

// Homogeneous responses of the two recurrences on each segment (host side),
// from the packed p, f, d, e, 1/c of a scheme of n points; A, B, C, D are n
// each, zero where unused (forward: segment 0; backward: the last segment).
inline void seg_tables(int n, const double *p, const double *f, const double *d,
                       const double *e, const double *rc, double alpha, double beta, double *A,
                       double *B, double *C, double *D) {
  const int K = seg_count(n);
  for (int i = 0; i < n; ++i) A[i] = B[i] = C[i] = D[i] = 0.0;
  for (int w = 0; w < K; ++w) {
    const int s0 = SEG_LEN * w, e0 = (w == K - 1) ? n : s0 + SEG_LEN;
    if (w > 0) {
      for (int u = 0; u < 2; ++u) {
        double *Y = u ? B : A;
        double m2 = u ? 0.0 : 1.0, m1 = u ? 1.0 : 0.0;  // y_{s0-2}, y_{s0-1}
        for (int i = s0; i < e0; ++i) {
          double yi = -p[i] * m2 - f[i] * m1;
          if (i == n - 1) yi -= beta * Y[n - 4];
          Y[i] = yi;
          m2 = m1;
          m1 = yi;
        }
      }
    }
    if (w < K - 1) {
      for (int u = 0; u < 2; ++u) {
        double *X = u ? D : C;
        double p1 = u ? 0.0 : 1.0, p2 = u ? 1.0 : 0.0;  // x_{e0}, x_{e0+1}
        for (int i = e0 - 1; i >= s0; --i) {
          double xi = -e[i] * p2 - d[i] * p1;
          if (i == 0) xi -= alpha * X[3];
          xi *= rc[i];
          X[i] = xi;
          p2 = p1;
          p1 = xi;
        }
      }
    }
  }
}

}  // namespace mf
