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
//
// The coefficients are wave-uniform: they are read through the constant
// address space, so they arrive by scalar loads into SGPRs (one 16-double
// row per point) and cost neither LDS bandwidth nor vector registers.
#pragma once
#include <hip/hip_runtime.h>

namespace mf {

constexpr int SEG_LEN = 32;   // segment length (all segments but the last)
constexpr int SEG_MAX = 38;   // longest last segment (it absorbs a short tail)
constexpr int SEG_KMAX = 16;  // segments per line (wavefronts per block)
constexpr int SEG_ROW = 16;   // coefficient doubles per point

// Segmentation of a line of n points: K segments starting at 32 j, the last
// running to n and at least 7 rows long (so that it holds the 7-point closure
// inputs and row n-4).
__host__ __device__ inline int seg_count(int n) {
  int K = (n + SEG_LEN - 1) / SEG_LEN;
  if (K > 1 && n - SEG_LEN * (K - 1) < 7) --K;
  return K;
}

// dynamic LDS doubles of a block running a scheme of n points: the forward
// and backward exchange arrays, [K][2][64] each
__host__ __device__ inline int seg_lds_doubles(int n) { return 4 * 64 * seg_count(n); }

// Row layout of the segment table (SEG_ROW doubles per point i, two 64-byte
// halves: the forward sweep of an interior row reads the first, the backward
// sweep the second):
//   0..4 dr(i, 0..4) | 5 p | 6 f | 7 dr(i, 5) || 8 dr(i, 6) | 9 e | 10 d | 11 1/c | 12..15 0
// The homogeneous responses sit in two arrays of their own, (A_i, B_i) and
// (C_i, D_i) pairs: the compiler merges adjacent scalar loads, and a fix-up
// coefficient merged into a load of the other sweep would be held in SGPRs
// from one sweep to the next.
enum { SR_P = 5, SR_F = 6, SR_E = 9, SR_D = 10, SR_RC = 11 };
__host__ __device__ constexpr int sr_dr(int m) { return m < 5 ? m : m + 2; }

struct SegTab {
  const double *rows;  // n x SEG_ROW
  const double *ab;    // n x 2: A, B
  const double *cd;    // n x 2: C, D
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

typedef const __attribute__((address_space(4))) double *seg_cptr;

// One wavefront's segment [s0, s0 + len).  LEN rows are unrolled; DYN: the
// length is the runtime len <= LEN (rows beyond it are skipped), else len =
// LEN.  HEAD: segment 0 (7-point rows 0 and 1, the alpha term of row 0);
// TAIL: the last segment (7-point rows n-2 and n-1, the beta term of row n-1).
// Ef, Eb: the block's exchange arrays; w, K: this segment and the count.
template <int LEN, bool HEAD, bool TAIL, bool DYN>
__device__ __forceinline__ void seg_wave(const double *__restrict__ xl, double *__restrict__ ol,
                                         long s, bool act, const SegTab &T, int w, int K,
                                         int s0, int len, double *Ef, double *Eb) {
  const int n = T.n;
  const int lane = threadIdx.x & 63;
  // row s0 + k sits at a compile-time offset from Rs (scalar-load immediates)
  const seg_cptr Rs = (seg_cptr)T.rows + (long)s0 * SEG_ROW;
  const seg_cptr AB = (seg_cptr)T.ab, CD = (seg_cptr)T.cd;
  const seg_cptr ABs = AB + 2L * s0, CDs = CD + 2L * s0;
  auto on = [&](int k) { return !DYN || k < len; };
  // tail rows: k = len - 2 (row n-2) and k = len - 1 (row n-1)
  auto is_n2 = [&](int k) { return TAIL && (DYN ? k == len - 2 : k == LEN - 2); };
  auto is_n1 = [&](int k) { return TAIL && (DYN ? k == len - 1 : k == LEN - 1); };
  // inputs V[k] = x[s0 - 2 + k]; the two before the line (segment 0) and the
  // two after it (last segment of exactly LEN rows) are never read; a
  // runtime-length segment clamps into the line
  double V[LEN + 4];
#pragma unroll
  for (int k = 0; k < LEN + 4; ++k) {
    if ((HEAD && k < 2) || (TAIL && !DYN && k >= LEN + 2)) {
      V[k] = 0.0;
      continue;
    }
    int q = s0 - 2 + k;
    if (DYN) q = q > n - 1 ? n - 1 : q;
    V[k] = xl[(long)q * s];
  }
  // forward particular solution from zero incoming values (segment 0 starts
  // at the true start)
  double Y[LEN];
  double ym1 = 0.0, ym2 = 0.0, ym3 = 0.0, yl0 = 0.0, yl1 = 0.0;
#pragma unroll
  for (int k = 0; k < LEN; ++k) {
    if (on(k)) {
      const seg_cptr c = Rs + k * SEG_ROW;
      double rr = 0.0, yi;
      if (HEAD && k < 2) {  // 7-point head rows on x[0..6] = V[2..8]
#pragma unroll
        for (int m = 0; m < 7; ++m) rr = fma(c[sr_dr(m)], V[2 + m], rr);
        yi = (k == 0) ? rr : rr - c[SR_F] * ym1;
      } else if (k >= 4 && (is_n2(k) || is_n1(k))) {
        // 7-point tail rows on x[n-7..n-1] = V[k-3..k+3] (row n-2) or V[k-4..k+2] (row n-1)
        if (is_n2(k)) {
#pragma unroll
          for (int m = 0; m < 7; ++m) rr = fma(c[sr_dr(m)], V[k - 3 + m], rr);
        } else {
#pragma unroll
          for (int m = 0; m < 7; ++m) rr = fma(c[sr_dr(m)], V[k - 4 + m], rr);
        }
        yi = rr - c[SR_P] * ym2 - c[SR_F] * ym1;
        if (is_n1(k)) yi -= T.beta * ym3;
      } else {
        rr = c[0] * V[k] + c[1] * V[k + 1] + c[2] * V[k + 2] + c[3] * V[k + 3] + c[4] * V[k + 4];
        yi = rr - c[SR_P] * ym2 - c[SR_F] * ym1;
      }
      Y[k] = yi;
      ym3 = ym2; ym2 = ym1; ym1 = yi;
      if (DYN ? k == len - 2 : k == LEN - 2) yl0 = yi;
      if (DYN ? k == len - 1 : k == LEN - 1) yl1 = yi;
    }
    // keep the scalar coefficient loads near their rows (hoisting them all
    // to the top exhausts the SGPRs)
    if (k % 4 == 3) __builtin_amdgcn_sched_barrier(0);
  }
  // publish the segment's last two particular values
  Ef[(w * 2 + 0) * 64 + lane] = yl0;
  Ef[(w * 2 + 1) * 64 + lane] = yl1;
  __syncthreads();
  if (!HEAD) {
    // incoming (y_{s0-2}, y_{s0-1}): segment 0 ends on its own values, later
    // segments map theirs through their homogeneous responses
    double v0 = Ef[lane], v1 = Ef[64 + lane];
    for (int j = 1; j < w; ++j) {
      const seg_cptr c = AB + 2L * (SEG_LEN * (j + 1) - 2);  // rows e_j-2, e_j-1
      const double p0 = Ef[(j * 2 + 0) * 64 + lane], p1 = Ef[(j * 2 + 1) * 64 + lane];
      const double n0 = p0 + c[0] * v0 + c[1] * v1;
      const double n1 = p1 + c[2] * v0 + c[3] * v1;
      v0 = n0; v1 = n1;
    }
#pragma unroll
    for (int k = 0; k < LEN; ++k) {
      const seg_cptr c = ABs + 2 * k;
      if (on(k)) Y[k] = Y[k] + c[0] * v0 + c[1] * v1;
      if (k % 8 == 7) __builtin_amdgcn_sched_barrier(0);
    }
  }
  // backward particular solution from zero incoming values (the last segment
  // ends at the true end), in place in Y
  double xp1 = 0.0, xp2 = 0.0, x1 = 0.0, x2 = 0.0, x3 = 0.0;
#pragma unroll
  for (int k = LEN - 1; k >= 0; --k) {
    if (on(k)) {
      const seg_cptr c = Rs + k * SEG_ROW;
      double xi;
      if (is_n1(k)) {
        xi = Y[k] * c[SR_RC];
      } else if (HEAD && k == 0) {
        xi = (Y[0] - c[SR_E] * x2 - c[SR_D] * x1 - T.alpha * x3) * c[SR_RC];
      } else {
        xi = (Y[k] - c[SR_E] * xp2 - c[SR_D] * xp1) * c[SR_RC];
      }
      Y[k] = xi;
      xp2 = xp1; xp1 = xi;
      if (HEAD && k == 3) x3 = xi;
      if (HEAD && k == 2) x2 = xi;
      if (HEAD && k == 1) x1 = xi;
    }
    if (k % 4 == 0) __builtin_amdgcn_sched_barrier(0);
  }
  Eb[(w * 2 + 0) * 64 + lane] = Y[0];
  Eb[(w * 2 + 1) * 64 + lane] = Y[1];
  __syncthreads();
  if (!TAIL) {
    // incoming (x_{e0}, x_{e0+1}) from the segments after this one
    double v0 = Eb[((K - 1) * 2 + 0) * 64 + lane], v1 = Eb[((K - 1) * 2 + 1) * 64 + lane];
    for (int j = K - 2; j > w; --j) {
      const seg_cptr c = CD + 2L * SEG_LEN * j;  // rows s_j, s_j+1
      const double q0 = Eb[(j * 2 + 0) * 64 + lane], q1 = Eb[(j * 2 + 1) * 64 + lane];
      const double n0 = q0 + c[0] * v0 + c[1] * v1;
      const double n1 = q1 + c[2] * v0 + c[3] * v1;
      v0 = n0; v1 = n1;
    }
#pragma unroll
    for (int k = 0; k < LEN; ++k) {
      const seg_cptr c = CDs + 2 * k;
      if (on(k)) Y[k] = Y[k] + c[0] * v0 + c[1] * v1;
      if (k % 8 == 7) __builtin_amdgcn_sched_barrier(0);
    }
  }
  if (act) {
    // a fresh copy of the stride, opaque to the compiler, so that the store
    // offsets are computed here rather than shared with the load offsets and
    // held in SGPRs through the whole segment
    long so = s;
    asm volatile("" : "+s"(so));
    double *po = ol + (long)s0 * so;
#pragma unroll
    for (int k = 0; k < LEN; ++k)
      if (on(k)) po[(long)k * so] = Y[k];
  }
}

// Solve the scheme on 64 lines per block, wave w = segment w.
// x, out: the field and result; lines per SegLines (one family of lines).
// lds: forward exchange [K][2][64] | backward [K][2][64].
// Every wave of the block passes the same two barriers.
__device__ __forceinline__ void seg_solve(const double *__restrict__ x, double *__restrict__ out,
                                          const SegTab &T, const SegLines &L, int lblock,
                                          double *lds) {
  const int n = T.n;
  const int K = seg_count(n);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (w >= K) {
    // waves beyond K (blocks sized for a longer scheme) only join the barriers
    __syncthreads();
    __syncthreads();
    return;
  }
  double *Ef = lds;
  double *Eb = Ef + 128 * K;
  // this lane's line
  int l = lblock * 64 + (int)(threadIdx.x & 63);
  const bool act = l < L.nlines;
  l = act ? l : L.nlines - 1;
  const long base = (long)(L.a0 + l % L.na) * L.sa + (long)(L.b0 + l / L.na) * L.sb;
  const double *xl = x + base;
  double *ol = out + base;
  const int s0 = SEG_LEN * w;
  const int len = (w == K - 1) ? n - s0 : SEG_LEN;
  if (K == 1)
    seg_wave<SEG_MAX, true, true, true>(xl, ol, L.s, act, T, w, K, s0, len, Ef, Eb);
  else if (w == 0)
    seg_wave<SEG_LEN, true, false, false>(xl, ol, L.s, act, T, w, K, s0, len, Ef, Eb);
  else if (w < K - 1)
    seg_wave<SEG_LEN, false, false, false>(xl, ol, L.s, act, T, w, K, s0, len, Ef, Eb);
  else if (len == SEG_LEN)
    seg_wave<SEG_LEN, false, true, false>(xl, ol, L.s, act, T, w, K, s0, len, Ef, Eb);
  else
    seg_wave<SEG_MAX, false, true, true>(xl, ol, L.s, act, T, w, K, s0, len, Ef, Eb);
}

// Segment tables of a scheme of n points (host side) from its dr (n,7)
// column-major and packed p, f, d, e, 1/c: rows[i*SEG_ROW + ...] as SR_*, and
// the homogeneous responses ab = (A_i, B_i) (forward, zero on segment 0) and
// cd = (C_i, D_i) (backward, zero on the last segment).
inline void seg_tables(int n, const double *dr, const double *p, const double *f,
                       const double *d, const double *e, const double *rc, double alpha,
                       double beta, double *rows, double *ab, double *cd) {
  const int K = seg_count(n);
  for (int i = 0; i < n; ++i) {
    double *r = rows + (size_t)i * SEG_ROW;
    for (int m = 0; m < 7; ++m) r[sr_dr(m)] = dr[i + (size_t)n * m];
    r[SR_P] = p[i];
    r[SR_F] = f[i];
    r[SR_E] = e[i];
    r[SR_D] = d[i];
    r[SR_RC] = rc[i];
    for (int m = 12; m < SEG_ROW; ++m) r[m] = 0.0;
    ab[2 * (size_t)i] = ab[2 * (size_t)i + 1] = cd[2 * (size_t)i] = cd[2 * (size_t)i + 1] = 0.0;
  }
  // slot 0..3: A, B, C, D
  auto at = [&](int i, int slot) -> double & {
    return (slot < 2 ? ab : cd)[2 * (size_t)i + (slot & 1)];
  };
  for (int w = 0; w < K; ++w) {
    const int s0 = SEG_LEN * w, e0 = (w == K - 1) ? n : s0 + SEG_LEN;
    if (w > 0) {
      for (int u = 0; u < 2; ++u) {
        const int slot = u;  // A or B
        double m2 = u ? 0.0 : 1.0, m1 = u ? 1.0 : 0.0;  // y_{s0-2}, y_{s0-1}
        for (int i = s0; i < e0; ++i) {
          double yi = -p[i] * m2 - f[i] * m1;
          if (i == n - 1) yi -= beta * at(n - 4, slot);
          at(i, slot) = yi;
          m2 = m1;
          m1 = yi;
        }
      }
    }
    if (w < K - 1) {
      for (int u = 0; u < 2; ++u) {
        const int slot = 2 + u;  // C or D
        double p1 = u ? 0.0 : 1.0, p2 = u ? 1.0 : 0.0;  // x_{e0}, x_{e0+1}
        for (int i = e0 - 1; i >= s0; --i) {
          double xi = -e[i] * p2 - d[i] * p1;
          if (i == 0) xi -= alpha * at(3, slot);
          xi *= rc[i];
          at(i, slot) = xi;
          p2 = p1;
          p1 = xi;
        }
      }
    }
  }
}

}  // namespace mf
