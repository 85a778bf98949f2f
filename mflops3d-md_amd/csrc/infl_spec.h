// Influence blocks by spectral contraction (md_influence_matrix_set,
// class_field.f90:1396-1510, computed without one solve per impulse).
//
// The reference fills the influence matrix column by column: a unit
// Dirichlet value at one interface point, a full 3-D solve, the normal
// derivatives on the interface faces.  With the diagonalisation solver the
// response is U = (Vx (x) Vy (x) Vz) R (Wx (x) Wy (x) Wz) F, R the inverse
// eigenvalue sums, and for an impulse F is one line: the fold vector of the
// impulse face along its normal axis s, unit vectors along the two face
// axes.  The output is a row vector w (the boundary derivative row, with
// the Neumann extrapolation of the line folded in) along the output face's
// normal t and the identity along its two face axes.  So every block
// (impulse face s, output face t) is a product of one-axis factors:
//   s = t (face axes b, c):
//     D[(ob,oc),(cb,cc)] = sum_{be,ga} Vb(ob,be) Wb(be,cb) Vc(oc,ga) Wc(ga,cc) G(be,ga)
//     G(be,ga) = sum_al what(al) ahat(al) R(al,be,ga)      [+ a diagonal term]
//   s != t (u the third axis; the impulse face spans t, u, the output s, u):
//     D[(os,ou),(ct,cu)] = sum_{as,au} Vs(os,as) ahat(as) Vu(ou,au) Wu(au,cu) H(as,au,ct)
//     H(as,au,ct) = sum_at R(as,at,au) what(at) Wt(at,ct)
// with ahat = Ws (-fold) and what = w^T Vt.  G and H are O(n^3)/O(n^4)
// kernels, one O(n^4) GEMM contracts the first face axis, and the block
// itself is one O(n^5) GEMM per face pair against the tables
//   T[(o,c), a] = V(o,a) W(a,c)   and   TT[a, (o,c)] (the same, transposed),
// whose epilogue writes the block into the subdomain's dense matrix.  All
// those GEMMs of all subdomains run in one persistent grouped launch.  An
// impulse batch costs O(n^6) per subdomain; this costs O(n^5).
#pragma once
#include <hip/hip_runtime.h>

#include "gemm.h"

namespace mf {

// R = 1 / (lx + ly + lz + sigma), 0 where max(|sigma|, |l|) < 1e-11 (the
// eigen-divide's null-mode guard, class_solver_3d.f90:202-225)
__device__ __forceinline__ double inv_eig(double lx, double ly, double lz, double sigma) {
  const double big = fmax(fmax(fabs(sigma), fmax(fabs(lx), fabs(ly))), fabs(lz));
  return big < 1e-11 ? 0.0 : 1.0 / (((lx + ly) + lz) + sigma);
}

// T[(o + n c) + n^2 a] = V(o,a) W(a,c) and TT[a + n (o + n c)] = the same
// (V, W column-major n x n); one thread per element
__global__ void __launch_bounds__(256) infl_t_kernel(const double *__restrict__ V,
                                                     const double *__restrict__ W, int n,
                                                     double *__restrict__ T,
                                                     double *__restrict__ TT) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  const long n2 = (long)n * n;
  if (t >= n2 * n) return;
  const int o = (int)(t % n);
  const long r = t / n;
  const int c = (int)(r % n), a = (int)(r / n);
  const double v = V[o + (long)n * a] * W[a + (long)n * c];
  T[t] = v;
  TT[a + (long)n * (o + (long)n * c)] = v;
}

// Eigenvalues of the three physical axes and sigma; index helpers map an
// (impulse axis, output axis, third axis) role triple to (x, y, z).
struct EigArgs {
  const double *l[3];  // lx, ly, lz
  double sigma;
};

// G[(be + nb ga) + nb nc side] = sum_al what_side(al) ahat(al) R, for the two
// output sides; axis a (the common normal), b < c the face axes.
// One thread per (be, ga, side).
__global__ void __launch_bounds__(256) infl_g_kernel(EigArgs e, int a, int b, int c, int na,
                                                     int nb, int nc,
                                                     const double *__restrict__ what0,
                                                     const double *__restrict__ what1,
                                                     const double *__restrict__ ahat,
                                                     double *__restrict__ G) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  const long nbc = (long)nb * nc;
  if (t >= 2 * nbc) return;
  const int side = (int)(t / nbc);
  const long r = t % nbc;
  const int be = (int)(r % nb), ga = (int)(r / nb);
  const double *w = side ? what1 : what0;
  double idx_l[3];
  idx_l[b] = e.l[b][be];
  idx_l[c] = e.l[c][ga];
  double acc = 0.0;
  for (int al = 0; al < na; ++al) {
    idx_l[a] = e.l[a][al];
    acc = fma(w[al] * ahat[al], inv_eig(idx_l[0], idx_l[1], idx_l[2], e.sigma), acc);
  }
  G[t] = acc;
}

// H = sum_at R(as, at, au) Om[at + nt ct] (Om = diag(what) Wt) in one of two
// layouts:
//   layout 0: H[as + ns (ct + nt au)]   (thread index = output index)
//   layout 1: H[au + nu (as + ns ct)]
__global__ void __launch_bounds__(256) infl_h_kernel(EigArgs e, int s, int t_ax, int u, int ns,
                                                     int nt, int nu, int layout,
                                                     const double *__restrict__ Om,
                                                     double *__restrict__ H) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  const long tot = (long)ns * nt * nu;
  if (t >= tot) return;
  int as, au, ct;
  if (layout == 0) {
    as = (int)(t % ns);
    const long r = t / ns;
    ct = (int)(r % nt);
    au = (int)(r / nt);
  } else {
    au = (int)(t % nu);
    const long r = t / nu;
    as = (int)(r % ns);
    ct = (int)(r / ns);
  }
  double lv[3];
  lv[s] = e.l[s][as];
  lv[u] = e.l[u][au];
  const double *om = Om + (long)nt * ct;
  double acc = 0.0;
  for (int at = 0; at < nt; ++at) {
    lv[t_ax] = e.l[t_ax][at];
    acc = fma(inv_eig(lv[0], lv[1], lv[2], e.sigma), om[at], acc);
  }
  H[t] = acc;
}

// One block product of the grouped launch: C = A B (A: M x K column-major,
// B: K x N with K contiguous), written into the subdomain matrix at
//   row = m1 + r2 n1,  col = c1 m2 + c2 n2,  m = m1 + P m2,  n = n1 + Q n2.
// For s = t blocks the impulse's own boundary value on its line adds dcoef
// on the block diagonal (infl_diag_kernel, after the products).
struct InflProblem {
  const double *A;
  const double *B;
  double *D;  // block origin in the subdomain matrix
  long lda, ldb, ldd;
  int M, N, K;
  int P, Q, r2, c1, c2;
  int tiles_n;
  long tile0;  // first global tile of this problem
  double dcoef;
};

// D[q + ldd q] += dcoef for q < count (the diagonal of an s = t block);
// grid (ceil(max count / 256), blocks)
struct InflDiag {
  double *D;
  long ldd;
  int count;
  double dcoef;
};
__global__ void __launch_bounds__(256) infl_diag_kernel(const InflDiag *__restrict__ dg) {
  const InflDiag g = dg[blockIdx.y];
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= g.count) return;
  g.D[q + g.ldd * q] += g.dcoef;
}

// Pointers read from the problem list are generic to the compiler; these
// casts let it emit global (not flat) memory instructions.
typedef const double __attribute__((address_space(1))) gcdouble;
typedef double __attribute__((address_space(1))) gdouble;

// Persistent grouped FP64 MFMA GEMM over a list of block products: each
// block walks the global tile list with stride gridDim.x, locates its
// problem by binary search over tile0 and computes one BM x BN tile (four
// 64 x 64 wave tiles, k-steps of 16 staged through LDS).  Every tile
// writes distinct matrix entries, so no ordering between blocks is needed.
// The matrix is written with non-temporal stores (it is far larger than
// the caches and read only by later solves).
template <int BM, int BN>
__global__ void __launch_bounds__(256) infl_gemm_kernel(const InflProblem *__restrict__ probs,
                                                        int nprob, long total_tiles) {
  constexpr int BK = 16;
  constexpr int NA = BM * BK / 256;
  constexpr int NB = BN * BK / 256;
  static_assert(BM * BN == 256 * 64, "4 waves of 64x64");
  __shared__ double As[BK][BM];
  __shared__ double Bs[BN][BK + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr int WM = BM / 64;
  const int wm = (wave % WM) * 64;
  const int wn = (wave / WM) * 64;
  for (long tile = blockIdx.x; tile < total_tiles; tile += gridDim.x) {
    // problem of this tile: the last with tile0 <= tile
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (probs[mid].tile0 <= tile) lo = mid; else hi = mid - 1;
    }
    const InflProblem p = probs[lo];
    gcdouble *A = (gcdouble *)p.A;
    gcdouble *B = (gcdouble *)p.B;
    const int M = p.M, N = p.N, K = p.K;
    const int lda = (int)p.lda, ldb = (int)p.ldb;
    const long lt = tile - p.tile0;
    const int m0 = (int)(lt / p.tiles_n) * BM;
    const int n0 = (int)(lt % p.tiles_n) * BN;
    d4_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = (d4_t){0, 0, 0, 0};
    double ra[NA], rb[NB];
#pragma unroll
    for (int q = 0; q < NA; ++q) {
      const int e = tid + 256 * q;
      const int m = m0 + (e % BM), k = e / BM;
      ra[q] = (m < M && k < K) ? A[m + (long)lda * k] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int e = tid + 256 * q;
      const int k = e % BK, n = n0 + e / BK;
      rb[q] = (n < N && k < K) ? B[k + (long)ldb * n] : 0.0;
    }
    for (int kt = 0; kt < K; kt += BK) {
      __syncthreads();
#pragma unroll
      for (int q = 0; q < NA; ++q) {
        const int e = tid + 256 * q;
        As[e / BM][e % BM] = ra[q];
      }
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const int e = tid + 256 * q;
        Bs[e / BK][e % BK] = rb[q];
      }
      __syncthreads();
      if (kt + BK < K) {
#pragma unroll
        for (int q = 0; q < NA; ++q) {
          const int e = tid + 256 * q;
          const int m = m0 + (e % BM), k = kt + BK + e / BM;
          ra[q] = (m < M && k < K) ? A[m + (long)lda * k] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < NB; ++q) {
          const int e = tid + 256 * q;
          const int k = kt + BK + e % BK, n = n0 + e / BK;
          rb[q] = (n < N && k < K) ? B[k + (long)ldb * n] : 0.0;
        }
      }
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        double af[4], bf[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          af[i] = As[kk + (lane >> 4)][wm + i * 16 + (lane & 15)];
          bf[i] = Bs[wn + i * 16 + (lane & 15)][kk + (lane >> 4)];
        }
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
#pragma unroll
          for (int ni = 0; ni < 4; ++ni)
            acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[ni], af[mi], acc[mi][ni], 0, 0, 0);
      }
    }
    // epilogue: lane holds C(mb + 16 mi, nb + 16 ni + 4 r), mb = m0+wm+lane%16,
    // nb = n0+wn+lane/16; the (m1, m2) and (n1, n2) splits advance with the
    // offsets (one division each per tile)
    gdouble *D = (gdouble *)p.D;
    const long ldd = p.ldd;
    const int P = p.P, Q = p.Q;
    const int nb = n0 + wn + (lane >> 4);
    const int q1b = nb % Q, q2b = nb / Q;
    const int mb = m0 + wm + (lane & 15);
    int m1 = mb % P, m2 = mb / P;
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
      if (mi > 0) {
        m1 += 16;
        while (m1 >= P) { m1 -= P; ++m2; }
      }
      if (mb + 16 * mi >= M) continue;
      const long colm = (long)p.c1 * m2;
      int a1 = q1b, a2 = q2b;  // column split of nb + 4 d, advanced with d
#pragma unroll
      for (int d = 0; d < 16; ++d) {
        if (nb + 4 * d < N)
          __builtin_nontemporal_store(
              acc[mi][d >> 2][d & 3], D + ((m1 + (long)p.r2 * a1) + ldd * (colm + (long)p.c2 * a2)));
        a1 += 4;
        while (a1 >= Q) { a1 -= Q; ++a2; }
      }
    }
    __syncthreads();  // LDS reuse by the next tile
  }
}

}  // namespace mf
