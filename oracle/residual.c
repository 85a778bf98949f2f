/*
 * CPU restatement of the solver operator's residual, multiplied through by
 * P = lhs_x lhs_y lhs_z, in x87 80-bit arithmetic with OpenMP.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ (through oracle/oracle_c.py) as
 * the checker of full-size solves, never by the product path.
 *
 * The operator the diagonalisation solve inverts (class_solver_3d.f90:104-225,
 * with the dense lhs/rhs of solver_matrix_init :484-521) is
 *     L u = sum_a inv(lhs_a) (rhs_a u + bcol_a bc_a) + sigma u.
 * P commutes with every term, so
 *     P (g - L u) = P (g - sigma u) - sum_a (prod_{b != a} lhs_b)(rhs_a u + bcol_a bc_a),
 * a sum of banded 1-D products: no line solves.  bcol_a holds the boundary
 * weights ddr(1,1), ddr(2,1) on the first two rows and ddr(n-1,8), ddr(n,7) on
 * the last two (class_solver_3d.f90:113-124).  The same formula is
 * mf_oracle.multiplied_residual (numpy, longdouble); this file evaluates it
 * with 80-bit sums over every band row and rounds once at the end, fast
 * enough for 256^3 fields.
 *
 * Fields are Fortran-ordered interior arrays (n1,n2,n3), i fastest.  Matrices
 * are dense row-major n x n (row i, column j at i*n+j); only each row's
 * nonzero range is visited.
 */
#include <stdlib.h>
#include <string.h>

typedef long double ld;

typedef struct {
  int n;
  const double *m;  /* n x n row-major */
  int *lo, *hi;     /* nonzero column range of each row (hi exclusive) */
} band_t;

static int band_init(band_t *b, const double *m, int n) {
  b->n = n;
  b->m = m;
  b->lo = (int *)malloc(sizeof(int) * (size_t)n);
  b->hi = (int *)malloc(sizeof(int) * (size_t)n);
  if (!b->lo || !b->hi) return -1;
  for (int i = 0; i < n; ++i) {
    int lo = n, hi = 0;
    for (int j = 0; j < n; ++j)
      if (m[(size_t)i * n + j] != 0.0) {
        if (j < lo) lo = j;
        hi = j + 1;
      }
    if (lo > hi) lo = hi = 0;
    b->lo[i] = lo;
    b->hi[i] = hi;
  }
  return 0;
}

static void band_free(band_t *b) {
  free(b->lo);
  free(b->hi);
}

/* out = M applied along axis `ax` of in (n1,n2,n3). */
static void band_apply(const band_t *b, int ax, const ld *in, ld *out, int n1, int n2, int n3) {
  const long s1 = 1, s2 = n1, s3 = (long)n1 * n2;
  const long stride = ax == 0 ? s1 : (ax == 1 ? s2 : s3);
  const int n = b->n;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < n3; ++k)
    for (int j = 0; j < n2; ++j)
      for (int i = 0; i < n1; ++i) {
        const int idx[3] = {i, j, k};
        const int r = idx[ax];
        const long base = i * s1 + j * s2 + k * s3 - r * stride;
        ld acc = 0.0L;
        const double *row = b->m + (size_t)r * n;
        for (int c = b->lo[r]; c < b->hi[r]; ++c) acc += (ld)row[c] * in[base + c * stride];
        out[base + r * stride] = acc;
      }
}

/*
 * res = P (g - L u) rounded to FP64.
 *   lhs[a], rhs[a]: dense n_a x n_a row-major (a = x, y, z)
 *   bw[a][4]: ddr(1,1), ddr(2,1), ddr(n-1,8), ddr(n,7) of axis a
 *   bcx(n2,n3,2), bcy(n1,n3,2), bcz(n1,n2,2) Fortran order (NULL = zero)
 * Returns 0, or -1 when memory runs out.
 */
int mr_residual(int n1, int n2, int n3, const double *lhs_x, const double *lhs_y,
                const double *lhs_z, const double *rhs_x, const double *rhs_y,
                const double *rhs_z, const double *bw, const double *u, const double *g,
                double sigma, const double *bcx, const double *bcy, const double *bcz,
                double *res) {
  const long N = (long)n1 * n2 * n3;
  const int nn[3] = {n1, n2, n3};
  const double *lhs[3] = {lhs_x, lhs_y, lhs_z}, *rhs[3] = {rhs_x, rhs_y, rhs_z};
  const double *bcs[3] = {bcx, bcy, bcz};
  band_t L[3], R[3];
  ld *acc = (ld *)malloc(sizeof(ld) * (size_t)N);
  ld *t0 = (ld *)malloc(sizeof(ld) * (size_t)N);
  ld *t1 = (ld *)malloc(sizeof(ld) * (size_t)N);
  int rc = (acc && t0 && t1) ? 0 : -1;
  for (int a = 0; a < 3 && rc == 0; ++a)
    if (band_init(&L[a], lhs[a], nn[a]) || band_init(&R[a], rhs[a], nn[a])) rc = -1;
  if (rc) {
    free(acc);
    free(t0);
    free(t1);
    return rc;
  }
  /* P (g - sigma u) */
#pragma omp parallel for schedule(static)
  for (long p = 0; p < N; ++p) t0[p] = (ld)g[p] - (ld)sigma * (ld)u[p];
  band_apply(&L[0], 0, t0, t1, n1, n2, n3);
  band_apply(&L[1], 1, t1, t0, n1, n2, n3);
  band_apply(&L[2], 2, t0, acc, n1, n2, n3);
  /* - prod_{b != a} lhs_b (rhs_a u + bcol_a bc_a) */
  for (int a = 0; a < 3; ++a) {
#pragma omp parallel for schedule(static)
    for (long p = 0; p < N; ++p) t0[p] = (ld)u[p];
    band_apply(&R[a], a, t0, t1, n1, n2, n3);
    if (bcs[a]) {
      const int n = nn[a];
      const double *w = bw + 4 * a;
      /* the face plane: the two transverse axes in increasing order */
      const int ta = a == 0 ? 1 : 0, tb = a == 2 ? 1 : 2;
      const int na = nn[ta], nb = nn[tb];
      const long st[3] = {1, n1, (long)n1 * n2};
      for (int q = 0; q < nb; ++q)
        for (int pp = 0; pp < na; ++pp) {
          const long base = pp * st[ta] + q * st[tb];
          const double b1 = bcs[a][pp + (long)na * q];
          const double b2 = bcs[a][pp + (long)na * (q + (long)nb)];
          t1[base] += (ld)w[0] * b1;
          t1[base + 1 * st[a]] += (ld)w[1] * b1;
          t1[base + (n - 2) * st[a]] += (ld)w[2] * b2;
          t1[base + (n - 1) * st[a]] += (ld)w[3] * b2;
        }
    }
    ld *src = t1, *dst = t0;
    for (int b = 0; b < 3; ++b) {
      if (b == a) continue;
      band_apply(&L[b], b, src, dst, n1, n2, n3);
      ld *t = src;
      src = dst;
      dst = t;
    }
#pragma omp parallel for schedule(static)
    for (long p = 0; p < N; ++p) acc[p] -= src[p];
  }
#pragma omp parallel for schedule(static)
  for (long p = 0; p < N; ++p) res[p] = (double)acc[p];
  for (int a = 0; a < 3; ++a) {
    band_free(&L[a]);
    band_free(&R[a]);
  }
  free(acc);
  free(t0);
  free(t1);
  return 0;
}
