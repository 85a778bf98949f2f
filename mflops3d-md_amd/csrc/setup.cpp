// Host-side setup for the compact derivatives and the diagonalisation solver:
// Taylor-system coefficients, the packed quasi-pentadiagonal LU, the dense
// solver matrices and their real eigen-decomposition.  Setup runs once per
// grid; everything here is FP64 on the host.
//
// The image carries no system LAPACK, so this file has its own small dense
// LU (partial pivoting), inverse, and a Hessenberg + shifted-QR eigensolver
// with Schur-vector back-substitution (the classical real-Schur route that
// LAPACK's dgeev also takes: Householder Hessenberg reduction, QR sweeps,
// triangular eigenvectors mapped back through the accumulated Q).
#include "internal.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

namespace mf {

// ---------------------------------------------------------------------------
// Small dense solve (dgesv replacement): Gaussian elimination with partial
// pivoting on a column-major n x n matrix a, right-hand side r (overwritten).
// ---------------------------------------------------------------------------
bool dense_solve(int n, double *a, double *r) {
  std::vector<int> piv(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::fabs(a[k + (size_t)n * k]);
    for (int i = k + 1; i < n; ++i) {
      double v = std::fabs(a[i + (size_t)n * k]);
      if (v > best) { best = v; p = i; }
    }
    if (best == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(a[k + (size_t)n * j], a[p + (size_t)n * j]);
      std::swap(r[k], r[p]);
    }
    const double inv = 1.0 / a[k + (size_t)n * k];
    for (int i = k + 1; i < n; ++i) {
      double l = a[i + (size_t)n * k] * inv;
      a[i + (size_t)n * k] = l;
      if (l == 0.0) continue;
      for (int j = k + 1; j < n; ++j) a[i + (size_t)n * j] -= l * a[k + (size_t)n * j];
      r[i] -= l * r[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = r[i];
    for (int j = i + 1; j < n; ++j) s -= a[i + (size_t)n * j] * r[j];
    r[i] = s / a[i + (size_t)n * i];
  }
  return true;
}

// Dense inverse by LU with partial pivoting (dgetrf+dgetri replacement).
bool dense_inverse(int n, const double *a_in, double *inv) {
  std::vector<double> a(a_in, a_in + (size_t)n * n);
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::fabs(a[k + (size_t)n * k]);
    for (int i = k + 1; i < n; ++i) {
      double v = std::fabs(a[i + (size_t)n * k]);
      if (v > best) { best = v; p = i; }
    }
    if (best == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(a[k + (size_t)n * j], a[p + (size_t)n * j]);
      std::swap(perm[k], perm[p]);
    }
    const double d = 1.0 / a[k + (size_t)n * k];
    for (int i = k + 1; i < n; ++i) a[i + (size_t)n * k] *= d;
    for (int j = k + 1; j < n; ++j) {
      const double akj = a[k + (size_t)n * j];
      if (akj == 0.0) continue;
      double *cj = &a[(size_t)n * j];
      const double *ck = &a[(size_t)n * k];
      for (int i = k + 1; i < n; ++i) cj[i] -= ck[i] * akj;
    }
  }
  // solve L U X = P for each unit column
  std::vector<double> col(n);
  for (int c = 0; c < n; ++c) {
    for (int i = 0; i < n; ++i) col[i] = (perm[i] == c) ? 1.0 : 0.0;
    for (int i = 0; i < n; ++i) {
      double s = col[i];
      for (int j = 0; j < i; ++j) s -= a[i + (size_t)n * j] * col[j];
      col[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = col[i];
      for (int j = i + 1; j < n; ++j) s -= a[i + (size_t)n * j] * col[j];
      col[i] = s / a[i + (size_t)n * i];
    }
    std::memcpy(inv + (size_t)n * c, col.data(), sizeof(double) * n);
  }
  return true;
}

// ---------------------------------------------------------------------------
// Taylor matrix shared by both coefficient generators
// (class_derivatives_coefficient.f90:135-151, class_solver_coefficient_3d.f90:176-192):
// implicit column j: a(der+1,j) = -1, a(i,j) = a(i-1,j) h(j)/(i-der-1);
// explicit column j: a(1,j) = 1,     a(i,j) = a(i-1,j) h(j)/(i-1).
//
// These small systems are badly conditioned for the high-order second
// derivative (the spread of an FP64 LU across LAPACK builds reaches 5e-11
// relative for so=8), so they are built and solved in x87 80-bit extended
// precision and rounded once to FP64.  The elimination order is fixed
// (partial pivoting on the first maximal |a(i,k)|, row updates left to
// right) so that any restatement using the same steps gets the same bits.
// ---------------------------------------------------------------------------
typedef long double xreal;

static void taylor_matrix(const double *h, int n1, int nt, int der, xreal *a) {
  std::fill(a, a + (size_t)nt * nt, (xreal)0);
  for (int j = 0; j < n1; ++j) {
    a[der + (size_t)nt * j] = -1;
    for (int i = der + 1; i < nt; ++i)
      a[i + (size_t)nt * j] = a[i - 1 + (size_t)nt * j] * (xreal)h[j] / (xreal)(i - der);
  }
  for (int j = n1; j < nt; ++j) {
    a[(size_t)nt * j] = 1;
    for (int i = 1; i < nt; ++i)
      a[i + (size_t)nt * j] = a[i - 1 + (size_t)nt * j] * (xreal)h[j] / (xreal)i;
  }
}

// Gaussian elimination with partial pivoting in extended precision; a is
// column-major nt x nt, r the right-hand side; the solution is rounded to
// FP64 into out.
static bool taylor_solve(int nt, xreal *a, xreal *r, double *out) {
  auto A = [&](int i, int j) -> xreal & { return a[i + (size_t)nt * j]; };
  for (int k = 0; k < nt; ++k) {
    int p = k;
    xreal best = std::fabs(A(k, k));
    for (int i = k + 1; i < nt; ++i)
      if (std::fabs(A(i, k)) > best) { best = std::fabs(A(i, k)); p = i; }
    if (best == 0) return false;
    if (p != k) {
      for (int j = 0; j < nt; ++j) std::swap(A(k, j), A(p, j));
      std::swap(r[k], r[p]);
    }
    for (int i = k + 1; i < nt; ++i) {
      const xreal l = A(i, k) / A(k, k);
      for (int j = k; j < nt; ++j) A(i, j) = A(i, j) - l * A(k, j);
      r[i] = r[i] - l * r[k];
    }
  }
  for (int i = nt - 1; i >= 0; --i) {
    xreal s = r[i];
    for (int j = i + 1; j < nt; ++j) s = s - A(i, j) * r[j];
    r[i] = s / A(i, i);
  }
  for (int i = 0; i < nt; ++i) out[i] = (double)r[i];
  return true;
}

static inline int fdiv(int a, int b) {  // Fortran integer division
  int q = std::abs(a) / std::abs(b);
  return ((a >= 0) == (b >= 0)) ? q : -q;
}

// Distribute the solved unknowns into coef_imp/coef_exp
// (class_derivatives_coefficient.f90:167-182, class_solver_coefficient_3d.f90:220-236).
static void export_coeffs(const double *r, int cl, int n1, int n2, int exp_shift,
                          double *imp, double *exp_, int nexp) {
  std::fill(imp, imp + 5, 0.0);
  std::fill(exp_, exp_ + nexp, 0.0);
  int j = 0;
  int k = (cl == 3) ? (4 - n1) / 2 : 0;
  int l = (n1 + k >= cl) ? n1 + 1 : n1;
  for (int i = 1; i <= l; ++i)
    if (i != cl - k) imp[i + k - 1] = r[j++];
  imp[cl - 1] = 1.0;
  for (int m = 0; m < n2; ++m) exp_[exp_shift + m] = r[n1 + m];
}

// der_coeffs_generic (class_derivatives_coefficient.f90:32-191)
bool der_coeffs_generic(const double *h0, int cl, int der, int so, double imp[5],
                        double exp_[7]) {
  static const int tab[3][4][2] = {  // [cl-1][so index] {n1,n2}; so = 2,4,6,8
      {{0, 3}, {1, 4}, {1, 6}, {3, 6}},
      {{0, 3}, {1, 5}, {2, 6}, {1, 5}},
      {{0, 3}, {2, 3}, {2, 5}, {4, 5}}};
  const int si = so / 2 - 1;
  const int n1 = tab[cl - 1][si][0], n2 = tab[cl - 1][si][1], nt = n1 + n2;
  double h[10] = {0};
  if (cl == 1) {
    for (int m = 0; m < n1; ++m) h[m] = h0[m + 1];
    for (int m = 0; m < n2; ++m) h[n1 + m] = h0[m];
  } else if (cl == 2) {
    h[0] = h0[0];
    for (int m = 1; m < n1; ++m) h[m] = h0[m + 1];
    for (int m = 0; m < n2; ++m) h[n1 + m] = h0[m];
  } else {
    int i = (4 - n1) / 2;
    h[0] = h0[i];
    h[1] = h0[1 + i];
    for (int m = 3 - i; m <= n1; ++m) h[m - 1] = h0[m + i];  // h(3-i:n1)=h0(4:n1+1+i)
    i = (5 - n2) / 2;
    for (int m = 0; m < n2; ++m) h[n1 + m] = h0[i + m];
  }
  xreal a[100], r[10] = {0};
  double sol[10];
  taylor_matrix(h, n1, nt, der, a);
  r[der] = 1;
  if (!taylor_solve(nt, a, r, sol)) return false;
  export_coeffs(sol, cl, n1, n2, cl == 3 ? (5 - n2) / 2 : 0, imp, exp_, 7);
  return true;
}

// Fortran CSHIFT(a, s) on 5 elements into out.
static void cshift5(const double *a, int s, double *out) {
  for (int i = 0; i < 5; ++i) out[i] = a[(((i + s) % 5) + 5) % 5];
}

// der_coeffs_generic_gobal (class_derivatives.f90:330-377); dl(n,5), dr(n,7)
// column-major.
bool der_coeffs_global(const double *grid, int n, int der, int so, double *dl, double *dr) {
  if (n < 8) return false;
  std::fill(dl, dl + (size_t)5 * n, 0.0);
  std::fill(dr, dr + (size_t)7 * n, 0.0);
  double h[7], imp[5], ex[7], row[5], rev5[5];
  auto put = [&](int i, const double *l5, const double *r7, bool reverse_r) {
    for (int m = 0; m < 5; ++m) dl[i + (size_t)n * m] = l5[m];
    for (int m = 0; m < 7; ++m) dr[i + (size_t)n * m] = -(reverse_r ? r7[6 - m] : r7[m]);
  };
  for (int m = 0; m < 7; ++m) h[m] = grid[0] - grid[m];
  if (!der_coeffs_generic(h, 1, der, so, imp, ex)) return false;
  cshift5(imp, -2, row);
  put(0, row, ex, false);
  for (int m = 0; m < 7; ++m) h[m] = grid[1] - grid[m];
  if (!der_coeffs_generic(h, 2, der, so, imp, ex)) return false;
  cshift5(imp, -1, row);
  put(1, row, ex, false);
  for (int i = 2; i < n - 2; ++i) {
    for (int m = 0; m < 5; ++m) h[m] = grid[i] - grid[i - 2 + m];
    h[5] = h[6] = 0.0;
    if (!der_coeffs_generic(h, 3, der, so, imp, ex)) return false;
    put(i, imp, ex, false);
  }
  for (int m = 0; m < 7; ++m) h[m] = grid[n - 2] - grid[n - 1 - m];
  if (!der_coeffs_generic(h, 2, der, so, imp, ex)) return false;
  for (int m = 0; m < 5; ++m) rev5[m] = imp[4 - m];
  cshift5(rev5, 1, row);
  put(n - 2, row, ex, true);
  for (int m = 0; m < 7; ++m) h[m] = grid[n - 1] - grid[n - 1 - m];
  if (!der_coeffs_generic(h, 1, der, so, imp, ex)) return false;
  for (int m = 0; m < 5; ++m) rev5[m] = imp[4 - m];
  cshift5(rev5, 2, row);
  put(n - 1, row, ex, true);
  if (der == 2)
    for (size_t q = 0; q < (size_t)7 * n; ++q) dr[q] = -dr[q];
  return true;
}

// pentalu (class_derivatives.f90:379-423) on the packed dl(n,5), with the
// reference's aliasing alpha = dl(1,1) and beta = dl(n,5).
void pentalu(int n, double *dl) {
  double *p = dl, *f = dl + n, *c = dl + 2 * (size_t)n, *d = dl + 3 * (size_t)n,
         *e = dl + 4 * (size_t)n;
  const double alpha = p[0];
  f[1] = f[1] / c[0];
  c[1] = c[1] - d[0] * f[1];
  d[1] = d[1] - e[0] * f[1];
  e[1] = e[1] - alpha * f[1];
  int i = 2;
  p[i] = p[i] / c[i - 2];
  f[i] = (f[i] - p[i] * d[i - 2]) / c[i - 1];
  c[i] = c[i] - d[i - 1] * f[i] - e[i - 2] * p[i];
  d[i] = d[i] - e[i - 1] * f[i] - alpha * p[i];
  for (i = 3; i < n - 1; ++i) {
    p[i] = p[i] / c[i - 2];
    f[i] = (f[i] - p[i] * d[i - 2]) / c[i - 1];
    c[i] = c[i] - e[i - 2] * p[i] - d[i - 1] * f[i];
    d[i] = d[i] - e[i - 1] * f[i];
  }
  const double beta = e[n - 1] / c[n - 4];
  e[n - 1] = beta;
  p[n - 1] = p[n - 1] - beta * d[n - 4];
  f[n - 1] = f[n - 1] - beta * e[n - 4];
  p[n - 1] = p[n - 1] / c[n - 3];
  f[n - 1] = (f[n - 1] - p[n - 1] * d[n - 3]) / c[n - 2];
  c[n - 1] = c[n - 1] - e[n - 3] * p[n - 1] - d[n - 2] * f[n - 1];
}

// Apply one packed scheme to a line on the host: stencil right-hand side
// (class_derivatives.f90:262-295) and pentares (:425-466).
static void apply_scheme(int n, const double *dl, const double *dr, const double *f, double *x) {
  auto DR = [&](int i, int m) { return dr[i + (size_t)n * m]; };
  const double *p = dl, *fl = dl + n, *c = dl + 2 * (size_t)n, *d = dl + 3 * (size_t)n,
               *e = dl + 4 * (size_t)n;
  const double alpha = p[0], beta = e[n - 1];
  std::vector<double> b(n);
  for (int row = 0; row < 2; ++row) {
    double a = 0.0;
    for (int m = 0; m < 7; ++m) a += DR(row, m) * f[m];
    b[row] = a;
  }
  for (int i = 2; i < n - 2; ++i) {
    double a = 0.0;
    for (int m = 0; m < 5; ++m) a += DR(i, m) * f[i - 2 + m];
    b[i] = a;
  }
  for (int row = n - 2; row < n; ++row) {
    double a = 0.0;
    for (int m = 0; m < 7; ++m) a += DR(row, m) * f[n - 7 + m];
    b[row] = a;
  }
  x[0] = b[0];
  x[1] = b[1] - x[0] * fl[1];
  for (int i = 2; i < n - 1; ++i) x[i] = b[i] - p[i] * x[i - 2] - fl[i] * x[i - 1];
  x[n - 1] = b[n - 1] - p[n - 1] * x[n - 3] - fl[n - 1] * x[n - 2] - beta * x[n - 4];
  x[n - 1] = x[n - 1] / c[n - 1];
  x[n - 2] = (x[n - 2] - d[n - 2] * x[n - 1]) / c[n - 2];
  for (int i = n - 3; i >= 1; --i) x[i] = (x[i] - e[i] * x[i + 2] - d[i] * x[i + 1]) / c[i];
  x[0] = (x[0] - e[0] * x[2] - d[0] * x[1] - alpha * x[3]) / c[0];
}

void derivative_rows(int n, const double *dl_lu, const double *dr, double *row0, double *rown) {
  std::vector<double> f(n, 0.0), x(n);
  for (int m = 0; m < n; ++m) {
    f[m] = 1.0;
    apply_scheme(n, dl_lu, dr, f.data(), x.data());
    row0[m] = x[0];
    rown[m] = x[n - 1];
    f[m] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// Solver coefficients (class_solver_coefficient_3d.f90:30-245)
// ---------------------------------------------------------------------------
static void solver_stencil(int cl, int so, int bct, int &n1, int &n2) {
  if (cl == 1) {
    static const int t[4][2] = {{0, 3}, {1, 4}, {1, 6}, {3, 6}};
    n1 = t[so / 2 - 1][0]; n2 = t[so / 2 - 1][1];
  } else if (cl == 4) {
    static const int t[4] = {3, 5, 7, 7};
    n1 = 0; n2 = t[so / 2 - 1];
  } else if (cl == 2) {
    if (so == 2) { n1 = 0; n2 = 3; }
    else if (so == 8) {
      if (bct == 3) { n1 = 2; n2 = 7; } else { n1 = 1; n2 = 5; }
    } else {
      if (bct == 3) { n1 = 1; n2 = 3; } else { n1 = 1; n2 = 5; }
    }
  } else {
    static const int t[4][2] = {{0, 3}, {2, 3}, {2, 5}, {4, 5}};
    n1 = t[so / 2 - 1][0]; n2 = t[so / 2 - 1][1];
  }
}

bool solver_coeffs_generic(const double *h0, int cl, int der, int bct, int so,
                           double imp[5], double exp_[8]) {
  int n1, n2;
  solver_stencil(cl, so, bct, n1, n2);
  const int n3 = n1 + n2;
  double h[10] = {0};
  if (cl == 1) {
    for (int m = 0; m < n1; ++m) h[m] = h0[m + 2];
    for (int m = 0; m < n2; ++m) h[n1 + m] = h0[m];
  } else if (cl == 4) {
    for (int m = 0; m < n3; ++m) h[m] = h0[m];
  } else if (cl == 2) {
    h[0] = h0[1];
    for (int m = 1; m < n1; ++m) h[m] = h0[m + 2];  // h(2:n1)=h0(4:n1+2)
    int i = std::max(0, fdiv(5 - n2, 2));
    for (int m = 0; m < n2; ++m) h[n1 + m] = h0[i + m];
  } else {
    int i = (4 - n1) / 2;
    h[0] = h0[i];
    h[1] = h0[1 + i];
    for (int m = 3 - i; m <= n1; ++m) h[m - 1] = h0[m + i];
    i = (5 - n2) / 2;
    for (int m = 0; m < n2; ++m) h[n1 + m] = h0[i + m];
  }
  xreal a[100], r[10] = {0};
  double sol[10];
  taylor_matrix(h, n1, n3, der, a);
  // boundary column (:196-203)
  if ((cl == 2 && n2 > 4) || cl == 1) {
    const int c = n1;
    for (int i = 0; i < bct - 1; ++i) a[i + (size_t)n3 * c] = 0;
    a[bct - 1 + (size_t)n3 * c] = (bct == 3) ? -1 : 1;
    for (int i = bct; i < n3; ++i)
      a[i + (size_t)n3 * c] = a[i - 1 + (size_t)n3 * c] * (xreal)h[c] / (xreal)(i + 1 - bct);
  }
  r[der] = 1;
  if (!taylor_solve(n3, a, r, sol)) return false;
  int s = 0;
  if (cl == 3) s = fdiv(5 - n2, 2);
  if (cl == 2) s = std::max(0, fdiv(5 - n2, 2));
  export_coeffs(sol, cl, n1, n2, s, imp, exp_, 8);
  return true;
}

// dder_coeffs (class_solver_3d.f90:568-619): grid(0:n+1), ddl(n,5), ddr(n,8).
bool dder_coeffs(const double *grid, int n, const int bct[2], int so, double *ddl, double *ddr) {
  if (n < 7) return false;
  std::fill(ddl, ddl + (size_t)5 * n, 0.0);
  std::fill(ddr, ddr + (size_t)8 * n, 0.0);
  const int der = 2;
  double h[7], imp[5], ex[8], row[5], rev5[5];
  auto L = [&](int i, int m) -> double & { return ddl[i + (size_t)n * m]; };
  auto R = [&](int i, int m) -> double & { return ddr[i + (size_t)n * m]; };
  for (int m = 0; m < 7; ++m) h[m] = grid[1] - grid[m];
  if (!solver_coeffs_generic(h, 1, der, bct[0], so, imp, ex)) return false;
  cshift5(imp, -2, row);
  for (int m = 0; m < 5; ++m) L(0, m) = row[m];
  for (int m = 0; m < 8; ++m) R(0, m) = -ex[m];
  if (bct[0] == 2) R(0, 0) = -R(0, 0);
  for (int m = 0; m < 7; ++m) h[m] = grid[2] - grid[m];
  if (!solver_coeffs_generic(h, 2, der, bct[0], so, imp, ex)) return false;
  cshift5(imp, -1, row);
  for (int m = 0; m < 5; ++m) L(1, m) = row[m];
  for (int m = 0; m < 8; ++m) R(1, m) = -ex[m];
  if (bct[0] == 2) R(1, 0) = -R(1, 0);
  for (int i = 3; i <= n - 2; ++i) {  // reference index on grid(0:n+1)
    for (int m = 0; m < 5; ++m) h[m] = grid[i] - grid[i - 2 + m];
    h[5] = h[6] = 0.0;
    if (!solver_coeffs_generic(h, 3, der, 0, so, imp, ex)) return false;
    for (int m = 0; m < 5; ++m) L(i - 1, m) = imp[m];
    for (int m = 0; m < 8; ++m) R(i - 1, m) = -ex[m];
  }
  for (int m = 0; m < 7; ++m) h[m] = grid[n - 1] - grid[n + 1 - m];
  if (!solver_coeffs_generic(h, 2, der, bct[1], so, imp, ex)) return false;
  for (int m = 0; m < 5; ++m) rev5[m] = imp[4 - m];
  cshift5(rev5, 1, row);
  for (int m = 0; m < 5; ++m) L(n - 2, m) = row[m];
  for (int m = 0; m < 8; ++m) R(n - 2, m) = -ex[7 - m];
  if (bct[1] == 2) R(n - 2, 7) = -R(n - 2, 7);
  for (int m = 0; m < 7; ++m) h[m] = grid[n] - grid[n + 1 - m];
  if (!solver_coeffs_generic(h, 1, der, bct[1], so, imp, ex)) return false;
  for (int m = 0; m < 5; ++m) rev5[m] = imp[4 - m];
  cshift5(rev5, 2, row);
  for (int m = 0; m < 5; ++m) L(n - 1, m) = row[m];
  double rev8[8];
  for (int m = 0; m < 8; ++m) rev8[m] = ex[7 - m];
  for (int m = 0; m < 8; ++m) R(n - 1, m) = -rev8[(m + 1) % 8];  // -CSHIFT(rev, 1)
  if (bct[1] == 2) R(n - 1, 6) = -R(n - 1, 6);
  for (size_t q = 0; q < (size_t)8 * n; ++q) ddr[q] = -ddr[q];  // der == 2
  return true;
}

// Dense lhs, rhs (class_solver_3d.f90:484-521), column-major n x n.
void solver_lhs_rhs(int n, const double *ddl, const double *ddr, double *lhs, double *rhs) {
  std::fill(lhs, lhs + (size_t)n * n, 0.0);
  std::fill(rhs, rhs + (size_t)n * n, 0.0);
  auto L = [&](int i, int m) { return ddl[i + (size_t)n * m]; };
  auto R = [&](int i, int m) { return ddr[i + (size_t)n * m]; };
  auto A = [&](int i, int j) -> double & { return lhs[i + (size_t)n * j]; };
  auto B = [&](int i, int j) -> double & { return rhs[i + (size_t)n * j]; };
  A(0, 0) = L(0, 2); A(0, 1) = L(0, 3); A(0, 2) = L(0, 4); A(0, 3) = L(0, 0);
  A(1, 0) = L(1, 1); A(1, 1) = L(1, 2); A(1, 2) = L(1, 3); A(1, 3) = L(1, 4);
  for (int i = 2; i < n - 2; ++i)
    for (int m = 0; m < 5; ++m) A(i, i - 2 + m) = L(i, m);
  for (int m = 0; m < 4; ++m) A(n - 2, n - 4 + m) = L(n - 2, m);
  A(n - 1, n - 4) = L(n - 1, 4);
  for (int m = 0; m < 3; ++m) A(n - 1, n - 3 + m) = L(n - 1, m);
  for (int m = 0; m < 6; ++m) B(0, m) = R(0, m + 1);
  for (int m = 0; m < 7; ++m) B(1, m) = R(1, m + 1);
  for (int i = 2; i < n - 2; ++i)
    for (int m = 0; m < 5; ++m) B(i, i - 2 + m) = R(i, m);
  for (int m = 0; m < 7; ++m) B(n - 2, n - 7 + m) = R(n - 2, m);
  for (int m = 0; m < 6; ++m) B(n - 1, n - 6 + m) = R(n - 1, m);
}

// extrapolation_coeffs (class_solver_3d.f90:622-656); neum(2,7) column-major.
bool extrapolation_coeffs(const double *grid, int n, const int bct[2], int so, double *neum) {
  double h[7], imp[5], ex[8];
  for (int m = 0; m < 7; ++m) h[m] = grid[0] - grid[m];
  if (!solver_coeffs_generic(h, 4, 1, bct[0], so, imp, ex)) return false;
  for (int m = 0; m < 7; ++m) neum[0 + 2 * m] = -ex[m];
  for (int m = 0; m < 7; ++m) h[m] = grid[n + 1] - grid[n + 1 - m];
  if (!solver_coeffs_generic(h, 4, 1, bct[1], so, imp, ex)) return false;
  for (int m = 0; m < 7; ++m) neum[1 + 2 * m] = -ex[6 - m];
  return true;
}

// ---------------------------------------------------------------------------
// Real eigen-decomposition of a general matrix with a real spectrum.
// 1. Householder reduction to upper Hessenberg form H = Q0^T A Q0.
// 2. Implicit single-shift QR sweeps (Wilkinson shift) to real Schur form
//    T = Z^T A Z, 2x2 blocks with real eigenvalues split by a rotation.
// 3. Eigenvectors of the triangular T by back-substitution, mapped by Z.
// A pair of complex eigenvalues with |Im| above 1e-13 is an error, as in the
// reference (class_solver_3d.f90:537-540).
// ---------------------------------------------------------------------------
namespace {
struct Mat {
  int n;
  std::vector<double> v;
  explicit Mat(int n_) : n(n_), v((size_t)n_ * n_, 0.0) {}
  double &operator()(int i, int j) { return v[i + (size_t)n * j]; }
  double operator()(int i, int j) const { return v[i + (size_t)n * j]; }
};

inline void givens(double x, double z, double &c, double &s) {
  if (z == 0.0) { c = 1.0; s = 0.0; return; }
  double r = std::hypot(x, z);
  c = x / r;
  s = z / r;
}

// apply G = [c s; -s c] to rows k,k+1 (columns j0..j1) and G^T to columns
// k,k+1 (rows i0..i1) of T, and accumulate Z <- Z G^T.
inline void rotate(Mat &T, Mat &Z, int k, double c, double s, int j0, int j1, int i0, int i1) {
  for (int j = j0; j <= j1; ++j) {
    double t1 = T(k, j), t2 = T(k + 1, j);
    T(k, j) = c * t1 + s * t2;
    T(k + 1, j) = -s * t1 + c * t2;
  }
  for (int i = i0; i <= i1; ++i) {
    double t1 = T(i, k), t2 = T(i, k + 1);
    T(i, k) = c * t1 + s * t2;
    T(i, k + 1) = -s * t1 + c * t2;
  }
  for (int i = 0; i < Z.n; ++i) {
    double t1 = Z(i, k), t2 = Z(i, k + 1);
    Z(i, k) = c * t1 + s * t2;
    Z(i, k + 1) = -s * t1 + c * t2;
  }
}
}  // namespace

bool real_eig(int n, const double *A, double *wr, double *V, double *max_imag, const char **why) {
  *max_imag = 0.0;
  Mat T(n), Z(n);
  std::memcpy(T.v.data(), A, sizeof(double) * n * n);
  for (int i = 0; i < n; ++i) Z(i, i) = 1.0;
  // 1. Hessenberg reduction
  std::vector<double> u(n), w(n);
  for (int k = 0; k + 2 < n; ++k) {
    double alpha = 0.0;
    for (int i = k + 1; i < n; ++i) alpha += T(i, k) * T(i, k);
    alpha = std::sqrt(alpha);
    if (alpha == 0.0) continue;
    if (T(k + 1, k) > 0) alpha = -alpha;
    for (int i = 0; i < n; ++i) u[i] = 0.0;
    for (int i = k + 1; i < n; ++i) u[i] = T(i, k);
    u[k + 1] -= alpha;
    double un = 0.0;
    for (int i = k + 1; i < n; ++i) un += u[i] * u[i];
    if (un == 0.0) continue;
    const double beta = 2.0 / un;
    // T <- (I - beta u u^T) T
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int i = k + 1; i < n; ++i) s += u[i] * T(i, j);
      s *= beta;
      for (int i = k + 1; i < n; ++i) T(i, j) -= s * u[i];
    }
    // T <- T (I - beta u u^T); Z likewise
    for (int i = 0; i < n; ++i) {
      double s = 0.0, sz = 0.0;
      for (int j = k + 1; j < n; ++j) { s += T(i, j) * u[j]; sz += Z(i, j) * u[j]; }
      s *= beta;
      sz *= beta;
      for (int j = k + 1; j < n; ++j) { T(i, j) -= s * u[j]; Z(i, j) -= sz * u[j]; }
    }
    for (int i = k + 2; i < n; ++i) T(i, k) = 0.0;
  }
  double hnorm = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i <= std::min(n - 1, j + 1); ++i) hnorm = std::max(hnorm, std::fabs(T(i, j)));
  // 2. QR sweeps to real Schur form
  const double eps = DBL_EPSILON;
  int hi = n - 1, iter = 0, total = 0;
  while (hi >= 0) {
    if (hi == 0) { hi = -1; break; }
    int l = hi;
    while (l > 0) {
      double s = std::fabs(T(l - 1, l - 1)) + std::fabs(T(l, l));
      if (s == 0.0) s = hnorm;
      if (std::fabs(T(l, l - 1)) <= eps * s) { T(l, l - 1) = 0.0; break; }
      --l;
    }
    if (l == hi) { --hi; iter = 0; continue; }
    if (l == hi - 1) {
      // split the 2x2 block [a b; c d] into triangular form
      const double a = T(hi - 1, hi - 1), b = T(hi - 1, hi), c = T(hi, hi - 1), d = T(hi, hi);
      const double p = 0.5 * (a - d), disc = p * p + b * c;
      if (disc < 0.0) {
        double im = std::sqrt(-disc);
        *max_imag = std::max(*max_imag, im);
        if (im > 1e-13) { *why = "complex eigenvalues in solver matrix"; return false; }
      }
      const double sq = disc > 0.0 ? std::sqrt(disc) : 0.0;
      // an eigenvalue lam of the block and its eigenvector v; the rotation
      // with G^T e1 = v/|v| makes the block upper triangular
      const double z = p + (p >= 0 ? sq : -sq);
      double lam = (z != 0.0) ? d - b * c / z : d;
      double v1 = lam - d, v2 = c;       // from row 2 of (B - lam I) v = 0
      double w1 = b, w2 = lam - a;       // from row 1
      if (std::hypot(w1, w2) > std::hypot(v1, v2)) { v1 = w1; v2 = w2; }
      double cs, sn;
      if (v1 == 0.0 && v2 == 0.0) { cs = 1.0; sn = 0.0; } else givens(v1, v2, cs, sn);
      rotate(T, Z, hi - 1, cs, sn, hi - 1, n - 1, 0, hi);
      T(hi, hi - 1) = 0.0;
      hi -= 2;
      iter = 0;
      continue;
    }
    if (++iter > 80 || ++total > 40 * n) {
      *why = "QR iteration did not converge (complex or defective spectrum)";
      return false;
    }
    // Wilkinson shift from the trailing 2x2 of the active block
    double mu;
    {
      const double a = T(hi - 1, hi - 1), b = T(hi - 1, hi), c = T(hi, hi - 1), d = T(hi, hi);
      const double p = 0.5 * (a - d), disc = p * p + b * c;
      if (disc >= 0.0) {
        const double z = p + (p >= 0 ? std::sqrt(disc) : -std::sqrt(disc));
        mu = (z != 0.0) ? d - b * c / z : d;
      } else {
        mu = d;
      }
      if (iter % 11 == 10) mu = d + 0.75 * std::fabs(c);  // exceptional shift
    }
    double x = T(l, l) - mu, zz = T(l + 1, l);
    for (int k = l; k < hi; ++k) {
      double c, s;
      givens(x, zz, c, s);
      rotate(T, Z, k, c, s, std::max(l, k - 1), n - 1, 0, std::min(k + 2, hi));
      if (k > l) T(k + 1, k - 1) = 0.0;
      if (k < hi - 1) { x = T(k + 1, k); zz = T(k + 2, k); }
    }
  }
  for (int i = 0; i < n; ++i) wr[i] = T(i, i);
  // 3. eigenvectors of T by back-substitution, then V = Z Y
  double tnorm = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i <= j; ++i) tnorm = std::max(tnorm, std::fabs(T(i, j)));
  const double small = std::max(eps * tnorm, DBL_MIN);
  std::vector<double> y(n);
  for (int j = 0; j < n; ++j) {
    const double lam = T(j, j);
    std::fill(y.begin(), y.end(), 0.0);
    y[j] = 1.0;
    for (int i = j - 1; i >= 0; --i) {
      double s = T(i, j);
      for (int m = i + 1; m < j; ++m) s += T(i, m) * y[m];
      double den = T(i, i) - lam;
      if (std::fabs(den) < small) den = (den >= 0 ? small : -small);
      y[i] = -s / den;
      if (std::fabs(y[i]) > 1e100) {  // rescale to avoid overflow
        for (int m = i; m <= j; ++m) y[m] *= 1e-100;
      }
    }
    double nrm = 0.0;
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int m = 0; m <= j; ++m) s += Z(i, m) * y[m];
      V[i + (size_t)n * j] = s;
      nrm += s * s;
    }
    nrm = std::sqrt(nrm);
    // unit 2-norm, largest-magnitude component positive
    int imax = 0;
    for (int i = 1; i < n; ++i)
      if (std::fabs(V[i + (size_t)n * j]) > std::fabs(V[imax + (size_t)n * j])) imax = i;
    const double sc = (V[imax + (size_t)n * j] < 0 ? -1.0 : 1.0) / nrm;
    for (int i = 0; i < n; ++i) V[i + (size_t)n * j] *= sc;
  }
  return true;
}

// The null mode of an axis with Neumann conditions at both ends (the
// constant vector) has eigenvalue zero; dgeev or the QR iteration returns it
// as rounding noise of order eps * max|lambda| / 1e6 or so.  The solve's guard
// (class_solver_3d.f90:207-209) drops the mode when |sigma| and the three
// eigenvalues are all below the absolute threshold 1e-11, which such noise can
// exceed on strongly stretched grids (1.05e-11 on a 256-point tanh 2.5 axis,
// whose largest eigenvalue is 1.5e7): the mode is then divided by the noise
// and the pressure's constant becomes arbitrary.  Eigenvalues below
// 1e-13 max|lambda| are therefore set to exactly zero, so that the guard
// acts on every grid; where the reference's threshold already catches the
// mode the solve is unchanged.
void snap_null_eigenvalues(int n, double *eigen) {
  double emax = 0.0;
  for (int i = 0; i < n; ++i) emax = std::max(emax, std::fabs(eigen[i]));
  for (int i = 0; i < n; ++i)
    if (std::fabs(eigen[i]) <= 1e-13 * emax) eigen[i] = 0.0;
}

// Complete 1-D solver setup (solver_coeffs_init, class_solver_3d.f90:413-466).
bool solver_setup_1d(const double *grid, int nall, const int bct[2], int so, Solver1D &out,
                     std::string &err) {
  const int n = nall - 2;
  if (n < 7) { err = "solver needs at least 9 points per direction"; return false; }
  if (so != 2 && so != 4 && so != 6 && so != 8) { err = "so must be 2, 4, 6 or 8"; return false; }
  for (int s = 0; s < 2; ++s)
    if (bct[s] < 1 || bct[s] > 3) { err = "bct codes must be 1, 2 or 3"; return false; }
  out.n = n;
  out.bc[0] = bct[0];
  out.bc[1] = bct[1];
  out.ddl.assign((size_t)5 * n, 0.0);
  out.ddr.assign((size_t)8 * n, 0.0);
  if (!dder_coeffs(grid, n, bct, so, out.ddl.data(), out.ddr.data())) {
    err = "singular Taylor system in solver coefficients";
    return false;
  }
  std::vector<double> lhs((size_t)n * n), rhs((size_t)n * n);
  solver_lhs_rhs(n, out.ddl.data(), out.ddr.data(), lhs.data(), rhs.data());
  out.lhs = lhs;
  out.rhs = rhs;
  out.inv_lhs.assign((size_t)n * n, 0.0);
  if (!dense_inverse(n, lhs.data(), out.inv_lhs.data())) { err = "singular solver lhs"; return false; }
  // A = inv(lhs) rhs
  std::vector<double> A((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) {
      const double r = rhs[k + (size_t)n * j];
      if (r == 0.0) continue;
      for (int i = 0; i < n; ++i) A[i + (size_t)n * j] += out.inv_lhs[i + (size_t)n * k] * r;
    }
  out.eigen.assign(n, 0.0);
  out.vect.assign((size_t)n * n, 0.0);
  double mi = 0.0;
  const char *why = "";
  if (!real_eig(n, A.data(), out.eigen.data(), out.vect.data(), &mi, &why)) {
    err = why;
    return false;
  }
  snap_null_eigenvalues(n, out.eigen.data());
  out.inv_vect.assign((size_t)n * n, 0.0);
  if (!dense_inverse(n, out.vect.data(), out.inv_vect.data())) {
    err = "eigenvector matrix is singular";
    return false;
  }
  out.neum.assign(14, 0.0);
  if (!extrapolation_coeffs(grid, n, bct, so, out.neum.data())) {
    err = "singular Taylor system in extrapolation coefficients";
    return false;
  }
  return true;
}

}  // namespace mf
