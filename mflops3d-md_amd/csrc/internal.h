// Internal declarations shared by the host setup, the HIP kernels and the
// C-ABI layer.  Not part of the public interface (include/mflops/mf.h).
#pragma once

#include <string>
#include <vector>

struct mf_der;
struct mf_solver3d;
struct mf_solve_report;

namespace mf {

// One direction of the diagonalisation solver (the reference's
// type(solver_coeffs), class_solver_3d.f90:28-44), plus the dense lhs/rhs the
// iterative-refinement residual needs.
struct Solver1D {
  int n = 0;      // interior points
  int bc[2] = {1, 1};
  std::vector<double> ddl, ddr;        // (n,5), (n,8) column-major
  std::vector<double> lhs, rhs;        // dense (n,n)
  std::vector<double> inv_lhs;         // (n,n)
  std::vector<double> eigen;           // (n)
  std::vector<double> vect, inv_vect;  // (n,n), not transposed
  std::vector<double> neum;            // (2,7)
};

bool dense_solve(int n, double *a, double *r);
bool dense_inverse(int n, const double *a, double *inv);
bool der_coeffs_generic(const double *h0, int cl, int der, int so, double imp[5], double exp_[7]);
bool der_coeffs_global(const double *grid, int n, int der, int so, double *dl, double *dr);
void pentalu(int n, double *dl);
bool solver_coeffs_generic(const double *h0, int cl, int der, int bct, int so, double imp[5],
                           double exp_[8]);
bool dder_coeffs(const double *grid, int n, const int bct[2], int so, double *ddl, double *ddr);
void solver_lhs_rhs(int n, const double *ddl, const double *ddr, double *lhs, double *rhs);
bool extrapolation_coeffs(const double *grid, int n, const int bct[2], int so, double *neum);
bool real_eig(int n, const double *A, double *wr, double *V, double *max_imag, const char **why);
bool solver_setup_1d(const double *grid, int nall, const int bct[2], int so, Solver1D &out,
                     std::string &err);

// Rows 0 and n-1 of the dense operator D = lhs^-1 rhs of one packed compact
// scheme (dl_lu after pentalu, dr), i.e. the boundary-point derivative as a
// dot product with the whole line.
void derivative_rows(int n, const double *dl_lu, const double *dr, double *row0, double *rown);

// Influence blocks of a set of subdomains by spectral contraction
// (solve.hip): for every pair of interface faces (f_in, f_out) of a
// subdomain, the block of its dense u x u matrix D (column-major) at rows
// loc_off[f_out].., columns loc_off[f_in]..  Phase-A kernels per subdomain,
// then one persistent grouped GEMM launch per tile class, captured in a
// HIP graph and launched once when graph is true.
struct InflSub {
  struct ::mf_solver3d *sc;
  const struct ::mf_der *d[3];
  int nf;
  int face[6];
  long loc_off[7];
  long u;
  double *D;
};
struct InflStats {
  double ms = 0.0;         // device time of the whole sequence (events around it)
  double gemm_flop = 0.0;  // 2 M N K of the grouped block products
  double block_bytes = 0.0;  // bytes of the blocks written
};
int influence_assemble(const std::vector<InflSub> &subs, double sigma, bool graph, void *stream,
                       InflStats *stats = nullptr);

// host copies of a derivative handle's tables (order 1|2, sub 0|1)
bool der_host_tables(const struct ::mf_der *dc, int order, int sub, const double **dl_lu,
                     const double **dr, int *n);

// mf_solver3d_solve in two halves, for callers that overlap solves of
// different handles on different streams.  _begin enqueues the solve on the
// stream (FP64: all of it; mixed: the right-hand side, the speculative
// sweeps, u's faces and the copy of the residual norms to the host) and does
// not wait; _end waits for that stream and finishes the refinement (further
// sweeps if the tolerance is not met), reporting as mf_solver3d_solve.
int solver3d_solve_begin(struct ::mf_solver3d *h, const double *g, double sigma, double *u,
                         const double *bcx, const double *bcy, const double *bcz, int precision,
                         double tol, int max_iter, void *stream);
int solver3d_solve_end(struct ::mf_solver3d *h, struct ::mf_solve_report *rep);

// Squared spacings of one grid for the eddy viscosity's filter width
// (one-sided on the first and last points, centred inside).
void les_spacing2(const double *g, int n, double *out);
// Eddy-viscosity kernel launch (les.hip): du[3 i + j] = d u_i / d x_j, d2x/y/z
// device arrays of squared spacings, stream a hipStream_t.
void les_nu_launch(const double *const du[9], const double *d2x, const double *d2y,
                   const double *d2z, int nx, int ny, int nz, int type, double c, double *nu,
                   void *stream);

// Learned-closure kernels (les.hip): features X (n x 10 float) from the
// gradients, and o_i = -tscale tau_ij (edges zero) of direction j from the
// closure's outputs Y (n x 6 float: xx, yy, zz, xy, xz, yz).
void sgs_features_launch(const double *const du[9], const double *d2x, const double *d2y,
                         const double *d2z, int nx, int ny, int nz, double gscale, double lscale,
                         float *X, void *stream);
void sgs_aux_launch(const float *Y, int j, double tscale, double *o0, double *o1, double *o2,
                    int nx, int ny, int nz, void *stream);
// nu = max(0, -(tau : S)) / (2 S : S) of the closure's outputs (tau = tscale Y)
// and the gradients du (du[3 i + j] = d u_i / d x_j); 0 on edges and where S = 0
void sgs_nu_launch(const float *Y, const double *const du[9], double tscale, double *nu, int nx,
                   int ny, int nz, void *stream);

// NetCDF classic field files (io.cpp): one 3-D variable f(nx,ny,nz) of a
// host array, FP64 (dbl) or FP32 on disk; reading converts to FP64.
int nc_write_field(const char *path, const char *var_name, const char *const dim_names[3], int nx,
                   int ny, int nz, const double *f, bool dbl, std::string &err);
int nc_read_field(const char *path, const char *var_name, int dims[3], std::string dim_names[3],
                  bool *is_double, double *f, std::string &err);

// thread-local last error message (mf_last_error)
int fail(int code, const std::string &msg);

}  // namespace mf
