/*
 * mflops: MI355X-native compact derivatives and matrix-diagonalisation
 * Helmholtz/Poisson solver -- C ABI.
 *
 * Conventions
 *   - Fields are FP64, column-major (Fortran order) f(nx,ny,nz), i fastest,
 *     exactly the memory layout of the reference's type(field)%f.
 *   - Boundary arrays follow the reference's type(boundary_condition):
 *     bcx(ny-2,nz-2,2), bcy(nx-2,nz-2,2), bcz(nx-2,ny-2,2).
 *   - Functions taking a `stream` work on DEVICE pointers and enqueue on that
 *     hipStream_t (NULL = default stream); they return once enqueued unless
 *     stated otherwise.  *_host variants take host pointers and block.
 *   - Every function returns MF_OK (0) or an error code; mf_last_error()
 *     gives the message of the last failure on the calling thread.
 *   - bct codes: 1 = Dirichlet, 2 = Neumann (3 = second-derivative variant
 *     of the reference's extrapolation table).
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   mf_der_create        <- der_coeffs_init      src/class_derivatives/class_derivatives.f90:111
 *   mf_der_lines         <- der/dder/der_s/dder_s src/class_derivatives/class_derivatives.f90:186-260
 *   mf_der3d             <- derx..dderz          src/class_field/class_field.f90:629-978
 *   mf_solver3d_create   <- solver_coeffs_init_3d src/class_solver/class_solver_3d.f90:380
 *   mf_solver3d_solve    <- solver_3d / solve_3d src/class_field/class_field.f90:1072,
 *                                               src/class_solver/class_solver_3d.f90:83
 *   mf_der_coeffs_1d     <- der_coeffs_generic_gobal + pentalu
 *                                               src/class_derivatives/class_derivatives.f90:330-423
 *   mf_solver_coeffs_1d  <- solver_coeffs_init   src/class_solver/class_solver_3d.f90:413-466
 */
#ifndef MFLOPS_MF_H
#define MFLOPS_MF_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MF_OK = 0,
  MF_ERR_ARG = 1,      /* invalid argument (sizes, codes, null pointers) */
  MF_ERR_SETUP = 2,    /* coefficient setup failed (singular Taylor system,
                          complex spectrum: class_solver_3d.f90:537-540) */
  MF_ERR_DEVICE = 3,   /* HIP runtime error */
  MF_ERR_NOCONV = 4    /* mixed-precision refinement did not reach tol */
};

/* MF_MIXED: FP32 MFMA transforms.  MF_MIXED_BF16X3: the same transforms with
 * every operand split into three bf16 parts (six bf16 MFMA products, FP32
 * accumulation: FP32-class accuracy on the bf16 matrix cores).
 * MF_MIXED_FP16X2: operands scaled by powers of two and split into two fp16
 * parts (three fp16 MFMA products: 22-bit accuracy at half the matrix-core
 * work of BF16X3).  All refine with the FP64 residual to tol. */
enum { MF_FP64 = 0, MF_MIXED = 1, MF_MIXED_BF16X3 = 2, MF_MIXED_FP16X2 = 3 };

typedef struct mf_der mf_der;
typedef struct mf_solver3d mf_solver3d;

typedef struct mf_solve_report {
  int iterations;    /* low-precision solves done (0 for MF_FP64)            */
  double residual;   /* final ||r||_2/||b||_2 of the FP64 operator (MIXED)   */
  double residual0;  /* same after the first low-precision solve (MIXED)     */
} mf_solve_report;

const char *mf_last_error(void);
int mf_version(void);

/* ---------------------------------------------------------------------
 * Host-side coefficient tables (no GPU needed).
 * --------------------------------------------------------------------- */

/* Packed compact-scheme tables for one line of n points: dl_lu(n,5) is the
 * pentalu-factored (p,f,c,d,e) with alpha=dl(1,1), beta=dl(n,5); dr(n,7) is
 * the explicit stencil.  Column-major like the reference's dl(n,5), dr(n,7).
 * der = 1|2, so = 2|4|6|8. */
int mf_der_coeffs_1d(const double *grid, int n, int der, int so,
                     double *dl_lu, double *dr);

/* Solver tables for one direction of nall points (interior n = nall-2):
 * ddl(n,5), ddr(n,8), eigen(n), vect(n,n), inv_vect(n,n), inv_lhs(n,n),
 * neum(2,7); all column-major; vect/inv_vect NOT transposed.  Any output
 * pointer may be NULL. */
int mf_solver_coeffs_1d(const double *grid, int nall, const int bct[2], int so,
                        double *ddl, double *ddr, double *eigen, double *vect,
                        double *inv_vect, double *inv_lhs, double *neum);

/* ---------------------------------------------------------------------
 * Compact derivatives.
 * --------------------------------------------------------------------- */

/* solver_mode != 0 also builds the sub-line (_s) tables on grid(2:n-1),
 * used by the 3-D driver on lines lying in boundary planes. */
int mf_der_create(const double *grid, int n, int solver_mode, int so, mf_der **out);
int mf_der_destroy(mf_der *dc);

/* Per-line derivative of nlines lines of length n (n = the handle's size),
 * line l at x + l*ld.  order 1|2; sub != 0 applies der_s/dder_s (points
 * 2..n-1 only; out(1), out(n) are left untouched). */
int mf_der_lines(const mf_der *dc, int order, int sub, const double *x, double *out,
                 int nlines, long long ld, void *stream);

/* 3-D derivative along axis 0|1|2 (x|y|z) of x(nx,ny,nz) into out.
 * Solver mode: interior lines use the full scheme, lines in one boundary
 * plane use the sub-line scheme; points on two boundary planes are set 0. */
int mf_der3d(const mf_der *dc, int axis, int order, const double *x, double *out,
             int nx, int ny, int nz, void *stream);
int mf_der3d_host(const mf_der *dc, int axis, int order, const double *x, double *out,
                  int nx, int ny, int nz);
/* mf_der_lines on host arrays (the Fortran der/dder/der_s/dder_s of one
 * line, class_derivatives.f90:186-260).  The host-array entry points stage
 * through a device buffer kept in the handle (grown to the largest call),
 * so repeated calls allocate nothing; calls on one handle are serialised. */
int mf_der_lines_host(const mf_der *dc, int order, int sub, const double *x, double *out,
                      int nlines, long long ld);
/* mf_der3d of nfields (1..8) fields of one shape in one launch: x[f] -> out[f]
 * (host arrays of device pointers; no output may alias an input).  The
 * line recurrences are latency-bound, so fields batched together run in
 * about the time of one. */
int mf_der3d_batch(const mf_der *dc, int axis, int order, int nfields, const double *const *x,
                   double *const *out, int nx, int ny, int nz, void *stream);

/* ---------------------------------------------------------------------
 * Helmholtz / Poisson solver:  d2u/dx2 + d2u/dy2 + d2u/dz2 + sigma u = g.
 * --------------------------------------------------------------------- */

/* Full grids gx(nx), gy(ny), gz(nz) including boundary points;
 * bct = {x1,x2,y1,y2,z1,z2}. */
int mf_solver3d_create(const double *gx, int nx, const double *gy, int ny,
                       const double *gz, int nz, const int bct[6], int so,
                       mf_solver3d **out);
int mf_solver3d_destroy(mf_solver3d *sc);

/* Solve with g(nx,ny,nz) (only the interior is read; g is not modified),
 * u(nx,ny,nz) inout (interior and the 6 faces written, the 12 edges left
 * untouched).  precision MF_FP64: the reference algorithm in FP64.
 * MF_MIXED / MF_MIXED_BF16X3: FP32 / split-bf16 MFMA transforms with
 * FP64-residual iterative refinement until ||r||/||b|| <= tol (tol <= 0
 * selects 1e-12) or max_iter solves (max_iter <= 0 selects 8, at most 32);
 * for tol < 1e-6 at least two solves are made (one FP32-class solve leaves
 * ~1e-7).  Both synchronise the stream (they read back residual norms). */
int mf_solver3d_solve(mf_solver3d *sc, const double *g, double sigma, double *u,
                      const double *bcx, const double *bcy, const double *bcz,
                      int precision, double tol, int max_iter,
                      mf_solve_report *rep, void *stream);
int mf_solver3d_solve_host(mf_solver3d *sc, const double *g, double sigma, double *u,
                           const double *bcx, const double *bcy, const double *bcz,
                           int precision, double tol, int max_iter,
                           mf_solve_report *rep);

/* Interior sizes of a solver handle (n1 = nx-2, ...). */
int mf_solver3d_dims(const mf_solver3d *sc, int *n1, int *n2, int *n3);

/* Unit-impulse responses for the influence matrix of a subdomain
 * (md_influence_matrix_set, src/class_field/class_field.f90:1396-1510).
 * For impulses first .. first+count-1 on face `face` (0..5 = x1,x2,y1,y2,z1,z2;
 * face points numbered column-major over the face's interior, as the
 * reference's loop `do j; do i`), solve with g = 0, a unit Dirichlet value at
 * that point and zero elsewhere, then write the inward-normal first
 * derivatives on the six faces (+d/dn on the low face, -d/dn on the high
 * face, the reference's sign rule :1465-1490) as
 *   out[b*stride ...] = [bcx(n2,n3,2) | bcy(n1,n3,2) | bcz(n1,n2,2)],
 * stride = 2(n2 n3 + n1 n3 + n1 n2) (interior sizes).  dx, dy, dz are
 * solver-mode derivative handles on the subdomain's grids; the impulse face
 * must be Dirichlet.  out is a device pointer; enqueued on stream. */
int mf_influence_impulses(mf_solver3d *sc, const mf_der *dx, const mf_der *dy,
                          const mf_der *dz, double sigma, int face, long long first,
                          int count, double *out, void *stream);

/* ---- Multidomain (influence-matrix) solve --------------------------------
 * nd[0] x nd[1] x nd[2] subdomains sharing their interface planes
 * (md_influence_matrix_param, src/class_md/class_md.f90:1450-1630;
 * multidomain_solve, src/class_field/class_field.f90:1295-1388; md_solve,
 * src/class_md/class_md.f90:379-438).  npts: the point count of every slab,
 * x slabs then y then z (nd[0] + nd[1] + nd[2] ints, each >= 10); grids: the
 * slab grids concatenated in the same order; consecutive slabs share their
 * end point.  Subdomain s = a + nd[0] (b + nd[1] c) has the grids of slabs a,
 * b, c.  bct: outer boundary codes (x1,x2,y1,y2,z1,z2; 1 Dirichlet, 2
 * Neumann); interfaces carry Dirichlet data.  Interfaces are numbered as
 * the reference does (x-subdomain index outermost, then y, then z; the x-,
 * y-, z-interface above each subdomain in turn), their unknowns are the
 * interior points of the interface plane, column-major.  Create builds one
 * solver and three solver-mode derivative handles per subdomain and
 * assembles the influence matrix (md_influence_matrix_set) on the device, as
 * one dense block per subdomain over the unknowns of its interfaces; sigma
 * is fixed at creation (the matrix depends on it).  Synchronous. */
typedef struct mf_md mf_md;
typedef struct mf_md_report {
  int iterations;   /* GMRES iterations                                  */
  double residual;  /* final ||rhs - M x||_2 / ||rhs||_2                 */
} mf_md_report;
int mf_md_create(const int nd[3], const int *npts, const double *grids, const int bct[6],
                 int so, double sigma, mf_md **out);
int mf_md_destroy(mf_md *h);
/* Subdomain count, interface unknowns and bytes of the influence blocks. */
int mf_md_info(const mf_md *h, int *nsub, long long *ninf, long long *matrix_bytes);
/* y = M x for device vectors of ninf interface unknowns (the influence
 * matrix, KSP's operator). */
int mf_md_matvec(mf_md *h, const double *x, double *y, void *stream);
/* Solve: g[s], u[s] are device full-size fields of subdomain s; bc[s] its
 * boundary arrays [bcx(n2,n3,2) | bcy(n1,n3,2) | bcz(n1,n2,2)] (outer faces
 * used, interface faces ignored).  The interface system is solved by
 * restarted GMRES(30) from x = 0 until the relative residual is <= tol
 * (<= 0 selects 1e-10) or max_iter iterations (<= 0: 1000); then every
 * subdomain is solved with the interface values, FP64.  x_iface (optional,
 * device, ninf doubles) receives the interface values.  Returns
 * MF_ERR_NOCONV (with u written) when GMRES stops above 10 tol.
 * Synchronises the stream. */
int mf_md_solve(mf_md *h, const double *const *g, double *const *u, const double *const *bc,
                double tol, int max_iter, double *x_iface, mf_md_report *rep, void *stream);
/* Initial guess of the interface solve (md_solve with inf_sol,
 * src/class_md/class_md.f90:389-424): when enabled, every solve starts GMRES
 * from s1 - 3 s2 + 3 s3 of the last three interface solutions (oldest s1;
 * zero before there are three) and then shifts its own solution in.
 * Enabling or disabling clears the history. */
int mf_md_set_guess(mf_md *h, int enable);
/* Influence-matrix cache (the role of md_influence_matrix_write/read,
 * src/class_md/class_md.f90:950-985, a PETSc binary file there): a binary
 * file of this library's own format, a header with the layout's identity
 * (nd, bct, so, sigma, a hash of the slab sizes and grids, sizes) then the
 * dense blocks.  mf_md_create_cached reads the blocks from path when its
 * header matches the arguments, else assembles them as mf_md_create and
 * writes path; *loaded (optional) tells which. */
int mf_md_write_matrix(const mf_md *h, const char *path);
int mf_md_create_cached(const int nd[3], const int *npts, const double *grids, const int bct[6],
                        int so, double sigma, const char *path, int *loaded, mf_md **out);
/* <- mpid%periods (md_mpi_init, class_md.f90) as used by
 *    md_influence_matrix_param (class_md.f90:1482-1560): mf_md_create with
 *    periodic axes (periods[a] != 0; NULL = none).  Along a periodic axis the
 *    last subdomain's high face and the first one's low face are one more
 *    interface (numbered where the last subdomain's interface of that axis
 *    stands; one subdomain along the axis: its two faces carry the same
 *    interface); the grids of such an axis run over one period.  With path
 *    non-NULL, as mf_md_create_cached (the cache records the periods). */
int mf_md_create_ex(const int nd[3], const int *npts, const double *grids, const int bct[6],
                    const int periods[3], int so, double sigma, const char *path, int *loaded,
                    mf_md **out);
/* <- psm (command_line.f90:27-29; navier_initialization :2119-2121), passed
 *    as null= to md_solve_init (class_md.f90:318-328) and multidomain_solve
 *    (class_field.f90:1338-1352) for singular interface systems (sigma = 0,
 *    outer faces Neumann or periodic): 0 solve as is; 1 the constant is the
 *    null space: removed from the right-hand side and from every Krylov
 *    vector, so x keeps the mean of its start (zero without the guess
 *    history) and M x = rhs - c 1; 2 the first right-hand-side entry is set
 *    to zero (md_vector_zero_lastpoint :251-267; the reference does so on the
 *    pressure solve, the one caller that passes null).  Default 0. */
int mf_md_set_null_space(mf_md *h, int psm);
/* No counterpart in the reference: device time of the influence-block
 * assembly of mf_md_create* (0 when the blocks came from a cache file or the
 * unit-impulse path), the FP64 flops of its grouped block products and the
 * bytes of blocks they wrote. */
int mf_md_assembly_stats(const mf_md *h, double *device_ms, double *gemm_flop, double *block_bytes);
/* No counterpart in the reference: the measurement of the autotuner
 * (mflops3d-md_amd/tune/infl_autotune.py) that chooses the tile shape of the
 * grouped influence-block GEMM: one M x N x K block product through that
 * kernel with tile class tile (0: 256x64, 1: 128x128, 2: 64x256) and
 * blocks_per_cu persistent blocks per compute unit; mean ms per launch over
 * reps launches.  Allocates its operands. */
int mf_tune_infl_gemm(int tile, int blocks_per_cu, int M, int N, int K, int reps, double *ms);

/* Kernel timing (no counterpart in the reference; its solve_3d has
 * commented-out cpu_time probes, class_solver_3d.f90:106-259).  While
 * enabled, every launch of the solve is bracketed by HIP events on its
 * stream.  mf_solver3d_profile_read synchronises and returns, per kernel
 * class, the summed milliseconds, launch counts and algorithmic work since
 * the last reset (flops for the transforms, minimum HBM bytes for the
 * memory-bound kernels):
 * classes 0 BC fold, 1 FP64 transform, 2 FP32 transform, 3 FP64 residual,
 * 4 right-hand side multiply-through, 5 faces. */
#define MF_PROF_CLASSES 6
int mf_solver3d_profile(mf_solver3d *sc, int enable);
int mf_solver3d_profile_read(mf_solver3d *sc, double ms[MF_PROF_CLASSES],
                             int count[MF_PROF_CLASSES], double work[MF_PROF_CLASSES],
                             int reset);
/* Kernel durations of the transforms alone (no counterpart in the
 * reference): the first five transforms of a solve at `precision` (x, y, z
 * with the divide, z and y backward) launched reps times back to back on the
 * handle's own work arrays (their contents are overwritten; no caller data
 * is touched), between two events.  They are five launches, or four for
 * MF_MIXED_FP16X2 when its z pair runs as one kernel (n3 in 129..256); each
 * launch reads one FP32 field and writes one.  When MF_MIXED_FP16X2 runs its
 * x and y transforms as pairs per z-plane (n1, n2 in 129..256, n3 >= half
 * the CUs), the timed launches are those of a whole inner solve instead:
 * the x+y forward pair, the z transforms (one launch when fused) and the
 * y+x backward pair, which writes an FP64 scratch field of the full size
 * (allocated for the call).  ms_per_launch = elapsed / (launches reps);
 * flops_per_launch (may be NULL) = the algorithmic 2 M N K of the timed
 * transforms over the launches.  Per-launch event pairs
 * (mf_solver3d_profile) also count the host's enqueue gaps; this does not.
 * Synchronises the stream. */
int mf_solver3d_time_transforms(mf_solver3d *sc, int precision, int reps, double *ms_per_launch,
                                double *flops_per_launch, void *stream);

/* ---------------------------------------------------------------------
 * Prediction-projection Navier-Stokes step (navier3d.f90:48-286,
 * class_navier_3D.f90), single domain, on the device.  The reference's
 * options are nlt (nonlinear form: 2 skew-symmetric, 1 advective), tou (BDF
 * order 1..3) and top (pressure extrapolation order top-1, 1..3); pt=1
 * (non-incremental rotational correction with the tangential boundary
 * gradient, add_boundary_gradient :695-791), no mapping, one sub-iteration.
 * Velocity is Dirichlet and phi Neumann on all six faces
 * (navier_initialization :2166-2169).  Fields are full (nx,ny,nz) device
 * arrays, bc arrays as for the solver.
 * --------------------------------------------------------------------- */
typedef struct mf_ns mf_ns;

/* <- navier_initialization (class_navier_3D.f90:2084-2345): grids, Re, dt,
 *    BDF factors, derivative and solver coefficients; fields zero. */
int mf_ns_create(const double *gx, int nx, const double *gy, int ny, const double *gz, int nz,
                 double re, double dt, int tou, int top, int nlt, int so, mf_ns **out);
int mf_ns_destroy(mf_ns *ns);

/* Velocity boundary values of component c (0,1,2 = u,v,w), kept until set
 * again (<- navier_bc_velocity :395-433, whose values come from sol()). */
int mf_ns_set_bc(mf_ns *ns, int c, const double *bcx, const double *bcy, const double *bcz,
                 void *stream);
/* Constant body force of component c (full field), or NULL for none
 * (<- the sol()-based part of f(nav,var) :1043-1070). */
int mf_ns_set_forcing(mf_ns *ns, int c, const double *f, void *stream);
/* Constant-flow-rate force on u (<- f(nav,'rhsu') :1124-1164): uniform g
 * with g = g0 at the first step (g0 <= 0: the reference's 3 * flow rate),
 * then g -= (Q - Q0) fac1, Q the weighted volume mean of extrapolated u. */
int mf_ns_set_flow_rate(mf_ns *ns, int enable, double g0, void *stream);
/* All time levels of u, v, w := the given fields (edges zeroed); p = phi = 0. */
int mf_ns_set_state(mf_ns *ns, const double *u, const double *v, const double *w, void *stream);
/* nsteps time steps.  The three Helmholtz velocity solves use `precision`
 * (MF_FP64, MF_MIXED, MF_MIXED_BF16X3 or MF_MIXED_FP16X2, refined to tol);
 * the all-Neumann pressure Poisson solve is always FP64 (its null mode makes
 * the right-hand side's incompatible part an irreducible residual). */
int mf_ns_step(mf_ns *ns, int nsteps, int precision, double tol, void *stream);
/* Latest time level of u, v, w, p, phi into device arrays (any may be NULL). */
int mf_ns_get(const mf_ns *ns, double *u, double *v, double *w, double *p, double *phi,
              void *stream);
/* Simulated time and the current uniform flow-rate force (synchronises). */
int mf_ns_info(const mf_ns *ns, double *time, double *flow_force, void *stream);
/* <- les_type / les_c (navier_initialization :2137-2138), navier_LES
 *    (:1225-1452) and the LES part of f(nav,var) (:1083-1134): from the next
 *    step on, each step computes the eddy viscosity nu of model les_type
 *    (0 = off, 1 = Smagorinsky, 2 = WALE, 3 = sigma; constant c, squared
 *    as the reference does) from the latest level's velocity gradients and
 *    adds sum_j d/dx_j (nu (du_i/dx_j + du_j/dx_i)) to the force of u_i. */
int mf_ns_set_les(mf_ns *ns, int les_type, double c, void *stream);
/* <- pt and nsubite (command_line.f90:22,32; navier_initialization
 *    class_navier_3D.f90:2133,2136): projection type 1 (default:
 *    non-incremental, tangential boundary gradient), 2 (incremental: the
 *    velocity sees grad extrapol_p(p)), 3 (velocity correction: pressure
 *    first, rot-rot Neumann term, add_boundary_rotrot :793-846) or 4
 *    (incremental velocity correction), and at most nsubite sub-iterations
 *    per step (navier3d.f90:56-147; from the second on every extrapolation
 *    returns the current iterate and testconv :366-412 stops the loop when
 *    the relative changes of |u| and of p fall below 1e-9; each test
 *    synchronises the stream).  Clears the stored presolve right-hand sides
 *    (rhs_u(1:2), zero after navier_initialization); call before stepping.
 *    MF_ERR_ARG for pt outside 1..4 or nsubite < 1. */
int mf_ns_set_scheme(mf_ns *ns, int pt, int nsubite, void *stream);
/* Sub-iterations made by the last step and the velocity and pressure
 * changes of its last convergence test (0 when it made one). */
int mf_ns_subiterations(const mf_ns *ns, int *count, double *dvel, double *dp);
/* nu of the last step with LES into a device field (<- the les_nu output
 * field, class_navier_3D.f90:101). */
int mf_ns_les_nu(const mf_ns *ns, double *nu, void *stream);
/* <- restart_write (class_navier_3D.f90:186-257): directory dirname (NULL:
 *    "restart_" // tn(time), tn = floor(time) and 8 decimals, 8 digits each,
 *    :169-183; its name into dir_out when given) holding param.dat (time as
 *    real(8), then it(1:4) as integer(4), unformatted stream, native byte
 *    order) and NetCDF files: for each level i of it(:) vel_u_i, vel_v_i,
 *    vel_w_i, vel_p_i, vel_phi_i (FP64; variables U, V, W, P, P) and
 *    rhs_u_i .. rhs_p_i, then rhs_phi_4 (RHS_*; written as zero fields: the
 *    reference rebuilds them from zero before each use, :1615, :1810), and the
 *    latest level in FP32 as vel_u, vel_v, vel_w, vel_p (the reference puts
 *    w's data in vel_p, :256; here vel_p holds p).  Synchronises. */
int mf_ns_restart_write(const mf_ns *ns, const char *dirname, char *dir_out, int dir_out_len,
                        void *stream);
/* <- restart_read (:259-340): time, it(:) and every level's u, v, w, p, phi
 *    from dirname (NULL: "restart"); the phi gradients of the boundary
 *    correction are recomputed and the flow-rate force re-initialises at the
 *    next step (its state is not part of the files).  MF_ERR_ARG when
 *    param.dat is missing or a file does not match the handle's sizes. */
int mf_ns_restart_read(mf_ns *ns, const char *dirname, void *stream);

/* ---------------------------------------------------------------------
 * Field files (NetCDF classic, host arrays; no GPU needed).
 * --------------------------------------------------------------------- */
/* <- write_field / write_var3d (class_field.f90:528-579, class_io.f90:23-124):
 *    f(nx,ny,nz) (host, column-major) as variable var_name on dimensions
 *    dim_names (NULL: "resolution_x", "resolution_y", "resolution_z", the
 *    type(field) defaults) in FP64 (dbl != 0) or FP32.  An existing file
 *    keeps its other variables; dimensions of the same name must have the
 *    same length (the reference's "wrong dimensions" stop -> MF_ERR_ARG). */
int mf_io_write_field(const char *path, const char *var_name, const char *const dim_names[3],
                      int nx, int ny, int nz, const double *f, int dbl);
/* <- get_var3d_info (class_io.f90:191-223): sizes (x, y, z) of var_name,
 *    the dimension names into dim_names (3 strings of name_len bytes, may be
 *    NULL) and whether it is stored in FP64. */
int mf_io_field_info(const char *path, const char *var_name, int dims[3], char *dim_names,
                     int name_len, int *is_double);
/* <- read_var3d (class_io.f90:126-189): var_name into f(nx,ny,nz) (host,
 *    FP64; FP32 files are converted); the sizes must match. */
int mf_io_read_field(const char *path, const char *var_name, int nx, int ny, int nz, double *f);

/* ---------------------------------------------------------------------
 * Spline filter, eddy viscosities and the learned SGS closure.
 * --------------------------------------------------------------------- */
typedef struct mf_filter mf_filter;

/* <- LSQUnivariateSpline (src/class_filter/class_filter.f90:130-174): the
 *    least-squares spline of degree k (1..5) on the knots of curfit's
 *    iopt = -1 call (k+1 copies of xi(1) and xi(nin), then xn(2:nn-1)),
 *    fitted to values on xi(nin) with weights wi (curfit minimises
 *    sum (wi (yi - s(xi)))^2; NULL = unit weights), then evaluated at
 *    xo(nout) (deriv 0: splev, the reference's yo) or differentiated there
 *    (deriv 1: splder, its ypo); nout <= 256.  The fit is linear in the
 *    values: the handle holds F (nout x nin).  xi and xn increase,
 *    xn(1) = xi(1) and xn(nn) = xi(nin) (to rounding: 16 ulps of the
 *    interval; the knots take xi's end points); MF_ERR_SETUP when a spline has no
 *    data point in its support (curfit's Schoenberg-Whitney condition). */
int mf_filter_create(const double *xi, const double *wi, int nin, const double *xo, int nout,
                     const double *xn, int nn, int k, int deriv, mf_filter **out);
int mf_filter_destroy(mf_filter *h);
/* F into a host array of nout*nin doubles, column-major. */
int mf_filter_matrix(const mf_filter *h, double *F);
/* The filter of a 3-D field, one 1-D filter per axis (the reference applies
 * LSQUnivariateSpline line by line, ls_spline_1d.f90): in(nxi,nyi,nzi) ->
 * out(nxo,nyo,nzo) with nxi = fx.nin, nxo = fx.nout and so on.  work holds
 * nxo*nyi*nzi + nxo*nyo*nzi doubles.  Device pointers. */
int mf_filter3d(const mf_filter *fx, const mf_filter *fy, const mf_filter *fz, const double *in,
                double *out, double *work, void *stream);
/* mf_filter3d with no work argument and no allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * two intermediate fields live in a library-owned device array of
 * MF_FILTER_SCRATCH_DOUBLES doubles (48 MiB, enough for 256^3 -> 64^3), one
 * per GPU, held until the process ends; MF_ERR_ARG when
 * nxo*nyi*nzi + nxo*nyo*nzi exceeds it.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (other streams, other threads) race.
 * Same kernels and results as mf_filter3d; enqueued on stream. */
#define MF_FILTER_SCRATCH_DOUBLES 6291456
int mf_filter3d_noalloc(const mf_filter *fx, const mf_filter *fy, const mf_filter *fz,
                        const double *in, double *out, void *stream);

/* <- navier_LES (src/class_navier/class_navier_3D.f90:1225-1449), les_sigma
 *    (:1453-1497): the eddy viscosity nu(nx,ny,nz) from the nine velocity
 *    gradients du[3 i + j] = d u_i / d x_j (device fields), les_type
 *    1 = Smagorinsky c sqrt(2 S:S), 2 = WALE, 3 = sigma model, times
 *    (dx^2 + dy^2 + dz^2) / 3 (one-sided spacings on the first and last
 *    points, centred ones inside; host grids), zero on the box edges.
 *    c is the model constant squared (the reference's (C_s)^2 etc.).
 *    Synchronises the stream. */
int mf_les_nu(const double *const *du, const double *gx, int nx, const double *gy, int ny,
              const double *gz, int nz, int les_type, double c, double *nu, void *stream);

/* Learned SGS closure: MLP 10 -> 64 -> 64 -> 6 (tanh hidden layers) from
 * (the nine filtered velocity gradients, the filter width) to the six SGS
 * stresses (xx, yy, zz, xy, xz, yz).  5254 FP32 master parameters in the
 * order W1 (64x10, row-major), b1, W2 (64x64), b2, W3 (6x64), b3; the
 * kernels use bf16 copies of W1..W3 (round to nearest even) and FP32
 * arithmetic.  No counterpart in the reference (its filter module prepares
 * the data for such a closure, class_filter.f90). */
typedef struct mf_mlp mf_mlp;
int mf_mlp_create(const float *params, mf_mlp **out);  /* host params */
int mf_mlp_destroy(mf_mlp *h);
int mf_mlp_params(const mf_mlp *h, float *params);     /* into a host array */
/* Y(n,6) = MLP(X(n,10)), row-major device arrays. */
int mf_mlp_forward(const mf_mlp *h, const float *X, float *Y, long long n, void *stream);
/* loss = mean over the n samples of sum_c (Y - T)^2 and its gradient with
 * respect to the parameters (taken at the bf16 weights) into grad (device,
 * 5254 floats; may be NULL).  loss (host, may be NULL) synchronises.  The
 * per-block partial gradients are summed in a fixed order: bitwise
 * reproducible. */
int mf_mlp_loss_grad(mf_mlp *h, const float *X, const float *T, long long n, float *loss,
                     float *grad, void *stream);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected, rate lr)
 * on the full batch; loss (host, may be NULL) is the loss before the step. */
int mf_mlp_adam_step(mf_mlp *h, const float *X, const float *T, long long n, float lr,
                     float *loss, void *stream);

/* Training data of the closure from a fine field (the use the reference's
 * filter module serves, class_filter.f90): ubar = F u, vbar, wbar on the
 * filter's output grid and tau (n x 6 float, row-major, n the output
 * points in Fortran order) with tau_ij = (F(u_i u_j) - F(u_i) F(u_j)) /
 * tscale, columns xx, yy, zz, xy, xz, yz.  work holds nin_x nin_y nin_z +
 * nout_x nout_y nout_z + nout_x nin_y nin_z + nout_x nout_y nin_z doubles. */
int mf_sgs_targets(const mf_filter *fx, const mf_filter *fy, const mf_filter *fz, const double *u,
                   const double *v, const double *w, double *ubar, double *vbar, double *wbar,
                   float *tau, double tscale, double *work, void *stream);
/* Closure inputs X (n x 10 float) of a velocity field (nx,ny,nz): the
 * gradients du_i/dx_j (row-major ij, derivative handles in solver mode, as
 * navier_LES takes them, zero on the box edges) times gscale, and
 * lscale * Delta with Delta^2 = (dx^2+dy^2+dz^2)/3 (host grids).  work holds
 * 9 nx ny nz doubles.  Synchronises the stream. */
int mf_sgs_features(const mf_der *dx, const mf_der *dy, const mf_der *dz, const double *u,
                    const double *v, const double *w, int nx, int ny, int nz, const double *gx,
                    const double *gy, const double *gz, double gscale, double lscale, float *X,
                    double *work, void *stream);
/* No counterpart in the reference: a-posteriori-consistent training targets
 * for the learned closure.  For feature rows X (n x 10, as mf_sgs_features
 * with gscale = lscale = 1) the Smagorinsky stress of constant c,
 * T = -2 c^2 Delta^2 |S| S (n x 6: xx, yy, zz, xy, xz, yz), S the strain
 * rate of the nine gradients, |S| = sqrt(2 S:S); with c fitted so that
 * coarse runs follow the filtered fine run, and the closure coupled as an
 * eddy viscosity (mf_ns_set_sgs_coupling 1), whose projection of this
 * stress is exactly nu = c^2 Delta^2 |S|. */
int mf_sgs_eddy_targets(const float *X, long long n, double c, float *T, void *stream);
/* The learned closure in the NS step (tau enters the momentum equation as
 * -div tau, as the eddy-viscosity stress does in f(nav,var) :1083-1134):
 * each step takes X of the latest level as mf_sgs_features does, tau =
 * tscale * mlp(X), and adds -sum_j d_j tau_ij (tau zero on the box edges)
 * to the force of u_i, on top of any eddy viscosity.  mlp NULL turns it
 * off; the handle must outlive its use. */
int mf_ns_set_sgs(mf_ns *ns, const mf_mlp *mlp, double gscale, double lscale, double tscale,
                  void *stream);
/* No counterpart in the reference (its closures are eddy viscosities): how
 * the learned stress enters the momentum force.  0 (default): -div tau.
 * 1: as the eddy viscosity nu = max(0, -tau:S) / (2 S:S) of the strain rate
 * S of the same gradients (the stress projected on the dissipative
 * direction, backscatter clipped), added as the LES force is; the velocity
 * stays bounded where the raw stress's divergence is noisy at the grid
 * scale.  mf_ns_les_nu then returns this nu. */
int mf_ns_set_sgs_coupling(mf_ns *ns, int coupling);

#ifdef __cplusplus
}
#endif
#endif /* MFLOPS_MF_H */
