// Matrix-diagonalisation Helmholtz/Poisson solver on gfx950
// (solve_3d, class_solver_3d.f90:83-299; solver_3d, class_field.f90:1072-1109).
//
// FP64 path (the reference algorithm):
//   fold    R = g_int - sum_a inv(lhs_a) cl_a              (BC fold, :104-175)
//   X/Y/Z   forward eigenvector transforms (MFMA FP64), divide fused into Z
//   Z/Y/X   backward transforms, X scatters into u's interior
//   faces   Dirichlet copy / Neumann extrapolation          (:261-297, :658-776)
//
// Mixed path: the same six transforms in FP32 on MFMA inside an iterative
// refinement whose residual is taken in FP64 with the solver's own discrete
// operator.  The operator sum_a inv(lhs_a) rhs_a + sigma is multiplied
// through by P = lhs_x (x) lhs_y (x) lhs_z, which commutes with every term, so
// the residual r' = P g_f - B u with
//   B = rhs_x lhs_y lhs_z + lhs_x rhs_y lhs_z + lhs_x lhs_y rhs_z + sigma P
// is a product of banded 1-D operators (no line solves), and the inner solve
// uses W'_a = inv(V_a) inv(lhs_a) in place of inv(V_a), which absorbs P^-1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gemm.h"
#include "gemm_bf3.h"
#include "gemm_h2.h"
#include "gemm_h2r.h"
#include "gemm_h2p.h"
#include "gemm_h2z.h"
#include "infl_spec.h"
#include "infl_tune.h"
#include "influence.h"
#include "internal.h"
#include "residual.h"
#include "residual_zs.h"
#include "stage.h"
#include "mflops/mf.h"

namespace mf {

// ---------------------------------------------------------------------------
// Banded 1-D operator rows: row i acts on columns c0[i] .. c0[i]+6.
// ---------------------------------------------------------------------------
struct Band {
  const int *c0;
  const double *w;  // (n,8) row-major, 7 used
};

// BC fold: R(i,j,k) = g(i+1,j+1,k+1) - sum over the 6 faces of
// fold_a(idx)*bc_a  with fold vectors fx1(i) = inv_lhs(i,1) ddr(1,1) +
// inv_lhs(i,2) ddr(2,1), fx2(i) = inv_lhs(i,n-1) ddr(n-1,8) + inv_lhs(i,n) ddr(n,7).
__global__ void __launch_bounds__(256)
    fold_kernel(const double *__restrict__ g, double *__restrict__ R, int n1, int n2, int n3,
                const double *__restrict__ fold,  // fx1 fx2 fy1 fy2 fz1 fz2
                const double *__restrict__ bcx, const double *__restrict__ bcy,
                const double *__restrict__ bcz) {
  // launch: block (64,4), grid (ceil(n1/64), ceil(n2/4), n3)
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= n1 || j >= n2) return;
  const long p = i + (long)n1 * (j + (long)n2 * k);
  const int nx = n1 + 2, ny = n2 + 2;
  const double *fx1 = fold, *fx2 = fx1 + n1, *fy1 = fx2 + n1, *fy2 = fy1 + n2;
  const double *fz1 = fy2 + n2, *fz2 = fz1 + n3;
  double v = g[(i + 1) + (long)nx * ((j + 1) + (long)ny * (k + 1))];
  const long sx = (long)n2 * n3, sy = (long)n1 * n3, sz = (long)n1 * n2;
  v -= fx1[i] * bcx[j + (long)n2 * k];
  v -= fx2[i] * bcx[sx + j + (long)n2 * k];
  v -= fy1[j] * bcy[i + (long)n1 * k];
  v -= fy2[j] * bcy[sy + i + (long)n1 * k];
  v -= fz1[k] * bcz[i + (long)n1 * j];
  v -= fz2[k] * bcz[sz + i + (long)n1 * j];
  R[p] = v;
}

// Null component of a singular problem (sigma = 0 and a zero eigenvalue on
// every axis): c0 = (w0x (x) w0y (x) w0z) . R / (sum w0x sum w0y sum w0z), with
// w0a the row of inv(V_a) of the zero eigenvalue, is removed from the folded
// right-hand side before the transforms.  The constant is the operator's
// right null vector and every other left eigenvector is orthogonal to it,
// so the modes the divide keeps are those of the reference in exact
// arithmetic; in floating point the incompatible part would otherwise leak
// into them through the rounding of their eigenvectors.
// Pass 1: NULL_SLICES blocks per z-plane, each over a contiguous run of
// rows: part[k NULL_SLICES + s] = w0z(k) sum_(ij in slice s) w0x(i) w0y(j) R(i,j,k).
constexpr int NULL_SLICES = 8;
__global__ void __launch_bounds__(256)
    null_dot_kernel(const double *__restrict__ R, int n1, int n2, const double *__restrict__ w0,
                    int n3, double *__restrict__ part) {
  const int k = blockIdx.x, sl = blockIdx.y;
  const double *wx = w0, *wy = w0 + n1, *wz = w0 + n1 + n2;
  const int rows = (n2 + NULL_SLICES - 1) / NULL_SLICES;
  const int j0 = sl * rows, j1 = min(n2, j0 + rows);
  const double *Rk = R + (long)n1 * n2 * k + (long)n1 * j0;
  const int np = j1 > j0 ? (j1 - j0) * n1 : 0;
  double acc = 0.0;
  for (int p = threadIdx.x; p < np; p += 256) {
    const int jj = p / n1, i = p - jj * n1;
    acc += wx[i] * wy[j0 + jj] * Rk[p];
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[k * NULL_SLICES + sl] = red[0] * wz[k];
}

// Pass 2: every block sums the npart partials in the same order (so all see
// the same c0) and subtracts c0 from its points.
__global__ void __launch_bounds__(256)
    null_sub_kernel(double *__restrict__ R, long N, const double *__restrict__ part, int npart,
                    double inv_wsum) {
  __shared__ double red[256];
  double t = 0.0;
  for (int q = threadIdx.x; q < npart; q += 256) t += part[q];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double c0 = red[0] * inv_wsum;
  for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < N; p += (long)gridDim.x * 256) R[p] -= c0;
}

// Faces of u: Dirichlet copies bc; Neumann solves the one-sided 7-point
// first-derivative formula for the boundary value
// (class_solver_3d.f90:261-297, :658-776).  Edges are not touched.
__global__ void __launch_bounds__(256)
    faces_kernel(double *__restrict__ u, int n1, int n2, int n3, const int *__restrict__ bct,
                 const double *__restrict__ neum,  // 3 x (2,7) column-major
                 const double *__restrict__ bcx, const double *__restrict__ bcy,
                 const double *__restrict__ bcz) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long fx = (long)n2 * n3, fy = (long)n1 * n3, fz = (long)n1 * n2;
  const long nx = n1 + 2, ny = n2 + 2;
  auto U = [&](long i, long j, long k) -> double & { return u[i + nx * (j + ny * k)]; };
  int axis, side;
  long a, b;  // interior face indices (0-based)
  if (t < 2 * fx) { axis = 0; side = (int)(t / fx); long r = t % fx; a = r % n2; b = r / n2; }
  else if ((t -= 2 * fx) < 2 * fy) { axis = 1; side = (int)(t / fy); long r = t % fy; a = r % n1; b = r / n1; }
  else if ((t -= 2 * fy) < 2 * fz) { axis = 2; side = (int)(t / fz); long r = t % fz; a = r % n1; b = r / n1; }
  else return;
  const int code = bct[2 * axis + side];
  const double *nm = neum + 14 * axis;  // nm[s + 2*m] = neum(s+1, m+1)
  const int n = axis == 0 ? n1 : (axis == 1 ? n2 : n3);
  const double *bcarr = axis == 0 ? bcx : (axis == 1 ? bcy : bcz);
  const long fsz = axis == 0 ? fx : (axis == 1 ? fy : fz);
  const long lenA = axis == 0 ? n2 : n1;
  const double bcv = bcarr[side * fsz + a + lenA * b];
  // position along the axis -> full-array index
  auto at = [&](long q) -> double & {
    if (axis == 0) return U(q, a + 1, b + 1);
    if (axis == 1) return U(a + 1, q, b + 1);
    return U(a + 1, b + 1, q);
  };
  if (code == 1) {
    at(side ? n + 1 : 0) = bcv;
  } else if (code == 2 || code == 3) {
    if (side == 0) {
      double acc = bcv;
      for (int m = 1; m < 7; ++m) acc -= nm[0 + 2 * m] * at(m);
      at(0) = acc * (1.0 / nm[0]);
    } else {
      double acc = bcv;
      for (int m = 5; m >= 0; --m) acc -= nm[1 + 2 * m] * at(n - 5 + m);
      at(n + 1) = acc * (1.0 / nm[1 + 2 * 6]);
    }
  }
}

// Several norms at once, one block of 1024 threads each, in a fixed order:
// block i sums the n partials at part + i stride into *out.p[i].
constexpr int NORM_SLOTS = 3;
struct NormOuts {
  double *p[NORM_SLOTS];
};
__global__ void __launch_bounds__(1024) sum_slots_kernel(const double *__restrict__ part,
                                                         long stride, long n, NormOuts out) {
  __shared__ double s[1024];
  const double *pp = part + stride * blockIdx.x;
  double a = 0.0;
  for (long q = threadIdx.x; q < n; q += 1024) a += pp[q];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out.p[blockIdx.x][0] = s[0];
}

}  // namespace mf

// ---------------------------------------------------------------------------
// Handle
// ---------------------------------------------------------------------------
struct mf_solver3d {
  int nx = 0, ny = 0, nz = 0, n1 = 0, n2 = 0, n3 = 0, so = 6;
  mf::HostStage stage;  // device staging of mf_solver3d_solve_host
  int bct[6] = {1, 1, 1, 1, 1, 1};
  mf::Solver1D ax[3];
  // device memory
  void *dmem = nullptr;
  // FP64 eigen transforms: Wx = inv_vect_x, Vx = vect_x; Wy_t = inv_vect_y^T ...
  double *Wx = nullptr, *Vx = nullptr, *Wyt = nullptr, *Vyt = nullptr, *Wzt = nullptr, *Vzt = nullptr;
  float *Wx32 = nullptr, *Vx32 = nullptr, *Wyt32 = nullptr, *Vyt32 = nullptr, *Wzt32 = nullptr,
        *Vzt32 = nullptr;  // mixed: W' = inv(V) inv(lhs)
  // the same six matrices split into three bf16 parts, [part][k/8][row][k%8]
  // (row = the eigenvector-dimension index, zero-padded to wrows[axis] rows
  // and kp[axis] columns)
  __bf16 *Wsp[3] = {nullptr, nullptr, nullptr}, *Vsp[3] = {nullptr, nullptr, nullptr};
  // and into two fp16 parts of the matrix scaled by wsc (a power of two),
  // same layout
  _Float16 *Whs[3] = {nullptr, nullptr, nullptr}, *Vhs[3] = {nullptr, nullptr, nullptr};
  float wsc[3][2] = {{1, 1}, {1, 1}, {1, 1}};
  // the forward y and backward x fp16 splits again with k permuted within
  // each 32-wide k-step (h2p_perm), the second operands of the plane pairs
  _Float16 *Whp1 = nullptr, *Vhp0 = nullptr, *Vhp2 = nullptr;
  // largest magnitudes (float bits) of the fp16 path's fields: [0] r' / b',
  // [1..5] the outputs of the first five transforms
  unsigned *fmx = nullptr;
  int kp[3] = {0, 0, 0}, wrows[3] = {0, 0, 0};
  double *lam = nullptr;   // lx(n1) ly(n2) lz(n3)
  double *fold = nullptr;  // fx1 fx2 fy1 fy2 fz1 fz2
  double *neum = nullptr;  // 3 x 14
  int *bctd = nullptr;     // 6
  int *bandc = nullptr;    // [axis][M|R] row starts
  double *bandw = nullptr; // [axis][M|R] (n,8)
  mf::Band M[3], R[3];
  double *T1 = nullptr, *T2 = nullptr;  // interior FP64 work
  float *F1 = nullptr, *F2 = nullptr;   // interior FP32 work (mixed)
  float *R32 = nullptr;                 // b' / r' in FP32 (mixed)
  double *ebd = nullptr;                // ex1 ex2 ey1 ey2 ez1 ez2: P-folded boundary rows
  double *fbf = nullptr;                // face products fbx fby fbz (mixed)
  double *norms = nullptr;              // device scalars
  double *part = nullptr;               // per-block partial sums of norms, NORM_SLOTS arrays
  size_t part_stride = 0;               // doubles per partials array
  // residual norms whose partials are written but not yet summed (slots 0..)
  int npend = 0;
  unsigned pend_nb = 0;
  double *pend_out[3] = {nullptr, nullptr, nullptr};
  double *zq = nullptr;                 // Mz, Rz rows on planes b(k)..b(k)+4, e_z1, e_z2 (n3, 12)
  int res_cz = 0;                       // output planes per z-streaming residual block; 0: plane pipeline
  double *hnorms = nullptr;             // pinned host mirror
  // a mixed solve enqueued by solver3d_solve_begin and not yet finished by
  // solver3d_solve_end: its arguments and the sweeps enqueued so far
  struct Pending {
    bool active = false;
    const double *g = nullptr, *bcx = nullptr, *bcy = nullptr, *bcz = nullptr;
    double *u = nullptr;
    double sigma = 0.0, tol = 0.0;
    int precision = 0, max_iter = 0, it = 0;
    hipStream_t st = nullptr;
  } pending;
  // singular problems (sigma = 0, a zero eigenvalue on every axis): the
  // null rows w0x | w0y | w0z of inv(V), 1 / (sum w0x sum w0y sum w0z), and
  // n3 plane partials
  bool h2_resident = true;  // fp16x2 transforms with the eigenvectors in registers (MF_H2_RESIDENT=0: off)
  bool h2_plane = true;     // fp16x2 x/y transform pairs per z-plane (MF_H2_PLANE=0: off)
  bool h2_cols = true;      // fp16x2 z pair on column blocks (MF_H2_COLS=0: the row-tile kernel)
  int ncu = 256;        // compute units of the device the handle was created on
  bool null_mode = false;
  double *w0 = nullptr, *npart = nullptr;
  double inv_wsum = 0.0;
  void *iws = nullptr;                  // unit-impulse batch workspace (grown on demand)
  size_t iws_bytes = 0;
  // kernel timing (mf_solver3d_profile): event pairs recorded on the launch
  // stream, resolved at read time
  bool prof = false;
  struct Mark {
    int cls;
    double work;
  };
  std::vector<hipEvent_t> ev_pool;
  std::vector<Mark> marks;
  double prof_ms[MF_PROF_CLASSES] = {0};
  double prof_work[MF_PROF_CLASSES] = {0};
  int prof_cnt[MF_PROF_CLASSES] = {0};
};

namespace {
#define HIPCHK(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return mf::fail(MF_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

enum { P_FOLD = 0, P_T64 = 1, P_T32 = 2, P_RES = 3, P_MULT = 4, P_FACES = 5 };
constexpr int MF_MAX_REFINE = 32;  // cap on max_iter (residual norms kept per solve)

// Brackets the launches made during its lifetime with a pair of events.
struct ProfScope {
  mf_solver3d *h;
  hipStream_t st;
  bool on;
  ProfScope(mf_solver3d *h_, int cls, double work, hipStream_t st_) : h(h_), st(st_), on(h_->prof) {
    if (!on) return;
    const size_t k = 2 * h->marks.size();
    while (h->ev_pool.size() < k + 2) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) { on = false; return; }
      h->ev_pool.push_back(e);
    }
    h->marks.push_back({cls, work});
    (void)hipEventRecord(h->ev_pool[k], st);
  }
  ~ProfScope() {
    if (on) (void)hipEventRecord(h->ev_pool[2 * h->marks.size() - 1], st);
  }
};

// Side of the GEMM block tile along an eigenvector dimension of n: the
// smallest of 64/128/256 covering it (256 beyond).
inline int gemm_tile(int n) { return n <= 64 ? 64 : (n <= 128 ? 128 : 256); }

// float -> bf16 bits, round to nearest even (finite inputs)
inline uint16_t bf16_bits(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
inline float bf16_float(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float x;
  std::memcpy(&x, &u, 4);
  return x;
}
// x = p0 + p1 + p2 in bf16 (the device split3 in gemm_bf3.h, on the host)
inline void bf16_split3(float x, uint16_t p[3]) {
  p[0] = bf16_bits(x);
  const float r1 = x - bf16_float(p[0]);
  p[1] = bf16_bits(r1);
  const float r2 = r1 - bf16_float(p[1]);
  p[2] = bf16_bits(r2);
}

void transpose(int n, const std::vector<double> &a, std::vector<double> &t) {
  t.resize((size_t)n * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) t[j + (size_t)n * i] = a[i + (size_t)n * j];
}

// Rows of a dense banded matrix as (start column, BAND_W weights).  The
// window of row i is centred, [i-2, i+2], where the row's nonzeros allow, and
// always lies in [i-4, i+4] inside the line; starts are non-decreasing in i
// (the residual kernel's halo and plane range rely on both).
bool to_band(int n, const std::vector<double> &a, std::vector<int> &c0, std::vector<double> &w) {
  constexpr int W = mf::BAND_W, R = mf::RZ_H;
  c0.assign(n, 0);
  w.assign((size_t)8 * n, 0.0);
  for (int i = 0; i < n; ++i) {
    int first = n, last = -1;
    for (int j = 0; j < n; ++j)
      if (a[i + (size_t)n * j] != 0.0) { first = std::min(first, j); last = std::max(last, j); }
    if (last < 0) first = last = i;
    const int lo_min = std::max(std::max(0, last - (W - 1)), i - R);
    const int lo_max = std::min(std::min(first, n - W), i);
    if (lo_min > lo_max) return false;
    const int lo = std::min(std::max(i - W / 2, lo_min), lo_max);
    if (i > 0 && lo < c0[i - 1]) return false;
    c0[i] = lo;
    for (int m = 0; m < W; ++m) w[(size_t)8 * i + m] = a[i + (size_t)n * (lo + m)];
  }
  return true;
}

}  // namespace

extern "C" int mf_solver_coeffs_1d(const double *grid, int nall, const int bct[2], int so,
                                   double *ddl, double *ddr, double *eigen, double *vect,
                                   double *inv_vect, double *inv_lhs, double *neum) {
  if (!grid || !bct) return mf::fail(MF_ERR_ARG, "mf_solver_coeffs_1d: null pointer");
  mf::Solver1D s;
  std::string err;
  if (!mf::solver_setup_1d(grid, nall, bct, so, s, err))
    return mf::fail(err.find("at least") != std::string::npos || err.find("must") != std::string::npos
                        ? MF_ERR_ARG
                        : MF_ERR_SETUP,
                    "mf_solver_coeffs_1d: " + err);
  const size_t n = s.n;
  if (ddl) std::memcpy(ddl, s.ddl.data(), sizeof(double) * 5 * n);
  if (ddr) std::memcpy(ddr, s.ddr.data(), sizeof(double) * 8 * n);
  if (eigen) std::memcpy(eigen, s.eigen.data(), sizeof(double) * n);
  if (vect) std::memcpy(vect, s.vect.data(), sizeof(double) * n * n);
  if (inv_vect) std::memcpy(inv_vect, s.inv_vect.data(), sizeof(double) * n * n);
  if (inv_lhs) std::memcpy(inv_lhs, s.inv_lhs.data(), sizeof(double) * n * n);
  if (neum) std::memcpy(neum, s.neum.data(), sizeof(double) * 14);
  return MF_OK;
}

extern "C" int mf_solver3d_dims(const mf_solver3d *sc, int *n1, int *n2, int *n3) {
  if (!sc) return mf::fail(MF_ERR_ARG, "mf_solver3d_dims: null handle");
  if (n1) *n1 = sc->n1;
  if (n2) *n2 = sc->n2;
  if (n3) *n3 = sc->n3;
  return MF_OK;
}

extern "C" int mf_solver3d_profile(mf_solver3d *sc, int enable) {
  if (!sc) return mf::fail(MF_ERR_ARG, "mf_solver3d_profile: null handle");
  sc->prof = enable != 0;
  return MF_OK;
}

extern "C" int mf_solver3d_profile_read(mf_solver3d *sc, double ms[MF_PROF_CLASSES],
                                        int count[MF_PROF_CLASSES],
                                        double work[MF_PROF_CLASSES], int reset) {
  if (!sc) return mf::fail(MF_ERR_ARG, "mf_solver3d_profile_read: null handle");
  for (size_t m = 0; m < sc->marks.size(); ++m) {
    HIPCHK(hipEventSynchronize(sc->ev_pool[2 * m + 1]));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, sc->ev_pool[2 * m], sc->ev_pool[2 * m + 1]));
    const int c = sc->marks[m].cls;
    sc->prof_ms[c] += t;
    sc->prof_work[c] += sc->marks[m].work;
    sc->prof_cnt[c] += 1;
  }
  sc->marks.clear();
  for (int c = 0; c < MF_PROF_CLASSES; ++c) {
    if (ms) ms[c] = sc->prof_ms[c];
    if (count) count[c] = sc->prof_cnt[c];
    if (work) work[c] = sc->prof_work[c];
    if (reset) { sc->prof_ms[c] = 0; sc->prof_work[c] = 0; sc->prof_cnt[c] = 0; }
  }
  return MF_OK;
}

extern "C" int mf_solver3d_destroy(mf_solver3d *sc) {
  if (!sc) return MF_OK;
  for (hipEvent_t e : sc->ev_pool) (void)hipEventDestroy(e);
  if (sc->iws) (void)hipFree(sc->iws);
  if (sc->dmem) (void)hipFree(sc->dmem);
  if (sc->hnorms) (void)hipHostFree(sc->hnorms);
  delete sc;
  return MF_OK;
}

extern "C" int mf_solver3d_create(const double *gx, int nx, const double *gy, int ny,
                                  const double *gz, int nz, const int bct[6], int so,
                                  mf_solver3d **out) {
  if (!gx || !gy || !gz || !bct || !out) return mf::fail(MF_ERR_ARG, "mf_solver3d_create: null pointer");
  if (nx < 9 || ny < 9 || nz < 9)
    return mf::fail(MF_ERR_ARG, "mf_solver3d_create: at least 9 points per direction are needed");
  mf_solver3d *h = new mf_solver3d();
  h->nx = nx; h->ny = ny; h->nz = nz;
  h->n1 = nx - 2; h->n2 = ny - 2; h->n3 = nz - 2;
  h->so = so;
  std::memcpy(h->bct, bct, sizeof(int) * 6);
  const double *grids[3] = {gx, gy, gz};
  const int nall[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a) {
    std::string err;
    if (!mf::solver_setup_1d(grids[a], nall[a], bct + 2 * a, so, h->ax[a], err)) {
      delete h;
      return mf::fail(err.find("at least") != std::string::npos || err.find("must") != std::string::npos
                          ? MF_ERR_ARG
                          : MF_ERR_SETUP,
                      std::string("mf_solver3d_create: ") + "xyz"[a] + "-direction: " + err);
    }
  }
  const int n[3] = {h->n1, h->n2, h->n3};
  // host images of the device tables
  std::vector<double> W[3], V[3], Wp[3];
  for (int a = 0; a < 3; ++a) {
    const mf::Solver1D &s = h->ax[a];
    const int m = n[a];
    // W' = inv(V) inv(lhs) for the mixed path
    std::vector<double> wp((size_t)m * m, 0.0);
    for (int j = 0; j < m; ++j)
      for (int k = 0; k < m; ++k) {
        const double b = s.inv_lhs[k + (size_t)m * j];
        if (b == 0.0) continue;
        for (int i = 0; i < m; ++i) wp[i + (size_t)m * j] += s.inv_vect[i + (size_t)m * k] * b;
      }
    if (a == 0) {
      W[a] = s.inv_vect;
      V[a] = s.vect;
      Wp[a] = wp;
    } else {  // y and z act from the right: store transposes (class_solver_3d.f90:403-409)
      transpose(m, s.inv_vect, W[a]);
      transpose(m, s.vect, V[a]);
      transpose(m, wp, Wp[a]);
    }
  }
  std::vector<double> lam, fold, neum, ebd;
  for (int a = 0; a < 3; ++a) lam.insert(lam.end(), h->ax[a].eigen.begin(), h->ax[a].eigen.end());
  for (int a = 0; a < 3; ++a) {
    const mf::Solver1D &s = h->ax[a];
    const int m = n[a];
    auto IL = [&](int i, int j) { return s.inv_lhs[i + (size_t)m * j]; };
    auto DR = [&](int i, int c) { return s.ddr[i + (size_t)m * c]; };
    std::vector<double> f1(m), f2(m);
    for (int i = 0; i < m; ++i) {
      f1[i] = IL(i, 0) * DR(0, 0) + IL(i, 1) * DR(1, 0);
      f2[i] = IL(i, m - 2) * DR(m - 2, 7) + IL(i, m - 1) * DR(m - 1, 6);
    }
    fold.insert(fold.end(), f1.begin(), f1.end());
    fold.insert(fold.end(), f2.begin(), f2.end());
    // lhs f1 and lhs f2: the boundary columns of rhs (two rows each)
    std::vector<double> e1(m, 0.0), e2(m, 0.0);
    e1[0] = DR(0, 0);
    e1[1] = DR(1, 0);
    e2[m - 2] = DR(m - 2, 7);
    e2[m - 1] = DR(m - 1, 6);
    ebd.insert(ebd.end(), e1.begin(), e1.end());
    ebd.insert(ebd.end(), e2.begin(), e2.end());
    neum.insert(neum.end(), s.neum.begin(), s.neum.end());
  }
  std::vector<int> bandc;
  std::vector<double> bandw;
  for (int a = 0; a < 3; ++a) {
    std::vector<int> c;
    std::vector<double> w;
    if (!to_band(n[a], h->ax[a].lhs, c, w) ) {
      delete h;
      return mf::fail(MF_ERR_SETUP, "mf_solver3d_create: lhs rows do not fit 5-wide windows");
    }
    bandc.insert(bandc.end(), c.begin(), c.end());
    bandw.insert(bandw.end(), w.begin(), w.end());
    if (!to_band(n[a], h->ax[a].rhs, c, w)) {
      delete h;
      return mf::fail(MF_ERR_SETUP, "mf_solver3d_create: rhs rows do not fit 5-wide windows");
    }
    bandc.insert(bandc.end(), c.begin(), c.end());
    bandw.insert(bandw.end(), w.begin(), w.end());
  }
  // z rows on the planes b(k) .. b(k)+4, b(k) = clamp(k-2, 0, n3-5), for the
  // z-streaming residual (used when every nonzero of Mz and Rz falls there)
  std::vector<double> zq;
  bool zs_ok = n[0] >= 4 && n[1] >= 4 && n[2] >= 5;
  if (zs_ok) {
    const int m3 = n[2];
    constexpr int PT = mf::ZS_PT;
    zq.assign((size_t)PT * m3, 0.0);
    for (int k = 0; k < m3; ++k) {  // e_z1, e_z2
      zq[(size_t)PT * k + 2 * mf::BAND_W] = ebd[2 * (size_t)(n[0] + n[1]) + k];
      zq[(size_t)PT * k + 2 * mf::BAND_W + 1] = ebd[2 * (size_t)(n[0] + n[1]) + m3 + k];
    }
    const size_t cb = 2 * (size_t)(n[0] + n[1]), wb = 8 * cb;  // axis 2 in bandc / bandw
    for (int q = 0; q < 2 && zs_ok; ++q)
      for (int k = 0; k < m3 && zs_ok; ++k) {
        const int b = std::min(std::max(k - 2, 0), m3 - 5);
        const int c0 = bandc[cb + (size_t)q * m3 + k];
        for (int m = 0; m < mf::BAND_W; ++m) {
          const double v = bandw[wb + 8 * ((size_t)q * m3 + k) + m];
          const int d = c0 + m - b;
          if (d >= 0 && d < mf::BAND_W) zq[(size_t)PT * k + mf::BAND_W * q + d] = v;
          else if (v != 0.0) zs_ok = false;
        }
      }
  }
  if (const char *env = std::getenv("MF_H2_RESIDENT")) h->h2_resident = std::atoi(env) != 0;
  if (const char *env = std::getenv("MF_H2_PLANE")) h->h2_plane = std::atoi(env) != 0;
  if (const char *env = std::getenv("MF_H2_COLS")) h->h2_cols = std::atoi(env) != 0;
  {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
      h->ncu = cus;
  }
  size_t nparts_zs = 0;
  if (zs_ok) {
    // about 5 blocks per CU over the (i,j) tiles, 32 planes per block at 256^3
    // (fewer planes per block: more warm-up reads; more: fewer blocks in
    // flight; measured at 256^3 with three resident blocks per CU: 16 and 32
    // equal, 48 is 7% slower); MF_RESIDUAL_CZ sets the planes per block (0:
    // plane pipeline)
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
      cus = 256;
    const int tiles = ((n[0] + mf::ZS_TX - 1) / mf::ZS_TX) * ((n[1] + mf::ZS_TY - 1) / mf::ZS_TY);
    const int nch = std::max(1, (int)std::lround(5.0 * cus / tiles));
    int cz = (n[2] + nch - 1) / nch;
    if (const char *env = std::getenv("MF_RESIDUAL_CZ")) cz = std::atoi(env);

    h->res_cz = cz <= 0 ? 0 : std::min(std::max(cz, mf::ZS_CZMIN), mf::ZS_CZMAX);
    if (h->res_cz > 0) nparts_zs = (size_t)tiles * ((n[2] + h->res_cz - 1) / h->res_cz);
  }
  // device layout: one allocation carved into 256-byte aligned arrays; the
  // carve runs twice, first to size the allocation, then to assign pointers
  const size_t N = (size_t)n[0] * n[1] * n[2];
  const size_t nparts = std::max(nparts_zs, (size_t)((n[0] + mf::RZ_TX - 1) / mf::RZ_TX) *
                                                ((n[1] + mf::RZ_TY - 1) / mf::RZ_TY) *
                                                ((n[2] + mf::RZ_CZ - 1) / mf::RZ_CZ));
  double *dW[3], *dV[3];
  float *fW[3], *fV[3];
  auto carve = [&](double *base) -> size_t {
    size_t off = 0;
    auto take = [&](size_t cnt) {
      double *r = base ? base + off : nullptr;
      off += (cnt + 31) / 32 * 32;
      return r;
    };
    for (int a = 0; a < 3; ++a) { dW[a] = take((size_t)n[a] * n[a]); dV[a] = take((size_t)n[a] * n[a]); }
    for (int a = 0; a < 3; ++a) {
      fW[a] = (float *)take(((size_t)n[a] * n[a] + 1) / 2);
      fV[a] = (float *)take(((size_t)n[a] * n[a] + 1) / 2);
    }
    h->lam = take(lam.size());
    h->fold = take(fold.size());
    h->neum = take(neum.size());
    h->bandw = take(bandw.size());
    h->norms = take(MF_MAX_REFINE + 1);
    h->part = take(mf::NORM_SLOTS * nparts);
    h->part_stride = nparts;
    h->T1 = take(N); h->T2 = take(N);
    h->F1 = (float *)take((N + 1) / 2);
    h->F2 = (float *)take((N + 1) / 2);
    h->R32 = (float *)take((N + 1) / 2);
    h->ebd = take(ebd.size());
    h->zq = take(std::max<size_t>(zq.size(), 1));
    h->fbf = take(2 * ((size_t)n[1] * n[2] + (size_t)n[0] * n[2] + (size_t)n[0] * n[1]));
    h->bandc = (int *)take((bandc.size() + 1) / 2);
    h->bctd = (int *)take(4);
    for (int a = 0; a < 3; ++a) {
      h->kp[a] = (n[a] + 63) / 64 * 64;  // an even number of 32-wide k-steps
      const int t = gemm_tile(n[a]);
      h->wrows[a] = (n[a] + t - 1) / t * t;
      const size_t cnt = (3 * (size_t)h->wrows[a] * h->kp[a] + 3) / 4;  // bf16 -> doubles
      h->Wsp[a] = (__bf16 *)take(cnt);
      h->Vsp[a] = (__bf16 *)take(cnt);
      const size_t cnh = (2 * (size_t)h->wrows[a] * h->kp[a] + 3) / 4;  // fp16 -> doubles
      h->Whs[a] = (_Float16 *)take(cnh);
      h->Vhs[a] = (_Float16 *)take(cnh);
    }
    h->Whp1 = (_Float16 *)take((2 * (size_t)h->wrows[1] * h->kp[1] + 3) / 4);
    h->Vhp0 = (_Float16 *)take((2 * (size_t)h->wrows[0] * h->kp[0] + 3) / 4);
    h->Vhp2 = (_Float16 *)take((2 * (size_t)h->wrows[2] * h->kp[2] + 3) / 4);
    h->fmx = (unsigned *)take(4);
    h->w0 = take((size_t)n[0] + n[1] + n[2]);
    h->npart = take((size_t)n[2] * mf::NULL_SLICES);
    return off;
  };
  const size_t nd = carve(nullptr);
  hipError_t e = hipMalloc(&h->dmem, nd * sizeof(double));
  if (e != hipSuccess) {
    delete h;
    return mf::fail(MF_ERR_DEVICE, std::string("mf_solver3d_create: hipMalloc: ") + hipGetErrorString(e));
  }
  carve((double *)h->dmem);
  h->Wx = dW[0]; h->Vx = dV[0]; h->Wyt = dW[1]; h->Vyt = dV[1]; h->Wzt = dW[2]; h->Vzt = dV[2];
  h->Wx32 = fW[0]; h->Vx32 = fV[0]; h->Wyt32 = fW[1]; h->Vyt32 = fV[1]; h->Wzt32 = fW[2]; h->Vzt32 = fV[2];
  auto up = [&](void *dst, const void *src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  };
  std::vector<float> tmp;
  auto upf = [&](float *dst, const std::vector<double> &src) {
    tmp.assign(src.begin(), src.end());
    return up(dst, tmp.data(), tmp.size() * sizeof(float));
  };
  e = hipSuccess;
  for (int a = 0; a < 3 && e == hipSuccess; ++a) {
    e = up(dW[a], W[a].data(), W[a].size() * sizeof(double));
    if (e == hipSuccess) e = up(dV[a], V[a].data(), V[a].size() * sizeof(double));
    if (e == hipSuccess) e = upf(fW[a], Wp[a]);
    if (e == hipSuccess) e = upf(fV[a], V[a]);
  }
  // three-part bf16 splits: [part][k/8][row][k%8]; the x matrices are the left
  // operand (row = m of W(m,k)), the stored y/z transposes the right one
  // (row = n of B(k,n))
  for (int a = 0; a < 3 && e == hipSuccess; ++a) {
    const int m = n[a], kp = h->kp[a], wr = h->wrows[a];
    for (int which = 0; which < 2 && e == hipSuccess; ++which) {
      const std::vector<double> &src = which == 0 ? Wp[a] : V[a];
      std::vector<uint16_t> sp((size_t)3 * wr * kp, 0);
      for (int r = 0; r < m; ++r)
        for (int k = 0; k < m; ++k) {
          const double v = (a == 0) ? src[r + (size_t)m * k] : src[k + (size_t)m * r];
          uint16_t p[3];
          bf16_split3((float)v, p);
          for (int q = 0; q < 3; ++q)
            sp[(((size_t)q * (kp / 8) + k / 8) * wr + r) * 8 + k % 8] = p[q];
        }
      e = up(which == 0 ? (void *)h->Wsp[a] : (void *)h->Vsp[a], sp.data(), sp.size() * 2);
      // two fp16 parts of the matrix scaled so that its largest entry lands
      // in [2^13, 2^14)
      double big = 0.0;
      for (size_t q = 0; q < (size_t)m * m; ++q) big = std::max(big, std::fabs(src[q]));
      int ex = 0;
      (void)std::frexp(big, &ex);  // big in [2^(ex-1), 2^ex)
      const float sc = big > 0.0 ? (float)std::ldexp(1.0, 14 - ex) : 1.0f;
      h->wsc[a][which] = sc;
      std::vector<uint16_t> hp((size_t)2 * wr * kp, 0);
      for (int r = 0; r < m; ++r)
        for (int k = 0; k < m; ++k) {
          const double v = (a == 0) ? src[r + (size_t)m * k] : src[k + (size_t)m * r];
          const float xs = (float)v * sc;
          const _Float16 h0 = (_Float16)xs, h1 = (_Float16)(xs - (float)h0);
          uint16_t b0, b1;
          std::memcpy(&b0, &h0, 2);
          std::memcpy(&b1, &h1, 2);
          hp[(((size_t)0 * (kp / 8) + k / 8) * wr + r) * 8 + k % 8] = b0;
          hp[(((size_t)1 * (kp / 8) + k / 8) * wr + r) * 8 + k % 8] = b1;
        }
      if (e == hipSuccess)
        e = up(which == 0 ? (void *)h->Whs[a] : (void *)h->Vhs[a], hp.data(), hp.size() * 2);
      if (e == hipSuccess && ((a == 1 && which == 0) || (a != 1 && which == 1))) {
        // k permuted within each 32-wide k-step: position kk holds k = h2p_perm(kk)
        std::vector<uint16_t> pp(hp.size(), 0);
        for (int part = 0; part < 2; ++part)
          for (int kk = 0; kk < kp; ++kk) {
            const int k = 32 * (kk / 32) + mf::h2p_perm(kk % 32);
            for (int r = 0; r < wr; ++r)
              pp[(((size_t)part * (kp / 8) + kk / 8) * wr + r) * 8 + kk % 8] =
                  hp[(((size_t)part * (kp / 8) + k / 8) * wr + r) * 8 + k % 8];
          }
        e = up(a == 1 ? (void *)h->Whp1 : a == 0 ? (void *)h->Vhp0 : (void *)h->Vhp2, pp.data(),
               pp.size() * 2);
      }
    }
  }
  // null rows of a singular problem (the zero eigenvalues are exact zeros)
  {
    std::vector<double> w0;
    double wsum = 1.0;
    bool all = true;
    for (int a = 0; a < 3 && all; ++a) {
      const mf::Solver1D &s = h->ax[a];
      const int m = n[a];
      int idx = -1, cnt = 0;
      for (int i = 0; i < m; ++i)
        if (s.eigen[i] == 0.0) { idx = i; ++cnt; }
      if (cnt != 1) { all = false; break; }
      double sum = 0.0;
      for (int j = 0; j < m; ++j) {
        const double v = s.inv_vect[idx + (size_t)m * j];
        w0.push_back(v);
        sum += v;
      }
      wsum *= sum;
    }
    if (all && wsum != 0.0 && std::isfinite(wsum)) {
      h->null_mode = true;
      h->inv_wsum = 1.0 / wsum;
      if (e == hipSuccess) e = up(h->w0, w0.data(), w0.size() * sizeof(double));
    }
  }
  if (e == hipSuccess) e = up(h->lam, lam.data(), lam.size() * sizeof(double));
  if (e == hipSuccess) e = up(h->fold, fold.data(), fold.size() * sizeof(double));
  if (e == hipSuccess) e = up(h->ebd, ebd.data(), ebd.size() * sizeof(double));
  if (e == hipSuccess) e = up(h->neum, neum.data(), neum.size() * sizeof(double));
  if (e == hipSuccess) e = up(h->bandw, bandw.data(), bandw.size() * sizeof(double));
  if (e == hipSuccess) e = up(h->bandc, bandc.data(), bandc.size() * sizeof(int));
  if (e == hipSuccess && !zq.empty()) e = up(h->zq, zq.data(), zq.size() * sizeof(double));
  if (e == hipSuccess) e = up(h->bctd, h->bct, 6 * sizeof(int));
  if (e == hipSuccess)
    e = hipHostMalloc((void **)&h->hnorms, (MF_MAX_REFINE + 1) * sizeof(double), hipHostMallocDefault);
  if (e != hipSuccess) {
    mf_solver3d_destroy(h);
    return mf::fail(MF_ERR_DEVICE, std::string("mf_solver3d_create: upload: ") + hipGetErrorString(e));
  }
  size_t coff = 0, woff = 0;
  for (int a = 0; a < 3; ++a) {
    h->M[a].c0 = h->bandc + coff; h->M[a].w = h->bandw + woff;
    coff += n[a]; woff += (size_t)8 * n[a];
    h->R[a].c0 = h->bandc + coff; h->R[a].w = h->bandw + woff;
    coff += n[a]; woff += (size_t)8 * n[a];
  }
  *out = h;
  return MF_OK;
}

namespace {
using mf::EpiArgs;

template <typename T, typename TA, typename TB, typename TC, int BM, int BN, int EPI>
void launch_tile(mf_solver3d *h, const TA *A, long lda, long sA, const TB *B, long ldb, TC *C,
                 long ldc, long sC, int M, int N, int K, int batch, const EpiArgs &ep,
                 hipStream_t st) {
  dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN, batch);
  ProfScope ps(h, sizeof(T) == 8 ? P_T64 : P_T32, 2.0 * M * N * K * batch, st);
  hipLaunchKernelGGL((mf::gemm_mode_kernel<T, TA, TB, TC, BM, BN, EPI>), grid, dim3(256), 0, st, A,
                     lda, sA, B, ldb, C, ldc, sC, M, N, K, ep);
}

// Block tile by the eigenvector dimension n (M for the x-transform, whose
// eigenvector matrix is the left operand; N for y/z): the tile side along n
// is the smallest of 64/128/192/256 covering it, so every field element is
// read once and few MFMA rows are padding.  WLEFT selects the orientation.
inline int gemm_mode_tile(int n) { return n > 128 && n <= 192 ? 192 : gemm_tile(n); }

template <typename T, typename TA, typename TB, typename TC, bool WLEFT, int EPI>
void launch_gemm(mf_solver3d *h, const TA *A, long lda, long sA, const TB *B, long ldb, TC *C,
                 long ldc, long sC, int M, int N, int K, int batch, const EpiArgs &ep,
                 hipStream_t st) {
  const int t = gemm_mode_tile(WLEFT ? M : N);
  // the batched y-transform with a 256-wide N tile and 129..192 rows: one
  // 192-row M tile and four 64-wide N tiles (the field slice re-read from L2)
  // give 4 blocks per slice where 64-row M tiles give 3, which at 190 slices
  // leaves the last round of blocks a quarter full
  if (!WLEFT && t == 256 && M > 128 && M <= 192 && batch > 1) {
    launch_tile<T, TA, TB, TC, 192, 64, EPI>(h, A, lda, sA, B, ldb, C, ldc, sC, M, N, K, batch, ep, st);
    return;
  }
  if (t == 192) {
    if (WLEFT)
      launch_tile<T, TA, TB, TC, 192, 64, EPI>(h, A, lda, sA, B, ldb, C, ldc, sC, M, N, K, batch, ep, st);
    else
      launch_tile<T, TA, TB, TC, 64, 192, EPI>(h, A, lda, sA, B, ldb, C, ldc, sC, M, N, K, batch, ep, st);
    return;
  }
  if (WLEFT) {
    if (t == 64)
      launch_tile<T, TA, TB, TC, 64, 256, EPI>(h, A, lda, sA, B, ldb, C, ldc, sC, M, N, K, batch, ep, st);
    else if (t == 128)
      launch_tile<T, TA, TB, TC, 128, 128, EPI>(h, A, lda, sA, B, ldb, C, ldc, sC, M, N, K, batch, ep, st);
    else
      launch_tile<T, TA, TB, TC, 256, 64, EPI>(h, A, lda, sA, B, ldb, C, ldc, sC, M, N, K, batch, ep, st);
  } else {
    if (t == 64)
      launch_tile<T, TA, TB, TC, 256, 64, EPI>(h, A, lda, sA, B, ldb, C, ldc, sC, M, N, K, batch, ep, st);
    else if (t == 128)
      launch_tile<T, TA, TB, TC, 128, 128, EPI>(h, A, lda, sA, B, ldb, C, ldc, sC, M, N, K, batch, ep, st);
    else
      launch_tile<T, TA, TB, TC, 64, 256, EPI>(h, A, lda, sA, B, ldb, C, ldc, sC, M, N, K, batch, ep, st);
  }
}

// Six transforms: in (interior, type TIn) -> u interior (scatter or add).
template <typename T, typename TIn>
void transforms(mf_solver3d *h, const TIn *in, T *w1, T *w2, const T *Wx, const T *Wyt,
                const T *Wzt, const T *Vx, const T *Vyt, const T *Vzt, double sigma, double *u,
                bool accumulate, hipStream_t st, bool last = true) {
  const int n1 = h->n1, n2 = h->n2, n3 = h->n3;
  const long n12 = (long)n1 * n2;
  EpiArgs ep;
  std::memset(&ep, 0, sizeof(ep));
  ep.lx = h->lam;
  ep.ly = h->lam + n1;
  ep.lz = h->lam + n1 + n2;
  ep.sigma = sigma;
  ep.n1 = n1;
  ep.n2 = n2;
  ep.u = u;
  ep.nx = h->nx;
  ep.ny = h->ny;
  // forward X: w1 = Wx . in
  launch_gemm<T, T, TIn, T, true, mf::EPI_STORE>(h, Wx, n1, 0, in, n1, w1, n1, 0, n1, n2 * n3, n1, 1, ep, st);
  // forward Y: w2_k = w1_k . Wyt
  launch_gemm<T, T, T, T, false, mf::EPI_STORE>(h, w1, n1, n12, Wyt, n2, w2, n1, n12, n1, n2, n2, n3, ep, st);
  // forward Z with the eigenvalue divide: w1 = (w2 . Wzt) / lambda
  launch_gemm<T, T, T, T, false, mf::EPI_DIVIDE>(h, w2, n12, 0, Wzt, n3, w1, n12, 0, (int)n12, n3, n3, 1, ep, st);
  // backward Z: w2 = w1 . Vzt
  launch_gemm<T, T, T, T, false, mf::EPI_STORE>(h, w1, n12, 0, Vzt, n3, w2, n12, 0, (int)n12, n3, n3, 1, ep, st);
  // backward Y: w1_k = w2_k . Vyt
  launch_gemm<T, T, T, T, false, mf::EPI_STORE>(h, w2, n1, n12, Vyt, n2, w1, n1, n12, n1, n2, n2, n3, ep, st);
  // backward X into u's interior
  if (!last) return;
  if (accumulate)
    launch_gemm<T, T, T, double, true, mf::EPI_SCATTER_ADD>(h, Vx, n1, 0, w1, n1, (double *)nullptr, 0, 0, n1, n2 * n3, n1, 1, ep, st);
  else
    launch_gemm<T, T, T, double, true, mf::EPI_SCATTER>(h, Vx, n1, 0, w1, n1, (double *)nullptr, 0, 0, n1, n2 * n3, n1, 1, ep, st);
}

// ---- split-bf16 transforms (MF_MIXED_BF16X3)
template <typename TF, typename TC, int BM, int BN, bool WLEFT, int EPI>
void launch_bf3_tile(mf_solver3d *h, const TF *F, long ldf, long sF, const __bf16 *W, int KP,
                     int WROWS, TC *C, long ldc, long sC, int M, int N, int K, int batch,
                     const EpiArgs &ep, hipStream_t st) {
  dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN, batch);
  ProfScope ps(h, P_T32, 2.0 * M * N * K * batch, st);
  hipLaunchKernelGGL((mf::gemm_bf3_kernel<TF, TC, BM, BN, WLEFT, EPI>), grid, dim3(256), 0, st, F,
                     ldf, sF, W, KP, WROWS, C, ldc, sC, M, N, K, ep);
}

// Tile choice as launch_gemm: the side along the eigenvector dimension is the
// smallest of 64/128/256 that covers it.
template <typename TF, typename TC, bool WLEFT, int EPI>
void launch_bf3(mf_solver3d *h, const TF *F, long ldf, long sF, const __bf16 *W, int KP, int WROWS,
                TC *C, long ldc, long sC, int M, int N, int K, int batch, const EpiArgs &ep,
                hipStream_t st) {
  const int t = gemm_tile(WLEFT ? M : N);
  if (WLEFT) {
    if (t == 64)
      launch_bf3_tile<TF, TC, 64, 256, WLEFT, EPI>(h, F, ldf, sF, W, KP, WROWS, C, ldc, sC, M, N, K, batch, ep, st);
    else if (t == 128)
      launch_bf3_tile<TF, TC, 128, 128, WLEFT, EPI>(h, F, ldf, sF, W, KP, WROWS, C, ldc, sC, M, N, K, batch, ep, st);
    else
      launch_bf3_tile<TF, TC, 256, 64, WLEFT, EPI>(h, F, ldf, sF, W, KP, WROWS, C, ldc, sC, M, N, K, batch, ep, st);
  } else {
    if (t == 64)
      launch_bf3_tile<TF, TC, 256, 64, WLEFT, EPI>(h, F, ldf, sF, W, KP, WROWS, C, ldc, sC, M, N, K, batch, ep, st);
    else if (t == 128)
      launch_bf3_tile<TF, TC, 128, 128, WLEFT, EPI>(h, F, ldf, sF, W, KP, WROWS, C, ldc, sC, M, N, K, batch, ep, st);
    else
      launch_bf3_tile<TF, TC, 64, 256, WLEFT, EPI>(h, F, ldf, sF, W, KP, WROWS, C, ldc, sC, M, N, K, batch, ep, st);
  }
}

// The six transforms of transforms<float, double> with the pre-split
// matrices W' = inv(V) inv(lhs) and V.
void transforms_bf3(mf_solver3d *h, const float *in, float *w1, float *w2, double sigma,
                    double *u, bool accumulate, hipStream_t st, bool last = true) {
  const int n1 = h->n1, n2 = h->n2, n3 = h->n3;
  const long n12 = (long)n1 * n2;
  const int kx = h->kp[0], ky = h->kp[1], kz = h->kp[2];
  const int rx = h->wrows[0], ry = h->wrows[1], rz = h->wrows[2];
  EpiArgs ep;
  std::memset(&ep, 0, sizeof(ep));
  ep.lx = h->lam;
  ep.ly = h->lam + n1;
  ep.lz = h->lam + n1 + n2;
  ep.sigma = sigma;
  ep.n1 = n1;
  ep.n2 = n2;
  ep.u = u;
  ep.nx = h->nx;
  ep.ny = h->ny;
  using mf::EPI_STORE;
  launch_bf3<float, float, true, EPI_STORE>(h, in, n1, 0, h->Wsp[0], kx, rx, w1, n1, 0, n1, n2 * n3, n1, 1, ep, st);
  launch_bf3<float, float, false, EPI_STORE>(h, w1, n1, n12, h->Wsp[1], ky, ry, w2, n1, n12, n1, n2, n2, n3, ep, st);
  launch_bf3<float, float, false, mf::EPI_DIVIDE>(h, w2, n12, 0, h->Wsp[2], kz, rz, w1, n12, 0, (int)n12, n3, n3, 1, ep, st);
  launch_bf3<float, float, false, EPI_STORE>(h, w1, n12, 0, h->Vsp[2], kz, rz, w2, n12, 0, (int)n12, n3, n3, 1, ep, st);
  launch_bf3<float, float, false, EPI_STORE>(h, w2, n1, n12, h->Vsp[1], ky, ry, w1, n1, n12, n1, n2, n2, n3, ep, st);
  if (!last) return;
  if (accumulate)
    launch_bf3<float, double, true, mf::EPI_SCATTER_ADD>(h, w1, n1, 0, h->Vsp[0], kx, rx, (double *)nullptr, 0, 0, n1, n2 * n3, n1, 1, ep, st);
  else
    launch_bf3<float, double, true, mf::EPI_SCATTER>(h, w1, n1, 0, h->Vsp[0], kx, rx, (double *)nullptr, 0, 0, n1, n2 * n3, n1, 1, ep, st);
}

// ---- two-part fp16 transforms (MF_MIXED_FP16X2)
template <typename TF, typename TC, int BM, int BN, bool WLEFT, int EPI>
void launch_h2_tile(mf_solver3d *h, const TF *F, long ldf, long sF, const _Float16 *W, float wsc,
                    int KP, int WROWS, TC *C, long ldc, long sC, int M, int N, int K, int batch,
                    const EpiArgs &ep, int slot_in, int slot_out, bool zero0, hipStream_t st) {
  dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN, batch);
  ProfScope ps(h, P_T32, 2.0 * M * N * K * batch, st);
  hipLaunchKernelGGL((mf::gemm_h2_kernel<TF, TC, BM, BN, WLEFT, EPI>), grid, dim3(256), 0, st, F,
                     ldf, sF, W, KP, WROWS, C, ldc, sC, M, N, K, ep,
                     (const unsigned *)(h->fmx + slot_in), wsc,
                     slot_out >= 0 ? h->fmx + slot_out : (unsigned *)nullptr,
                     zero0 ? h->fmx : (unsigned *)nullptr);
}

template <typename TF, typename TC, bool WLEFT, int EPI>
void launch_h2(mf_solver3d *h, const TF *F, long ldf, long sF, const _Float16 *W, float wsc,
               int KP, int WROWS, TC *C, long ldc, long sC, int M, int N, int K, int batch,
               const EpiArgs &ep, int slot_in, int slot_out, bool zero0, hipStream_t st) {
  const int t = gemm_tile(WLEFT ? M : N);
  if (WLEFT && t == 256 && WROWS == 256 && KP <= mf::H2R_KP && h->h2_resident) {
    // eigenvector matrix resident in registers, persistent blocks (gemm_h2r.h);
    // x transforms only: 43.6 us against 45.6 us for gemm_h2_kernel at 254^3,
    // while the y orientation is slower there (43.5 against 37.0 us)
    const long fr = WLEFT ? N : M;
    const long tiles = (fr + mf::H2R_RF - 1) / mf::H2R_RF * batch;
    const int grid = (int)std::min<long>(tiles, h->ncu);
    ProfScope ps(h, P_T32, 2.0 * M * N * K * batch, st);
    hipLaunchKernelGGL((mf::gemm_h2r_kernel<TC, WLEFT, EPI>), dim3(grid), dim3(512), 0, st,
                       (const float *)F, ldf, sF, W, KP, C, ldc, sC, M, N, K, batch, ep,
                       (const unsigned *)(h->fmx + slot_in), wsc,
                       slot_out >= 0 ? h->fmx + slot_out : (unsigned *)nullptr,
                       zero0 ? h->fmx : (unsigned *)nullptr);
    return;
  }
#define MF_H2(BM, BN)                                                                      \
  launch_h2_tile<TF, TC, BM, BN, WLEFT, EPI>(h, F, ldf, sF, W, wsc, KP, WROWS, C, ldc, sC, M, N, \
                                             K, batch, ep, slot_in, slot_out, zero0, st)
  if (WLEFT) {
    if (t == 64) MF_H2(64, 256);
    else if (t == 128) MF_H2(128, 128);
    else MF_H2(256, 64);
  } else {
    if (t == 64) MF_H2(256, 64);
    else if (t == 128) MF_H2(128, 128);
    else MF_H2(64, 256);
  }
#undef MF_H2
}

// The z-forward transform, the divide and the z-backward transform run as
// one kernel (gemm_h2z_kernel) when one block holds every z of both products.
bool h2_zfused(const mf_solver3d *h) { return h->kp[2] <= 256 && h->wrows[2] == mf::H2Z_BN; }

// The x and y transforms of each direction run as one kernel per z-plane
// (gemm_h2p.h) when both eigenvector dimensions take 256-row splits (129..256
// points, k padded to 192 or 256) and there are planes for at least half
// the CUs (one block per plane: at 38 planes the pair was 18% slower than
// the separate transforms, at 98 equal, at 148 4% faster).
bool h2_planes(const mf_solver3d *h) {
  return h->h2_plane && h->wrows[0] == mf::H2P_N && h->wrows[1] == mf::H2P_N &&
         (h->kp[0] == 192 || h->kp[0] == 256) && (h->kp[1] == 192 || h->kp[1] == 256) &&
         2 * h->n3 >= h->ncu;
}

// forward x+y pair: in (slot 0) -> out (slot 2)
template <int KPX, int KPY>
void launch_h2p_fwd_k(mf_solver3d *h, const float *in, float *out, hipStream_t st) {
  const int n1 = h->n1, n2 = h->n2, n3 = h->n3;
  ProfScope ps(h, P_T32, 2.0 * n1 * n2 * n3 * (n1 + n2), st);
  hipLaunchKernelGGL((mf::h2p_fwd_kernel<KPX, KPY>), dim3((unsigned)n3),
                     dim3(512), 0, st, in, n1, n2, n3, (const _Float16 *)h->Whs[0],
                     (const _Float16 *)h->Whp1, h->wsc[0][0], h->wsc[1][0], out,
                     (const unsigned *)(h->fmx + 0), h->fmx + 2);
}
void launch_h2p_fwd(mf_solver3d *h, const float *in, float *out, hipStream_t st) {
  if (h->kp[0] == 256) {
    if (h->kp[1] == 256) launch_h2p_fwd_k<256, 256>(h, in, out, st);
    else launch_h2p_fwd_k<256, 192>(h, in, out, st);
  } else {
    if (h->kp[1] == 256) launch_h2p_fwd_k<192, 256>(h, in, out, st);
    else launch_h2p_fwd_k<192, 192>(h, in, out, st);
  }
}

// backward y+x pair: in (the largest magnitude in slot 4) -> u (written or
// added); clears slot 0
template <int KPX, int KPY, bool ADD>
void launch_h2p_bwd_k(mf_solver3d *h, const float *in, double *u, hipStream_t st) {
  const int n1 = h->n1, n2 = h->n2, n3 = h->n3;
  ProfScope ps(h, P_T32, 2.0 * n1 * n2 * n3 * (n1 + n2), st);
  hipLaunchKernelGGL((mf::h2p_bwd_kernel<KPX, KPY, ADD>), dim3((unsigned)n3),
                     dim3(512), 0, st, in, n1, n2, n3, (const _Float16 *)h->Vhs[1],
                     (const _Float16 *)h->Vhp0, h->wsc[1][1], h->wsc[0][1], u, h->nx, h->ny,
                     (const unsigned *)(h->fmx + 4), h->fmx);
}
template <bool ADD>
void launch_h2p_bwd_a(mf_solver3d *h, const float *in, double *u, hipStream_t st) {
  if (h->kp[0] == 256) {
    if (h->kp[1] == 256) launch_h2p_bwd_k<256, 256, ADD>(h, in, u, st);
    else launch_h2p_bwd_k<256, 192, ADD>(h, in, u, st);
  } else {
    if (h->kp[1] == 256) launch_h2p_bwd_k<192, 256, ADD>(h, in, u, st);
    else launch_h2p_bwd_k<192, 192, ADD>(h, in, u, st);
  }
}
void launch_h2p_bwd(mf_solver3d *h, const float *in, double *u, bool accumulate, hipStream_t st) {
  if (accumulate) launch_h2p_bwd_a<true>(h, in, u, st);
  else launch_h2p_bwd_a<false>(h, in, u, st);
}

template <int KPZ>
void launch_h2c(mf_solver3d *h, const float *in, float *out, const EpiArgs &ep, hipStream_t st) {
  const long n12 = (long)h->n1 * h->n2;
  hipLaunchKernelGGL((mf::h2c_kernel<KPZ>), dim3((unsigned)((n12 + 255) / 256)), dim3(512), 0, st,
                     in, n12, h->n3, (const _Float16 *)h->Whs[2], (const _Float16 *)h->Vhp2,
                     h->wsc[2][0], h->wsc[2][1], out, ep, (const unsigned *)(h->fmx + 2),
                     h->fmx + 4);
}

// the fused z pair: column blocks (gemm_h2p.h h2c_kernel) when k is padded to
// 192 or 256, else row tiles (gemm_h2z_kernel)
void launch_h2z(mf_solver3d *h, const float *in, float *out, const EpiArgs &ep, hipStream_t st) {
  const long n12 = (long)h->n1 * h->n2;
  ProfScope ps(h, P_T32, 4.0 * n12 * h->n3 * h->n3, st);
  if (h->h2_cols && (h->kp[2] == 192 || h->kp[2] == 256)) {
    if (h->kp[2] == 256) launch_h2c<256>(h, in, out, ep, st);
    else launch_h2c<192>(h, in, out, ep, st);
    return;
  }
  hipLaunchKernelGGL(mf::gemm_h2z_kernel, dim3((unsigned)((n12 + mf::H2Z_BM - 1) / mf::H2Z_BM)),
                     dim3(256), 0, st, in, n12, (const _Float16 *)h->Whs[2],
                     (const _Float16 *)h->Vhs[2], h->kp[2], h->wrows[2], out, n12, (int)n12, h->n3,
                     ep, (const unsigned *)(h->fmx + 2), h->wsc[2][0], h->wsc[2][1], h->fmx + 4);
}

// The six transforms of transforms_bf3 on the two-part fp16 matrices.  Field
// scales: slot 0 holds max |r'| (written by the residual kernel), slot q the
// largest output of transform q (the fused z pair writes slot 4); the last
// transform clears slot 0 for the residual that follows (the residual clears
// slots 1-5).
void transforms_h2(mf_solver3d *h, const float *in, float *w1, float *w2, double sigma,
                   double *u, bool accumulate, hipStream_t st, bool last = true) {
  const int n1 = h->n1, n2 = h->n2, n3 = h->n3;
  const long n12 = (long)n1 * n2;
  const int kx = h->kp[0], ky = h->kp[1], kz = h->kp[2];
  const int rx = h->wrows[0], ry = h->wrows[1], rz = h->wrows[2];
  EpiArgs ep;
  std::memset(&ep, 0, sizeof(ep));
  ep.lx = h->lam;
  ep.ly = h->lam + n1;
  ep.lz = h->lam + n1 + n2;
  ep.sigma = sigma;
  ep.n1 = n1;
  ep.n2 = n2;
  ep.u = u;
  ep.nx = h->nx;
  ep.ny = h->ny;
  using mf::EPI_STORE;
  if (h2_planes(h) && u) {
    // x+y forward pair, the z transforms (one kernel when fused), y+x
    // backward pair into u
    launch_h2p_fwd(h, in, w2, st);
    if (h2_zfused(h)) {
      launch_h2z(h, w2, w1, ep, st);
      launch_h2p_bwd(h, w1, u, accumulate, st);
    } else {
      launch_h2<float, float, false, mf::EPI_DIVIDE>(h, w2, n12, 0, h->Whs[2], h->wsc[2][0], kz, rz, w1, n12, 0, (int)n12, n3, n3, 1, ep, 2, 3, false, st);
      launch_h2<float, float, false, EPI_STORE>(h, w1, n12, 0, h->Vhs[2], h->wsc[2][1], kz, rz, w2, n12, 0, (int)n12, n3, n3, 1, ep, 3, 4, false, st);
      launch_h2p_bwd(h, w2, u, accumulate, st);
    }
    return;
  }
  launch_h2<float, float, true, EPI_STORE>(h, in, n1, 0, h->Whs[0], h->wsc[0][0], kx, rx, w1, n1, 0, n1, n2 * n3, n1, 1, ep, 0, 1, false, st);
  launch_h2<float, float, false, EPI_STORE>(h, w1, n1, n12, h->Whs[1], h->wsc[1][0], ky, ry, w2, n1, n12, n1, n2, n2, n3, ep, 1, 2, false, st);
  if (h2_zfused(h)) {
    launch_h2z(h, w2, w1, ep, st);
    launch_h2<float, float, false, EPI_STORE>(h, w1, n1, n12, h->Vhs[1], h->wsc[1][1], ky, ry, w2, n1, n12, n1, n2, n2, n3, ep, 4, 5, false, st);
    if (!last) return;
    if (accumulate)
      launch_h2<float, double, true, mf::EPI_SCATTER_ADD>(h, w2, n1, 0, h->Vhs[0], h->wsc[0][1], kx, rx, (double *)nullptr, 0, 0, n1, n2 * n3, n1, 1, ep, 5, -1, true, st);
    else
      launch_h2<float, double, true, mf::EPI_SCATTER>(h, w2, n1, 0, h->Vhs[0], h->wsc[0][1], kx, rx, (double *)nullptr, 0, 0, n1, n2 * n3, n1, 1, ep, 5, -1, true, st);
    return;
  }
  launch_h2<float, float, false, mf::EPI_DIVIDE>(h, w2, n12, 0, h->Whs[2], h->wsc[2][0], kz, rz, w1, n12, 0, (int)n12, n3, n3, 1, ep, 2, 3, false, st);
  launch_h2<float, float, false, EPI_STORE>(h, w1, n12, 0, h->Vhs[2], h->wsc[2][1], kz, rz, w2, n12, 0, (int)n12, n3, n3, 1, ep, 3, 4, false, st);
  launch_h2<float, float, false, EPI_STORE>(h, w2, n1, n12, h->Vhs[1], h->wsc[1][1], ky, ry, w1, n1, n12, n1, n2, n2, n3, ep, 4, 5, false, st);
  if (!last) return;
  if (accumulate)
    launch_h2<float, double, true, mf::EPI_SCATTER_ADD>(h, w1, n1, 0, h->Vhs[0], h->wsc[0][1], kx, rx, (double *)nullptr, 0, 0, n1, n2 * n3, n1, 1, ep, 5, -1, true, st);
  else
    launch_h2<float, double, true, mf::EPI_SCATTER>(h, w1, n1, 0, h->Vhs[0], h->wsc[0][1], kx, rx, (double *)nullptr, 0, 0, n1, n2 * n3, n1, 1, ep, 5, -1, true, st);
}

int launch_faces(mf_solver3d *h, double *u, const double *bcx, const double *bcy,
                 const double *bcz, hipStream_t st) {
  const long nf = 2 * ((long)h->n2 * h->n3 + (long)h->n1 * h->n3 + (long)h->n1 * h->n2);
  ProfScope ps(h, P_FACES, 16.0 * nf, st);
  hipLaunchKernelGGL(mf::faces_kernel, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, st, u,
                     h->n1, h->n2, h->n3, h->bctd, h->neum, bcx, bcy, bcz);
  HIPCHK(hipGetLastError());
  return MF_OK;
}

mf::BandRows rows(const mf::Band &b) { return mf::BandRows{b.c0, b.w}; }
}  // namespace

// Back-to-back launches of the first five transforms of a solve (x, y, z
// with the divide, z and y backward) on the handle's own work arrays, reps
// times, between two events: the kernels' own durations, without the
// host-side gaps that per-launch event pairs add.  For the fp16x2 plane
// pairs: the three launches of an inner solve (x+y forward pair, z pair,
// y+x backward pair written into a scratch field).
extern "C" int mf_solver3d_time_transforms(mf_solver3d *h, int precision, int reps,
                                           double *ms_per_launch, double *flops_per_launch,
                                           void *stream) {
  if (!h || reps < 1 || !ms_per_launch)
    return mf::fail(MF_ERR_ARG, "mf_solver3d_time_transforms: invalid argument");
  if (precision < MF_FP64 || precision > MF_MIXED_FP16X2)
    return mf::fail(MF_ERR_ARG, "mf_solver3d_time_transforms: unknown precision");
  hipStream_t st = (hipStream_t)stream;
  const bool planes = precision == MF_MIXED_FP16X2 && h2_planes(h);
  double *uscr = nullptr;  // the backward pair's target
  if (planes) {
    const size_t bytes = (size_t)h->nx * h->ny * h->nz * sizeof(double);
    HIPCHK(hipMalloc(&uscr, bytes));
  }
  const bool prof = h->prof;
  h->prof = false;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  auto five = [&] {
    if (precision == MF_FP64)
      transforms<double, double>(h, h->T2, h->T1, h->T2, h->Wx, h->Wyt, h->Wzt, h->Vx, h->Vyt,
                                 h->Vzt, 0.0, nullptr, false, st, false);
    else if (precision == MF_MIXED)
      transforms<float, float>(h, h->R32, h->F1, h->F2, h->Wx32, h->Wyt32, h->Wzt32, h->Vx32,
                               h->Vyt32, h->Vzt32, 0.0, nullptr, false, st, false);
    else if (precision == MF_MIXED_BF16X3)
      transforms_bf3(h, h->R32, h->F1, h->F2, 0.0, nullptr, false, st, false);
    else
      transforms_h2(h, h->R32, h->F1, h->F2, 0.0, uscr, false, st, false);
  };
  // launches per pass: five, four when the fp16 z pair is one kernel, three
  // (or four, with separate z transforms) for the plane pairs
  const int launches = planes ? (h2_zfused(h) ? 3 : 4)
                              : (precision == MF_MIXED_FP16X2 && h2_zfused(h)) ? 4 : 5;
  five();  // warm
  hipError_t e = hipEventRecord(e0, st);
  for (int r = 0; r < reps && e == hipSuccess; ++r) five();
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (uscr) (void)hipFree(uscr);
  h->prof = prof;
  if (e != hipSuccess)
    return mf::fail(MF_ERR_DEVICE, std::string("mf_solver3d_time_transforms: ") + hipGetErrorString(e));
  *ms_per_launch = ms / ((double)launches * reps);
  if (flops_per_launch)  // 2 M N K per transform, K = n1, n2, n3, n3, n2 (, n1) in turn
    *flops_per_launch = 2.0 * h->n1 * h->n2 * h->n3 *
                        ((planes ? 2.0 : 1.0) * h->n1 + 2.0 * h->n2 + 2.0 * h->n3) / launches;
  return MF_OK;
}

namespace {

mf::ResArgs res_args(mf_solver3d *h, const double *u, const double *g, double sigma) {
  mf::ResArgs a;
  std::memset(&a, 0, sizeof(a));
  const int n1 = h->n1, n2 = h->n2, n3 = h->n3;
  const long ldj = h->nx, ldk = (long)h->nx * h->ny;
  a.n1 = n1; a.n2 = n2; a.n3 = n3;
  a.u = u ? u + 1 + ldj + ldk : nullptr;
  a.g = g + 1 + ldj + ldk;
  a.ldj = ldj; a.ldk = ldk;
  a.sigma = sigma;
  a.Mx = rows(h->M[0]); a.Rx = rows(h->R[0]); a.My = rows(h->M[1]); a.Ry = rows(h->R[1]);
  a.Mz = rows(h->M[2]); a.Rz = rows(h->R[2]);
  a.ex = h->ebd; a.ey = h->ebd + 2 * n1; a.ez = h->ebd + 2 * (n1 + n2);
  a.fbx = h->fbf;
  a.fby = a.fbx + 2L * n2 * n3;
  a.fbz = a.fby + 2L * n1 * n3;
  a.out = h->R32;
  a.part = h->part;
  a.zq = h->zq;
  a.cz = h->res_cz;
  return a;
}

// The norms of the residual launches since the last flush, summed by one
// launch (block i: the partials of pending norm i).
int flush_norms(mf_solver3d *h, hipStream_t st) {
  if (h->npend == 0) return MF_OK;
  mf::NormOuts o;
  for (int i = 0; i < mf::NORM_SLOTS; ++i) o.p[i] = h->pend_out[i];
  hipLaunchKernelGGL(mf::sum_slots_kernel, dim3(h->npend), dim3(1024), 0, st,
                     (const double *)h->part, (long)h->part_stride, (long)h->pend_nb, o);
  h->npend = 0;
  HIPCHK(hipGetLastError());
  return MF_OK;
}

// One residual (or right-hand side) launch; its partial sums go to the next
// free slot and are summed into norm2_dev at the next flush_norms (at once
// when the slots are full), so that a solve's first norms take one launch.
template <bool RES>
int launch_residual(mf_solver3d *h, const mf::ResArgs &a, double *norm2_dev, hipStream_t st) {
  const bool zs = h->res_cz > 0;
  const int cz = zs ? h->res_cz : mf::RZ_CZ, ty = zs ? mf::ZS_TY : mf::RZ_TY;
  const unsigned nb = (unsigned)(((h->n1 + mf::RZ_TX - 1) / mf::RZ_TX) * ((h->n2 + ty - 1) / ty) *
                                 ((h->n3 + cz - 1) / cz));
  mf::ResArgs az = a;
  az.part = h->part + h->part_stride * h->npend;
  if (zs)
    hipLaunchKernelGGL(mf::residual_zs_kernel<RES>, dim3(nb), dim3(64, mf::ZS_NW), 0, st, az);
  else
    hipLaunchKernelGGL(mf::residual_kernel<RES>, dim3(nb), dim3(64, mf::RZ_NW), 0, st, az);
  HIPCHK(hipGetLastError());
  h->pend_out[h->npend] = norm2_dev;
  h->pend_nb = nb;
  if (++h->npend == mf::NORM_SLOTS) return flush_norms(h, st);
  return MF_OK;
}

// r' = P g_f - B u (FP32 copy -> R32) and ||r'||^2 -> norm2_dev, in one pass
// over u and g (plane pipeline; the face products must be current) or over u
// and b' = P g_f in FP64 (z-streaming kernel; b' from this solve's rhs_prime).
// scales: max |r'| into slot 0 of the fp16 path (cleared by the transforms
// before), slots 1-5 cleared for the transforms after
void with_scales(mf_solver3d *h, mf::ResArgs &a) {
  a.amax = h->fmx;
  a.zero = h->fmx + 1;
  a.nzero = 5;
}

int residual(mf_solver3d *h, const double *u, const double *g, double sigma, double *norm2_dev,
             bool scales, hipStream_t st) {
  const long N = (long)h->n1 * h->n2 * h->n3;
  ProfScope ps(h, P_RES, 20.0 * N, st);  // algorithmic: read u and g, write r' (FP32)
  mf::ResArgs a = res_args(h, u, g, sigma);
  if (scales) with_scales(h, a);
  if (h->res_cz > 0) a.bp = h->T1;  // b' of this solve (rhs_prime), in FP64
  return launch_residual<true>(h, a, norm2_dev, st);
}

// b' = P g_f (FP32 copy -> R32) with ||b'||^2, and the face products of the
// boundary fold for the residuals that follow.
int rhs_prime(mf_solver3d *h, const double *g, const double *bcx, const double *bcy,
              const double *bcz, double *norm2_dev, bool scales, hipStream_t st) {
  const int n1 = h->n1, n2 = h->n2, n3 = h->n3;
  const long N = (long)n1 * n2 * n3;
  // algorithmic: read g, write b' (FP32, and FP64 for the z-streaming residual)
  ProfScope ps(h, P_MULT, (h->res_cz > 0 ? 20.0 : 12.0) * N, st);
  mf::FaceFoldArgs f;
  f.n1 = n1; f.n2 = n2; f.n3 = n3;
  f.Mx = rows(h->M[0]); f.My = rows(h->M[1]); f.Mz = rows(h->M[2]);
  f.bcx = bcx; f.bcy = bcy; f.bcz = bcz;
  f.fbx = h->fbf;
  f.fby = f.fbx + 2L * n2 * n3;
  f.fbz = f.fby + 2L * n1 * n3;
  f.zero = scales ? h->fmx : nullptr;
  const long nf = 2 * ((long)n2 * n3 + (long)n1 * n3 + (long)n1 * n2);
  hipLaunchKernelGGL(mf::face_fold_kernel, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, st, f);
  mf::ResArgs a = res_args(h, nullptr, g, 0.0);
  if (scales) with_scales(h, a);
  if (h->res_cz > 0) a.out64 = h->T1;  // b' in FP64 for the residuals of this solve
  return launch_residual<false>(h, a, norm2_dev, st);
}
}  // namespace

namespace {
// One inner (low-precision) solve of the mixed path on r' (R32): u = S(r')
// or u += S(r').
void inner_solve(mf_solver3d *h, int precision, double sigma, double *u, bool accumulate,
                 hipStream_t st) {
  if (precision == MF_MIXED_BF16X3)
    transforms_bf3(h, h->R32, h->F1, h->F2, sigma, u, accumulate, st);
  else if (precision == MF_MIXED_FP16X2)
    transforms_h2(h, h->R32, h->F1, h->F2, sigma, u, accumulate, st);
  else
    transforms<float, float>(h, h->R32, h->F1, h->F2, h->Wx32, h->Wyt32, h->Wzt32, h->Vx32,
                             h->Vyt32, h->Vzt32, sigma, u, accumulate, st);
}

// u's faces from the latest sweep's interior and the norms of its residuals
// to the host mirror (pinned), both enqueued: the host reads the mirror only
// after a wait on the stream
int enqueue_check(mf_solver3d *h) {
  auto &p = h->pending;
  int rc = launch_faces(h, p.u, p.bcx, p.bcy, p.bcz, p.st);
  if (rc == MF_OK) rc = flush_norms(h, p.st);
  if (rc != MF_OK) return rc;
  HIPCHK(hipMemcpyAsync(h->hnorms, h->norms, (p.it + 1) * sizeof(double), hipMemcpyDeviceToHost,
                        p.st));
  return MF_OK;
}
}  // namespace

int mf::solver3d_solve_begin(mf_solver3d *h, const double *g, double sigma, double *u,
                             const double *bcx, const double *bcy, const double *bcz,
                             int precision, double tol, int max_iter, void *stream) {
  if (!h || !g || !u || !bcx || !bcy || !bcz)
    return mf::fail(MF_ERR_ARG, "mf_solver3d_solve: null pointer");
  if (precision != MF_FP64 && precision != MF_MIXED && precision != MF_MIXED_BF16X3 &&
      precision != MF_MIXED_FP16X2)
    return mf::fail(MF_ERR_ARG, "mf_solver3d_solve: precision must be MF_FP64, MF_MIXED, "
                                "MF_MIXED_BF16X3 or MF_MIXED_FP16X2");
  if (g == u) return mf::fail(MF_ERR_ARG, "mf_solver3d_solve: g and u must not alias");
  if (!std::isfinite(sigma)) return mf::fail(MF_ERR_ARG, "mf_solver3d_solve: sigma is not finite");
  hipStream_t st = (hipStream_t)stream;
  const long N = (long)h->n1 * h->n2 * h->n3;
  h->pending.active = false;
  h->npend = 0;
  if (precision == MF_FP64) {
    {
      ProfScope ps(h, P_FOLD, 16.0 * N, st);
      hipLaunchKernelGGL(mf::fold_kernel, dim3((h->n1 + 63) / 64, (h->n2 + 3) / 4, h->n3), dim3(64, 4), 0, st, g, h->T2, h->n1, h->n2,
                         h->n3, (const double *)h->fold, bcx, bcy, bcz);
    }
    if (h->null_mode && sigma == 0.0) {
      ProfScope ps(h, P_FOLD, 16.0 * N, st);
      hipLaunchKernelGGL(mf::null_dot_kernel, dim3(h->n3, mf::NULL_SLICES), dim3(256), 0, st,
                         (const double *)h->T2, h->n1, h->n2, (const double *)h->w0, h->n3,
                         h->npart);
      const long nb = std::min<long>((N + 255) / 256, 4096);
      hipLaunchKernelGGL(mf::null_sub_kernel, dim3((unsigned)nb), dim3(256), 0, st, h->T2, N,
                         (const double *)h->npart, h->n3 * mf::NULL_SLICES, h->inv_wsum);
    }
    transforms<double, double>(h, h->T2, h->T1, h->T2, h->Wx, h->Wyt, h->Wzt, h->Vx, h->Vyt,
                               h->Vzt, sigma, u, false, st);
    HIPCHK(hipGetLastError());
    return launch_faces(h, u, bcx, bcy, bcz, st);
  }
  // ---- mixed precision: FP32-class MFMA transforms, FP64 residual refinement
  if (!(tol > 0.0)) tol = 1e-12;
  if (max_iter <= 0) max_iter = 8;
  if (max_iter > MF_MAX_REFINE) max_iter = MF_MAX_REFINE;
  const bool scales = precision == MF_MIXED_FP16X2;
  int rc = rhs_prime(h, g, bcx, bcy, bcz, h->norms + 0, scales, st);
  if (rc != MF_OK) return rc;
  // first low-precision solve: u_int = S(b'); norms[it] = ||r'||^2 after it solves
  inner_solve(h, precision, sigma, u, false, st);
  rc = residual(h, u, g, sigma, h->norms + 1, scales, st);
  if (rc != MF_OK) return rc;
  int it = 1;
  // One FP32-class solve leaves a residual near 1e-7, so below that
  // tolerance the first correction is enqueued without waiting for the
  // first residual (one host round trip less).
  if (tol < 1e-6 && max_iter >= 2) {
    inner_solve(h, precision, sigma, u, true, st);
    rc = residual(h, u, g, sigma, h->norms + 2, scales, st);
    if (rc != MF_OK) return rc;
    it = 2;
  }
  auto &p = h->pending;
  p.g = g; p.u = u; p.bcx = bcx; p.bcy = bcy; p.bcz = bcz;
  p.sigma = sigma; p.tol = tol; p.precision = precision; p.max_iter = max_iter; p.it = it;
  p.st = st;
  rc = enqueue_check(h);
  if (rc != MF_OK) return rc;
  p.active = true;
  return MF_OK;
}

int mf::solver3d_solve_end(mf_solver3d *h, mf_solve_report *rep) {
  if (rep) { rep->iterations = 0; rep->residual = 0.0; rep->residual0 = 0.0; }
  if (!h || !h->pending.active) return MF_OK;  // FP64: nothing left to do
  auto &p = h->pending;
  p.active = false;
  const bool scales = p.precision == MF_MIXED_FP16X2;
  double rel = 0.0;
  for (;;) {
    HIPCHK(hipStreamSynchronize(p.st));
    const double b2 = h->hnorms[0];
    auto relres = [&](double r2) { return (b2 > 0.0) ? std::sqrt(r2 / b2) : std::sqrt(r2); };
    rel = relres(h->hnorms[p.it]);
    if (rep) {
      rep->residual0 = relres(h->hnorms[1]);
      rep->residual = rel;
      rep->iterations = p.it;
    }
    if (!(rel > p.tol) || p.it >= p.max_iter) break;
    inner_solve(h, p.precision, p.sigma, p.u, true, p.st);
    ++p.it;
    int rc = residual(h, p.u, p.g, p.sigma, h->norms + p.it, scales, p.st);
    if (rc != MF_OK) return rc;
    // a further sweep writes u's faces again
    rc = enqueue_check(h);
    if (rc != MF_OK) return rc;
  }
  if (!(rel <= p.tol))
    return mf::fail(MF_ERR_NOCONV, "mf_solver3d_solve: refinement stopped at relative residual " +
                                       std::to_string(rel));
  return MF_OK;
}

extern "C" int mf_solver3d_solve(mf_solver3d *h, const double *g, double sigma, double *u,
                                 const double *bcx, const double *bcy, const double *bcz,
                                 int precision, double tol, int max_iter, mf_solve_report *rep,
                                 void *stream) {
  if (rep) { rep->iterations = 0; rep->residual = 0.0; rep->residual0 = 0.0; }
  const int rc = mf::solver3d_solve_begin(h, g, sigma, u, bcx, bcy, bcz, precision, tol,
                                          max_iter, stream);
  if (rc != MF_OK) return rc;
  return mf::solver3d_solve_end(h, rep);
}

extern "C" int mf_solver3d_solve_host(mf_solver3d *h, const double *g, double sigma, double *u,
                                      const double *bcx, const double *bcy, const double *bcz,
                                      int precision, double tol, int max_iter,
                                      mf_solve_report *rep) {
  if (!h || !g || !u || !bcx || !bcy || !bcz)
    return mf::fail(MF_ERR_ARG, "mf_solver3d_solve_host: null pointer");
  const size_t full = (size_t)h->nx * h->ny * h->nz * sizeof(double);
  const size_t bx = (size_t)2 * h->n2 * h->n3 * sizeof(double);
  const size_t by = (size_t)2 * h->n1 * h->n3 * sizeof(double);
  const size_t bz = (size_t)2 * h->n1 * h->n2 * sizeof(double);
  std::lock_guard<std::mutex> lock(h->stage.mu);
  HIPCHK(h->stage.reserve(2 * full + bx + by + bz));
  char *d = (char *)h->stage.dev;
  double *dg = (double *)d, *du = (double *)(d + full);
  double *dbx = (double *)(d + 2 * full), *dby = (double *)(d + 2 * full + bx);
  double *dbz = (double *)(d + 2 * full + bx + by);
  hipError_t e = hipMemcpy(dg, g, full, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(du, u, full, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dbx, bcx, bx, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dby, bcy, by, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dbz, bcz, bz, hipMemcpyHostToDevice);
  int rc = MF_OK;
  if (e == hipSuccess) {
    rc = mf_solver3d_solve(h, dg, sigma, du, dbx, dby, dbz, precision, tol, max_iter, rep, nullptr);
    if (rc == MF_OK || rc == MF_ERR_NOCONV) {
      hipError_t e2 = hipMemcpy(u, du, full, hipMemcpyDeviceToHost);
      if (e2 != hipSuccess) e = e2;
    }
  }
  if (e != hipSuccess)
    return mf::fail(MF_ERR_DEVICE, std::string("mf_solver3d_solve_host: ") + hipGetErrorString(e));
  return rc;
}

// ---------------------------------------------------------------------------
// Unit-impulse batch (md_influence_matrix_set, class_field.f90:1396-1510)
// ---------------------------------------------------------------------------
extern "C" int mf_influence_impulses(mf_solver3d *h, const mf_der *dx, const mf_der *dy,
                                     const mf_der *dz, double sigma, int face, long long first,
                                     int count, double *out, void *stream) {
  if (!h || !dx || !dy || !dz || !out) return mf::fail(MF_ERR_ARG, "mf_influence_impulses: null pointer");
  if (face < 0 || face > 5) return mf::fail(MF_ERR_ARG, "mf_influence_impulses: face must be 0..5");
  if (h->bct[face] != 1)
    return mf::fail(MF_ERR_ARG, "mf_influence_impulses: the impulse face must be Dirichlet (an interface)");
  const int n1 = h->n1, n2 = h->n2, n3 = h->n3;
  const int s = face / 2;
  const long fpts = (s == 0) ? (long)n2 * n3 : (s == 1 ? (long)n1 * n3 : (long)n1 * n2);
  if (count < 0 || first < 0 || first + count > fpts)
    return mf::fail(MF_ERR_ARG, "mf_influence_impulses: impulse range outside the face");
  if (count == 0) return MF_OK;
  if (!std::isfinite(sigma)) return mf::fail(MF_ERR_ARG, "mf_influence_impulses: sigma is not finite");
  // derivative rows of the three full-line schemes (first derivative)
  const mf_der *ds[3] = {dx, dy, dz};
  const int nfull[3] = {h->nx, h->ny, h->nz};
  std::vector<double> drow;
  for (int a = 0; a < 3; ++a) {
    const double *dl = nullptr, *dr = nullptr;
    int n = 0;
    if (!mf::der_host_tables(ds[a], 1, 0, &dl, &dr, &n) || n != nfull[a])
      return mf::fail(MF_ERR_ARG, "mf_influence_impulses: derivative handle does not match the grid");
    std::vector<double> r0(n), rn(n);
    mf::derivative_rows(n, dl, dr, r0.data(), rn.data());
    drow.insert(drow.end(), r0.begin(), r0.end());
    drow.insert(drow.end(), rn.begin(), rn.end());
  }
  // ahat = W_s (-fold of the impulse face)
  const mf::Solver1D &S1 = h->ax[s];
  const int ns = S1.n;
  std::vector<double> fold(ns), ahat(ns, 0.0);
  {
    auto IL = [&](int i, int j) { return S1.inv_lhs[i + (size_t)ns * j]; };
    auto DR = [&](int i, int c) { return S1.ddr[i + (size_t)ns * c]; };
    for (int i = 0; i < ns; ++i)
      fold[i] = (face % 2 == 0) ? IL(i, 0) * DR(0, 0) + IL(i, 1) * DR(1, 0)
                                : IL(i, ns - 2) * DR(ns - 2, 7) + IL(i, ns - 1) * DR(ns - 1, 6);
    for (int p = 0; p < ns; ++p) {
      double acc = 0.0;
      for (int i = 0; i < ns; ++i) acc += S1.inv_vect[p + (size_t)ns * i] * (-fold[i]);
      ahat[p] = acc;
    }
  }
  // batch size: two interior fields per impulse, within a 2 GB workspace and
  // the launch-grid limits of the batched transforms
  const long N = (long)n1 * n2 * n3;
  long B = (2L << 30) / (2 * N * (long)sizeof(double));
  B = std::min<long>(B, 65535L / n3);
  B = std::min<long>(B, (65535L * 64) / ((long)n2 * n3));
  if (const char *cap = std::getenv("MFLOPS_IMPULSE_BATCH")) {  // smaller batches on request
    const long c = std::atol(cap);
    if (c > 0) B = std::min(B, c);
  }
  B = std::max<long>(1, std::min<long>(B, count));
  const size_t small = (size_t)(ns + drow.size() + 64);
  const size_t need = (2 * (size_t)B * N + small) * sizeof(double);
  hipStream_t st = (hipStream_t)stream;
  if (h->iws_bytes < need) {
    if (h->iws) {
      HIPCHK(hipStreamSynchronize(st));
      (void)hipFree(h->iws);
      h->iws = nullptr;
      h->iws_bytes = 0;
    }
    HIPCHK(hipMalloc(&h->iws, need));
    h->iws_bytes = need;
  }
  double *W1 = (double *)h->iws, *W2 = W1 + (size_t)B * N;
  double *d_ahat = W2 + (size_t)B * N;
  double *d_drow = d_ahat + ns;
  HIPCHK(hipMemcpyAsync(d_ahat, ahat.data(), ns * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(d_drow, drow.data(), drow.size() * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));  // host vectors go out of scope after return
  const long stride = 2L * n2 * n3 + 2L * n1 * n3 + 2L * n1 * n2;
  const long n12 = (long)n1 * n2;
  EpiArgs ep;
  std::memset(&ep, 0, sizeof(ep));
  for (long b0 = 0; b0 < count; b0 += B) {
    const int nb = (int)std::min<long>(B, count - b0);
    mf::ImpulseArgs ia;
    ia.n1 = n1; ia.n2 = n2; ia.n3 = n3; ia.face = face; ia.first = first + b0;
    ia.ahat = d_ahat; ia.Wx = h->Wx; ia.Wyt = h->Wyt; ia.Wzt = h->Wzt;
    ia.lx = h->lam; ia.ly = h->lam + n1; ia.lz = h->lam + n1 + n2; ia.sigma = sigma;
    {
      ProfScope ps(h, P_FOLD, 8.0 * N * nb, st);
      hipLaunchKernelGGL(mf::impulse_spectral_pq_kernel,
                         dim3((unsigned)(((long)n1 * n2 + 255) / 256), 1, nb), dim3(256), 0, st, W1, ia);
    }
    // backward transforms, batched over the impulses
    launch_gemm<double, double, double, double, false, mf::EPI_STORE>(
        h, W1, n12, N, h->Vzt, n3, W2, n12, N, (int)n12, n3, n3, nb, ep, st);
    launch_gemm<double, double, double, double, false, mf::EPI_STORE>(
        h, W2, n1, n12, h->Vyt, n2, W1, n1, n12, n1, n2, n2, n3 * nb, ep, st);
    launch_gemm<double, double, double, double, true, mf::EPI_STORE>(
        h, h->Vx, n1, 0, W1, n1, W2, n1, 0, n1, (int)((long)n2 * n3 * nb), n1, 1, ep, st);
    mf::FaceDerArgs fa;
    fa.n1 = n1; fa.n2 = n2; fa.n3 = n3; fa.face = face; fa.first = first + b0;
    fa.bct = h->bctd; fa.neum = h->neum; fa.drow = d_drow; fa.out_stride = stride;
    double *o = out + b0 * stride;
    {
      ProfScope ps(h, P_FACES, 24.0 * N * nb, st);
      hipLaunchKernelGGL(mf::face_der_x_kernel, dim3((n2 + 3) / 4, n3, nb), dim3(64, 4), 0, st,
                         (const double *)W2, o, fa);
      hipLaunchKernelGGL((mf::face_der_yz_kernel<1>),
                         dim3((unsigned)(((long)n1 * n3 + 255) / 256), 1, nb), dim3(256), 0, st,
                         (const double *)W2, o, fa);
      hipLaunchKernelGGL((mf::face_der_yz_kernel<2>),
                         dim3((unsigned)(((long)n1 * n2 + 255) / 256), 1, nb), dim3(256), 0, st,
                         (const double *)W2, o, fa);
    }
    HIPCHK(hipGetLastError());
  }
  return MF_OK;
}

// ---------------------------------------------------------------------------
// Influence blocks by spectral contraction (infl_spec.h): the influence
// matrix of md_influence_matrix_set (class_field.f90:1396-1510) without one
// solve per impulse
// ---------------------------------------------------------------------------
namespace {

// ahat = inv_vect_s (-fold) of impulse face (s, side): the spectral
// coefficients of a unit boundary value folded into the right-hand side
std::vector<double> impulse_ahat(const mf_solver3d *h, int s, int side) {
  const mf::Solver1D &S1 = h->ax[s];
  const int ns = S1.n;
  std::vector<double> fold(ns), ahat(ns, 0.0);
  auto IL = [&](int i, int j) { return S1.inv_lhs[i + (size_t)ns * j]; };
  auto DR = [&](int i, int c) { return S1.ddr[i + (size_t)ns * c]; };
  for (int i = 0; i < ns; ++i)
    fold[i] = side == 0 ? IL(i, 0) * DR(0, 0) + IL(i, 1) * DR(1, 0)
                        : IL(i, ns - 2) * DR(ns - 2, 7) + IL(i, ns - 1) * DR(ns - 1, 6);
  for (int p = 0; p < ns; ++p) {
    double acc = 0.0;
    for (int i = 0; i < ns; ++i) acc += S1.inv_vect[p + (size_t)ns * i] * (-fold[i]);
    ahat[p] = acc;
  }
  return ahat;
}

// Host tables of one axis: the boundary-derivative rows of the two faces
// with the Neumann extrapolation of the line folded in (as the face
// derivatives of the impulse path form them), in spectral space (what), the
// products Om = diag(what) W, and the Dirichlet end weights.
struct AxisTabs {
  int n = 0;
  std::vector<double> what[2], Om[2];
  double dcoef[2][2];  // [output side][impulse side]
};

bool axis_tabs(const mf_solver3d *h, const mf_der *d, int a, AxisTabs &t) {
  const mf::Solver1D &S = h->ax[a];
  const int n = S.n;
  t.n = n;
  const double *dl = nullptr, *dr = nullptr;
  int nall = 0;
  if (!mf::der_host_tables(d, 1, 0, &dl, &dr, &nall) || nall != n + 2) return false;
  std::vector<double> r0(nall), rn(nall);
  mf::derivative_rows(nall, dl, dr, r0.data(), rn.data());
  // boundary value of a Neumann end as a combination of the first/last six
  // interior values (extrapolation_coeffs, class_solver_3d.f90:622-656)
  std::vector<double> eL(n, 0.0), eH(n, 0.0);
  const double *nm = S.neum.data();
  if (S.bc[0] != 1)
    for (int m = 1; m < 7; ++m) eL[m - 1] = -nm[2 * m] / nm[0];
  if (S.bc[1] != 1)
    for (int m = 0; m < 6; ++m) eH[n - 6 + m] = -nm[1 + 2 * m] / nm[13];
  std::vector<double> w[2] = {std::vector<double>(n), std::vector<double>(n)};
  for (int i = 0; i < n; ++i) {
    w[0][i] = r0[i + 1] + r0[0] * eL[i] + r0[n + 1] * eH[i];
    w[1][i] = -(rn[i + 1] + rn[0] * eL[i] + rn[n + 1] * eH[i]);
  }
  for (int side = 0; side < 2; ++side) {
    t.what[side].assign(n, 0.0);
    for (int al = 0; al < n; ++al) {
      double acc = 0.0;
      for (int i = 0; i < n; ++i) acc += w[side][i] * S.vect[i + (size_t)n * al];
      t.what[side][al] = acc;
    }
    t.Om[side].resize((size_t)n * n);
    for (int c = 0; c < n; ++c)
      for (int al = 0; al < n; ++al)
        t.Om[side][al + (size_t)n * c] = t.what[side][al] * S.inv_vect[al + (size_t)n * c];
  }
  t.dcoef[0][0] = r0[0];
  t.dcoef[0][1] = r0[n + 1];
  t.dcoef[1][0] = -rn[0];
  t.dcoef[1][1] = -rn[n + 1];
  return true;
}

// device offsets (doubles) of one subdomain's uploaded tables
struct SubTabs {
  size_t V[3], W[3], what[3][2], Om[3][2];
  size_t ahat[6], VsA[6], VsT[6];
  double dcoef[3][2][2];
};

// one face pair of one subdomain
struct PairPlan {
  int sub, fi, fo;          // subdomain; impulse and output face (0..5)
  int ii, io;               // their positions among the subdomain's interface faces
  size_t buf1, buf2;        // work offsets: G or H, then Z or Y
  int tile;
};

int tile_bm(int tile) { return tile == 0 ? 256 : (tile == 1 ? 128 : 64); }

template <int BM, int BN>
void launch_infl(const mf::InflProblem *probs, int nprob, long tiles, int blocks, hipStream_t st) {
  hipLaunchKernelGGL((mf::infl_gemm_kernel<BM, BN>), dim3(blocks), dim3(256), 0, st, probs, nprob,
                     tiles);
}

void launch_infl_tile(int tile, const mf::InflProblem *probs, int nprob, long tiles, int blocks,
                      hipStream_t st) {
  if (tile == 0) launch_infl<256, 64>(probs, nprob, tiles, blocks, st);
  else if (tile == 1) launch_infl<128, 128>(probs, nprob, tiles, blocks, st);
  else launch_infl<64, 256>(probs, nprob, tiles, blocks, st);
}

}  // namespace

int mf::influence_assemble(const std::vector<InflSub> &subs, double sigma, bool graph, void *stream,
                           InflStats *stats) {
  using mf::InflProblem;
  if (subs.empty()) return MF_OK;
  const int nsub = (int)subs.size();
  // ---- host tables and their device layout
  std::vector<double> htab;
  std::vector<SubTabs> tabs(nsub);
  auto put = [&](const std::vector<double> &v) {
    const size_t o = htab.size();
    htab.insert(htab.end(), v.begin(), v.end());
    return o;
  };
  for (int i = 0; i < nsub; ++i) {
    const InflSub &S = subs[i];
    const mf_solver3d *h = S.sc;
    for (int a = 0; a < 3; ++a) {
      AxisTabs at;
      if (!axis_tabs(h, S.d[a], a, at)) return mf::fail(MF_ERR_SETUP, "influence_assemble: derivative tables");
      tabs[i].V[a] = put(h->ax[a].vect);
      tabs[i].W[a] = put(h->ax[a].inv_vect);
      for (int side = 0; side < 2; ++side) {
        tabs[i].what[a][side] = put(at.what[side]);
        tabs[i].Om[a][side] = put(at.Om[side]);
        for (int is = 0; is < 2; ++is) tabs[i].dcoef[a][side][is] = at.dcoef[side][is];
      }
    }
    for (int f = 0; f < S.nf; ++f) {
      const int face = S.face[f], s = face / 2, n = h->ax[s].n;
      const std::vector<double> ah = impulse_ahat(h, s, face % 2);
      std::vector<double> va((size_t)n * n), vt((size_t)n * n);
      for (int al = 0; al < n; ++al)
        for (int o = 0; o < n; ++o) {
          const double v = h->ax[s].vect[o + (size_t)n * al] * ah[al];
          va[o + (size_t)n * al] = v;
          vt[al + (size_t)n * o] = v;
        }
      tabs[i].ahat[face] = put(ah);
      tabs[i].VsA[face] = put(va);
      tabs[i].VsT[face] = put(vt);
    }
  }
  // ---- work layout per subdomain: T, TT per axis, then two buffers per pair
  std::vector<PairPlan> pairs;
  std::vector<size_t> sub_work(nsub), sub_T(nsub);
  for (int i = 0; i < nsub; ++i) {
    const InflSub &S = subs[i];
    const int n[3] = {S.sc->n1, S.sc->n2, S.sc->n3};
    size_t off = 0;
    sub_T[i] = 0;
    for (int a = 0; a < 3; ++a) off += 2 * (size_t)n[a] * n[a] * n[a];
    for (int fi = 0; fi < S.nf; ++fi)
      for (int fo = 0; fo < S.nf; ++fo) {
        PairPlan pp;
        pp.sub = i; pp.ii = fi; pp.io = fo; pp.fi = S.face[fi]; pp.fo = S.face[fo];
        const int s = pp.fi / 2, t = pp.fo / 2;
        size_t b1, b2;
        if (s == t) {
          const int b = s == 0 ? 1 : 0, c = s == 2 ? 1 : 2;
          b1 = 2 * (size_t)n[b] * n[c];
          b2 = (size_t)n[b] * n[c] * n[c];
        } else {
          b1 = b2 = (size_t)n[0] * n[1] * n[2];
        }
        pp.buf1 = off; off += b1;
        pp.buf2 = off; off += b2;
        pairs.push_back(pp);
      }
    sub_work[i] = off;
  }
  // ---- groups of consecutive subdomains within a work budget
  size_t free_b = 0, total_b = 0;
  HIPCHK(hipMemGetInfo(&free_b, &total_b));
  size_t budget = free_b / 4 / sizeof(double);
  size_t need_max = 0;
  for (int i = 0; i < nsub; ++i) need_max = std::max(need_max, sub_work[i]);
  budget = std::max(budget, need_max);
  std::vector<int> group_of(nsub);
  std::vector<size_t> sub_base(nsub);
  size_t work_doubles = 0;
  {
    int g = 0;
    size_t used = 0;
    for (int i = 0; i < nsub; ++i) {
      if (used + sub_work[i] > budget && used > 0) { ++g; used = 0; }
      group_of[i] = g;
      sub_base[i] = used;
      used += sub_work[i];
      work_doubles = std::max(work_doubles, used);
    }
  }
  const int ngroups = group_of[nsub - 1] + 1;
  // ---- problem descriptors: per group, per tile class, contiguous
  int ncu = 256;
  {
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
      int v = 0;
      if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) ncu = v;
    }
  }
  // pointers into buffers that do not exist yet: record offsets, patch below
  struct ProbRef { int pair; int cls; };
  std::vector<std::vector<ProbRef>> order(ngroups * 3);
  for (size_t q = 0; q < pairs.size(); ++q) {
    PairPlan &pp = pairs[q];
    const InflSub &S = subs[pp.sub];
    const int n[3] = {S.sc->n1, S.sc->n2, S.sc->n3};
    const int s = pp.fi / 2, t = pp.fo / 2;
    int K;
    if (s == t) K = n[s == 0 ? 1 : 0];
    else K = n[3 - s - t];
    pp.tile = mf::infl_tile_for(K).tile;
    order[group_of[pp.sub] * 3 + pp.tile].push_back({(int)q, pp.tile});
  }
  // ---- device buffers
  double *dtab = nullptr, *dwork = nullptr;
  InflProblem *dprob = nullptr;
  mf::InflDiag *ddiag = nullptr;
  struct Launch { int tile; long first; int nprob; long tiles; int blocks; };
  std::vector<std::vector<Launch>> launches(ngroups);
  std::vector<InflProblem> hprob;
  std::vector<std::vector<mf::InflDiag>> hdiag(ngroups);
  int rc = MF_OK;
  hipStream_t cs = nullptr;
  hipGraph_t gr = nullptr;
  hipGraphExec_t ge = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  auto cleanup = [&]() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ge) (void)hipGraphExecDestroy(ge);
    if (gr) (void)hipGraphDestroy(gr);
    if (cs) { (void)hipStreamSynchronize(cs); (void)hipStreamDestroy(cs); }
    if (dtab) (void)hipFree(dtab);
    if (dwork) (void)hipFree(dwork);
    if (dprob) (void)hipFree(dprob);
    if (ddiag) (void)hipFree(ddiag);
  };
  auto fail_dev = [&](hipError_t e, const char *what) {
    cleanup();
    return mf::fail(MF_ERR_DEVICE, std::string("influence_assemble: ") + what + ": " + hipGetErrorString(e));
  };
  hipError_t e = hipMalloc(&dtab, htab.size() * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dwork, std::max<size_t>(work_doubles, 1) * sizeof(double));
  if (e != hipSuccess) return fail_dev(e, "hipMalloc");
  for (int g = 0; g < ngroups; ++g)
    for (int cls = 0; cls < 3; ++cls) {
      const std::vector<ProbRef> &lst = order[g * 3 + cls];
      if (lst.empty()) continue;
      Launch L;
      L.tile = cls;
      L.first = (long)hprob.size();
      L.nprob = (int)lst.size();
      long tiles = 0;
      const int BM = tile_bm(cls), BN = 256 * 64 / BM;
      for (const ProbRef &r : lst) {
        const PairPlan &pp = pairs[r.pair];
        const InflSub &S = subs[pp.sub];
        const int n[3] = {S.sc->n1, S.sc->n2, S.sc->n3};
        const int s = pp.fi / 2, t = pp.fo / 2;
        double *W = dwork + sub_base[pp.sub];
        size_t toff[3], ttoff[3];
        {
          size_t o = 0;
          for (int a = 0; a < 3; ++a) {
            toff[a] = o; o += (size_t)n[a] * n[a] * n[a];
            ttoff[a] = o; o += (size_t)n[a] * n[a] * n[a];
          }
        }
        InflProblem p;
        std::memset(&p, 0, sizeof(p));
        if (s == t) {
          const int b = s == 0 ? 1 : 0, c = s == 2 ? 1 : 2;
          p.A = W + toff[b]; p.lda = (long)n[b] * n[b];
          p.B = W + pp.buf2; p.ldb = n[b];
          p.M = n[b] * n[b]; p.N = n[c] * n[c]; p.K = n[b];
          p.P = n[b]; p.Q = n[c]; p.r2 = n[b]; p.c1 = 1; p.c2 = n[b];
          p.dcoef = tabs[pp.sub].dcoef[s][pp.fo % 2][pp.fi % 2];
        } else {
          const int u = 3 - s - t;
          if (s < u) {
            p.A = W + pp.buf2; p.lda = (long)n[s] * n[t];
            p.B = W + ttoff[u]; p.ldb = n[u];
            p.M = n[s] * n[t]; p.N = n[u] * n[u]; p.K = n[u];
            p.P = n[s]; p.Q = n[u]; p.r2 = n[s];
            if (t < u) { p.c1 = 1; p.c2 = n[t]; } else { p.c1 = n[u]; p.c2 = 1; }
          } else {
            p.A = W + toff[u]; p.lda = (long)n[u] * n[u];
            p.B = W + pp.buf2; p.ldb = n[u];
            p.M = n[u] * n[u]; p.N = n[s] * n[t]; p.K = n[u];
            p.P = n[u]; p.Q = n[s]; p.r2 = n[u];
            if (t < u) { p.c1 = n[t]; p.c2 = 1; } else { p.c1 = 1; p.c2 = n[u]; }
          }
          p.dcoef = 0.0;
        }
        p.D = S.D + S.loc_off[pp.io] + S.u * S.loc_off[pp.ii];
        p.ldd = S.u;
        if (s == t) {
          mf::InflDiag dg;
          dg.D = p.D;
          dg.ldd = p.ldd;
          dg.count = p.P * p.Q;
          dg.dcoef = p.dcoef;
          hdiag[g].push_back(dg);
        }
        p.tiles_n = (p.N + BN - 1) / BN;
        p.tile0 = tiles;
        tiles += (long)((p.M + BM - 1) / BM) * p.tiles_n;
        hprob.push_back(p);
      }
      L.tiles = tiles;
      const mf::InflTile tt = mf::infl_tile_for(hprob[L.first].K);
      L.blocks = (int)std::max<long>(1, std::min<long>(tiles, (long)ncu * tt.blocks_per_cu));
      launches[g].push_back(L);
    }
  // diagonal terms of every group, contiguous
  std::vector<mf::InflDiag> alld;
  std::vector<size_t> diag_first(ngroups + 1, 0);
  int diag_max = 1;
  for (int g = 0; g < ngroups; ++g) {
    diag_first[g] = alld.size();
    for (const mf::InflDiag &d : hdiag[g]) {
      alld.push_back(d);
      diag_max = std::max(diag_max, d.count);
    }
  }
  diag_first[ngroups] = alld.size();
  e = hipMalloc(&ddiag, std::max<size_t>(alld.size(), 1) * sizeof(mf::InflDiag));
  if (e == hipSuccess && !alld.empty())
    e = hipMemcpy(ddiag, alld.data(), alld.size() * sizeof(mf::InflDiag), hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail_dev(e, "diagonal list");
  e = hipMalloc(&dprob, std::max<size_t>(hprob.size(), 1) * sizeof(InflProblem));
  if (e == hipSuccess) e = hipMemcpy(dtab, htab.data(), htab.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess && !hprob.empty())
    e = hipMemcpy(dprob, hprob.data(), hprob.size() * sizeof(InflProblem), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&ev0);
  if (e == hipSuccess) e = hipEventCreate(&ev1);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess) return fail_dev(e, "set-up");
  if (stats) {
    stats->gemm_flop = 0.0;
    stats->block_bytes = 0.0;
    for (const InflProblem &p : hprob) {
      stats->gemm_flop += 2.0 * p.M * p.N * p.K;
      stats->block_bytes += 8.0 * p.M * p.N;
    }
  }
  // ---- the launch sequence: phase A per subdomain, then the grouped GEMMs
  auto record = [&](hipStream_t st) -> int {
    for (int g = 0; g < ngroups; ++g) {
      for (int i = 0; i < nsub; ++i) {
        if (group_of[i] != g) continue;
        const InflSub &S = subs[i];
        mf_solver3d *h = S.sc;
        const int n[3] = {h->n1, h->n2, h->n3};
        double *W = dwork + sub_base[i];
        const double *tb = dtab;
        const SubTabs &tt = tabs[i];
        size_t toff[3], ttoff[3];
        {
          size_t o = 0;
          for (int a = 0; a < 3; ++a) {
            toff[a] = o; o += (size_t)n[a] * n[a] * n[a];
            ttoff[a] = o; o += (size_t)n[a] * n[a] * n[a];
          }
        }
        for (int a = 0; a < 3; ++a) {
          const long n3 = (long)n[a] * n[a] * n[a];
          hipLaunchKernelGGL(mf::infl_t_kernel, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st,
                             tb + tt.V[a], tb + tt.W[a], n[a], W + toff[a], W + ttoff[a]);
        }
        mf::EigArgs ea;
        ea.l[0] = h->lam; ea.l[1] = h->lam + n[0]; ea.l[2] = h->lam + n[0] + n[1];
        ea.sigma = sigma;
        EpiArgs ep;
        std::memset(&ep, 0, sizeof(ep));
        for (const PairPlan &pp : pairs) {
          if (pp.sub != i) continue;
          const int s = pp.fi / 2, t = pp.fo / 2;
          double *b1 = W + pp.buf1, *b2 = W + pp.buf2;
          if (s == t) {
            const int b = s == 0 ? 1 : 0, c = s == 2 ? 1 : 2;
            const long cnt = 2L * n[b] * n[c];
            hipLaunchKernelGGL(mf::infl_g_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st,
                               ea, s, b, c, n[s], n[b], n[c], tb + tt.what[s][0], tb + tt.what[s][1],
                               tb + tt.ahat[pp.fi], b1);
            // Z = G_side TT_c
            const double *G = b1 + (long)(pp.fo % 2) * n[b] * n[c];
            launch_gemm<double, double, double, double, true, mf::EPI_STORE>(
                h, G, n[b], 0, W + ttoff[c], n[c], b2, n[b], 0, n[b], n[c] * n[c], n[c], 1, ep, st);
          } else {
            const int u = 3 - s - t;
            const int layout = s < u ? 0 : 1;
            const long cnt = (long)n[0] * n[1] * n[2];
            hipLaunchKernelGGL(mf::infl_h_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st,
                               ea, s, t, u, n[s], n[t], n[u], layout, tb + tt.Om[t][pp.fo % 2], b1);
            if (layout == 0)  // Y = (Vs diag(ahat)) H
              launch_gemm<double, double, double, double, true, mf::EPI_STORE>(
                  h, tb + tt.VsA[pp.fi], n[s], 0, b1, n[s], b2, n[s], 0, n[s], n[t] * n[u], n[s], 1,
                  ep, st);
            else  // Y'_ct = H'_ct (Vs diag(ahat))^T, batched over ct
              launch_gemm<double, double, double, double, true, mf::EPI_STORE>(
                  h, b1, n[u], (long)n[u] * n[s], tb + tt.VsT[pp.fi], n[s], b2, n[u],
                  (long)n[u] * n[s], n[u], n[s], n[s], n[t], ep, st);
          }
        }
      }
      for (const Launch &L : launches[g])
        launch_infl_tile(L.tile, dprob + L.first, L.nprob, L.tiles, L.blocks, st);
      const int ndg = (int)(diag_first[g + 1] - diag_first[g]);
      if (ndg > 0)
        hipLaunchKernelGGL(mf::infl_diag_kernel, dim3((unsigned)((diag_max + 255) / 256), ndg), dim3(256), 0,
                           st, (const mf::InflDiag *)(ddiag + diag_first[g]));
      if (hipGetLastError() != hipSuccess) return mf::fail(MF_ERR_DEVICE, "influence_assemble: launch");
    }
    return MF_OK;
  };
  if (graph) {
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return fail_dev(e, "hipStreamBeginCapture");
    rc = record(cs);
    e = hipStreamEndCapture(cs, &gr);
    if (rc != MF_OK) { cleanup(); return rc; }
    if (e != hipSuccess) return fail_dev(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
    if (e == hipSuccess) e = hipEventRecord(ev0, cs);
    if (e == hipSuccess) e = hipGraphLaunch(ge, cs);
    if (e == hipSuccess) e = hipEventRecord(ev1, cs);
    if (e != hipSuccess) return fail_dev(e, "graph launch");
  } else {
    e = hipEventRecord(ev0, cs);
    if (e != hipSuccess) return fail_dev(e, "hipEventRecord");
    rc = record(cs);
    if (rc != MF_OK) { cleanup(); return rc; }
    e = hipEventRecord(ev1, cs);
    if (e != hipSuccess) return fail_dev(e, "hipEventRecord");
  }
  e = hipStreamSynchronize(cs);
  if (e != hipSuccess) return fail_dev(e, "hipStreamSynchronize");
  if (stats) {
    float t = 0.0f;
    if (hipEventElapsedTime(&t, ev0, ev1) == hipSuccess) stats->ms = t;
  }
  cleanup();
  return MF_OK;
}

// Autotuner measurement: one block product of M x N x K through the grouped
// kernel with tile class `tile` and blocks_per_cu blocks per compute unit;
// mean milliseconds per launch over reps launches into ms.
extern "C" int mf_tune_infl_gemm(int tile, int blocks_per_cu, int M, int N, int K, int reps,
                                 double *ms) {
  if (tile < 0 || tile > 2 || blocks_per_cu < 1 || blocks_per_cu > 16 || M < 1 || N < 1 || K < 1 ||
      reps < 1 || !ms)
    return mf::fail(MF_ERR_ARG, "mf_tune_infl_gemm: invalid argument");
  double *A = nullptr, *B = nullptr, *D = nullptr;
  mf::InflProblem *dp = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto done = [&]() {
    if (A) (void)hipFree(A);
    if (B) (void)hipFree(B);
    if (D) (void)hipFree(D);
    if (dp) (void)hipFree(dp);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  hipError_t e = hipMalloc(&A, (size_t)M * K * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&B, (size_t)K * N * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&D, (size_t)M * N * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dp, sizeof(mf::InflProblem));
  if (e == hipSuccess) e = hipMemset(A, 0, (size_t)M * K * sizeof(double));
  if (e == hipSuccess) e = hipMemset(B, 0, (size_t)K * N * sizeof(double));
  const int BM = tile_bm(tile), BN = 256 * 64 / BM;
  mf::InflProblem p;
  std::memset(&p, 0, sizeof(p));
  p.A = A; p.lda = M; p.B = B; p.ldb = K; p.D = D; p.ldd = M;
  p.M = M; p.N = N; p.K = K;
  p.P = M; p.Q = N; p.r2 = M; p.c1 = 0; p.c2 = 0;  // D[m + M n]
  p.tiles_n = (N + BN - 1) / BN;
  p.tile0 = 0;
  const long tiles = (long)((M + BM - 1) / BM) * p.tiles_n;
  int ncu = 256, dev = 0;
  if (e == hipSuccess) e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipMemcpy(dp, &p, sizeof(p), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e != hipSuccess) { done(); return mf::fail(MF_ERR_DEVICE, std::string("mf_tune_infl_gemm: ") + hipGetErrorString(e)); }
  const int blocks = (int)std::max<long>(1, std::min<long>(tiles, (long)ncu * blocks_per_cu));
  launch_infl_tile(tile, dp, 1, tiles, blocks, nullptr);  // warm-up
  e = hipEventRecord(e0, nullptr);
  for (int r = 0; r < reps && e == hipSuccess; ++r) launch_infl_tile(tile, dp, 1, tiles, blocks, nullptr);
  if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  if (e == hipSuccess) e = hipGetLastError();
  done();
  if (e != hipSuccess) return mf::fail(MF_ERR_DEVICE, std::string("mf_tune_infl_gemm: ") + hipGetErrorString(e));
  *ms = t / reps;
  return MF_OK;
}
