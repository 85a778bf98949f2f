// Multidomain (influence-matrix) Helmholtz/Poisson solve on one GPU.
//
// The reference (multidomain_solve, src/class_field/class_field.f90:1295-1388;
// md_solve, src/class_md/class_md.f90:379-438) makes the subdomain solutions
// C1 across shared interface planes: with the interface values x as the
// subdomains' Dirichlet data, the sum over both sides of the outward-normal
// first derivatives must vanish, which is the linear system M x = rhs with
// the influence matrix M of unit-impulse responses (md_influence_matrix_set,
// class_field.f90:1396-1510) and rhs from the solve with zero interface
// values.  PETSc assembles M as a sparse matrix of dense blocks and solves
// with KSP.
//
// Here every subdomain s owns one dense block D_s over the unknowns of its
// own interfaces (u_s x u_s, column-major in HBM): M = sum_s E_s D_s E_s^T,
// with E_s the injection of s's interface unknowns.  D_s is filled from the
// batched unit-impulse responses (mf_influence_impulses) by a scatter kernel.
// A matrix-vector product is then one gather, one memory-bound GEMV per
// subdomain (column chunks, fixed-order partial sums) and one combine of the
// two sides of every interface unknown, so M x is deterministic.  The system
// is solved by restarted GMRES on the device (classical Gram-Schmidt applied
// twice, fixed-order dot products; the small Hessenberg least-squares
// problem on the host).  Interface numbering and unknown order follow
// md_influence_matrix_param (class_md.f90:1450-1630), generalised to one
// point count per slab; a periodic axis adds the interface that joins the
// last subdomain's high face to the first one's low face (:1482-1560).
// Singular systems (sigma = 0 with only Neumann or periodic outer faces: the
// constant interface vector is a null vector of M) follow the pressure
// singular method psm: 1 removes the constant from the right-hand side and
// from every Krylov vector, as PETSc's constant null space does, so x keeps
// the mean of its start; 2 zeroes the first right-hand-side entry.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"
#include "mflops/mf.h"

namespace mf {
namespace md {

// D[(loc_row) + u * (col0 + b)] = out[b * stride + face_off + p] for every
// interface face of the subdomain; one thread per (row, impulse).
struct ScatterArgs {
  double *D;
  long u;               // block size
  long col0;            // first column (local unknown of the first impulse)
  const double *out;    // impulse responses (count x stride)
  long stride;
  int nf;               // interface faces of the subdomain
  long face_off[6];     // offset of each interface face in an output row
  long loc_off[7];      // local unknown offsets of the faces (loc_off[nf] = u)
};

__global__ void __launch_bounds__(256) scatter_kernel(ScatterArgs a, int count) {
  const long r = (long)blockIdx.x * 256 + threadIdx.x;
  const int b = blockIdx.y;
  if (r >= a.u || b >= count) return;
  int f = 0;
  while (f + 1 < a.nf && r >= a.loc_off[f + 1]) ++f;
  const long p = r - a.loc_off[f];
  a.D[r + a.u * (a.col0 + b)] = a.out[(long)b * a.stride + a.face_off[f] + p];
}

// xs[i] = x[map[i]]
__global__ void __launch_bounds__(256) gather_kernel(const double *__restrict__ x,
                                                     const long *__restrict__ map, long n,
                                                     double *__restrict__ xs) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) xs[i] = x[map[i]];
}

// part[c * u + i] = sum_{j in chunk c} D[i + u j] xs[j]; grid (ceil(u/256), nchunk)
__global__ void __launch_bounds__(256) gemv_part_kernel(const double *__restrict__ D, long u,
                                                        const double *__restrict__ xs, long chunk,
                                                        double *__restrict__ part) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  const long j0 = (long)blockIdx.y * chunk;
  const long j1 = min(j0 + chunk, u);
  if (i >= u) return;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  long j = j0;
  for (; j + 3 < j1; j += 4) {
    s0 = fma(D[i + u * j], xs[j], s0);
    s1 = fma(D[i + u * (j + 1)], xs[j + 1], s1);
    s2 = fma(D[i + u * (j + 2)], xs[j + 2], s2);
    s3 = fma(D[i + u * (j + 3)], xs[j + 3], s3);
  }
  for (; j < j1; ++j) s0 = fma(D[i + u * j], xs[j], s0);
  part[(long)blockIdx.y * u + i] = (s0 + s1) + (s2 + s3);
}

// ys[i] = sum_c part[c u + i] (c ascending)
__global__ void __launch_bounds__(256) gemv_sum_kernel(const double *__restrict__ part, long u,
                                                       int nchunk, double *__restrict__ ys) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= u) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[(long)c * u + i];
  ys[i] = s;
}

// y[g] = ys[pa[g]] + ys[pb[g]]: the two sides of every interface unknown
__global__ void __launch_bounds__(256) combine_kernel(const double *__restrict__ ys,
                                                      const long *__restrict__ pa,
                                                      const long *__restrict__ pb, long n,
                                                      double *__restrict__ y) {
  const long g = (long)blockIdx.x * 256 + threadIdx.x;
  if (g < n) y[g] = ys[pa[g]] + ys[pb[g]];
}

// Dot products of w with the k vectors V[:, j] (column stride n): block
// partials part[j * nb + blk] (fixed order), grid (nb, k)
__global__ void __launch_bounds__(256) dots_part_kernel(const double *__restrict__ V, long n,
                                                        const double *__restrict__ w,
                                                        double *__restrict__ part) {
  __shared__ double sh[256];
  const int j = blockIdx.y;
  const double *v = V + (long)j * n;
  double s = 0.0;
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256)
    s = fma(v[t], w[t], s);
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long)j * gridDim.x + blockIdx.x] = sh[0];
}

// out[j] = sum_b part[j nb + b]; one block per j
__global__ void __launch_bounds__(256) dots_sum_kernel(const double *__restrict__ part, int nb,
                                                       double *__restrict__ out) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) s += part[(long)blockIdx.x * nb + b];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// w -= sum_j V[:, j] h[j] (j = 0..k-1)
__global__ void __launch_bounds__(256) orth_kernel(const double *__restrict__ V, long n, int k,
                                                   const double *__restrict__ h,
                                                   double *__restrict__ w) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  double v = w[t];
  for (int j = 0; j < k; ++j) v = fma(-h[j], V[(long)j * n + t], v);
  w[t] = v;
}

// y = a x + b y (x may be y)
__global__ void __launch_bounds__(256) axpby_kernel(double a, const double *x, double b, double *y,
                                                    long n) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t < n) y[t] = fma(a, x[t], b * y[t]);
}

// x += sum_j V[:, j] c[j]
__global__ void __launch_bounds__(256) update_kernel(const double *__restrict__ V, long n, int k,
                                                     const double *__restrict__ c,
                                                     double *__restrict__ x) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  double v = x[t];
  for (int j = 0; j < k; ++j) v = fma(c[j], V[(long)j * n + t], v);
  x[t] = v;
}

// Face rows of the right-hand side (multidomain_solve :1317-1336): for each
// interface face f of the subdomain, the derivative along the face's normal
// at the face points of the full field u: -d on a low face, +d on a high
// face, written to ys[loc_off[f] + p].  drow: per axis row0 (n) then rown (n)
// of the full-line first-derivative operator (n = full points).
struct RhsFaceArgs {
  int nx, ny, nz;
  int nf;
  int face[6];
  long loc_off[7];
  const double *drow;  // x: 2 nx, y: 2 ny, z: 2 nz
};

__global__ void __launch_bounds__(256) rhs_face_kernel(const double *__restrict__ u,
                                                       double *__restrict__ ys, RhsFaceArgs a) {
  const long r = (long)blockIdx.x * 256 + threadIdx.x;
  if (r >= a.loc_off[a.nf]) return;
  int f = 0;
  while (f + 1 < a.nf && r >= a.loc_off[f + 1]) ++f;
  const long p = r - a.loc_off[f];
  const int face = a.face[f], ax = face / 2, hi = face % 2;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const long sy = nx, sz = (long)nx * ny;
  long base, s;
  int n;
  const double *row;
  if (ax == 0) {
    const int j = 1 + (int)(p % (ny - 2)), k = 1 + (int)(p / (ny - 2));
    base = j * sy + k * sz; s = 1; n = nx;
    row = a.drow + (hi ? nx : 0);
  } else if (ax == 1) {
    const int i = 1 + (int)(p % (nx - 2)), k = 1 + (int)(p / (nx - 2));
    base = i + k * sz; s = sy; n = ny;
    row = a.drow + 2 * nx + (hi ? ny : 0);
  } else {
    const int i = 1 + (int)(p % (nx - 2)), j = 1 + (int)(p / (nx - 2));
    base = i + j * sy; s = sz; n = nz;
    row = a.drow + 2 * nx + 2 * ny + (hi ? nz : 0);
  }
  double d = 0.0;
  for (int q = 0; q < n; ++q) d = fma(row[q], u[base + q * s], d);
  ys[r] = hi ? d : -d;
}

// Interface faces of the boundary arrays [bcx(n2,n3,2) | bcy(n1,n3,2) |
// bcz(n1,n2,2)]: copy from src, then each interface face set from x through
// map (or zero when x is null).
struct BcArgs {
  long fsz[3];          // face sizes n2 n3, n1 n3, n1 n2
  int nf;
  int face[6];
  long loc_off[7];
  const long *map;      // local unknown -> global (this subdomain's slice)
};

__global__ void __launch_bounds__(256) bc_iface_kernel(double *__restrict__ bc,
                                                       const double *__restrict__ x, BcArgs a) {
  const long r = (long)blockIdx.x * 256 + threadIdx.x;
  if (r >= a.loc_off[a.nf]) return;
  int f = 0;
  while (f + 1 < a.nf && r >= a.loc_off[f + 1]) ++f;
  const long p = r - a.loc_off[f];
  const int face = a.face[f], ax = face / 2, hi = face % 2;
  long off = 0;
  for (int q = 0; q < ax; ++q) off += 2 * a.fsz[q];
  bc[off + hi * a.fsz[ax] + p] = x ? x[a.map[r]] : 0.0;
}

inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace md
}  // namespace mf

struct mf_md {
  struct Sub {
    int idx[3];
    int nx, ny, nz;
    mf_solver3d *sc = nullptr;
    mf_der *d[3] = {nullptr, nullptr, nullptr};
    int nf = 0;
    int face[6];
    int iface[6];
    long loc_off[7];
    long u = 0;            // local unknowns
    long ys_off = 0;       // offset in the concatenated per-subdomain vectors
    size_t D_off = 0;      // offset of D_s in the block buffer
    long stride = 0;       // doubles per boundary-array set
    double *bc = nullptr;  // device boundary arrays (stride doubles)
    double *drow = nullptr;  // device derivative rows (2 (nx + ny + nz))
  };
  int nd[3] = {0, 0, 0};
  int bct[6];
  int periods[3] = {0, 0, 0};
  double sigma = 0.0;
  // pressure singular method (command_line.f90:27-29): 0 none, 1 constant
  // null space (md_rhs_nullspace + KSPSetNullSpace), 2 first right-hand-side
  // entry zeroed (md_vector_zero_lastpoint)
  int psm = 0;
  double *ones = nullptr;  // ninf ones, for psm = 1
  mf::InflStats asm_stats;  // the device assembly of the blocks
  std::vector<Sub> subs;
  std::vector<long> ioff;  // interface offsets (size it+1)
  long ninf = 0, nloc = 0;
  double *D = nullptr;     // all blocks
  size_t D_doubles = 0;
  long *map = nullptr;     // concatenated local -> global
  long *pa = nullptr, *pb = nullptr;
  double *ws = nullptr;    // GMRES and GEMV workspace
  size_t ws_doubles = 0;
  int restart = 30;
  // initial-guess history (md_solve with inf_sol, class_md.f90:389-400,
  // 419-424): the last three interface solutions, oldest first
  bool guess = false;
  double *hist = nullptr;  // 3 x ninf slots
  int hist_head = 0;       // slot of the oldest solution; the others follow cyclically
  double *slot(int age) const { return hist + (long)((hist_head + age) % 3) * ninf; }
  // identity of the layout for the matrix cache
  std::vector<int> npts;
  std::vector<double> grids;
  int so = 6;
  ~mf_md() {
    for (auto &s : subs) {
      if (s.sc) mf_solver3d_destroy(s.sc);
      for (int a = 0; a < 3; ++a)
        if (s.d[a]) mf_der_destroy(s.d[a]);
      if (s.bc) (void)hipFree(s.bc);
      if (s.drow) (void)hipFree(s.drow);
    }
    if (D) (void)hipFree(D);
    if (map) (void)hipFree(map);
    if (pa) (void)hipFree(pa);
    if (pb) (void)hipFree(pb);
    if (ws) (void)hipFree(ws);
    if (hist) (void)hipFree(hist);
    if (ones) (void)hipFree(ones);
  }
};

namespace {

#define MDCHK(call)                                                                  \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return mf::fail(MF_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define MDRC(expr)                \
  do {                            \
    int rc_ = (expr);             \
    if (rc_ != MF_OK) return rc_; \
  } while (0)

// workspace layout (doubles): xs / ys (nloc each), GEMV partials, the Krylov
// basis (ninf x (m+1)), w, x, r, dot partials, small vectors
struct Ws {
  double *xs, *ys, *part, *V, *w, *x, *r, *dpart, *hcol;
  long nchunk_max;
};

Ws carve(mf_md *h) {
  Ws w;
  double *p = h->ws;
  const long m1 = h->restart + 1;
  w.xs = p; p += h->nloc;
  w.ys = p; p += h->nloc;
  long pmax = 0;
  for (auto &s : h->subs) pmax = std::max(pmax, ((s.u + 511) / 512) * s.u);
  w.part = p; p += pmax;
  w.V = p; p += h->ninf * m1;
  w.w = p; p += h->ninf;
  w.x = p; p += h->ninf;
  w.r = p; p += h->ninf;
  w.dpart = p; p += 256 * m1;
  w.hcol = p; p += 2 * m1;
  w.nchunk_max = 0;
  return w;
}

size_t ws_size(const mf_md *h) {
  const long m1 = h->restart + 1;
  long pmax = 0;
  for (auto &s : h->subs) pmax = std::max(pmax, ((s.u + 511) / 512) * s.u);
  return 2 * (size_t)h->nloc + pmax + (size_t)h->ninf * m1 + 3 * (size_t)h->ninf + 256 * m1 + 2 * m1;
}

// y = M x (device vectors of ninf)
int matvec(mf_md *h, const Ws &w, const double *x, double *y, hipStream_t st) {
  for (auto &s : h->subs) {
    if (!s.u) continue;
    hipLaunchKernelGGL(mf::md::gather_kernel, dim3(mf::md::nblk(s.u)), dim3(256), 0, st, x,
                       (const long *)(h->map + s.ys_off), s.u, w.xs + s.ys_off);
    const long chunk = 512;
    const int nchunk = (int)((s.u + chunk - 1) / chunk);
    hipLaunchKernelGGL(mf::md::gemv_part_kernel, dim3(mf::md::nblk(s.u), nchunk), dim3(256), 0, st,
                       (const double *)(h->D + s.D_off), s.u, (const double *)(w.xs + s.ys_off),
                       chunk, w.part);
    hipLaunchKernelGGL(mf::md::gemv_sum_kernel, dim3(mf::md::nblk(s.u)), dim3(256), 0, st,
                       (const double *)w.part, s.u, nchunk, w.ys + s.ys_off);
  }
  hipLaunchKernelGGL(mf::md::combine_kernel, dim3(mf::md::nblk(h->ninf)), dim3(256), 0, st,
                     (const double *)w.ys, (const long *)h->pa, (const long *)h->pb, h->ninf, y);
  MDCHK(hipGetLastError());
  return MF_OK;
}

// out[0..k-1] = V[:, 0..k-1] . v (host), fixed order
int dots(mf_md *h, const Ws &w, const double *V, int k, const double *v, double *out,
         hipStream_t st) {
  const int nb = 256;
  hipLaunchKernelGGL(mf::md::dots_part_kernel, dim3(nb, k), dim3(256), 0, st, V, h->ninf, v, w.dpart);
  hipLaunchKernelGGL(mf::md::dots_sum_kernel, dim3(k), dim3(256), 0, st, (const double *)w.dpart, nb,
                     w.hcol);
  MDCHK(hipGetLastError());
  MDCHK(hipMemcpyAsync(out, w.hcol, k * sizeof(double), hipMemcpyDeviceToHost, st));
  MDCHK(hipStreamSynchronize(st));
  return MF_OK;
}

// v -= mean(v) (the constant null space of psm = 1)
int remove_constant(mf_md *h, const Ws &w, double *v, hipStream_t st) {
  double sum = 0.0;
  MDRC(dots(h, w, h->ones, 1, v, &sum, st));
  hipLaunchKernelGGL(mf::md::axpby_kernel, dim3(mf::md::nblk(h->ninf)), dim3(256), 0, st,
                     -sum / (double)h->ninf, (const double *)h->ones, 1.0, v, h->ninf);
  MDCHK(hipGetLastError());
  return MF_OK;
}

// Solve all subdomains with their boundary arrays (interface faces from x,
// or zero when x is null) into u[s].
int solve_subs(mf_md *h, const double *const *g, double *const *u, const double *const *bc,
               const double *x, hipStream_t st) {
  for (size_t i = 0; i < h->subs.size(); ++i) {
    auto &s = h->subs[i];
    MDCHK(hipMemcpyAsync(s.bc, bc[i], s.stride * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (s.u) {
      mf::md::BcArgs a;
      std::memset(&a, 0, sizeof(a));
      const long n1 = s.nx - 2, n2 = s.ny - 2, n3 = s.nz - 2;
      a.fsz[0] = n2 * n3; a.fsz[1] = n1 * n3; a.fsz[2] = n1 * n2;
      a.nf = s.nf;
      for (int f = 0; f < s.nf; ++f) a.face[f] = s.face[f];
      for (int f = 0; f <= s.nf; ++f) a.loc_off[f] = s.loc_off[f];
      a.map = h->map + s.ys_off;
      hipLaunchKernelGGL(mf::md::bc_iface_kernel, dim3(mf::md::nblk(s.u)), dim3(256), 0, st, s.bc, x, a);
      MDCHK(hipGetLastError());
    }
    const long n1 = s.nx - 2, n2 = s.ny - 2, n3 = s.nz - 2;
    const double *bx = s.bc, *by = bx + 2 * n2 * n3, *bz = by + 2 * n1 * n3;
    MDRC(mf_solver3d_solve(s.sc, g[i], h->sigma, u[i], bx, by, bz, MF_FP64, 0.0, 0, nullptr, st));
  }
  return MF_OK;
}

}  // namespace

namespace {
int md_create(const int nd[3], const int *npts, const double *grids, const int bct[6],
              const int periods[3], int so, double sigma, bool assemble, mf_md **out) {
  if (!nd || !npts || !grids || !bct || !out) return mf::fail(MF_ERR_ARG, "mf_md_create: null pointer");
  *out = nullptr;
  for (int a = 0; a < 3; ++a)
    if (nd[a] < 1) return mf::fail(MF_ERR_ARG, "mf_md_create: nd must be >= 1");
  for (int f = 0; f < 6; ++f)
    if (bct[f] != 1 && bct[f] != 2) return mf::fail(MF_ERR_ARG, "mf_md_create: bct must be 1 or 2");
  if (!std::isfinite(sigma)) return mf::fail(MF_ERR_ARG, "mf_md_create: sigma is not finite");
  // slab grids per axis
  std::vector<const double *> slab[3];
  std::vector<int> cnt[3];
  {
    long q = 0, gofs = 0;
    for (int a = 0; a < 3; ++a)
      for (int k = 0; k < nd[a]; ++k) {
        const int n = npts[q++];
        if (n < 10) return mf::fail(MF_ERR_ARG, "mf_md_create: every slab needs >= 10 points");
        slab[a].push_back(grids + gofs);
        cnt[a].push_back(n);
        gofs += n;
      }
    for (int a = 0; a < 3; ++a)
      for (int k = 0; k + 1 < nd[a]; ++k)
        if (slab[a][k][cnt[a][k] - 1] != slab[a][k + 1][0])
          return mf::fail(MF_ERR_ARG, "mf_md_create: consecutive slabs must share their end points");
  }
  mf_md *h = new mf_md();
  h->so = so;
  {
    long q = 0, gofs = 0;
    for (int a = 0; a < 3; ++a)
      for (int k = 0; k < nd[a]; ++k) {
        h->npts.push_back(npts[q]);
        gofs += npts[q++];
      }
    h->grids.assign(grids, grids + gofs);
  }
  for (int a = 0; a < 3; ++a) h->nd[a] = nd[a];
  for (int a = 0; a < 3; ++a) h->periods[a] = (periods && periods[a]) ? 1 : 0;
  for (int f = 0; f < 6; ++f) h->bct[f] = bct[f];
  h->sigma = sigma;
  auto bail = [&](int rc) { delete h; return rc; };
  // interfaces (class_md.f90:1525-1559): i (x) outermost, k fastest; x, y, z
  // interface of each subdomain in turn, the high face of the last
  // subdomain of a periodic axis included (it joins the first one's low face)
  struct Itf { int ax, i, j, k; long size; };
  std::vector<Itf> itfs;
  auto fsize = [&](int ax, int i, int j, int k) -> long {
    const long n1 = cnt[0][i] - 2, n2 = cnt[1][j] - 2, n3 = cnt[2][k] - 2;
    return ax == 0 ? n2 * n3 : (ax == 1 ? n1 * n3 : n1 * n2);
  };
  const int *per = h->periods;
  for (int i = 0; i < nd[0]; ++i)
    for (int j = 0; j < nd[1]; ++j)
      for (int k = 0; k < nd[2]; ++k) {
        if (i < nd[0] - 1 || per[0]) itfs.push_back({0, i, j, k, fsize(0, i, j, k)});
        if (j < nd[1] - 1 || per[1]) itfs.push_back({1, i, j, k, fsize(1, i, j, k)});
        if (k < nd[2] - 1 || per[2]) itfs.push_back({2, i, j, k, fsize(2, i, j, k)});
      }
  h->ioff.assign(itfs.size() + 1, 0);
  for (size_t m = 0; m < itfs.size(); ++m) h->ioff[m + 1] = h->ioff[m] + itfs[m].size;
  h->ninf = h->ioff.back();
  auto find_itf = [&](int ax, int i, int j, int k) -> int {
    for (size_t m = 0; m < itfs.size(); ++m)
      if (itfs[m].ax == ax && itfs[m].i == i && itfs[m].j == j && itfs[m].k == k) return (int)m;
    return -1;
  };
  // subdomains, x fastest; local unknowns = their interface faces, ascending
  std::vector<long> hmap;
  std::vector<std::vector<long>> owners(h->ninf);
  for (int c = 0; c < nd[2]; ++c)
    for (int b = 0; b < nd[1]; ++b)
      for (int a = 0; a < nd[0]; ++a) {
        mf_md::Sub s;
        s.idx[0] = a; s.idx[1] = b; s.idx[2] = c;
        s.nx = cnt[0][a]; s.ny = cnt[1][b]; s.nz = cnt[2][c];
        int sbct[6];
        for (int f = 0; f < 6; ++f) sbct[f] = bct[f];
        const int id3[3] = {a, b, c};
        s.nf = 0;
        s.loc_off[0] = 0;
        for (int f = 0; f < 6; ++f) {
          const int ax = f / 2, hi = f % 2;
          int m = -1;
          if (!hi && (id3[ax] > 0 || per[ax])) {
            int lo[3] = {a, b, c};
            lo[ax] = (lo[ax] + nd[ax] - 1) % nd[ax];
            m = find_itf(ax, lo[0], lo[1], lo[2]);
          } else if (hi && (id3[ax] < nd[ax] - 1 || per[ax])) {
            m = find_itf(ax, a, b, c);
          }
          if (m < 0) continue;
          sbct[f] = 1;  // interfaces carry Dirichlet data
          s.face[s.nf] = f;
          s.iface[s.nf] = m;
          s.loc_off[s.nf + 1] = s.loc_off[s.nf] + itfs[m].size;
          s.nf += 1;
        }
        s.u = s.loc_off[s.nf];
        s.ys_off = h->nloc;
        for (int f = 0; f < s.nf; ++f)
          for (long p = 0; p < s.loc_off[f + 1] - s.loc_off[f]; ++p) {
            const long gidx = h->ioff[s.iface[f]] + p;
            owners[gidx].push_back(h->nloc + s.loc_off[f] + p);
            hmap.push_back(gidx);
          }
        h->nloc += s.u;
        s.D_off = h->D_doubles;
        h->D_doubles += (size_t)s.u * s.u;
        const long n1 = s.nx - 2, n2 = s.ny - 2, n3 = s.nz - 2;
        s.stride = 2 * (n2 * n3 + n1 * n3 + n1 * n2);
        int rc = mf_solver3d_create(slab[0][a], s.nx, slab[1][b], s.ny, slab[2][c], s.nz, sbct, so, &s.sc);
        const double *g3[3] = {slab[0][a], slab[1][b], slab[2][c]};
        const int n3v[3] = {s.nx, s.ny, s.nz};
        for (int q = 0; q < 3 && rc == MF_OK; ++q) rc = mf_der_create(g3[q], n3v[q], 1, so, &s.d[q]);
        h->subs.push_back(s);
        if (rc != MF_OK) return bail(rc);
      }
  for (long gi = 0; gi < h->ninf; ++gi)
    if (owners[gi].size() != 2) return bail(mf::fail(MF_ERR_SETUP, "mf_md_create: interface bookkeeping"));
  std::vector<long> hpa(h->ninf), hpb(h->ninf);
  for (long gi = 0; gi < h->ninf; ++gi) { hpa[gi] = owners[gi][0]; hpb[gi] = owners[gi][1]; }
  // device allocations
  hipError_t e = hipSuccess;
  if (h->D_doubles) e = hipMalloc(&h->D, h->D_doubles * sizeof(double));
  if (e == hipSuccess && h->nloc) e = hipMalloc(&h->map, h->nloc * sizeof(long));
  if (e == hipSuccess && h->ninf) e = hipMalloc(&h->pa, h->ninf * sizeof(long));
  if (e == hipSuccess && h->ninf) e = hipMalloc(&h->pb, h->ninf * sizeof(long));
  h->ws_doubles = ws_size(h);
  if (e == hipSuccess) e = hipMalloc(&h->ws, h->ws_doubles * sizeof(double));
  if (e == hipSuccess && h->nloc) e = hipMemcpy(h->map, hmap.data(), h->nloc * sizeof(long), hipMemcpyHostToDevice);
  if (e == hipSuccess && h->ninf) e = hipMemcpy(h->pa, hpa.data(), h->ninf * sizeof(long), hipMemcpyHostToDevice);
  if (e == hipSuccess && h->ninf) e = hipMemcpy(h->pb, hpb.data(), h->ninf * sizeof(long), hipMemcpyHostToDevice);
  for (auto &s : h->subs) {
    if (e != hipSuccess) break;
    e = hipMalloc(&s.bc, s.stride * sizeof(double));
    // full-line first-derivative rows for the right-hand side faces
    std::vector<double> drow;
    for (int q = 0; q < 3 && e == hipSuccess; ++q) {
      const double *dl = nullptr, *dr = nullptr;
      int n = 0;
      if (!mf::der_host_tables(s.d[q], 1, 0, &dl, &dr, &n)) return bail(mf::fail(MF_ERR_SETUP, "mf_md_create: derivative tables"));
      std::vector<double> r0(n), rn(n);
      mf::derivative_rows(n, dl, dr, r0.data(), rn.data());
      drow.insert(drow.end(), r0.begin(), r0.end());
      drow.insert(drow.end(), rn.begin(), rn.end());
    }
    if (e == hipSuccess) e = hipMalloc(&s.drow, drow.size() * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(s.drow, drow.data(), drow.size() * sizeof(double), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) return bail(mf::fail(MF_ERR_DEVICE, std::string("mf_md_create: ") + hipGetErrorString(e)));
  if (e == hipSuccess && h->ninf) e = hipMalloc(&h->hist, 3 * (size_t)h->ninf * sizeof(double));
  if (e == hipSuccess && h->ninf) e = hipMemset(h->hist, 0, 3 * (size_t)h->ninf * sizeof(double));
  if (e != hipSuccess) return bail(mf::fail(MF_ERR_DEVICE, std::string("mf_md_create: ") + hipGetErrorString(e)));
  if (!assemble) {  // the caller fills the blocks (matrix cache)
    *out = h;
    return MF_OK;
  }
  // influence blocks by spectral contraction (one persistent grouped launch,
  // captured in a HIP graph), or, with MF_INFL_IMPULSES=1, from batched
  // unit-impulse solves as the reference forms them
  const char *imp_env = std::getenv("MF_INFL_IMPULSES");
  if (!(imp_env && imp_env[0] == '1')) {
    std::vector<mf::InflSub> is;
    for (auto &s : h->subs) {
      if (!s.u) continue;
      mf::InflSub q;
      q.sc = s.sc;
      for (int a = 0; a < 3; ++a) q.d[a] = s.d[a];
      q.nf = s.nf;
      for (int f = 0; f < s.nf; ++f) q.face[f] = s.face[f];
      for (int f = 0; f <= s.nf; ++f) q.loc_off[f] = s.loc_off[f];
      q.u = s.u;
      q.D = h->D + s.D_off;
      is.push_back(q);
    }
    const char *g_env = std::getenv("MF_INFL_GRAPH");
    const bool graph = !(g_env && g_env[0] == '0');
    const int rc = mf::influence_assemble(is, sigma, graph, nullptr, &h->asm_stats);
    if (rc != MF_OK) return bail(rc);
    *out = h;
    return MF_OK;
  }
  {
    const long chunk = 1024;
    long maxstride = 0;
    for (auto &s : h->subs) maxstride = std::max(maxstride, s.stride);
    double *out = nullptr;
    if (h->nloc) {
      e = hipMalloc(&out, (size_t)chunk * maxstride * sizeof(double));
      if (e != hipSuccess) return bail(mf::fail(MF_ERR_DEVICE, std::string("mf_md_create: ") + hipGetErrorString(e)));
    }
    int rc = MF_OK;
    for (auto &s : h->subs) {
      if (!s.u || rc != MF_OK) continue;
      mf::md::ScatterArgs sa;
      std::memset(&sa, 0, sizeof(sa));
      sa.D = h->D + s.D_off;
      sa.u = s.u;
      sa.out = out;
      sa.stride = s.stride;
      sa.nf = s.nf;
      const long n1 = s.nx - 2, n2 = s.ny - 2, n3 = s.nz - 2;
      const long fsz[3] = {n2 * n3, n1 * n3, n1 * n2};
      for (int f = 0; f < s.nf; ++f) {
        const int ax = s.face[f] / 2, hi = s.face[f] % 2;
        long off = 0;
        for (int q = 0; q < ax; ++q) off += 2 * fsz[q];
        sa.face_off[f] = off + hi * fsz[ax];
      }
      for (int f = 0; f <= s.nf; ++f) sa.loc_off[f] = s.loc_off[f];
      for (int f = 0; f < s.nf && rc == MF_OK; ++f) {
        const long np = s.loc_off[f + 1] - s.loc_off[f];
        for (long first = 0; first < np && rc == MF_OK; first += chunk) {
          const int c = (int)std::min(chunk, np - first);
          rc = mf_influence_impulses(s.sc, s.d[0], s.d[1], s.d[2], sigma, s.face[f], first, c, out, nullptr);
          if (rc != MF_OK) break;
          sa.col0 = s.loc_off[f] + first;
          hipLaunchKernelGGL(mf::md::scatter_kernel, dim3(mf::md::nblk(s.u), c), dim3(256), 0, nullptr, sa, c);
          if (hipGetLastError() != hipSuccess) rc = mf::fail(MF_ERR_DEVICE, "mf_md_create: scatter launch");
        }
      }
    }
    if (out) {
      (void)hipDeviceSynchronize();
      (void)hipFree(out);
    }
    if (rc != MF_OK) return bail(rc);
  }
  *out = h;
  return MF_OK;
}
}  // namespace

extern "C" int mf_md_create(const int nd[3], const int *npts, const double *grids, const int bct[6],
                            int so, double sigma, mf_md **out) {
  return md_create(nd, npts, grids, bct, nullptr, so, sigma, true, out);
}

extern "C" int mf_md_assembly_stats(const mf_md *h, double *device_ms, double *gemm_flop,
                                    double *block_bytes) {
  if (!h) return mf::fail(MF_ERR_ARG, "mf_md_assembly_stats: null handle");
  if (device_ms) *device_ms = h->asm_stats.ms;
  if (gemm_flop) *gemm_flop = h->asm_stats.gemm_flop;
  if (block_bytes) *block_bytes = h->asm_stats.block_bytes;
  return MF_OK;
}

extern "C" int mf_md_set_null_space(mf_md *h, int psm) {
  if (!h || psm < 0 || psm > 2) return mf::fail(MF_ERR_ARG, "mf_md_set_null_space: psm must be 0, 1 or 2");
  if (psm == 1 && !h->ones && h->ninf) {
    std::vector<double> one(h->ninf, 1.0);
    MDCHK(hipMalloc(&h->ones, h->ninf * sizeof(double)));
    MDCHK(hipMemcpy(h->ones, one.data(), h->ninf * sizeof(double), hipMemcpyHostToDevice));
  }
  h->psm = psm;
  return MF_OK;
}

extern "C" int mf_md_destroy(mf_md *h) {
  if (!h) return MF_OK;
  (void)hipDeviceSynchronize();
  delete h;
  return MF_OK;
}

extern "C" int mf_md_info(const mf_md *h, int *nsub, long long *ninf, long long *matrix_bytes) {
  if (!h) return mf::fail(MF_ERR_ARG, "mf_md_info: null handle");
  if (nsub) *nsub = (int)h->subs.size();
  if (ninf) *ninf = h->ninf;
  if (matrix_bytes) *matrix_bytes = (long long)(h->D_doubles * sizeof(double));
  return MF_OK;
}

extern "C" int mf_md_matvec(mf_md *h, const double *x, double *y, void *stream) {
  if (!h || !x || !y) return mf::fail(MF_ERR_ARG, "mf_md_matvec: null pointer");
  Ws w = carve(h);
  return matvec(h, w, x, y, (hipStream_t)stream);
}

extern "C" int mf_md_solve(mf_md *h, const double *const *g, double *const *u,
                           const double *const *bc, double tol, int max_iter, double *x_iface,
                           mf_md_report *rep, void *stream) {
  if (!h || !g || !u || !bc) return mf::fail(MF_ERR_ARG, "mf_md_solve: null pointer");
  for (size_t i = 0; i < h->subs.size(); ++i)
    if (!g[i] || !u[i] || !bc[i]) return mf::fail(MF_ERR_ARG, "mf_md_solve: null subdomain pointer");
  if (!(tol > 0.0)) tol = 1e-10;
  if (max_iter <= 0) max_iter = 1000;
  hipStream_t st = (hipStream_t)stream;
  if (rep) { rep->iterations = 0; rep->residual = 0.0; }
  Ws w = carve(h);
  const long n = h->ninf;
  // 1. zero interface values: rhs = -d/dn (low), +d/dn (high), both sides
  MDRC(solve_subs(h, g, u, bc, nullptr, st));
  if (n == 0) return MF_OK;
  for (auto &s : h->subs) {
    if (!s.u) continue;
    mf::md::RhsFaceArgs ra;
    std::memset(&ra, 0, sizeof(ra));
    ra.nx = s.nx; ra.ny = s.ny; ra.nz = s.nz;
    ra.nf = s.nf;
    for (int f = 0; f < s.nf; ++f) ra.face[f] = s.face[f];
    for (int f = 0; f <= s.nf; ++f) ra.loc_off[f] = s.loc_off[f];
    ra.drow = s.drow;
    const size_t i = &s - h->subs.data();
    hipLaunchKernelGGL(mf::md::rhs_face_kernel, dim3(mf::md::nblk(s.u)), dim3(256), 0, st,
                       (const double *)u[i], w.ys + s.ys_off, ra);
  }
  hipLaunchKernelGGL(mf::md::combine_kernel, dim3(mf::md::nblk(n)), dim3(256), 0, st,
                     (const double *)w.ys, (const long *)h->pa, (const long *)h->pb, n, w.r);
  MDCHK(hipGetLastError());
  // singular systems: the constant part of the right-hand side removed
  // (psm = 1, md_rhs_nullspace class_md.f90:347-357) or its first entry
  // zeroed (psm = 2, md_vector_zero_lastpoint :251-267)
  const bool nullsp = h->psm == 1;
  if (nullsp) MDRC(remove_constant(h, w, w.r, st));
  if (h->psm == 2) {
    const double zero = 0.0;
    MDCHK(hipMemcpyAsync(w.r, &zero, sizeof(double), hipMemcpyHostToDevice, st));
    MDCHK(hipStreamSynchronize(st));
  }
  // 2. GMRES(m) on M x = rhs, from x0 = 0 or, with the guess history on,
  //    x0 = s1 - 3 s2 + 3 s3 (class_md.f90:395-397)
  if (h->guess) {
    const double *s1 = h->slot(0), *s2 = h->slot(1), *s3 = h->slot(2);
    MDCHK(hipMemcpyAsync(w.x, s3, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(mf::md::axpby_kernel, dim3(mf::md::nblk(n)), dim3(256), 0, st, -3.0, s2, 3.0,
                       w.x, n);
    hipLaunchKernelGGL(mf::md::axpby_kernel, dim3(mf::md::nblk(n)), dim3(256), 0, st, 1.0, s1, 1.0,
                       w.x, n);
  } else {
    MDCHK(hipMemsetAsync(w.x, 0, n * sizeof(double), st));
  }
  double bnorm2 = 0.0;
  MDRC(dots(h, w, w.r, 1, w.r, &bnorm2, st));
  const double bnorm = std::sqrt(bnorm2);
  const int m = h->restart;
  double *b = w.r;  // the right-hand side, kept
  double *V = w.V;
  std::vector<double> H((m + 1) * m), cs(m), sn(m), gv(m + 1), hc(m + 1), hc2(m + 1);
  int iters = 0;
  double rel = bnorm > 0.0 ? 1.0 : 0.0;
  bool done = bnorm == 0.0;
  while (!done && iters < max_iter) {
    // r0 = b - M x -> V[:,0]
    MDRC(matvec(h, w, w.x, V, st));
    hipLaunchKernelGGL(mf::md::axpby_kernel, dim3(mf::md::nblk(n)), dim3(256), 0, st, 1.0,
                       (const double *)b, -1.0, V, n);
    // with the constant null space every Krylov vector is kept orthogonal
    // to the constant (KSPSetNullSpace, md_solve_init :318-328), so x keeps
    // the mean of its start
    if (nullsp) MDRC(remove_constant(h, w, V, st));
    double beta2 = 0.0;
    MDRC(dots(h, w, V, 1, V, &beta2, st));
    const double beta = std::sqrt(beta2);
    rel = beta / bnorm;
    if (rel <= tol) break;
    hipLaunchKernelGGL(mf::md::axpby_kernel, dim3(mf::md::nblk(n)), dim3(256), 0, st, 0.0,
                       (const double *)V, 1.0 / beta, V, n);
    std::fill(gv.begin(), gv.end(), 0.0);
    gv[0] = beta;
    int k = 0;
    for (; k < m && iters < max_iter; ++k) {
      double *vk1 = V + (long)(k + 1) * n;
      MDRC(matvec(h, w, V + (long)k * n, vk1, st));
      if (nullsp) MDRC(remove_constant(h, w, vk1, st));
      // classical Gram-Schmidt, applied twice
      for (int pass = 0; pass < 2; ++pass) {
        double *hh = pass ? hc2.data() : hc.data();
        MDRC(dots(h, w, V, k + 1, vk1, hh, st));
        MDCHK(hipMemcpyAsync(w.hcol, hh, (k + 1) * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(mf::md::orth_kernel, dim3(mf::md::nblk(n)), dim3(256), 0, st,
                           (const double *)V, n, k + 1, (const double *)w.hcol, vk1);
      }
      for (int j = 0; j <= k; ++j) hc[j] += hc2[j];
      double hn2 = 0.0;
      MDRC(dots(h, w, vk1, 1, vk1, &hn2, st));
      const double hn = std::sqrt(hn2);
      ++iters;
      // Givens rotations on column k
      for (int j = 0; j < k; ++j) {
        const double t = cs[j] * hc[j] + sn[j] * hc[j + 1];
        hc[j + 1] = -sn[j] * hc[j] + cs[j] * hc[j + 1];
        hc[j] = t;
      }
      const double a = hc[k], bb = hn;
      const double rr = std::hypot(a, bb);
      cs[k] = rr > 0.0 ? a / rr : 1.0;
      sn[k] = rr > 0.0 ? bb / rr : 0.0;
      hc[k] = rr;
      gv[k + 1] = -sn[k] * gv[k];
      gv[k] = cs[k] * gv[k];
      for (int j = 0; j <= k; ++j) H[j + (size_t)(m + 1) * k] = hc[j];
      rel = std::fabs(gv[k + 1]) / bnorm;
      if (hn > 0.0)
        hipLaunchKernelGGL(mf::md::axpby_kernel, dim3(mf::md::nblk(n)), dim3(256), 0, st, 0.0,
                           (const double *)vk1, 1.0 / hn, vk1, n);
      if (rel <= tol || hn == 0.0) { ++k; break; }
    }
    // x += V y, H y = g (upper triangular k x k)
    std::vector<double> yv(k, 0.0);
    for (int i = k - 1; i >= 0; --i) {
      double t = gv[i];
      for (int j = i + 1; j < k; ++j) t -= H[i + (size_t)(m + 1) * j] * yv[j];
      yv[i] = t / H[i + (size_t)(m + 1) * i];
    }
    if (k > 0) {
      MDCHK(hipMemcpyAsync(w.hcol, yv.data(), k * sizeof(double), hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(mf::md::update_kernel, dim3(mf::md::nblk(n)), dim3(256), 0, st,
                         (const double *)V, n, k, (const double *)w.hcol, w.x);
      MDCHK(hipStreamSynchronize(st));  // yv goes out of scope
    }
    if (rel <= tol) done = true;
  }
  // true residual of the final x
  MDRC(matvec(h, w, w.x, w.w, st));
  hipLaunchKernelGGL(mf::md::axpby_kernel, dim3(mf::md::nblk(n)), dim3(256), 0, st, 1.0,
                     (const double *)b, -1.0, w.w, n);
  if (nullsp) MDRC(remove_constant(h, w, w.w, st));
  double rn2 = 0.0;
  MDRC(dots(h, w, w.w, 1, w.w, &rn2, st));
  const double res = bnorm > 0.0 ? std::sqrt(rn2) / bnorm : 0.0;
  if (rep) { rep->iterations = iters; rep->residual = res; }
  if (x_iface) MDCHK(hipMemcpyAsync(x_iface, w.x, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  if (h->guess) {  // sol_old(1) <- sol_old(2) <- sol_old(3) <- x (:419-424):
    // the oldest slot takes x and becomes the newest (no overlapping copies)
    MDCHK(hipMemcpyAsync(h->slot(0), w.x, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    h->hist_head = (h->hist_head + 1) % 3;
  }
  // 3. final subdomain solves with the interface values
  MDRC(solve_subs(h, g, u, bc, w.x, st));
  if (res > tol * 10.0) return mf::fail(MF_ERR_NOCONV, "mf_md_solve: GMRES did not reach tol");
  return MF_OK;
}

extern "C" int mf_md_set_guess(mf_md *h, int enable) {
  if (!h) return mf::fail(MF_ERR_ARG, "mf_md_set_guess: null handle");
  h->guess = enable != 0;
  h->hist_head = 0;
  if (h->hist) MDCHK(hipMemset(h->hist, 0, 3 * (size_t)h->ninf * sizeof(double)));
  return MF_OK;
}

// ---- influence-matrix cache (the role of md_influence_matrix_write/read,
// class_md.f90:950-985, which use PETSc's binary viewer): a file of its own
// format holding the layout's identity and the dense blocks.
namespace {
struct CacheHeader {
  char magic[8];        // "MFMDBLK2"
  int nd[3], bct[6], periods[3], so, nslab;
  double sigma;
  unsigned long long grid_hash, ninf, doubles;
};

unsigned long long fnv1a(const void *p, size_t n, unsigned long long h = 1469598103934665603ull) {
  const unsigned char *c = (const unsigned char *)p;
  for (size_t i = 0; i < n; ++i) { h ^= c[i]; h *= 1099511628211ull; }
  return h;
}

CacheHeader header_of(const mf_md *h) {
  CacheHeader c;
  std::memset(&c, 0, sizeof(c));
  std::memcpy(c.magic, "MFMDBLK2", 8);
  for (int a = 0; a < 3; ++a) c.nd[a] = h->nd[a];
  for (int a = 0; a < 3; ++a) c.periods[a] = h->periods[a];
  for (int f = 0; f < 6; ++f) c.bct[f] = h->bct[f];
  c.so = h->so;
  c.nslab = (int)h->npts.size();
  c.sigma = h->sigma;
  c.grid_hash = fnv1a(h->npts.data(), h->npts.size() * sizeof(int));
  c.grid_hash = fnv1a(h->grids.data(), h->grids.size() * sizeof(double), c.grid_hash);
  c.ninf = (unsigned long long)h->ninf;
  c.doubles = (unsigned long long)h->D_doubles;
  return c;
}

constexpr size_t kCacheChunk = 32u << 20;  // doubles per transfer
}  // namespace

extern "C" int mf_md_write_matrix(const mf_md *h, const char *path) {
  if (!h || !path) return mf::fail(MF_ERR_ARG, "mf_md_write_matrix: null pointer");
  FILE *f = std::fopen(path, "wb");
  if (!f) return mf::fail(MF_ERR_ARG, std::string("mf_md_write_matrix: cannot open ") + path);
  const CacheHeader c = header_of(h);
  bool ok = std::fwrite(&c, sizeof(c), 1, f) == 1;
  std::vector<double> buf;
  for (size_t o = 0; ok && o < h->D_doubles; o += kCacheChunk) {
    const size_t cnt = std::min(kCacheChunk, h->D_doubles - o);
    buf.resize(cnt);
    ok = hipMemcpy(buf.data(), h->D + o, cnt * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess &&
         std::fwrite(buf.data(), sizeof(double), cnt, f) == cnt;
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return mf::fail(MF_ERR_DEVICE, std::string("mf_md_write_matrix: write failed: ") + path);
  return MF_OK;
}

extern "C" int mf_md_create_ex(const int nd[3], const int *npts, const double *grids,
                               const int bct[6], const int periods[3], int so, double sigma,
                               const char *path, int *loaded, mf_md **out) {
  if (!out) return mf::fail(MF_ERR_ARG, "mf_md_create_ex: null pointer");
  if (loaded) *loaded = 0;
  if (!path) return md_create(nd, npts, grids, bct, periods, so, sigma, true, out);
  mf_md *h = nullptr;
  int rc = md_create(nd, npts, grids, bct, periods, so, sigma, false, &h);
  if (rc != MF_OK) return rc;
  bool hit = false;
  if (FILE *f = std::fopen(path, "rb")) {
    CacheHeader c;
    const CacheHeader want = header_of(h);
    if (std::fread(&c, sizeof(c), 1, f) == 1 && std::memcmp(&c, &want, sizeof(c)) == 0) {
      hit = true;
      std::vector<double> buf;
      for (size_t o = 0; hit && o < h->D_doubles; o += kCacheChunk) {
        const size_t cnt = std::min(kCacheChunk, h->D_doubles - o);
        buf.resize(cnt);
        hit = std::fread(buf.data(), sizeof(double), cnt, f) == cnt &&
              hipMemcpy(h->D + o, buf.data(), cnt * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
      }
    }
    std::fclose(f);
  }
  if (hit) {
    if (loaded) *loaded = 1;
    *out = h;
    return MF_OK;
  }
  // no usable cache: assemble and write it
  mf_md_destroy(h);
  h = nullptr;
  rc = md_create(nd, npts, grids, bct, periods, so, sigma, true, &h);
  if (rc != MF_OK) return rc;
  rc = mf_md_write_matrix(h, path);
  if (rc != MF_OK) {
    mf_md_destroy(h);
    return rc;
  }
  *out = h;
  return MF_OK;
}

extern "C" int mf_md_create_cached(const int nd[3], const int *npts, const double *grids,
                                   const int bct[6], int so, double sigma, const char *path,
                                   int *loaded, mf_md **out) {
  if (!path || !out) return mf::fail(MF_ERR_ARG, "mf_md_create_cached: null pointer");
  return mf_md_create_ex(nd, npts, grids, bct, nullptr, so, sigma, path, loaded, out);
}
