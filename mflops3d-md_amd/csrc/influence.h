// Unit-impulse responses for the influence (capacitance) matrix
// (md_influence_matrix_set, class_field.f90:1396-1510).
//
// For a unit Dirichlet value at one point of face F (g = 0, every other
// boundary value 0), the BC fold leaves a right-hand side that is nonzero on
// a single line normal to F: rhs = a (x) e_c1 (x) e_c2 with a = -(fold vector
// of F).  Its three forward transforms are therefore separable, and the
// spectral coefficients of a whole batch of impulses are written by one
// kernel,
//     S(p,q,r; b) = ahat_p * W_y(q, c1_b) * W_z(r, c2_b) / (lx_p + ly_q + lz_r + sigma)
// (axes permuted for y- and z-faces), after which the three backward
// transforms run as batched MFMA GEMMs and the face-derivative kernels
// (face_der_x_kernel, face_der_yz_kernel) form the reference's outputs: the inward-normal first derivative on all six faces
// (+d/dn at the low face, -d/dn at the high face), in the layout of
// type(boundary_condition) [bcx(n2,n3,2) | bcy(n1,n3,2) | bcz(n1,n2,2)].
#pragma once
#include <hip/hip_runtime.h>

namespace mf {

struct ImpulseArgs {
  int n1, n2, n3;
  int face;        // 0..5: x1 x2 y1 y2 z1 z2 (the Dirac's face)
  long first;      // first impulse index in the face's column-major order
  const double *ahat;  // W_s (-fold_s) for the source axis s
  const double *Wx, *Wyt, *Wzt;  // inv_vect_x, inv_vect_y^T, inv_vect_z^T
  const double *lx, *ly, *lz;
  double sigma;
};

// S for a batch of impulses: one thread per (p, q) point of a plane, running
// over r: consecutive threads write consecutive points of each plane (every
// lane busy, coalesced), and the r-dependent factors are block-uniform.
// grid: (ceil(n1 n2 / 256), 1, B), block 256
__global__ void __launch_bounds__(256) impulse_spectral_pq_kernel(double *__restrict__ S,
                                                                  ImpulseArgs a) {
  const long pq = (long)blockIdx.x * 256 + threadIdx.x;
  const long n12 = (long)a.n1 * a.n2;
  if (pq >= n12) return;
  const int p = (int)(pq % a.n1), q = (int)(pq / a.n1);
  const int b = blockIdx.z;
  const int s = a.face / 2;
  const long t = a.first + b;
  const int L1 = (s == 0) ? a.n2 : a.n1;
  const int c1 = (int)(t % L1), c2 = (int)(t / L1);
  double vxy;
  if (s == 0) vxy = a.ahat[p] * a.Wyt[c1 + (long)a.n2 * q];
  else if (s == 1) vxy = a.Wx[p + (long)a.n1 * c1] * a.ahat[q];
  else vxy = a.Wx[p + (long)a.n1 * c1] * a.Wyt[c2 + (long)a.n2 * q];
  const double lxy = a.lx[p] + a.ly[q];
  const double bxy = fmax(fabs(a.lx[p]), fabs(a.ly[q]));
  double *o = S + (long)b * n12 * a.n3 + pq;
  for (int r = 0; r < a.n3; ++r) {
    const double vz = (s == 2) ? a.ahat[r] : a.Wzt[c2 + (long)a.n3 * r];
    const double lzr = a.lz[r];
    const double big = fmax(fmax(fabs(a.sigma), bxy), fabs(lzr));
    o[n12 * r] = (big < 1e-11) ? 0.0 : (vxy * vz) / (lxy + lzr + a.sigma);
  }
}

struct FaceDerArgs {
  int n1, n2, n3;
  int face;          // source face of the impulses
  long first;
  const int *bct;    // 6 boundary codes (device)
  const double *neum;  // 3 x (2,7) extrapolation weights (device)
  const double *drow;  // per axis: row0 (n+2) then rown (n+2), axes x, y, z
  long out_stride;     // doubles per impulse in out
};

// y- and z-lines: one thread per line over the flattened transverse plane
// (t = i + n1 t2: consecutive threads read consecutive i, every lane busy),
// the line walked with four independent partial sums (loads in flight).
// grid: (ceil(n1 T2 / 256), 1, B), block 256; T2 = n3 (y-lines) or n2.
template <int AX>
__global__ void __launch_bounds__(256) face_der_yz_kernel(const double *__restrict__ U,
                                                          double *__restrict__ out, FaceDerArgs a) {
  static_assert(AX == 1 || AX == 2, "y or z lines");
  const int n1 = a.n1, n2 = a.n2, n3 = a.n3;
  const int T2 = (AX == 1) ? n3 : n2;
  const long tt = (long)blockIdx.x * 256 + threadIdx.x;
  if (tt >= (long)n1 * T2) return;
  const int t1 = (int)(tt % n1), t2 = (int)(tt / n1);
  const int b = blockIdx.z;
  const int n = (AX == 1) ? n2 : n3;
  const long base = (AX == 1) ? t1 + (long)n1 * n2 * t2 : t1 + (long)n1 * t2;
  const long s = (AX == 1) ? (long)n1 : (long)n1 * n2;
  const long N = (long)n1 * n2 * n3;
  const double *u = U + (long)b * N + base;
  // boundary values of the line: Dirichlet -> the impulse (1 at its point
  // of its face, else 0); Neumann -> extrapolated with zero derivative
  const long t = a.first + b;
  const int sa = a.face / 2, side = a.face % 2;
  const bool here = (sa == AX) && (t % n1 == t1) && (t / n1 == t2);
  const double *nm = a.neum + 14 * AX;
  double bv0, bv1;
  if (a.bct[2 * AX] == 1) {
    bv0 = (here && side == 0) ? 1.0 : 0.0;
  } else {
    double acc = 0.0;
    for (int m = 1; m < 7; ++m) acc -= nm[2 * m] * u[(long)(m - 1) * s];
    bv0 = acc * (1.0 / nm[0]);
  }
  if (a.bct[2 * AX + 1] == 1) {
    bv1 = (here && side == 1) ? 1.0 : 0.0;
  } else {
    double acc = 0.0;
    for (int m = 5; m >= 0; --m) acc -= nm[1 + 2 * m] * u[(long)(n - 6 + m) * s];
    bv1 = acc * (1.0 / nm[13]);
  }
  const int off = (AX == 1) ? 2 * (n1 + 2) : 2 * (n1 + 2) + 2 * (n2 + 2);
  const double *r0 = a.drow + off, *rn = r0 + (n + 2);
  double d0[4] = {0.0, 0.0, 0.0, 0.0}, dn[4] = {0.0, 0.0, 0.0, 0.0};
  int q = 0;
  for (; q + 4 <= n; q += 4) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double v = u[(long)(q + c) * s];
      d0[c] = fma(r0[q + c + 1], v, d0[c]);
      dn[c] = fma(rn[q + c + 1], v, dn[c]);
    }
  }
  for (; q < n; ++q) {
    const double v = u[(long)q * s];
    d0[0] = fma(r0[q + 1], v, d0[0]);
    dn[0] = fma(rn[q + 1], v, dn[0]);
  }
  double e0 = (d0[0] + d0[1]) + (d0[2] + d0[3]), en = (dn[0] + dn[1]) + (dn[2] + dn[3]);
  e0 = fma(r0[0], bv0, fma(r0[n + 1], bv1, e0));
  en = fma(rn[0], bv0, fma(rn[n + 1], bv1, en));
  double *o = out + (long)b * a.out_stride;
  const long fsz = (AX == 1) ? (long)n1 * n3 : (long)n1 * n2;
  const long foff = (AX == 1) ? 2L * n2 * n3 : 2L * n2 * n3 + 2L * n1 * n3;
  o[foff + tt] = e0;
  o[foff + fsz + tt] = -en;
}

// x-lines (contiguous): one wavefront per line, lanes along the line, so the
// loads coalesce (a thread per x-line would make a wave's threads read 64
// lines a row apart); the two end-row dot products are reduced across the
// wave.  The boundary values of the line, d/dn at its ends: as in
// face_der_yz_kernel.  grid: (ceil(n2/4), n3, B), block (64, 4).
__global__ void __launch_bounds__(256) face_der_x_kernel(const double *__restrict__ U,
                                                         double *__restrict__ out, FaceDerArgs a) {
  const int n1 = a.n1, n2 = a.n2, n3 = a.n3;
  const int lane = threadIdx.x;
  const int t1 = blockIdx.x * 4 + threadIdx.y;  // j
  const int t2 = blockIdx.y;                     // k
  const int b = blockIdx.z;
  if (t1 >= n2) return;  // wave-uniform
  const int n = n1;
  const long N = (long)n1 * n2 * n3;
  const double *u = U + (long)b * N + (long)n1 * (t1 + (long)n2 * t2);
  const long t = a.first + b;
  const int sa = a.face / 2, side = a.face % 2;
  const bool here = (sa == 0) && (t % n2 == t1) && (t / n2 == t2);
  const double *nm = a.neum;
  double bv0, bv1;
  if (a.bct[0] == 1) {
    bv0 = (here && side == 0) ? 1.0 : 0.0;
  } else {
    double acc = 0.0;
    for (int m = 1; m < 7; ++m) acc -= nm[2 * m] * u[m - 1];
    bv0 = acc * (1.0 / nm[0]);
  }
  if (a.bct[1] == 1) {
    bv1 = (here && side == 1) ? 1.0 : 0.0;
  } else {
    double acc = 0.0;
    for (int m = 5; m >= 0; --m) acc -= nm[1 + 2 * m] * u[n - 6 + m];
    bv1 = acc * (1.0 / nm[13]);
  }
  const double *r0 = a.drow, *rn = r0 + (n + 2);
  double d0 = 0.0, dn = 0.0;
  for (int q = lane; q < n; q += 64) {
    const double v = u[q];
    d0 = fma(r0[q + 1], v, d0);
    dn = fma(rn[q + 1], v, dn);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    d0 += __shfl_xor(d0, off, 64);
    dn += __shfl_xor(dn, off, 64);
  }
  if (lane == 0) {
    d0 = fma(r0[0], bv0, fma(r0[n + 1], bv1, d0));
    dn = fma(rn[0], bv0, fma(rn[n + 1], bv1, dn));
    double *o = out + (long)b * a.out_stride;
    const long fsz = (long)n2 * n3, idx = t1 + (long)n2 * t2;
    o[idx] = d0;
    o[fsz + idx] = -dn;
  }
}

}  // namespace mf
