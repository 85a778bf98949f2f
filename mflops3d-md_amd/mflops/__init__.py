"""mflops: MI355X-native compact derivatives and diagonalisation solver.

Python mirror of the reference's module API for this path, over the C ABI
(include/mflops/mf.h) of the HIP library.  Device memory and streams come from
PyTorch; every computation runs in libmflops.so.

Reference interface -> here (paths relative to the reference root):
  der_coeffs_init          src/class_derivatives/class_derivatives.f90:111 -> DerivativeCoefficients
  der, dder, der_s, dder_s src/class_derivatives/class_derivatives.f90:186-260 -> .der/.dder/.der_s/.dder_s
  derx .. dderz            src/class_field/class_field.f90:629-978 -> derx .. dderz
  solver_coeffs_init_3d    src/class_solver/class_solver_3d.f90:380 -> SolverCoeffs3D
  solver_3d                src/class_field/class_field.f90:1072 -> solver_3d
  type(boundary_condition) src/class_field/class_field.f90:25-30 -> BoundaryCondition

Fields are FP64 tensors of shape (nx, ny, nz) with Fortran strides
(1, nx, nx*ny), i.e. the reference's column-major f(nx,ny,nz); use
``fortran_empty``/``to_device``/``to_host`` to make and move them.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import (MF_FP64, MF_MIXED, MF_MIXED_BF16X3, MF_MIXED_FP16X2, MFError,  # noqa: F401
                   SolveReport, check, load)

__all__ = [
    "load", "MFError", "MF_FP64", "MF_MIXED", "MF_MIXED_BF16X3", "MF_MIXED_FP16X2", "SolveReport",
    "fortran_empty", "fortran_zeros", "to_device", "to_host", "is_fortran",
    "DerivativeCoefficients", "derivative", "derx", "dderx", "dery", "ddery", "derz", "dderz",
    "BoundaryCondition", "SolverCoeffs3D", "solver_3d", "der_coeffs_1d", "solver_coeffs_1d",
]


def _torch():
    import torch

    return torch


# ---------------------------------------------------------------------------
# Fortran-ordered device fields
# ---------------------------------------------------------------------------

def fortran_empty(shape, device="cuda", dtype=None):
    torch = _torch()
    dtype = dtype or torch.float64
    rev = tuple(reversed(tuple(shape)))
    return torch.empty(rev, device=device, dtype=dtype).permute(*reversed(range(len(shape))))


def fortran_zeros(shape, device="cuda", dtype=None):
    t = fortran_empty(shape, device, dtype)
    t.zero_()
    return t


def is_fortran(t) -> bool:
    exp, s = [], 1
    for d in t.shape:
        exp.append(s)
        s *= d
    return all(st == e for st, e, d in zip(t.stride(), exp, t.shape) if d > 1)


def to_device(a: np.ndarray, device="cuda"):
    torch = _torch()
    a = np.asarray(a, dtype=np.float64)
    flat = torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).to(device)
    return flat.view(tuple(reversed(a.shape))).permute(*reversed(range(a.ndim)))


def to_host(t) -> np.ndarray:
    torch = _torch()
    flat = t.permute(*reversed(range(t.dim()))).contiguous().cpu()
    return np.asfortranarray(flat.numpy().reshape(tuple(reversed(t.shape))).transpose())


def _ptr(t) -> int:
    if not is_fortran(t):
        raise ValueError("tensor must have Fortran strides (use fortran_empty/to_device)")
    if t.dtype != _torch().float64:
        raise TypeError("tensor must be float64")
    return t.data_ptr()


def _stream(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------------------
# Host-side coefficient tables
# ---------------------------------------------------------------------------

def der_coeffs_1d(grid, der: int, so: int = 6):
    """(dl_lu (n,5), dr (n,7)) exactly as der_coeffs_generic_gobal + pentalu."""
    lib = load()
    g = np.ascontiguousarray(grid, dtype=np.float64)
    n = g.shape[0]
    dl = np.zeros((n, 5), order="F")
    dr = np.zeros((n, 7), order="F")
    check(lib.mf_der_coeffs_1d(_dptr(g), n, der, so, _dptr(dl), _dptr(dr)))
    return dl, dr


def solver_coeffs_1d(grid, bct, so: int = 6) -> dict:
    """One direction of solver_coeffs_init (class_solver_3d.f90:413-466)."""
    lib = load()
    g = np.ascontiguousarray(grid, dtype=np.float64)
    nall = g.shape[0]
    n = nall - 2
    out = {
        "ddl": np.zeros((n, 5), order="F"), "ddr": np.zeros((n, 8), order="F"),
        "eigen": np.zeros(n), "vect": np.zeros((n, n), order="F"),
        "inv_vect": np.zeros((n, n), order="F"), "inv_lhs": np.zeros((n, n), order="F"),
        "neum": np.zeros((2, 7), order="F"),
    }
    b = (C.c_int * 2)(*[int(v) for v in bct])
    check(lib.mf_solver_coeffs_1d(_dptr(g), nall, C.cast(b, C.c_void_p), so,
                                  *[_dptr(out[k]) for k in
                                    ("ddl", "ddr", "eigen", "vect", "inv_vect", "inv_lhs", "neum")]))
    return out


# ---------------------------------------------------------------------------
# Derivatives
# ---------------------------------------------------------------------------

class DerivativeCoefficients:
    """type(derivatives_coefficients) built by der_coeffs_init(grid, dc, n, solve, so)."""

    def __init__(self, grid, solver: bool = False, so: int = 6):
        self._lib = load()
        g = np.ascontiguousarray(grid, dtype=np.float64)
        self.n = int(g.shape[0])
        self.so = so
        self.solver = bool(solver)
        h = C.c_void_p()
        check(self._lib.mf_der_create(_dptr(g), self.n, int(self.solver), so, C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.mf_der_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def _lines(self, x, order: int, sub: bool, out=None, stream=None):
        torch = _torch()
        if x.dim() == 1:
            x2 = x.reshape(self.n, 1)
        else:
            x2 = x
        if x2.shape[0] != self.n:
            raise ValueError("line length %d != handle size %d" % (x2.shape[0], self.n))
        nl = x2.shape[1]
        if out is None:
            out = torch.zeros_like(x) if x.dim() == 1 else fortran_zeros(tuple(x2.shape), x.device)
        o2 = out.reshape(self.n, 1) if out.dim() == 1 else out
        ld = x2.stride(1) if nl > 1 else self.n
        if x2.stride(0) != 1 or (nl > 1 and (o2.stride(0) != 1 or o2.stride(1) != ld)):
            raise ValueError("lines must be contiguous columns with a common stride")
        check(self._lib.mf_der_lines(self._h, order, int(sub), C.c_void_p(x2.data_ptr()),
                                     C.c_void_p(o2.data_ptr()), nl, ld, _stream(stream)))
        return out

    # class_derivatives.f90:186-260
    def der(self, x, out=None):
        return self._lines(x, 1, False, out)

    def dder(self, x, out=None):
        return self._lines(x, 2, False, out)

    def der_s(self, x, out=None):
        return self._lines(x, 1, True, out)

    def dder_s(self, x, out=None):
        return self._lines(x, 2, True, out)


def derivative(dc: DerivativeCoefficients, x, axis: int, order: int, out=None, stream=None):
    """3-D derivative along axis (class_field.f90:629-978); returns a new field."""
    nx, ny, nz = x.shape
    if out is None:
        out = fortran_empty((nx, ny, nz), x.device)
    check(dc._lib.mf_der3d(dc.handle, axis, order, C.c_void_p(_ptr(x)), C.c_void_p(_ptr(out)),
                           nx, ny, nz, _stream(stream)))
    return out


def derx(dc, x, out=None):
    return derivative(dc, x, 0, 1, out)


def dderx(dc, x, out=None):
    return derivative(dc, x, 0, 2, out)


def dery(dc, x, out=None):
    return derivative(dc, x, 1, 1, out)


def ddery(dc, x, out=None):
    return derivative(dc, x, 1, 2, out)


def derz(dc, x, out=None):
    return derivative(dc, x, 2, 1, out)


def dderz(dc, x, out=None):
    return derivative(dc, x, 2, 2, out)


# ---------------------------------------------------------------------------
# Solver
# ---------------------------------------------------------------------------

@dataclass
class BoundaryCondition:
    """type(boundary_condition): bcx(ny-2,nz-2,2), bcy(nx-2,nz-2,2), bcz(nx-2,ny-2,2)."""
    bcx: object
    bcy: object
    bcz: object

    @staticmethod
    def zeros(nx, ny, nz, device="cuda"):
        return BoundaryCondition(fortran_zeros((ny - 2, nz - 2, 2), device),
                                 fortran_zeros((nx - 2, nz - 2, 2), device),
                                 fortran_zeros((nx - 2, ny - 2, 2), device))


class SolverCoeffs3D:
    """type(solver_coeffs_3d) built by solver_coeffs_init_3d(grid1,grid2,grid3,sc,n1,n2,n3,bct,so)."""

    def __init__(self, gx, gy, gz, bct, so: int = 6):
        self._lib = load()
        gs = [np.ascontiguousarray(g, dtype=np.float64) for g in (gx, gy, gz)]
        self.shape = tuple(int(g.shape[0]) for g in gs)
        self.bct = tuple(int(b) for b in bct)
        self.so = so
        b = (C.c_int * 6)(*self.bct)
        h = C.c_void_p()
        check(self._lib.mf_solver3d_create(_dptr(gs[0]), self.shape[0], _dptr(gs[1]), self.shape[1],
                                           _dptr(gs[2]), self.shape[2], C.cast(b, C.c_void_p), so,
                                           C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.mf_solver3d_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def profile(self, enable: bool = True) -> None:
        """Bracket every kernel launch of subsequent solves with HIP events."""
        check(self._lib.mf_solver3d_profile(self._h, int(bool(enable))))

    def profile_read(self, reset: bool = True) -> dict:
        """{class: (milliseconds, launches, algorithmic work)} since the last reset."""
        from ._lib import PROF_CLASSES

        n = len(PROF_CLASSES)
        ms = (C.c_double * n)()
        cnt = (C.c_int * n)()
        work = (C.c_double * n)()
        check(self._lib.mf_solver3d_profile_read(self._h, C.cast(ms, C.c_void_p),
                                                 C.cast(cnt, C.c_void_p),
                                                 C.cast(work, C.c_void_p), int(bool(reset))))
        return {k: (ms[i], cnt[i], work[i]) for i, k in enumerate(PROF_CLASSES)}

    def time_transforms(self, precision: int, reps: int = 10, stream=None):
        """(ms per launch, algorithmic flops per launch) of the first five
        transforms of a solve launched back to back, or, for the fp16x2
        plane pairs, of a whole inner solve's launches (overwrites the
        handle's work arrays only)."""
        ms, fl = C.c_double(), C.c_double()
        check(self._lib.mf_solver3d_time_transforms(self._h, int(precision), int(reps),
                                                    C.byref(ms), C.byref(fl), _stream(stream)))
        return ms.value, fl.value

    def solve(self, g, u, bc: BoundaryCondition, sigma: float, precision: int = MF_FP64,
              tol: float = 1e-12, max_iter: int = 8, stream=None) -> SolveReport:
        """solver_3d(sc, g, u, bc, sigma): u's interior and faces are overwritten."""
        if tuple(g.shape) != self.shape or tuple(u.shape) != self.shape:
            raise ValueError("field shape %s/%s != solver shape %s" % (g.shape, u.shape, self.shape))
        nx, ny, nz = self.shape
        for t, s in ((bc.bcx, (ny - 2, nz - 2, 2)), (bc.bcy, (nx - 2, nz - 2, 2)),
                     (bc.bcz, (nx - 2, ny - 2, 2))):
            if tuple(t.shape) != s:
                raise ValueError("boundary array shape %s != %s" % (tuple(t.shape), s))
        rep = SolveReport()
        code = self._lib.mf_solver3d_solve(
            self._h, C.c_void_p(_ptr(g)), float(sigma), C.c_void_p(_ptr(u)),
            C.c_void_p(_ptr(bc.bcx)), C.c_void_p(_ptr(bc.bcy)), C.c_void_p(_ptr(bc.bcz)),
            int(precision), float(tol), int(max_iter), C.byref(rep), _stream(stream))
        check(code)
        return rep


def solver_3d(sc: SolverCoeffs3D, g, u, bc: BoundaryCondition, sigma: float,
              precision: int = MF_FP64, **kw) -> SolveReport:
    """solver_3d (class_field.f90:1072-1109)."""
    return sc.solve(g, u, bc, sigma, precision, **kw)


FACES = ("x1", "x2", "y1", "y2", "z1", "z2")


def face_points(shape, face: int) -> int:
    nx, ny, nz = shape
    n1, n2, n3 = nx - 2, ny - 2, nz - 2
    return (n2 * n3, n1 * n3, n1 * n2)[face // 2]


def influence_impulses(sc: SolverCoeffs3D, dcs, sigma: float, face: int, first: int, count: int,
                       out=None, stream=None):
    """Unit-impulse responses of md_influence_matrix_set (class_field.f90:1396-1510).

    Returns (or fills) a device tensor (count, stride) whose row b is
    [bcx(n2,n3,2) | bcy(n1,n3,2) | bcz(n1,n2,2)] (each Fortran-ordered) of the
    inward-normal derivatives for impulse first+b on `face`.
    """
    torch = _torch()
    nx, ny, nz = sc.shape
    n1, n2, n3 = nx - 2, ny - 2, nz - 2
    stride = 2 * (n2 * n3 + n1 * n3 + n1 * n2)
    if out is None:
        out = torch.empty((count, stride), device="cuda", dtype=torch.float64)
    if out.shape != (count, stride) or not out.is_contiguous():
        raise ValueError("out must be a contiguous (count, %d) float64 tensor" % stride)
    dx, dy, dz = dcs
    check(sc._lib.mf_influence_impulses(sc.handle, dx.handle, dy.handle, dz.handle, float(sigma),
                                        int(face), int(first), int(count),
                                        C.c_void_p(out.data_ptr()), _stream(stream)))
    return out


def split_faces(row: np.ndarray, shape):
    """One output row -> (bcx, bcy, bcz) Fortran arrays."""
    nx, ny, nz = shape
    n1, n2, n3 = nx - 2, ny - 2, nz - 2
    a = 2 * n2 * n3
    b = a + 2 * n1 * n3
    bcx = row[:a].reshape((n2, n3, 2), order="F")
    bcy = row[a:b].reshape((n1, n3, 2), order="F")
    bcz = row[b:].reshape((n1, n2, 2), order="F")
    return bcx, bcy, bcz


# ---------------------------------------------------------------------------
# Navier-Stokes prediction-projection step
# ---------------------------------------------------------------------------


class NavierStokes:
    """navier3d state on the device (class_navier_3D.f90:26-76) with the
    reference's step: projection type pt (1..4, default 1) and up to nsubite
    sub-iterations per step (default 1); see mf_ns_* in include/mflops/mf.h.

    Fields are full (nx, ny, nz) Fortran-order float64 device tensors."""

    def __init__(self, gx, gy, gz, re: float, dt: float, tou: int = 2, top: int = 2,
                 nlt: int = 2, so: int = 6, pt: int = 1, nsubite: int = 1):
        lib = load()
        self._lib = lib
        gs = [np.ascontiguousarray(g, dtype=np.float64) for g in (gx, gy, gz)]
        self._grids = gs
        self.shape = tuple(len(g) for g in gs)
        h = C.c_void_p()
        check(lib.mf_ns_create(_dptr(gs[0]), len(gs[0]), _dptr(gs[1]), len(gs[1]),
                               _dptr(gs[2]), len(gs[2]), float(re), float(dt), int(tou),
                               int(top), int(nlt), int(so), C.byref(h)))
        self._h = h
        if pt != 1 or nsubite != 1:
            self.set_scheme(pt, nsubite)

    def set_scheme(self, pt: int, nsubite: int = 1, stream=None):
        """Projection type pt (1..4) and the maximum number of sub-iterations
        per step; clears the stored presolve right-hand sides."""
        check(self._lib.mf_ns_set_scheme(self._h, int(pt), int(nsubite), _stream(stream)))

    def subiterations(self):
        """(count, velocity change, pressure change) of the last step's
        sub-iterations (the changes of its last convergence test)."""
        n, dv, dp = C.c_int(), C.c_double(), C.c_double()
        check(self._lib.mf_ns_subiterations(self._h, C.byref(n), C.byref(dv), C.byref(dp)))
        return n.value, dv.value, dp.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.mf_ns_destroy(h)
            self._h = None

    def set_bc(self, comp: int, bc: "BoundaryCondition", stream=None):
        check(self._lib.mf_ns_set_bc(self._h, int(comp), C.c_void_p(_ptr(bc.bcx)),
                                     C.c_void_p(_ptr(bc.bcy)), C.c_void_p(_ptr(bc.bcz)),
                                     _stream(stream)))

    def set_forcing(self, comp: int, f=None, stream=None):
        p = C.c_void_p(_ptr(f)) if f is not None else C.c_void_p()
        check(self._lib.mf_ns_set_forcing(self._h, int(comp), p, _stream(stream)))

    def set_flow_rate(self, enable: bool = True, g0: float = 0.0, stream=None):
        check(self._lib.mf_ns_set_flow_rate(self._h, int(bool(enable)), float(g0),
                                            _stream(stream)))

    def set_state(self, u, v, w, stream=None):
        for t in (u, v, w):
            if tuple(t.shape) != self.shape:
                raise ValueError("field shape %s != %s" % (tuple(t.shape), self.shape))
        check(self._lib.mf_ns_set_state(self._h, C.c_void_p(_ptr(u)), C.c_void_p(_ptr(v)),
                                        C.c_void_p(_ptr(w)), _stream(stream)))

    def step(self, nsteps: int = 1, precision: int = MF_FP64, tol: float = 1e-12, stream=None):
        check(self._lib.mf_ns_step(self._h, int(nsteps), int(precision), float(tol),
                                   _stream(stream)))

    def get(self, stream=None):
        """(u, v, w, p, phi) of the latest time level, new device tensors."""
        out = [fortran_empty(self.shape) for _ in range(5)]
        check(self._lib.mf_ns_get(self._h, *[C.c_void_p(_ptr(t)) for t in out],
                                  _stream(stream)))
        return tuple(out)

    def info(self, stream=None):
        t, g = C.c_double(), C.c_double()
        check(self._lib.mf_ns_info(self._h, C.byref(t), C.byref(g), _stream(stream)))
        return t.value, g.value

    def set_les(self, les_type: int, c: float = 0.0, stream=None):
        """Eddy-viscosity model (0 off, 1 Smagorinsky, 2 WALE, 3 sigma) with
        constant c (squared inside, as navier_initialization does)."""
        check(self._lib.mf_ns_set_les(self._h, int(les_type), float(c), _stream(stream)))

    def restart_write(self, dirname=None, stream=None) -> str:
        """restart_write: the restart directory in the reference's layout; returns its name."""
        out = C.create_string_buffer(1024)
        check(self._lib.mf_ns_restart_write(self._h, dirname.encode() if dirname else None,
                                            C.cast(out, C.c_void_p), 1024, _stream(stream)))
        return out.value.decode()

    def restart_read(self, dirname=None, stream=None):
        """restart_read: every time level, time and level order from dirname."""
        check(self._lib.mf_ns_restart_read(self._h, dirname.encode() if dirname else None,
                                           _stream(stream)))

    def set_sgs(self, net, gscale: float = 1.0, lscale: float = 1.0, tscale: float = 1.0,
                stream=None):
        """Learned closure (an SGSNet, or None to turn it off): tau =
        tscale * net(gscale du, lscale Delta), added as -div tau."""
        self._sgs = net  # keep the network alive while the step uses it
        h = net._h if net is not None else C.c_void_p()
        check(self._lib.mf_ns_set_sgs(self._h, h, float(gscale), float(lscale), float(tscale),
                                      _stream(stream)))

    def set_sgs_coupling(self, coupling: int):
        """0: the closure's stress enters as -div tau; 1: as the eddy viscosity
        max(0, -tau:S)/(2 S:S)."""
        check(self._lib.mf_ns_set_sgs_coupling(self._h, int(coupling)))

    def les_nu(self, stream=None):
        out = fortran_empty(self.shape)
        check(self._lib.mf_ns_les_nu(self._h, C.c_void_p(_ptr(out)), _stream(stream)))
        return out


# ---------------------------------------------------------------------------
# Multidomain (influence-matrix) solve
# ---------------------------------------------------------------------------

class MDReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("residual", C.c_double)]


class Multidomain:
    """Influence-matrix multidomain solve (multidomain_solve,
    class_field.f90:1295-1388; md_influence_matrix_set :1396-1510; md_solve,
    class_md.f90:379-438); see mf_md_* in include/mflops/mf.h.

    grids[ax][k] is the 1-D grid of slab k along axis ax (consecutive slabs
    share their end point).  Subdomains are ordered x fastest; solve() takes
    and returns per-subdomain lists of full-size Fortran-order device
    fields; bc[s] is a device vector [bcx | bcy | bcz] (each Fortran-ordered)
    of subdomain s, outer faces used."""

    def __init__(self, grids, sigma: float, bct=(1,) * 6, so: int = 6, cache=None,
                 periods=(False, False, False), psm: int = 0):
        """cache: optional path of the influence-matrix file (read when it
        matches this layout, else written after assembly; self.loaded tells).
        periods: periodic axes (the last slab's high face joins the first
        one's low face).  psm: the pressure singular method for singular
        systems (0 none, 1 constant null space, 2 first rhs entry zeroed)."""
        lib = load()
        self._lib = lib
        self.nd = tuple(len(g) for g in grids)
        flat = [np.ascontiguousarray(g, dtype=np.float64) for ax in range(3) for g in grids[ax]]
        npts = np.array([len(g) for g in flat], dtype=np.int32)
        allg = np.ascontiguousarray(np.concatenate(flat))
        nd = np.array(self.nd, dtype=np.int32)
        bc = np.array(bct, dtype=np.int32)
        self.shapes = [tuple(len(grids[ax][s[ax]]) for ax in range(3))
                       for s in [(a, b, c) for c in range(self.nd[2]) for b in range(self.nd[1])
                                 for a in range(self.nd[0])]]
        h = C.c_void_p()
        self.loaded = False
        per = np.array([1 if p else 0 for p in periods], dtype=np.int32)
        ld = C.c_int()
        check(lib.mf_md_create_ex(_dptr(nd), _dptr(npts), _dptr(allg), _dptr(bc), _dptr(per),
                                  int(so), float(sigma),
                                  str(cache).encode() if cache is not None else None,
                                  C.byref(ld), C.byref(h)))
        self.loaded = bool(ld.value)
        self._h = h
        if psm:
            self.set_null_space(psm)
        nsub, ninf, nbytes = C.c_int(), C.c_longlong(), C.c_longlong()
        check(lib.mf_md_info(h, C.byref(nsub), C.byref(ninf), C.byref(nbytes)))
        self.nsub, self.ninf, self.matrix_bytes = nsub.value, ninf.value, nbytes.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.mf_md_destroy(h)
            self._h = None

    def set_guess(self, enable: bool = True):
        check(self._lib.mf_md_set_guess(self._h, int(bool(enable))))

    def assembly_stats(self):
        """(device ms, GEMM flops, block bytes) of the influence-block assembly."""
        ms, fl, by = C.c_double(), C.c_double(), C.c_double()
        check(self._lib.mf_md_assembly_stats(self._h, C.byref(ms), C.byref(fl), C.byref(by)))
        return ms.value, fl.value, by.value

    def set_null_space(self, psm: int):
        """Pressure singular method psm (0, 1, 2) of the following solves."""
        check(self._lib.mf_md_set_null_space(self._h, int(psm)))

    def write_matrix(self, path):
        check(self._lib.mf_md_write_matrix(self._h, str(path).encode()))

    def matvec(self, x, y=None, stream=None):
        torch = _torch()
        if y is None:
            y = torch.empty_like(x)
        check(self._lib.mf_md_matvec(self._h, C.c_void_p(_ptr(x)), C.c_void_p(_ptr(y)),
                                     _stream(stream)))
        return y

    def solve(self, g, bc, u=None, tol: float = 1e-10, max_iter: int = 0, x_iface=None,
              stream=None):
        if len(g) != self.nsub or len(bc) != self.nsub:
            raise ValueError("one g and one bc per subdomain (%d)" % self.nsub)
        if u is None:
            u = [fortran_zeros(sh, g[0].device) for sh in self.shapes]
        arr = C.c_void_p * self.nsub
        gp = arr(*[_ptr(t) for t in g])
        up = arr(*[_ptr(t) for t in u])
        bp = arr(*[_ptr(t) for t in bc])
        rep = MDReport()
        xp = C.c_void_p(_ptr(x_iface)) if x_iface is not None else C.c_void_p()
        check(self._lib.mf_md_solve(self._h, gp, up, bp, float(tol), int(max_iter), xp,
                                    C.byref(rep), _stream(stream)))
        return u, rep


# ---------------------------------------------------------------------------
# Spline filter, eddy viscosities, learned SGS closure
# ---------------------------------------------------------------------------

class SplineFilter:
    """1-D least-squares spline filter (LSQUnivariateSpline,
    class_filter.f90:130-174): degree-k spline on curfit's iopt = -1 knots
    built from the nodes xn, fitted to values on xi with weights wi,
    evaluated at xo (deriv=0) or differentiated there (deriv=1)."""

    def __init__(self, xi, xo, xn, k: int = 3, wi=None, deriv: int = 0):
        lib = load()
        self._lib = lib
        xi, xo, xn = (np.ascontiguousarray(a, dtype=np.float64) for a in (xi, xo, xn))
        w = C.c_void_p()
        if wi is not None:
            self._wi = np.ascontiguousarray(wi, dtype=np.float64)
            if self._wi.shape != xi.shape:
                raise ValueError("one weight per input point")
            w = _dptr(self._wi)
        self.nin, self.nout = len(xi), len(xo)
        h = C.c_void_p()
        check(lib.mf_filter_create(_dptr(xi), w, len(xi), _dptr(xo), len(xo), _dptr(xn),
                                   len(xn), int(k), int(deriv), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.mf_filter_destroy(h)
            self._h = None

    def matrix(self) -> np.ndarray:
        F = np.zeros((self.nout, self.nin), order="F")
        check(self._lib.mf_filter_matrix(self._h, _dptr(F)))
        return F


def filter3d(fx: SplineFilter, fy: SplineFilter, fz: SplineFilter, x, out=None, stream=None):
    """Filter a (nxi, nyi, nzi) Fortran-order device field to (nxo, nyo, nzo)."""
    torch = _torch()
    if tuple(x.shape) != (fx.nin, fy.nin, fz.nin):
        raise ValueError("field shape %s != (%d, %d, %d)" % (tuple(x.shape), fx.nin, fy.nin, fz.nin))
    if out is None:
        out = fortran_empty((fx.nout, fy.nout, fz.nout), x.device)
    work = torch.empty(fx.nout * fy.nin * fz.nin + fx.nout * fy.nout * fz.nin,
                       dtype=torch.float64, device=x.device)
    check(fx._lib.mf_filter3d(fx._h, fy._h, fz._h, C.c_void_p(_ptr(x)), C.c_void_p(_ptr(out)),
                              C.c_void_p(work.data_ptr()), _stream(stream)))
    return out


LES_SMAGORINSKY, LES_WALE, LES_SIGMA = 1, 2, 3


def les_nu(du, gx, gy, gz, les_type: int, c: float, out=None, stream=None):
    """Eddy viscosity (navier_LES, class_navier_3D.f90:1225-1449) from
    du[i][j] = d u_i / d x_j (device fields) on the grids gx, gy, gz."""
    lib = load()
    gs = [np.ascontiguousarray(g, dtype=np.float64) for g in (gx, gy, gz)]
    shape = tuple(len(g) for g in gs)
    flat = [du[i][j] for i in range(3) for j in range(3)]
    for t in flat:
        if tuple(t.shape) != shape:
            raise ValueError("gradient shape %s != %s" % (tuple(t.shape), shape))
    if out is None:
        out = fortran_empty(shape, flat[0].device)
    arr = (C.c_void_p * 9)(*[_ptr(t) for t in flat])
    check(lib.mf_les_nu(arr, _dptr(gs[0]), shape[0], _dptr(gs[1]), shape[1], _dptr(gs[2]),
                        shape[2], int(les_type), float(c), C.c_void_p(_ptr(out)), _stream(stream)))
    return out


MLP_SIZES = (10, 64, 64, 6)
MLP_NPARAM = sum(a * b + b for a, b in zip(MLP_SIZES[:-1], MLP_SIZES[1:]))


class SGSNet:
    """Learned SGS closure 10 -> 64 -> 64 -> 6 (tanh) on the device; see
    mf_mlp_* in include/mflops/mf.h.  X (n, 10) and T (n, 6) are row-major
    float32 device tensors."""

    def __init__(self, params):
        lib = load()
        self._lib = lib
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != MLP_NPARAM:
            raise ValueError("expected %d parameters, got %d" % (MLP_NPARAM, p.size))
        h = C.c_void_p()
        check(lib.mf_mlp_create(p.ctypes.data_as(C.c_void_p), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.mf_mlp_destroy(h)
            self._h = None

    @staticmethod
    def _f32(t, cols):
        torch = _torch()
        if t.dtype != torch.float32 or not t.is_contiguous() or t.dim() != 2 or t.shape[1] != cols:
            raise ValueError("expected a contiguous float32 (n, %d) tensor" % cols)
        return C.c_void_p(t.data_ptr())

    def params(self) -> np.ndarray:
        p = np.zeros(MLP_NPARAM, dtype=np.float32)
        check(self._lib.mf_mlp_params(self._h, p.ctypes.data_as(C.c_void_p)))
        return p

    def forward(self, X, out=None, stream=None):
        torch = _torch()
        if out is None:
            out = torch.empty((X.shape[0], 6), dtype=torch.float32, device=X.device)
        check(self._lib.mf_mlp_forward(self._h, self._f32(X, 10), self._f32(out, 6), X.shape[0],
                                       _stream(stream)))
        return out

    def loss_grad(self, X, T, stream=None):
        torch = _torch()
        g = torch.empty(MLP_NPARAM, dtype=torch.float32, device=X.device)
        loss = C.c_float()
        check(self._lib.mf_mlp_loss_grad(self._h, self._f32(X, 10), self._f32(T, 6), X.shape[0],
                                         C.byref(loss), C.c_void_p(g.data_ptr()), _stream(stream)))
        return loss.value, g

    def adam_step(self, X, T, lr: float = 1e-3, stream=None) -> float:
        loss = C.c_float()
        check(self._lib.mf_mlp_adam_step(self._h, self._f32(X, 10), self._f32(T, 6), X.shape[0],
                                         float(lr), C.byref(loss), _stream(stream)))
        return loss.value


def sgs_targets(fx: SplineFilter, fy: SplineFilter, fz: SplineFilter, u, v, w,
                tscale: float = 1.0, stream=None):
    """(ubar, vbar, wbar), tau (n x 6 float32 / tscale) of a fine field."""
    torch = _torch()
    shape_o = (fx.nout, fy.nout, fz.nout)
    nf = fx.nin * fy.nin * fz.nin
    nc = fx.nout * fy.nout * fz.nout
    bars = [fortran_empty(shape_o, u.device) for _ in range(3)]
    tau = torch.empty((nc, 6), dtype=torch.float32, device=u.device)
    work = torch.empty(nf + nc + fx.nout * fy.nin * fz.nin + fx.nout * fy.nout * fz.nin,
                       dtype=torch.float64, device=u.device)
    for t in (u, v, w):
        if tuple(t.shape) != (fx.nin, fy.nin, fz.nin):
            raise ValueError("field shape %s != filter input" % (tuple(t.shape),))
    check(fx._lib.mf_sgs_targets(fx._h, fy._h, fz._h, *[C.c_void_p(_ptr(t)) for t in (u, v, w)],
                                 *[C.c_void_p(_ptr(t)) for t in bars], C.c_void_p(tau.data_ptr()),
                                 float(tscale), C.c_void_p(work.data_ptr()), _stream(stream)))
    return bars, tau


def sgs_eddy_targets(X, c: float, stream=None):
    """Smagorinsky stress -2 c^2 Delta^2 |S| S (n x 6 float32) of unscaled
    feature rows X (n x 10 float32)."""
    torch = _torch()
    if X.dtype != torch.float32 or X.dim() != 2 or X.shape[1] != 10 or not X.is_contiguous():
        raise ValueError("X must be a contiguous (n, 10) float32 tensor")
    T = torch.empty((X.shape[0], 6), dtype=torch.float32, device=X.device)
    check(load().mf_sgs_eddy_targets(C.c_void_p(X.data_ptr()), int(X.shape[0]), float(c),
                                     C.c_void_p(T.data_ptr()), _stream(stream)))
    return T


def sgs_features(dcs, u, v, w, grids, gscale: float = 1.0, lscale: float = 1.0, stream=None):
    """Closure inputs (n x 10 float32) of a velocity field; dcs are solver-mode
    DerivativeCoefficients of the three grids."""
    torch = _torch()
    lib = load()
    gs = [np.ascontiguousarray(g, dtype=np.float64) for g in grids]
    shape = tuple(len(g) for g in gs)
    for t in (u, v, w):
        if tuple(t.shape) != shape:
            raise ValueError("field shape %s != %s" % (tuple(t.shape), shape))
    n = shape[0] * shape[1] * shape[2]
    X = torch.empty((n, 10), dtype=torch.float32, device=u.device)
    work = torch.empty(9 * n, dtype=torch.float64, device=u.device)
    check(lib.mf_sgs_features(*[d._h for d in dcs], *[C.c_void_p(_ptr(t)) for t in (u, v, w)],
                              *shape, *[_dptr(g) for g in gs], float(gscale), float(lscale),
                              C.c_void_p(X.data_ptr()), C.c_void_p(work.data_ptr()),
                              _stream(stream)))
    return X


# ---------------------------------------------------------------------------
# Field files (NetCDF classic)
# ---------------------------------------------------------------------------

FIELD_DIMS = ("resolution_x", "resolution_y", "resolution_z")


def write_field(path, var_name: str, f: np.ndarray, dbl: bool = True, dim_names=FIELD_DIMS):
    """write_field: host array f(nx,ny,nz) as var_name in a NetCDF classic file."""
    lib = load()
    a = np.asfortranarray(f, dtype=np.float64)
    names = (C.c_char_p * 3)(*[d.encode() for d in dim_names])
    check(lib.mf_io_write_field(str(path).encode(), var_name.encode(), C.cast(names, C.c_void_p),
                                *a.shape, a.ctypes.data_as(C.c_void_p), int(bool(dbl))))


def field_info(path, var_name: str):
    """get_var3d_info: (sizes, dimension names, stored in FP64)."""
    lib = load()
    dims = (C.c_int * 3)()
    names = C.create_string_buffer(3 * 256)
    dbl = C.c_int()
    check(lib.mf_io_field_info(str(path).encode(), var_name.encode(), C.cast(dims, C.c_void_p),
                               C.cast(names, C.c_void_p), 256, C.byref(dbl)))
    raw = names.raw
    dn = tuple(raw[256 * a:256 * (a + 1)].split(b"\0")[0].decode() for a in range(3))
    return tuple(dims), dn, bool(dbl.value)


def read_field(path, var_name: str) -> np.ndarray:
    """read_field: the variable as a Fortran-ordered FP64 host array."""
    lib = load()
    dims, _, _ = field_info(path, var_name)
    out = np.zeros(dims, order="F")
    check(lib.mf_io_read_field(str(path).encode(), var_name.encode(), *dims,
                               out.ctypes.data_as(C.c_void_p)))
    return out
