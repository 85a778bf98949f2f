"""CPU restatement of the reference's prediction-projection Navier-Stokes step.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker, never by the product path.

Restates one time step of navier3d.f90:48-286 for a single domain with the
reference's defaults (command_line.f90:20-35): skew-symmetric nonlinear term
(nlt=2), non-incremental rotational pressure correction with tangential
boundary-gradient correction (pt=1), BDF order tou and pressure
extrapolation order top-1 (tou, top in 1..3), no mapping (mapt=0), one
sub-iteration.  LES (set_les) follows navier_LES (:1225-1452, via
les_oracle.les_nu) and the LES part of f(nav,var) (:1083-1134): nu from the
latest level's gradients du_ij = d u_i / d x_j (edges zeroed), then
f_i += sum_j d_j (nu (du_ij + du_ji)).  All derivatives are the
solver-mode compact derivatives (derivatives_coefficients_init(...,
solver='yes'), class_navier_3D.f90:2287-2292) and all solves are solver_3d in
FP64 (mf_oracle.solver_3d).

Step (class_navier_3D.f90 line numbers):
  time += dt                                      navier_time :2052-2064
  bc_u(it1) = boundary values                     navier_bc_velocity :395-433
  rhs_q = f_q - extrapol_v(N_q)                   navier_presolve_u/v/w :1528-1595,
      N_q = (u q_x + v q_y + w q_z)/2               navier_nonlinear :1169-1222
            + ((q u)_x + (q v)_y + (q w)_z)/2
  tangential bc += grad(extrapol_p(phi))/fac1     add_boundary_gradient :695-791
  q(it1) = solver_3d(Re (fac2 q^n + fac3 q^n-1    navier_solve_u/v/w :1597-1761
           + fac4 q^n-2 - rhs_q), sigma=-Re fac1)
  rhs_pq = fac1 q(it1)                            navier_presolve_phi/phi_rhs :1764-1785,1961-2007
  phi(it1) = solver_3d(div rhs_p, Neumann bc      navier_solve_phi :1787-1845,
           rhs_p.n - fac1 bc_q.n, sigma=0)          navier_bc_pressure :342-393
  p(it1) = phi - div(q(it1))/Re                   navier_projection :1847-1959
  q(it1) = (rhs_pq - grad phi)/fac1
  rotate it                                       navier3d.f90:269-273

The constant-flow-rate forcing of f(nav,'rhsu') (:1124-1164) is optional
here (FlowRate); the reference applies it to every run.
"""
from __future__ import annotations

import numpy as np

import les_oracle as les
import mf_oracle as o

NT = 4


def bdf_factors(tou: int, dt: float) -> np.ndarray:
    """fac(1:4) of navier_initialization (class_navier_3D.f90:2145-2160)."""
    if tou == 1:
        f = [1.0, -1.0, 0.0, 0.0]
        return np.array(f) / dt
    if tou == 2:
        return np.array([1.5, -2.0, 0.5, 0.0]) / dt
    if tou == 3:
        return np.array([11.0, -18.0, 9.0, -2.0]) / (6.0 * dt)
    raise ValueError("tou must be 1, 2 or 3")


def extrapol(levels, it, order: int):
    """navier_extrapol (class_navier_3D.f90:2010-2050) for subite == 1:
    order -1: level it(1); 1: it(nt); 2: 2 it(nt) - it(nt-1);
    3: 3 it(nt) - 3 it(nt-1) + it(nt-2); anything else: 0."""
    a = levels
    if order == -1:
        out = 1.0 * a[it[0]]
    elif order == 1:
        out = 1.0 * a[it[NT - 1]]
    elif order == 2:
        out = 2.0 * a[it[NT - 1]] - 1.0 * a[it[NT - 2]]
    elif order == 3:
        out = 3.0 * a[it[NT - 1]] - 3.0 * a[it[NT - 2]] + 1.0 * a[it[NT - 3]]
    else:
        out = np.zeros_like(a[it[0]])
    return zero_edges(out)


def zero_edges(f: np.ndarray) -> np.ndarray:
    """field_zero_edges (class_field.f90:366-391): the 12 edge lines -> 0."""
    f[0, 0, :] = f[-1, 0, :] = f[0, -1, :] = f[-1, -1, :] = 0.0
    f[0, :, 0] = f[-1, :, 0] = f[0, :, -1] = f[-1, :, -1] = 0.0
    f[:, 0, 0] = f[:, -1, 0] = f[:, 0, -1] = f[:, -1, -1] = 0.0
    return f


def integral_weights(gx, gy, gz) -> np.ndarray:
    """Weights of integrale (class_navier_3D.f90:2419-2507): half-width dual
    cells on interior points and faces, edges and corners excluded."""
    def w1(g):
        g = np.asarray(g, dtype=np.float64)
        w = np.empty_like(g)
        w[1:-1] = (g[2:] - g[:-2]) * 0.5
        w[0] = (g[1] - g[0]) * 0.5
        w[-1] = (g[-1] - g[-2]) * 0.5
        return w
    wx, wy, wz = w1(gx), w1(gy), w1(gz)
    W = wx[:, None, None] * wy[None, :, None] * wz[None, None, :]
    nb = np.zeros(W.shape, dtype=int)
    nb[[0, -1], :, :] += 1
    nb[:, [0, -1], :] += 1
    nb[:, :, [0, -1]] += 1
    W[nb >= 2] = 0.0
    return np.asfortranarray(W)


class FlowRate:
    """Constant-flow-rate forcing of f(nav,'rhsu') (class_navier_3D.f90:1124-1164):
    q = integral(extrapol_v(u)) / integral(1); on the first call
    debit = q (1 if q <= 1e-10) and g = g0 (the reference: 3 debit);
    then g -= (q - debit) fac1 and f_u += g."""

    def __init__(self, g0=None):
        self.g0 = g0
        self.g = None
        self.debit = None

    def update(self, q: float, fac1: float) -> float:
        if self.g is None:
            self.debit = q if q > 1e-10 else 1.0
            self.g = 3.0 * self.debit if self.g0 is None else float(self.g0)
        self.g -= (q - self.debit) * fac1
        return self.g


class NSOracle:
    """Single-domain NS state with nt=4 time levels (class_navier_3D.f90:26-76)."""

    def __init__(self, gx, gy, gz, re: float, dt: float, tou: int = 2, top: int = 2,
                 nlt: int = 2, so: int = 6):
        self.grids = [np.asarray(g, dtype=np.float64) for g in (gx, gy, gz)]
        self.shape = tuple(len(g) for g in self.grids)
        self.re, self.dt, self.tou, self.top, self.nlt = float(re), float(dt), tou, top, nlt
        self.fac = bdf_factors(tou, dt)
        self.sigmau = -self.re * self.fac[0]
        self.dc = [o.DerCoeffs(g, solver=True, so=so) for g in self.grids]
        self.scu = o.SolverCoeffs3D(*self.grids, (1,) * 6, so)
        self.scp = o.SolverCoeffs3D(*self.grids, (2,) * 6, so)
        z = lambda: np.zeros(self.shape, order="F")
        self.q = {c: [z() for _ in range(NT)] for c in "uvw"}
        self.p = [z() for _ in range(NT)]
        self.phi = [z() for _ in range(NT)]
        self.nl = {c: [None] * NT for c in "uvw"}
        nx, ny, nz = self.shape
        self.bc = {c: (np.zeros((ny - 2, nz - 2, 2), order="F"),
                       np.zeros((nx - 2, nz - 2, 2), order="F"),
                       np.zeros((nx - 2, ny - 2, 2), order="F")) for c in "uvw"}
        self.force = {c: None for c in "uvw"}
        self.flow = None
        self.W = integral_weights(*self.grids)
        self.it = list(range(NT))
        self.time = 0.0
        self.les_type, self.les_c2, self.les_nu = 0, 0.0, None
        self.sgs = None
        # called with each predicted velocity right after its solve (tests
        # use it to perturb the step and measure the result's sensitivity)
        self.velocity_hook = None

    def set_sgs(self, params, gscale: float, lscale: float, tscale: float):
        """Learned closure: tau = tscale MLP(X), X as les.sgs_features (rounded
        to float32, as the device stores it), added as -div tau."""
        self.sgs = (np.asarray(params, dtype=np.float32), gscale, lscale, tscale)

    def set_les(self, les_type: int, c: float):
        """les_type 0..3 and the command-line constant (squared, :2138)."""
        self.les_type, self.les_c2 = int(les_type), float(c) ** 2

    def les_force(self, L: int):
        """-les_f of the three components at level L (navier_LES + f), plus
        -div tau of the learned closure when set."""
        q = [self.q[c][L] for c in "uvw"]
        du = [[zero_edges(self.der(q[i], j)) for j in range(3)] for i in range(3)]
        out = {c: np.zeros(self.shape, order="F") for c in "uvw"}
        if self.les_type:
            nu = les.les_nu(du, *self.grids, self.les_type, self.les_c2)
            self.les_nu = nu
            for i, c in enumerate("uvw"):
                for j in range(3):
                    out[c] = out[c] + self.der(nu * (du[i][j] + du[j][i]), j)
        if self.sgs is not None:
            p, gs, ls, ts = self.sgs
            X = les.sgs_features(du, *self.grids, gs, ls).astype(np.float32)
            Y = les.forward(p, X)[0]
            tau = [ts * Y[:, k].reshape(self.shape, order="F") for k in range(6)]
            f = self.minus_div_tau(tau)
            for c in "uvw":
                out[c] = out[c] + f[c]
        return out

    def minus_div_tau(self, tau):
        """-sum_j d_j tau_ij for i = u, v, w; tau holds the six columns
        (xx, yy, zz, xy, xz, yz) as fields, zeroed on the box edges before
        differentiation (the stress enters as the LES stress does in
        f(nav,var), class_navier_3D.f90:1083-1134)."""
        out = {c: np.zeros(self.shape, order="F") for c in "uvw"}
        for i, c in enumerate("uvw"):
            for j in range(3):
                t = np.array(tau[les.pair_col(i, j)], dtype=np.float64, order="F", copy=True)
                out[c] = out[c] - self.der(zero_edges(t), j)
        return out

    # ---- set-up
    def set_state(self, u, v, w):
        """All time levels = the given velocity; p = phi = 0."""
        for c, f in zip("uvw", (u, v, w)):
            for L in range(NT):
                self.q[c][L] = zero_edges(np.array(f, dtype=np.float64, order="F", copy=True))
                self.nl[c][L] = None

    def der(self, f, axis):
        return o.derivative_3d(self.dc[axis], f, axis, 1)

    # ---- pieces of the step
    def nonlinear(self, L: int, c: str):
        """navier_nonlinear (:1169-1222), level L, transported component c."""
        u, v, w = self.q["u"][L], self.q["v"][L], self.q["w"][L]
        x = self.q[c][L]
        if self.nlt == 1:
            return u * self.der(x, 0) + v * self.der(x, 1) + w * self.der(x, 2)
        if self.nlt == 2:
            t = (u * self.der(x, 0) + v * self.der(x, 1) + w * self.der(x, 2)) * 0.5
            t = t + self.der(x * u, 0) * 0.5
            t = t + self.der(x * v, 1) * 0.5
            t = t + self.der(x * w, 2) * 0.5
            return t
        return np.zeros(self.shape, order="F")

    def _nl(self, L, c):
        if self.nl[c][L] is None:
            self.nl[c][L] = self.nonlinear(L, c)
        return self.nl[c][L]

    def step(self):
        it, fac, re = self.it, self.fac, self.re
        n1 = it[0]
        self.time += self.dt
        # navier_bc_velocity: boundary values of level it(1)
        bc = {c: [b.copy() for b in self.bc[c]] for c in "uvw"}
        # navier_LES on the latest level (navier3d.f90:68)
        lesf = self.les_force(it[NT - 1]) if (self.les_type or self.sgs is not None) else None
        # presolve: rhs_c = f_c - extrapol_v(N_c)
        rhs = {}
        for c in "uvw":
            lv = [None] * NT
            for L in (it[NT - 1], it[NT - 2], it[NT - 3]):
                lv[L] = self._nl(L, c)
            for L in range(NT):
                if lv[L] is None:
                    lv[L] = np.zeros(self.shape, order="F")
            f = np.zeros(self.shape, order="F") if self.force[c] is None else self.force[c].copy()
            if c == "u" and self.flow is not None:
                ue = extrapol(self.q["u"], it, self.tou)
                qf = float((self.W * ue).sum() / self.W.sum())
                f = f + self.flow.update(qf, fac[0])
            if lesf is not None:
                f = f + lesf[c]
            rhs[c] = zero_edges(f - extrapol(lv, it, self.tou))
        # add_boundary_gradient (pt=1): tangential bc += d(extrapol_p(phi))/fac1
        pe = extrapol(self.phi, it, self.top - 1)
        d = [self.der(pe, a) for a in range(3)]
        nx, ny, nz = self.shape
        for m, s in ((0, 0), (1, -1)):
            bc["v"][0][:, :, m] += d[1][s, 1:-1, 1:-1] / fac[0]
            bc["w"][0][:, :, m] += d[2][s, 1:-1, 1:-1] / fac[0]
            bc["u"][1][:, :, m] += d[0][1:-1, s, 1:-1] / fac[0]
            bc["w"][1][:, :, m] += d[2][1:-1, s, 1:-1] / fac[0]
            bc["u"][2][:, :, m] += d[0][1:-1, 1:-1, s] / fac[0]
            bc["v"][2][:, :, m] += d[1][1:-1, 1:-1, s] / fac[0]
        # velocity solves (navier_solve_u/v/w, pt=1: no pressure term)
        for c in "uvw":
            q = self.q[c]
            g = fac[1] * q[it[NT - 1]] + fac[2] * q[it[NT - 2]]
            if self.tou == 3:
                g = g + fac[3] * q[it[NT - 3]]
            g = re * (g - rhs[c])
            o.solver_3d(self.scu, g, q[n1], *bc[c], self.sigmau)
            if self.velocity_hook is not None:
                self.velocity_hook(q[n1])
        # presolve_phi: rhs_p = fac1 q(it1)
        rp = {c: zero_edges(fac[0] * self.q[c][n1]) for c in "uvw"}
        # solve_phi (kmax = 1 without mapping)
        fphi = 0.0 * extrapol(self.phi, it, self.top - 1)
        fphi = fphi + self.der(rp["u"], 0) + self.der(rp["v"], 1) + self.der(rp["w"], 2)
        bcphi = (
            rp["u"][[0, -1]][:, 1:-1, 1:-1].transpose(1, 2, 0) - fac[0] * bc["u"][0],
            rp["v"][:, [0, -1]][1:-1, :, 1:-1].transpose(0, 2, 1) - fac[0] * bc["v"][1],
            rp["w"][:, :, [0, -1]][1:-1, 1:-1, :] - fac[0] * bc["w"][2],
        )
        bcphi = tuple(np.asfortranarray(b) for b in bcphi)
        o.solver_3d(self.scp, fphi, self.phi[n1], *bcphi, 0.0)
        # projection
        zero_edges(self.phi[n1])
        p = self.phi[n1].copy()
        zero_edges(p)
        div = (self.der(self.q["u"][n1], 0) + self.der(self.q["v"][n1], 1)
               + self.der(self.q["w"][n1], 2)) / re
        p = zero_edges(p - div)
        self.p[n1] = p
        for a, c in enumerate("uvw"):
            self.q[c][n1] = zero_edges((rp[c] - self.der(self.phi[n1], a)) / fac[0])
            self.nl[c][n1] = None
        # rotate the time levels
        self.it = it[1:] + it[:1]

    # ---- accessors (latest level)
    def velocity(self):
        L = self.it[NT - 1]
        return self.q["u"][L], self.q["v"][L], self.q["w"][L]

    def pressure(self):
        return self.p[self.it[NT - 1]]

    def phi_latest(self):
        return self.phi[self.it[NT - 1]]
