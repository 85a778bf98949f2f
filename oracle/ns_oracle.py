"""CPU restatement of the reference's prediction-projection Navier-Stokes step.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker, never by the product path.

Restates one time step of navier3d.f90:48-286 for a single domain, no
mapping (mapt=0): nonlinear form nlt (2 skew-symmetric, the default of
command_line.f90:20-35, or 1 advective), BDF order tou and pressure
extrapolation order top-1 (tou, top in 1..3), projection type pt 1..4 and up
to nsubite sub-iterations.  LES (set_les) follows navier_LES (:1225-1452, via
les_oracle.les_nu) and the LES part of f(nav,var) (:1083-1134): nu from the
latest level's gradients du_ij = d u_i / d x_j (edges zeroed), then
f_i += sum_j d_j (nu (du_ij + du_ji)).  All derivatives are the
solver-mode compact derivatives (derivatives_coefficients_init(...,
solver='yes'), class_navier_3D.f90:2287-2292) and all solves are solver_3d in
FP64 (mf_oracle.solver_3d).

Sub-iteration s (class_navier_3D.f90 line numbers; e_v, e_p = extrapol of
order tou, top-1 over it(nt), it(nt-1), it(nt-2) when s = 1, level it(1)
when s > 1, navier_extrapol :2010-2050):
  time += dt (once per step)                      navier_time :2052-2064
  bc_u(it1) = boundary values                     navier_bc_velocity :395-433
  rhs2_q = rhs1_q; rhs1_q = f_q - e_v(N_q)        navier_presolve_u/v/w :1528-1595,
      N_q = (u q_x + v q_y + w q_z)/2               navier_nonlinear :1169-1222
            + ((q u)_x + (q v)_y + (q w)_z)/2
  pt <= 2:
    pt = 1: tangential bc += grad(e_p(phi))/fac1  add_boundary_gradient :695-791
    q(it1) = solver_3d(Re (fac2 q^n + fac3 q^n-1  navier_solve_u/v/w :1597-1761
             [+ fac4 q^n-2] [+ d_q e_p(p), pt=2]
             - rhs1_q), sigma=-Re fac1)
  rhs_pq = pt <= 2: fac1 q(it1)                   navier_phi_rhs :1961-2007
           pt = 3: rhs1 - fac2 q^n - fac3 q^n-1 - fac4 q^n-2
           pt = 4, s = 1: that - rhs2 + fac1 q^n + fac2 q^n-1 + fac3 q^n-2
                          + fac4 q(it1)
           pt = 4, s > 1: rhs1 - rhs2 + fac1 q(it1)
  phi(it1) = solver_3d(div rhs_p, Neumann bc      navier_solve_phi :1787-1845,
           rhs_p.n - fac1 bc_q.n, sigma=0;          navier_bc_pressure :342-393
           pt = 3: minus (rot rot e_p(q)).n / Re    add_boundary_rotrot :793-846
  p(it1) = phi [+ e_p(p), pt = 2, 4]              navier_projection :1847-1959
           - div(q(it1))/Re (pt <= 2) or - div(e_p(q))/Re (pt = 4)
  q(it1) = (rhs_pq - grad phi)/fac1
  pt >= 3: q(it1) = velocity solves with + d_q p(it1)
  s > 1: testconv; stop when both changes < 1e-9  navier3d.f90:123-131, 366-412
rotate it                                         navier3d.f90:269-273

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


def extrapol(levels, it, order: int, subite: int = 1):
    """navier_extrapol (class_navier_3D.f90:2010-2050).  subite == 1:
    order -1: level it(1); 1: it(nt); 2: 2 it(nt) - it(nt-1);
    3: 3 it(nt) - 3 it(nt-1) + it(nt-2); anything else: 0.  From the second
    sub-iteration on, level it(1) (the current iterate) whatever the order."""
    a = levels
    if subite > 1:
        out = 1.0 * a[it[0]]
    elif order == -1:
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


def sgs_eddy_viscosity(tau, du):
    """nu = max(0, -tau:S) / (2 S:S), S_ij = (du_ij + du_ji)/2, tau the six
    columns (xx, yy, zz, xy, xz, yz); 0 on the box edges and where S = 0."""
    sxx, syy, szz = du[0][0], du[1][1], du[2][2]
    sxy = 0.5 * (du[0][1] + du[1][0])
    sxz = 0.5 * (du[0][2] + du[2][0])
    syz = 0.5 * (du[1][2] + du[2][1])
    tS = (tau[0] * sxx + tau[1] * syy + tau[2] * szz
          + 2.0 * (tau[3] * sxy + tau[4] * sxz + tau[5] * syz))
    SS = sxx * sxx + syy * syy + szz * szz + 2.0 * (sxy * sxy + sxz * sxz + syz * syz)
    with np.errstate(divide="ignore", invalid="ignore"):
        nu = np.where(SS > 0.0, np.maximum(0.0, -tS) / (2.0 * SS), 0.0)
    return zero_edges(np.asfortranarray(nu))


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
                 nlt: int = 2, so: int = 6, pt: int = 1, nsubite: int = 1):
        self.grids = [np.asarray(g, dtype=np.float64) for g in (gx, gy, gz)]
        self.shape = tuple(len(g) for g in self.grids)
        self.re, self.dt, self.tou, self.top, self.nlt = float(re), float(dt), tou, top, nlt
        if pt not in (1, 2, 3, 4) or nsubite < 1:
            raise ValueError("pt must be 1..4 and nsubite >= 1")
        self.pt, self.nsubite = pt, nsubite
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
        # rhs_q(1), rhs_q(2) = f - extrapol_v(N) of this and the previous
        # presolve (zero before the first, navier_initialization)
        self.rhs1 = {c: z() for c in "uvw"}
        self.rhs2 = {c: z() for c in "uvw"}
        self.sub_q, self.sub_p = None, None
        self.conv = None          # (velocity, pressure) change of the last testconv
        self.subite_done = 0      # sub-iterations made by the last step
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

    def set_sgs(self, params, gscale: float, lscale: float, tscale: float, coupling: int = 0):
        """Learned closure: tau = tscale MLP(X), X as les.sgs_features (rounded
        to float32, as the device stores it), added as -div tau (coupling 0)
        or as the eddy viscosity nu = max(0, -tau:S)/(2 S:S) (coupling 1)."""
        self.sgs = (np.asarray(params, dtype=np.float32), gscale, lscale, tscale)
        self.sgs_coupling = coupling

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
            if getattr(self, "sgs_coupling", 0) == 1:
                nu = sgs_eddy_viscosity(tau, du)
                self.les_nu = nu
                for i, c in enumerate("uvw"):
                    for j in range(3):
                        out[c] = out[c] + self.der(nu * (du[i][j] + du[j][i]), j)
            else:
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
            self.rhs1[c] = np.zeros(self.shape, order="F")
            self.rhs2[c] = np.zeros(self.shape, order="F")

    def der(self, f, axis):
        return o.derivative_3d(self.dc[axis], f, axis, 1)

    def dder(self, f, axis):
        return o.derivative_3d(self.dc[axis], f, axis, 2)

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
        """One time step: time += dt, then up to nsubite sub-iterations
        (navier3d.f90:56-273); the last sub-iteration's level it(1) becomes
        the latest level."""
        self.time += self.dt
        self.subite_done = 0
        for sub in range(1, self.nsubite + 1):
            self.subite_done = sub
            if self._substep(sub):
                break
        self.it = self.it[1:] + self.it[:1]

    def _ex(self, levels, order, sub):
        return extrapol(levels, self.it, order, sub)

    def _substep(self, sub: int) -> bool:
        """One pass of the sub-iteration loop (navier3d.f90:73-147, 254-255);
        True when testconv ends the loop."""
        it, fac, re, pt = self.it, self.fac, self.re, self.pt
        n1 = it[0]
        nd = len(self.shape)
        # navier_bc_velocity: boundary values of level it(1)
        bc = {c: [b.copy() for b in self.bc[c]] for c in "uvw"}
        # navier_LES on the latest level (navier3d.f90:68)
        lesf = self.les_force(it[NT - 1]) if (self.les_type or self.sgs is not None) else None
        # presolve: rhs2 = rhs1; rhs1 = f - extrapol_v(N)
        self.rhs2 = self.rhs1
        self.rhs1 = {}
        for c in "uvw":
            lv = [np.zeros(self.shape, order="F") for _ in range(NT)]
            if sub == 1:
                for L in (it[NT - 1], it[NT - 2], it[NT - 3]):
                    lv[L] = self._nl(L, c)
            else:  # extrapol returns level it(1): the current iterate
                lv[n1] = self.nonlinear(n1, c)
            f = np.zeros(self.shape, order="F") if self.force[c] is None else self.force[c].copy()
            if c == "u" and self.flow is not None:
                ue = self._ex(self.q["u"], self.tou, sub)
                qf = float((self.W * ue).sum() / self.W.sum())
                f = f + self.flow.update(qf, fac[0])
            if lesf is not None:
                f = f + lesf[c]
            self.rhs1[c] = zero_edges(f - self._ex(lv, self.tou, sub))
        rhs1, rhs2 = self.rhs1, self.rhs2
        if pt <= 2:
            if pt == 1:
                # add_boundary_gradient: tangential bc += d(extrapol_p(phi))/fac1
                pe = self._ex(self.phi, self.top - 1, sub)
                d = [self.der(pe, a) for a in range(3)]
                for m, s in ((0, 0), (1, -1)):
                    bc["v"][0][:, :, m] += d[1][s, 1:-1, 1:-1] / fac[0]
                    bc["w"][0][:, :, m] += d[2][s, 1:-1, 1:-1] / fac[0]
                    bc["u"][1][:, :, m] += d[0][1:-1, s, 1:-1] / fac[0]
                    bc["w"][1][:, :, m] += d[2][1:-1, s, 1:-1] / fac[0]
                    bc["u"][2][:, :, m] += d[0][1:-1, 1:-1, s] / fac[0]
                    bc["v"][2][:, :, m] += d[1][1:-1, 1:-1, s] / fac[0]
            self._velocity_solves(bc, sub)
        # presolve_phi (navier_phi_rhs :1961-2007)
        q = self.q
        rp = {}
        for c in "uvw":
            qc = q[c]
            if pt <= 2:
                r = fac[0] * qc[n1]
            elif pt == 3 or sub == 1:
                r = (rhs1[c] - fac[1] * qc[it[NT - 1]] - fac[2] * qc[it[NT - 2]]
                     - fac[3] * qc[it[NT - 3]])
                if pt == 4:
                    r = (r - rhs2[c] + fac[0] * qc[it[NT - 1]] + fac[1] * qc[it[NT - 2]]
                         + fac[2] * qc[it[NT - 3]] + fac[3] * qc[it[0]])
            else:
                r = rhs1[c] - rhs2[c] + fac[0] * qc[n1]
            rp[c] = zero_edges(np.asfortranarray(r))
        # solve_phi (kmax = 1 without mapping): fphi = sigmap extrapol_p(phi) + div rhs_p
        fphi = 0.0 * self._ex(self.phi, self.top - 1, sub)
        fphi = fphi + self.der(rp["u"], 0) + self.der(rp["v"], 1) + self.der(rp["w"], 2)
        # navier_bc_pressure: rhs_p.n - fac1 bc.n, then the rot-rot term for pt=3
        bcphi = [
            np.asfortranarray(rp["u"][[0, -1]][:, 1:-1, 1:-1].transpose(1, 2, 0) - fac[0] * bc["u"][0]),
            np.asfortranarray(rp["v"][:, [0, -1]][1:-1, :, 1:-1].transpose(0, 2, 1) - fac[0] * bc["v"][1]),
            np.asfortranarray(rp["w"][:, :, [0, -1]][1:-1, 1:-1, :] - fac[0] * bc["w"][2]),
        ]
        if pt == 3:
            self.add_boundary_rotrot(bcphi, sub)
        o.solver_3d(self.scp, fphi, self.phi[n1], *bcphi, 0.0)
        # navier_projection (:1847-1959)
        zero_edges(self.phi[n1])
        pe = self._ex(self.p, self.top - 1, sub) if pt in (2, 4) else None
        p = self.phi[n1].copy()
        if pe is not None:
            p = p + pe
        zero_edges(p)
        if pt <= 2:
            div = (self.der(q["u"][n1], 0) + self.der(q["v"][n1], 1) + self.der(q["w"][n1], 2)) / re
            p = p - div
        elif pt == 4:
            div = (self.der(self._ex(q["u"], self.top - 1, sub), 0)
                   + self.der(self._ex(q["v"], self.top - 1, sub), 1)
                   + self.der(self._ex(q["w"], self.top - 1, sub), 2)) / re
            p = p - div
        self.p[n1] = zero_edges(p)
        for a, c in enumerate("uvw"):
            q[c][n1] = zero_edges((rp[c] - self.der(self.phi[n1], a)) / fac[0])
            self.nl[c][n1] = None
        # velocity correction (pt >= 3) with the new pressure
        if pt >= 3:
            self._velocity_solves(bc, sub)
        # testconv (navier3d.f90:366-412) from the second sub-iteration on
        done = False
        if sub > 1:
            done = self.testconv()
        self.sub_q = {c: q[c][n1].copy() for c in "uvw"}
        self.sub_p = self.p[n1].copy()
        return done

    def _velocity_solves(self, bc, sub):
        """navier_solve_u/v/w (:1597-1761): Re (fac2 q^n + fac3 q^n-1
        [+ fac4 q^n-2 for tou=3] + pressure gradient - rhs1), sigma = -Re fac1;
        the gradient is of extrapol_p(p) for pt=2 and of p(it(1)) for pt>=3."""
        it, fac, re, pt = self.it, self.fac, self.re, self.pt
        n1 = it[0]
        pg = None
        if pt == 2:
            pg = self._ex(self.p, self.top - 1, sub)
        elif pt >= 3:
            pg = self.p[n1]
        for a, c in enumerate("uvw"):
            q = self.q[c]
            g = fac[1] * q[it[NT - 1]] + fac[2] * q[it[NT - 2]]
            if self.tou == 3:
                g = g + fac[3] * q[it[NT - 3]]
            if pg is not None:
                g = g + self.der(pg, a)
            g = re * (g - self.rhs1[c])
            o.solver_3d(self.scu, g, q[n1], *bc[c], self.sigmau)
            self.nl[c][n1] = None
            if self.velocity_hook is not None:
                self.velocity_hook(q[n1])

    def add_boundary_rotrot(self, bcphi, sub):
        """add_boundary_rotrot (:793-846): on the faces normal to x,
        bcphi -= (d_x(d_y ev + d_z ew) - dd_y eu - dd_z eu)/Re, e. the
        extrapol_p velocities; likewise on the y and z faces."""
        e = [self._ex(self.q[c], self.top - 1, sub) for c in "uvw"]
        re = self.re
        for a in range(3):
            b, c = [x for x in range(3) if x != a]
            t = self.der(e[b], b) + self.der(e[c], c)
            t = self.der(t, a) - self.dder(e[a], b) - self.dder(e[a], c)
            if a == 0:
                face = np.stack([t[0, 1:-1, 1:-1], t[-1, 1:-1, 1:-1]], -1)
            elif a == 1:
                face = np.stack([t[1:-1, 0, 1:-1], t[1:-1, -1, 1:-1]], -1)
            else:
                face = np.stack([t[1:-1, 1:-1, 0], t[1:-1, 1:-1, -1]], -1)
            bcphi[a] -= face / re

    def testconv(self, eps: float = 1e-9) -> bool:
        """testconv (navier3d.f90:366-412): relative change of |q| and of p
        (mean removed) since the previous sub-iteration, both below eps.
        norme2 is sqrt(integral of x^2) with the integral weights."""
        n1 = self.it[0]
        W = self.W
        q = [self.q[c][n1] for c in "uvw"]
        s = [self.sub_q[c] for c in "uvw"]
        norm = lambda x: np.sqrt((W * x * x).sum())
        ref = norm(np.sqrt(q[0] ** 2 + q[1] ** 2 + q[2] ** 2))
        e1 = norm(np.sqrt((s[0] - q[0]) ** 2 + (s[1] - q[1]) ** 2 + (s[2] - q[2]) ** 2)) / ref
        p = self.p[n1]
        sp_ = self.sub_p + (W * (p - self.sub_p)).sum() / W.sum()
        e2 = norm(p - sp_) / norm(sp_)
        self.conv = (e1, e2)
        return bool(e1 < eps and e2 < eps)

    # ---- accessors (latest level)
    def velocity(self):
        L = self.it[NT - 1]
        return self.q["u"][L], self.q["v"][L], self.q["w"][L]

    def pressure(self):
        return self.p[self.it[NT - 1]]

    def phi_latest(self):
        return self.phi[self.it[NT - 1]]
