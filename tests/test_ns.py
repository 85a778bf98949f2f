"""Navier-Stokes prediction-projection step: the oracle pinned by closed forms
and invariants (CPU), then the device step against the oracle (GPU).

Oracle pins (each fails for a plausible restatement mistake):
* BDF factors differentiate t^k exactly for k <= tou, and the extrapolation of
  order k reproduces polynomials of degree k-1 (wrong factor or index fails);
* the nonlinear operator equals the analytic advective (nlt=1) and
  skew-symmetric (nlt=2) forms on smooth non-solenoidal fields, to the
  compact schemes' accuracy (a dropped 1/2 or product term fails);
* integral weights equal a direct loop over the reference's four loops;
* the laminar channel (parabola, body force 2/Re, or the flow-rate force)
  is a discrete steady state to 1e-12 (sign of the viscous term, sigma,
  Re scaling, BDF factors, pressure coupling);
* a 3-D manufactured steady flow whose advection is not a gradient is kept
  to the scheme's splitting error, which falls with the grid, while the same
  run with the advective forcing's sign flipped is off by O(1);
* the projection removes most of the divergence of a perturbed field;
* the LES force of u = (y^2, 0, 0) on a uniform grid is the closed form
  (8 c^2 Delta^2 y, 0, 0) for Smagorinsky (the compact schemes differentiate
  these polynomials exactly; a wrong index, sign, factor or width fails) and
  zero for WALE (pure shear).
"""
import numpy as np
import pytest

import mf_oracle as o
import ns_oracle as ns

sp = pytest.importorskip("sympy")


# ------------------------------------------------------------------ oracle --

@pytest.mark.parametrize("tou", [1, 2, 3])
def test_bdf_factors_and_extrapolation(tou):
    dt = 0.01
    fac = ns.bdf_factors(tou, dt)
    t = np.array([1.0, 1.0 - dt, 1.0 - 2 * dt, 1.0 - 3 * dt])  # t^{n+1}, t^n, t^{n-1}, t^{n-2}
    for k in range(tou + 1):  # d/dt t^k at t = 1 is k
        d = sum(fac[m] * t[m] ** k for m in range(4))
        assert abs(d - k) < 1e-9 * max(1, k)
    # and BDF-tou is not exact one degree higher
    k = tou + 1
    assert abs(sum(fac[m] * t[m] ** k for m in range(4)) - k) > 1e-6
    # extrapolation of order tou to t^{n+1} from levels it(nt), it(nt-1), it(nt-2)
    it = [3, 0, 1, 2]  # it(1) = 3 (new), it(nt) = 2, it(nt-1) = 1, it(nt-2) = 0
    for deg in range(tou):
        levels = [np.full((3, 3, 3), 0.0) for _ in range(4)]
        levels[2][:] = (1.0 - dt) ** deg
        levels[1][:] = (1.0 - 2 * dt) ** deg
        levels[0][:] = (1.0 - 3 * dt) ** deg
        e = ns.extrapol(levels, it, tou)
        assert abs(e[1, 1, 1] - 1.0) < 1e-12
    assert ns.extrapol([np.ones((3, 3, 3))] * 4, it, 0).max() == 0.0


def _fields(n=24):
    g = o.grid_uniform(n, 0.0, 1.0)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    x, y, z = sp.symbols("x y z")
    U = [sp.sin(x) * sp.cos(2 * y) + z ** 2, sp.exp(x / 2) * sp.sin(z) + y * x,
         sp.cos(x + y) * (1 + z) + sp.sin(3 * y)]   # not divergence-free
    ev = lambda e: np.asfortranarray(np.broadcast_to(
        sp.lambdify((x, y, z), e, "numpy")(X, Y, Z), X.shape).astype(float))
    return g, (x, y, z), U, ev


@pytest.mark.parametrize("nlt", [1, 2])
def test_nonlinear_operator(nlt):
    g, X, U, ev = _fields()
    s = ns.NSOracle(g, g, g, 1.0, 0.01, nlt=nlt)
    s.set_state(*[ev(e) for e in U])
    inner = (slice(3, -3),) * 3
    for c, name in enumerate("uvw"):
        got = s.nonlinear(0, name)
        q = U[c]
        adv = sum(U[a] * sp.diff(q, X[a]) for a in range(3))
        cons = sum(sp.diff(q * U[a], X[a]) for a in range(3))
        want = ev(adv if nlt == 1 else (adv + cons) / 2)
        other = ev(cons if nlt == 1 else adv)
        scale = np.abs(want[inner]).max()
        err = np.abs(got - want)[inner].max() / scale
        assert err < 1e-5, (name, err)
        # the fields are not solenoidal, so the two forms differ
        assert np.abs(got - other)[inner].max() / scale > 1e-2


def test_integral_weights():
    gx = np.array([0.0, 0.1, 0.3, 0.6, 1.0])
    gy = np.array([-1.0, -0.5, 0.2, 0.4, 0.9, 1.0])
    gz = np.array([0.0, 0.2, 0.5, 0.55, 0.9, 1.3, 2.0])
    W = ns.integral_weights(gx, gy, gz)
    f = np.random.default_rng(3).standard_normal(W.shape)
    nx, ny, nz = W.shape
    tot = 0.0
    for k in range(1, nz - 1):           # interior (class_navier_3D.f90:2437-2446)
        for j in range(1, ny - 1):
            for i in range(1, nx - 1):
                tot += f[i, j, k] * (gx[i + 1] - gx[i - 1]) * (gy[j + 1] - gy[j - 1]) * (gz[k + 1] - gz[k - 1]) / 8
    for j in range(1, ny - 1):           # x faces
        for k in range(1, nz - 1):
            dyz = (gy[j + 1] - gy[j - 1]) * (gz[k + 1] - gz[k - 1]) / 4
            tot += f[0, j, k] * dyz * (gx[1] - gx[0]) / 2 + f[-1, j, k] * dyz * (gx[-1] - gx[-2]) / 2
    for i in range(1, nx - 1):           # y faces
        for k in range(1, nz - 1):
            dxz = (gx[i + 1] - gx[i - 1]) * (gz[k + 1] - gz[k - 1]) / 4
            tot += f[i, 0, k] * dxz * (gy[1] - gy[0]) / 2 + f[i, -1, k] * dxz * (gy[-1] - gy[-2]) / 2
    for i in range(1, nx - 1):           # z faces
        for j in range(1, ny - 1):
            dxy = (gx[i + 1] - gx[i - 1]) * (gy[j + 1] - gy[j - 1]) / 4
            tot += f[i, j, 0] * dxy * (gz[1] - gz[0]) / 2 + f[i, j, -1] * dxy * (gz[-1] - gz[-2]) / 2
    assert abs((W * f).sum() - tot) < 1e-13 * max(1.0, abs(tot))


def _channel(n=(16, 24, 12), amp=0.1):
    from mflops import synthetic
    return synthetic.channel_problem(*n, amp=amp)


@pytest.mark.parametrize("flow", [False, True])
def test_laminar_channel_is_steady(flow):
    gx, gy, gz, u, v, w, bc = _channel()
    U = np.asfortranarray(np.broadcast_to(1.0 - gy[None, :, None] ** 2, u.shape).copy())
    re, dt = 3000.0, 1e-3
    s = ns.NSOracle(gx, gy, gz, re, dt)
    s.set_state(U, 0 * U, 0 * U)
    for c in range(3):
        s.bc["uvw"[c]] = bc[c]
    if flow:
        s.flow = ns.FlowRate(g0=2.0 / re)
    else:
        s.force["u"] = np.asfortranarray(np.full(U.shape, 2.0 / re))
    for _ in range(8):
        s.step()
    uu, vv, ww = s.velocity()
    inner = (slice(1, -1),) * 3
    assert np.abs(uu - U)[inner].max() < 1e-12
    assert max(np.abs(vv)[inner].max(), np.abs(ww)[inner].max()) < 1e-12
    if flow:
        assert abs(s.flow.g - 2.0 / re) < 1e-9


def _mms(n, sign, re=10.0, dt=1e-2, steps=30, pt=1):
    x, y, z = sp.symbols("x y z")
    X = (x, y, z)
    psi = sp.sin(sp.pi * x) ** 2 * sp.sin(sp.pi * y) ** 2 * sp.cos(sp.pi * z / 2) / sp.pi
    U = [sp.diff(psi, y), -sp.diff(psi, x), sp.Integer(0)]
    P = sp.cos(sp.pi * x) * sp.cos(sp.pi * y) * sp.sin(sp.pi * z)
    F = []
    for c in range(3):
        adv = sum(U[a] * sp.diff(U[c], X[a]) for a in range(3))
        lap = sum(sp.diff(U[c], X[a], 2) for a in range(3))
        F.append(sign * adv + sp.diff(P, X[c]) - lap / re)
    g = o.grid_uniform(n, 0.0, 1.0)
    XX, YY, ZZ = np.meshgrid(g, g, g, indexing="ij")
    ev = lambda e: np.asfortranarray(np.broadcast_to(
        sp.lambdify(X, e, "numpy")(XX, YY, ZZ), XX.shape).astype(float))
    u = [ev(e) for e in U]
    s = ns.NSOracle(g, g, g, re, dt, pt=pt)
    s.set_state(*u)
    for c in range(3):
        s.force["uvw"[c]] = ev(F[c])
        q = u[c]
        s.bc["uvw"[c]] = (np.asfortranarray(np.stack([q[0, 1:-1, 1:-1], q[-1, 1:-1, 1:-1]], -1)),
                          np.asfortranarray(np.stack([q[1:-1, 0, 1:-1], q[1:-1, -1, 1:-1]], -1)),
                          np.asfortranarray(np.stack([q[1:-1, 1:-1, 0], q[1:-1, 1:-1, -1]], -1)))
    for _ in range(steps):
        s.step()
    vel = s.velocity()
    inner = (slice(1, -1),) * 3
    return max(np.abs(vel[c] - u[c])[inner].max() for c in range(3)) / max(np.abs(q).max() for q in u)


@pytest.mark.slow
def test_manufactured_steady_flow():
    e16, e24 = _mms(16, 1), _mms(24, 1)
    flipped = _mms(16, -1)
    assert e16 < 0.02 and e24 < 0.6 * e16, (e16, e24)
    assert flipped > 0.1, flipped


def _face_bc(q):
    return (np.asfortranarray(np.stack([q[0, 1:-1, 1:-1], q[-1, 1:-1, 1:-1]], -1)),
            np.asfortranarray(np.stack([q[1:-1, 0, 1:-1], q[1:-1, -1, 1:-1]], -1)),
            np.asfortranarray(np.stack([q[1:-1, 1:-1, 0], q[1:-1, 1:-1, -1]], -1)))


def _unsteady_mms(n, pt, dt, T, nsubite=1, top=2, cls=None, re=10.0):
    """u = (psi_y, -psi_x, 0) (1 + 4 t^3), p = P 4 t^3 on [0, 1]^3, forced so
    that they solve the Navier-Stokes equations; the forcing and the boundary
    values are set at t^{n+1} before each step (f(nav,'rhsu') and
    navier_bc_velocity read the already advanced time).  Both time factors
    are flat at t = 0, so the constant start history and p = 0 are exact to
    O(dt^3).  Returns the oracle after T / dt steps and the exact velocity."""
    x, y, z, t = sp.symbols("x y z t")
    X = (x, y, z)
    psi = sp.sin(sp.pi * x) ** 2 * sp.sin(sp.pi * y) ** 2 * sp.cos(sp.pi * z / 2) / sp.pi
    U = [sp.diff(psi, y) * (1 + 4 * t ** 3), -sp.diff(psi, x) * (1 + 4 * t ** 3), sp.Integer(0)]
    P = sp.cos(sp.pi * x) * sp.cos(sp.pi * y) * sp.sin(sp.pi * z) * 4 * t ** 3
    F = [sp.diff(U[c], t) + sum(U[a] * sp.diff(U[c], X[a]) for a in range(3)) + sp.diff(P, X[c])
         - sum(sp.diff(U[c], X[a], 2) for a in range(3)) / re for c in range(3)]
    g = o.grid_uniform(n, 0.0, 1.0)
    XX, YY, ZZ = np.meshgrid(g, g, g, indexing="ij")
    fU = [sp.lambdify((x, y, z, t), e, "numpy") for e in U]
    fF = [sp.lambdify((x, y, z, t), e, "numpy") for e in F]
    ev = lambda f, tt: np.asfortranarray(np.broadcast_to(f(XX, YY, ZZ, tt), XX.shape).astype(float))
    s = (cls or ns.NSOracle)(g, g, g, re, dt, pt=pt, nsubite=nsubite, top=top)
    s.set_state(*[ev(f, 0.0) for f in fU])
    steps = int(round(T / dt))
    for k in range(steps):
        tn = (k + 1) * dt
        for c in range(3):
            s.force["uvw"[c]] = ev(fF[c], tn)
            s.bc["uvw"[c]] = _face_bc(ev(fU[c], tn))
        s.step()
    return s, [ev(f, steps * dt) for f in fU]


@pytest.mark.slow
@pytest.mark.parametrize("pt", [2, 3, 4])
def test_incremental_schemes_keep_the_steady_flow(pt):
    """pt = 2, 3, 4 carry the pressure from step to step, so the manufactured
    steady flow is kept to the spatial error, which falls with the grid, far
    below the non-incremental scheme's splitting error (7.7e-3 at 16^3)."""
    e16, e24 = _mms(16, 1, pt=pt), _mms(24, 1, pt=pt)
    assert e16 < 1e-3 and e24 < 0.3 * e16, (e16, e24)


class _RotRotFlipped(ns.NSOracle):
    """The oracle with a planted mistake: the rot-rot boundary term of pt=3
    added instead of subtracted."""

    def add_boundary_rotrot(self, bcphi, sub):
        before = [b.copy() for b in bcphi]
        super().add_boundary_rotrot(bcphi, sub)
        for a in range(3):
            bcphi[a][...] = 2.0 * before[a] - bcphi[a]


@pytest.mark.slow
@pytest.mark.parametrize("pt", [1, 2, 3, 4])
def test_unsteady_manufactured_flow_converges_in_time(pt):
    """Every projection type follows an unsteady exact solution: at 12^3 the
    error at t = 0.4 is at the spatial floor (~2e-3) for dt = 0.01, and the
    differences between dt = 0.04, 0.02, 0.01 shrink by more than 2 per
    halving.  A wrong increment (the rot-rot term's sign flipped) is off by
    more than 1e-2 (measured 2.9e-2)."""
    runs = [_unsteady_mms(12, pt, dt, 0.4) for dt in (0.04, 0.02, 0.01)]
    v = [np.stack(s.velocity()) for s, _ in runs]
    exact = np.stack(runs[-1][1])
    err = np.abs(v[-1] - exact).max() / np.abs(exact).max()
    ratio = np.abs(v[0] - v[1]).max() / np.abs(v[1] - v[2]).max()
    assert err < 3e-3 and ratio > 2.0, (err, ratio)
    if pt == 3:
        s, ex = _unsteady_mms(12, 3, 0.01, 0.4, cls=_RotRotFlipped)
        bad = np.abs(np.stack(s.velocity()) - np.stack(ex)).max() / np.abs(exact).max()
        assert bad > 1e-2, bad


@pytest.mark.parametrize("pt", [1, 3])
def test_subiterations_reach_a_fixed_point_independent_of_top(pt):
    """From the second sub-iteration on every extrapolation returns the
    current iterate, so converged sub-iterations solve a problem that no
    longer depends on the pressure extrapolation order top: top = 1, 2, 3
    agree to the 1e-9 stopping test, while single sub-iterations differ by
    more than 1e-4.  (pt = 2 and 4 accumulate phi - div(u*)/Re into p at
    every sub-iteration, which does not vanish, so their sub-iterations run
    to nsubite.)"""
    def run(top, nsub):
        s, _ = _unsteady_mms(12, pt, 0.02, 0.04, nsubite=nsub, top=top)
        W = s.W
        p = s.pressure() - (W * s.pressure()).sum() / W.sum()
        return np.stack(s.velocity()), p, s.subite_done
    one = [run(top, 1) for top in (1, 2, 3)]
    many = [run(top, 40) for top in (1, 2, 3)]
    spread = lambda r, k: max(np.abs(r[0][k] - r[m][k])[1:-1, 1:-1, 1:-1].max() for m in (1, 2)) \
        / np.abs(r[0][k]).max()
    assert spread(one, 0) > 1e-4
    assert spread(many, 0) < 1e-10 and spread(many, 1) < 1e-8, (spread(many, 0), spread(many, 1))
    assert all(1 < r[2] < 40 for r in many)


def test_projection_reduces_divergence():
    gx, gy, gz, u, v, w, bc = _channel()
    s = ns.NSOracle(gx, gy, gz, 3000.0, 1e-3)
    s.set_state(u, v, w)
    for c in range(3):
        s.bc["uvw"[c]] = bc[c]
    s.flow = ns.FlowRate(g0=2.0 / 3000.0)
    d0 = s.der(u, 0) + s.der(v, 1) + s.der(w, 2)
    for _ in range(3):
        s.step()
    uu, vv, ww = s.velocity()
    d = s.der(uu, 0) + s.der(vv, 1) + s.der(ww, 2)
    inner = (slice(2, -2),) * 3
    assert np.abs(d[inner]).max() < 0.2 * np.abs(d0[inner]).max()


def test_oracle_les_force_closed_form():
    gx, gy, gz = o.grid_uniform(10, 0.0, 1.0), o.grid_uniform(12, 0.1, 1.0), o.grid_uniform(9, 0.0, 0.8)
    s = ns.NSOracle(gx, gy, gz, 100.0, 1e-3)
    X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
    zero = np.zeros(X.shape)
    s.set_state(np.asfortranarray(Y ** 2), zero, zero)
    cs = 0.165
    d2 = ((gx[1] - gx[0]) ** 2 + (gy[1] - gy[0]) ** 2 + (gz[1] - gz[0]) ** 2) / 3.0
    inner = (slice(1, -1),) * 3
    s.set_les(1, cs)
    f = s.les_force(0)
    ref = 8.0 * cs ** 2 * d2 * Y
    assert np.abs((f["u"] - ref)[inner]).max() < 1e-12 * np.abs(ref).max()
    assert np.abs(f["v"][inner]).max() < 1e-12 * np.abs(ref).max()
    assert np.abs(f["w"][inner]).max() < 1e-12 * np.abs(ref).max()
    assert np.abs((s.les_nu - 2.0 * cs ** 2 * d2 * Y)[inner]).max() < 1e-13
    s.set_les(2, 0.5)
    f = s.les_force(0)
    assert max(np.abs(f[c]).max() for c in "uvw") == 0.0


# --------------------------------------------------------------------- GPU --

def _gpu_ns(mf, gx, gy, gz, u, v, w, bc, re, dt, **kw):
    s = mf.NavierStokes(gx, gy, gz, re, dt, **kw)
    for c in range(3):
        s.set_bc(c, mf.BoundaryCondition(*[mf.to_device(b) for b in bc[c]]))
    s.set_state(*[mf.to_device(f) for f in (u, v, w)])
    return s


@pytest.mark.gpu
@pytest.mark.parametrize("tou,top", [(2, 2), (3, 3), (1, 1)])
def test_ns_steps_match_oracle(tou, top):
    import mflops as mf
    import torch
    assert torch.cuda.is_available()
    mf.load()
    gx, gy, gz, u, v, w, bc = _channel((20, 24, 18))
    re, dt, steps = 3000.0, 1e-3, 5
    print("run ns tou=%d top=%d shape=%s steps=%d" % (tou, top, u.shape, steps), flush=True)
    orc = ns.NSOracle(gx, gy, gz, re, dt, tou=tou, top=top)
    orc.set_state(u, v, w)
    for c in range(3):
        orc.bc["uvw"[c]] = bc[c]
    orc.flow = ns.FlowRate(g0=2.0 / re)
    for _ in range(steps):
        orc.step()
    dev = _gpu_ns(mf, gx, gy, gz, u, v, w, bc, re, dt, tou=tou, top=top)
    dev.set_flow_rate(True, 2.0 / re)
    dev.step(steps, mf.MF_FP64)
    got = [mf.to_host(t) for t in dev.get()]
    want = list(orc.velocity()) + [orc.pressure(), orc.phi_latest()]
    for name, a, b in zip(("u", "v", "w", "p", "phi"), got, want):
        scale = max(np.abs(b).max(), 1e-30)
        err = np.abs(a - b).max() / scale
        print("  %s: max|gpu-oracle|/max|oracle| = %.2e" % (name, err), flush=True)
        assert err < 1e-10, (name, err)
    t, g = dev.info()
    assert abs(t - steps * dt) < 1e-12
    assert abs(g - orc.flow.g) < 1e-10 * abs(orc.flow.g)


@pytest.mark.gpu
def test_ns_advective_form_matches_oracle():
    """nlt=1 (advective nonlinear term) against the oracle."""
    import mflops as mf
    mf.load()
    gx, gy, gz, u, v, w, bc = _channel((20, 24, 18))
    re, dt, steps = 3000.0, 1e-3, 4
    print("run ns nlt=1 shape=%s steps=%d" % (u.shape, steps), flush=True)
    orc = ns.NSOracle(gx, gy, gz, re, dt, nlt=1)
    orc.set_state(u, v, w)
    for c in range(3):
        orc.bc["uvw"[c]] = bc[c]
    orc.flow = ns.FlowRate(g0=2.0 / re)
    for _ in range(steps):
        orc.step()
    dev = _gpu_ns(mf, gx, gy, gz, u, v, w, bc, re, dt, nlt=1)
    dev.set_flow_rate(True, 2.0 / re)
    dev.step(steps, mf.MF_FP64)
    _compare(mf, dev, orc, 1e-10, "nlt=1")


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [0, 3], ids=["fp64", "fp16x2"])
def test_ns_concurrent_velocity_solves_match_sequential(monkeypatch, prec):
    """The three velocity solves on three handles and streams (the default)
    against one handle in turn (MF_NS_CONCURRENT=0): the same kernels on the
    same data, so the fields agree bit for bit."""
    import mflops as mf
    mf.load()
    gx, gy, gz, u, v, w, bc = _channel((40, 48, 36))
    re, dt, steps = 3000.0, 1e-3, 3
    runs = {}
    for conc in ("1", "0"):
        print("run ns concurrent=%s precision=%d" % (conc, prec), flush=True)
        monkeypatch.setenv("MF_NS_CONCURRENT", conc)
        dev = _gpu_ns(mf, gx, gy, gz, u, v, w, bc, re, dt)
        dev.set_flow_rate(True, 2.0 / re)
        dev.step(steps, prec, tol=1e-13)
        runs[conc] = [mf.to_host(t) for t in dev.get()]
    for name, a, b in zip(("u", "v", "w", "p", "phi"), runs["1"], runs["0"]):
        assert np.array_equal(a, b), (name, np.abs(a - b).max())


@pytest.mark.gpu
@pytest.mark.parametrize("tou", [1, 2, 3])
def test_ns_fused_right_hand_sides_match_separate(monkeypatch, tou):
    """The first presolve's velocity right-hand sides formed inside the
    nonlinear-term kernel (pt = 1, the default) against rhs_kernel
    (MF_NS_FUSED_RHS=0): the same arithmetic in the same order, so the
    fields agree bit for bit."""
    import mflops as mf
    mf.load()
    gx, gy, gz, u, v, w, bc = _channel((40, 48, 36))
    re, dt, steps = 3000.0, 1e-3, 4
    runs = {}
    for fused in ("1", "0"):
        print("run ns fused_rhs=%s tou=%d" % (fused, tou), flush=True)
        monkeypatch.setenv("MF_NS_FUSED_RHS", fused)
        dev = _gpu_ns(mf, gx, gy, gz, u, v, w, bc, re, dt, tou=tou, top=tou)
        dev.set_flow_rate(True, 2.0 / re)
        dev.step(steps, mf.MF_FP64)
        runs[fused] = [mf.to_host(t) for t in dev.get()]
    for name, a, b in zip(("u", "v", "w", "p", "phi"), runs["1"], runs["0"]):
        assert np.array_equal(a, b), (name, np.abs(a - b).max())


def _orc_fields(orc):
    return list(orc.velocity()) + [orc.pressure(), orc.phi_latest()]


def _rel_errs(got, want):
    return {name: np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
            for name, a, b in zip(("u", "v", "w", "p", "phi"), got, want)}


def _compare(mf, dev, orc, tol, label):
    """Fields of dev against the oracle, relative to the oracle's maximum;
    tol is a number or a dict per field."""
    errs = _rel_errs([mf.to_host(t) for t in dev.get()], _orc_fields(orc))
    print("  %s: %s" % (label, {k: "%.2e" % v for k, v in errs.items()}), flush=True)
    bar = tol if isinstance(tol, dict) else {k: tol for k in errs}
    assert all(errs[k] < bar[k] for k in errs), (errs, bar)


@pytest.mark.gpu
@pytest.mark.parametrize("pt,tou,top", [(2, 2, 2), (3, 2, 2), (4, 2, 2), (3, 3, 3), (4, 3, 3),
                                        (2, 1, 1)])
def test_ns_projection_types_match_oracle(pt, tou, top):
    """Projection types 2 (incremental), 3 (velocity correction with the
    rot-rot pressure boundary term) and 4 (incremental velocity correction)
    over 5 steps of the perturbed channel with the flow-rate force: every
    field within 1e-10 of the oracle (relative to its maximum)."""
    import mflops as mf
    import torch
    assert torch.cuda.is_available()
    mf.load()
    gx, gy, gz, u, v, w, bc = _channel((20, 24, 18))
    re, dt, steps = 3000.0, 1e-3, 5
    print("run ns pt=%d tou=%d top=%d shape=%s steps=%d" % (pt, tou, top, u.shape, steps), flush=True)
    orc = ns.NSOracle(gx, gy, gz, re, dt, tou=tou, top=top, pt=pt)
    orc.set_state(u, v, w)
    for c in range(3):
        orc.bc["uvw"[c]] = bc[c]
    orc.flow = ns.FlowRate(g0=2.0 / re)
    for _ in range(steps):
        orc.step()
    dev = _gpu_ns(mf, gx, gy, gz, u, v, w, bc, re, dt, tou=tou, top=top, pt=pt)
    dev.set_flow_rate(True, 2.0 / re)
    dev.step(steps, mf.MF_FP64)
    _compare(mf, dev, orc, 1e-10, "pt=%d tou=%d top=%d" % (pt, tou, top))
    t, g = dev.info()
    assert abs(g - orc.flow.g) < 1e-10 * abs(orc.flow.g)


@pytest.mark.gpu
@pytest.mark.parametrize("pt,nsub,les", [(1, 20, 0), (3, 20, 0), (2, 4, 0), (4, 4, 1), (3, 20, 1)])
def test_ns_subiterations_match_oracle(pt, nsub, les):
    """Sub-iterations (nsubite) on the channel with a uniform body force:
    pt = 1 and 3 stop on the convergence test (after 7 and 5 sub-iterations
    in the oracle), pt = 2 and 4 run to nsubite; 3 steps, the same number of
    sub-iterations, the last test's changes alike (to 1e-6 where they are far
    above rounding), fields within 1e-10 of the oracle.  The sub-iterations
    of pt = 2 and 4 do not contract (each moves p by 20-50%), so they amplify
    rounding: there the bar of each field is the larger of 1e-10 and three
    times how far the oracle itself moves when every velocity solve's result
    is perturbed by 1e-13 (relative, random; one draw, while the GPU's
    rounding differs from the oracle's in every operation, not only the
    solves: measured GPU/oracle distances reach 1.7x that movement)."""
    import mflops as mf
    import torch
    assert torch.cuda.is_available()
    mf.load()
    gx, gy, gz, u, v, w, bc = _channel((20, 24, 18))
    re, dt, steps = 3000.0, 1e-3, 3
    print("run ns subiterations pt=%d nsubite=%d les=%d steps=%d" % (pt, nsub, les, steps), flush=True)
    force = np.asfortranarray(np.full(u.shape, 2.0 / re))
    orc = ns.NSOracle(gx, gy, gz, re, dt, pt=pt, nsubite=nsub)
    orc.set_state(u, v, w)
    for c in range(3):
        orc.bc["uvw"[c]] = bc[c]
    orc.force["u"] = force
    if les:
        orc.set_les(1, 0.165)
    dev = _gpu_ns(mf, gx, gy, gz, u, v, w, bc, re, dt, pt=pt, nsubite=nsub)
    dev.set_forcing(0, mf.to_device(force))
    if les:
        dev.set_les(1, 0.165)
    for k in range(steps):
        orc.step()
        dev.step(1, mf.MF_FP64)
        n, dv, dp = dev.subiterations()
        print("  step %d: sub-iterations gpu %d oracle %d, changes gpu (%.2e, %.2e) oracle (%.2e, %.2e)"
              % (k, n, orc.subite_done, dv, dp, *orc.conv), flush=True)
        assert n == orc.subite_done
        for a, b in zip((dv, dp), orc.conv):
            if b > 1e-6:
                assert abs(a - b) <= 1e-6 * b, (a, b)
            else:
                assert a < 1e-9 and b < 1e-9 or n == nsub
    bar = 1e-10
    if pt in (2, 4):
        pert = ns.NSOracle(gx, gy, gz, re, dt, pt=pt, nsubite=nsub)
        pert.set_state(u, v, w)
        for c in range(3):
            pert.bc["uvw"[c]] = bc[c]
        pert.force["u"] = force
        if les:
            pert.set_les(1, 0.165)
        rng = np.random.default_rng(1)

        def hook(q):
            q *= 1.0 + 1e-13 * rng.standard_normal(q.shape)

        pert.velocity_hook = hook
        for _ in range(steps):
            pert.step()
        sens = _rel_errs(_orc_fields(pert), _orc_fields(orc))
        print("  oracle moved by 1e-13 perturbations: %s" % {k: "%.2e" % v for k, v in sens.items()},
              flush=True)
        bar = {k: max(1e-10, 3.0 * v) for k, v in sens.items()}
    _compare(mf, dev, orc, bar, "pt=%d nsubite=%d" % (pt, nsub))


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [1, 2, 3])
def test_ns_mixed_velocity_solves(prec):
    import mflops as mf
    mf.load()
    gx, gy, gz, u, v, w, bc = _channel((20, 24, 18))
    re, dt, steps = 3000.0, 1e-3, 5
    print("run ns mixed precision=%d" % prec, flush=True)
    a = _gpu_ns(mf, gx, gy, gz, u, v, w, bc, re, dt)
    b = _gpu_ns(mf, gx, gy, gz, u, v, w, bc, re, dt)
    for s in (a, b):
        s.set_flow_rate(True, 2.0 / re)
    a.step(steps, mf.MF_FP64)
    b.step(steps, prec, tol=1e-13)
    for name, x, y in zip(("u", "v", "w", "p"), a.get()[:4], b.get()[:4]):
        xa, ya = mf.to_host(x), mf.to_host(y)
        err = np.abs(xa - ya).max() / max(np.abs(xa).max(), 1e-30)
        print("  %s: mixed vs fp64 %.2e" % (name, err), flush=True)
        assert err < 1e-10, (name, err)


@pytest.mark.gpu
@pytest.mark.parametrize("les_type,cs", [(1, 0.165), (2, 0.5), (3, 1.35)])
def test_ns_les_matches_oracle(les_type, cs):
    import mflops as mf
    mf.load()
    gx, gy, gz, u, v, w, bc = _channel((20, 24, 18))
    re, dt, steps = 3000.0, 1e-3, 3
    print("run ns les_type=%d steps=%d" % (les_type, steps), flush=True)
    orc = ns.NSOracle(gx, gy, gz, re, dt)
    orc.set_state(u, v, w)
    for c in range(3):
        orc.bc["uvw"[c]] = bc[c]
    orc.flow = ns.FlowRate(g0=2.0 / re)
    orc.set_les(les_type, cs)
    for _ in range(steps):
        orc.step()
    dev = _gpu_ns(mf, gx, gy, gz, u, v, w, bc, re, dt)
    dev.set_flow_rate(True, 2.0 / re)
    dev.set_les(les_type, cs)
    dev.step(steps, mf.MF_FP64)
    got = [mf.to_host(t) for t in dev.get()[:4]] + [mf.to_host(dev.les_nu())]
    want = list(orc.velocity()) + [orc.pressure(), orc.les_nu]
    assert np.abs(orc.les_nu).max() > 0
    # the sigma model's s3 = sqrt(i1 - s1 - s2) carries eps i1 / (2 s3^2)
    # relative error, large where the flow is nearly two-dimensional (the
    # channel's base shear): nu itself (proportional to s3) agrees to about
    # sqrt(eps) of its largest value, the flow it drives to 1e-7
    tol = {"u": 1e-10, "v": 1e-10, "w": 1e-10, "p": 1e-10, "nu": 1e-10}
    if les_type == 3:
        tol = {"u": 1e-7, "v": 1e-7, "w": 1e-7, "p": 1e-7, "nu": 1e-6}
    errs = {}
    for name, a, b in zip(("u", "v", "w", "p", "nu"), got, want):
        errs[name] = np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)
        print("  %s: max|gpu-oracle|/max|oracle| = %.2e" % (name, errs[name]), flush=True)
    assert all(errs[k] < tol[k] for k in errs), errs


def test_oracle_minus_div_tau_closed_form():
    """The learned closure's force -sum_j d_j tau_ij on cubic stress fields,
    which every compact row differentiates exactly: equal to the analytic
    divergence at interior points (any swapped column, index or sign
    fails)."""
    shape = (12, 14, 10)
    gx, gy, gz = o.grid_uniform(12, 0, 1), o.grid_tanh(14, -1, 1, 1.5), o.grid_asin(10, 0, 2)
    s = ns.NSOracle(gx, gy, gz, 100.0, 1e-3)
    X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
    # columns xx, yy, zz, xy, xz, yz and their gradients (d/dx, d/dy, d/dz)
    tau = [X ** 3 + Y, Y ** 2 * Z, Z ** 3 - X * Y, X * Y * Z, X ** 2 * Z, Y ** 3 + X]
    grad = [(3 * X ** 2, 1 + 0 * X, 0 * X), (0 * X, 2 * Y * Z, Y ** 2), (-Y, -X, 3 * Z ** 2),
            (Y * Z, X * Z, X * Y), (2 * X * Z, 0 * X, X ** 2), (1 + 0 * X, 3 * Y ** 2, 0 * X)]
    cols = {(0, 0): 0, (1, 1): 1, (2, 2): 2, (0, 1): 3, (0, 2): 4, (1, 2): 5}
    f = s.minus_div_tau([np.asfortranarray(t) for t in tau])
    inner = (slice(1, -1),) * 3
    for i, c in enumerate("uvw"):
        ref = -sum(grad[cols[(min(i, j), max(i, j))]][j] for j in range(3))
        err = np.abs(f[c] - ref)[inner].max() / np.abs(ref[inner]).max()
        assert err < 1e-11, (c, err)
