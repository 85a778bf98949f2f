"""Multidomain (influence-matrix) solve: oracle pins (CPU) and GPU parity.

The oracle (oracle/md_oracle.py) is pinned by what the mathematics fixes:
* a polynomial that the compact schemes differentiate exactly (degree 3 in
  each variable; every row is exact to degree 5 or more) is reproduced on
  every subdomain: any error in the interface numbering, the sign rules of
  md_influence_matrix_set / multidomain_solve or the assembly would show;
* after the solve the flux jumps (sum over both sides of the outward-normal
  first derivatives) vanish, for a non-polynomial solution too;
* with one subdomain there are no interfaces and the solve is the plain one;
* periodic layouts and the null-space options: see the oracle tests below.
The GPU solve (mf_md_*) is compared with the oracle on the same layouts: the
influence matrix through M x, the interface values and every subdomain
field, to 1e-9 relative (GMRES stops at 1e-12).
"""
import numpy as np
import pytest

import mf_oracle as o
import md_oracle as md


def layout_small():
    # 2 x 2 x 1 subdomains, uneven point counts, one stretched slab
    return [[o.grid_uniform(10, 0, 0.5), o.grid_uniform(12, 0.5, 1.0)],
            [o.grid_uniform(11, 0, 0.4), o.grid_asin(10, 0.4, 1.0)],
            [o.grid_uniform(10, 0, 1)]]


def layout_3d():
    # 2 x 1 x 2 subdomains with interfaces along x and z
    return [[o.grid_uniform(11, 0, 0.3), o.grid_asin(12, 0.3, 1.0)],
            [o.grid_uniform(10, 0, 1)],
            [o.grid_uniform(10, 0, 0.6), o.grid_uniform(11, 0.6, 1.0)]]


def poly(X, Y, Z):
    return X ** 3 + 2 * Y ** 3 - Z ** 2 + X * Y * Z + 0.5 * X * Y


def poly_lap(X, Y, Z):
    return 6 * X + 12 * Y - 2


def sines(X, Y, Z):
    return np.sin(np.pi * X) * np.cos(0.5 * np.pi * Y) * np.sin(np.pi * Z + 0.3)


def sines_lap(X, Y, Z):
    return -(np.pi ** 2 + 0.25 * np.pi ** 2 + np.pi ** 2) * sines(X, Y, Z)


def _exact_err(L, u, fun):
    err = 0.0
    for s in L.subs:
        gx, gy, gz = (np.asarray(L.grids[ax][s[ax]]) for ax in range(3))
        X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
        err = max(err, np.abs(u[s] - fun(X, Y, Z))[1:-1, 1:-1, 1:-1].max())
    return err


@pytest.mark.parametrize("lay,sigma", [(layout_small, 2.0), (layout_3d, 0.0), (layout_3d, -30.0)])
def test_oracle_reproduces_polynomial(lay, sigma):
    L = md.Layout(lay())
    mdo = md.Multidomain(L, sigma)
    g, bc = md.manufactured(L, poly, poly_lap, sigma)
    x, u = mdo.solve(g, bc)
    err = _exact_err(L, u, poly)
    print("ninf %d  max |u - poly| %.2e" % (L.ninf, err))
    assert err < 1e-10


def test_oracle_flux_jumps_vanish():
    L = md.Layout(layout_small())
    sigma = 5.0
    mdo = md.Multidomain(L, sigma)
    g, bc = md.manufactured(L, sines, sines_lap, sigma)
    x, u = mdo.solve(g, bc)
    r = mdo.rhs(g, bc)
    j = mdo.flux_jumps(u)
    err = _exact_err(L, u, sines)
    print("flux jumps %.2e (rhs %.2e), discretisation error %.2e" % (np.abs(j).max(), np.abs(r).max(), err))
    assert np.abs(j).max() < 1e-10 * np.abs(r).max()
    assert err < 1e-3  # discretisation level on 10-12 points per slab
    # the solve with zero interface values leaves jumps of the size of rhs
    u0 = {}
    for s in L.subs:
        u0[s] = np.zeros(L.shape(s), order="F")
        o.solver_3d(mdo.sc[s], g[s], u0[s], *mdo._bc_with(s, bc[s], None), sigma)
    assert np.abs(mdo.flux_jumps(u0) - r).max() < 1e-12 * np.abs(r).max()


def test_oracle_single_domain_is_plain_solve():
    grids = [[o.grid_uniform(12)], [o.grid_asin(11, 0, 1)], [o.grid_uniform(10)]]
    L = md.Layout(grids)
    assert L.ninf == 0
    mdo = md.Multidomain(L, 3.0)
    g, bc = md.manufactured(L, sines, sines_lap, 3.0)
    _, u = mdo.solve(g, bc)
    s = (0, 0, 0)
    sc = o.SolverCoeffs3D(*[grids[a][0] for a in range(3)], (1,) * 6, 6)
    ref = np.zeros(L.shape(s), order="F")
    o.solver_3d(sc, g[s], ref, *bc[s], 3.0)
    assert np.abs(u[s] - ref).max() == 0.0


def layout_periodic(n=12):
    # periodic along x (2 slabs, the wrap interface joins x = 1 to x = 0) and
    # z (one slab whose two faces are one interface), walls along y
    return [[o.grid_uniform(n, 0, 0.5), o.grid_uniform(n, 0.5, 1.0)],
            [o.grid_asin(n, -1.0, 1.0)],
            [o.grid_uniform(n + 1, 0, 1.0)]]


def trig(X, Y, Z):
    return np.sin(2 * np.pi * X + 0.4) * np.cos(0.5 * np.pi * Y) * np.cos(2 * np.pi * Z) + 0.3 * np.cos(2 * np.pi * X)


def trig_lap(X, Y, Z):
    a = np.sin(2 * np.pi * X + 0.4) * np.cos(0.5 * np.pi * Y) * np.cos(2 * np.pi * Z)
    return -(8.25 * np.pi ** 2) * a - 0.3 * 4 * np.pi ** 2 * np.cos(2 * np.pi * X)


def neumann_trig(X, Y, Z):
    # d/dy = 0 at y = -1, 1: a compatible all-Neumann/periodic pressure
    return np.cos(2 * np.pi * X) * np.cos(np.pi * Y) * np.cos(2 * np.pi * Z + 0.2) + np.sin(2 * np.pi * X) * np.cos(np.pi * Y)


def neumann_trig_lap(X, Y, Z):
    a = np.cos(2 * np.pi * X) * np.cos(np.pi * Y) * np.cos(2 * np.pi * Z + 0.2)
    b = np.sin(2 * np.pi * X) * np.cos(np.pi * Y)
    return -9.0 * np.pi ** 2 * a - 5.0 * np.pi ** 2 * b


def _neumann_data(L, fun):
    """Outer-face arrays with the exact normal derivative on the y walls (the
    solver's Neumann data, class_solver_3d.f90: du/dn values), interface
    faces left for the solve."""
    g, bc = md.manufactured(L, fun, lambda X, Y, Z: 0.0 * X, 0.0)
    h = 1e-6
    for s in L.subs:
        gx, gy, gz = (np.asarray(L.grids[ax][s[ax]]) for ax in range(3))
        X, Y, Z = np.meshgrid(gx[1:-1], gy[[0, -1]], gz[1:-1], indexing="ij")
        dy = (fun(X, Y + h, Z) - fun(X, Y - h, Z)) / (2 * h)
        bcx, bcy, bcz = bc[s]
        bcy = np.asfortranarray(dy.transpose(0, 2, 1))
        bc[s] = (bcx, bcy, bcz)
    return bc


@pytest.mark.parametrize("sigma", [-5.0, 3.0])
def test_oracle_periodic_converges(sigma):
    """Periodic along x and z: the solution of a periodic trigonometric
    problem converges with the grid (error below 5e-3 at 12 points per slab
    and falling by > 6 from 12 to 20, as the compact boundary closures
    allow; the same solution with Dirichlet data on every face falls by 9.5),
    and the flux jumps vanish at every interface, the wrap interfaces
    included."""
    errs = []
    for n in (12, 20):
        L = md.Layout(layout_periodic(n), bct=(1, 1, 1, 1, 1, 1), periods=(True, False, True))
        assert L.ninf == 2 * (n - 2) * (n - 1) + 2 * (n - 2) * (n - 2)
        mdo = md.Multidomain(L, sigma)
        g, bc = md.manufactured(L, trig, trig_lap, sigma)
        x, u = mdo.solve(g, bc)
        j = mdo.flux_jumps(u)
        assert np.abs(j).max() < 1e-10 * np.abs(mdo.rhs(g, bc)).max()
        errs.append(_exact_err(L, u, trig))
    print("periodic errors", errs)
    assert errs[0] < 5e-3 and errs[0] / errs[1] > 6.0, errs


def test_oracle_periodic_shift():
    """Data shifted by one slab along the periodic x axis give subdomain
    solutions shifted by one slab (x values differ by 0.5 only through the
    slabs' grids, so to rounding)."""
    grids = layout_periodic(12)
    L = md.Layout(grids, periods=(True, False, True))
    mdo = md.Multidomain(L, -5.0)
    g, bc = md.manufactured(L, trig, trig_lap, -5.0)
    shifted = lambda X, Y, Z: trig(X + 0.5, Y, Z)
    shifted_lap = lambda X, Y, Z: trig_lap(X + 0.5, Y, Z)
    g2, bc2 = md.manufactured(L, shifted, shifted_lap, -5.0)
    _, u = mdo.solve(g, bc)
    _, u2 = mdo.solve(g2, bc2)
    scale = max(np.abs(v).max() for v in u.values())
    for s in L.subs:
        t = ((s[0] + 1) % 2, s[1], s[2])
        assert np.abs(u2[s] - u[t]).max() < 1e-11 * scale


def test_oracle_null_space_psm1():
    """All outer faces Neumann or periodic with sigma = 0: M annihilates the
    constant; with psm = 1 the interface values have zero mean, the flux
    jumps after the solve are one constant (the right-hand side's constant
    part, ~0 for compatible data) and the solution is the exact one up to a
    constant and the discretisation error."""
    grids = layout_periodic(14)
    L = md.Layout(grids, bct=(2, 2, 2, 2, 2, 2), periods=(True, False, True))
    mdo = md.Multidomain(L, 0.0, psm=1)
    ones = np.ones(L.ninf)
    assert np.abs(mdo.M @ ones).max() < 1e-10 * np.abs(mdo.M).max()
    g, _ = md.manufactured(L, neumann_trig, neumann_trig_lap, 0.0)
    bc = _neumann_data(L, neumann_trig)
    r = mdo.rhs(g, bc)
    x, u = mdo.solve(g, bc)
    assert abs(x.mean()) < 1e-13 * np.abs(x).max()
    j = mdo.flux_jumps(u)
    assert np.abs(j - j.mean()).max() < 1e-10 * np.abs(r).max()
    assert abs(j.mean()) < 1e-3 * np.abs(r).max()
    # compare with the exact solution up to a constant
    W = []
    for s in L.subs:
        gx, gy, gz = (np.asarray(L.grids[ax][s[ax]]) for ax in range(3))
        X, Y, Z = np.meshgrid(gx, gy, gz, indexing="ij")
        W.append((u[s] - neumann_trig(X, Y, Z))[1:-1, 1:-1, 1:-1])
    off = np.mean(np.concatenate([w.ravel() for w in W]))
    err = max(np.abs(w - off).max() for w in W)
    print("psm=1: flux-jump constant %.2e, error up to a constant %.2e" % (j.mean(), err))
    assert err < 5e-3  # discretisation level, as the periodic case above


def test_oracle_psm2_zeroes_first_rhs_entry():
    """psm = 2 sets the first right-hand-side entry to zero: on a nonsingular
    layout, the flux jumps of the resulting solution are zero except at
    the first interface unknown, where the jump is the removed entry."""
    L = md.Layout(layout_small())
    mdo = md.Multidomain(L, 2.0, psm=2)
    g, bc = md.manufactured(L, sines, sines_lap, 2.0)
    r = mdo.rhs(g, bc)
    _, u = mdo.solve(g, bc)
    j = mdo.flux_jumps(u)
    # the jumps are r - M x (zero when M x = r), so here (r[0], 0, ..., 0)
    assert abs(j[0] - r[0]) < 1e-10 * np.abs(r).max() and abs(r[0]) > 1e-3 * np.abs(r).max()
    assert np.abs(j[1:]).max() < 1e-10 * np.abs(r).max()


# ------------------------------------------------------------------- GPU ----

def _mf():
    import mflops
    import torch
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    mflops.load()
    return mflops, torch


def _bc_vec(bc):
    return np.concatenate([b.ravel(order="F") for b in bc])


@pytest.mark.gpu
@pytest.mark.parametrize("lay,sigma,fun,lap", [(layout_small, 2.0, poly, poly_lap),
                                               (layout_3d, -30.0, sines, sines_lap)])
def test_gpu_md_matches_oracle(lay, sigma, fun, lap):
    mf, torch = _mf()
    grids = lay()
    L = md.Layout(grids)
    mdo = md.Multidomain(L, sigma)
    g, bc = md.manufactured(L, fun, lap, sigma)
    xo, uo = mdo.solve(g, bc)
    print("run md layout %s sigma %g ninf %d" % ([len(v) for v in grids], sigma, L.ninf), flush=True)
    h = mf.Multidomain(grids, sigma)
    assert h.ninf == L.ninf and h.nsub == len(L.subs)
    # the influence matrix, through M x
    rng = np.random.default_rng(7)
    xr = rng.standard_normal(L.ninf)
    y = mf.to_host(h.matvec(torch.tensor(xr, device="cuda")))
    ref = mdo.M @ xr
    e_m = np.abs(y - ref).max() / np.abs(ref).max()
    # the solve
    gd = [mf.to_device(g[s]) for s in L.subs]
    bd = [torch.tensor(_bc_vec(bc[s]), device="cuda") for s in L.subs]
    xi = torch.zeros(L.ninf, dtype=torch.float64, device="cuda")
    u, rep = h.solve(gd, bd, tol=1e-12, x_iface=xi)
    e_x = np.abs(mf.to_host(xi) - xo).max() / np.abs(xo).max()
    e_u = max(np.abs(mf.to_host(u[i]) - uo[s])[1:-1, 1:-1, 1:-1].max() for i, s in enumerate(L.subs))
    scale = max(np.abs(uo[s]).max() for s in L.subs)
    print("M x %.2e  x %.2e  u %.2e  GMRES %d its, residual %.2e"
          % (e_m, e_x, e_u / scale, rep.iterations, rep.residual), flush=True)
    assert rep.residual <= 1e-11
    assert e_m < 1e-11
    assert e_x < 1e-9
    assert e_u / scale < 1e-9
    if fun is poly:
        assert _exact_err(L, {s: mf.to_host(u[i]) for i, s in enumerate(L.subs)}, poly) < 1e-9


def _gpu_vs_oracle(mf, torch, grids, sigma, g, bc, L, mdo, tol=1e-12, **kw):
    """Device solve of the layout against the oracle's: (e_M, e_x, e_u, report, u)."""
    xo, uo = mdo.solve(g, bc)
    h = mf.Multidomain(grids, sigma, **kw)
    assert h.ninf == L.ninf and h.nsub == len(L.subs)
    rng = np.random.default_rng(7)
    xr = rng.standard_normal(L.ninf)
    y = mf.to_host(h.matvec(torch.tensor(xr, device="cuda")))
    ref = mdo.M @ xr
    e_m = np.abs(y - ref).max() / np.abs(ref).max()
    gd = [mf.to_device(g[s]) for s in L.subs]
    bd = [torch.tensor(_bc_vec(bc[s]), device="cuda") for s in L.subs]
    xi = torch.zeros(L.ninf, dtype=torch.float64, device="cuda")
    u, rep = h.solve(gd, bd, tol=tol, x_iface=xi)
    e_x = np.abs(mf.to_host(xi) - xo).max() / np.abs(xo).max()
    uh = {s: mf.to_host(u[i]) for i, s in enumerate(L.subs)}
    scale = max(np.abs(uo[s]).max() for s in L.subs)
    e_u = max(np.abs(uh[s] - uo[s])[1:-1, 1:-1, 1:-1].max() for s in L.subs) / scale
    print("M x %.2e  x %.2e  u %.2e  GMRES %d its, residual %.2e"
          % (e_m, e_x, e_u, rep.iterations, rep.residual), flush=True)
    return e_m, e_x, e_u, rep, uh, h


@pytest.mark.gpu
def test_gpu_md_periodic_matches_oracle(tmp_path):
    """Periodic along x (two slabs, wrap interface) and z (one slab whose two
    faces share an interface), Dirichlet walls along y: the device matrix,
    interface values and fields against the oracle, as for the non-periodic
    layouts; the matrix cache records the periods."""
    mf, torch = _mf()
    grids = layout_periodic(12)
    per = (True, False, True)
    sigma = -5.0
    L = md.Layout(grids, periods=per)
    mdo = md.Multidomain(L, sigma)
    g, bc = md.manufactured(L, trig, trig_lap, sigma)
    print("run md periodic layout ninf %d" % L.ninf, flush=True)
    path = tmp_path / "periodic.bin"
    e_m, e_x, e_u, rep, uh, h = _gpu_vs_oracle(mf, torch, grids, sigma, g, bc, L, mdo,
                                               periods=per, cache=path)
    assert rep.residual <= 1e-11 and e_m < 1e-11 and e_x < 1e-9 and e_u < 1e-9
    j = mdo.flux_jumps(uh)
    assert np.abs(j).max() < 1e-9 * np.abs(mdo.rhs(g, bc)).max()
    h2 = mf.Multidomain(grids, sigma, periods=per, cache=path)
    assert h2.loaded
    h3 = mf.Multidomain(grids, sigma, periods=(True, False, False), cache=path)
    assert not h3.loaded and h3.ninf != h2.ninf


@pytest.mark.gpu
def test_gpu_md_null_space_psm1():
    """sigma = 0 with Neumann walls and periodic x, z: the interface system
    is singular; psm = 1 solves it orthogonally to the constant on the
    device as in the oracle (zero-mean interface values), the flux jumps of
    the device solution are one constant."""
    mf, torch = _mf()
    grids = layout_periodic(14)
    per = (True, False, True)
    L = md.Layout(grids, bct=(2, 2, 2, 2, 2, 2), periods=per)
    mdo = md.Multidomain(L, 0.0, psm=1)
    g, _ = md.manufactured(L, neumann_trig, neumann_trig_lap, 0.0)
    bc = _neumann_data(L, neumann_trig)
    print("run md psm=1 layout ninf %d" % L.ninf, flush=True)
    e_m, e_x, e_u, rep, uh, h = _gpu_vs_oracle(mf, torch, grids, 0.0, g, bc, L, mdo,
                                               bct=(2, 2, 2, 2, 2, 2), periods=per, psm=1)
    assert rep.residual <= 1e-11 and e_m < 1e-11 and e_x < 1e-9 and e_u < 1e-9
    j = mdo.flux_jumps(uh)
    r = mdo.rhs(g, bc)
    assert np.abs(j - j.mean()).max() < 1e-9 * np.abs(r).max()


@pytest.mark.gpu
def test_gpu_md_psm2():
    """psm = 2 zeroes the first right-hand-side entry: on a nonsingular
    layout the device solution equals the oracle's solution of that
    modified system."""
    mf, torch = _mf()
    grids = layout_small()
    L = md.Layout(grids)
    mdo = md.Multidomain(L, 2.0, psm=2)
    g, bc = md.manufactured(L, sines, sines_lap, 2.0)
    print("run md psm=2 layout ninf %d" % L.ninf, flush=True)
    e_m, e_x, e_u, rep, uh, h = _gpu_vs_oracle(mf, torch, grids, 2.0, g, bc, L, mdo, psm=2)
    assert rep.residual <= 1e-11 and e_x < 1e-9 and e_u < 1e-9


def _with_env(name, value, fn):
    import os
    old = os.environ.get(name)
    os.environ[name] = value
    try:
        return fn()
    finally:
        if old is None:
            del os.environ[name]
        else:
            os.environ[name] = old


@pytest.mark.gpu
@pytest.mark.parametrize("lay,bct,per", [(layout_3d, (1, 2, 1, 1, 2, 1), (False, False, False)),
                                         (lambda: layout_periodic(12), (2, 2, 1, 2, 2, 2),
                                          (True, False, True))])
def test_gpu_md_spectral_assembly(lay, bct, per):
    """The influence blocks by spectral contraction (default) against the
    blocks from unit-impulse solves (MF_INFL_IMPULSES=1), which the oracle
    pins: M x within 1e-12 of the largest entry of the product.  Built with
    and without the HIP graph, and twice with it: bitwise equal."""
    mf, torch = _mf()
    grids = lay()
    sigma = -7.0
    print("run md spectral assembly %s bct %s periods %s" % ([len(v) for v in grids], bct, per), flush=True)
    mk = lambda: mf.Multidomain(grids, sigma, bct=bct, periods=per)
    h_imp = _with_env("MF_INFL_IMPULSES", "1", mk)
    h_eager = _with_env("MF_INFL_GRAPH", "0", mk)
    h_g1, h_g2 = mk(), mk()
    x = torch.tensor(np.random.default_rng(5).standard_normal(h_imp.ninf), device="cuda")
    y_imp = h_imp.matvec(x)
    ys = [h.matvec(x) for h in (h_eager, h_g1, h_g2)]
    err = (ys[1] - y_imp).abs().max().item() / y_imp.abs().max().item()
    print("  spectral vs impulse M x: %.2e" % err, flush=True)
    assert err < 1e-12
    assert torch.equal(ys[0], ys[1]) and torch.equal(ys[1], ys[2])


@pytest.mark.gpu
def test_gpu_md_rejects_bad_layout():
    mf, torch = _mf()
    grids = layout_small()
    grids[0][1] = o.grid_uniform(12, 0.55, 1.0)  # slabs do not touch
    with pytest.raises(mf.MFError):
        mf.Multidomain(grids, 1.0)


@pytest.mark.gpu
def test_gpu_md_matrix_cache(tmp_path):
    mf, torch = _mf()
    grids = layout_small()
    path = tmp_path / "influence.bin"
    h1 = mf.Multidomain(grids, 2.0, cache=path)
    assert not h1.loaded and path.exists()
    h2 = mf.Multidomain(grids, 2.0, cache=path)
    assert h2.loaded
    x = torch.tensor(np.random.default_rng(3).standard_normal(h1.ninf), device="cuda")
    assert torch.equal(h1.matvec(x), h2.matvec(x))
    h3 = mf.Multidomain(grids, 3.0, cache=path)  # another sigma: assembled again
    assert not h3.loaded
    assert not torch.equal(h3.matvec(x), h1.matvec(x))


@pytest.mark.gpu
def test_gpu_md_guess_extrapolation():
    """s1 - 3 s2 + 3 s3 of three equal solutions is the solution itself (up
    to 7 times their own GMRES error), so the fourth solve needs only a few
    iterations."""
    mf, torch = _mf()
    grids = layout_3d()
    L = md.Layout(grids)
    sigma = -30.0
    g, bc = md.manufactured(L, sines, sines_lap, sigma)
    h = mf.Multidomain(grids, sigma)
    h.set_guess(True)
    gd = [mf.to_device(g[s]) for s in L.subs]
    bd = [torch.tensor(_bc_vec(bc[s]), device="cuda") for s in L.subs]
    its, us = [], []
    for _ in range(4):
        u, rep = h.solve(gd, bd, tol=1e-11)
        its.append(rep.iterations)
        us.append([mf.to_host(t) for t in u])
    print("GMRES iterations with the guess history:", its, flush=True)
    assert its[3] * 4 < its[0]
    for a, b in zip(us[0], us[3]):
        assert np.abs(a - b).max() <= 1e-9 * np.abs(a).max()


@pytest.mark.gpu
def test_gpu_md_guess_three_distinct_solutions():
    """The history holds three different solutions: right-hand sides scaled
    by t = 1, 4, 9 give interface solutions t x, and s1 - 3 s2 + 3 s3 =
    (1 - 12 + 27) x = 16 x is exactly the solution for t = 16.  Any slot
    mix-up in the shift (sol_old(1) <- sol_old(2) <- sol_old(3) <- x,
    class_md.f90:419-424) leaves a guess off by a multiple of x, which GMRES
    then needs its full iteration count to remove."""
    mf, torch = _mf()
    grids = layout_3d()
    L = md.Layout(grids)
    sigma = -30.0
    g, bc = md.manufactured(L, sines, sines_lap, sigma)
    gd = [mf.to_device(g[s]) for s in L.subs]
    bd = [torch.tensor(_bc_vec(bc[s]), device="cuda") for s in L.subs]
    h = mf.Multidomain(grids, sigma)
    x1 = torch.zeros(L.ninf, dtype=torch.float64, device="cuda")
    _, rep1 = h.solve(gd, bd, tol=1e-13, x_iface=x1)
    h.set_guess(True)
    its = []
    for t in (1.0, 4.0, 9.0, 16.0):
        xi = torch.zeros_like(x1)
        _, rep = h.solve([a * t for a in gd], [b * t for b in bd], tol=1e-11, x_iface=xi)
        its.append(rep.iterations)
        err = (xi - t * x1).abs().max().item() / (t * x1.abs().max().item())
        assert err < 1e-9, "t=%g interface error %.2e" % (t, err)
    print("GMRES iterations: cold %d, with history %s" % (rep1.iterations, its), flush=True)
    # the fourth guess is exact up to the GMRES error of the three before it
    assert its[3] <= 2 and its[0] >= 4 * max(its[3], 1)
