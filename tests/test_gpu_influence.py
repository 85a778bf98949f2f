"""GPU unit-impulse batch (influence-matrix columns) against the oracle's
restatement of md_influence_matrix_set (class_field.f90:1396-1510).

Bar as for solves: |gpu - exact| <= max(1e-12, 2 |oracle64 - exact|) relative
to the largest response, with exact = the refined discrete solution and
80-bit derivatives.
"""
import os

import numpy as np
import pytest

import mf_oracle as o

pytestmark = pytest.mark.gpu


def _mf():
    import mflops
    import torch

    assert torch.cuda.is_available()
    mflops.load()
    return mflops, torch


def _case(shape, bct, sigma, faces, first, count, exact=True):
    mf, torch = _mf()
    gs = (o.grid_asin(shape[0], 0.0, 1.0), o.grid_tanh(shape[1], 0.0, 0.5, 1.4),
          o.grid_uniform(shape[2], 0.0, 0.75))
    osc = o.SolverCoeffs3D(*gs, bct, 6)
    odc = [o.DerCoeffs(g, solver=True, so=6) for g in gs]
    sc = mf.SolverCoeffs3D(*gs, bct, 6)
    dcs = [mf.DerivativeCoefficients(g, solver=True, so=6) for g in gs]
    for face in faces:
        tag = "impulses shape=%s bct=%s face=%s first=%d count=%d" % (
            shape, bct, mf.FACES[face], first, count)
        print("run " + tag, flush=True)
        got = mf.influence_impulses(sc, dcs, sigma, face, first, count).cpu().numpy()
        ora = o.influence_impulses(osc, odc, sigma, face, first, count)
        if exact:
            ex = o.influence_impulses(osc, odc, sigma, face, first, count, exact=True)
        else:
            ex = ora
        scale = np.abs(ex).max()
        e_gpu = np.abs(got - ex).max() / scale
        e_ora = np.abs(ora - ex).max() / scale
        bar = max(1e-12, 2 * e_ora)
        print("%s: |gpu-exact| %.2e |oracle64-exact| %.2e (bar %.1e)" % (tag, e_gpu, e_ora, bar),
              flush=True)
        assert e_gpu <= bar, (tag, e_gpu, e_ora)


def test_impulses_all_dirichlet():
    _case((12, 11, 13), (1,) * 6, -40.0, range(6), 0, 6)


def test_impulses_neumann_faces():
    # Neumann faces only receive responses; the impulse faces are Dirichlet
    _case((14, 12, 11), (1, 2, 2, 1, 1, 1), 3.0, (0, 3, 4, 5), 3, 5)


def test_impulses_long_x_lines():
    """x-lines longer than a wavefront (78 interior points: the x-face
    derivative kernel reduces two 64-lane chunks per line), Neumann x faces
    included."""
    _case((80, 10, 10), (2, 1, 1, 2, 1, 1), -10.0, (1, 2, 5), 0, 4)


def test_impulses_batched_chunks(monkeypatch):
    """A whole face in several batches (batch cap 7) equals the oracle."""
    monkeypatch.setenv("MFLOPS_IMPULSE_BATCH", "7")
    _case((11, 10, 12), (1,) * 6, -1e3, (2,), 0, 9 * 10, exact=False)


def test_impulses_linearity():
    """Responses are linear: the sum over a face's impulses equals the response
    to a unit value on the whole face (one solve through the solver path)."""
    mf, torch = _mf()
    shape = (13, 12, 11)
    gs = (o.grid_uniform(13), o.grid_tanh(12, 0, 1, 1.2), o.grid_asin(11, 0, 1))
    sc = mf.SolverCoeffs3D(*gs, (1,) * 6, 6)
    dcs = [mf.DerivativeCoefficients(g, solver=True, so=6) for g in gs]
    odc = [o.DerCoeffs(g, solver=True, so=6) for g in gs]
    face, npts = 1, 10 * 9
    rows = mf.influence_impulses(sc, dcs, 7.0, face, 0, npts).cpu().numpy()
    total = rows.sum(0)
    bc = mf.BoundaryCondition.zeros(*shape)
    bcx = np.zeros((10, 9, 2), order="F")
    bcx[:, :, 1] = 1.0
    bc.bcx = mf.to_device(bcx)
    u = mf.fortran_zeros(shape)
    sc.solve(mf.fortran_zeros(shape), u, bc, 7.0)
    uh = mf.to_host(u)
    dx = o.derivative_3d(odc[0], uh, 0, 1)
    np.testing.assert_allclose(total[:2 * 10 * 9].reshape((10, 9, 2), order="F")[:, :, 1],
                               -dx[-1, 1:-1, 1:-1], rtol=0, atol=1e-11 * np.abs(dx).max())
