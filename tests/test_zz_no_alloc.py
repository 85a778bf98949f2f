"""mf_filter3d_noalloc (no allocation, library-owned scratch) against
mf_filter3d (caller workspace), both through the C ABI: two calls in a row,
the second on a smaller field, bitwise equal outputs; sizes beyond the
scratch are refused on the host before anything is launched."""
import ctypes as C

import numpy as np
import pytest

import les_oracle as les
import mf_oracle as o

pytestmark = pytest.mark.gpu

MF_FILTER_SCRATCH_DOUBLES = 6291456


def _setup():
    import mflops
    import torch

    assert torch.cuda.is_available(), "GPU test run without a GPU"
    lib = mflops.load()
    lib.mf_filter3d_noalloc.restype = C.c_int
    lib.mf_filter3d.restype = C.c_int
    return mflops, torch, lib


def _filters(mf, shape_in, shape_out):
    fs = []
    for a in range(3):
        xi = o.grid_tanh(shape_in[a], 0.0, 1.0, 1.2 + 0.3 * a)
        xo = o.grid_tanh(shape_out[a], 0.0, 1.0, 1.2 + 0.3 * a)
        fs.append(mf.SplineFilter(xi, xo, xo, 3))
    return fs


def _both(mf, torch, lib, shape_in, shape_out, seed):
    fx, fy, fz = _filters(mf, shape_in, shape_out)
    rng = np.random.default_rng(seed)
    x = mf.to_device(np.asfortranarray(rng.standard_normal(shape_in)))
    nwork = shape_out[0] * shape_in[1] * shape_in[2] + shape_out[0] * shape_out[1] * shape_in[2]
    assert nwork <= MF_FILTER_SCRATCH_DOUBLES
    work = torch.empty(nwork, dtype=torch.float64, device="cuda")
    a = mf.fortran_zeros(shape_out)
    b = mf.fortran_zeros(shape_out)
    h = [f._h for f in (fx, fy, fz)]
    rc = lib.mf_filter3d(*h, C.c_void_p(x.data_ptr()), C.c_void_p(a.data_ptr()),
                         C.c_void_p(work.data_ptr()), None)
    assert rc == 0, lib.mf_last_error()
    rc = lib.mf_filter3d_noalloc(*h, C.c_void_p(x.data_ptr()), C.c_void_p(b.data_ptr()), None)
    assert rc == 0, lib.mf_last_error()
    torch.cuda.synchronize()
    ref = les.filter3d(*[f.matrix() for f in (fx, fy, fz)], mf.to_host(x))
    return mf.to_host(a), mf.to_host(b), ref


def test_noalloc_filter_matches_workspace_filter():
    mf, torch, lib = _setup()
    # the 256^3 -> 64^3 filter of the SGS workflow fills most of the scratch;
    # the second, smaller call reuses what the first left there
    for shape_in, shape_out, seed in (((256, 256, 256), (64, 64, 64), 1),
                                      ((40, 36, 30), (12, 11, 10), 2)):
        print("run filter3d_noalloc %s -> %s" % (shape_in, shape_out), flush=True)
        a, b, ref = _both(mf, torch, lib, shape_in, shape_out, seed)
        err = np.abs(b - ref).max() / np.abs(ref).max()
        print("  noalloc == workspace: %s, |noalloc - oracle| %.2e" % (np.array_equal(a, b), err),
              flush=True)
        assert np.array_equal(a, b)
        assert err < 1e-12


def test_noalloc_filter_refuses_oversized_fields():
    mf, torch, lib = _setup()
    # 256^3 -> 128 x 64 x 64 needs 128*256*256 + 128*64*256 = 10.5M doubles
    fx, fy, fz = _filters(mf, (256, 256, 256), (128, 64, 64))
    x = torch.zeros(1, dtype=torch.float64, device="cuda")  # never read: refused first
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    rc = lib.mf_filter3d_noalloc(fx._h, fy._h, fz._h, C.c_void_p(x.data_ptr()),
                                 C.c_void_p(out.data_ptr()), None)
    assert rc == 1
    assert b"MF_FILTER_SCRATCH_DOUBLES" in lib.mf_last_error()
