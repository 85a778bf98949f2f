"""Field files and NS restarts in the reference's formats.

The NetCDF classic reader/writer (mflops3d-md_amd/csrc/io.cpp) is pinned
against an independent implementation of the format, scipy.io.netcdf_file:
files written here read back in scipy with the reference's dimension names,
dimension order (z, y, x: Fortran's x-fastest layout) and values, and files
scipy writes in that layout read back here.  Host-only: no GPU needed.
The restart test (GPU) checks the reference's file set and that a run
continued from a restart is bitwise identical to an uninterrupted one.
"""
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

import mflops as mf


def _field(shape, seed=0):
    return np.asfortranarray(np.random.default_rng(seed).standard_normal(shape))


def test_write_reads_in_scipy(tmp_path):
    f = _field((7, 5, 4), 1)
    path = tmp_path / "vel_u_1.nc"
    mf.write_field(path, "U", f, dbl=True)
    with netcdf_file(str(path), "r", mmap=False) as nc:
        assert nc.version_byte == 1
        assert nc.dimensions == {"resolution_x": 7, "resolution_y": 5, "resolution_z": 4}
        v = nc.variables["U"]
        assert v.dimensions == ("resolution_z", "resolution_y", "resolution_x")
        assert v.data.dtype == np.dtype(">f8")
        assert np.array_equal(np.asarray(v.data).transpose(), f)


def test_single_precision_and_second_variable(tmp_path):
    f, g = _field((6, 3, 5), 2), _field((6, 3, 5), 3)
    path = tmp_path / "both.nc"
    mf.write_field(path, "U", f, dbl=False)
    mf.write_field(path, "RHS_U", g, dbl=True)  # added to the existing file
    with netcdf_file(str(path), "r", mmap=False) as nc:
        assert set(nc.variables) == {"U", "RHS_U"}
        assert nc.variables["U"].data.dtype == np.dtype(">f4")
        assert np.array_equal(np.asarray(nc.variables["U"].data).transpose(), f.astype(np.float32))
        assert np.array_equal(np.asarray(nc.variables["RHS_U"].data).transpose(), g)
    dims, names, dbl = mf.field_info(path, "U")
    assert dims == (6, 3, 5) and names == mf.FIELD_DIMS and not dbl
    assert np.array_equal(mf.read_field(path, "U"), f.astype(np.float32).astype(np.float64))
    # rewriting a variable replaces its values and keeps its type
    mf.write_field(path, "RHS_U", 2.0 * g, dbl=False)
    assert mf.field_info(path, "RHS_U")[2]
    assert np.array_equal(mf.read_field(path, "RHS_U"), 2.0 * g)
    assert np.array_equal(mf.read_field(path, "U"), f.astype(np.float32).astype(np.float64))


def test_wrong_dimensions_and_missing_variable(tmp_path):
    path = tmp_path / "f.nc"
    mf.write_field(path, "U", _field((4, 4, 4)))
    with pytest.raises(mf.MFError):
        mf.write_field(path, "V", _field((4, 5, 4)))  # resolution_y is 4 in the file
    with pytest.raises(mf.MFError):
        mf.read_field(path, "W")
    with pytest.raises(mf.MFError):
        mf.read_field(tmp_path / "absent.nc", "U")
    bad = tmp_path / "bad.nc"
    bad.write_bytes(b"HDF5 not a classic file")
    with pytest.raises(mf.MFError):
        mf.read_field(bad, "U")


def test_scipy_written_file_reads_here(tmp_path):
    """A file in the reference's layout written by another implementation,
    with a global attribute and a variable attribute (skipped on read), the
    variables defined in another order and an extra 1-D variable."""
    f = _field((5, 6, 3), 4)
    path = tmp_path / "ref.nc"
    with netcdf_file(str(path), "w") as nc:
        nc.title = "written elsewhere"
        for name, n in zip(mf.FIELD_DIMS, f.shape):
            nc.createDimension(name, n)
        x = nc.createVariable("gridx", "d", ("resolution_x",))
        x[:] = np.linspace(0, 1, 5)
        p = nc.createVariable("P", "d", ("resolution_z", "resolution_y", "resolution_x"))
        p.units = "none"
        p[:] = f.transpose()
        q = nc.createVariable("Q", "f", ("resolution_z", "resolution_y", "resolution_x"))
        q[:] = (3 * f).transpose()
    assert mf.field_info(path, "P") == ((5, 6, 3), mf.FIELD_DIMS, True)
    assert np.array_equal(mf.read_field(path, "P"), f)
    assert np.array_equal(mf.read_field(path, "Q"), (3 * f).astype(np.float32).astype(np.float64))


# ------------------------------------------------------------------- GPU ----

REF_FILES = (["param.dat", "rhs_phi_4.nc", "vel_u.nc", "vel_v.nc", "vel_w.nc", "vel_p.nc"]
             + ["%s_%d.nc" % (b, i) for i in range(1, 5)
                for b in ("vel_u", "vel_v", "vel_w", "vel_p", "vel_phi",
                          "rhs_u", "rhs_v", "rhs_w", "rhs_p")])


@pytest.mark.gpu
def test_ns_restart_continues_bitwise(tmp_path):
    import torch
    from mflops import synthetic
    assert torch.cuda.is_available()
    mf.load()
    gx, gy, gz, u, v, w, bc = synthetic.channel_problem(20, 24, 18, amp=0.1)
    re, dt = 3000.0, 1e-3

    def make():
        s = mf.NavierStokes(gx, gy, gz, re, dt)
        for c in range(3):
            s.set_bc(c, mf.BoundaryCondition(*[mf.to_device(b) for b in bc[c]]))
        s.set_state(*[mf.to_device(f) for f in (u, v, w)])
        return s

    print("run ns restart round trip", flush=True)
    a = make()
    a.step(6, mf.MF_FP64)
    ref = [mf.to_host(t) for t in a.get()]
    b = make()
    b.step(3, mf.MF_FP64)
    d = b.restart_write(str(tmp_path / "restart"))
    assert sorted(os.listdir(d)) == sorted(REF_FILES)
    raw = open(os.path.join(d, "param.dat"), "rb").read()
    t = np.frombuffer(raw[:8], dtype=np.float64)[0]
    it = np.frombuffer(raw[8:], dtype=np.int32)
    assert abs(t - 3 * dt) < 1e-15 and sorted(it) == [1, 2, 3, 4]
    # the latest level, single precision, as the reference writes it
    assert np.array_equal(mf.read_field(os.path.join(d, "vel_u.nc"), "U"),
                          mf.to_host(b.get()[0]).astype(np.float32).astype(np.float64))
    c = mf.NavierStokes(gx, gy, gz, re, dt)
    for k in range(3):
        c.set_bc(k, mf.BoundaryCondition(*[mf.to_device(x) for x in bc[k]]))
    c.restart_read(d)
    assert abs(c.info()[0] - 3 * dt) < 1e-15
    c.step(3, mf.MF_FP64)
    got = [mf.to_host(t) for t in c.get()]
    for name, x, y in zip(("u", "v", "w", "p", "phi"), got, ref):
        print("  %s: max |restarted - uninterrupted| = %.1e" % (name, np.abs(x - y).max()), flush=True)
        assert np.array_equal(x, y), name
    # a default-named directory is tn(time): 8 + 8 digits
    name = b.restart_write()
    try:
        assert os.path.basename(name) == "restart_00000000.00300000"
    finally:
        import shutil
        shutil.rmtree(name)
