"""The GPU driver's input handling, which ends before any GPU call: the
parameter file in the reference's line order (allabout.f), the master-grid
and nested-grid files, and the error exits of cliffs_main.f:167-187 and
get_boundary_feed.f (no input of any kind, missing feed files, size
mismatch, a feed that never reaches the threshold).  On a machine with a
GPU the valid case runs to the end; without one it stops at
cliffs_create -- both are accepted, the checks read the log it wrote first.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "cliffs-src_amd", "fortran", "cliffs_gpu_driver")
SLP = os.path.join(ROOT, "tests", "golden", "slpbeach")
NC3 = os.path.join(ROOT, "cliffs-src_amd", "netcdf3", "libcliffs_nc3.so")


@pytest.fixture(autouse=True)
def _built():
    if not (os.path.exists(DRIVER) and os.path.exists(NC3)):
        pytest.skip("driver not built (cliffs-src_amd/fortran/build.sh)")


def _run(args, log):
    out = subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=120)
    text = open(log).read() if os.path.exists(log) else ""
    return out.returncode, out.stdout + out.stderr, text


def _feed(path, npts, times, eta):
    """A boundary-feed file as output_feed.f writes it: vals(pnt, uvq, tim), time(tim)."""
    lib = C.CDLL(NC3)
    ncid, dp, du, dt, v1, v2 = (C.c_int() for _ in range(6))
    assert lib.nc3_create(path.encode(), 0, C.byref(ncid)) == 0
    assert lib.nc3_def_dim(ncid, b"pnt", C.c_size_t(npts), C.byref(dp)) == 0
    assert lib.nc3_def_dim(ncid, b"uvq", C.c_size_t(3), C.byref(du)) == 0
    assert lib.nc3_def_dim(ncid, b"tim", C.c_size_t(0), C.byref(dt)) == 0
    assert lib.nc3_def_var(ncid, b"vals", 6, 3, (C.c_int * 3)(dt.value, du.value, dp.value), C.byref(v1)) == 0
    assert lib.nc3_def_var(ncid, b"time", 6, 1, (C.c_int * 1)(dt.value), C.byref(v2)) == 0
    assert lib.nc3_enddef(ncid) == 0
    nt = len(times)
    vals = np.zeros((nt, 3, npts))
    vals[:, 2, :] = np.asarray(eta)[:, None]
    sz = C.POINTER(C.c_size_t)
    lib.nc3_put_vara.argtypes = [C.c_int, C.c_int, sz, sz, C.c_void_p, C.c_int]
    assert lib.nc3_put_vara(ncid, v1.value, (C.c_size_t * 3)(0, 0, 0), (C.c_size_t * 3)(nt, 3, npts),
                            vals.ctypes.data_as(C.c_void_p), 6) == 0
    t = np.asarray(times, dtype=np.float64)
    assert lib.nc3_put_vara(ncid, v2.value, (C.c_size_t * 1)(0), (C.c_size_t * 1)(nt),
                            t.ctypes.data_as(C.c_void_p), 6) == 0
    assert lib.nc3_close(ncid) == 0


def test_parameters_and_grid(tmp_path):
    out = str(tmp_path) + "/"
    rc, msg, log = _run([out + "slp", SLP + "/", "0", "solitary.0185", SLP + "/slpbeach_params.txt",
                         "a", "note"], out + "slp_log.txt")
    assert rc in (0, 3), msg + log
    if rc == 3:
        assert "cliffs_create failed" in log
    assert "a note" in log
    assert "CARTESIAN COORDINATE SYSTEM" in log
    assert "grid size  384 1" in log
    assert "Number of grids enclosed in Master Grid:  0" in log
    # slpbeach_params.txt: cuke 0, ground 0.003, crough 0, topo 1, dt 0.03,
    # 2000 steps, quake 0, "1" on the stop-when-forcing-stops line (no freeze)
    assert "cuke, ground, crough: 0. 3.E-03 0." in log
    assert "dt, steps: 3.E-02 2000  topo 1  quake 0  freeze  F" in log
    assert "outputs every (steps): sea 3  feed 1  max 500  gauges 1" in log
    assert "Initial h: " + SLP + "/solitary.0185_h.nc T" in log
    assert "Initial v: " + SLP + "/solitary.0185_v.nc F" in log


def test_no_input_of_any_kind(tmp_path):
    out = str(tmp_path) + "/"
    rc, msg, log = _run([out + "x", SLP + "/", "0", "0", SLP + "/slpbeach_params.txt"], out + "x_log.txt")
    assert rc == 1 and "no input of any kind" in log


def test_missing_and_bad_boundary_feed(tmp_path):
    d = str(tmp_path) + "/"
    shutil.copy(os.path.join(SLP, "slpbeach1D.nc"), d)
    shutil.copy(os.path.join(SLP, "slpbeach_params.txt"), d)
    # missing west feed file
    rc, msg, log = _run([d + "a", d, "bnd", "0", d + "slpbeach_params.txt"], d + "a_log.txt")
    assert rc == 1 and "bnd_slpbeach1D_west.nc not found" in log
    # the 1-D x grid reads west and east; a feed with the wrong point count
    _feed(d + "bnd_slpbeach1D_west.nc", 2, [0.0, 1.0, 2.0], [0.0, 0.5, 0.0])
    _feed(d + "bnd_slpbeach1D_east.nc", 1, [0.0, 1.0, 2.0], [0.0, 0.0, 0.0])
    rc, msg, log = _run([d + "b", d, "bnd", "0", d + "slpbeach_params.txt"], d + "b_log.txt")
    assert rc == 1 and "size mismatch" in log
    # right sizes, but the signal never reaches cuke before the last record
    _feed(d + "bnd_slpbeach1D_west.nc", 1, [0.0, 1.0, 2.0], [0.0, 0.0, 0.0])
    with open(d + "slpbeach_params.txt") as f:
        lines = f.read().splitlines()
    lines[3] = "0.01\tStill sea threshold"
    with open(d + "p2.txt", "w") as f:
        f.write("\n".join(lines) + "\n")
    rc, msg, log = _run([d + "c", d, "bnd", "0", d + "p2.txt"], d + "c_log.txt")
    assert rc == 1 and "bndr signal below threshold" in log


def _params(d, edit):
    """slpbeach_params.txt with its lines passed through edit(lines) -> lines."""
    with open(os.path.join(SLP, "slpbeach_params.txt")) as f:
        lines = f.read().splitlines()
    with open(d + "p.txt", "w") as f:
        f.write("\n".join(edit(lines)) + "\n")
    return d + "p.txt"


def test_parameter_file_errors(tmp_path):
    """Missing, short and malformed parameter files end with a message that
    names the line and status 1 (the reference's list-directed reads would
    stop the program at the first of these)."""
    d = str(tmp_path) + "/"
    shutil.copy(os.path.join(SLP, "slpbeach1D.nc"), d)
    args = lambda p: [d + "x", SLP + "/", "0", "solitary.0185", p]   # noqa: E731
    rc, msg, log = _run(args(d + "nosuch.txt"), d + "x_log.txt")
    assert rc == 1 and "parameter file not found" in log
    rc, msg, log = _run(args(_params(d, lambda l: l[:9])), d + "x_log.txt")     # ends after dt
    assert rc == 1 and "missing or malformed value on line 10" in log
    rc, msg, log = _run(args(_params(d, lambda l: l[:8] + ["abc\ttime step"] + l[9:])), d + "x_log.txt")
    assert rc == 1 and "missing or malformed value on line 9" in log
    rc, msg, log = _run(args(_params(d, lambda l: l[:1] + [""] + l[2:])), d + "x_log.txt")  # no grid name
    assert rc == 1 and "missing or malformed value on line 2" in log
    # two gauges announced, one listed
    rc, msg, log = _run(args(_params(d, lambda l: l[:17] + ["2\tN gages", "1\tevery", "10 1"])),
                        d + "x_log.txt")
    assert rc == 1 and "missing or malformed value on line 21" in log


def test_nesting_limits(tmp_path):
    """Nested grids inside a 1-D master grid (allabout.f:35-39) and more
    nested grids than the nest tables hold are refused."""
    d = str(tmp_path) + "/"
    shutil.copy(os.path.join(SLP, "slpbeach1D.nc"), d)
    args = lambda p: [d + "x", SLP + "/", "0", "solitary.0185", p]   # noqa: E731
    rc, msg, log = _run(args(_params(d, lambda l: l[:2] + ["1\tnested grids"] + l[3:])), d + "x_log.txt")
    assert rc == 1 and "nesting in 1D is not permitted" in log
    rc, msg, log = _run(args(_params(d, lambda l: l[:2] + ["100\tnested grids"] + l[3:])), d + "x_log.txt")
    assert rc == 1 and "must be 0..99" in log
