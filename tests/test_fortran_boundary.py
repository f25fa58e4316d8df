"""The Fortran boundary (cliffs-src_amd/fortran/): ISO_C_BINDING module and
drop-in `cliffs(iu,irw,n)`, driven the way cliffs_main.f:303-355 drives the
reference, against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDRIVER = os.path.join(ROOT, "tests", "fortran", "fdriver")


def _write_input(path, case, nsteps, nline):
    ny, nx = case.shape
    s1, s2, zeta = case.metrics()
    with open(path, "wb") as f:
        np.array([nx, ny, nsteps, nline], dtype=np.int32).tofile(f)
        np.array([case.dt, case.ground, case.crough], dtype=np.float64).tofile(f)
        for a in (case.dep, case.cel, case.xvel, case.yvel):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
        if nx > 2:
            np.ascontiguousarray(s1, dtype=np.float64).tofile(f)
        if ny > 2:
            np.ascontiguousarray(s2, dtype=np.float64).tofile(f)
            np.ascontiguousarray(zeta, dtype=np.float64).tofile(f)


def test_fortran_driver_built_and_linked():
    if not os.path.exists(FDRIVER):
        subprocess.run(["sh", os.path.join(ROOT, "cliffs-src_amd", "fortran", "build.sh"), "test"],
                       check=True)
    out = subprocess.run(["ldd", FDRIVER], capture_output=True, text=True).stdout
    assert "libcliffs_gpu.so" in out and "not found" not in out.split("libcliffs_gpu.so")[1].split("\n")[0]


@pytest.mark.gpu
@pytest.mark.parametrize("cartesian", [True, False])
def test_fortran_driver_matches_oracle(tmp_path, gpu_lib, cartesian):
    case = synth.coastal(nx=150, ny=110, seed=21, cartesian=cartesian, crough=0.002)
    nsteps, nline = 8, 2
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(fin, case, nsteps, nline)
    print(f"[fdriver] coastal 150x110 cartesian={cartesian}", flush=True)
    subprocess.run([FDRIVER, str(fin), str(fout)], check=True, timeout=120)
    ny, nx = case.shape
    got = np.fromfile(fout, dtype=np.float64).reshape(5, ny, nx)
    s1, s2, zeta = case.metrics()
    o = O.Oracle(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground, crough=case.crough)
    o.set_state(case.cel, case.xvel, case.yvel)
    o.steps(nsteps)
    ref = o.get_state()
    cm = o.celmin
    wet_o, wet_g = ref[0] > cm, got[0] > cm
    assert int((wet_o != wet_g).sum()) <= 1
    both = wet_o & wet_g
    eo = ref[0] ** 2 / o.grav - case.dep
    eg = got[0] ** 2 / o.grav - case.dep
    assert np.abs(eo - eg)[both].max() <= 1e-5 * case.meta["amplitude"]
    assert np.abs(ref[1] - got[1])[both].max() < 1e-4
    hw = (ref[3] > cm) & (got[3] > cm)
    assert np.abs((ref[3] ** 2 - got[3] ** 2) / o.grav)[hw].max() <= 1e-5 * case.meta["amplitude"]


FNEST = os.path.join(ROOT, "tests", "fortran", "fnest")


def _read_feed(path):
    raw = open(path, "rb").read()
    off, recs = 0, []
    while True:
        ncid, nn, it = np.frombuffer(raw, dtype=np.int32, count=3, offset=off)
        off += 12
        if ncid == -1:
            return recs, bool(nn), int(it)
        t = np.frombuffer(raw, dtype=np.float64, count=1, offset=off)[0]
        off += 8
        vals = np.frombuffer(raw, dtype=np.float64, count=3 * nn, offset=off).reshape(3, nn).T
        off += 24 * nn
        recs.append((int(ncid), int(it), float(t), vals.copy()))


@pytest.mark.gpu
def test_fortran_feed_children_grid_matches_oracle(tmp_path, gpu_lib):
    case = synth.coastal(nx=180, ny=140, seed=41)
    s1, s2, zeta = case.metrics()
    xc = np.linspace(case.xcrd[90], case.xcrd[170], 241)
    yc = np.linspace(case.ycrd[20], case.ycrd[110], 271)
    cuke = 0.01
    fin, fout = tmp_path / "in.bin", tmp_path / "feed.bin"
    with open(fin, "wb") as f:
        np.array([case.shape[1], case.shape[0], len(xc), len(yc)], dtype=np.int32).tofile(f)
        np.array([case.dt, case.ground, case.crough, cuke], dtype=np.float64).tofile(f)
        for a in (case.dep, case.cel, case.xvel, case.yvel, s1, s2, zeta, case.xcrd, case.ycrd,
                  xc, yc):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
    print("[fnest] coastal 180x140, child 241x271", flush=True)
    subprocess.run([FNEST, str(fin), str(fout)], check=True, timeout=120)
    recs, arrived, irec = _read_feed(fout)
    o = O.Oracle(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground)
    o.set_state(case.cel, case.xvel, case.yvel)
    maps = O.nest_maps(case.xcrd, case.ycrd, xc, yc)
    want = O.feed_child(o, maps, cuke, case.ground, False)
    assert arrived and want[0] and irec == 1
    assert [r[0] for r in recs] == [11, 12, 13, 14] and all(r[1] == 1 for r in recs)
    for k in range(4):
        assert np.abs(recs[k][3] - want[1 + k]).max() <= 1e-5 * case.meta["amplitude"]
