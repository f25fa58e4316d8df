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
