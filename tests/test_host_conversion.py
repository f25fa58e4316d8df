"""The host-side conversions of cliffs_create, cliffs_upload_state,
cliffs_download_state, cliffs_download_max and cliffs_reset_max split a
grid's rows over host threads from 2^20 cells on; each cell converts on
its own.  A 2.1 M-cell grid with land, thin water and random state must
come back bit-identical to the same rows held by grids below the threshold,
whose conversions run row by row on the calling thread."""
import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu


def _roundtrip(Grid, dep, cel, u, v, ground):
    ny, nx = dep.shape
    s1, s2, zeta = synth.metrics(np.arange(nx) * 500.0, np.arange(ny) * 500.0, True)
    g = Grid(dep, s1, s2, zeta, dt=1.0, ground=ground)
    g.upload(cel, u, v)
    out = g.download() + g.download_max()
    g.close()
    return out


def test_threaded_conversions_match_serial(gpu_lib):
    rng = np.random.default_rng(3)
    nx, ny = 2047, 1031                      # > 2^20 cells, odd sizes: uneven row split
    assert nx * ny >= 1 << 20 and 400 * nx < 1 << 20
    dep = rng.uniform(-20.0, 4000.0, (ny, nx))
    dep[rng.random((ny, nx)) < 0.2] = -0.5
    dep[rng.random((ny, nx)) < 0.05] = 0.005   # below ground: dry or thin water
    eta = rng.normal(0.0, 0.5, (ny, nx))
    ground = 0.01
    cel = synth.depth_to_celerity(dep + eta, ground)
    u = rng.normal(0.0, 1.0, (ny, nx))
    v = rng.normal(0.0, 1.0, (ny, nx))
    whole = _roundtrip(gpu_lib.Grid, dep, cel, u, v, ground)
    # maxima as cliffs_create leaves them (cliffs_reset_max): c = 0, u^2+v^2 = 0
    assert not np.any(whole[3].view(np.int64)) and not np.any(whole[4].view(np.int64))
    for j0 in range(0, ny, 400):
        j1 = min(j0 + 400, ny)
        band = _roundtrip(gpu_lib.Grid, dep[j0:j1], cel[j0:j1], u[j0:j1], v[j0:j1], ground)
        for a, b, name in zip(whole, band, ("cel", "xvel", "yvel", "hmax", "velmax")):
            bad = int((a[j0:j1].view(np.int64) != b.view(np.int64)).sum())
            assert bad == 0, f"{name}, rows {j0}..{j1}: {bad} cells differ"


def test_reset_max_after_steps(gpu_lib):
    """cliffs_reset_max on a stepped grid (a device kernel from the device's
    S): hmax back to c = 0 and velmax to 0, bit for bit; one step later the
    maxima are that step's c (where positive) and u^2 + v^2."""
    case = synth.coastal(nx=300, ny=260, seed=5)
    s1, s2, zeta = case.metrics()
    g = gpu_lib.Grid(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground)
    g.upload(case.cel, case.xvel, case.yvel)
    g.run(40)
    hm, vm = g.download_max()
    assert (hm > 0).any() and (vm > 0).any()
    g.reset_max()
    hm, vm = g.download_max()
    assert not np.any(hm.view(np.int64)) and not np.any(vm.view(np.int64))
    g.run(1)
    c, u, v = g.download()
    hm, vm = g.download_max()
    assert np.array_equal(hm, np.where(c > 0, c, 0.0))
    sp2 = (u.astype(np.float32) ** 2 + v.astype(np.float32) ** 2).astype(np.float64)
    assert np.allclose(vm, sp2, rtol=1e-6, atol=0)
    g.close()
