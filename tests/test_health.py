"""cliffs_health: the largest Courant number (|u| + c) dt / dx over wet
cells and the count of non-finite cells, computed on the device, against
the same quantities computed in numpy from the downloaded state and the
grid's own spacings (x where rows are swept, y where columns are; dx the
smaller spacing next to the cell)."""
import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu


def _spacing_inv(s):
    """1 / min(spacing left, spacing right) at each node of a line with
    n - 1 spacings s (end nodes: their one spacing)."""
    left = np.concatenate([s[:1], s])
    right = np.concatenate([s, s[-1:]])
    return 1.0 / np.minimum(left, right)


def _courant(case, c, u, v, celmin):
    ny, nx = case.dep.shape
    s1, s2, _ = case.metrics()
    s1 = np.asarray(s1).reshape(ny, nx - 1) if nx > 2 else None
    wet = np.isfinite(c) & np.isfinite(u) & np.isfinite(v) & (c > celmin)
    cx = np.zeros_like(c)
    if nx > 2:
        rx = np.stack([_spacing_inv(s1[j]) for j in range(ny)])
        cx = (np.abs(u) + c) * case.dt * rx
    cy = np.zeros_like(c)
    if ny > 2:
        ry = _spacing_inv(np.asarray(s2))
        cy = (np.abs(v) + c) * case.dt * ry[:, None]
    return float(np.max(np.where(wet, np.maximum(cx, cy), 0.0)))


def _grid(gpu_lib, case):
    s1, s2, zeta = case.metrics()
    g = gpu_lib.Grid(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground, crough=case.crough)
    g.upload(case.cel, case.xvel, case.yvel)
    return g


@pytest.mark.parametrize("which", ["open_ocean", "coastal_spherical", "beach_1d"])
def test_courant_matches_numpy(gpu_lib, which):
    case = {"open_ocean": lambda: synth.open_ocean(n=300, seed=1, dx=500.0, ny=260),
            "coastal_spherical": lambda: synth.coastal(nx=280, ny=230, seed=4, cartesian=False,
                                                       lat0=-30.0),
            "beach_1d": lambda: synth.beach_1d(seed=1, n=800)}[which]()
    g = _grid(gpu_lib, case)
    for nsteps in (0, 40):
        g.run(nsteps)
        cour, bad = g.health()
        c, u, v = g.download()
        want = _courant(case, c, u, v, g.celmin)
        assert bad == 0
        assert want > 0.05
        assert cour == pytest.approx(want, rel=2e-5)
    g.close()


def test_nonfinite_cells_are_counted(gpu_lib):
    case = synth.coastal(nx=200, ny=150, seed=6)
    wet = np.argwhere(case.cel > 1.0)
    (j1, i1), (j2, i2) = wet[10], wet[-10]
    case.xvel = case.xvel.copy()
    case.yvel = case.yvel.copy()
    case.xvel[j1, i1] = np.nan
    case.yvel[j2, i2] = np.inf
    g = _grid(gpu_lib, case)
    cour, bad = g.health()
    c, u, v = g.download()
    assert bad == 2
    assert np.isfinite(cour) and cour == pytest.approx(_courant(case, c, u, v, g.celmin), rel=2e-5)
    g.close()
