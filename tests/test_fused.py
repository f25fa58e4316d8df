"""The fused 2-D step (one kernel per step plus its flag-check kernel,
step_fused.hip) against the two-kernel step (x-sweep kernel, then y-sweep
kernel) on the same GPU: state and maxima must be bit-identical.

The cases exercise every branch of the fused step's line-flag handling:
all-dry rows (x-sweep skipped), all-dry columns (y-sweep skipped, max-wave
deferred to the fixup kernel), columns that flood during the x-sweep (the
predicted column flag is wrong and the fixup kernel recomputes the tile),
dry cells holding non-canonical values (which an unwanted sweep would
reset), time-varying edges, one-strip and one-chunk grids, switching
between the fused step and single sweeps, a grid with no wet node and the
smallest 2-D grid.
"""
import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu

NAMES = ("cel", "xvel", "yvel", "hmax", "velmax")


def _grid(gpu_lib, case, fused):
    s1, s2, zeta = case.metrics()
    g = gpu_lib.Grid(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground, crough=case.crough)
    g.set_fused(fused)
    g.upload(case.cel, case.xvel, case.yvel)
    return g


def _result(g):
    c, u, v = g.download()
    hm, vm = g.download_max()
    return c, u, v, hm, vm


def _run(gpu_lib, case, fused, nsteps, chunk=1):
    g = _grid(gpu_lib, case, fused)
    done = 0
    while done < nsteps:
        n = min(chunk, nsteps - done)
        g.run(n)
        done += n
    out = _result(g)
    g.close()
    return out


def _same(a, b):
    for x, y, name in zip(a, b, NAMES):
        assert x.shape == y.shape
        bad = int((x.view(np.int64) != y.view(np.int64)).sum())
        assert bad == 0, f"{name}: {bad} cells differ"


def _transposed(case):
    """The same case with x and y swapped (land on the high-y side)."""
    return synth.Case(case.name + "_T", case.dep.T.copy(), case.cel.T.copy(), case.yvel.T.copy(),
                      case.xvel.T.copy(), case.ycrd, case.xcrd, case.cartesian, case.dt,
                      case.ground, case.crough, case.steps, dict(case.meta))


def _plane_beach(nx=400, ny=64, amp=2.0, seed=0):
    """2-D plane beach along x (land on the high-x side, every land column
    dry in every row) with a wave running up it: columns flood during the
    x-sweep."""
    rng = np.random.default_rng(seed)
    dx = 10.0
    x = np.arange(nx) * dx
    y = np.arange(ny) * dx
    X, _ = np.meshgrid(x, y)
    dep = 20.0 - X * (30.0 / x[-1]) + rng.uniform(-0.05, 0.05, (ny, nx))
    x0 = 0.45 * x[-1]
    k = np.sqrt(3 * amp / (4 * 20.0 ** 3))
    eta = amp * np.cosh(k * (X - x0)) ** -2 * (1 + 0.1 * np.sin(2 * np.pi * np.arange(ny) / ny))[:, None]
    eta = np.where(dep > 0, eta, 0.0)
    ground = 0.01
    cel = synth.depth_to_celerity(dep + eta, ground)
    H = np.maximum(dep, 0.0)
    u = np.where(dep > 0, eta * np.sqrt(synth.GRAV / np.maximum(H, 1.0)), 0.0)
    dt = 0.5 * dx / np.sqrt(synth.GRAV * (20.0 + 2 * amp))
    return synth.Case("plane_beach", dep, cel, u, np.zeros_like(dep), x, y, True, dt, ground, 0.0,
                      0, dict(amplitude=amp))


def test_open_ocean_three_strips(gpu_lib):
    case = synth.open_ocean(n=600, seed=3, dx=500.0, ny=500)
    _same(_run(gpu_lib, case, True, 60, chunk=20), _run(gpu_lib, case, False, 60, chunk=20))


@pytest.mark.parametrize("transpose", [False, True])
def test_coastal_dry_lines(gpu_lib, transpose):
    case = synth.coastal(nx=300, ny=260, seed=5, crough=0.0025)
    if transpose:
        case = _transposed(case)
    dry_lines = ~(case.cel > np.sqrt(synth.GRAV * case.ground))
    assert (dry_lines.all(axis=1).any() if transpose else dry_lines.all(axis=0).any())
    _same(_run(gpu_lib, case, True, 120, chunk=30), _run(gpu_lib, case, False, 120, chunk=30))


def test_columns_flood_during_x_sweep(gpu_lib):
    case = _plane_beach()
    cm = np.sqrt(synth.GRAV * case.ground)
    a = _run(gpu_lib, case, True, 600, chunk=50)
    b = _run(gpu_lib, case, False, 600, chunk=50)
    _same(a, b)
    dry0 = (case.cel <= cm).all(axis=0).sum()
    dry1 = (a[3] <= cm).all(axis=0).sum()   # columns never wet during the run
    assert dry1 < dry0, "no column flooded: the fixup path was not exercised"


def test_noncanonical_dry_values(gpu_lib):
    case = synth.coastal(nx=280, ny=200, seed=7)
    cm = np.sqrt(synth.GRAV * case.ground)
    dry = case.cel <= cm
    rng = np.random.default_rng(1)
    case.cel = np.where(dry, cm * rng.uniform(0.2, 0.9, dry.shape), case.cel)
    case.xvel = np.where(dry, rng.uniform(-0.5, 0.5, dry.shape), case.xvel)
    case.yvel = np.where(dry, rng.uniform(-0.5, 0.5, dry.shape), case.yvel)
    a = _run(gpu_lib, case, True, 50, chunk=1)
    _same(a, _run(gpu_lib, case, False, 50, chunk=1))
    # all-dry columns are never y-swept: their maxima (raised by the film
    # and velocities above) come from the fixup kernel's deferred update,
    # which runs only in tiles the step kernel flagged
    dry_cols = dry.all(axis=0)
    assert dry_cols.any()
    assert (a[3][:, dry_cols] > 0).any() and (a[4][:, dry_cols] > 0).any()


def test_time_varying_edges(gpu_lib):
    case = synth.coastal(nx=260, ny=230, seed=9)
    rng = np.random.default_rng(2)
    mx = max(case.shape)
    edges = [[rng.normal(0, 0.05, (mx, 3)) for _ in range(4)] for _ in range(40)]
    out = []
    for fused in (True, False):
        g = _grid(gpu_lib, case, fused)
        for e in edges:
            g.step(*e)
        out.append(_result(g))
        g.close()
    _same(*out)


@pytest.mark.parametrize("shape", [(40, 700), (1000, 12), (249, 33)])
def test_thin_grids(gpu_lib, shape):
    nx, ny = shape
    case = synth.coastal(nx=nx, ny=ny, seed=11)
    _same(_run(gpu_lib, case, True, 60, chunk=7), _run(gpu_lib, case, False, 60, chunk=7))


def test_switching_paths(gpu_lib):
    """Fused steps around single sweeps and unfused steps (flag hand-over)."""
    case = synth.coastal(nx=270, ny=210, seed=13)
    ny, nx = case.dep.shape
    out = []
    for fused in (True, False):
        g = _grid(gpu_lib, case, fused)
        g.run(5)
        g.sweep(1, 2, ny - 1)
        g.sweep(2, 2, nx - 1)
        g.maxwave()
        g.run(5)
        g.set_fused(not fused)
        g.run(3)
        g.set_fused(fused)
        g.run(4)
        g.sweep(2, 40, 120)
        g.run(6)
        out.append(_result(g))
        g.close()
    _same(*out)


@pytest.mark.parametrize("rows", [8, 16, 24, 40])
def test_short_tiles(gpu_lib, rows):
    """cliffs_set_tile_rows: short tiles (many chunks, halos larger than the
    tile) give the two-kernel step's results, on a coastal case and on a
    plane beach whose columns flood during the x-sweep."""
    for case, n in ((synth.coastal(nx=300, ny=260, seed=11), 40), (_plane_beach(nx=400, ny=100, seed=12), 60)):
        want = _run(gpu_lib, case, False, n)
        g = _grid(gpu_lib, case, True)
        g.set_tile_rows(rows)
        g.run(n // 2)
        g.set_tile_rows(0)            # back to the automatic tiles mid-run
        g.run(n - n // 2 - 5)
        g.set_tile_rows(rows)
        g.run(5)
        got = _result(g)
        g.close()
        _same(got, want)


def test_tile_rows_rejects_bad_values(gpu_lib):
    g = _grid(gpu_lib, synth.coastal(nx=120, ny=100, seed=1), True)
    for bad in (-1, 4, 7, 1025):
        with pytest.raises(gpu_lib.CliffsError):
            g.set_tile_rows(bad)
    g.close()


def test_run_timed_matches_run(gpu_lib):
    """cliffs_run_timed (events around every 8th step, in a ring of 256
    bracketed steps) advances the grid exactly as cliffs_run does, for step
    counts below, at and past the sampling stride and the ring."""
    case = synth.coastal(nx=300, ny=260, seed=7)
    for n in (1, 8, 9, 2100):
        a = _grid(gpu_lib, case, True)
        a.run(n)
        want = _result(a)
        a.close()
        b = _grid(gpu_lib, case, True)
        total, k1, k2 = b.run_timed(n)
        got = _result(b)
        b.close()
        _same(got, want)
        assert total > 0 and k1 > 0 and k2 > 0 and k1 + k2 <= total / n * 8 + 1e-3


def test_all_dry_grid_is_left_untouched(gpu_lib):
    """A grid with no wet node: every line is all-dry, which the reference
    leaves untouched (cliffs.f:113), so steps change no state, even where
    dry cells hold a film below the threshold and small velocities; the
    fused and the two-kernel step agree."""
    rng = np.random.default_rng(4)
    ny, nx = 90, 130
    dep = rng.uniform(-5.0, -0.1, (ny, nx))
    ground = 0.01
    cm = np.sqrt(synth.GRAV * ground)
    cel = cm * rng.uniform(0.0, 0.9, (ny, nx))
    u = rng.uniform(-0.1, 0.1, (ny, nx))
    v = rng.uniform(-0.1, 0.1, (ny, nx))
    x, y = np.arange(nx) * 10.0, np.arange(ny) * 10.0
    case = synth.Case("all_dry", dep, cel, u, v, x, y, True, 0.1, ground, 0.0, 5, {})
    out = []
    for fused in (True, False):
        g = _grid(gpu_lib, case, fused)
        before = g.download()
        g.run(5)
        after = g.download()
        for a, b in zip(before, after):
            assert np.array_equal(a.view(np.int64), b.view(np.int64))
        out.append(_result(g))
        g.close()
    _same(*out)


def test_smallest_2d_grid(gpu_lib):
    """3 x 3 cells: one interior node, every line an edge line."""
    case = synth.coastal(nx=3, ny=3, seed=2)
    case.cel = np.full((3, 3), 10.0)
    case.xvel = np.array([[0.0, 0.1, 0.0], [0.2, 0.3, -0.1], [0.0, 0.0, 0.1]])
    _same(_run(gpu_lib, case, True, 12, chunk=5), _run(gpu_lib, case, False, 12, chunk=5))


def _narrow_cases():
    nc = synth.coastal(nx=280, ny=200, seed=7)
    cm = np.sqrt(synth.GRAV * nc.ground)
    dry = nc.cel <= cm
    rng = np.random.default_rng(1)
    nc.cel = np.where(dry, cm * rng.uniform(0.2, 0.9, dry.shape), nc.cel)
    nc.xvel = np.where(dry, rng.uniform(-0.5, 0.5, dry.shape), nc.xvel)
    return [("coastal", synth.coastal(nx=300, ny=240, seed=6), 120, 30),
            ("coastal_T", _transposed(synth.coastal(nx=240, ny=300, seed=5)), 120, 30),
            ("plane_beach", _plane_beach(), 400, 50),
            ("noncanonical", nc, 50, 1),
            ("friction", synth.coastal(nx=250, ny=190, seed=4, cartesian=False, crough=0.0025), 150, 25),
            ("thin", synth.open_ocean(n=121, seed=3, ny=33), 60, 20)]


@pytest.mark.parametrize("name,case,nsteps,chunk", _narrow_cases(), ids=[c[0] for c in _narrow_cases()])
def test_narrow_tiles(gpu_lib, name, case, nsteps, chunk):
    """120-column tiles (a 128-node x window, two nodes per lane, six-wave
    blocks) against the two-kernel step: bit-identical, on the branches the
    248-column cases above cover."""
    g = _grid(gpu_lib, case, True)
    g.set_tile_width(120)
    done = 0
    while done < nsteps:
        n = min(chunk, nsteps - done)
        g.run(n)
        done += n
    a = _result(g)
    g.close()
    _same(a, _run(gpu_lib, case, False, nsteps, chunk))


def test_tile_width_rejects_bad_values(gpu_lib):
    case = synth.open_ocean(n=64, seed=1)
    g = _grid(gpu_lib, case, True)
    for w in (0, 100, 256):
        with pytest.raises(gpu_lib.CliffsError):
            g.set_tile_width(w)
    g.set_tile_width(120)
    g.set_tile_width(248)
    g.close()
