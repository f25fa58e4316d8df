"""FP64 line mode (include/cliffs_fp64.h): the reference's state and line
procedure in double precision on the GPU, against the FP64 oracle.

The bar is bit-identity of state and maxima: both restate cliffs.f in the
reference's operation order with IEEE double arithmetic and no fused
multiply-adds, and the oracle is itself bit-identical to the reference's
own cliffs.f (tests/test_oracle.py).  Cases: run-up over a shelf coast
with friction, frictionless films running up steep land, the 1-D beach
(x-sweep only), forced edges, and one of the full-size ensemble scenarios
whose FP32 trajectory parts from FP64 (its oracle run-up is a committed
fixture).  CPU tests: the header is plain C and the library exports it.
"""
import ctypes
import json
import os
import re
import subprocess

import numpy as np
import pytest

import cliffs_gpu
import oracle as O
import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "cliffs_fp64.h")
NAMES = ("cel", "xvel", "yvel", "hmax", "velmax")


def _symbols():
    code = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    return sorted(set(re.findall(r"^\s*int\s+(cliffs64_\w+)\s*\(", code, re.M)))


def test_fp64_header_plain_c_and_exported():
    src = '#include "%s"\nint main(void){return 0;}\n' % HDR
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-x", "c", "-", "-o", os.devnull],
                   input=src, text=True, check=True)
    names = _symbols()
    assert len(names) == 13, names
    lib = ctypes.CDLL(cliffs_gpu.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n


def _pair(gpu_lib, case):
    s1, s2, zeta = case.metrics()
    orc = O.Oracle(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground, crough=case.crough)
    orc.set_state(case.cel, case.xvel, case.yvel)
    g = gpu_lib.Grid64(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground, crough=case.crough)
    g.upload(case.cel, case.xvel, case.yvel)
    return orc, g


def _same(orc, g):
    want = orc.get_state()
    got = tuple(g.download()) + tuple(g.download_max())
    for x, y, name in zip(got, want, NAMES):
        assert x.shape == y.shape
        bad = int((x.view(np.int64) != y.view(np.int64)).sum())
        assert bad == 0, f"{name}: {bad} cells differ (max abs {float(np.abs(x - y).max()):.3e})"


def _flooded(g, orc, case):
    hm = g.download_max()[0]
    return int(((hm > orc.celmin) & (case.dep < 0)).sum())


@pytest.mark.gpu
def test_fp64_shelf_coast_runup_with_friction(gpu_lib):
    """Run-up with almost-Manning friction and a 10 cm minimal flow depth."""
    case = synth.shelf_coast(320, 240, 7, 400.0, 6.0, 60e3, steps=600, ground=0.1, crough=0.0025)
    orc, g = _pair(gpu_lib, case)
    orc.steps(600)
    g.run(600)
    _same(orc, g)
    # the run floods land, so the expansion and limiter paths ran
    assert _flooded(g, orc, case) > 50
    g.close()


@pytest.mark.gpu
def test_fp64_coastal_fast_thin_films(gpu_lib):
    """A frictionless coastal grid whose films run up steep land at up to
    ~100 m/s (a 1 cm minimal flow depth)."""
    case = synth.coastal(nx=220, ny=170, seed=5)
    orc, g = _pair(gpu_lib, case)
    orc.steps(400)
    g.run(400)
    _same(orc, g)
    assert _flooded(g, orc, case) > 100
    # gauges (output_gages.f:18-29) and health of the same state
    c, u, v = orc.get_state()[:3]
    rng = np.random.default_rng(2)
    ny, nx = c.shape
    I = rng.integers(1, nx + 1, 200)
    J = rng.integers(1, ny + 1, 200)
    eta, gu, gv = g.gauges(I, J)
    cc = c[J - 1, I - 1]
    want = np.where(cc > orc.celmin, cc * cc / orc.grav - case.dep[J - 1, I - 1], -1e32)
    assert (cc <= orc.celmin).any() and (cc > orc.celmin).any()
    assert np.array_equal(eta, want) and np.array_equal(gu, u[J - 1, I - 1]) and np.array_equal(gv, v[J - 1, I - 1])
    s1, s2, _ = case.metrics()
    rdx = np.zeros((ny, nx))
    rdx[:, 1:] = 1.0 / s1
    rdx[:, :-1] = np.maximum(rdx[:, :-1], 1.0 / s1)
    rdy = np.zeros(ny)
    rdy[1:] = 1.0 / s2
    rdy[:-1] = np.maximum(rdy[:-1], 1.0 / s2)
    wet = c > orc.celmin
    cour = max(float(((np.abs(u) + c) * case.dt * rdx)[wet].max()),
               float(((np.abs(v) + c) * case.dt * rdy[:, None])[wet].max()))
    got, nbad = g.health()
    assert nbad == 0 and np.isclose(got, cour, rtol=1e-14, atol=0), (got, cour)
    g.close()


@pytest.mark.gpu
def test_fp64_shelf_coast_thin_fast_flow(gpu_lib):
    """A small shelf coast with a 6 m dipole, frictionless, 1 cm minimal flow
    depth: the regime where FP32 state and FP64 part after ~1,000 steps."""
    case = synth.shelf_coast(320, 240, 7, 400.0, 6.0, 60e3, steps=600, ground=0.01, crough=0.0)
    orc, g = _pair(gpu_lib, case)
    orc.steps(600)
    g.run(600)
    _same(orc, g)
    assert _flooded(g, orc, case) > 50
    g.close()


@pytest.mark.gpu
def test_fp64_beach_1d(gpu_lib):
    case = synth.beach_1d(seed=3, n=1200)
    orc, g = _pair(gpu_lib, case)
    orc.steps(800)
    g.run(800)
    _same(orc, g)
    g.close()


@pytest.mark.gpu
def test_fp64_beach_1d_along_y(gpu_lib):
    """The 1-D beach turned along y (one column: the y-sweep only, through
    the column-major copy of a one-column grid)."""
    c = synth.beach_1d(seed=4, n=900)
    case = synth.Case("beach_1d_T", c.dep.T.copy(), c.cel.T.copy(), c.yvel.T.copy(), c.xvel.T.copy(),
                      c.ycrd, c.xcrd, c.cartesian, c.dt, c.ground, c.crough, c.steps, dict(c.meta))
    orc, g = _pair(gpu_lib, case)
    orc.steps(600)
    g.run(600)
    _same(orc, g)
    assert _flooded(g, orc, case) > 0
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(3, 60), (60, 3), (130, 4)])
def test_fp64_thin_grids(gpu_lib, shape):
    nx, ny = shape
    case = synth.coastal(nx=nx, ny=ny, seed=11)
    orc, g = _pair(gpu_lib, case)
    orc.steps(60)
    g.run(60)
    _same(orc, g)
    g.close()


@pytest.mark.gpu
def test_fp64_forced_edges_and_line_ranges(gpu_lib):
    """Sweeps of chosen line ranges with forced edge values (elevation,
    normal and tangential velocity at both ends of each line)."""
    case = synth.coastal(nx=160, ny=130, seed=9)
    orc, g = _pair(gpu_lib, case)
    ny, nx = case.dep.shape
    rng = np.random.default_rng(4)
    for k in range(30):
        e1 = rng.normal(0, 0.05, (g.mx, 3))
        e2 = rng.normal(0, 0.05, (g.mx, 3))
        orc.set_edges(e1, e2)
        orc.sweep(1, 2, ny - 1)
        g.sweep(1, 2, ny - 1, e1, e2)
        orc.sweep(2, 3 + k % 5, nx - 1 - k % 7)
        g.sweep(2, 3 + k % 5, nx - 1 - k % 7, e1, e2)
        orc.maxwave()
        g.maxwave()
    _same(orc, g)
    g.close()


def _random_sweeps(gpu_lib, rng, nx, ny):
    """One x-sweep and one y-sweep of the dense random wet/dry grid nx x ny
    from the same state, forced edges, random metric term, friction on or
    off: each bit-identical to the oracle's."""
    from test_gpu_parity import _random_wetdry
    dep, cel, u, v, x, y = _random_wetdry(rng, nx, ny)
    s1, s2, zeta = synth.metrics(x, y, True)
    zeta = rng.normal(0, 1e-3, ny)
    crough = float(rng.choice([0.0, 0.003]))
    orc = O.Oracle(dep, s1, s2, zeta, dt=0.05, ground=0.01, crough=crough)
    g = gpu_lib.Grid64(dep, s1, s2, zeta, dt=0.05, ground=0.01, crough=crough)
    mx = max(nx, ny)
    for iu, nl in ((1, ny), (2, nx)):
        e1 = rng.normal(0, 0.2, (mx, 3))
        e2 = rng.normal(0, 0.2, (mx, 3))
        orc.set_state(cel, u, v)
        orc.set_edges(e1, e2)
        g.upload(cel, u, v)
        orc.sweep(iu, 1, nl)
        g.sweep(iu, 1, nl, e1, e2)
        for a, b, name in zip(g.download(), orc.get_state()[:3], NAMES):
            bad = int((a.view(np.int64) != b.view(np.int64)).sum())
            assert bad == 0, f"{nx}x{ny} sweep {iu}, {name}: {bad} cells differ"
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(8))
def test_fp64_single_sweeps_random_wetdry(gpu_lib, seed):
    """Dense random wet/dry lines (every run-up expansion, segment merge,
    ghost, limiter and edge branch of the segment scan, on lines longer and
    shorter than a wave): bit-identical."""
    rng = np.random.default_rng(300 + seed)
    _random_sweeps(gpu_lib, rng, int(rng.integers(40, 300)), int(rng.integers(40, 200)))


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(1025, 48), (48, 513), (257, 130), (640, 65), (1100, 1100)])
def test_fp64_single_sweeps_random_wetdry_long_lines(gpu_lib, shape):
    """Dense random wet/dry lines whose lengths fall on and just past the
    kernel's 64-node chunks and its 256-node load batches, over several
    batches (segments, expansions and merges that straddle a chunk edge
    meet the carried neighbour and the look-ahead): bit-identical."""
    rng = np.random.default_rng(sum(shape))
    _random_sweeps(gpu_lib, rng, *shape)


@pytest.mark.gpu
def test_fp64_nest_feed_matches_oracle(gpu_lib):
    """feed_children_grid on the FP64 parent: the arrival latch and the four
    edges bit-identical to the oracle's, before the wave reaches the child
    footprint and every few steps after it; a grid with a live nest cannot
    be destroyed."""
    case = synth.coastal(nx=200, ny=160, seed=31)
    orc, g = _pair(gpu_lib, case)
    x, y = case.xcrd, case.ycrd
    xc = np.linspace(x[100], x[185], 3 * 85 + 1)
    yc = np.linspace(y[30], y[120], 3 * 90 + 1)
    maps = gpu_lib.nest_maps(x, y, xc, yc)
    nest = gpu_lib.Nest64(g, maps)
    assert gpu_lib.load_library().cliffs64_destroy(g._h) == -6     # CLIFFS_ERR_STATE
    # no arrival while no cell reaches the threshold: the latch stays open
    ao, *_ = O.feed_child(orc, maps, 1e3, case.ground, False)
    ag, fg, *_ = nest.feed(1e3, case.ground, False)
    assert not ao and not ag and not fg
    arrived_o = arrived_g = False
    fed = 0
    for k in range(40):
        orc.steps(3)
        g.run(3)
        ao, *eo = O.feed_child(orc, maps, 0.01, case.ground, arrived_o)
        ag, fg, *eg = nest.feed(0.01, case.ground, arrived_g)
        assert ao == ag and fg == ag
        arrived_o, arrived_g = ao, ag
        if ag:
            fed += 1
            for a, b, side in zip(eg, eo, ("west", "east", "south", "north")):
                assert np.array_equal(a.view(np.int64), b.view(np.int64)), f"step {3 * (k + 1)}, {side}"
    assert fed >= 5
    nest.close()
    g.close()


@pytest.mark.gpu
def test_fp64_rejects_bad_arguments(gpu_lib):
    case = synth.coastal(nx=40, ny=30, seed=1)
    s1, s2, zeta = case.metrics()
    g = gpu_lib.Grid64(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground)
    with pytest.raises(gpu_lib.CliffsError):
        g.run(1)                      # no state uploaded
    g.upload(case.cel, case.xvel, case.yvel)
    for bad in ((3, 2, 5), (1, 0, 5), (1, 5, 4), (2, 2, 41)):
        with pytest.raises(gpu_lib.CliffsError):
            g.sweep(*bad)
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("key", ["131f", "0f", "1f", "3f", "5f"])
def test_fp64_full_size_scenario(gpu_lib, key):
    """Ensemble scenarios at full size (767-1610 cells a side, 4000 steps),
    frictionless with a 1 cm minimal flow depth: there the FP32 state's
    trajectory leaves FP64's (on scenario 131 after ~1,250 steps) and ends
    with another run-up; the FP64 line mode gives the FP64 oracle's run-up
    and flooded area exactly (tests/golden/ensemble_runup.json)."""
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ensemble_runup.json")))[key]
    case = synth.ensemble_scenario(int(key.rstrip("f")), ground=fx["ground"], crough=fx["crough"])
    assert list(case.dep.shape) == fx["shape"]
    s1, s2, zeta = case.metrics()
    g = gpu_lib.Grid64(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground, crough=case.crough)
    g.upload(case.cel, case.xvel, case.yvel)
    for k in range(8):
        g.run(500)
        print(f"scenario {key}, FP64 line mode: {500 * (k + 1)} steps", flush=True)
    hm = g.download_max()[0]
    cm = float(np.sqrt(np.float64(np.float32(9.8)) * case.ground))
    assert synth.max_runup(hm, case.dep, cm) == fx["max_runup_m"]
    assert int(((hm > cm) & (case.dep < 0)).sum()) == fx["flooded_land_cells"]
    g.close()
