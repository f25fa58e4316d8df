"""One-way nest feed (feed_children_grid.f:10-179) and its parent maps
(NestedGrid.f:98-148): the library against the oracle."""
import numpy as np
import pytest

import cliffs_gpu
import oracle as O
import synth


def _lattices(rng):
    xp = np.cumsum(rng.uniform(50, 150, 120))
    yp = np.cumsum(rng.uniform(50, 150, 90))
    x0, x1 = np.sort(rng.uniform(xp[3], xp[-4], 2))
    y0, y1 = np.sort(rng.uniform(yp[3], yp[-4], 2))
    xc = np.linspace(x0, x1, int(rng.integers(5, 200)))
    yc = np.linspace(y0, y1, int(rng.integers(5, 200)))
    return xp, yp, xc, yc


@pytest.mark.parametrize("seed", range(20))
def test_nest_maps_bit_exact(seed):
    rng = np.random.default_rng(seed)
    xp, yp, xc, yc = _lattices(rng)
    if seed % 4 == 0:   # child nodes exactly on parent nodes (pinpoint ties)
        xc = xp[10:60:3].copy()
        yc = yp[5:40:2].copy()
    a = cliffs_gpu.nest_maps(xp, yp, xc, yc)
    b = O.nest_maps(xp, yp, xc, yc)
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
        if np.asarray(a[k]).dtype == np.float64:
            assert np.array_equal(np.asarray(a[k]).view(np.int64), np.asarray(b[k]).view(np.int64))


def test_nest_maps_reject_child_outside_parent():
    xp = np.arange(10.0)
    yp = np.arange(8.0)
    with pytest.raises(cliffs_gpu.CliffsError):
        cliffs_gpu.nest_maps(xp, yp, np.linspace(2, 12, 5), np.linspace(1, 5, 4))


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(3))
def test_nest_feed_matches_oracle(gpu_lib, seed):
    case = synth.coastal(nx=200, ny=160, seed=30 + seed)
    s1, s2, zeta = case.metrics()
    rng = np.random.default_rng(seed)
    # a child box straddling the coastline (dry parent cells contribute 0)
    x0, x1 = case.xcrd[100], case.xcrd[185]
    y0, y1 = case.ycrd[30], case.ycrd[120]
    xc = np.linspace(x0, x1, 3 * 85 + 1)
    yc = np.linspace(y0, y1, 3 * 90 + 1)
    maps = O.nest_maps(case.xcrd, case.ycrd, xc, yc)
    g = gpu_lib.Grid(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground)
    o = O.Oracle(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground)
    # before the wave reaches the box: no arrival on either side
    z = np.zeros_like(case.dep)
    rest = synth.depth_to_celerity(case.dep, case.ground)
    g.upload(rest, z, z)
    o.set_state(rest, z, z)
    nest = gpu_lib.Nest(g, maps)
    cuke = 0.01
    got = nest.feed(cuke, case.ground, False)
    want = O.feed_child(o, maps, cuke, case.ground, False)
    assert got[0] is False and want[0] is False and got[1] is False
    # identical state with the wave in the box
    g.upload(case.cel, case.xvel + rng.normal(0, 0.01, z.shape), case.yvel)
    o.set_state(case.cel, case.xvel + rng.normal(0, 0.01, z.shape) * 0, case.yvel)
    c, u, v = g.download()
    o.set_state(c, u, v)           # the oracle sees exactly the device state
    got = nest.feed(cuke, case.ground, False)
    want = O.feed_child(o, maps, cuke, case.ground, False)
    assert got[0] and want[0] and got[1]
    amp = case.meta["amplitude"]
    for k in range(4):
        gk, wk = got[2 + k], want[1 + k]
        assert gk.shape == wk.shape
        assert np.abs(gk - wk).max() <= 1e-5 * amp, (k, np.abs(gk - wk).max())
        # edge nodes fed only by dry parent cells are exactly zero on both sides
        # (at-rest wet cells give eta = 0 on the device but ~1e-13 in FP64)
        assert np.array_equal(gk[:, 0] == 0.0, wk[:, 0] == 0.0)
