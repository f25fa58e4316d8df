"""Ensembles (cliffs_ensemble_*): all members stepped by one fused launch per
step, replayed from a HIP graph.  Every member must end bit-identical to the
same grid stepped alone, whatever the tile heights the autotuner picks, the
step count (graph pairs plus a single step), the members' starting step
parity and the hand-back to single-grid calls."""
import numpy as np
import pytest

import synth
from test_fused import NAMES, _plane_beach, _transposed

pytestmark = pytest.mark.gpu


def _cases():
    return [
        synth.coastal(nx=300, ny=260, seed=1),
        synth.open_ocean(n=256, seed=2, dx=500.0, ny=200),
        synth.coastal(nx=520, ny=300, seed=3, crough=0.0025),
        _plane_beach(nx=400, ny=64, seed=4),
        _transposed(synth.coastal(nx=240, ny=300, seed=5)),
    ]


def _grid(gpu_lib, case):
    s1, s2, zeta = case.metrics()
    g = gpu_lib.Grid(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground, crough=case.crough)
    g.upload(case.cel, case.xvel, case.yvel)
    return g


def _state(g):
    c, u, v = g.download()
    hm, vm = g.download_max()
    return c, u, v, hm, vm


def _same(a, b, k):
    for x, y, name in zip(a, b, NAMES):
        bad = int((x.view(np.int64) != y.view(np.int64)).sum())
        assert bad == 0, f"member {k} {name}: {bad} cells differ"


@pytest.mark.parametrize("tune", [False, True])
def test_ensemble_matches_members_alone(gpu_lib, tune):
    cases = _cases()
    pre = [0, 0, 1, 0, 0]            # member 2 starts one step ahead (other parity)
    steps, after = 31, 3
    want = []
    for case, p in zip(cases, pre):
        g = _grid(gpu_lib, case)
        g.run(p + steps + after)
        want.append(_state(g))
        g.close()
    grids = [_grid(gpu_lib, c) for c in cases]
    for g, p in zip(grids, pre):
        g.run(p)
    ens = gpu_lib.Ensemble(grids)
    if tune:
        rows = ens.tune(2)
        assert ens.shape()[0] in (248, 120) and ens.shape()[1] == rows
        assert 16 <= rows <= 256
    ens.run(20)
    ens.run(steps - 20)              # odd remainder: one step outside the graph
    for g in grids:                  # hand-back: single-grid steps after the ensemble
        g.run(after)
    got = [_state(g) for g in grids]
    ens.close()
    for g in grids:
        g.close()
    for k, (a, b) in enumerate(zip(got, want)):
        _same(a, b, k)


def test_ensemble_run_timed_and_repeat(gpu_lib):
    cases = _cases()[:3]
    grids = [_grid(gpu_lib, c) for c in cases]
    ens = gpu_lib.Ensemble(grids)
    ms = ens.run_timed(8)
    assert ms > 0.0
    ens.run(8)
    ens.sync()
    got = [_state(g) for g in grids]
    ens.close()
    for g in grids:
        g.close()
    for k, case in enumerate(cases):
        g = _grid(gpu_lib, case)
        g.run(16)
        _same(got[k], _state(g), k)
        g.close()


def test_ensemble_rejects_bad_members(gpu_lib):
    beach = synth.beach_1d(seed=0, n=300)
    s1, s2, zeta = beach.metrics()
    g1 = gpu_lib.Grid(beach.dep, s1, s2, zeta, dt=beach.dt, ground=beach.ground)
    g1.upload(beach.cel, beach.xvel, beach.yvel)
    with pytest.raises(gpu_lib.CliffsError):
        gpu_lib.Ensemble([g1])       # 1-D grids step with the sweep kernels only
    g2 = _grid(gpu_lib, synth.coastal(nx=100, ny=90, seed=6))
    with pytest.raises(gpu_lib.CliffsError):
        gpu_lib.Ensemble([g2, g2])
    g1.close()
    g2.close()


def test_ensemble_gauges_match_members_alone(gpu_lib):
    """cliffs_ensemble_run_gauges: each member's series equals the same grid's
    cliffs_run_gauges series, and the state after it is the same."""
    cases = _cases()[:4]
    every, ns = 4, 9
    rng = np.random.default_rng(7)
    I = np.stack([rng.integers(1, c.dep.shape[1] + 1, 12) for c in cases]).astype(np.int32)
    J = np.stack([rng.integers(1, c.dep.shape[0] + 1, 12) for c in cases]).astype(np.int32)
    want_s, want = [], []
    for k, case in enumerate(cases):
        g = _grid(gpu_lib, case)
        want_s.append(g.run_gauges(every, ns, I[k], J[k]))
        want.append(_state(g))
        g.close()
    grids = [_grid(gpu_lib, c) for c in cases]
    ens = gpu_lib.Ensemble(grids)
    got = ens.run_gauges(every, ns, I, J)
    ens.close()
    for k, g in enumerate(grids):
        assert np.array_equal(got[k].view(np.int32), want_s[k].view(np.int32)), f"member {k} series"
        _same(_state(g), want[k], k)
        g.close()
    assert any(np.abs(s).max() > 0 for s in want_s)


def test_use_counts_guard_destroy_and_retiling(gpu_lib):
    """cliffs_destroy refuses a grid that a nest or an ensemble still uses
    (CLIFFS_ERR_STATE, nothing freed), cliffs_set_tile_rows refuses an
    ensemble member, and an ensemble whose shared member another ensemble's
    tuner re-tiled rebuilds its launch map (members stay bit-identical)."""
    import oracle as O
    lib = gpu_lib.load_library()
    cases = _cases()[:2]
    want = []
    for case, n in zip(cases, (12, 10)):   # member 0 also steps in e2
        g = _grid(gpu_lib, case)
        g.run(n)
        want.append(_state(g))
        g.close()
    grids = [_grid(gpu_lib, c) for c in cases]
    e1 = gpu_lib.Ensemble(grids)
    e2 = gpu_lib.Ensemble(grids[:1])
    assert lib.cliffs_destroy(grids[0]._h) == -6
    assert lib.cliffs_destroy(grids[1]._h) == -6
    with pytest.raises(gpu_lib.CliffsError) as ei:
        grids[0].set_tile_rows(64)
    assert ei.value.code == -6
    e1.run(4)
    e2.tune(1)                     # re-tiles member 0, which e1 shares
    e1.run(4)
    e2.run(2)
    e1.run(2)
    e1.sync()
    e2.sync()
    for k, (g, w) in enumerate(zip(grids, want)):
        _same(_state(g), w, k)
    # a nest keeps its parent alive the same way
    case = cases[0]
    maps = O.nest_maps(case.xcrd, case.ycrd, case.xcrd[10:41], case.ycrd[10:41])
    nest = gpu_lib.Nest(grids[0], maps)
    e1.close()
    e2.close()
    assert lib.cliffs_destroy(grids[0]._h) == -6   # still the nest's parent
    assert lib.cliffs_destroy(grids[1]._h) == 0
    grids[1]._h = None
    nest.close()
    assert lib.cliffs_destroy(grids[0]._h) == 0
    grids[0]._h = None


def test_ensemble_runup_parity(gpu_lib):
    """Full-size scenarios of the batched-ensemble workload (4000 steps each,
    512-2048 cells a side, friction and minimal flow depth of the workload)
    stepped together in one Ensemble: maximum run-up within one land cell's
    ground height of the FP64 oracle (tests/golden/ensemble_runup.json,
    written by tests/golden/make_ensemble_runup.py), flooded land cells
    within 1 %."""
    import json
    import os
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ensemble_runup.json")))
    seeds = [0, 1, 3, 5, 131]
    assert all(str(s) in fx for s in seeds)
    cases = [synth.ensemble_scenario(s) for s in seeds]
    grids = [_grid(gpu_lib, c) for c in cases]
    ens = gpu_lib.Ensemble(grids)
    ens.tune(2)
    ens.run(4000)
    ens.sync()
    rows = []
    for s, c, g in zip(seeds, cases, grids):
        ref = fx[str(s)]
        assert list(c.dep.shape) == ref["shape"] and ref["steps"] == 4000
        hm = g.download_max()[0]
        ru = synth.max_runup(hm, c.dep, g.celmin)
        fl = int(((hm > g.celmin) & (c.dep < 0)).sum())
        rows.append((s, ru, ref["max_runup_m"], fl, ref["flooded_land_cells"]))
        assert abs(ru - ref["max_runup_m"]) <= synth.land_cell_rise(c), rows[-1]
        assert abs(fl - ref["flooded_land_cells"]) <= max(1, 0.01 * ref["flooded_land_cells"]), rows[-1]
    print("seed, run-up device / oracle (m), flooded device / oracle:", rows)
    ens.close()
    for g in grids:
        g.close()


def test_ensemble_narrow_tiles(gpu_lib):
    """An ensemble whose members step with 120-column tiles: bit-identical to
    the members stepped alone; members of different tile widths are refused."""
    cases = _cases()[:3]
    want = []
    for case in cases:
        g = _grid(gpu_lib, case)
        g.run(9)
        want.append(_state(g))
        g.close()
    grids = [_grid(gpu_lib, c) for c in cases]
    grids[0].set_tile_width(120)
    with pytest.raises(gpu_lib.CliffsError):
        gpu_lib.Ensemble(grids)
    for g in grids[1:]:
        g.set_tile_width(120)
    ens = gpu_lib.Ensemble(grids)
    assert ens.shape()[0] == 120
    ens.run(9)
    ens.sync()
    for k, (g, w) in enumerate(zip(grids, want)):
        _same(_state(g), w, k)
    ens.close()
    for g in grids:
        g.close()
