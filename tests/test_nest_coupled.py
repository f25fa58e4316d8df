"""Parent and child grids stepped together on one device
(cliffs_nest_couple/record/child_step): the feed records stay in device
memory and the child's edges are interpolated in time on the device.  The
child must end bit-identical to the reference's data path -- feed records
brought to the host (cliffs_nest_feed), interpolated in FP64 as
cliffs_main.f:305-339 does, and handed to cliffs_step as edge arrays."""
import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

import oracle as O
import synth

pytestmark = pytest.mark.gpu


def _setup(gpu_lib):
    case = synth.coastal(nx=200, ny=160, seed=31)
    s1, s2, zeta = case.metrics()
    x0, x1 = case.xcrd[100], case.xcrd[185]
    y0, y1 = case.ycrd[30], case.ycrd[120]
    xc = np.linspace(x0, x1, 3 * 85 + 1)
    yc = np.linspace(y0, y1, 3 * 90 + 1)
    maps = O.nest_maps(case.xcrd, case.ycrd, xc, yc)
    interp = RegularGridInterpolator((case.ycrd, case.xcrd), case.dep)
    Yc, Xc = np.meshgrid(yc, xc, indexing="ij")
    depc = interp(np.stack([Yc, Xc], axis=-1))
    cs1, cs2, czeta = synth.metrics(xc, yc, True)

    def parent():
        g = gpu_lib.Grid(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground)
        g.upload(case.cel, case.xvel, case.yvel)
        return g

    def child():
        c = gpu_lib.Grid(depc, cs1, cs2, czeta, dt=case.dt / 3, ground=case.ground)
        z = np.zeros_like(depc)
        c.upload(synth.depth_to_celerity(depc, case.ground), z, z)
        return c

    return case, maps, parent, child


def _host_edges(recs, state, t, mx):
    """cliffs_main.f:305-339 in FP64: records (t, west, east, south, north)."""
    while state["bndr"] and t >= recs[state["irec"] - 1][0]:
        assert state["irec"] < len(recs), "child ahead of the feed"
        state["irec"] += 1
    k = state["irec"]
    t0, t1 = recs[k - 2][0], recs[k - 1][0]
    cc = (t1 - t) / (t1 - t0)
    out = []
    for side in range(1, 5):
        e = recs[k - 2][side] * cc + recs[k - 1][side] * (1 - cc)
        pad = np.zeros((mx, 3))
        pad[:e.shape[0]] = e
        out.append(pad)
    return out


def test_coupled_child_matches_host_feed_path(gpu_lib):
    case, maps, parent, child = _setup(gpu_lib)
    cuke, bndout, cycles = 0.01, 2, 30
    gA, cA = parent(), child()
    gB, cB = parent(), child()
    nA = gpu_lib.Nest(gA, maps)
    nB = gpu_lib.Nest(gB, maps)
    nB.couple(cB, capacity=4)
    dtp, dtc = case.dt, case.dt / 3
    recs, state = [], {"irec": 1, "bndr": True}
    iarvA = iarvB = False
    tp, tc, child_steps = 0.0, None, 0
    for _ in range(cycles):
        gA.run(bndout)
        gB.run(bndout)
        tp += bndout * dtp
        iarvA, fed, w, e, s, n = nA.feed(cuke, case.ground, iarvA)
        if fed:
            recs.append((tp, w, e, s, n))
        iarvB, nrec = nB.record(cuke, case.ground, tp, iarvB)
        assert iarvA == iarvB and nrec == len(recs)
        if len(recs) < 2:
            continue
        if tc is None:
            tc = recs[0][0]
        while tc + dtc < recs[-1][0]:
            tc += dtc
            cA.step(*_host_edges(recs, state, tc, cA.mx))
            nB.child_step(tc)
            child_steps += 1
    assert len(recs) >= 10 and child_steps >= 100
    a = cA.download() + cA.download_max()
    b = cB.download() + cB.download_max()
    for x, y, name in zip(a, b, ("cel", "xvel", "yvel", "hmax", "velmax")):
        bad = int((x.view(np.int64) != y.view(np.int64)).sum())
        assert bad == 0, f"child {name}: {bad} cells differ"
    # the parent's wave came in through the child's edges
    c, u, v = b[:3]
    assert float(np.abs(u).max()) > 1e-4
    # the child cannot run ahead of the parent's last record ...
    with pytest.raises(gpu_lib.CliffsError):
        nB.child_step(recs[-1][0] + dtc)
    # ... unless the feed has ended: then its edges are 0 (cliffs_main.f:306-310)
    nB.end_feed()
    nB.child_step(recs[-1][0] + dtc)
    for g in (nA, nB):
        g.close()
    for g in (gA, gB, cA, cB):
        g.close()


def test_ring_capacity_guard(gpu_lib):
    case, maps, parent, child = _setup(gpu_lib)
    g, c = parent(), child()
    n = gpu_lib.Nest(g, maps)
    n.couple(c, capacity=2)
    t, iarv, k = 0.0, False, 0
    with pytest.raises(gpu_lib.CliffsError):
        for _ in range(40):           # the child never steps: the ring fills up
            g.run(2)
            t += 2 * case.dt
            iarv, k = n.record(0.01, case.ground, t, iarv)
    assert k == 2
    n.close()
    g.close()
    c.close()


def _record_wave(rec):
    """Largest |eta| over the four edges of one feed record
    (get_boundary_feed.f:147-157)."""
    return max(float(np.abs(e[:, 2]).max()) for e in rec[1:5])


def _host_start(recs, cuke):
    """get_boundary_feed.f:148-163: the first record (1-based) whose wave
    reaches cuke; None while no record does."""
    for k, r in enumerate(recs, start=1):
        if _record_wave(r) >= cuke:
            return k
    return None


def test_child_start_rule(gpu_lib):
    """The child starts at the first record whose edge |eta| reaches its own
    threshold (get_boundary_feed.f:144-167) and its clock at that record's
    time (cliffs_main.f:188); bit-identical to a host restatement of the
    rule feeding cliffs_step.  Records found below the threshold are released
    from the ring, so a ring of 4 holds the wait."""
    case, maps, parent, child = _setup(gpu_lib)
    cuke, cuke_child, bndout, cycles = 0.01, 1.3, 2, 30
    gA, cA = parent(), child()
    gB, cB = parent(), child()
    nA, nB = gpu_lib.Nest(gA, maps), gpu_lib.Nest(gB, maps)
    nB.couple(cB, capacity=4)
    dtp, dtc = case.dt, case.dt / 3
    recs, state = [], {"irec": 1, "bndr": True}
    iarvA = iarvB = False
    tp, tc, tcB, child_steps, startA = 0.0, None, None, 0, None
    for _ in range(cycles):
        gA.run(bndout)
        gB.run(bndout)
        tp += bndout * dtp
        iarvA, fed, w, e, s, n = nA.feed(cuke, case.ground, iarvA)
        if fed:
            recs.append((tp, w, e, s, n))
        iarvB, nrec = nB.record(cuke, case.ground, tp, iarvB)
        assert nrec == len(recs)
        if startA is None:
            startA = _host_start(recs, cuke_child)
            if startA is not None:
                tc = recs[startA - 1][0]
                state["irec"] = startA
        if tcB is None and nrec:
            try:
                tcB = nB.child_start(cuke_child)
            except gpu_lib.CliffsError as err:
                assert err.code == -6 and startA is None
        assert (tcB is None) == (startA is None)
        if tc is None:
            continue
        assert tcB == recs[startA - 1][0]
        while tc + dtc < recs[-1][0]:
            tc += dtc
            cA.step(*_host_edges(recs, state, tc, cA.mx))
            nB.child_step(tc)
            child_steps += 1
    assert startA is not None and startA >= 10, startA
    assert child_steps >= 10
    # a time before the start record is refused
    with pytest.raises(gpu_lib.CliffsError):
        nB.child_step(recs[startA - 1][0] - dtc)
    a = cA.download() + cA.download_max()
    b = cB.download() + cB.download_max()
    for x, y, name in zip(a, b, ("cel", "xvel", "yvel", "hmax", "velmax")):
        bad = int((x.view(np.int64) != y.view(np.int64)).sum())
        assert bad == 0, f"child {name}: {bad} cells differ"
    for g in (nA, nB, gA, gB, cA, cB):
        g.close()


def test_child_start_below_threshold(gpu_lib):
    """A feed that ends without reaching the child's threshold is an error
    (get_boundary_feed.f:160-162, 'bndr signal below threshold'), and so is
    stepping the child before any record reaches it."""
    case, maps, parent, child = _setup(gpu_lib)
    g, c = parent(), child()
    n = gpu_lib.Nest(g, maps)
    n.couple(c, capacity=2)
    with pytest.raises(gpu_lib.CliffsError):
        n.child_start(1e3)        # no record yet; the threshold is now known
    t, iarv = 0.0, False
    for _ in range(6):            # records below the threshold are released: no ring overflow
        g.run(2)
        t += 2 * case.dt
        iarv, k = n.record(0.01, case.ground, t, iarv)
    assert k == 6
    with pytest.raises(gpu_lib.CliffsError) as ei:
        n.child_start(1e3)
    assert ei.value.code == -6
    with pytest.raises(gpu_lib.CliffsError):
        n.child_step(t)
    n.end_feed()
    with pytest.raises(gpu_lib.CliffsError) as ei:
        n.child_start(1e3)
    assert "ended below" in str(ei.value)
    for x in (n, g, c):
        x.close()


def test_nest_run_matches_record_and_child_step_loop(gpu_lib):
    """cliffs_nest_run (the coupled loop in one call, split here over
    several calls) ends bit-identical to the same loop written around
    cliffs_run, cliffs_nest_record and cliffs_nest_child_step."""
    case, maps, parent, child = _setup(gpu_lib)
    cuke, every, per, nrec = 0.01, 2, 6, 40
    dtp, dtc = case.dt, case.dt / 3
    gA, cA = parent(), child()
    nA = gpu_lib.Nest(gA, maps)
    nA.couple(cA, capacity=4)
    iarv, trec, kc = False, [], 0
    for k in range(nrec):
        gA.run(every)
        tp = (k + 1) * every * dtp
        iarv, nr = nA.record(cuke, case.ground, tp, iarv)
        if nr > len(trec):
            trec.append(tp)
        while kc + 1 < per * (len(trec) - 1):
            kc += 1
            j = kc // per
            nA.child_step(trec[j] + (kc - j * per) * dtc)
    gB, cB = parent(), child()
    nB = gpu_lib.Nest(gB, maps)
    nB.couple(cB, capacity=4)
    iarvB = False
    for chunk in (7, 1, 20, 12):
        iarvB = nB.run(chunk, every, dtp, per, dtc, cuke, case.ground, iarvB)
    assert iarvB == iarv and len(trec) > 10 and kc > 50
    for a, b in ((gA, gB), (cA, cB)):
        for fa, fb in zip(a.download() + a.download_max(), b.download() + b.download_max()):
            assert np.array_equal(fa, fb)
    for n in (nA, nB):
        n.close()
    for g in (cA, gA, cB, gB):
        g.close()


def test_concurrent_solves_match_solo(gpu_lib):
    """Coupled solves created, tiled and stepped from 8 host threads at once
    end bit-identical to the same solve run alone.  (Grids used to clear
    their fields and flag buffers on the null stream, which is not ordered
    with a grid's own non-blocking stream: under concurrent load a late
    clear wiped state already uploaded.)"""
    from concurrent.futures import ThreadPoolExecutor
    case, maps, parent, child = _setup(gpu_lib)

    def solve(rows):
        g, c = parent(), child()
        g.set_tile_rows(rows)
        c.set_tile_rows(rows)
        n = gpu_lib.Nest(g, maps)
        n.couple(c, capacity=4)
        n.run(40, 2, case.dt, 6, case.dt / 3, 0.01, case.ground)
        out = c.download() + c.download_max() + g.download()
        n.close()
        c.close()
        g.close()
        return out

    want = solve(16)
    with ThreadPoolExecutor(max_workers=8) as ex:
        got = list(ex.map(solve, [16] * 16))
    for k, res in enumerate(got):
        for a, b in zip(want, res):
            assert np.array_equal(a, b), f"solve {k} differs"
