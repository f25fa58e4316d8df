"""The workloads at their full sizes, through properties that do not need
the oracle to run the whole thing where it cannot:

* 1-D sloping beach, 2000 cells, 20,000 steps (the full run): surface
  elevation within 1e-5 of the amplitude of the FP64 oracle on cells wet in
  both runs, at most one wet/dry flip, run-up limit within one cell;
* 4096 x 4096 open ocean, 100 steps of the full grid against the oracle at
  the bar of test_gpu_parity (the oracle runs on all host threads);
* spherical 8192 x 6144 parent with run-up and a 2047 x 2047 child at 3x
  refinement on its coast: the child stepped from device feed records
  (cliffs_nest_couple/record/child_step) ends bit-identical to the child
  stepped from feed records brought to the host and blended there as
  cliffs_main.f:305-339 does.
"""
import functools

import numpy as np
import pytest

import oracle as O
import synth
from test_gpu_parity import _compare

pytestmark = pytest.mark.gpu
G = synth.GRAV


def _wet_eta(c, dep, celmin):
    return c > celmin, c * c / G - dep


def test_c1_beach_full_run(gpu_lib):
    case = synth.beach_1d(seed=1)                 # 2000 cells, 20,000 steps
    s1, s2, zeta = case.metrics()
    orc = O.Oracle(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground, crough=case.crough)
    orc.set_state(case.cel, case.xvel, case.yvel)
    g = gpu_lib.Grid(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground, crough=case.crough)
    g.upload(case.cel, case.xvel, case.yvel)
    amp, worst, flips = case.meta["amplitude"], 0.0, 0
    for _ in range(40):
        orc.steps(500)
        g.run(500)
        wo, eo = _wet_eta(orc.get_state()[0], case.dep, orc.celmin)
        wg, eg = _wet_eta(g.download()[0], case.dep, orc.celmin)
        both = wo & wg
        worst = max(worst, float(np.abs(eg - eo)[both].max()))
        flips = max(flips, int((wo != wg).sum()))
    hm_o, hm_g = orc.get_state()[3][0], g.download_max()[0][0]
    lim = abs(int(np.argmax(hm_o > orc.celmin)) - int(np.argmax(hm_g > orc.celmin)))
    g.close()
    print(f"C1: eta err {worst:.3e} (bar {1e-5 * amp:.3e}), flips {flips}, run-up limit cells {lim}")
    assert worst <= 1e-5 * amp and flips <= 1 and lim <= 1


def test_c2_open_ocean_full_grid(gpu_lib):
    case = synth.open_ocean(n=4096, seed=0)
    s1, s2, zeta = case.metrics()
    orc = O.Oracle(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground, crough=case.crough)
    orc.set_state(case.cel, case.xvel, case.yvel)
    g = gpu_lib.Grid(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground, crough=case.crough)
    g.upload(case.cel, case.xvel, case.yvel)
    orc.steps(100)
    g.run(100)
    err = _compare(case, orc, g, amp=1.0)          # the 1 m hump
    g.close()
    print(f"C2: eta err {err:.3e} after 100 steps of the full grid (bar 1e-5 m)")


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


@functools.lru_cache(maxsize=1)
def _c3():
    """The nested workload's parent and child (synth.coastal_nest), generated
    once per session (about a minute of numpy)."""
    return synth.coastal_nest()


def test_c3_nest_full_size(gpu_lib):
    # child: 2048^2 nodes at 3x over the parent's coast opposite the source
    case, xc, yc, depc, meta = _c3()
    x, y = case.xcrd, case.ycrd
    maps = gpu_lib.nest_maps(x, y, xc, yc)
    s1, s2, zeta = case.metrics()
    cs1, cs2, czeta = synth.metrics(xc, yc, False)

    def parent():
        p = gpu_lib.Grid(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground, crough=case.crough)
        p.upload(case.cel, case.xvel, case.yvel)
        return p

    def child():
        cg = meta["child_ground"]
        c = gpu_lib.Grid(depc, cs1, cs2, czeta, dt=case.dt / 3, ground=cg, crough=case.crough)
        z = np.zeros_like(depc)
        c.upload(synth.depth_to_celerity(depc, cg), z, z)
        return c

    gA, cA = parent(), child()
    gB, cB = parent(), child()
    nA = gpu_lib.Nest(gA, maps)
    nB = gpu_lib.Nest(gB, maps)
    nB.couple(cB, capacity=4)
    dtp, dtc, bndout, cuke = case.dt, case.dt / 3, 2, 0.01
    recs, state = [], {"irec": 1, "bndr": True}
    iarvA = iarvB = False
    tp, tc, child_steps = 0.0, None, 0
    for _ in range(8):
        gA.run(bndout)
        gB.run(bndout)
        tp += bndout * dtp
        iarvA, fed, w, e, s, nn = nA.feed(cuke, case.ground, iarvA)
        if fed:
            recs.append((tp, w, e, s, nn))
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
    assert len(recs) >= 6 and child_steps >= 25
    assert max(float(np.abs(r[1][:, 2]).max()) for r in recs) > 1e-3     # a wave on the west edge
    a = cA.download() + cA.download_max()
    b = cB.download() + cB.download_max()
    for arr_a, arr_b, name in zip(a, b, ("cel", "xvel", "yvel", "hmax", "velmax")):
        bad = int((arr_a.view(np.int64) != arr_b.view(np.int64)).sum())
        assert bad == 0, f"child {name}: {bad} cells differ"
    print(f"C3: parent {case.dep.shape}, child {depc.shape}, {len(recs)} records, {child_steps} child steps, "
          f"child max |u| {float(np.abs(b[1]).max()):.3e}")
    for h in (nA, nB):
        h.close()
    for h in (gA, gB, cA, cB):
        h.close()


def test_c3_parent_window_parity(gpu_lib):
    """A window of the nested workload's parent (spherical, 30 arc-seconds,
    its shelf, coast and dipole source, about 600 x 520 cells from the source to
    the land behind the coast) stepped through the wave's arrival and run-up
    on the device and by the FP64 oracle: first-step wet/dry masks equal,
    surface elevation within 1e-5 of the source amplitude on cells wet in
    both, run-up within one land cell's ground height."""
    import oracle as O
    case, xc, yc, depc, meta = _c3()
    jc, ic = meta["source_row"], meta["coast_col"]
    src_lon = case.meta["source"][0]
    i_src = int(np.searchsorted(case.xcrd, src_lon))
    i0, i1 = max(i_src - 120, 0), min(ic + 80, case.dep.shape[1])
    j0, j1 = jc - 260, jc + 260
    blk = (slice(j0, j1), slice(i0, i1))
    dep = np.ascontiguousarray(case.dep[blk])
    x, y = case.xcrd[i0:i1], case.ycrd[j0:j1]
    s1, s2, zeta = synth.metrics(x, y, False)
    cel = np.ascontiguousarray(case.cel[blk])
    z = np.zeros_like(dep)
    orc = O.Oracle(dep, s1, s2, zeta, dt=case.dt, ground=case.ground, crough=case.crough)
    orc.set_state(cel, z, z)
    g = gpu_lib.Grid(dep, s1, s2, zeta, dt=case.dt, ground=case.ground, crough=case.crough)
    g.upload(cel, z, z)
    orc.steps(1)
    g.run(1)
    cm = orc.celmin
    assert np.array_equal(orc.get_state()[0] > cm, g.download()[0] > cm)
    orc.steps(1799)
    g.run(1799)
    c_o, u_o, v_o, hm_o, vm_o = orc.get_state()
    c_g = g.download()[0]
    hm_g = g.download_max()[0]
    both = (c_o > cm) & (c_g > cm)
    err = float(np.abs((c_o ** 2 - c_g ** 2) / orc.grav)[both].max())
    assert err <= 1e-5 * 3.0, err
    ru_o = synth.max_runup(hm_o, dep, cm)
    ru_g = synth.max_runup(hm_g, dep, cm)
    rise = 300.0 / 50e3 * float(s1[j1 - j0 - 1].min())   # ground height of one land cell, 300 m over 50 km
    assert ru_o > 0.5, "the wave must run up the coast in the window"
    assert abs(ru_g - ru_o) <= rise, (ru_o, ru_g, rise)
    print(f"C3 window {dep.shape}: eta error {err:.2e} m, run-up device {ru_g:.2f} m / oracle {ru_o:.2f} m")
    g.close()


def test_largest_grids_step_as_two_kernels(gpu_lib):
    """A grid past the fused step's 32-bit offsets (11008^2 = 121 M cells,
    nine fields of 4.4 GB) steps with the two sweep kernels and 64-bit
    offsets.  Ten steps of a hump in its middle must equal, bit for bit, the
    same ten steps of the 600^2 block around the hump as a grid of its own
    (fused step), away from that block's edges: a step moves information at
    most three cells."""
    n, dx, ground = 11008, 500.0, 0.01
    assert 9 * (n * n + 64) * 4 >= 1 << 32
    x = np.arange(n) * dx
    y = np.arange(n) * dx
    dep = 4000.0 - 1500.0 * np.outer(np.sin(y / 4.0e5) ** 2, np.cos(x / 3.0e5) ** 2)
    i0 = j0 = 5200
    m = 600
    eta = np.zeros_like(dep)
    xb, yb = x[i0:i0 + m], y[j0:j0 + m]
    eta[j0:j0 + m, i0:i0 + m] = np.exp(-((xb[None, :] - xb[m // 2]) ** 2 + (yb[:, None] - yb[m // 2]) ** 2)
                                       / (2 * 20000.0 ** 2))
    cel = synth.depth_to_celerity(dep + eta, ground)
    del eta
    z = np.zeros_like(dep)
    dt = 0.8 * dx / np.sqrt(synth.GRAV * 4001.0)
    s1, s2, zeta = synth.metrics(x, y, True)
    big = gpu_lib.Grid(dep, s1, s2, zeta, dt=dt, ground=ground)
    big.upload(cel, z, z)
    big.run(10)
    want = big.download() + big.download_max()
    big.close()
    blk = (slice(j0, j0 + m), slice(i0, i0 + m))
    bs1, bs2, bzeta = synth.metrics(xb, yb, True)
    small = gpu_lib.Grid(dep[blk].copy(), bs1, bs2, bzeta, dt=dt, ground=ground)
    small.upload(cel[blk].copy(), z[blk].copy(), z[blk].copy())
    small.run(10)
    got = small.download() + small.download_max()
    small.close()
    inner = (slice(40, m - 40), slice(40, m - 40))
    for a, b, name in zip(want, got, ("cel", "xvel", "yvel", "hmax", "velmax")):
        wa = np.ascontiguousarray(a[blk][inner])
        wb = np.ascontiguousarray(b[inner])
        bad = int((wa.view(np.int64) != wb.view(np.int64)).sum())
        assert bad == 0, f"{name}: {bad} cells differ"
    assert float(np.abs(want[1][blk]).max()) > 1e-4      # the hump moved
