"""Parity of the device arithmetic against the FP64 CPU oracle.

Each test runs on two backends with one interface:
  * ``gpu``  -- the HIP path through the C ABI (marked ``gpu``);
  * ``emul`` -- the CPU emulator of the same FP32 node arithmetic and local
               wet/dry rules (tests/native/line_emul.cpp), so the parity
               bars are exercised without a GPU.

Bars (from the project's parity targets):
  * wet/dry masks of the first step bit-exact;
  * surface elevation on cells wet in both runs within 1e-5 of the wave
    amplitude (max-abs) over the run, for the non-breaking cases;
  * maximum run-up within one grid cell.
"""
import numpy as np
import pytest

import oracle as O
import synth
from conftest import eta_of

ETA_TOL = 1e-5   # x amplitude


@pytest.fixture(params=["emul", pytest.param("gpu", marks=pytest.mark.gpu)])
def gpu(request):
    """Backend module exposing ``Grid``."""
    if request.param == "gpu":
        return request.getfixturevalue("gpu_lib")
    import emul

    class _M:
        Grid = emul.EmulGrid
    return _M


def _pair(case, gpu):
    s1, s2, zeta = case.metrics()
    orc = O.Oracle(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground, crough=case.crough)
    orc.set_state(case.cel, case.xvel, case.yvel)
    g = gpu.Grid(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground, crough=case.crough)
    g.upload(case.cel, case.xvel, case.yvel)
    return orc, g


def _compare(case, orc, g, amp, eta_tol=ETA_TOL, max_mask_flips=0):
    c_o, u_o, v_o, hm_o, vm_o = orc.get_state()
    c_g, u_g, v_g = g.download()
    hm_g, vm_g = g.download_max()
    cm = orc.celmin
    wet_o, wet_g = c_o > cm, c_g > cm
    flips = int((wet_o != wet_g).sum())
    assert flips <= max_mask_flips, f"{flips} wet/dry flips"
    both = wet_o & wet_g
    e_o = eta_of(c_o, case.dep, orc.grav)
    e_g = eta_of(c_g, case.dep, orc.grav)
    err = float(np.abs(e_o - e_g)[both].max()) if both.any() else 0.0
    assert err <= eta_tol * amp, f"eta max-abs error {err:.3e} > {eta_tol * amp:.3e}"
    # velocities on wet cells, scaled by the long-wave speed of the amplitude
    uscale = amp * np.sqrt(orc.grav / max(float(case.dep.max()), 1e-3)) + 1e-12
    du = float(np.abs(u_o - u_g)[both].max()) if both.any() else 0.0
    dv = float(np.abs(v_o - v_g)[both].max()) if both.any() else 0.0
    assert du <= 20 * eta_tol * max(uscale, float(np.abs(u_o).max())), f"u error {du:.3e}"
    assert dv <= 20 * eta_tol * max(uscale, float(np.abs(v_o).max()) + 1e-12), f"v error {dv:.3e}"
    # max-wave: max surface elevation on cells wet in both
    hw = (hm_o > cm) & (hm_g > cm)
    me = float(np.abs(eta_of(hm_o, case.dep, orc.grav) - eta_of(hm_g, case.dep, orc.grav))[hw].max())
    assert me <= eta_tol * amp, f"MaxE error {me:.3e}"
    return err


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_open_ocean_parity(gpu, seed):
    case = synth.open_ocean(n=256, seed=seed, dx=500.0, ny=192)
    orc, g = _pair(case, gpu)
    for _ in range(3):
        orc.steps(40)
        g.run(40)
    g.sync()
    _compare(case, orc, g, amp=1.0)


def test_open_ocean_nonmultiple_sizes(gpu):
    # row length not a multiple of 4 or of the 248-node strip, odd rows
    case = synth.open_ocean(n=509, seed=5, dx=500.0, ny=131)
    orc, g = _pair(case, gpu)
    orc.steps(60)
    g.run(60)
    _compare(case, orc, g, amp=1.0)


def test_lake_at_rest(gpu):
    case = synth.open_ocean(n=200, seed=3, ny=160)
    case.cel = synth.depth_to_celerity(case.dep, case.ground)
    orc, g = _pair(case, gpu)
    orc.steps(50)
    g.run(50)
    c_g, u_g, v_g = g.download()
    # the FP64 input is at rest only to rounding (sqrt(g d)^2 != g d in the
    # last bit); the device keeps that residue, it does not amplify it
    assert np.abs(u_g).max() < 1e-12 and np.abs(v_g).max() < 1e-12
    e_g = eta_of(c_g, case.dep, orc.grav)
    assert np.abs(e_g).max() < 1e-9


@pytest.mark.parametrize("cartesian,crough", [(True, 0.0), (False, 0.0025)])
def test_coastal_runup_parity(gpu, cartesian, crough):
    case = synth.coastal(nx=240, ny=160, seed=4, cartesian=cartesian, crough=crough)
    orc, g = _pair(case, gpu)
    # first step: masks bit-exact
    orc.steps(1)
    g.run(1)
    c_o = orc.get_state()[0]
    c_g = g.download()[0]
    assert np.array_equal(c_o > orc.celmin, c_g > orc.celmin)
    assert np.array_equal(c_o == 0.0, c_g == 0.0)
    for _ in range(4):
        orc.steps(20)
        g.run(20)
    _compare(case, orc, g, amp=case.meta["amplitude"], max_mask_flips=2)


def test_beach_1d_runup(gpu):
    case = synth.beach_1d(seed=1, steps=3000)
    orc, g = _pair(case, gpu)
    amp = case.meta["amplitude"]
    for _ in range(6):
        orc.steps(500)
        g.run(500)
    _compare(case, orc, g, amp=amp, max_mask_flips=1)
    hm_o = orc.get_state()[3][0]
    hm_g = g.download_max()[0][0]
    # inundation limit: first ever-wet cell from the land side
    lim_o = int(np.argmax(hm_o > orc.celmin))
    lim_g = int(np.argmax(hm_g > orc.celmin))
    assert abs(lim_o - lim_g) <= 1
    ru_o = -case.dep[0, lim_o]
    ru_g = -case.dep[0, lim_g]
    assert abs(ru_o - ru_g) <= 0.25 / 20 + 1e-12


def test_beach_1d_along_y(gpu):
    case = synth.beach_1d(seed=2, steps=600, n=600)
    t = synth.Case("beach_y", case.dep.T.copy(), case.cel.T.copy(), case.yvel.T.copy(),
                   case.xvel.T.copy(), case.ycrd, case.xcrd, True, case.dt, case.ground, 0.0,
                   600, case.meta)
    orc, g = _pair(t, gpu)
    orc.steps(600)
    g.run(600)
    _compare(t, orc, g, amp=case.meta["amplitude"], max_mask_flips=1)


def _random_wetdry(rng, nx, ny, ground=0.01):
    dep = rng.uniform(-1.5, 3.0, (ny, nx))
    eta = rng.normal(0, 0.3, (ny, nx))
    cel = synth.depth_to_celerity(dep + eta, ground)
    u = np.where(cel > 0, rng.normal(0, 0.5, (ny, nx)), 0.0)
    v = np.where(cel > 0, rng.normal(0, 0.5, (ny, nx)), 0.0)
    x = np.cumsum(rng.uniform(0.5, 1.5, nx))
    y = np.cumsum(rng.uniform(0.5, 1.5, ny))
    return dep, cel, u, v, x, y


@pytest.mark.parametrize("seed", range(4))
def test_single_sweeps_random_wetdry(gpu, seed):
    """One x-sweep and one y-sweep over dense random wet/dry patterns with
    boundary edges: every expansion / ghost / limiter / edge branch."""
    rng = np.random.default_rng(100 + seed)
    nx, ny = int(rng.integers(40, 300)), int(rng.integers(40, 200))
    dep, cel, u, v, x, y = _random_wetdry(rng, nx, ny)
    s1, s2, zeta = synth.metrics(x, y, True)
    zeta = rng.normal(0, 1e-3, ny)
    crough = float(rng.choice([0.0, 0.003]))
    orc = O.Oracle(dep, s1, s2, zeta, dt=0.05, ground=0.01, crough=crough)
    g = gpu.Grid(dep, s1, s2, zeta, dt=0.05, ground=0.01, crough=crough)
    mx = max(nx, ny)
    for iu, nl in ((1, ny), (2, nx)):
        e1 = rng.normal(0, 0.2, (mx, 3))
        e2 = rng.normal(0, 0.2, (mx, 3))
        orc.set_state(cel, u, v)
        orc.set_edges(e1, e2)
        g.upload(cel, u, v)
        over = orc.sweep(iu, 1, nl)
        assert over == 0
        g.sweep(iu, 1, nl, e1, e2)
        c_o, u_o, v_o = orc.get_state()[:3]
        c_g, u_g, v_g = g.download()
        # expansion / zeroing decisions are exact
        assert np.array_equal(c_o == 0.0, c_g == 0.0)
        assert np.array_equal(c_o == orc.celmin, np.isclose(c_g, orc.celmin, rtol=1e-6, atol=0))
        scale = 1.0 + np.abs(c_o)
        assert np.abs(c_o - c_g).max() <= 2e-5 * scale.max()
        assert np.abs(u_o - u_g).max() <= 2e-4
        assert np.abs(v_o - v_g).max() <= 2e-4


def test_partial_line_ranges(gpu):
    rng = np.random.default_rng(7)
    nx, ny = 130, 90
    dep, cel, u, v, x, y = _random_wetdry(rng, nx, ny)
    s1, s2, zeta = synth.metrics(x, y, True)
    orc = O.Oracle(dep, s1, s2, zeta, dt=0.05, ground=0.01)
    g = gpu.Grid(dep, s1, s2, zeta, dt=0.05, ground=0.01)
    for iu, first, last in ((1, 17, 17), (1, 3, 60), (2, 70, 70), (2, 65, 129)):
        orc.set_state(cel, u, v)
        g.upload(cel, u, v)
        orc.sweep(iu, first, last)
        g.sweep(iu, first, last)
        c_o, u_o, v_o = orc.get_state()[:3]
        c_g, u_g, v_g = g.download()
        untouched = np.ones_like(c_o, bool)
        if iu == 1:
            untouched[first - 1:last, :] = False
        else:
            untouched[:, first - 1:last] = False
        assert np.array_equal(c_g[untouched] > orc.celmin, cel[untouched] > orc.celmin)
        assert np.abs(u_g - u_o).max() <= 2e-4
        assert np.abs(c_g - c_o).max() <= 2e-4


def test_step_with_edges(gpu):
    case = synth.open_ocean(n=160, seed=9, ny=120)
    case.cel = synth.depth_to_celerity(case.dep, case.ground)   # start at rest
    orc, g = _pair(case, gpu)
    rng = np.random.default_rng(3)
    mx = orc.mx
    amp = 0.5
    for k in range(30):
        t = k / 10.0
        xe1 = np.zeros((mx, 3)); xe2 = np.zeros((mx, 3))
        ye1 = np.zeros((mx, 3)); ye2 = np.zeros((mx, 3))
        xe1[:, 2] = amp * np.sin(t); xe1[:, 0] = 0.01 * np.cos(t)
        ye2[:, 2] = -amp * np.sin(t); ye2[:, 1] = 0.01
        orc.set_edges(xe1, xe2)
        orc.sweep(1, 2, case.shape[0] - 1)
        orc.set_edges(ye1, ye2)
        orc.sweep(2, 2, case.shape[1] - 1)
        orc.maxwave()
        g.step(xe1, xe2, ye1, ye2)
    _compare(case, orc, g, amp=amp)
    del rng


def test_gauges(gpu):
    case = synth.coastal(nx=120, ny=96, seed=2)
    orc, g = _pair(case, gpu)
    orc.steps(10)
    g.run(10)
    I = np.array([1, 5, 60, 120, 100])
    J = np.array([1, 50, 48, 96, 10])
    eta, u, v = g.gauges(I, J)
    c_o, u_o, v_o = orc.get_state()[:3]
    for q in range(len(I)):
        i, j = I[q] - 1, J[q] - 1
        if c_o[j, i] > orc.celmin:
            assert abs(eta[q] - (c_o[j, i] ** 2 / orc.grav - case.dep[j, i])) < 1e-4
        else:
            assert eta[q] == float(np.float32(-1e32))
        assert abs(u[q] - u_o[j, i]) < 1e-4 and abs(v[q] - v_o[j, i]) < 1e-4


def test_shelf_coast_runup_parity(gpu):
    """A small scenario of the ensemble workload (SSL-conditioned shelf and
    coastline, dipole source, run-up with the workload's minimal flow depth
    and friction): maximum run-up within one land cell's ground height of the
    FP64 oracle, flooded land cells within 1 %."""
    case = synth.shelf_coast(nx=220, ny=180, seed=8, dx=700.0, source_amp=4.0,
                             source_len=60e3, steps=900, ground=synth.ENSEMBLE_GROUND,
                             crough=synth.ENSEMBLE_CROUGH)
    orc, g = _pair(case, gpu)
    for _ in range(3):
        orc.steps(300)
        g.run(300)
    hm_o = orc.get_state()[3]
    hm_g = g.download_max()[0]
    cm = orc.celmin
    ru_o = synth.max_runup(hm_o, case.dep, cm)
    ru_g = synth.max_runup(hm_g, case.dep, cm)
    assert ru_o > 2.0, "the scenario must flood land"
    assert abs(ru_g - ru_o) <= synth.land_cell_rise(case), (ru_o, ru_g)
    fl_o = int(((hm_o > cm) & (case.dep < 0)).sum())
    fl_g = int(((hm_g > cm) & (case.dep < 0)).sum())
    assert abs(fl_g - fl_o) <= max(1, 0.01 * fl_o), (fl_o, fl_g)
