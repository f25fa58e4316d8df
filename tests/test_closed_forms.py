"""Oracle checks against what the mathematics fixes (no reference run needed).

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails it.
"""
import math

import numpy as np
import pytest

import oracle as O
import synth

G = synth.GRAV


# -- grid metrics (MasterGrid.f:85-116) ------------------------------------
def test_metrics_are_sphere_distances():
    rng = np.random.default_rng(0)
    lon = 140 + np.cumsum(rng.uniform(0.001, 0.02, 40))
    lat = -60 + np.cumsum(rng.uniform(0.5, 3.0, 50))
    s1, s2, zeta = O.metrics(lon, lat, cartesian=False)
    R = 6371009.0
    for j in range(len(lat)):
        phi = math.radians(lat[j])
        for i in range(len(lon) - 1):
            arc = R * math.cos(phi) * math.radians(lon[i + 1] - lon[i])
            assert abs(s1[j, i] - arc) <= 1e-6 * arc
    for j in range(len(lat) - 1):
        arc = R * math.radians(lat[j + 1] - lat[j])
        assert abs(s2[j] - arc) <= 1e-6 * arc
    # zeta = tan(phi) / (8 R): the spherical source term of cliffs.f:214
    for j in range(len(lat)):
        z = math.tan(math.radians(lat[j])) / (8 * R)
        assert abs(zeta[j] - z) <= 1e-6 * abs(z) + 1e-30
    # Cartesian: plain differences, zeta = 0
    c1, c2, cz = O.metrics(lon, lat, cartesian=True)
    assert np.array_equal(c1[3], np.diff(lon)) and np.array_equal(c2, np.diff(lat))
    assert not cz.any()


# -- pinpoint (NestedGrid.f:129-148) ---------------------------------------
def test_pinpoint_brute_force():
    rng = np.random.default_rng(1)
    for _ in range(400):
        n = int(rng.integers(1, 40))
        xx = np.sort(rng.uniform(-5, 5, n))
        if n > 1 and len(np.unique(xx)) < n:
            continue
        cands = list(rng.uniform(-6, 6, 4)) + list(xx[rng.integers(0, n, 2)])
        for xp in cands:
            ip, cp = O.pinpoint(xx, xp)
            if n == 1:
                assert ip == 1 and cp == 0.5
                continue
            # largest ip in [1, n-1] with xx(ip) < xp, else 1
            below = [k for k in range(1, n) if xx[k - 1] < xp]
            want = below[-1] if below else 1
            assert ip == want
            w = (xx[ip] - xp) / (xx[ip] - xx[ip - 1])
            assert cp == w
            if xx[0] < xp <= xx[-1]:
                assert 0.0 <= cp < 1.0 + 1e-15
                # the interpolant of a linear function is exact
                assert abs((cp * xx[ip - 1] + (1 - cp) * xx[ip]) - xp) < 1e-12


# -- well-balanced lake at rest, with a shoreline --------------------------
def test_lake_at_rest_with_shore_stays_at_rest():
    nx, ny = 60, 40
    x = np.arange(nx) * 10.0
    y = np.arange(ny) * 10.0
    dep = 5.0 - 0.02 * x[None, :] * np.ones((ny, 1)) + 0.3 * np.sin(y)[:, None]
    cel = synth.depth_to_celerity(np.maximum(dep, 0) * (dep > 0), 0.01)
    s1, s2, zeta = O.metrics(x, y, True)
    o = O.Oracle(dep, s1, s2, zeta, dt=0.5, ground=0.01)
    o.set_state(cel, np.zeros_like(dep), np.zeros_like(dep))
    o.steps(200)
    c, u, v = o.get_state()[:3]
    wet = c > o.celmin
    eta = c * c / G - dep
    assert np.abs(eta[wet]).max() < 1e-10
    assert np.abs(u).max() < 1e-10 and np.abs(v).max() < 1e-10
    assert np.array_equal(wet, cel > o.celmin)


# -- symmetries -------------------------------------------------------------
def _gauss_case(nx, ny, depth=50.0, dx=20.0):
    x = np.arange(nx) * dx
    y = np.arange(ny) * dx
    X, Y = np.meshgrid(x, y)
    dep = depth + 10 * np.cos(X / 400.0) * np.cos(Y / 300.0)
    eta = 0.5 * np.exp(-((X - x[nx // 3]) ** 2 + (Y - y[ny // 2]) ** 2) / (2 * 120.0 ** 2))
    cel = synth.depth_to_celerity(dep + eta, 0.01)
    return x, y, dep, cel


def test_mirror_symmetry_x():
    x, y, dep, cel = _gauss_case(80, 50)
    u0 = 0.05 * np.sin(np.arange(80) / 7.0)[None, :] * np.ones((50, 1))
    v0 = 0.02 * np.cos(np.arange(50) / 5.0)[:, None] * np.ones((1, 80))
    s1, s2, zeta = O.metrics(x, y, True)
    a = O.Oracle(dep, s1, s2, zeta, dt=0.3, ground=0.01)
    b = O.Oracle(dep[:, ::-1].copy(), s1[:, ::-1].copy(), s2, zeta, dt=0.3, ground=0.01)
    a.set_state(cel, u0, v0)
    b.set_state(cel[:, ::-1], -u0[:, ::-1], v0[:, ::-1])
    a.steps(60)
    b.steps(60)
    ca, ua, va = a.get_state()[:3]
    cb, ub, vb = b.get_state()[:3]
    assert np.abs(ca - cb[:, ::-1]).max() < 1e-11
    assert np.abs(ua + ub[:, ::-1]).max() < 1e-11
    assert np.abs(va - vb[:, ::-1]).max() < 1e-11
    # the mirrored run is not the same run: the wave does move
    assert np.abs(ua).max() > 1e-3


def test_y_sweep_is_x_sweep_of_transpose():
    x, y, dep, cel = _gauss_case(40, 64)
    rng = np.random.default_rng(3)
    u0 = rng.normal(0, 0.05, dep.shape)
    v0 = rng.normal(0, 0.05, dep.shape)
    s1, s2, zeta = O.metrics(x, y, True)
    a = O.Oracle(dep, s1, s2, zeta, dt=0.3, ground=0.01)
    a.set_state(cel, u0, v0)
    a.sweep(2, 2, 39)
    t1, t2, tz = O.metrics(y, x, True)
    b = O.Oracle(dep.T.copy(), t1, t2, tz, dt=0.3, ground=0.01)
    b.set_state(cel.T, v0.T, u0.T)
    b.sweep(1, 2, 39)
    ca, ua, va = a.get_state()[:3]
    cb, ub, vb = b.get_state()[:3]
    assert np.array_equal(ca, cb.T) and np.array_equal(ua, vb.T) and np.array_equal(va, ub.T)


# -- physics ----------------------------------------------------------------
def test_linear_long_wave_speed():
    """A small hump in uniform depth splits into waves moving at sqrt(g d)."""
    n, dx, d = 1600, 10.0, 20.0
    x = np.arange(n) * dx
    # 40-cell half-width: well resolved, so numerical dispersion is negligible
    eta = 0.01 * np.exp(-((x - x[n // 2]) / 400.0) ** 2)
    dep = np.full((1, n), d)
    cel = synth.depth_to_celerity(dep + eta[None, :], 0.01)
    s1, s2, zeta = O.metrics(x, np.zeros(1), True)
    c0 = math.sqrt(G * d)
    dt = 0.5 * dx / c0
    o = O.Oracle(dep, s1, s2, zeta, dt=dt, ground=0.01)
    o.set_state(cel, np.zeros_like(dep), np.zeros_like(dep))
    nsteps = 800
    o.steps(nsteps)
    c = o.get_state()[0][0]
    e = c * c / G - d
    right = n // 2 + int(np.argmax(e[n // 2:]))
    travelled = (right - n // 2) * dx
    assert abs(travelled - c0 * nsteps * dt) <= 2 * dx
    # each half carries half the amplitude
    assert 0.4 * 0.01 < e[right] < 0.55 * 0.01


def test_synolakis_runup_law():
    """Maximum run-up of a non-breaking solitary wave on a plane beach:
    R/d = 2.831 sqrt(cot beta) (A/d)^(5/4) (Synolakis 1987); slpbeach's
    0.0185 case, slope 1:19.85."""
    from test_oracle import slpbeach_case
    case = slpbeach_case("0185")
    o = O.Oracle(case["dep"], case["s1"], case["s2"], case["zeta"], dt=case["dt"],
                 ground=case["ground"])
    o.set_state(case["cel"], case["xvel"], case["yvel"])
    o.steps(1500)
    hm = o.get_state()[3][0]
    ever_wet = hm > o.celmin
    runup = float(-case["dep"][0][ever_wet].min())
    law = 2.831 * math.sqrt(19.85) * 0.0185 ** 1.25
    assert abs(runup - law) <= 0.12 * law, (runup, law)


# -- nest feed (feed_children_grid.f) ---------------------------------------
def test_feed_reproduces_linear_fields_and_masks_dry():
    nx, ny = 50, 40
    xp = np.linspace(0, 49, nx) * 100.0
    yp = np.linspace(0, 39, ny) * 120.0
    X, Y = np.meshgrid(xp, yp)
    dep = np.full((ny, nx), 30.0)
    dep[:, :5] = -1.0                     # dry strip in the parent
    a, b, c = 0.1, 2e-4, -3e-4
    u = a + b * X + c * Y
    v = -a + c * X + b * Y
    eta = 0.2 + 1e-4 * X - 2e-4 * Y
    cel = np.where(dep > 0, np.sqrt(G * (dep + eta)), 0.0)
    s1, s2, zeta = O.metrics(xp, yp, True)
    o = O.Oracle(dep, s1, s2, zeta, dt=1.0, ground=0.01)
    o.set_state(cel, u, v)
    xc = np.linspace(1230.0, 3870.0, 67)   # child lattice inside the wet part
    yc = np.linspace(700.0, 3500.0, 55)
    maps = O.nest_maps(xp, yp, xc, yc)
    arrived, w, e, s, nn = O.feed_child(o, maps, cuke=0.001, hmin=0.01, iarv=False)
    assert arrived
    f = {"u": lambda x, y: a + b * x + c * y, "v": lambda x, y: -a + c * x + b * y,
         "q": lambda x, y: 0.2 + 1e-4 * x - 2e-4 * y}
    for arr, xs, ys in ((w, np.full_like(yc, xc[0]), yc), (e, np.full_like(yc, xc[-1]), yc),
                        (s, xc, np.full_like(xc, yc[0])), (nn, xc, np.full_like(xc, yc[-1]))):
        for k, key in enumerate("uvq"):
            assert np.abs(arr[:, k] - f[key](xs, ys)).max() < 1e-9
    # a child whose west edge samples the dry strip gets zeros there
    maps2 = O.nest_maps(xp, yp, np.linspace(150.0, 900.0, 20), yc)
    _, w2, _, _, _ = O.feed_child(o, maps2, cuke=0.001, hmin=0.01, iarv=True)
    assert np.all(w2 == 0.0)


def test_feed_arrival_gate():
    nx, ny = 30, 30
    xp = np.arange(nx) * 100.0
    yp = np.arange(ny) * 100.0
    dep = np.full((ny, nx), 20.0)
    cel = np.sqrt(G * dep)
    s1, s2, zeta = O.metrics(xp, yp, True)
    o = O.Oracle(dep, s1, s2, zeta, dt=1.0, ground=0.01)
    o.set_state(cel, np.zeros_like(dep), np.zeros_like(dep))
    maps = O.nest_maps(xp, yp, np.linspace(800, 1500, 15), np.linspace(900, 1600, 15))
    cuke = 0.05
    assert not O.feed_child(o, maps, cuke=cuke, hmin=0.01, iarv=False)[0]
    # |eta| just above 0.1*cuke inside the footprint (0.1 is an FP32 literal)
    thr = float(np.float32(0.1)) * cuke
    eta = np.zeros_like(dep)
    eta[12, 12] = thr * 1.001
    o.set_state(np.sqrt(G * (dep + eta)), np.zeros_like(dep), np.zeros_like(dep))
    assert O.feed_child(o, maps, cuke=cuke, hmin=0.01, iarv=False)[0]
    # outside the footprint: no arrival
    eta[:] = 0
    eta[2, 2] = 1.0
    o.set_state(np.sqrt(G * (dep + eta)), np.zeros_like(dep), np.zeros_like(dep))
    assert not O.feed_child(o, maps, cuke=cuke, hmin=0.01, iarv=False)[0]
    # momentum criterion (|u|+|v|) c > cuke
    u = np.zeros_like(dep)
    u[13, 11] = 1.1 * cuke / math.sqrt(G * 20.0)
    o.set_state(np.sqrt(G * dep), u, np.zeros_like(dep))
    assert O.feed_child(o, maps, cuke=cuke, hmin=0.01, iarv=False)[0]


def test_maxwave_fields_recover_surface_maximum():
    """output_maxwave.f:14-24: MaxE = hmax^2/g - d returns the maximum surface
    elevation for a celerity maximum c = sqrt(g (d + eta)); cells whose
    maximum never exceeded celmin are the real*4 NAN in both fields."""
    import cliffs_gpu
    rng = np.random.default_rng(4)
    g = synth.GRAV
    ground = 0.01
    cm = np.sqrt(g * ground)
    dep = rng.uniform(-5.0, 4000.0, (40, 50))
    eta = rng.uniform(-0.5, 3.0, dep.shape)
    H = dep + eta
    hmax = np.where(H > ground, np.sqrt(g * np.maximum(H, 0.0)), 0.0)
    u = rng.normal(0, 0.3, dep.shape)
    v = rng.normal(0, 0.3, dep.shape)
    E, V = cliffs_gpu.maxwave_fields(hmax, u * u + v * v, dep, g, cm)
    wet = hmax > cm
    assert E.dtype == np.float32 and V.dtype == np.float32
    assert np.all(E[~wet] == np.float32(-1e32)) and np.all(V[~wet] == np.float32(-1e32))
    assert np.allclose(E[wet], eta[wet], atol=1e-9 * 4000 + 4e-4)   # FP32 rounding of ~4000 m terms
    assert np.allclose(V[wet], np.hypot(u, v)[wet].astype(np.float32), rtol=1e-6)
