"""The GPU driver (cliffs-src_amd/fortran/cliffs_gpu_driver): the
reference's command line, parameter file and NetCDF inputs and outputs
(cliffs_main.f), time steps on the GPU, files through the NetCDF-3 shim.

1. The reference's slpbeach example end to end (its own files and parameter
   file): every sea-state snapshot and the max-wave file against the FP64
   oracle at the parity bar (surface elevation within 1e-5 of the amplitude
   on cells wet in both runs, at most one wet/dry flip, run-up within one
   cell).
2. A two-level nest cascade on synthetic grid and source files: the
   parent run's snapshots, gauges, max-wave and child feed files, and the
   child run driven by those feed files, each equal bit for bit to the same
   steps driven through the C ABI from Python (cliffs_main.f:190-378 restated
   in the test).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator
from scipy.io import netcdf_file

import oracle as O
import synth
from test_oracle import slpbeach_case

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "cliffs-src_amd", "fortran", "cliffs_gpu_driver")
SLP = os.path.join(ROOT, "tests", "golden", "slpbeach")
NAN4 = np.float32(-1.0e32)
G = synth.GRAV


def _drive(args, log, env=None):
    if not os.path.exists(DRIVER):
        pytest.fail("cliffs_gpu_driver not built (cliffs-src_amd/fortran/build.sh)")
    print("driver:", " ".join(args), flush=True)
    out = subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=300,
                         env=None if env is None else dict(os.environ, **env))
    text = open(log).read() if os.path.exists(log) else ""
    assert out.returncode == 0, out.stdout + out.stderr + text
    return text


def _var(path, name):
    """A variable's values in native byte order."""
    with netcdf_file(path, "r", mmap=False) as f:
        v = np.array(f.variables[name][:])
    return v.astype(v.dtype.newbyteorder("="))


def _same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    bad = int((a.view(np.uint8) != b.view(np.uint8)).reshape(a.size, -1).any(axis=1).sum()) if a.size else 0
    assert a.dtype == b.dtype and bad == 0, f"{what}: {bad} values differ"


# --------------------------------------------------------------------------
# 1. slpbeach

def test_slpbeach_against_oracle(tmp_path, gpu_lib):
    out = str(tmp_path) + "/"
    log = _drive([out + "slp", SLP + "/", "0", "solitary.0185", SLP + "/slpbeach_params.txt"],
                 out + "slp_log.txt")
    # each screenshot logs the state's largest Courant number (cliffs_health)
    cour = [float(l.split()[-1]) for l in log.splitlines() if "max Courant number" in l]
    assert len(cour) == 2000 // 3 + 1 and 0.0 < max(cour) < 1.0
    assert "not finite" not in log
    case = slpbeach_case("0185")
    amp = 0.0185 * case["dep"][0, -1]
    o = O.Oracle(case["dep"], case["s1"], case["s2"], case["zeta"], dt=case["dt"], ground=case["ground"])
    o.set_state(case["cel"], case["xvel"], case["yvel"])
    ha = _var(out + "slp_sea_h.nc", "ha")[:, 0, :]           # (shots, x)
    tt = _var(out + "slp_sea_h.nc", "time")
    assert ha.shape == (2000 // 3 + 1, 384) and np.allclose(np.diff(tt), 3 * case["dt"])
    dep = case["dep"][0]
    celmin = np.sqrt(G * case["ground"])
    worst, flips = 0.0, 0
    for k in range(ha.shape[0]):
        if k:
            o.steps(3)
        c = o.get_state()[0][0]
        wet_o = c > celmin
        eta_o = c * c / G - dep
        wet_d = ha[k] != NAN4
        flips = max(flips, int((wet_o != wet_d).sum()))
        both = wet_o & wet_d
        worst = max(worst, float(np.abs(ha[k][both].astype(np.float64) - eta_o[both]).max()))
    o.steps(2000 - 3 * (ha.shape[0] - 1))
    hm = o.get_state()[3][0]
    maxe = _var(out + "slp_maxwave.nc", "MaxE")[0]
    wet_o = hm > celmin
    wet_d = maxe != NAN4
    both = wet_o & wet_d
    me = float(np.abs(maxe[both].astype(np.float64) - (hm * hm / G - dep)[both]).max())
    runup_cells = abs(int(np.flatnonzero(wet_o).max()) - int(np.flatnonzero(wet_d).max()))
    print(f"slpbeach: eta err {worst:.3e}, MaxE err {me:.3e} (bar {1e-5 * amp:.3e}), flips {flips}, "
          f"run-up cells {runup_cells}")
    assert worst <= 1e-5 * amp and me <= 1e-5 * amp
    assert flips <= 1 and runup_cells <= 1


def test_slpbeach_fp64_line_mode_bit_identical(tmp_path, gpu_lib):
    """CLIFFS_PRECISION=fp64: the driver steps the FP64 line mode; every
    screenshot and the max-wave file hold the FP64 oracle's values rounded
    to the files' real*4, bit for bit (wet cells; the NAN value on dry
    ones)."""
    out = str(tmp_path) + "/"
    log = _drive([out + "slp", SLP + "/", "0", "solitary.0185", SLP + "/slpbeach_params.txt"],
                 out + "slp_log.txt", env={"CLIFFS_PRECISION": "fp64"})
    assert "FP64 line mode" in log and "not finite" not in log
    cour = [float(l.split()[-1]) for l in log.splitlines() if "max Courant number" in l]
    assert len(cour) == 2000 // 3 + 1 and 0.0 < max(cour) < 1.0
    case = slpbeach_case("0185")
    # the initial fields as the reference reads them: real*4
    # (get_initial_state.f:128-150), then added to depth and velocity
    # (apply_initial_conditions.f) and turned into celerity
    z = np.load(os.path.join(ROOT, "tests", "golden", "slpbeach.npz"))
    hq = z["h0185"].astype(np.float32).astype(np.float64).reshape(1, -1)
    uq = z["u0185"].astype(np.float32).astype(np.float64).reshape(1, -1)
    zero = np.zeros(1)
    xvel = O.apply_initial_conditions(z["xq"], zero, uq, z["x"], z["y"], np.zeros_like(case["dep"]))
    H = O.apply_initial_conditions(z["xq"], zero, hq, z["x"], z["y"], case["dep"])
    o = O.Oracle(case["dep"], case["s1"], case["s2"], case["zeta"], dt=case["dt"], ground=case["ground"])
    o.set_state(synth.depth_to_celerity(H, case["ground"]), xvel, np.zeros_like(case["dep"]))
    ha = _var(out + "slp_sea_h.nc", "ha")[:, 0, :]
    dep = case["dep"][0]
    celmin = np.sqrt(G * case["ground"])

    def field(c):
        return np.where(c > celmin, (c * c / G - dep).astype(np.float32), NAN4)

    for k in range(ha.shape[0]):
        if k:
            o.steps(3)
        _same(ha[k], field(o.get_state()[0][0]), f"screenshot {k}")
    o.steps(2000 - 3 * (ha.shape[0] - 1))
    _same(_var(out + "slp_maxwave.nc", "MaxE")[0], field(o.get_state()[3][0]), "MaxE")


# --------------------------------------------------------------------------
# 2. two-level cascade

_NC3 = os.path.join(ROOT, "cliffs-src_amd", "netcdf3", "libcliffs_nc3.so")


def _nc_write(path, dims, variables):
    """A CDF-1 file with dimensions and double variables in the given order
    (the reference's readers take variables by position: scipy's writer
    reorders them by shape, so the shim writes these; its writer is checked
    against scipy's reader in test_netcdf3.py)."""
    lib = C.CDLL(_NC3)
    ncid = C.c_int()
    assert lib.nc3_create(path.encode(), 0, C.byref(ncid)) == 0
    did = {}
    for name, ln in dims:
        d = C.c_int()
        assert lib.nc3_def_dim(ncid, name.encode(), C.c_size_t(ln), C.byref(d)) == 0
        did[name] = d.value
    vid = []
    for name, vdims, _ in variables:
        v = C.c_int()
        ids = (C.c_int * len(vdims))(*[did[n] for n in vdims])
        assert lib.nc3_def_var(ncid, name.encode(), 6, len(vdims), ids, C.byref(v)) == 0
        vid.append(v.value)
    assert lib.nc3_enddef(ncid) == 0
    for v, (_, _, val) in zip(vid, variables):
        buf = np.ascontiguousarray(val, dtype=np.float64)
        assert lib.nc3_put_vara(ncid, v, None, None, buf.ctypes.data_as(C.c_void_p), 6) == 0
    assert lib.nc3_close(ncid) == 0


def _write_grid(path, x, y, dep):
    _nc_write(path, [("xxx", len(x)), ("yyy", len(y))],
              [("xxx", ("xxx",), x), ("yyy", ("yyy",), y), ("bathy", ("yyy", "xxx"), dep)])


def _write_field(path, name, x, y, val):
    _nc_write(path, [("xxx", len(x)), ("yyy", len(y)), ("time", 1)],
              [("xxx", ("xxx",), x), ("yyy", ("yyy",), y), ("time", ("time",), [0.0]),
               (name, ("time", "yyy", "xxx"), val[None])])


def _params(path, grid, nest, cuke, ground, dt, steps, freeze_flag, seaout, sub, bndout, maxout, gauges):
    lines = ["1", grid, str(1 if nest else 0)] + ([nest] if nest else [])
    lines += [repr(cuke), repr(ground), "0.0", "1", "5", repr(dt), str(steps), "0", str(freeze_flag),
              str(seaout), str(sub), str(sub), str(bndout), str(maxout), str(len(gauges))]
    if gauges:
        lines += ["5"] + [f"{i} {j}" for i, j in gauges]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def test_nest_parent_fp64_line_mode_matches_oracle(tmp_path, gpu_lib):
    """CLIFFS_PRECISION=fp64 on a parent grid with a nest: screenshots,
    gauges, max-wave file and the four feed files equal, bit for bit, the
    FP64 oracle stepped the same way with feed_children_grid every bndout
    steps (the feed from the FP64 parent state, cliffs64_nest_feed)."""
    d = str(tmp_path) + "/"
    case = synth.coastal(nx=200, ny=160, seed=31)
    x, y, dep = case.xcrd, case.ycrd, case.dep
    xc = np.linspace(x[100], x[185], 3 * 85 + 1)
    yc = np.linspace(y[30], y[120], 3 * 90 + 1)
    interp = RegularGridInterpolator((y, x), dep)
    Yc, Xc = np.meshgrid(yc, xc, indexing="ij")
    depc = interp(np.stack([Yc, Xc], axis=-1))
    eta = np.where(dep > 0, case.cel ** 2 / G - dep, 0.0).astype(np.float32).astype(np.float64)
    _write_grid(d + "parent.nc", x, y, dep)
    _write_grid(d + "child.nc", xc, yc, depc)
    _write_field(d + "src_h.nc", "ha", x, y, eta)
    steps, bndout, seaout, gout, maxout, sub = 80, 2, 20, 5, 25, 2
    gauges = [(150, 80), (120, 40)]
    cuke, ground, dt = 0.01, case.ground, case.dt
    _params(d + "pp.txt", "parent.nc", "child.nc", cuke, ground, dt, steps, 1, seaout, sub, bndout,
            maxout, gauges)
    log = _drive([d + "par", d, "0", "src", d + "pp.txt"], d + "par_log.txt", env={"CLIFFS_PRECISION": "fp64"})
    assert "FP64 line mode" in log

    s1, s2, zeta = synth.metrics(x, y, True)
    o = O.Oracle(dep, s1, s2, zeta, dt=dt, ground=ground)
    z = np.zeros_like(dep)
    o.set_state(synth.depth_to_celerity(dep + eta, ground), z, z)
    maps = O.nest_maps(x, y, xc, yc)
    celmin = np.sqrt(G * ground)
    shots, gts, feeds = [], [], []
    I = np.array([i for i, _ in gauges])
    J = np.array([j for _, j in gauges])
    arrived = False
    t = 0.0

    def shot():
        c, u, v = o.get_state()[:3]
        c, u, v = c[::sub, ::sub], u[::sub, ::sub], v[::sub, ::sub]
        wet = c > celmin
        sh = [np.where(wet, c * c / G - dep[::sub, ::sub], 0).astype(np.float32), u.astype(np.float32),
              v.astype(np.float32)]
        shots.append((t, [np.where(wet, a, NAN4) for a in sh]))

    def gauge():
        c, u, v = o.get_state()[:3]
        cg = c[J - 1, I - 1]
        gts.append((t, (np.where(cg > celmin, cg * cg / G - dep[J - 1, I - 1], -1e32), u[J - 1, I - 1],
                        v[J - 1, I - 1])))

    def feed():
        nonlocal arrived
        arrived, w, e, s_, n_ = O.feed_child(o, maps, cuke, ground, arrived)
        if arrived:
            feeds.append((t, w, e, s_, n_))

    shot()
    gauge()
    feed()
    for k in range(1, steps + 1):
        t = t + dt
        o.steps(1)
        if k % bndout == 0:
            feed()
        if k % seaout == 0:
            shot()
        if k % gout == 0:
            gauge()
    hm, vm = o.get_state()[3:5]
    assert len(feeds) >= 10
    for q, name in enumerate(("h", "u", "v")):
        path = d + f"par_sea_{name}.nc"
        _same(_var(path, "time"), np.array([s[0] for s in shots]), f"sea_{name} time")
        _same(_var(path, name + "a"), np.array([s[1][q] for s in shots]), f"sea_{name}")
    for q, name in enumerate(("gage", "u", "v")):
        _same(_var(d + "par_gages.nc", name), np.array([s[1][q] for s in gts]), f"gauge {name}")
    wet = hm > celmin
    _same(_var(d + "par_maxwave.nc", "MaxE"), np.where(wet, hm * hm / G - dep, NAN4).astype(np.float32), "MaxE")
    _same(_var(d + "par_maxwave.nc", "MaxV"), np.where(wet, np.sqrt(vm), NAN4).astype(np.float32), "MaxV")
    for q, side in enumerate(("west", "east", "south", "north")):
        path = d + f"par_child_{side}.nc"
        _same(_var(path, "time"), np.array([f[0] for f in feeds]), f"feed {side} time")
        _same(_var(path, "vals"), np.array([f[1 + q].T for f in feeds]), f"feed {side}")


def test_nest_cascade_bit_identical_to_c_abi(tmp_path, gpu_lib):
    d = str(tmp_path) + "/"
    case = synth.coastal(nx=200, ny=160, seed=31)
    x, y, dep = case.xcrd, case.ycrd, case.dep
    xc = np.linspace(x[100], x[185], 3 * 85 + 1)
    yc = np.linspace(y[30], y[120], 3 * 90 + 1)
    interp = RegularGridInterpolator((y, x), dep)
    Yc, Xc = np.meshgrid(yc, xc, indexing="ij")
    depc = interp(np.stack([Yc, Xc], axis=-1))
    # the source: the case's initial surface, in values a real*4 read keeps
    eta = np.where(dep > 0, case.cel ** 2 / G - dep, 0.0).astype(np.float32).astype(np.float64)
    _write_grid(d + "parent.nc", x, y, dep)
    _write_grid(d + "child.nc", xc, yc, depc)
    _write_field(d + "src_h.nc", "ha", x, y, eta)
    steps, bndout, seaout, gout, maxout, sub = 80, 2, 20, 5, 25, 2
    gauges = [(150, 80), (120, 40)]
    cuke, ground, dt = 0.01, case.ground, case.dt
    _params(d + "pp.txt", "parent.nc", "child.nc", cuke, ground, dt, steps, 1, seaout, sub, bndout,
            maxout, gauges)
    _drive([d + "par", d, "0", "src", d + "pp.txt"], d + "par_log.txt")

    # ---- the parent through the C ABI
    s1, s2, zeta = synth.metrics(x, y, True)
    g = gpu_lib.Grid(dep, s1, s2, zeta, dt=dt, ground=ground)
    cel = synth.depth_to_celerity(dep + eta, ground)
    z = np.zeros_like(dep)
    g.upload(cel, z, z)
    nest = gpu_lib.Nest(g, gpu_lib.nest_maps(x, y, xc, yc))
    celmin = np.sqrt(G * ground)
    shots, gts, feeds = [], [], []
    I = np.array([i for i, _ in gauges])
    J = np.array([j for _, j in gauges])
    arrived = False
    t = 0.0

    def shot():
        c, u, v = g.download()
        c, u, v = c[::sub, ::sub], u[::sub, ::sub], v[::sub, ::sub]
        wet = c > celmin
        sh = [np.where(wet, c * c / G - dep[::sub, ::sub], 0).astype(np.float32), u.astype(np.float32),
              v.astype(np.float32)]
        shots.append((t, [np.where(wet, a, NAN4) for a in sh]))

    def feed():
        nonlocal arrived
        arrived, fed, w, e, s, n = nest.feed(cuke, ground, arrived)
        if fed:
            feeds.append((t, w, e, s, n))

    shot()
    gts.append((t, g.gauges(I, J)))
    feed()
    for k in range(1, steps + 1):
        t = t + dt
        g.step()
        if k % bndout == 0:
            feed()
        if k % seaout == 0:
            shot()
        if k % gout == 0:
            gts.append((t, g.gauges(I, J)))
    hm, vm = g.download_max()
    nest.close()
    g.close()
    assert len(feeds) >= 10

    for q, name in enumerate(("h", "u", "v")):
        path = d + f"par_sea_{name}.nc"
        _same(_var(path, "time"), np.array([s[0] for s in shots]), f"sea_{name} time")
        _same(_var(path, name + "a"), np.array([s[1][q] for s in shots]), f"sea_{name}")
    _same(_var(d + "par_gages.nc", "time"), np.array([s[0] for s in gts]), "gauge time")
    for q, name in enumerate(("gage", "u", "v")):
        _same(_var(d + "par_gages.nc", name), np.array([s[1][q] for s in gts]), f"gauge {name}")
    wet = hm > celmin
    _same(_var(d + "par_maxwave.nc", "MaxE"), np.where(wet, hm * hm / G - dep, NAN4).astype(np.float32), "MaxE")
    _same(_var(d + "par_maxwave.nc", "MaxV"), np.where(wet, np.sqrt(vm), NAN4).astype(np.float32), "MaxV")
    for q, side in enumerate(("west", "east", "south", "north")):
        path = d + f"par_child_{side}.nc"
        _same(_var(path, "time"), np.array([f[0] for f in feeds]), f"feed {side} time")
        _same(_var(path, "vals"), np.array([f[1 + q].T for f in feeds]), f"feed {side}")

    # ---- the child from the parent's feed files (get_boundary_feed.f, cliffs_main.f:303-345)
    dtc = dt / 3
    cseaout = 20
    _params(d + "pc.txt", "child.nc", None, cuke, ground, dtc, 10_000, 0, cseaout, 1, 1, 200, [])
    log = _drive([d + "chi", d, "par", "0", d + "pc.txt"], d + "chi_log.txt")
    assert "bndr feed ended" in log
    print([ln for ln in log.splitlines() if "ended" in ln or "threshold" in ln])
    tin = np.array([f[0] for f in feeds])
    sides = [np.stack([f[1 + q] for f in feeds], axis=-1) for q in range(4)]   # (npts, 3, nt)
    nt = len(tin)
    first = next(k for k in range(nt) if max(np.abs(s[:, 2, k]).max() for s in sides) >= cuke)
    assert first < nt - 1
    cs1, cs2, czeta = synth.metrics(xc, yc, True)
    c = gpu_lib.Grid(depc, cs1, cs2, czeta, dt=dtc, ground=ground)
    zc = np.zeros_like(depc)
    c.upload(synth.depth_to_celerity(depc, ground), zc, zc)
    irec, t, nsteps = first, tin[first], 0      # 0-based irec: tin[irec-1] <= t < tin[irec]
    mx = c.mx
    cshots = []

    def cshot():
        cc_, uu, vv = c.download()
        cshots.append((t, np.where(cc_ > celmin, cc_ * cc_ / G - depc, NAN4).astype(np.float32)))

    cshot()
    while True:
        t = t + dtc
        while t >= tin[irec]:
            if irec == nt - 1:
                break
            irec += 1
        if t >= tin[irec]:            # the feed ended: the run stops (freeze)
            break
        cc = (tin[irec] - t) / (tin[irec] - tin[irec - 1])
        edges = []
        for s in sides:
            e = np.zeros((mx, 3))
            e[:s.shape[0]] = s[:, :, irec - 1] * cc + s[:, :, irec] * (1 - cc)
            edges.append(e)
        c.step(*edges)
        nsteps += 1
        if nsteps % cseaout == 0:
            cshot()
    hmc, vmc = c.download_max()
    got_t = _var(d + "chi_sea_h.nc", "time")
    got_h = _var(d + "chi_sea_h.nc", "ha")
    print(f"child: {nsteps} steps, first record {first} of {nt}; {len(cshots)} snapshots here, "
          f"{len(got_t)} from the driver")
    for k, (tk, hk) in enumerate(cshots[:len(got_t)]):
        if tk != got_t[k] or not np.array_equal(hk, got_h[k]):
            print(f"first differing child snapshot {k}: t {tk!r} vs {got_t[k]!r}, "
                  f"{int((hk != got_h[k]).sum())} cells")
            break
    _same(got_t, np.array([s_[0] for s_ in cshots]), "child snapshot times")
    _same(got_h, np.array([s_[1] for s_ in cshots]), "child snapshots")
    c.close()
    assert nsteps > 50
    wet = hmc > celmin
    _same(_var(d + "chi_maxwave.nc", "MaxE"), np.where(wet, hmc * hmc / G - depc, NAN4).astype(np.float32),
          "child MaxE")
    _same(_var(d + "chi_maxwave.nc", "MaxV"), np.where(wet, np.sqrt(vmc), NAN4).astype(np.float32),
          "child MaxV")
    assert float(np.where(wet & (depc > 0), hmc * hmc / G - depc, 0).max()) > 1e-3   # the wave came in
