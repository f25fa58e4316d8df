"""Benchmark: VTCS-2 step throughput on one MI355X.

Workload: 2-D Cartesian open-ocean propagation, 4096 x 4096 FP32 grid,
dx = dy = 500 m, depth 4000 m minus seeded fractal seamounts (never dry),
Gaussian hump (1 m, sigma 20 km), zero velocity, open boundaries,
non-dispersive; default 10,000 timed steps.  One step = x-sweep over rows,
y-sweep over columns, max-wave update (cliffs_main.f:317-355).  State is
resident in HBM before the timed region; no host copies inside it.

Prints ONE JSON line.  roofline.achieved counts the algorithmic bytes of the
dominant kernel over its mean HIP-event duration in the timed region.  The
default step is one fused kernel (x-sweep, y-sweep and max-wave of a step in
one pass), counted at 28 B per cell-update: the state read once (s, cp, u,
v) and written once (cp, u, v) in FP32.  The max-wave update it also
carries (read and write hmax, velmax: 16 B) is reported separately, as
roofline.with_maxwave (44 B).  With --unfused the step is two sweep kernels
of 28 B each; the y-sweep carries the max-wave update.  --coastal
times a spherical 8192 x 6144 coastal grid with run-up instead (land, dry
lines and wet/dry fronts in the tiles); the default is the line above.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cliffs-src_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

METRIC = ("Gcell-updates/s & %HBM roofline; ensemble scenarios/hr at ≤2% "
          "max-runup error")
HBM_PEAK_GBS = 8000.0
SWEEP_BYTES = 28  # one pass over the state: read s, cp, u, v; write cp, u, v (FP32)
MW_BYTES = 16     # max-wave update: read + write hmax, velmax (FP32)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pmc_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on the workload ("open_ocean_4096",
    "coastal_8192x6144"), from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (profiles/), already scaled by the plain-copy
    calibration; None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{workload}.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            t = line.split()
            if len(t) >= 2 and t[0] == f"{kernel}_bytes_per_launch":
                return int(t[1])
    return None


def ensemble_bench(nscen, steps, seed0, workers=8):
    """Batched ensemble (config 4 at a sample size): nscen scenarios of
    uneven size (512..2048 per side, dx 200-1000 m, shelf and coastline,
    dipole sources), all advanced `steps` steps in one fused launch per step
    replayed from a HIP graph, tile height picked by the autotuner.
    scenarios_per_hr counts the wall time from device allocation and upload
    to the per-scenario maxima on the host (inputs generated beforehand, by
    `workers` processes, or in this process when 0)."""
    import cliffs_gpu
    import synth
    log(f"[bench] ensemble: generating {nscen} scenarios")
    import functools
    import multiprocessing as mp
    gen = functools.partial(synth.ensemble_scenario, steps=steps)
    if workers > 0:
        # fresh interpreters (spawn): the workers run numpy only and never touch the GPU
        with mp.get_context("spawn").Pool(min(workers, os.cpu_count() or 1)) as pool:
            cases = pool.map(gen, range(seed0, seed0 + nscen), chunksize=4)
    else:
        cases = [gen(s) for s in range(seed0, seed0 + nscen)]
    cells = sum(c.dep.size for c in cases)
    cliffs_gpu.load_library()

    def make(c):
        s1, s2, zeta = c.metrics()
        g = cliffs_gpu.Grid(c.dep, s1, s2, zeta, dt=c.dt, ground=c.ground, crough=c.crough)
        g.upload(c.cel, c.xvel, c.yvel)
        return g

    def runup(gc):
        g, c = gc
        return synth.max_runup(g.download_max()[0], c.dep, g.celmin)

    # grids are created, uploaded and read back from host threads, one grid
    # (handle) per task: the library's calls release the interpreter lock, so
    # one grid's conversions overlap another's copies
    from concurrent.futures import ThreadPoolExecutor
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=8) as ex:
        grids = list(ex.map(make, cases))
    ens = cliffs_gpu.Ensemble(grids)
    rows = ens.tune(2)
    width = ens.shape()[0]
    log(f"[bench] ensemble: {cells / 1e6:.1f} M cells, tiles {width} x {rows}, {steps} steps")
    ms = ens.run_timed(steps)
    with ThreadPoolExecutor(max_workers=8) as ex:
        runups = list(ex.map(runup, zip(grids, cases)))
    wall = time.perf_counter() - t0
    ens.close()
    for g in grids:
        g.close()
    parity = runup_parity(seed0, nscen, steps, runups, cases)
    return {"scenarios": nscen, "steps": steps, "cells": cells, "tile_width": width, "tile_rows": rows,
            "wall_s": round(wall, 3), "device_ms": round(ms, 1),
            "scenarios_per_hr": round(nscen / (wall / 3600.0), 1),
            "gcell_updates_per_s": round(cells * steps / (ms * 1e-3) / 1e9, 2),
            "max_runup_m": [round(r, 2) for r in runups],
            "runup_vs_oracle": parity,
            "sample": f"{nscen} scenarios (seeds {seed0}..{seed0 + nscen - 1}), {steps} steps each, "
                      f"friction n^2 = {cases[0].crough}, minimal flow depth {cases[0].ground} m"}


def runup_parity(seed0, nscen, steps, runups, cases):
    """Max run-up of the scenarios that the FP64 oracle has run at full size
    (tests/golden/ensemble_runup.json, tests/golden/make_ensemble_runup.py)
    against the device's: worst difference in land cells and relative."""
    path = os.path.join(ROOT, "tests", "golden", "ensemble_runup.json")
    if not os.path.exists(path):
        return None
    import json
    import synth
    fx = json.load(open(path))
    rows = []
    for k in range(nscen):
        ref = fx.get(str(seed0 + k))
        if ref is None or ref["steps"] != steps:
            continue
        d = runups[k] - ref["max_runup_m"]
        rows.append((seed0 + k, round(runups[k], 2), round(ref["max_runup_m"], 2),
                     round(abs(d) / synth.land_cell_rise(cases[k]), 3),
                     round(abs(d) / max(ref["max_runup_m"], 1e-9), 4)))
    if not rows:
        return None
    return {"seeds": [r[0] for r in rows], "device_m": [r[1] for r in rows],
            "oracle_m": [r[2] for r in rows],
            "max_cells_off": max(r[3] for r in rows), "max_rel_err": max(r[4] for r in rows),
            "within_one_cell": sum(1 for r in rows if r[3] <= 1.0), "of": len(rows)}


def nested_bench(hours, every=900):
    """The nested workload: the spherical 8192 x 6144 coastal parent (30
    arc-seconds, 100 km shelf, run-up, 200 x 100 km dipole; synth.coastal_nest)
    and its 2048^2 child at 3x on the coast, coupled on the device: a feed
    record after every parent step (cliffs_nest_record), the child started at
    the first record whose edge |eta| reaches 1 mm (cliffs_nest_child_start,
    the reference examples' threshold) and stepped at dt/3 behind the feed
    (cliffs_nest_child_step) -- for `hours` simulated hours.  Every `every`
    parent steps the health of both grids (cliffs_health) and the child's
    maximum run-up are logged.  value counts parent and child cell-updates
    over the whole run's wall time."""
    import numpy as np
    import cliffs_gpu
    import synth
    log("[bench] nested: generating the 8192x6144 parent and its 2048^2 child")
    case, xc, yc, depc, meta = synth.coastal_nest()
    maps = cliffs_gpu.nest_maps(case.xcrd, case.ycrd, xc, yc)
    s1, s2, zeta = case.metrics()
    g = cliffs_gpu.Grid(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground,
                        crough=case.crough, device=0)
    g.upload(case.cel, case.xvel, case.yvel)
    cs1, cs2, czeta = synth.metrics(xc, yc, False)
    dtc = case.dt / 3
    cg = meta["child_ground"]
    c = cliffs_gpu.Grid(depc, cs1, cs2, czeta, dt=dtc, ground=cg, crough=case.crough, device=0)
    z = np.zeros_like(depc)
    c.upload(synth.depth_to_celerity(depc, cg), z, z)
    nest = cliffs_gpu.Nest(g, maps)
    nest.couple(c, capacity=8)
    nsteps = int(round(hours * 3600.0 / case.dt))
    tp, iarv, trec, tc, nchild, t0c = 0.0, False, [], None, 0, None
    health = []
    log(f"[bench] nested: {nsteps} parent steps of {case.dt:.3f} s ({hours} h)")
    t0 = time.perf_counter()
    for k in range(1, nsteps + 1):
        g.run(1)
        tp += case.dt
        iarv, nrec = nest.record(0.01, case.ground, tp, iarv)
        if nrec > len(trec):
            trec.append(tp)
        if tc is None and nrec:
            try:
                tc = t0c = nest.child_start(0.001)
            except cliffs_gpu.CliffsError:
                pass
        while tc is not None and tc + dtc < trec[-1]:
            tc += dtc
            nest.child_step(tc)
            nchild += 1
        if k % every == 0 or k == nsteps:
            cp, bp = g.health()
            cc, bc = c.health()
            ru = synth.max_runup(c.download_max()[0], depc, c.celmin)
            health.append({"step": k, "t_s": round(tp, 1), "parent_courant": round(cp, 4),
                           "parent_nonfinite": bp, "child_courant": round(cc, 4),
                           "child_nonfinite": bc, "child_steps": nchild,
                           "child_max_runup_m": round(ru, 2)})
            log(f"[bench] nested: {health[-1]}")
    g.sync()
    c.sync()
    wall = time.perf_counter() - t0
    hm_p = g.download_max()[0]
    finite = all(h["parent_nonfinite"] == 0 and h["child_nonfinite"] == 0 for h in health)
    _, ms_p, ms_fix = g.run_timed(200)    # the parent's step kernels alone, for their roofline
    cells_p, cells_c = case.dep.size, depc.size
    value = (cells_p * nsteps + cells_c * nchild) / wall / 1e9
    log(f"[bench] nested: {nsteps} parent steps and {nchild} child steps in {wall:.3f} s, "
        f"{value:.2f} Gcell-updates/s")
    out = {"parent_steps": nsteps, "child_steps": nchild, "records": len(trec),
           "child_start_s": None if t0c is None else round(t0c, 1),
           "parent_max_runup_m": round(synth.max_runup(hm_p, case.dep, g.celmin), 2),
           "health": health, "meta": meta}
    nest.close()
    c.close()
    g.close()
    return value, wall, out, (ms_p, ms_fix), cells_p, cells_c, case.dt, finite


def cpu_baseline(case, steps):
    """FP64 C restatement (oracle/) on the host cores, on a bounded sample of
    the same workload: `steps` steps of the full grid."""
    import oracle
    if not os.path.exists(oracle.ORACLE_SO):
        oracle.build(with_ref=False)
    s1, s2, zeta = case.metrics()
    o = oracle.Oracle(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground)
    o.set_state(case.cel, case.xvel, case.yvel)
    o.steps(1)
    t0 = time.perf_counter()
    o.steps(steps)
    dt = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ny, nx = case.dep.shape
    name = case.name.replace("_", "-")
    return {"value": nx * ny * steps / dt / 1e9, "unit": "Gcell-updates/s", "cores": cores,
            "kind": "port",
            "sample": f"{steps} steps of the {nx}x{ny} {name} grid after 1 warm-up step, "
                      f"FP64 C restatement of cliffs.f with OpenMP over lines, "
                      f"{dt:.1f} s wall"}, o.get_state()


def fp64_bench(case, steps, want=None):
    """The FP64 line mode (Grid64) on the same sample as the CPU baseline:
    1 warm-up step, then `steps` timed steps; with `want` (the CPU
    baseline's final state and maxima), whether they are bit-identical."""
    import numpy as np

    import cliffs_gpu
    s1, s2, zeta = case.metrics()
    g = cliffs_gpu.Grid64(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground)
    g.upload(case.cel, case.xvel, case.yvel)
    g.run(1)
    t0 = time.perf_counter()
    g.run(steps)
    dt = time.perf_counter() - t0
    got = tuple(g.download()) + tuple(g.download_max())
    g.close()
    ny, nx = case.dep.shape
    same = None
    if want is not None:
        same = all(np.array_equal(a.view(np.int64), b.view(np.int64)) for a, b in zip(got, want))
    return {"value": round(nx * ny * steps / dt / 1e9, 4), "unit": "Gcell-updates/s", "steps": steps,
            "ms_per_step": round(dt / steps * 1e3, 3), "dtype": "f64",
            "bit_identical_to_cpu_baseline": same}


def copy_bandwidth(mib=1024, reps=5):
    """GB/s of a plain device-to-device copy of `mib` MiB (bytes read plus
    bytes written, per second; HIP events around `reps` hipMemcpyAsync on
    the null stream after one untimed copy): the bandwidth a streaming
    kernel can reach on this device, beside the 8 TB/s spec."""
    import ctypes as C
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    vp, n = C.c_void_p, C.c_size_t(mib << 20)
    a, b, e0, e1, ms = vp(), vp(), vp(), vp(), C.c_float()

    def chk(rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc})")
    hip.hipMemcpyAsync.argtypes = [vp, vp, C.c_size_t, C.c_int, vp]
    hip.hipMemset.argtypes = [vp, C.c_int, C.c_size_t]
    hip.hipEventRecord.argtypes = [vp, vp]
    chk(hip.hipMalloc(C.byref(a), n), "hipMalloc")
    try:
        chk(hip.hipMalloc(C.byref(b), n), "hipMalloc")
        chk(hip.hipMemset(a, 1, n), "hipMemset")
        chk(hip.hipEventCreate(C.byref(e0)), "hipEventCreate")
        chk(hip.hipEventCreate(C.byref(e1)), "hipEventCreate")
        chk(hip.hipMemcpyAsync(b, a, n, 3, None), "hipMemcpyAsync")      # 3: device to device
        chk(hip.hipEventRecord(e0, None), "hipEventRecord")
        for _ in range(reps):
            chk(hip.hipMemcpyAsync(b, a, n, 3, None), "hipMemcpyAsync")
        chk(hip.hipEventRecord(e1, None), "hipEventRecord")
        chk(hip.hipEventSynchronize(e1), "hipEventSynchronize")
        chk(hip.hipEventElapsedTime(C.byref(ms), e0, e1), "hipEventElapsedTime")
    finally:
        for e in (e0, e1):
            if e.value:
                hip.hipEventDestroy(e)
        for q in (a, b):
            if q.value:
                hip.hipFree(q)
    return 2 * (mib << 20) * reps / (ms.value * 1e-3) / 1e9


def roofline(cells, value, ms_x, ms_y, unfused, traffic_key, copy_gbs=None):
    """The roofline object of the JSON line for a step timed at `value`
    Gcell-updates/s whose two kernels took ms_x and ms_y per launch (the
    x- and y-sweep, or the fused step and its flag-check kernel).

    Algorithmic bytes per cell-update: 28 B of state per pass over the grid
    (read s, cp, u, v; write cp, u, v), also for the pass that fuses both
    sweeps; the max-wave update's 16 B (read and write hmax, velmax) counted
    separately, in `with_maxwave` and `step_frac_with_maxwave`."""
    def gbs(nbytes, ms):
        return nbytes * cells / (ms * 1e-3) / 1e9

    if unfused:
        dom = "sweep_y" if ms_y >= ms_x else "sweep_x"
        ms, oth, oms = (ms_y, "sweep_x", ms_x) if dom == "sweep_y" else (ms_x, "sweep_y", ms_y)
        mw = MW_BYTES if dom == "sweep_y" else 0        # the y-sweep carries the max-wave
        other = {oth: {"ms_per_launch": round(oms, 4), "bytes_per_cell": SWEEP_BYTES,
                       "achieved": round(gbs(SWEEP_BYTES, oms), 1)}}
        step_bytes = 2 * SWEEP_BYTES
    else:
        dom, ms, mw = "step_fused", ms_x, MW_BYTES
        other = {"step_fixup": {"ms_per_launch": round(ms_y, 4)}}
        step_bytes = SWEEP_BYTES
    ach = gbs(SWEEP_BYTES, ms)
    ach_mw = gbs(SWEEP_BYTES + mw, ms)
    return {
        "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(dom, traffic_key),
        "kernel": dom, "bytes_per_cell": SWEEP_BYTES, "ms_per_launch": round(ms, 4),
        "algorithmic_bytes": SWEEP_BYTES * cells,
        # traffic (measured HBM bytes per launch) includes the max-wave update
        "with_maxwave": {"bytes_per_cell": SWEEP_BYTES + mw, "achieved": round(ach_mw, 1),
                         "frac": round(ach_mw / HBM_PEAK_GBS, 4),
                         "algorithmic_bytes": (SWEEP_BYTES + mw) * cells},
        "other_kernel": other,
        "step_bytes_per_cell": step_bytes,
        "step_frac": round(step_bytes * value / HBM_PEAK_GBS, 4),
        "step_frac_with_maxwave": round((step_bytes + MW_BYTES) * value / HBM_PEAK_GBS, 4),
        "copy_gbs": round(copy_gbs, 1) if copy_gbs else None,
        "frac_of_copy": round(ach / copy_gbs, 4) if copy_gbs else None,
    }


def bench_line(value, steps, warmup, wall, workload, nx, ny, dt, finite, roof, ensemble, cpu):
    """The one JSON object bench.py prints."""
    return {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": steps, "warmup": warmup,
        "ms_per_step": round(wall / steps * 1e3, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload, "nx": nx, "ny": ny, "dt_s": round(dt, 6),
                   "state_finite": finite},
        "roofline": roof,
        "ensemble": ensemble,
        "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=None,
                    help="seed of the grid (open ocean: 0, coastal: 3) and the first ensemble "
                         "scenario (0)")
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true",
                    help="skip the FP64 line mode's run of the CPU baseline's sample")
    ap.add_argument("--ensemble-scenarios", type=int, default=256,
                    help="scenarios of the batched-ensemble measurement (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=4000)
    ap.add_argument("--ensemble-workers", type=int, default=8,
                    help="processes generating the ensemble's inputs (0: in this process, "
                         "as for a profiled run)")
    ap.add_argument("--coastal", action="store_true",
                    help="instead of the open ocean: the spherical 8192 x 6144 coastal grid "
                         "of the nested workload (30 arc-seconds, shelf, coastline, run-up), "
                         "whose tiles take the wet/dry paths")
    ap.add_argument("--nested-hours", type=float, default=6.0)
    ap.add_argument("--nested", action="store_true",
                    help="the full-size nested workload instead: the 8192 x 6144 coastal "
                         "parent with a 2048^2 child at 3x coupled on the device, for "
                         "--nested-hours simulated hours")
    ap.add_argument("--unfused", action="store_true",
                    help="run the step as two sweep kernels instead of one fused kernel")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (x5) to calibrate "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE counters, then exit")
    args = ap.parse_args()

    if args.copy_calib:
        import torch
        n = args.copy_calib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 5
        print(json.dumps({"copy_calib_mib": args.copy_calib, "launches": 6,
                          "bytes_read_per_launch": 4 * n, "bytes_written_per_launch": 4 * n,
                          "ms_per_launch": round(ms, 4),
                          "copy_GBps": round(8 * n / (ms * 1e-3) / 1e9, 1)}))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU; the path shards by scenario or by "
              "y-slab across GPUs, but no multi-GPU run is made here",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import cliffs_gpu
    import synth

    if args.nested:
        value, wall, run, ms_p, cells_p, cells_c, dt, finite = nested_bench(args.nested_hours)
        roof = roofline(cells_p, cells_p / (ms_p[0] * 1e-3) / 1e9, ms_p[0], ms_p[1], False,
                        "coastal_8192x6144", copy_bandwidth())
        roof["kernel"] = "step_fused (parent alone)"
        out = bench_line(value, run["parent_steps"], 0, wall,
                         f"nested: 2D spherical coastal parent 8192x6144 (30 arc-seconds, 100 km "
                         f"shelf at 200 m, land to +300 m, run-up, 200x100 km dipole +3/-1.5 m, "
                         f"seed 3) and a 2048x2048 child at 3x on its coast, coupled on the device, "
                         f"{args.nested_hours} simulated hours", 8192, 6144, dt, finite, roof,
                         None, None)
        out["config"]["child"] = {"nx": 2048, "ny": 2048}
        out["nested"] = run
        print(json.dumps(out), flush=True)
        return

    n = args.n
    if args.coastal:
        seed = 3 if args.seed is None else args.seed
        log(f"[bench] generating the 8192x6144 spherical coastal parent, seed {seed}")
        case = synth.coastal_nest(seed=seed)[0]
        workload = (f"2D spherical coastal grid 8192x6144 at 30 arc-seconds from 25 S (the nested "
                    f"workload's parent alone): 100 km shelf at 200 m, land to +300 m, SSL "
                    f"conditioning, 200x100 km dipole +3/-1.5 m, run-up with minimal flow depth "
                    f"{case.ground:g} m and Manning n^2 {case.crough:g}, seed {seed}")
    else:
        args.seed = 0 if args.seed is None else args.seed
        log(f"[bench] generating {n}x{n} open-ocean case, seed {args.seed}")
        case = synth.open_ocean(n=n, seed=args.seed, steps=args.steps)
        workload = (f"2D Cartesian open ocean {n}x{n}, dx=dy=500 m, depth 4000 m minus fractal "
                    f"seamounts <=2500 m, Gaussian hump 1 m sigma 20 km, open boundaries, "
                    f"non-dispersive, seed {args.seed}")
    ny, nx = case.dep.shape
    s1, s2, zeta = case.metrics()
    g = cliffs_gpu.Grid(case.dep, s1, s2, zeta, dt=case.dt, ground=case.ground,
                        crough=case.crough, device=0)
    g.set_fused(not args.unfused)
    g.upload(case.cel, case.xvel, case.yvel)
    log(f"[bench] warmup {args.warmup} steps")
    g.run(args.warmup)
    g.sync()

    log(f"[bench] timed {args.steps} steps (HIP events around the launches of every 8th step)")
    g.sync()
    t0 = time.perf_counter()
    ms_dev, ms_x, ms_y = g.run_timed(args.steps)
    g.sync()
    wall = time.perf_counter() - t0
    cells = nx * ny
    value = cells * args.steps / wall / 1e9
    log(f"[bench] {wall:.3f} s wall, {ms_dev:.1f} ms device, {value:.2f} Gcell-updates/s")
    c, u, v = g.download()
    finite = bool(np.isfinite(c).all() and np.isfinite(u).all() and np.isfinite(v).all())
    g.close()

    traffic_key = "coastal_8192x6144" if args.coastal else f"open_ocean_{n}"
    copy_gbs = copy_bandwidth()
    log(f"[bench] device copy: {copy_gbs:.0f} GB/s")
    roof = roofline(nx * ny, value, ms_x, ms_y, args.unfused, traffic_key, copy_gbs)
    ensemble = None
    if args.ensemble_scenarios > 0:
        ensemble = ensemble_bench(args.ensemble_scenarios, args.ensemble_steps,
                                  0 if args.seed is None else args.seed, args.ensemble_workers)
        log(f"[bench] ensemble: {ensemble['scenarios_per_hr']} scenarios/hr, "
            f"{ensemble['gcell_updates_per_s']} Gcell-updates/s")
    cpu, want = None, None
    if not args.no_cpu:
        log(f"[bench] CPU baseline: {args.cpu_steps} steps of the FP64 oracle")
        cpu, want = cpu_baseline(case, args.cpu_steps)
        log(f"[bench] CPU {cpu['value']:.4f} Gcell-updates/s on {cpu['cores']} cores")
    fp64 = None
    if not args.no_fp64:
        fp64 = fp64_bench(case, args.cpu_steps, want)
        log(f"[bench] FP64 line mode: {fp64['value']} Gcell-updates/s, bit-identical to the CPU "
            f"baseline: {fp64['bit_identical_to_cpu_baseline']}")
    out = bench_line(value, args.steps, args.warmup, wall, workload, nx, ny, case.dt, finite, roof,
                     ensemble, cpu)
    out["fp64_line_mode"] = fp64
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
