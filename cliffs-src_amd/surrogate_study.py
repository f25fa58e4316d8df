"""Train and evaluate the learned nest surrogate on solver-generated scenarios.

    python3 cliffs-src_amd/surrogate_study.py [--train 896] [--calib 128] [--test 128]

Scenarios are sources at one nested site (surrogate.Site).  For each, the
parents of all scenarios are stepped as an ensemble with gauge recording
(the network's input) and the full nested solve (parent + coupled child) is
run for the truth.  Several networks (--nets) with different seeds are trained on the
MFMA kernels, with early stopping on the calibration split; the gate's
tolerance is calibrated on the same split for each run-up error budget;
everything is reported on the held-out test split.  A scenario is answered
by the surrogate when the networks' run-ups agree (surrogate.gate), otherwise by
the full nested solve.

Prints one JSON line: map error against the full nested solve (and against
the training mean map, the method without the network), run-up error, and
per run-up error budget the gate's acceptance and scenarios/hr with the
surrogate against the nested solve alone, with the time of each leg.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import surrogate as S  # noqa: E402

RUNUP_MIN = S.RUNUP_MIN   # relative run-up errors are quoted for run-ups >= 1 m
BUDGETS = (0.02, 0.05, 0.1, 0.2, 0.5)   # run-up error budgets (95th percentile)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def solve_split(name, seeds, batch, solvers):
    """Inputs from the ensemble parents, truth from nested solves run
    `solvers` at a time from host threads (each on its own grids and
    streams; the library's calls release the interpreter lock).  The nested
    solve's cost per scenario is the wall time of a batch over its size."""
    from concurrent.futures import ThreadPoolExecutor
    xs, maps, runups, t_par, t_nest = [], [], [], 0.0, []
    seeds = list(seeds)
    for b0 in range(0, len(seeds), batch):
        cases = [S.scenario(int(s)) for s in seeds[b0:b0 + batch]]
        t0 = time.perf_counter()
        xs.append(S.parent_gauges(cases) * S.X_SCALE)
        t_par += time.perf_counter() - t0
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=solvers) as ex:
            out = list(ex.map(S.nested_solve, cases))
        t_nest += [(time.perf_counter() - t0) / len(cases)] * len(cases)
        for m, r in out:
            maps.append(m)
            runups.append(r)
        log(f"[study] {name}: {min(b0 + batch, len(seeds))}/{len(seeds)} scenarios, "
            f"nested solve {np.mean(t_nest):.3f} s, parents {t_par / len(t_nest) * 1e3:.2f} ms per scenario")
    maps = np.array(maps, dtype=np.float32)
    return dict(x=np.concatenate(xs).astype(np.float32), y=S.to_quadtree(maps * S.Y_SCALE), maps=maps,
                runup=np.array(runups), t_parent=t_par / len(seeds), t_nested=np.array(t_nest))


def rmse(pred, maps):
    return np.sqrt(((pred - maps) ** 2).mean(axis=(1, 2)))


def runup_err(est, true):
    """Relative error for run-ups >= RUNUP_MIN, absolute / RUNUP_MIN below."""
    return np.abs(est - true) / np.maximum(true, RUNUP_MIN)


def train_early_stop(seed, tr, ca, steps, chunk, batch, lr):
    net = S.Surrogate(batch_max=max(batch, 64), seed=seed)
    best, best_p, hist = np.inf, None, []
    for k in range(0, steps, chunk):
        net.train(tr["x"], tr["y"], batch=batch, steps=chunk, lr=lr, seed=100 * seed + k)
        e = float(rmse(S.predict_maps(net, ca["x"]), ca["maps"]).mean())
        hist.append(round(e, 4))
        if e < best:
            best, best_p = e, net.params()
    net.set_params(best_p)
    log(f"[study] network {seed}: calibration map RMSE by chunk {hist}")
    return net, hist


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=896)
    ap.add_argument("--calib", type=int, default=128)
    ap.add_argument("--test", type=int, default=128)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--chunk", type=int, default=250)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--ens", type=int, default=64, help="parents per ensemble")
    ap.add_argument("--nets", type=int, default=4, help="networks trained from different seeds")
    ap.add_argument("--solvers", type=int, default=4, help="nested solves run at a time")
    args = ap.parse_args()

    t0 = time.perf_counter()
    S.nested_solve(S.scenario(0))        # tunes the site's tile heights once
    tr = solve_split("train", range(args.train), args.ens, args.solvers)
    ca = solve_split("calib", range(100_000, 100_000 + args.calib), args.ens, args.solvers)
    te = solve_split("test", range(200_000, 200_000 + args.test), args.ens, args.solvers)
    t_gen = time.perf_counter() - t0

    nets, hists = [], []
    t0 = time.perf_counter()
    for k in range(1, args.nets + 1):
        net, h = train_early_stop(k, tr, ca, args.steps, args.chunk, args.batch, args.lr)
        nets.append(net)
        hists.append(h)
    t_train = time.perf_counter() - t0

    def infer(split):
        t0 = time.perf_counter()
        preds = [S.predict_maps(n, split["x"]) for n in nets]
        t = (time.perf_counter() - t0) / len(split["x"])
        return preds, t

    # gate calibration per run-up error budget: the loosest tolerance whose
    # accepted calibration scenarios keep the 95th percentile of the run-up
    # error within the budget; then applied to the test split
    preds_c, _ = infer(ca)
    err_c = runup_err(S.runup_from_map(np.mean(preds_c, axis=0)), ca["runup"])
    dis_c = S.gate_score(*preds_c)
    preds, t_net = infer(te)
    pred = np.mean(preds, axis=0)
    dis_t = S.gate_score(*preds)
    ru_pred = S.runup_from_map(pred)
    err = runup_err(ru_pred, te["runup"])
    t_nested = float(te["t_nested"].mean())

    def gate_curve(e_cal, e_test, budgets):
        """Per budget: the loosest gate tolerance whose accepted calibration
        scenarios keep the 95th percentile of the error e_cal within the
        budget, applied to the test split."""
        curve = []
        for budget in budgets:
            tol = -1.0
            for t in np.sort(dis_c):
                ok = dis_c <= t
                if np.quantile(e_cal[ok], 0.95) <= budget:
                    tol = float(t)
            ok = dis_t <= tol
            per_with = te["t_parent"] + t_net + (1.0 - ok.mean()) * t_nested
            curve.append({"budget": budget, "gate_tol": round(tol, 4), "accept": round(float(ok.mean()), 3),
                          "err_p95_accepted": round(float(np.quantile(e_test[ok], 0.95)), 4) if ok.any() else None,
                          "speedup": round(t_nested / per_with, 2),
                          "scenarios_per_hr": round(3600.0 / per_with, 1)})
        return curve

    curve = gate_curve(err_c, err, BUDGETS)
    # the solver's run-up is the ground height of its highest flooded cell, so
    # it moves in steps of one land cell's rise; budgets in those steps
    cell = S.site().land_step
    cells_c = np.abs(S.runup_from_map(np.mean(preds_c, axis=0)) - ca["runup"]) / cell
    cells_t = np.abs(ru_pred - te["runup"]) / cell
    curve_cells = gate_curve(cells_c, cells_t, (0.5, 1.0, 2.0))
    ru_map = S.runup_from_map(te["maps"])            # the run-up reading alone, on the true map
    err_map = runup_err(ru_map, te["runup"])
    mean_map = tr["maps"].mean(axis=0)
    r_net = rmse(pred, te["maps"])
    r_mean = rmse(np.broadcast_to(mean_map, te["maps"].shape), te["maps"])
    err_mean = runup_err(S.runup_from_map(mean_map[None]).repeat(len(pred)), te["runup"])
    area_t, area_p = S.flooded_area(te["maps"]), S.flooded_area(pred)

    def q(a, p):
        return round(float(np.quantile(a, p)), 4) if len(a) else None

    out = {
        "site": {"parent": [S.PNX, S.PNY, S.PDX], "refine": S.REFINE, "child_cells": int(S.site().depc.size),
                 "parent_steps": S.EVERY * S.T, "gauges": S.G, "samples": S.T},
        "networks": args.nets, "solvers": args.solvers, "train_scenarios": args.train, "calib_scenarios": args.calib, "test_scenarios": args.test,
        "train_steps_max": args.steps, "calib_rmse_by_chunk": hists,
        "nested_solve_s": round(t_nested, 4), "ensemble_parent_s": round(te["t_parent"], 5),
        "surrogate_infer_s": round(t_net, 6),
        "scenarios_per_hr_nested": round(3600.0 / t_nested, 1),
        "gate_curve": curve,
        "land_cell_rise_m": round(float(cell), 4), "gate_curve_land_cells": curve_cells,
        "runup_within_one_cell_all": round(float((cells_t <= 1.0).mean()), 3),
        "runup_median_m": round(float(np.median(te["runup"])), 3),
        "map_rmse_m": round(float(r_net.mean()), 4),
        "map_rmse_mean_map_m": round(float(r_mean.mean()), 4),
        "runup_err_median_all": q(err, 0.5), "runup_err_p95_all": q(err, 0.95),
        "runup_within_2pct_all": round(float((err <= 0.02).mean()), 3),
        "runup_err_median_mean_map": q(err_mean, 0.5),
        "runup_err_median_true_map": q(err_map, 0.5), "runup_err_p95_true_map": q(err_map, 0.95),
        "flooded_area_rel_err_median": q(np.abs(area_p - area_t)[area_t > 0] / area_t[area_t > 0], 0.5),
        "data_generation_s": round(t_gen, 1), "training_s": round(t_train, 1),
    }
    print(json.dumps(out), flush=True)
    for n in nets:
        n.close()


if __name__ == "__main__":
    main()
