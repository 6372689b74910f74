#!/usr/bin/env python3
"""Benchmark: pairs/s of batched collocate + integrate on one MI355X.

Workload: 20,000 primitive pairs, la, lb uniform in 0..3, exponents
log-uniform in [0.1, 50] bohr^-2, centres uniform in a 15 bohr cubic cell,
128^3 periodic grid, coefficients uniform[-1, 1], cube radii from eps = 1e-10
(synthetic, seeded). One step = zero the grid, collocate every pair onto it,
zero the blocks, integrate a fixed uniform[-1, 1] potential back onto every
pair's block; all inputs resident in HBM. Prints one JSON line.

Also measured in the same run: the plain one-launch-sequence-per-pair GPU
form on the same task list, the CPU oracle (OpenMP) on a bounded sample of
the pairs (cpu_baseline), and the error of the GPU result against the
oracle on that sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pairs/s collocate+integrate at ≤1e-10 rel err: batched vs plain GPU, 1 MI355X"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--splits", type=int, default=int(os.environ.get("GPW_SPLITS", "4")))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain GPU leg")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep --splits for every kernel class instead of measured choices")
    ap.add_argument("--cpu-pairs", type=int, default=3000)
    ap.add_argument("--plain-pairs", type=int, default=2000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of 1 GiB (x3) to calibrate HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import libgpw_amd as G
    from libgpw_amd import workloads as Wk

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        # a plain copy of known size: reads and writes 2^30 bytes per launch
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 1 << 30, "launches": 3}), flush=True)
        return
    cfg = Wk.config(2)
    npairs = len(cfg["pairs"]["la_max"])
    V_h = Wk.potential(cfg)
    grid_desc = G.grid_desc(cfg["npts"], cfg["dh"])
    t0 = time.perf_counter()
    tl = G.TaskList(grid_desc, cfg["pairs"], splits=args.splits)
    plan_s = time.perf_counter() - t0
    tune = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        before, after = tl.autotune(max_splits=8, reps=3)
        tune = {"seconds": round(time.perf_counter() - t0, 3), "kernel_ms_before": round(before, 4),
                "kernel_ms_after": round(after, 4)}
    st = tl.stats()

    n = cfg["npts"]
    pab = torch.from_numpy(cfg["pab"]).to(dev)
    V = torch.from_numpy(V_h).to(dev)
    grid = torch.zeros((n[2], n[1], n[0]), dtype=torch.float64, device=dev)
    hab = torch.zeros(cfg["nblock"], dtype=torch.float64, device=dev)

    def step():
        grid.zero_()
        hab.zero_()
        tl.collocate(pab, grid)
        tl.integrate(V, hab)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # per-phase HIP events recorded by the library on the launch stream
    tl.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ph = tl.timings()
    tl.set_timing(False)
    phase_ms = {k: v[0] / max(v[1], 1) for k, v in ph.items()}
    coll_ms = phase_ms["collocate"]
    integ_ms = phase_ms["integrate"]
    ms_per_step = wall * 1e3 / args.steps
    value = npairs * args.steps / wall

    # algorithmic bytes: collocate reads+writes every sphere point (16 B),
    # integrate reads it (8 B)
    pts = st["cube_points"]
    coll_bytes, integ_bytes = 16.0 * pts, 8.0 * pts
    if coll_ms >= integ_ms:
        dom = ("k_collocate_tiles", coll_bytes, coll_ms)
    else:
        dom = ("k_integrate_tiles", integ_bytes, integ_ms)
    achieved = dom[1] / (dom[2] * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_128cubed_20000pairs.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get(dom[0])
        except (OSError, ValueError):
            traffic = None
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom[0],
            "kernel_ms": round(dom[2], 4), "algorithmic_bytes": dom[1],
            "step_frac": round((coll_bytes + integ_bytes) / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}

    out = {"metric": METRIC, "value": round(value, 1), "unit": "pairs/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": cfg["name"], "pairs": npairs, "grid": list(n), "cell_bohr": 15.0,
                      "l_range": [0, 3], "exponents": [0.1, 50.0], "eps": Wk.EPS,
                      "splits": args.splits, "autotune": tune,
                      "collocate_splits": st["csplits"], "integrate_splits": st["isplits"][0],
                      "tile": st["tile"], "sub_boxes": st["nsub"],
                      "work_items": st["nentries"], "sphere_points": pts, "plan_s": round(plan_s, 3),
                      "phase_ms": {k: round(v, 4) for k, v in phase_ms.items()}},
           "roofline": roof}

    # ---- plain GPU form (same task-list code family, one launch sequence per pair)
    if not args.no_plain:
        sub = Wk.subset(cfg, np.arange(min(args.plain_pairs, npairs)))
        tlp = G.TaskList(grid_desc, sub["pairs"])
        pab_s = torch.from_numpy(sub["pab"]).to(dev)
        hab_s = torch.zeros(sub["nblock"], dtype=torch.float64, device=dev)
        tlp.collocate(pab_s, grid, plain=True)
        tlp.integrate(V, hab_s, plain=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        grid.zero_(); hab_s.zero_()
        tlp.collocate(pab_s, grid, plain=True)
        tlp.integrate(V, hab_s, plain=True)
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0
        plain_v = len(sub["pairs"]["la_max"]) / tp
        out["plain_gpu"] = {"value": round(plain_v, 1), "unit": "pairs/s",
                            "sample": f"first {len(sub['pairs']['la_max'])} pairs of the workload, one step",
                            "speedup_batched": round(value / plain_v, 2)}
        tlp.close()

    # ---- CPU oracle baseline on a bounded sample + error of the GPU result
    if not args.no_cpu:
        from oracle import oracle as O
        O.build()
        ncpu = min(16, os.cpu_count() or 1)
        rng = np.random.default_rng(7)
        idx = np.sort(rng.choice(npairs, size=min(args.cpu_pairs, npairs), replace=False))
        sub = Wk.subset(cfg, idx)
        og = O.make_grid(sub["npts"], sub["dh"])
        op = O.make_pairs(sub["pairs"])
        reps = 0
        t0 = time.perf_counter()
        while True:
            ref_grid = O.batch_collocate(og, op, sub["pab"], nthreads=ncpu)
            ref_hab, _ = O.batch_integrate(og, op, V_h, sub["nblock"], nthreads=ncpu)
            reps += 1
            if time.perf_counter() - t0 > 10.0 or reps >= 5:
                break
        tc = (time.perf_counter() - t0) / reps
        out["cpu_baseline"] = {"value": round(len(idx) / tc, 1), "unit": "pairs/s", "cores": ncpu,
                               "kind": "port",
                               "sample": f"{len(idx)} random pairs of the workload (seed 7), "
                                         f"collocate+integrate x{reps}, OpenMP C oracle"}
        tls = G.TaskList(grid_desc, sub["pairs"], splits=args.splits)
        g_s = torch.zeros_like(grid)
        h_s = torch.zeros(sub["nblock"], dtype=torch.float64, device=dev)
        tls.collocate(torch.from_numpy(sub["pab"]).to(dev), g_s)
        tls.integrate(V, h_s)
        torch.cuda.synchronize()
        gg, hh = g_s.cpu().numpy(), h_s.cpu().numpy()
        out["rel_err"] = {"grid": float(np.abs(gg - ref_grid).max() / np.abs(ref_grid).max()),
                          "hab": float(np.abs(hh - ref_hab).max() / np.abs(ref_hab).max()),
                          "sample": "the cpu_baseline pairs"}
        tls.close()

    tl.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
