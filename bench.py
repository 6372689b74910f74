#!/usr/bin/env python3
"""Benchmark: pairs/s of batched collocate + integrate on one MI355X.

Default workload (--config 2): 20,000 primitive pairs, la, lb uniform in
0..3, exponents log-uniform in [0.1, 50] bohr^-2, centres uniform in a 15 bohr
cubic cell, 128^3 periodic grid, coefficients uniform[-1, 1], cube radii from
eps = 1e-10 (synthetic, seeded). One step = zero the grid, collocate every
pair onto it, zero the blocks, integrate a fixed uniform[-1, 1] potential back
onto every pair's block; all inputs resident in HBM. The step is timed as
eager launches and as replays of one HIP graph of it; the faster form is
the value (both are reported). Prints one JSON line.

Also measured in the same run: the plain one-launch-sequence-per-pair GPU
form on the same task list, the CPU oracle (OpenMP) on a bounded sample of
the pairs (cpu_baseline), and the error of the GPU result against the
oracle on that sample.

--config 1, 3, 4 run the other single-grid workloads the same way (3 and 4
with forces in the integrate step); --config 5 runs the four-level triclinic
multigrid workload (500,000 pairs), collocate + integrate + forces over all
levels, timed both as one HIP graph (levels in sequence) and as eager
launches with one stream per level; the faster form is the value.
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


def hbm_copy_peak(torch, dev, nbytes=1 << 30, reps=5):
    """Attainable HBM bandwidth measured in this run: best of `reps` device
    copies of nbytes (read + write = 2 nbytes of traffic), timed with events
    on the current stream. GB/s."""
    n = nbytes // 8
    src = torch.ones(n, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = 1e30
    for _ in range(reps):
        e0.record()
        dst.copy_(src)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1))
    del src, dst
    return 2.0 * nbytes / (best * 1e-3) / 1e9


def roofline(st, coll_ms, integ_ms, ms_per_step, workload, forces, peak_measured):
    """Two rooflines of the dominant kernel and of the whole step.

    Bytes (the contract's "hbm" line): a per-pair algorithmic MODEL --
    collocate reads and writes every sphere point of every pair (16 B),
    integrate reads it (8 B). A grid point covered by several pairs counts
    once per pair, so the model can exceed what crosses HBM (the tile
    kernels keep a tile in registers while all its pairs pass); `traffic`
    is the HBM bytes rocprofv3 counted for the same kernel and workload.
    FLOPs: the algorithmic FP64 flops of the z, y, x contraction order
    (stats flops_*: C3(lp) FMAs per z-plane, C2(lp) per row, lp + 1 per
    point of each sphere) against the FP64 peak."""
    pts = st["cube_points"]
    coll_bytes, integ_bytes = 16.0 * pts, 8.0 * pts
    coll_fl = st["flops_collocate"]
    integ_fl = st["flops_integrate"][1 if forces else 0]
    if coll_ms >= integ_ms:
        dom = ("k_collocate_tiles", coll_bytes, coll_ms, coll_fl)
    else:
        dom = ("k_integrate_tiles", integ_bytes, integ_ms, integ_fl)
    achieved = dom[1] / (dom[2] * 1e-3) / 1e9
    traffic, tsrc = None, None
    tpath = os.path.join(ROOT, "profiles", "traffic_128cubed_20000pairs.json")
    if os.path.exists(tpath):
        try:
            prof = json.load(open(tpath))
            if prof.get("workload") == workload:
                traffic = prof.get(dom[0])
                tsrc = prof.get("source")
        except (OSError, ValueError):
            traffic = None

    def tf(fl, ms):
        return fl / (ms * 1e-3) / 1e12 if ms > 0 else 0.0

    flops = {
        "collocate": {"tflops": round(tf(coll_fl, coll_ms), 3),
                      "frac": round(tf(coll_fl, coll_ms) / FP64_PEAK_TFLOPS, 4),
                      "algorithmic_flops": coll_fl, "kernel_ms": round(coll_ms, 4)},
        "integrate": {"tflops": round(tf(integ_fl, integ_ms), 3),
                      "frac": round(tf(integ_fl, integ_ms) / FP64_PEAK_TFLOPS, 4),
                      "algorithmic_flops": integ_fl, "kernel_ms": round(integ_ms, 4)},
        "step": {"tflops": round(tf(coll_fl + integ_fl, ms_per_step), 3),
                 "frac": round(tf(coll_fl + integ_fl, ms_per_step) / FP64_PEAK_TFLOPS, 4)},
        "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s"}
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": tsrc,
            "model": "per-pair algorithmic bytes (16 B collocate / 8 B integrate per sphere point); "
                     "a grid point counts once per pair covering it",
            "peak_measured": round(peak_measured, 1),
            "frac_of_measured": round(achieved / peak_measured, 4),
            "kernel": dom[0] if dom[0] != "k_integrate_tiles" else "k_integrate_tiles + k_moments_small",
            "kernel_ms": round(dom[2], 4), "algorithmic_bytes": dom[1],
            "collocate_byte_model": {"achieved": round(coll_bytes / (coll_ms * 1e-3) / 1e9, 1),
                                     "frac": round(coll_bytes / (coll_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
            "integrate_byte_model": {"achieved": round(integ_bytes / (integ_ms * 1e-3) / 1e9, 1),
                                     "frac": round(integ_bytes / (integ_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
            "step_frac": round((coll_bytes + integ_bytes) / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "flop_frac": round(tf(dom[3], dom[2]) / FP64_PEAK_TFLOPS, 4),
            "flops": flops}


def host_cores():
    """CPU cores this process may use: `nproc` (honours the affinity mask and
    OMP_NUM_THREADS), else the affinity mask."""
    import subprocess
    try:
        return max(1, int(subprocess.run(["nproc"], capture_output=True, text=True,
                                         check=True).stdout.strip()))
    except (OSError, ValueError, subprocess.CalledProcessError):
        return max(1, len(os.sched_getaffinity(0)))


def cpu_leg(O, legs, nthreads, budget_s=10.0):
    """Time the oracle on [(grid, pairs, pab, potential, nblock)] (collocate +
    integrate), repeated until ~budget_s; returns (seconds per pass, outputs, reps)."""
    reps, outs = 0, None
    t0 = time.perf_counter()
    while True:
        outs = []
        for og, op, pab, pot, nblock in legs:
            g = O.batch_collocate(og, op, pab, nthreads=nthreads)
            h, _ = O.batch_integrate(og, op, pot, nblock, nthreads=nthreads)
            outs.append((g, h))
        reps += 1
        if time.perf_counter() - t0 > budget_s or reps >= 5:
            break
    return (time.perf_counter() - t0) / reps, outs, reps


def dropin_per_call(G, torch, cfg, V_h, n):
    """Per-call time of the single-task entry points (the reference's
    grid_collocate_pgf_product_rspace argument list) on the first n pairs:
    host arrays (only the cube's points cross PCIe) and device arrays, each
    a collocate then an integrate with forces per pair."""
    import math
    p = cfg["pairs"]
    g = G.grid_desc(cfg["npts"], cfg["dh"])
    npts = list(cfg["npts"])
    tasks = []
    for i in range(n):
        geo = G.pair_geometry(g, p, i)
        la, lb = int(p["la_max"][i]), int(p["lb_max"][i])
        maxco = G.ncoset(max(la, lb))
        t = {"compute_tau": False, "use_ortho": True, "la_max": la, "la_min": int(p["la_min"][i]),
             "lb_max": lb, "lb_min": int(p["lb_min"][i]), "zeta": float(p["zeta"][i]),
             "zetb": float(p["zetb"][i]), "rscale": float(p["rscale"][i]), "dh": cfg["dh"],
             "dh_inv": np.linalg.inv(cfg["dh"]), "ra": p["ra"][i], "rab": p["rab"][i], "npts": npts,
             "ngrid": npts, "lb_grid": [0, 0, 0], "periodic": [1, 1, 1], "lmax": max(la, lb),
             "radius": float(p["radius"][i]), "lb_cube": geo["lb_cube"], "ub_cube": geo["ub_cube"],
             "sphere_bounds": geo["sphere_bounds"], "nspheres": len(geo["sphere_bounds"]),
             "maxco": maxco, "o1": 0, "o2": 0}
        tasks.append(t)
    rng = np.random.default_rng(3)
    pabs = [rng.uniform(-1, 1, (t["maxco"], t["maxco"])) for t in tasks]
    n3 = (npts[2], npts[1], npts[0])
    grid_h = np.zeros(n3)
    out = {"pairs": n, "sample": f"first {n} pairs of the workload, collocate + integrate with forces"}
    for mode in ("host", "device"):
        if mode == "device":
            gd = torch.zeros(n3, dtype=torch.float64, device="cuda")
            Vd = torch.from_numpy(V_h).cuda()
            pds = [torch.from_numpy(q).cuda() for q in pabs]
            hds = [torch.zeros_like(q) for q in pds]
            fd = torch.zeros(6, dtype=torch.float64, device="cuda")
        for rep in range(2):  # the first pass grows the library's workspace
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for t, q, k in zip(tasks, pabs, range(n)):
                if mode == "host":
                    G.collocate_pgf_product_rspace(t, q, grid_h)
                    G.integrate_pgf_product_rspace(t, V_h, pab=q, forces=True)
                else:
                    G.collocate_pgf_product_rspace_dev(t, pds[k], gd)
                    G.integrate_pgf_product_rspace_dev(t, Vd, hds[k], pab=pds[k], forces=fd)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
        out[f"{mode}_arrays_us_per_pair"] = round(dt / n * 1e6, 1)
    return out


def rel(a, b):
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / s) if s > 0 else float(np.abs(a).max())


def library_load(G, Wk, cfg, grid_desc):
    """Seconds of the first task list of the process (one pair of the workload):
    the HIP runtime loads the library's code object and sets up its allocator
    at the first launch and allocation. plan_s is timed after this, so it
    measures the planner."""
    one = {k: np.asarray(v)[:1] for k, v in cfg["pairs"].items()}
    one["block_offset"] = np.zeros(1, dtype=np.int64)
    t0 = time.perf_counter()
    G.TaskList(grid_desc, one).close()
    return time.perf_counter() - t0


def run_single(args, torch, G, Wk, dev, cfg):
    npairs = len(cfg["pairs"]["la_max"])
    with_forces = cfg["index"] in (3, 4)
    V_h = Wk.potential(cfg)
    grid_desc = G.grid_desc(cfg["npts"], cfg["dh"])
    load_s = library_load(G, Wk, cfg, grid_desc)
    t0 = time.perf_counter()
    tl = G.TaskList(grid_desc, cfg["pairs"], splits=args.splits)
    plan_s = time.perf_counter() - t0
    tune = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        before, after = tl.autotune(max_splits=0, reps=3)
        tune = {"seconds": round(time.perf_counter() - t0, 3), "kernel_ms_before": round(before, 4),
                "kernel_ms_after": round(after, 4)}
    st = tl.stats()

    n = cfg["npts"]
    pab = torch.from_numpy(cfg["pab"]).to(dev)
    V = torch.from_numpy(V_h).to(dev)
    grid = torch.zeros((n[2], n[1], n[0]), dtype=torch.float64, device=dev)
    hab = torch.zeros(cfg["nblock"], dtype=torch.float64, device=dev)
    forces = torch.zeros((npairs, 6), dtype=torch.float64, device=dev) if with_forces else None

    def step():
        grid.zero_()
        hab.zero_()
        if forces is not None:
            forces.zero_()
        tl.collocate(pab, grid)
        tl.integrate(V, hab, pab=pab if forces is not None else None, forces=forces)

    def timed(run):
        for _ in range(args.warmup):
            run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            run()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    # two forms of the same step, each timed over the same K steps: eager
    # launches, and replays of one HIP graph of the step (the fills and the
    # task-list calls allocate nothing, so the step captures as it is); the
    # faster one is the line's value, both are reported
    wall_eager = timed(step)
    graph = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(graph):
        step()
    torch.cuda.synchronize()
    wall_graph = timed(graph.replay)
    wall = min(wall_eager, wall_graph)
    form = "eager launches" if wall_eager <= wall_graph else "HIP graph replay"
    # per-phase HIP events recorded by the library on the launch stream, in a
    # separate pass of K eager steps (the events are not in the timed steps)
    tl.set_timing(True)
    torch.cuda.synchronize()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ph = tl.timings()
    tl.set_timing(False)
    phase_ms = {k: v[0] / max(v[1], 1) for k, v in ph.items()}
    ms_per_step = wall * 1e3 / args.steps
    value = npairs * args.steps / wall
    pts = st["cube_points"]
    roof = roofline(st, phase_ms["collocate"], phase_ms["integrate"], ms_per_step, cfg["name"],
                    with_forces, args.hbm_peak)

    out = {"metric": METRIC, "value": round(value, 1), "unit": "pairs/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": cfg["name"], "pairs": npairs, "grid": list(n),
                      "cell_bohr": [round(float(x), 6) for x in np.diag(cfg["dh"]) * np.asarray(n)],
                      "forces": with_forces, "eps": Wk.EPS,
                      "splits": args.splits, "autotune": tune,
                      "collocate_splits": st["csplits"],
                      "integrate_splits": st["isplits"][1 if with_forces else 0],
                      "tile": st["tile"], "sub_boxes": st["nsub"],
                      "work_items": st["nentries"], "sphere_points": pts, "plan_s": round(plan_s, 4),
                      "step_form": form,
                      "ms_per_step_eager": round(wall_eager * 1e3 / args.steps, 4),
                      "ms_per_step_graph": round(wall_graph * 1e3 / args.steps, 4),
                      "library_load_s": round(load_s, 4),
                      "plan_ms": {k: round(v, 2) for k, v in st["plan_ms"].items()},
                      "phase_ms": {k: round(v, 4) for k, v in phase_ms.items()}},
           "roofline": roof,
           # one pass including the host + device planning of the task list
           "end_to_end": {"value": round(npairs / (plan_s + ms_per_step * 1e-3), 1), "unit": "pairs/s",
                          "plan_s": round(plan_s, 4), "ms_per_step": round(ms_per_step, 4)}}

    # ---- plain GPU form (same task-list code family, one launch sequence per pair)
    if not args.no_plain:
        sub = Wk.subset(cfg, np.arange(min(args.plain_pairs, npairs)))
        tlp = G.TaskList(grid_desc, sub["pairs"])
        pab_s = torch.from_numpy(sub["pab"]).to(dev)
        hab_s = torch.zeros(sub["nblock"], dtype=torch.float64, device=dev)
        tlp.collocate(pab_s, grid, plain=True)
        tlp.integrate(V, hab_s, plain=True)
        torch.cuda.synchronize()
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            grid.zero_(); hab_s.zero_()
            tlp.collocate(pab_s, grid, plain=True)
            tlp.integrate(V, hab_s, plain=True)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        tp = float(np.median(times))
        plain_v = len(sub["pairs"]["la_max"]) / tp
        out["plain_gpu"] = {"value": round(plain_v, 1), "unit": "pairs/s",
                            "sample": f"first {len(sub['pairs']['la_max'])} pairs of the workload, "
                                      "median of 3 steps",
                            "speedup_batched": round(value / plain_v, 2)}
        tlp.close()

    # ---- the reference's one-call-per-pair interface on the same pairs
    if not args.no_plain and cfg["index"] != 5:
        out["single_task"] = dropin_per_call(G, torch, cfg, V_h, min(200, npairs))

    # ---- CPU oracle baseline on a bounded sample + error of the GPU result
    if not args.no_cpu:
        from oracle import oracle as O
        O.build()
        ncpu = host_cores()
        rng = np.random.default_rng(7)
        idx = np.sort(rng.choice(npairs, size=min(args.cpu_pairs, npairs), replace=False))
        sub = Wk.subset(cfg, idx)
        og = O.make_grid(sub["npts"], sub["dh"])
        op = O.make_pairs(sub["pairs"])
        leg = [(og, op, sub["pab"], V_h, sub["nblock"])]
        tc, outs, reps = cpu_leg(O, leg, ncpu)
        t1, _, reps1 = cpu_leg(O, leg, 1, budget_s=0.0)
        ref_grid, ref_hab = outs[0]
        out["cpu_baseline"] = {"value": round(len(idx) / tc, 1), "unit": "pairs/s", "cores": ncpu,
                               "kind": "port",
                               "sample": f"{len(idx)} random pairs of the workload (seed 7), "
                                         f"collocate+integrate x{reps}, OpenMP C oracle",
                               "one_core": {"value": round(len(idx) / t1, 1), "cores": 1,
                                            "sample": f"the same pairs, x{reps1}"}}
        tls = G.TaskList(grid_desc, sub["pairs"], splits=args.splits)
        g_s = torch.zeros_like(grid)
        h_s = torch.zeros(sub["nblock"], dtype=torch.float64, device=dev)
        tls.collocate(torch.from_numpy(sub["pab"]).to(dev), g_s)
        tls.integrate(V, h_s)
        torch.cuda.synchronize()
        out["rel_err"] = {"grid": rel(g_s.cpu().numpy(), ref_grid),
                          "hab": rel(h_s.cpu().numpy(), ref_hab), "sample": "the cpu_baseline pairs"}
        tls.close()
    tl.close()
    return out


def run_multigrid(args, torch, G, Wk, dev, cfg):
    from libgpw_amd.multigrid import MultiGrid
    lv0 = cfg["levels"][0]
    load_s = library_load(G, Wk, {"pairs": {k: np.asarray(v)[lv0["index"]] for k, v in cfg["pairs"].items()}},
                          G.grid_desc(lv0["npts"], lv0["dh"]))
    t0 = time.perf_counter()
    mg = MultiGrid(cfg, splits=args.splits)
    plan_s = time.perf_counter() - t0
    tune = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        res = mg.autotune(max_splits=0, reps=3)
        tune = {"seconds": round(time.perf_counter() - t0, 3),
                "kernel_ms_before": round(sum(r[0] for r in res), 4),
                "kernel_ms_after": round(sum(r[1] for r in res), 4)}
    pab = torch.from_numpy(cfg["pab"]).to(dev)
    grids = mg.grids(dev)
    pots_h = [np.random.default_rng(cfg["seed"] + 7 + k).uniform(-1.0, 1.0, tuple(g.shape))
              for k, g in enumerate(grids)]
    pots = [torch.from_numpy(p).to(dev) for p in pots_h]
    hab = torch.zeros(cfg["nblock"], dtype=torch.float64, device=dev)
    forces = mg.forces_buffers(dev)
    # per-level phase times from one timed eager step
    for lv in mg.levels:
        lv["tl"].set_timing(True)
    torch.cuda.synchronize()
    # levels one after the other here, so each level's phase times are its own
    mg.step(pab, grids, pots, hab, forces, concurrent=False)
    torch.cuda.synchronize()
    level_ms, pts, coll, integ, fl_c, fl_i = [], 0, 0.0, 0.0, 0.0, 0.0
    for lv in mg.levels:
        ph = lv["tl"].timings()
        lv["tl"].set_timing(False)
        st = lv["tl"].stats()
        pts += st["cube_points"]
        fl_c += st["flops_collocate"]
        fl_i += st["flops_integrate"][1]
        pm = {k: round(v[0] / max(v[1], 1), 4) for k, v in ph.items()}
        coll += pm["collocate"] + pm["reduce"]
        integ += pm["integrate"]
        level_ms.append({"grid": list(lv["npts"]), "pairs": int(len(lv["index"])),
                         "sphere_points": st["cube_points"], "collocate_splits": st["csplits"],
                         "integrate_splits": st["isplits"][1], "phase_ms": pm,
                         "plan_ms": {k: round(v, 2) for k, v in st["plan_ms"].items()}})
    # two forms of the same step: one HIP graph (levels one after the other on
    # one stream) and eager launches with each level on a stream of its own;
    # the faster one is the line's value, both are reported
    graph = mg.capture(pab, grids, pots, hab, forces)

    def timed(run):
        for _ in range(args.warmup):
            run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            run()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    wall_graph = timed(graph.replay)
    wall_conc = timed(lambda: mg.step(pab, grids, pots, hab, forces, concurrent=True))
    wall = min(wall_graph, wall_conc)
    form = ("eager, one stream per level" if wall_conc <= wall_graph
            else "HIP graph, levels in sequence on one stream")
    ms_per_step = wall * 1e3 / args.steps
    value = mg.npairs * args.steps / wall
    out = {"metric": METRIC, "value": round(value, 1), "unit": "pairs/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": cfg["name"], "pairs": mg.npairs, "forces": True, "eps": Wk.EPS,
                      "step_form": form,
                      "ms_per_step_graph": round(wall_graph * 1e3 / args.steps, 4),
                      "ms_per_step_concurrent": round(wall_conc * 1e3 / args.steps, 4),
                      "autotune": tune, "plan_s": round(plan_s, 3),
                      "library_load_s": round(load_s, 4),
                      "levels": level_ms},
           "roofline": roofline({"cube_points": pts, "flops_collocate": fl_c,
                                 "flops_integrate": [fl_i, fl_i]}, coll, integ, ms_per_step,
                                cfg["name"], True, args.hbm_peak),
           "end_to_end": {"value": round(mg.npairs / (plan_s + ms_per_step * 1e-3), 1),
                          "unit": "pairs/s", "plan_s": round(plan_s, 4),
                          "ms_per_step": round(ms_per_step, 4)}}
    if not args.no_cpu:
        from oracle import oracle as O
        O.build()
        ncpu = host_cores()
        rng = np.random.default_rng(7)
        frac = min(1.0, args.cpu_pairs / mg.npairs)
        legs, tls, sample = [], [], 0
        for k, lv in enumerate(mg.levels):
            m = max(1, int(round(frac * len(lv["index"]))))
            idx = np.sort(rng.choice(lv["index"], size=m, replace=False))
            sub = {kk: np.asarray(v)[idx] for kk, v in cfg["pairs"].items()}
            og = O.make_grid(lv["npts"], cfg["levels"][k]["dh"])
            legs.append((og, O.make_pairs(sub), cfg["pab"], pots_h[k], cfg["nblock"]))
            tls.append(G.TaskList(lv["grid_desc"], sub))
            sample += m
        tc, outs, reps = cpu_leg(O, legs, ncpu)
        t1, _, reps1 = cpu_leg(O, legs, 1, budget_s=0.0)
        out["cpu_baseline"] = {"value": round(sample / tc, 1), "unit": "pairs/s", "cores": ncpu,
                               "kind": "port",
                               "sample": f"{sample} pairs drawn per level in proportion (seed 7), "
                                         f"collocate+integrate x{reps}, OpenMP C oracle",
                               "one_core": {"value": round(sample / t1, 1), "cores": 1,
                                            "sample": f"the same pairs, x{reps1}"}}
        eg = 0.0
        h_s = torch.zeros(cfg["nblock"], dtype=torch.float64, device=dev)
        h_ref = np.zeros(cfg["nblock"])
        for k, (tl, (g_ref, h_r)) in enumerate(zip(tls, outs)):
            g_s = torch.zeros_like(grids[k])
            tl.collocate(pab, g_s)
            tl.integrate(pots[k], h_s)
            torch.cuda.synchronize()
            eg = max(eg, rel(g_s.cpu().numpy(), g_ref))
            h_ref += h_r
            tl.close()
        out["rel_err"] = {"grid": eg, "hab": rel(h_s.cpu().numpy(), h_ref),
                          "sample": "the cpu_baseline pairs"}
    mg.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4, 5],
                    help="workload (2: the 20,000-pair 128^3 configuration of the metric)")
    ap.add_argument("--pairs", type=int, default=None, help="override the pair count")
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
    # attainable HBM bandwidth of this GPU, measured before the workload
    args.hbm_peak = hbm_copy_peak(torch, dev)
    cfg = Wk.config(args.config, npairs=args.pairs)
    if args.config == 5:
        out = run_multigrid(args, torch, G, Wk, dev, cfg)
    else:
        out = run_single(args, torch, G, Wk, dev, cfg)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
