"""GPU tests of the full-size workloads and of argument handling.

* The 100,000-pair mixed workload (l <= 4, exponents down to 0.05, cubes
  wider than the 25 bohr cell, 200^3) and the 500,000-pair triclinic
  multigrid workload run whole on the GPU: the adjoint identity
  <collocate(P), W> = <P, integrate(W)> over every pair, plus oracle parity
  of hab and forces on a random sample of the pairs (integrate is per pair,
  so the sample's blocks of the full run are compared directly) and of the
  grid of the sample collocated on its own.
* Task lists whose pairs have la_min < la_max and lb_min < lb_max (rho and
  tau, with forces) against the oracle.
* The single-task drop-in with use_ortho = 1 on a cell with off-diagonal
  steps takes the orthorhombic path on dh[i][i], as the reference does.
* Short buffers are rejected before any device work.
Tolerance 1e-10 relative (max|x - ref| / max|ref|).
"""
import sys

import numpy as np
import pytest

from helpers import random_task, rel_err
from test_gpu import dev, oracle_batch, run_tasklist, same_support

pytestmark = pytest.mark.gpu
TOL = 1e-10


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _blocks_of(pairs, idx):
    """flat indices of the packed blocks of pairs idx, in order"""
    out = []
    for i in idx:
        sz = (int(pairs["la_max"][i]) + 1) * (int(pairs["la_max"][i]) + 2) * (int(pairs["la_max"][i]) + 3) // 6
        sz *= (int(pairs["lb_max"][i]) + 1) * (int(pairs["lb_max"][i]) + 2) * (int(pairs["lb_max"][i]) + 3) // 6
        bo = int(pairs["block_offset"][i])
        out.append(np.arange(bo, bo + sz))
    return np.concatenate(out)


def test_config4_full_size(gpw, oracle, torch_cuda):
    """All 100,000 pairs (la, lb 0..4, exponents 0.05..100, 200^3, with
    forces, so L up to 11): adjoint identity over every pair; hab and forces
    of 3,000 sampled pairs from the full run against the oracle; the
    sample's grid, collocated on its own, against the oracle."""
    from libgpw_amd import workloads as Wk
    torch = torch_cuda
    cfg = Wk.config(4)
    npairs = len(cfg["pairs"]["la_max"])
    say("config4 full", npairs, "pairs")
    W = Wk.potential(cfg)
    got = run_tasklist(gpw, torch, cfg, W=W, forces=True, splits=4)
    st = got["stats"]
    assert st["npairs"] == npairs and st["lp_max"] == 8
    # cubes wider than the period are split into sub-boxes
    assert st["nsub"] > npairs
    lhs = float((got["grid"] * W).sum())
    rhs = float((cfg["pab"] * got["hab"]).sum())
    assert abs(lhs - rhs) < 1e-11 * np.abs(got["grid"]).sum(), (lhs, rhs)
    rng = np.random.default_rng(44)
    idx = np.sort(rng.choice(npairs, size=3000, replace=False))
    sub = Wk.subset(cfg, idx)
    say("config4 oracle sample", len(idx))
    ref = oracle_batch(oracle, sub, W=W, forces=True)
    blk = _blocks_of(cfg["pairs"], idx)
    assert rel_err(got["hab"][blk], ref["hab"]) < TOL
    assert rel_err(got["forces"][idx], ref["forces"]) < TOL
    part = run_tasklist(gpw, torch, sub, splits=2)
    assert same_support(part["grid"], ref["grid"])
    assert rel_err(part["grid"], ref["grid"]) < TOL


def test_config5_full_size(gpw, oracle, torch_cuda):
    """All 500,000 pairs of the triclinic multigrid workload (four levels):
    per level, the adjoint identity over the level's pairs and oracle parity
    of hab and forces on 500 sampled pairs of the full run; the sample's
    grid, collocated on its own, against the oracle."""
    from libgpw_amd import workloads as Wk
    torch = torch_cuda
    cfg = Wk.config(5)
    say("config5 full", len(cfg["pairs"]["la_max"]), "pairs",
        [len(lv["index"]) for lv in cfg["levels"]])
    rng = np.random.default_rng(55)
    for k in range(len(cfg["levels"])):
        lvc = Wk.level(cfg, k)
        n = len(lvc["pairs"]["la_max"])
        if n == 0:
            continue
        say("config5 level", k, n, "pairs", lvc["npts"])
        W = Wk.potential(lvc, seed_offset=20 + k)
        got = run_tasklist(gpw, torch, lvc, W=W, forces=True, splits=2)
        lhs = float((got["grid"] * W).sum())
        rhs = float((lvc["pab"] * got["hab"]).sum())
        assert abs(lhs - rhs) < 1e-11 * np.abs(got["grid"]).sum(), (k, lhs, rhs)
        idx = np.sort(rng.choice(n, size=min(500, n), replace=False))
        sub = Wk.subset(lvc, idx)
        ref = oracle_batch(oracle, sub, W=W, forces=True)
        blk = _blocks_of(lvc["pairs"], idx)
        assert rel_err(got["hab"][blk], ref["hab"]) < TOL, k
        assert rel_err(got["forces"][idx], ref["forces"]) < TOL, k
        part = run_tasklist(gpw, torch, sub, splits=1)
        assert same_support(part["grid"], ref["grid"]), k
        assert rel_err(part["grid"], ref["grid"]) < TOL, k


def _ranged_pairs(rng, n, cell_m):
    """n pairs with la_min < la_max and lb_min < lb_max (mixed ranges)"""
    la_max = rng.integers(1, 4, n)
    lb_max = rng.integers(1, 4, n)
    la_min = np.array([rng.integers(0, a) for a in la_max])
    lb_min = np.array([rng.integers(0, b) for b in lb_max])
    zeta = np.exp(rng.uniform(np.log(0.3), np.log(20.0), n))
    zetb = np.exp(rng.uniform(np.log(0.3), np.log(20.0), n))
    ra = rng.uniform(0.0, 1.0, (n, 3)) @ cell_m
    rab = rng.normal(0.0, 0.7, (n, 3))
    sizes = np.array([((a + 1) * (a + 2) * (a + 3) // 6) * ((b + 1) * (b + 2) * (b + 3) // 6)
                      for a, b in zip(la_max, lb_max)])
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    pairs = {"la_max": la_max.astype(np.int32), "la_min": la_min.astype(np.int32),
             "lb_max": lb_max.astype(np.int32), "lb_min": lb_min.astype(np.int32),
             "zeta": zeta, "zetb": zetb, "rscale": rng.uniform(0.5, 2.0, n), "ra": ra, "rab": rab,
             "radius": np.sqrt(np.log(1e10) / (zeta + zetb)), "block_offset": off}
    return pairs, int(sizes.sum())


@pytest.mark.parametrize("tau", [False, True])
@pytest.mark.parametrize("shear", [0.0, 0.25])
def test_tasklist_angular_momentum_ranges(gpw, oracle, torch_cuda, tau, shear):
    """la_min < la_max and lb_min < lb_max through the tile kernels and the
    small-sphere integrate kernel (collocate, integrate, forces; rho and
    tau; orthorhombic and sheared cells), batched with splits and plain."""
    torch = torch_cuda
    rng = np.random.default_rng(71 + 2 * tau + int(10 * shear))
    npts = (40, 44, 48)
    cell = np.array([9.0, 9.9, 10.8])
    dh = np.diag(cell / np.asarray(npts))
    dh[1, 0] = shear * dh[0, 0]
    cell_m = dh * np.asarray(npts)[:, None]
    pairs, nblock = _ranged_pairs(rng, 150, cell_m)
    cfg = {"npts": npts, "dh": dh, "pairs": pairs, "pab": rng.uniform(-1.0, 1.0, nblock),
           "nblock": nblock}
    W = rng.uniform(-1.0, 1.0, (npts[2], npts[1], npts[0]))
    say("l ranges", "tau" if tau else "rho", "shear", shear)
    ref = oracle_batch(oracle, cfg, W=W, forces=True, tau=tau)
    assert np.count_nonzero(ref["grid"]) > 1000
    for plain, splits in ((False, 1), (False, 3), (True, 1)):
        got = run_tasklist(gpw, torch, cfg, plain=plain, splits=splits, forces=True, W=W, tau=tau)
        assert same_support(got["grid"], ref["grid"]), (plain, splits)
        assert rel_err(got["grid"], ref["grid"]) < TOL, (plain, splits)
        assert rel_err(got["hab"], ref["hab"]) < TOL, (plain, splits)
        assert rel_err(got["forces"], ref["forces"]) < TOL, (plain, splits)


@pytest.mark.parametrize("tau", [False, True])
def test_dropin_use_ortho_on_sheared_cell(gpw, oracle, torch_cuda, tau):
    """use_ortho = 1 with off-diagonal dh: the orthorhombic path on the
    diagonal steps and the caller's sphere_bounds (grid_base_ref_c.cc:981,
    :546), not the general path, as the oracle does."""
    say("drop-in use_ortho on sheared cell", tau)
    rng = np.random.default_rng(808 + tau)
    t, pab = random_task(rng, oracle, la=(0, 2), lb=(1, 2), tau=tau, ortho=True, shear=0.15)
    assert t["dh"][1][0] != 0.0 and t["use_ortho"]
    ref = oracle.collocate(t, pab)
    # the general path on the same cell writes a different point set
    gen = oracle.collocate(dict(t, use_ortho=False), pab)
    assert not same_support(gen, ref) or rel_err(gen, ref) > 1e-3
    grid = np.zeros_like(ref)
    gpw.collocate_pgf_product_rspace(t, pab, grid)
    assert same_support(grid, ref)
    assert rel_err(grid, ref) < TOL
    W = rng.uniform(-1, 1, ref.shape)
    h_ref, f_ref = oracle.integrate(t, W, pab=pab, forces=True)
    h, f = gpw.integrate_pgf_product_rspace(t, W, pab=pab, forces=True)
    assert rel_err(h, h_ref) < TOL
    assert rel_err(f, f_ref) < TOL


def test_tasklist_rejects_short_buffers(gpw, torch_cuda):
    """Buffers shorter than the task list needs raise ValueError before any
    launch (grid of the wrong size, pab / hab shorter than the largest
    block end, forces shorter than 6 per pair)."""
    from libgpw_amd import workloads as Wk
    torch = torch_cuda
    say("short buffers")
    cfg = Wk.subset(Wk.config(2), np.arange(50))
    tl = gpw.TaskList(gpw.grid_desc(cfg["npts"], cfg["dh"]), cfg["pairs"])
    n = cfg["npts"]
    grid = torch.zeros((n[2], n[1], n[0]), dtype=torch.float64, device="cuda")
    pab = dev(torch, cfg["pab"])
    hab = torch.zeros(cfg["nblock"], dtype=torch.float64, device="cuda")
    f = torch.zeros((50, 6), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        tl.collocate(pab, grid[:-1])
    with pytest.raises(ValueError):
        tl.collocate(pab[:-1], grid)
    with pytest.raises(ValueError):
        tl.integrate(grid, hab[:-1])
    with pytest.raises(ValueError):
        tl.integrate(grid, hab, pab=pab, forces=f[:-1])
    with pytest.raises(ValueError):
        tl.integrate(grid, hab, forces=f)
    with pytest.raises(ValueError):
        tl.collocate(pab.float(), grid)
    # right sizes still work
    tl.collocate(pab, grid)
    tl.integrate(grid, hab, pab=pab, forces=f)
    torch.cuda.synchronize()
    assert torch.isfinite(hab).all()
    tl.close()


@pytest.mark.parametrize("shear", [0.0, 0.25])
def test_tile_index_sets_per_pair(gpw, oracle, torch_cuda, shear):
    """Per pair, the tile kernels write exactly the oracle's point set on a
    grid whose spacing (0.225 bohr) is not a binary fraction, so that many
    points lie within rounding of a sphere or row bound: the in-kernel
    bounds must round as the host rule does (no fused multiply-add). Pairs
    of several radii, one task list per pair; collocate support compared
    exactly, and integrate of a random potential against the oracle."""
    from libgpw_amd import workloads as Wk
    torch = torch_cuda
    rng = np.random.default_rng(91 + int(10 * shear))
    npts = (40, 44, 48)
    cell = np.array([9.0, 9.9, 10.8])
    dh = np.diag(cell / np.asarray(npts))
    dh[1, 0] = shear * dh[0, 0]
    pairs, nblock = _ranged_pairs(rng, 40, dh * np.asarray(npts)[:, None])
    cfg = {"npts": npts, "dh": dh, "pairs": pairs, "pab": rng.uniform(-1.0, 1.0, nblock),
           "nblock": nblock, "name": "index-sets", "seed": 0}
    W = rng.uniform(-1.0, 1.0, (npts[2], npts[1], npts[0]))
    say("per-pair index sets, shear", shear)
    bad = []
    for i in range(40):
        sub = Wk.subset(cfg, [i])
        ref = oracle_batch(oracle, sub, W=W)
        got = run_tasklist(gpw, torch, sub, W=W)
        if not same_support(got["grid"], ref["grid"]):
            bad.append(i)
            continue
        assert rel_err(got["grid"], ref["grid"]) < TOL, i
        assert rel_err(got["hab"], ref["hab"]) < TOL, i
    assert not bad, bad


@pytest.mark.parametrize("case", ["window", "wide", "nonperiodic", "general-window", "general-wide"])
@pytest.mark.parametrize("tau", [False, True])
def test_dropin_host_and_device_arrays(gpw, oracle, torch_cuda, case, tau):
    """The single-task calls on host arrays (only the cube's points travel;
    the whole grid when the cube holds more points than the grid: "wide")
    and on device arrays (the _dev entry points, pab / hab read and written
    in place as maxco x maxco), against the oracle, with forces; the same
    call twice accumulates twice."""
    torch = torch_cuda
    say("drop-in host/device arrays", case, tau)
    rng = np.random.default_rng(900 + 7 * tau + ["window", "wide", "nonperiodic", "general-window",
                                                   "general-wide"].index(case))
    kw = dict(la=(0, 2), lb=(1, 3), tau=tau, ortho=not case.startswith("general"))
    if case in ("window", "general-window"):
        kw.update(npts=(48, 50, 52), cell=(12.0, 12.5, 13.0), zeta=1.5, zetb=2.5)
    elif case in ("wide", "general-wide"):
        kw.update(npts=(16, 18, 20), cell=(4.0, 4.5, 5.0), zeta=0.3, zetb=0.4)
    else:
        kw.update(npts=(40, 36, 44), cell=(10.0, 9.0, 11.0), periodic=(0, 1, 0),
                  window=(30, 36, 32), zeta=2.0, zetb=2.0)
    if case.startswith("general"):
        kw["shear"] = 0.2
    t, pab = random_task(rng, oracle, **kw)
    if case == "nonperiodic":
        t["ra"] = np.array([5.0, 4.0, 5.5]); t["rab"] = np.array([0.3, -0.2, 0.1])
        t["lb_grid"] = [-5, 0, -6]
    ref = oracle.collocate(t, pab)
    assert np.count_nonzero(ref) > 100
    W = rng.uniform(-1, 1, ref.shape)
    h_ref, f_ref = oracle.integrate(t, W, pab=pab, forces=True)
    # host arrays, twice: accumulates
    base = rng.uniform(-1, 1, ref.shape)
    grid = base.copy()
    gpw.collocate_pgf_product_rspace(t, pab, grid)
    assert rel_err(grid - base, ref) < TOL
    gpw.collocate_pgf_product_rspace(t, pab, grid)
    assert rel_err(grid - base, 2 * ref) < TOL
    g0 = np.zeros_like(ref)
    gpw.collocate_pgf_product_rspace(t, pab, g0)
    assert same_support(g0, ref)
    h, f = gpw.integrate_pgf_product_rspace(t, W, pab=pab, forces=True)
    assert rel_err(h, h_ref) < TOL and rel_err(f, f_ref) < TOL
    # device arrays
    gd = dev(torch, base)
    pd = dev(torch, pab)
    gpw.collocate_pgf_product_rspace_dev(t, pd, gd)
    gpw.collocate_pgf_product_rspace_dev(t, pd, gd)
    torch.cuda.synchronize()
    assert rel_err(gd.cpu().numpy() - base, 2 * ref) < TOL
    hd = torch.zeros((t["maxco"], t["maxco"]), dtype=torch.float64, device="cuda")
    fd = torch.zeros(6, dtype=torch.float64, device="cuda")
    gpw.integrate_pgf_product_rspace_dev(t, dev(torch, W), hd, pab=pd, forces=fd)
    torch.cuda.synchronize()
    assert rel_err(hd.cpu().numpy(), h_ref) < TOL
    assert rel_err(fd.cpu().numpy().reshape(2, 3), f_ref) < TOL
    # only the block (o2.., o1..) of hab is written
    mask = np.zeros_like(h_ref, dtype=bool)
    na, nb = (t["la_max"] + 1) * (t["la_max"] + 2) * (t["la_max"] + 3) // 6, \
        (t["lb_max"] + 1) * (t["lb_max"] + 2) * (t["lb_max"] + 3) // 6
    mask[t["o2"]:t["o2"] + nb, t["o1"]:t["o1"] + na] = True
    assert not hd.cpu().numpy()[~mask].any()


@pytest.mark.parametrize("which", ["config2", "config3-tau", "config4-subset", "config5-level0",
                                   "windowed-general"])
def test_device_planner_matches_host_planner(gpw, torch_cuda, which):
    """Task lists are planned on the device (items, offsets, tile lists,
    work statistics); gpw_tasklist_plan_stats runs the host planner. Both
    give the same item count, sub-boxes, sphere points and algorithmic
    flops, for orthorhombic and general cells, windows and tau."""
    from libgpw_amd import workloads as Wk
    say("device vs host planner", which)
    tau = which.endswith("tau")
    if which == "config2":
        cfg = Wk.config(2)
    elif which == "config3-tau":
        cfg = Wk.config(3)
    elif which == "config4-subset":
        cfg = Wk.subset(Wk.config(4), np.arange(0, 100000, 10))
    elif which == "config5-level0":
        cfg = Wk.level(Wk.config(5, npairs=40000), 0)
    else:
        rng = np.random.default_rng(5)
        npts, ngrid, lb_grid, per = (40, 36, 44), (34, 36, 32), (-3, 0, -6), (0, 1, 0)
        dh = np.diag(np.array([10.0, 9.0, 11.0]) / np.asarray(npts))
        dh[1, 0] = 0.2 * dh[0, 0]
        pairs, nblock = _ranged_pairs(rng, 30, dh * np.asarray(npts)[:, None])
        pairs["zeta"] = rng.uniform(2.5, 5.0, 30)
        pairs["zetb"] = rng.uniform(2.5, 5.0, 30)
        pairs["radius"] = np.sqrt(np.log(1e10) / (pairs["zeta"] + pairs["zetb"]))
        # centres inside the window, away from its non-periodic edges
        lo = (12 - np.asarray(lb_grid)) * np.diag(dh)
        hi = (np.asarray(ngrid) - 12 - np.asarray(lb_grid)) * np.diag(dh)
        pairs["ra"] = rng.uniform(lo, hi, (30, 3))
        pairs["rab"] = rng.normal(0.0, 0.3, (30, 3))
        g = gpw.grid_desc(npts, dh, ngrid=ngrid, lb_grid=lb_grid, periodic=per)
        cfg = None
    if cfg is not None:
        g = gpw.grid_desc(cfg["npts"], cfg["dh"])
        pairs = cfg["pairs"]
    host = gpw.plan_stats(g, pairs, compute_tau=tau, splits=3)
    tl = gpw.TaskList(g, pairs, compute_tau=tau, splits=3)
    devs = tl.stats()
    tl.close()
    for k in ("npairs", "nsub", "ntiles", "nentries", "cube_points", "tile_points", "lp_max",
              "flops_collocate"):
        assert devs[k] == host[k], (k, devs[k], host[k])
    assert devs["flops_integrate"] == host["flops_integrate"]


def test_collocator_replans_in_place(gpw, oracle, torch_cuda):
    """A Collocator keeps one task list (tuned splits included) while the
    centres move: each call matches the oracle at its own centres (grid,
    dL/dpab, forces), a call at unchanged centres does not re-plan."""
    from libgpw_amd import workloads as Wk
    from libgpw_amd.autograd import Collocator
    torch = torch_cuda
    cfg = Wk.subset(Wk.config(3), np.arange(600))
    p = cfg["pairs"]
    W = Wk.potential(cfg)
    Wd = dev(torch, W)
    col = Collocator(p, cfg["npts"], cfg["dh"], splits=2, autotune=True)
    rng = np.random.default_rng(3)
    ra0, rb0 = p["ra"].copy(), p["ra"] + p["rab"]
    splits0 = None
    for step in range(3):
        shift = rng.normal(0.0, 0.05, ra0.shape) if step else 0.0
        ra_h, rb_h = ra0 + shift, rb0 + shift
        say("collocator step", step)
        pab = dev(torch, cfg["pab"]).requires_grad_(True)
        ra = torch.tensor(ra_h, dtype=torch.float64, requires_grad=True)
        rb = torch.tensor(rb_h, dtype=torch.float64, requires_grad=True)
        grid = col(pab, ra, rb)
        gpab, gra, grb = torch.autograd.grad((grid * Wd).sum(), (pab, ra, rb))
        moved = dict(cfg, pairs=dict(p, ra=ra_h, rab=rb_h - ra_h))
        ref = oracle_batch(oracle, moved, W=W, forces=True)
        assert rel_err(grid.detach().cpu().numpy(), ref["grid"]) < TOL, step
        assert rel_err(gpab.cpu().numpy(), ref["hab"]) < TOL, step
        f = np.concatenate([gra.numpy(), grb.numpy()], axis=1).reshape(-1, 2, 3)
        assert rel_err(f, ref["forces"]) < TOL, step
        st = col._tl.stats()
        if splits0 is None:
            splits0 = (st["csplits"], st["isplits"])
        assert (st["csplits"], st["isplits"]) == splits0
    assert col.replans == 2
    grid = col(dev(torch, cfg["pab"]), torch.tensor(ra_h), torch.tensor(rb_h))
    assert col.replans == 2
    col.close()


def test_small_sphere_collocate_general_cell(gpw, oracle, torch_cuda):
    """General cell with steep and soft pairs: the steep ones (cube one
    sub-box of at most 16 points per axis) are collocated into per-pair boxes
    and added tile by tile in list order, the soft ones by the tile kernels.
    Both against the oracle (same support, 1e-10), with one and three
    workgroups per tile; repeated calls give bitwise equal grids (every grid
    point is summed in one fixed order)."""
    from libgpw_amd import workloads as Wk
    torch = torch_cuda
    rng = np.random.default_rng(123)
    npts = (48, 40, 44)
    cell = Wk.triclinic_cell(9.6, 8.0, 8.8, 90.0, 90.0, 75.0)
    dh = cell / np.asarray(npts, dtype=np.float64)[:, None]
    steep, _, _ = Wk.make_pairs(rng, 300, 3, 3, 4.0, 40.0, cell)
    soft, _, _ = Wk.make_pairs(rng, 60, 2, 2, 0.2, 1.0, cell)
    pairs = {k: np.concatenate([np.asarray(steep[k]), np.asarray(soft[k])]) for k in steep}
    off, nblock = Wk._blocks(pairs["la_max"], pairs["lb_max"])
    pairs["block_offset"] = off
    cfg = {"npts": npts, "dh": dh, "pairs": pairs, "pab": rng.uniform(-1.0, 1.0, nblock), "nblock": nblock}
    say("small-sphere collocate, general cell", len(pairs["la_max"]), "pairs")
    st = gpw.plan_stats(gpw.grid_desc(npts, dh), pairs)
    assert st["csmall_pairs"] >= 200 and st["csmall_pairs"] < len(pairs["la_max"]), st["csmall_pairs"]
    assert st["csmall_points"] > 0
    ref = oracle_batch(oracle, cfg)
    first = None
    for splits in (1, 1, 3):
        got = run_tasklist(gpw, torch, cfg, splits=splits)
        assert got["stats"]["csmall_pairs"] == st["csmall_pairs"]
        assert same_support(got["grid"], ref["grid"]), splits
        assert rel_err(got["grid"], ref["grid"]) < TOL, splits
        if first is None:
            first = got["grid"]
        elif splits == 1:
            assert np.array_equal(got["grid"], first)
