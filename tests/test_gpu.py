"""GPU parity tests: the HIP product, called through the C ABI, against the
CPU oracle on the same seeded inputs (FP64; tolerance 1e-10 relative,
max|x - ref| / max|ref|, unless stated). Index work (which grid point a
cube point lands on) must be exact: the grid points holding a
value are the same set of points (values above 1e-250 in magnitude: a pair whose Gaussian-product
prefactor nearly underflows leaves values in the subnormal range, where the
last-bit differences of FMA contraction decide between 0 and a subnormal)."""
import sys

import numpy as np
import pytest

from helpers import FIXTURES, fixture, fixture_plane, random_task, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-10


def support(a):
    """flat indices of the grid points holding a value (see the module docstring)"""
    return np.flatnonzero(np.abs(np.asarray(a)) > 1e-250)


def same_support(a, b):
    """the two grids hold values on exactly the same set of points"""
    sa, sb = support(a), support(b)
    return sa.size == sb.size and bool(np.array_equal(sa, sb))


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


# ---------------------------------------------------------------- drop-in


@pytest.mark.parametrize("name", FIXTURES)
def test_dropin_fixtures(gpw, oracle, torch_cuda, name):
    say("dropin fixture", name)
    t = fixture(name)
    grid = np.zeros((t["ngrid"][2], t["ngrid"][1], t["ngrid"][0]))
    gpw.collocate_pgf_product_rspace(t, t["pab"], grid)
    ref = oracle.collocate(t, t["pab"])
    assert same_support(grid, ref)
    assert rel_err(grid, ref) < 1e-13
    plane = fixture_plane(t)
    if plane.any():
        assert rel_err(grid[:, :, 0], plane) < 1e-13
    else:
        assert not grid[:, :, 0].any()


@pytest.mark.parametrize("tau", [False, True])
@pytest.mark.parametrize("window", [False, True])
@pytest.mark.parametrize("la,lb", [((0, 0), (0, 0)), ((0, 2), (1, 2)), ((3, 3), (2, 3))])
def test_dropin_random(gpw, oracle, torch_cuda, tau, window, la, lb):
    say("dropin random", tau, window, la, lb)
    rng = np.random.default_rng(100 + 10 * tau + 20 * window + la[1] + 4 * lb[1])
    kw = {}
    if window:
        kw = dict(npts=(30, 20, 32), cell=(9.0, 6.0, 9.5), periodic=(0, 1, 0),
                  window=(24, 20, 24), zeta=2.0, zetb=2.0)
    t, pab = random_task(rng, oracle, la=la, lb=lb, tau=tau, **kw)
    if window:
        t["ra"] = np.array([4.5, 3.0, 4.7]); t["rab"] = np.array([0.2, -0.3, 0.1])
        t["lb_grid"] = [-2, 0, -5]
    ref = oracle.collocate(t, pab)
    grid = np.zeros_like(ref)
    gpw.collocate_pgf_product_rspace(t, pab, grid)
    assert same_support(grid, ref)
    assert rel_err(grid, ref) < TOL
    W = rng.uniform(-1, 1, ref.shape)
    h_ref, f_ref = oracle.integrate(t, W, pab=pab, forces=True)
    h, f = gpw.integrate_pgf_product_rspace(t, W, pab=pab, forces=True)
    assert rel_err(h, h_ref) < TOL
    assert rel_err(f, f_ref) < TOL


@pytest.mark.parametrize("tau", [False, True])
@pytest.mark.parametrize("shear,window", [(0.0, False), (0.35, False), (-0.4, True)])
@pytest.mark.parametrize("la,lb", [((0, 1), (0, 0)), ((2, 3), (1, 3))])
def test_dropin_random_general(gpw, oracle, torch_cuda, tau, shear, window, la, lb):
    # use_ortho = 0: the general path, on a sheared cell or (shear 0) forced
    # on an orthorhombic one
    say("dropin general", tau, shear, window, la, lb)
    rng = np.random.default_rng(300 + 10 * tau + int(100 * shear) + la[1] + 4 * lb[1])
    kw = dict(ortho=False, shear=shear)
    if window:
        kw.update(npts=(30, 20, 32), cell=(9.0, 6.0, 9.5), periodic=(0, 1, 0),
                  window=(28, 20, 28), zeta=2.0, zetb=2.0)
    t, pab = random_task(rng, oracle, la=la, lb=lb, tau=tau, **kw)
    if window:
        t["ra"] = np.array([4.5, 3.0, 4.7]); t["rab"] = np.array([0.2, -0.3, 0.1])
        t["lb_grid"] = [-5, 0, -5]  # the sheared box spans 25 x-planes
    ref = oracle.collocate(t, pab)
    assert np.count_nonzero(ref) > 100
    grid = np.zeros_like(ref)
    gpw.collocate_pgf_product_rspace(t, pab, grid)
    assert same_support(grid, ref)
    assert rel_err(grid, ref) < TOL
    W = rng.uniform(-1, 1, ref.shape)
    h_ref, f_ref = oracle.integrate(t, W, pab=pab, forces=True)
    h, f = gpw.integrate_pgf_product_rspace(t, W, pab=pab, forces=True)
    assert rel_err(h, h_ref) < TOL
    assert rel_err(f, f_ref) < TOL


# ---------------------------------------------------------------- task lists


def run_tasklist(gpw, torch, cfg, plain=False, splits=1, forces=False, tau=False, W=None,
                 autotune=False):
    g = gpw.grid_desc(cfg["npts"], cfg["dh"])
    tl = gpw.TaskList(g, cfg["pairs"], compute_tau=tau, splits=splits)
    if autotune:
        before, after = tl.autotune(max_splits=8, reps=2)
        assert after > 0.0 and after <= before * 1.10 + 0.05, (before, after)
    n = cfg["npts"]
    grid = torch.zeros((n[2], n[1], n[0]), dtype=torch.float64, device="cuda")
    pab = dev(torch, cfg["pab"])
    tl.collocate(pab, grid, plain=plain)
    out = {"grid": grid.cpu().numpy(), "stats": tl.stats()}
    if W is not None:
        hab = torch.zeros(cfg["nblock"], dtype=torch.float64, device="cuda")
        fr = torch.zeros((len(cfg["pairs"]["la_max"]), 6), dtype=torch.float64, device="cuda") if forces else None
        tl.integrate(dev(torch, W), hab, pab=pab, forces=fr, plain=plain)
        out["hab"] = hab.cpu().numpy()
        out["forces"] = None if fr is None else fr.cpu().numpy().reshape(-1, 2, 3)
    torch.cuda.synchronize()
    tl.close()
    return out


def oracle_batch(oracle, cfg, W=None, forces=False, tau=False):
    g = oracle.make_grid(cfg["npts"], cfg["dh"])
    pairs = oracle.make_pairs(cfg["pairs"])
    grid = oracle.batch_collocate(g, pairs, cfg["pab"], compute_tau=tau)
    out = {"grid": grid}
    if W is not None:
        hab, f = oracle.batch_integrate(g, pairs, W, cfg["nblock"], pab=cfg["pab"], forces=forces,
                                        compute_tau=tau)
        out["hab"] = hab
        out["forces"] = f
    return out


@pytest.mark.parametrize("level", [0, 1, 2, 3])
def test_config5_triclinic_levels(gpw, oracle, torch_cuda, level):
    # multigrid workload on its triclinic cell (gamma = 80 deg): each level's
    # task list (collocate, integrate, forces), batched and plain, vs the oracle
    from libgpw_amd import workloads as Wk
    cfg = Wk.level(Wk.config(5, npairs=1500), level)
    say("config5 level", level, len(cfg["pairs"]["la_max"]), "pairs")
    W = Wk.potential(cfg)
    ref = oracle_batch(oracle, cfg, W=W, forces=True)
    assert np.count_nonzero(ref["grid"]) > 0
    for plain, splits in ((False, 1), (False, 4), (True, 1)):
        got = run_tasklist(gpw, torch_cuda, cfg, plain=plain, splits=splits, forces=True, W=W)
        assert same_support(got["grid"], ref["grid"]), (plain, splits)
        assert rel_err(got["grid"], ref["grid"]) < TOL, (plain, splits)
        assert rel_err(got["hab"], ref["hab"]) < TOL, (plain, splits)
        assert rel_err(got["forces"], ref["forces"]) < TOL, (plain, splits)


def test_config5_tau_general(gpw, oracle, torch_cuda):
    from libgpw_amd import workloads as Wk
    cfg = Wk.level(Wk.config(5, npairs=800), 1)
    say("config5 tau level 1", len(cfg["pairs"]["la_max"]), "pairs")
    W = Wk.potential(cfg)
    ref = oracle_batch(oracle, cfg, W=W, forces=True, tau=True)
    got = run_tasklist(gpw, torch_cuda, cfg, splits=2, forces=True, W=W, tau=True)
    assert rel_err(got["grid"], ref["grid"]) < TOL
    assert rel_err(got["hab"], ref["hab"]) < TOL
    assert rel_err(got["forces"], ref["forces"]) < TOL


@pytest.mark.parametrize("plain", [False, True])
def test_config1_single_pair(gpw, oracle, torch_cuda, plain):
    from libgpw_amd import workloads as Wk
    say("config1", "plain" if plain else "batched")
    cfg = Wk.config(1)
    W = Wk.potential(cfg)
    got = run_tasklist(gpw, torch_cuda, cfg, plain=plain, forces=True, W=W)
    ref = oracle_batch(oracle, cfg, W=W, forces=True)
    assert same_support(got["grid"], ref["grid"])
    assert rel_err(got["grid"], ref["grid"]) < TOL
    assert rel_err(got["hab"], ref["hab"]) < TOL
    assert rel_err(got["forces"], ref["forces"]) < TOL


@pytest.mark.parametrize("splits", [1, 3])
def test_config2_subset_batched_and_plain(gpw, oracle, torch_cuda, splits):
    from libgpw_amd import workloads as Wk
    say("config2 subset 600 pairs splits", splits)
    cfg = Wk.subset(Wk.config(2), np.arange(600))
    W = Wk.potential(cfg)
    ref = oracle_batch(oracle, cfg, W=W)
    b = run_tasklist(gpw, torch_cuda, cfg, splits=splits, W=W)
    assert rel_err(b["grid"], ref["grid"]) < TOL
    assert rel_err(b["hab"], ref["hab"]) < TOL
    if splits == 1:
        p = run_tasklist(gpw, torch_cuda, cfg, plain=True, W=W)
        assert rel_err(p["grid"], ref["grid"]) < TOL
        assert rel_err(p["hab"], ref["hab"]) < TOL


def test_autotuned_splits_keep_results(gpw, oracle, torch_cuda):
    # measured per-class splits (mixed, so partial-grid slots are shared
    # between classes with different splits) give the oracle's results
    from libgpw_amd import workloads as Wk
    cfg = Wk.subset(Wk.config(2), np.arange(3000))
    say("autotune config2 subset", len(cfg["pairs"]["la_max"]))
    W = Wk.potential(cfg)
    ref = oracle_batch(oracle, cfg, W=W, forces=True)
    got = run_tasklist(gpw, torch_cuda, cfg, splits=1, forces=True, W=W, autotune=True)
    st = got["stats"]
    allowed = {1, 2, 3, 4, 6, 8}
    assert set(st["csplits"]) <= allowed and set(st["isplits"][0] + st["isplits"][1]) <= allowed
    assert same_support(got["grid"], ref["grid"])
    assert rel_err(got["grid"], ref["grid"]) < TOL
    assert rel_err(got["hab"], ref["hab"]) < TOL
    assert rel_err(got["forces"], ref["forces"]) < TOL


def test_config2_full_size(gpw, oracle, torch_cuda):
    """All 20,000 pairs on 128^3: oracle parity plus the adjoint identity."""
    from libgpw_amd import workloads as Wk
    say("config2 full 20000 pairs")
    cfg = Wk.config(2)
    W = Wk.potential(cfg)
    got = run_tasklist(gpw, torch_cuda, cfg, W=W)
    ref = oracle_batch(oracle, cfg, W=W)
    assert rel_err(got["grid"], ref["grid"]) < TOL
    assert rel_err(got["hab"], ref["hab"]) < TOL
    lhs = float((got["grid"] * W).sum())
    rhs = float((cfg["pab"] * got["hab"]).sum())
    assert abs(lhs - rhs) < 1e-11 * np.abs(got["grid"]).sum()


def test_config3_forces(gpw, oracle, torch_cuda):
    from libgpw_amd import workloads as Wk
    say("config3 forces 512 pairs")
    cfg = Wk.subset(Wk.config(3), np.arange(512))
    W = Wk.potential(cfg)
    got = run_tasklist(gpw, torch_cuda, cfg, W=W, forces=True)
    ref = oracle_batch(oracle, cfg, W=W, forces=True)
    assert rel_err(got["grid"], ref["grid"]) < TOL
    assert rel_err(got["hab"], ref["hab"]) < TOL
    assert rel_err(got["forces"], ref["forces"]) < TOL


def test_tau_batched(gpw, oracle, torch_cuda):
    from libgpw_amd import workloads as Wk
    say("tau batched 300 pairs")
    cfg = Wk.subset(Wk.config(3), np.arange(300))
    W = Wk.potential(cfg)
    got = run_tasklist(gpw, torch_cuda, cfg, W=W, tau=True, forces=True)
    ref = oracle_batch(oracle, cfg, W=W, tau=True, forces=True)
    assert rel_err(got["grid"], ref["grid"]) < TOL
    assert rel_err(got["hab"], ref["hab"]) < TOL
    assert rel_err(got["forces"], ref["forces"]) < TOL


def test_batched_is_deterministic(gpw, torch_cuda):
    from libgpw_amd import workloads as Wk
    say("determinism 2000 pairs")
    cfg = Wk.subset(Wk.config(2), np.arange(2000))
    W = Wk.potential(cfg)
    a = run_tasklist(gpw, torch_cuda, cfg, W=W)
    b = run_tasklist(gpw, torch_cuda, cfg, W=W)
    assert np.array_equal(a["grid"], b["grid"])
    assert np.array_equal(a["hab"], b["hab"])


def test_collocate_accumulates(gpw, oracle, torch_cuda):
    from libgpw_amd import workloads as Wk
    say("accumulate")
    cfg = Wk.subset(Wk.config(2), np.arange(100))
    g = gpw.grid_desc(cfg["npts"], cfg["dh"])
    tl = gpw.TaskList(g, cfg["pairs"])
    n = cfg["npts"]
    base = np.random.default_rng(1).uniform(-1, 1, (n[2], n[1], n[0]))
    grid = dev(torch_cuda, base)
    tl.collocate(dev(torch_cuda, cfg["pab"]), grid)
    ref = oracle_batch(oracle, cfg)["grid"] + base
    assert rel_err(grid.cpu().numpy(), ref) < 1e-14


@pytest.mark.parametrize("shape", [(14, 23, 17, 3), (14, 23, 17, 1), (1, 12, 29, 5), (13, 45, 29, 7)])
def test_libgpw_core(gpw, oracle, torch_cuda, shape):
    say("libgpw core", shape)
    i, j, k, L = shape
    rng = np.random.default_rng(sum(shape))
    P = rng.uniform(-1, 1, (3, L, max(i, j, k)))
    co = rng.uniform(-1, 1, (L, L, L))
    V = gpw.collocate_core((i, j, k), dev(torch_cuda, co), dev(torch_cuda, P)).cpu().numpy()
    assert np.abs(V - oracle.collocate_core((i, j, k), co, P)).max() < 1e-14 * max(1, np.abs(V).max())
    Wg = rng.uniform(-1, 1, (i, j, k))
    out = gpw.integrate_core((i, j, k), dev(torch_cuda, P), dev(torch_cuda, Wg)).cpu().numpy()
    ref = oracle.integrate_core((i, j, k), P, Wg)
    assert np.abs(out - ref).max() < 1e-13 * max(1, np.abs(ref).max())


@pytest.mark.parametrize("name", FIXTURES)
def test_replay_driver_on_fixture(gpw, torch_cuda, name):
    import os
    import subprocess
    say("replay driver", name)
    exe = os.path.join(os.path.dirname(gpw.LIB_PATH), "gpw_replay")
    here = os.path.dirname(os.path.abspath(__file__))
    r = subprocess.run([exe, os.path.join(here, "golden", name + ".task"), "3"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.parametrize("name", FIXTURES)
def test_replay_driver_full_goldens(gpw, torch_cuda, name):
    """The replay driver on the full-3D goldens: every grid point compared
    (--complete: every point the file does not list must come out zero)."""
    import os
    import subprocess
    say("replay driver full golden", name)
    exe = os.path.join(os.path.dirname(gpw.LIB_PATH), "gpw_replay")
    here = os.path.dirname(os.path.abspath(__file__))
    r = subprocess.run([exe, os.path.join(here, "golden", name + ".full.task"), "2", "--complete"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "and every other point" in r.stdout


def test_replay_driver_batch(gpw, torch_cuda):
    """--batch: files on one grid become one device task list, collocated
    `cycles` times; repeated files are repeated pairs. Every fixture is
    reproduced by the task list's own cube rule, so nothing falls back to
    the drop-in."""
    import os
    import subprocess
    say("replay driver batch")
    exe = os.path.join(os.path.dirname(gpw.LIB_PATH), "gpw_replay")
    g = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    files = [os.path.join(g, n + ".full.task") for n in
             ["collocate_ortho_density"] * 3 + ["collocate_ortho_tau", "collocate_general_density"] +
             ["collocate_general_tau"] * 2]
    r = subprocess.run([exe, "--batch", "3", "--complete"] + files, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("batch of")]
    assert len(lines) == 4, r.stdout
    assert "batch of 3 tasks" in lines[0] and "(3 in the task list, 0 drop-in)" in lines[0]
    assert "batch of 2 tasks" in lines[3] and "(2 in the task list, 0 drop-in)" in lines[3]


@pytest.mark.parametrize("name", FIXTURES)
def test_tasklist_full_goldens(gpw, torch_cuda, name):
    """Each fixture as a pair of a task list (three copies of it): every
    point of the grid equals three times the full-3D golden, and the points
    holding a value are the golden's."""
    from libgpw_amd.taskfile import batch, expected_grid, read_task
    import os
    torch = torch_cuda
    say("task list on full golden", name)
    here = os.path.dirname(os.path.abspath(__file__))
    t = read_task(os.path.join(here, "golden", name + ".full.task"))
    b = batch([t, t, t])
    g = gpw.grid_desc(b["npts"], b["dh"], ngrid=b["ngrid"], lb_grid=b["lb_grid"], periodic=b["periodic"])
    tl = gpw.TaskList(g, b["pairs"], compute_tau=b["compute_tau"])
    n = b["ngrid"]
    grid = torch.zeros((n[2], n[1], n[0]), dtype=torch.float64, device="cuda")
    tl.collocate(dev(torch, b["pab"]), grid)
    torch.cuda.synchronize()
    ref = 3.0 * expected_grid(t)
    got = grid.cpu().numpy()
    assert same_support(got, ref)
    assert rel_err(got, ref) < 1e-13
    tl.close()


def test_libgpw_core_host_arrays(gpw, oracle, torch_cuda):
    import ctypes as C
    say("libgpw core host arrays")
    rng = np.random.default_rng(9)
    length, L = (13, 45, 29), 7
    P = rng.uniform(-1, 1, (3, L, 45))
    co = rng.uniform(-1, 1, (L, L, L))
    V = np.zeros(length)
    dp = C.POINTER(C.c_double)
    lng = (C.c_int * 3)(*length)
    assert gpw.lib().gpw_collocate_core_host(lng, L, co.ctypes.data_as(dp), P.ctypes.data_as(dp), 45,
                                             V.ctypes.data_as(dp)) == 0
    ref = oracle.collocate_core(length, co, P)
    assert np.abs(V - ref).max() < 1e-14 * np.abs(ref).max()
    out = np.zeros((L, L, L))
    assert gpw.lib().gpw_integrate_core_host(lng, L, P.ctypes.data_as(dp), 45, V.ctypes.data_as(dp),
                                             out.ctypes.data_as(dp)) == 0
    ref2 = oracle.integrate_core(length, P, V)
    assert np.abs(out - ref2).max() < 1e-13 * np.abs(ref2).max()


# ---------------------------------------------------------------- autograd, graphs


def test_autograd_matches_oracle_and_fd(gpw, oracle, torch_cuda):
    """Differentiation workload (4,096 pairs, 96x96x112): grad pab = integrate,
    grad centres = forces; forces also against a per-pair central finite
    difference (h = 1e-5 bohr) of <W, collocate>."""
    from libgpw_amd import workloads as Wk
    from libgpw_amd.autograd import collocate
    torch = torch_cuda
    say("autograd config3 4096 pairs")
    cfg = Wk.config(3)
    p = cfg["pairs"]
    W = Wk.potential(cfg)
    Wd = dev(torch, W)
    pab = dev(torch, cfg["pab"]).requires_grad_(True)
    ra = torch.tensor(p["ra"], dtype=torch.float64, requires_grad=True)
    rb = torch.tensor(p["ra"] + p["rab"], dtype=torch.float64, requires_grad=True)
    grid = collocate(pab, ra, rb, p, cfg["npts"], cfg["dh"])
    loss = (grid * Wd).sum()
    gpab, gra, grb = torch.autograd.grad(loss, (pab, ra, rb))
    ref = oracle_batch(oracle, cfg, W=W, forces=True)
    assert rel_err(grid.detach().cpu().numpy(), ref["grid"]) < TOL
    assert rel_err(gpab.cpu().numpy(), ref["hab"]) < TOL
    f = np.concatenate([gra.numpy(), grb.numpy()], axis=1).reshape(-1, 2, 3)
    assert rel_err(f, ref["forces"]) < TOL

    # per-pair energies E_p = <pab_p, integrate_p(W)> at displaced centres
    g = gpw.grid_desc(cfg["npts"], cfg["dh"])
    pab_h = cfg["pab"]
    off = p["block_offset"]
    sizes = [gpw.ncoset(int(b)) * gpw.ncoset(int(a)) for a, b in zip(p["la_max"], p["lb_max"])]
    seg = np.repeat(np.arange(len(sizes)), sizes)

    def energies(ra_, rab_):
        q = dict(p, ra=ra_, rab=rab_)
        tl = gpw.TaskList(g, q)
        hab = torch.zeros(cfg["nblock"], dtype=torch.float64, device="cuda")
        tl.integrate(Wd, hab)
        h = hab.cpu().numpy()
        tl.close()
        return np.bincount(seg, weights=pab_h * h, minlength=len(sizes))

    h = 1e-5
    dh_inv = np.linalg.inv(cfg["dh"])
    fr = p["zetb"] / (p["zeta"] + p["zetb"])

    def centres(ra_, rab_):
        rp = ra_ + fr[:, None] * rab_
        return np.floor(rp @ dh_inv).astype(int)

    checked = 0
    scale = np.abs(f).max()
    for which in range(2):
        for d in range(3):
            e = np.zeros(3)
            e[d] = h
            if which == 0:
                ap, am = (p["ra"] + e, p["rab"] - e), (p["ra"] - e, p["rab"] + e)
            else:
                ap, am = (p["ra"], p["rab"] + e), (p["ra"], p["rab"] - e)
            same = np.all(centres(*ap) == centres(*am), axis=1)
            fd = (energies(*ap) - energies(*am)) / (2 * h)
            err = np.abs(fd - f[:, which, d])[same]
            assert err.max() <= 1e-6 * scale, (which, d, err.max(), scale)
            checked += int(same.sum())
    assert checked > 6 * 4000


def test_hip_graph_replays(gpw, torch_cuda):
    """collocate + integrate + forces captured in one HIP graph: every replay,
    with eager work in between, equals the eager result bit for bit."""
    from libgpw_amd import workloads as Wk
    from libgpw_amd.autograd import capture_step
    torch = torch_cuda
    say("hip graph capture 1500 pairs")
    cfg = Wk.subset(Wk.config(2), np.arange(1500))
    tl = gpw.TaskList(gpw.grid_desc(cfg["npts"], cfg["dh"]), cfg["pairs"], splits=2)
    n = cfg["npts"]
    pab = dev(torch, cfg["pab"])
    pot = dev(torch, Wk.potential(cfg))
    grid = torch.zeros((n[2], n[1], n[0]), dtype=torch.float64, device="cuda")
    hab = torch.zeros(cfg["nblock"], dtype=torch.float64, device="cuda")
    forces = torch.zeros((1500, 6), dtype=torch.float64, device="cuda")
    # eager reference
    tl.collocate(pab, grid)
    tl.integrate(pot, hab, pab=pab, forces=forces)
    torch.cuda.synchronize()
    ref = (grid.clone(), hab.clone(), forces.clone())
    graph = capture_step(tl, pab, grid, pot, hab, forces)
    for rep in range(4):
        grid.fill_(7.0)  # eager work between replays
        hab.fill_(-3.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(grid, ref[0]), rep
        assert torch.equal(hab, ref[1]), rep
        assert torch.equal(forces, ref[2]), rep
    tl.close()


def test_multigrid_graph_matches_oracle(gpw, oracle, torch_cuda):
    """Config 5 (triclinic, four levels, pairs by exponent): collocate +
    integrate + forces of all levels in one HIP graph, replayed twice, equals
    the oracle level by level; hab blocks of all levels share one buffer."""
    from libgpw_amd import workloads as Wk
    from libgpw_amd.multigrid import MultiGrid
    torch = torch_cuda
    cfg = Wk.config(5, npairs=2000)
    say("multigrid graph", [len(lv["index"]) for lv in cfg["levels"]])
    mg = MultiGrid(cfg, splits=2)
    mg.autotune(max_splits=4, reps=1)
    pab = dev(torch, cfg["pab"])
    grids = mg.grids()
    pots_h = [np.random.default_rng(50 + k).uniform(-1, 1, tuple(g.shape)) for k, g in enumerate(grids)]
    pots = [dev(torch, p) for p in pots_h]
    hab = torch.zeros(cfg["nblock"], dtype=torch.float64, device="cuda")
    forces = mg.forces_buffers()
    graph = mg.capture(pab, grids, pots, hab, forces)
    for rep in range(2):
        for g in grids:
            g.fill_(5.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
    hab_ref = np.zeros(cfg["nblock"])
    f_ref = np.zeros((mg.npairs, 2, 3))
    for k, lv in enumerate(mg.levels):
        sub = {kk: np.asarray(v)[lv["index"]] for kk, v in cfg["pairs"].items()}
        og = oracle.make_grid(lv["npts"], cfg["levels"][k]["dh"])
        op = oracle.make_pairs(sub)
        ref = oracle.batch_collocate(og, op, cfg["pab"])
        got = grids[k].cpu().numpy()
        assert same_support(got, ref), k
        assert rel_err(got, ref) < TOL, k
        h, f = oracle.batch_integrate(og, op, pots_h[k], cfg["nblock"], pab=cfg["pab"], forces=True)
        hab_ref += h
        f_ref[lv["index"]] = f
    assert rel_err(hab.cpu().numpy(), hab_ref) < TOL
    assert rel_err(mg.forces_by_pair(forces), f_ref) < TOL
    mg.close()


@pytest.mark.gpu
def test_multigrid_concurrent_levels_match_sequential(gpw, torch_cuda):
    """Levels on streams of their own give the same grids, hab and forces,
    bit for bit, as the levels one after the other and as the single-stream
    graph; two concurrent steps in a row agree too."""
    from libgpw_amd import workloads as Wk
    from libgpw_amd.multigrid import MultiGrid
    torch = torch_cuda
    cfg = Wk.config(5, npairs=3000)
    say("multigrid concurrent", [len(lv["index"]) for lv in cfg["levels"]])
    mg = MultiGrid(cfg, splits=2)
    pab = dev(torch, cfg["pab"])
    grids = mg.grids()
    pots = [dev(torch, np.random.default_rng(60 + k).uniform(-1, 1, tuple(g.shape)))
            for k, g in enumerate(grids)]
    hab = torch.zeros(cfg["nblock"], dtype=torch.float64, device="cuda")
    forces = mg.forces_buffers()

    def snap():
        torch.cuda.synchronize()
        return ([g.cpu().numpy().copy() for g in grids], hab.cpu().numpy().copy(),
                [f.cpu().numpy().copy() for f in forces])

    mg.step(pab, grids, pots, hab, forces, concurrent=False)
    ref = snap()
    graph = mg.capture(pab, grids, pots, hab, forces)
    graph.replay()
    runs = [snap()]
    for _ in range(2):
        for g in grids:
            g.fill_(3.0)
        mg.step(pab, grids, pots, hab, forces, concurrent=True)
        runs.append(snap())
    for got in runs:
        for a, b in zip(got[0], ref[0]):
            assert np.array_equal(a, b)
        assert np.array_equal(got[1], ref[1])
        for a, b in zip(got[2], ref[2]):
            assert np.array_equal(a, b)
    mg.close()


@pytest.mark.parametrize("plain", [False, True])
def test_contracted_shell_pairs(gpw, oracle, torch_cuda, plain):
    """Contracted front end: primitive pairs share their shell pair's block
    with rscale = c_a c_b. Against the oracle on the expanded form (own block
    per primitive pair, coefficients c_a c_b P, rscale 1): grid, contracted
    hab = sum over primitives of c_a c_b hab_prim, and per-primitive forces."""
    from libgpw_amd import workloads as Wk
    from libgpw_amd.contracted import ShellPairs
    torch = torch_cuda
    shells, cell = Wk.contracted_shells(60, lmax=2)
    sp = ShellPairs(shells, cell)
    say("contracted", sp.nshells, "shell pairs", sp.nprimitive, "primitive pairs", "plain" if plain else "")
    npts = (48, 48, 48)
    dh = cell / np.asarray(npts)[:, None]
    rng = np.random.default_rng(5)
    P = rng.uniform(-1.0, 1.0, sp.nblock)
    W = rng.uniform(-1.0, 1.0, (npts[2], npts[1], npts[0]))
    # expanded form for the oracle
    ex = {k: np.array(v, copy=True) for k, v in sp.pairs.items()}
    sizes = np.array([(la + 1) * (la + 2) * (la + 3) // 6 * ((lb + 1) * (lb + 2) * (lb + 3) // 6)
                      for la, lb in zip(ex["la_max"], ex["lb_max"])])
    ex_off = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    pab_ex = np.zeros(int(sizes.sum()))
    for i in range(sp.nprimitive):
        bo = sp.pairs["block_offset"][i]
        pab_ex[ex_off[i]:ex_off[i] + sizes[i]] = sp.pairs["rscale"][i] * P[bo:bo + sizes[i]]
    ex["rscale"] = np.ones(sp.nprimitive)
    ex["block_offset"] = ex_off
    og = oracle.make_grid(npts, dh)
    op = oracle.make_pairs(ex)
    grid_ref = oracle.batch_collocate(og, op, pab_ex)
    hab_ex, f_ref = oracle.batch_integrate(og, op, W, int(sizes.sum()), pab=pab_ex, forces=True)
    hab_ref = np.zeros(sp.nblock)
    for i in range(sp.nprimitive):
        bo = sp.pairs["block_offset"][i]
        hab_ref[bo:bo + sizes[i]] += sp.pairs["rscale"][i] * hab_ex[ex_off[i]:ex_off[i] + sizes[i]]
    # product
    tl = sp.tasklist(gpw.grid_desc(npts, dh), splits=2)
    grid = torch.zeros((npts[2], npts[1], npts[0]), dtype=torch.float64, device="cuda")
    hab = torch.zeros(sp.nblock, dtype=torch.float64, device="cuda")
    forces = torch.zeros((sp.nprimitive, 6), dtype=torch.float64, device="cuda")
    Pd = dev(torch, P)
    tl.collocate(Pd, grid, plain=plain)
    tl.integrate(dev(torch, W), hab, pab=Pd, forces=forces, plain=plain)
    torch.cuda.synchronize()
    assert same_support(grid.cpu().numpy(), grid_ref)
    assert rel_err(grid.cpu().numpy(), grid_ref) < TOL
    assert rel_err(hab.cpu().numpy(), hab_ref) < TOL
    assert rel_err(forces.cpu().numpy().reshape(-1, 2, 3), f_ref) < TOL
    fs = sp.forces_by_shell(forces)
    assert fs.shape == (sp.nshells, 2, 3)
    tl.close()


@pytest.mark.parametrize("shear", [0.0, 0.2])
def test_tasklist_windowed_nonperiodic(gpw, oracle, torch_cuda, shear):
    """Task list on a local window with non-periodic x and z (lb_grid, ngrid <
    npts), several pairs, l <= 2, collocate + integrate + forces, against the
    oracle's single-task path pair by pair; shear > 0: a general cell."""
    torch = torch_cuda
    rng = np.random.default_rng(17)
    npts, ngrid, lb_grid, per = (40, 36, 44), (34, 36, 32), (-3, 0, -6), (0, 1, 0)
    cell = np.array([10.0, 9.0, 11.0])
    dh = np.diag(cell / np.asarray(npts))
    dh[1, 0] = shear * dh[0, 0]
    n = 14
    la = rng.integers(0, 3, n); lb = rng.integers(0, 3, n)
    zeta = rng.uniform(2.5, 5.0, n); zetb = rng.uniform(2.5, 5.0, n)
    # centres inside the window, away from its non-periodic edges
    lo = (12 - np.asarray(lb_grid)) * np.diag(dh)  # window = indices -lb_grid .. -lb_grid + ngrid - 1
    hi = (np.asarray(ngrid) - 12 - np.asarray(lb_grid)) * np.diag(dh)
    ra = rng.uniform(lo, hi, (n, 3))
    rab = rng.normal(0.0, 0.3, (n, 3))
    radius = np.sqrt(np.log(1e10) / (zeta + zetb))
    sizes = np.array([gpw.ncoset(int(b)) * gpw.ncoset(int(a)) for a, b in zip(la, lb)])
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    pab = rng.uniform(-1.0, 1.0, int(sizes.sum()))
    pairs = {"la_max": la.astype(np.int32), "la_min": la.astype(np.int32),
             "lb_max": lb.astype(np.int32), "lb_min": lb.astype(np.int32),
             "zeta": zeta, "zetb": zetb, "rscale": np.ones(n), "ra": ra, "rab": rab,
             "radius": radius, "block_offset": off}
    W = rng.uniform(-1.0, 1.0, (ngrid[2], ngrid[1], ngrid[0]))
    grid_ref = np.zeros_like(W)
    hab_ref = np.zeros(int(sizes.sum()))
    f_ref = np.zeros((n, 2, 3))
    for i in range(n):
        lbc, ubc, sb = oracle.sphere_bounds(radius[i], np.diag(dh))
        t = {"compute_tau": False, "use_ortho": not shear, "la_max": int(la[i]), "la_min": int(la[i]),
             "lb_max": int(lb[i]), "lb_min": int(lb[i]), "zeta": zeta[i], "zetb": zetb[i],
             "rscale": 1.0, "dh": dh, "dh_inv": np.linalg.inv(dh), "ra": ra[i], "rab": rab[i],
             "npts": list(npts), "ngrid": list(ngrid), "lb_grid": list(lb_grid),
             "periodic": list(per), "lmax": 4, "radius": radius[i], "lb_cube": lbc,
             "ub_cube": ubc, "sphere_bounds": sb, "nspheres": len(sb),
             "maxco": gpw.ncoset(int(la[i])), "pab_ld": gpw.ncoset(int(la[i])),
             "pab_rows": gpw.ncoset(int(lb[i])), "o1": 0, "o2": 0}
        blk = pab[off[i]:off[i] + sizes[i]]
        oracle.collocate(t, blk, grid_ref)
        h, f = oracle.integrate(t, W, pab=blk, forces=True)
        hab_ref[off[i]:off[i] + sizes[i]] = np.asarray(h).reshape(-1)
        f_ref[i] = f
    say("windowed task list", n, "pairs, shear", shear)
    g = gpw.grid_desc(npts, dh, ngrid=ngrid, lb_grid=lb_grid, periodic=per)
    for splits in (1, 3):
        tl = gpw.TaskList(g, pairs, splits=splits)
        grid = torch.zeros((ngrid[2], ngrid[1], ngrid[0]), dtype=torch.float64, device="cuda")
        hab = torch.zeros(int(sizes.sum()), dtype=torch.float64, device="cuda")
        forces = torch.zeros((n, 6), dtype=torch.float64, device="cuda")
        pd = dev(torch, pab)
        tl.collocate(pd, grid)
        tl.integrate(dev(torch, W), hab, pab=pd, forces=forces)
        torch.cuda.synchronize()
        assert same_support(grid.cpu().numpy(), grid_ref), splits
        assert rel_err(grid.cpu().numpy(), grid_ref) < TOL, splits
        assert rel_err(hab.cpu().numpy(), hab_ref) < TOL, splits
        assert rel_err(forces.cpu().numpy().reshape(-1, 2, 3), f_ref) < TOL, splits
        tl.close()


@pytest.mark.parametrize("case", ["wrap", "steep"])
def test_general_tasklist_row_recursion_edges(gpw, oracle, torch_cuda, case):
    """General-cell task lists where the tile kernels cannot run the Gaussian
    recursion along a whole 8-point x block and take the per-point path:
    "wrap": x period 20 (not a multiple of 8) and Gaussians wider than the
    period, so blocks hold points on both sides of a periodic wrap of one
    sub-box; "steep": zetp |h0|^2 of 18 to 40, so the Gaussian of a block's
    first point underflows while later points are inside the sphere.
    Collocate, integrate and forces against the oracle."""
    torch = torch_cuda
    rng = np.random.default_rng(41 if case == "wrap" else 42)
    if case == "wrap":
        npts, cell, n = (20, 24, 28), np.array([8.0, 9.0, 10.0]), 12
        zeta = rng.uniform(0.25, 0.5, n); zetb = rng.uniform(0.25, 0.5, n)
    else:
        npts, cell, n = (48, 40, 44), np.array([12.0, 10.0, 11.0]), 60
        zeta = rng.uniform(150.0, 320.0, n); zetb = rng.uniform(150.0, 320.0, n)
    dh = np.diag(cell / np.asarray(npts))
    dh[1, 0] = 0.3 * dh[0, 0]
    dh[2, 1] = -0.2 * dh[1, 1]
    la = rng.integers(0, 3, n); lb = rng.integers(0, 3, n)
    ra = rng.uniform(0.0, 1.0, (n, 3)) @ (dh * np.asarray(npts)[:, None])
    rab = rng.normal(0.0, 0.3 if case == "wrap" else 0.02, (n, 3))
    sizes = np.array([gpw.ncoset(int(b)) * gpw.ncoset(int(a)) for a, b in zip(la, lb)])
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    pairs = {"la_max": la.astype(np.int32), "la_min": la.astype(np.int32),
             "lb_max": lb.astype(np.int32), "lb_min": lb.astype(np.int32),
             "zeta": zeta, "zetb": zetb, "rscale": np.ones(n), "ra": ra, "rab": rab,
             "radius": np.sqrt(np.log(1e10) / (zeta + zetb)), "block_offset": off}
    cfg = {"npts": npts, "dh": dh, "pairs": pairs, "pab": rng.uniform(-1.0, 1.0, int(sizes.sum())),
           "nblock": int(sizes.sum())}
    W = rng.uniform(-1.0, 1.0, (npts[2], npts[1], npts[0]))
    say("general row recursion edges", case, n, "pairs")
    ref = oracle_batch(oracle, cfg, W=W, forces=True)
    assert np.count_nonzero(ref["grid"]) > 100
    for splits in (1, 2):
        got = run_tasklist(gpw, torch, cfg, splits=splits, forces=True, W=W)
        assert same_support(got["grid"], ref["grid"]), splits
        assert rel_err(got["grid"], ref["grid"]) < TOL, splits
        assert rel_err(got["hab"], ref["hab"]) < TOL, splits
        assert rel_err(got["forces"], ref["forces"]) < TOL, splits


@pytest.mark.parametrize("cell_kind", ["ortho", "general"])
def test_tasklist_small_sphere_limits(gpw, oracle, torch_cuda, cell_kind):
    """Pairs whose sub-boxes lie on both sides of the size limits of the
    small-sphere integrate kernel (32 rows per axis on orthorhombic cells, 48
    on general ones; several passes of z-planes for the larger ones) in one
    task list: collocate, integrate and forces against the oracle."""
    torch = torch_cuda
    rng = np.random.default_rng(61 if cell_kind == "ortho" else 62)
    npts, cell, n = (80, 72, 76), np.array([16.0, 14.4, 15.2]), 12
    dh = np.diag(cell / np.asarray(npts))
    if cell_kind == "general":
        dh[1, 0] = 0.25 * dh[0, 0]
        zp = rng.uniform(0.8, 1.6, n)    # lattice boxes about 40 to 56 rows wide
    else:
        zp = rng.uniform(1.9, 3.4, n)    # boxes about 27 to 36 rows wide
    f = rng.uniform(0.3, 0.7, n)
    zeta, zetb = zp * f, zp * (1.0 - f)
    la = rng.integers(0, 3, n); lb = rng.integers(0, 3, n)
    ra = rng.uniform(0.0, 1.0, (n, 3)) @ (dh * np.asarray(npts)[:, None])
    rab = rng.normal(0.0, 0.3, (n, 3))
    sizes = np.array([gpw.ncoset(int(b)) * gpw.ncoset(int(a)) for a, b in zip(la, lb)])
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    pairs = {"la_max": la.astype(np.int32), "la_min": la.astype(np.int32),
             "lb_max": lb.astype(np.int32), "lb_min": lb.astype(np.int32),
             "zeta": zeta, "zetb": zetb, "rscale": np.ones(n), "ra": ra, "rab": rab,
             "radius": np.sqrt(np.log(1e10) / (zeta + zetb)), "block_offset": off}
    cfg = {"npts": npts, "dh": dh, "pairs": pairs, "pab": rng.uniform(-1.0, 1.0, int(sizes.sum())),
           "nblock": int(sizes.sum())}
    W = rng.uniform(-1.0, 1.0, (npts[2], npts[1], npts[0]))
    say("small-sphere limits", cell_kind, n, "pairs")
    ref = oracle_batch(oracle, cfg, W=W, forces=True)
    got = run_tasklist(gpw, torch, cfg, splits=2, forces=True, W=W)
    assert same_support(got["grid"], ref["grid"])
    assert rel_err(got["grid"], ref["grid"]) < TOL
    assert rel_err(got["hab"], ref["hab"]) < TOL
    assert rel_err(got["forces"], ref["forces"]) < TOL
