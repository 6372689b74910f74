"""Pins the CPU oracle (oracle/gpw_oracle.c) before anything is compared to it.

Each check is chosen so a plausible slip in the restatement fails it:
  * the reference's four CP2K task fixtures (x = 0 plane values and the exact
    count of non-zero points: map, cube centre, sphere bounds, tau prep);
  * the cube rule regenerates the fixtures' lb_cube / ub_cube / sphere_bounds;
  * a brute-force evaluator (direct phi_a phi_b, no coefficient transform, no
    1D tables) on the same point set, rho and tau, several l;
  * closed forms: the grid sum of a Gaussian and of its second moment;
  * ortho and general paths agree when the radius covers every significant
    point (pins the multinomial transform and the general point set);
  * integrate is the exact adjoint of collocate; forces match central finite
    differences of <collocate, W>;
  * the libgpw contractions against numpy einsum at the reference test shapes.
"""
import math

import numpy as np
import pytest

from helpers import FIXTURES, fixture, fixture_plane, random_task, rel_err


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_x0_plane(oracle, name):
    t = fixture(name)
    grid = oracle.collocate(t, t["pab"])
    plane = grid[:, :, 0]
    ref = fixture_plane(t)
    assert np.count_nonzero(plane) == len(t["grid_entries"])
    if len(t["grid_entries"]):
        assert rel_err(plane, ref) < 1e-13
    else:
        assert np.abs(grid).max() > 0.0  # the task does collocate somewhere else


@pytest.mark.parametrize("name", ["collocate_ortho_density", "collocate_ortho_tau"])
def test_cube_rule_reproduces_fixture_bounds(oracle, name):
    t = fixture(name)
    dr = [t["dh"][i][i] for i in range(3)]
    lb, ub, sb = oracle.sphere_bounds(t["radius"], dr)
    assert lb == t["lb_cube"] and ub == t["ub_cube"]
    ref = np.asarray(t["sphere_bounds"])
    assert np.array_equal(sb, ref[: len(sb)])
    assert not ref[len(sb):].any()


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_brute_force(oracle, name):
    t = fixture(name)
    g = oracle.collocate(t, t["pab"])
    b = oracle.collocate_brute(t, t["pab"])
    assert rel_err(g, b) < 1e-12


@pytest.mark.parametrize("tau", [False, True])
@pytest.mark.parametrize("ortho", [True, False])
@pytest.mark.parametrize("la,lb", [((0, 0), (0, 0)), ((1, 1), (0, 1)), ((0, 2), (1, 2)), ((3, 3), (2, 3))])
def test_brute_force_random(oracle, tau, ortho, la, lb):
    rng = np.random.default_rng(hash((tau, ortho, la, lb)) % 2**32)
    t, pab = random_task(rng, oracle, la=la, lb=lb, tau=tau, ortho=ortho,
                         shear=0.0 if ortho else 0.3)
    g = oracle.collocate(t, pab)
    b = oracle.collocate_brute(t, pab)
    assert np.abs(g).max() > 0
    assert rel_err(g, b) < 1e-11


def test_gaussian_closed_forms(oracle):
    # s-s pair: grid sum * dV == rscale exp(-zeta zetb/zetp rab^2) (pi/zetp)^1.5
    # px-px pair with ra == rb: sum of (x-p)^2 exp(-zetp r^2) dV == (pi/zetp)^1.5 / (2 zetp)
    rng = np.random.default_rng(5)
    n = 64
    cell = 9.0
    for la in (0, 1):
        t, pab = random_task(rng, oracle, la=(la, la), lb=(la, la), npts=(n, n, n),
                             cell=(cell, cell, cell), zeta=1.1, zetb=0.9, radius=cell * 0.49)
        t["rab"] = np.zeros(3) if la else np.array([0.3, -0.2, 0.5])
        pab[:] = 0.0
        if la == 0:
            pab[t["o2"], t["o1"]] = 1.0
        else:
            pab[t["o2"] + 1, t["o1"] + 1] = 1.0  # coset(1,0,0) = 1: px * px
        grid = oracle.collocate(t, pab)
        dV = (cell / n) ** 3
        zetp = 2.0
        if la == 0:
            exact = t["rscale"] * math.exp(-1.1 * 0.9 / zetp * float(t["rab"] @ t["rab"])) * (math.pi / zetp) ** 1.5
        else:
            exact = t["rscale"] * (math.pi / zetp) ** 1.5 / (2 * zetp)
        assert abs(grid.sum() * dV - exact) / exact < 1e-9


@pytest.mark.parametrize("la,lb", [((0, 0), (0, 0)), ((1, 2), (0, 1)), ((2, 2), (3, 3))])
def test_ortho_equals_general_on_orthorhombic_cell(oracle, la, lb):
    rng = np.random.default_rng(11 + la[1] + lb[1])
    t, pab = random_task(rng, oracle, la=la, lb=lb, zeta=1.3, zetb=1.6, eps=1e-36)
    g_ortho = oracle.collocate(t, pab)
    tg = dict(t, use_ortho=False)
    g_gen = oracle.collocate(tg, pab)
    assert rel_err(g_gen, g_ortho) < 1e-12


@pytest.mark.parametrize("tau", [False, True])
@pytest.mark.parametrize("ortho", [True, False])
@pytest.mark.parametrize("window", [None, (24, 20, 24)])
def test_integrate_is_adjoint(oracle, tau, ortho, window):
    rng = np.random.default_rng(21 + tau + 2 * ortho + 4 * (window is None))
    periodic = (1, 1, 1) if window is None else (0, 1, 0)
    npts = (30, 20, 32)
    win = window
    ran = 0
    for _ in range(5):
        t, pab = random_task(rng, oracle, la=(0, 2), lb=(1, 2), tau=tau, ortho=ortho, npts=npts,
                             cell=(9.0, 6.0, 9.5), periodic=periodic, window=win,
                             shear=0.0 if ortho else 0.25,
                             zeta=None if win is None else 2.0, zetb=None if win is None else 2.0)
        if win is not None:
            # a pair near x = z = 4.5 bohr, inside the local window [2, 20)
            t["ra"] = np.array([4.5, 3.0, 4.7]) + rng.uniform(-0.2, 0.2, 3)
            t["rab"] = np.array([0.2, -0.3, 0.1])
            t["lb_grid"] = [-2, 0, -5]
        try:
            g = oracle.collocate(t, pab)
        except oracle.OracleError:
            continue  # cube outside the window for this draw
        W = rng.uniform(-1, 1, g.shape)
        hab, _ = oracle.integrate(t, W)
        lhs = float((g * W).sum())
        rhs = float((pab * hab).sum())
        assert abs(lhs - rhs) <= 1e-12 * (np.abs(g).sum() * 1.0 + 1e-300)
        ran += 1
    assert ran >= 3


@pytest.mark.parametrize("tau", [False, True])
@pytest.mark.parametrize("ortho", [True, False])
def test_forces_central_fd(oracle, tau, ortho):
    rng = np.random.default_rng(31 + tau + 2 * ortho)
    h = 1e-5
    checked = 0
    for _ in range(3):
        t, pab = random_task(rng, oracle, la=(0, 2), lb=(1, 1), tau=tau, ortho=ortho,
                             shear=0.0 if ortho else 0.25)
        W = rng.uniform(-1, 1, (t["ngrid"][2], t["ngrid"][1], t["ngrid"][0]))
        _, f = oracle.integrate(t, W, pab=pab, forces=True)

        def energy(ra, rab):
            tt = dict(t, ra=ra, rab=rab)
            return float((oracle.collocate(tt, pab) * W).sum())

        def centre(ra, rab):
            f_ = t["zetb"] / (t["zeta"] + t["zetb"])
            rp = ra + f_ * rab
            return np.floor(rp @ t["dh_inv"]).astype(int) if ortho else None

        scale = np.abs(f).max()
        for which in range(2):  # 0: Ra (Rb fixed), 1: Rb (Ra fixed)
            for d in range(3):
                e = np.zeros(3); e[d] = h
                if which == 0:
                    args_p = (t["ra"] + e, t["rab"] - e); args_m = (t["ra"] - e, t["rab"] + e)
                else:
                    args_p = (t["ra"], t["rab"] + e); args_m = (t["ra"], t["rab"] - e)
                if ortho and not np.array_equal(centre(*args_p), centre(*args_m)):
                    continue  # index set changes inside the stencil
                fd = (energy(*args_p) - energy(*args_m)) / (2 * h)
                assert abs(fd - f[which, d]) <= 2e-6 * scale + 1e-12
                checked += 1
    assert checked >= 12


SHAPES = [(14, 23, 17, 3), (14, 23, 17, 1), (1, 12, 29, 5), (13, 45, 29, 7)]


@pytest.mark.parametrize("shape", SHAPES)
def test_libgpw_core_contractions(oracle, shape):
    i, j, k, L = shape
    rng = np.random.default_rng(i * 1000 + j * 10 + k + L)
    P = rng.uniform(-1, 1, (3, L, max(i, j, k)))
    co = rng.uniform(-1, 1, (L, L, L))
    V = oracle.collocate_core((i, j, k), co, P)
    ref = np.einsum("agb,gz,by,ax->zyx", co, P[0, :, :i], P[1, :, :j], P[2, :, :k])
    assert np.abs(V - ref).max() < 1e-14 * max(1.0, np.abs(ref).max())
    W = rng.uniform(-1, 1, (i, j, k))
    out = oracle.integrate_core((i, j, k), P, W)
    ref2 = np.einsum("zyx,gz,by,ax->gba", W, P[0, :, :i], P[1, :, :j], P[2, :, :k])
    assert np.abs(out - ref2).max() < 1e-13 * max(1.0, np.abs(ref2).max())
    # adjointness of the two contractions
    assert abs((V * W).sum() - (co.transpose(1, 2, 0) * out).sum()) < 1e-11 * np.abs(V).sum()


def test_batch_driver_matches_single_tasks(oracle):
    from libgpw_amd import workloads as Wk
    cfg = Wk.subset(Wk.config(3), np.arange(12))
    g = oracle.make_grid(cfg["npts"], cfg["dh"])
    pairs = oracle.make_pairs(cfg["pairs"])
    grid = oracle.batch_collocate(g, pairs, cfg["pab"], nthreads=2)
    ref = np.zeros_like(grid)
    p = cfg["pairs"]
    for i in range(12):
        la, lb = int(p["la_max"][i]), int(p["lb_max"][i])
        lbc, ubc, sb = oracle.sphere_bounds(p["radius"][i], np.diag(cfg["dh"]))
        t = {"compute_tau": False, "use_ortho": True, "la_max": la, "la_min": la, "lb_max": lb,
             "lb_min": lb, "zeta": p["zeta"][i], "zetb": p["zetb"][i], "rscale": 1.0,
             "dh": cfg["dh"], "dh_inv": np.linalg.inv(cfg["dh"]), "ra": p["ra"][i], "rab": p["rab"][i],
             "npts": cfg["npts"], "ngrid": cfg["npts"], "lb_grid": [0, 0, 0], "periodic": [1, 1, 1],
             "lmax": 8, "radius": p["radius"][i], "lb_cube": lbc, "ub_cube": ubc,
             "sphere_bounds": sb, "nspheres": len(sb), "maxco": ncos(la), "o1": 0, "o2": 0}
        blk = cfg["pab"][p["block_offset"][i]:p["block_offset"][i] + ncos(la) * ncos(lb)]
        pab = np.zeros((ncos(la) if ncos(la) >= ncos(lb) else ncos(lb),) * 2)
        t["maxco"] = pab.shape[0]
        pab[: ncos(lb), : ncos(la)] = blk.reshape(ncos(lb), ncos(la))
        oracle.collocate(t, pab, ref)
    assert rel_err(grid, ref) < 1e-13


def ncos(l):
    return (l + 1) * (l + 2) * (l + 3) // 6


def test_oracle_error_codes(oracle):
    rng = np.random.default_rng(3)
    t, pab = random_task(rng, oracle, la=(0, 2), lb=(0, 2))
    with pytest.raises(oracle.OracleError):
        oracle.collocate(dict(t, lmax=1), pab)  # lmax below la_max
    bad = np.array(t["sphere_bounds"]).copy()
    bad[0] = t["lb_cube"][2] - 1
    with pytest.raises(oracle.OracleError):
        oracle.collocate(dict(t, sphere_bounds=bad), pab)
    with pytest.raises(oracle.OracleError):
        oracle.collocate(dict(t, la_min=3), pab)


@pytest.mark.parametrize("name", FIXTURES)
def test_full_goldens_agree_with_fixtures_and_oracle(oracle, name):
    """The full-3D goldens (tests/golden/make_full_tasks.py) hold the
    fixture's own task and, on the x = 0 plane, exactly the fixture's dumped
    values; every other plane is the oracle's collocation."""
    import os
    from libgpw_amd.taskfile import expected_grid, read_task
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    t = fixture(name)
    f = read_task(os.path.join(here, name + ".full.task"))
    for k in ("la_max", "la_min", "lb_max", "lb_min", "zeta", "zetb", "rscale", "radius", "maxco"):
        assert f[k] == t[k], k
    assert np.array_equal(f["pab"], t["pab"]) and np.array_equal(f["ra"], t["ra"])
    full = expected_grid(f)
    plane = fixture_plane(t)
    assert np.count_nonzero(full[:, :, 0]) == np.count_nonzero(plane)
    if plane.any():
        assert rel_err(full[:, :, 0], plane) < 1e-13
    ref = oracle.collocate(t, t["pab"])
    assert np.array_equal(full != 0.0, ref != 0.0)
    assert np.array_equal(full, ref)
