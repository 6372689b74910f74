"""GPU tests of the two forms of the general-cell lattice transform.

Cells whose steps are triangular (a along x, b in the xy plane) run the
transform in three substitution passes from the six steps; every other
general cell reads the expanded tables. Checked against the oracle:

* a triclinic cell with alpha, beta, gamma all off 90 degrees (the passes,
  b0, c0 and c1 non-zero) and the same cell rotated about a skew axis (no
  zero step: the tables), through task lists (tile kernels with and without
  splits, and the plain kernels; rho and tau; forces; la_min < la_max) and
  the single-task drop-in;
* task lists on cells with only c0 or only c1 non-zero (the passes whose
  step is zero skip their sums).
Tolerance 1e-10 relative (max|x - ref| / max|ref|).
"""
import sys

import numpy as np
import pytest

from helpers import random_task, rel_err
from test_gpu import oracle_batch, run_tasklist, same_support
from test_gpu_configs import _ranged_pairs

pytestmark = pytest.mark.gpu
TOL = 1e-10
NPTS = (40, 44, 48)


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _rotation(axis, angle):
    axis = np.asarray(axis, dtype=np.float64)
    axis /= np.linalg.norm(axis)
    k = np.array([[0.0, -axis[2], axis[1]], [axis[2], 0.0, -axis[0]], [-axis[1], axis[0], 0.0]])
    return np.eye(3) + np.sin(angle) * k + (1.0 - np.cos(angle)) * (k @ k)


def cell_steps(form):
    """dh (rows: steps) of the test cell in the given form: "stepped" (a
    triclinic cell, a along x, b in the xy plane), "rotated" (the same cell
    turned about a skew axis: no zero step), "c0" / "c1" (an orthorhombic cell
    with only dh[2][0], or only dh[2][1], non-zero: the passes with zero
    steps b0 and c1, or b0 and c0)"""
    from libgpw_amd import workloads as Wk
    if form in ("c0", "c1"):
        dh = np.diag(np.array([9.0, 9.9, 10.8]) / np.asarray(NPTS, dtype=np.float64))
        if form == "c0":
            dh[2, 0] = 0.3 * dh[0, 0]
        else:
            dh[2, 1] = -0.25 * dh[1, 1]
        return dh
    hmat = Wk.triclinic_cell(9.0, 9.9, 10.8, 75.0, 80.0, 70.0)
    if form == "rotated":
        hmat = hmat @ _rotation([1.0, 2.0, 0.5], 0.6).T
    dh = hmat / np.asarray(NPTS, dtype=np.float64)[:, None]
    upper = (dh[0, 1], dh[0, 2], dh[1, 2])
    if form == "stepped":
        assert all(v == 0.0 for v in upper) and dh[2, 0] != 0.0 and dh[2, 1] != 0.0
    else:
        assert all(v != 0.0 for v in upper)
    return dh


@pytest.mark.parametrize("tau", [False, True])
@pytest.mark.parametrize("form", ["stepped", "rotated", "c0", "c1"])
def test_tasklist_cell_forms(gpw, oracle, torch_cuda, form, tau):
    torch = torch_cuda
    rng = np.random.default_rng(515 + 2 * tau + 10 * ["stepped", "rotated", "c0", "c1"].index(form))
    dh = cell_steps(form)
    cell_m = dh * np.asarray(NPTS)[:, None]
    pairs, nblock = _ranged_pairs(rng, 150, cell_m)
    cfg = {"npts": NPTS, "dh": dh, "pairs": pairs, "pab": rng.uniform(-1.0, 1.0, nblock),
           "nblock": nblock}
    W = rng.uniform(-1.0, 1.0, (NPTS[2], NPTS[1], NPTS[0]))
    say("cell form", form, "tau" if tau else "rho")
    ref = oracle_batch(oracle, cfg, W=W, forces=True, tau=tau)
    assert np.count_nonzero(ref["grid"]) > 1000
    for plain, splits in ((False, 1), (False, 3), (True, 1)):
        got = run_tasklist(gpw, torch, cfg, plain=plain, splits=splits, forces=True, W=W, tau=tau)
        assert same_support(got["grid"], ref["grid"]), (plain, splits)
        assert rel_err(got["grid"], ref["grid"]) < TOL, (plain, splits)
        assert rel_err(got["hab"], ref["hab"]) < TOL, (plain, splits)
        assert rel_err(got["forces"], ref["forces"]) < TOL, (plain, splits)


@pytest.mark.parametrize("tau", [False, True])
@pytest.mark.parametrize("form", ["stepped", "rotated"])
def test_dropin_cell_forms(gpw, oracle, torch_cuda, form, tau):
    say("drop-in cell form", form, "tau" if tau else "rho")
    rng = np.random.default_rng(616 + 2 * tau + (form == "rotated"))
    for la, lb in (((0, 2), (1, 3)), ((3, 4), (2, 2))):
        t, pab = random_task(rng, oracle, la=la, lb=lb, tau=tau, ortho=False, npts=NPTS)
        t["dh"] = cell_steps(form)
        t["dh_inv"] = np.linalg.inv(t["dh"])
        ref = oracle.collocate(t, pab)
        assert np.count_nonzero(ref) > 100
        grid = np.zeros_like(ref)
        gpw.collocate_pgf_product_rspace(t, pab, grid)
        assert same_support(grid, ref), (la, lb)
        assert rel_err(grid, ref) < TOL, (la, lb)
        W = rng.uniform(-1, 1, ref.shape)
        h_ref, f_ref = oracle.integrate(t, W, pab=pab, forces=True)
        h, f = gpw.integrate_pgf_product_rspace(t, W, pab=pab, forces=True)
        assert rel_err(h, h_ref) < TOL, (la, lb)
        assert rel_err(f, f_ref) < TOL, (la, lb)
