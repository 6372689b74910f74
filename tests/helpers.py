"""Shared helpers for the test-suite: random single tasks in the reference's
argument layout (grid_base_ref_c.h:40-67) and error metrics."""
from __future__ import annotations

import math

import numpy as np

from libgpw_amd.taskfile import read_task


def ncoset(l):
    return 0 if l < 0 else (l + 1) * (l + 2) * (l + 3) // 6


def rel_err(x, ref):
    """max|x - ref| / max|ref| (the error metric of the benchmark)."""
    x = np.asarray(x)
    ref = np.asarray(ref)
    scale = np.abs(ref).max()
    if scale == 0.0:
        return float(np.abs(x).max())
    return float(np.abs(x - ref).max() / scale)


def random_task(rng, oracle, la=(0, 1), lb=(0, 1), tau=False, ortho=True, npts=(18, 20, 22),
                cell=(6.0, 6.5, 7.0), periodic=(1, 1, 1), zeta=None, zetb=None, radius=None,
                eps=1e-10, window=None, shear=0.0):
    """A single collocation task with bounds from the cube rule.

    la / lb: (min, max). window: local grid sizes for non-periodic axes
    (ngrid), lb_grid = 0. shear: fraction of dh[0] added to dh[1] (general).
    """
    la_min, la_max = la
    lb_min, lb_max = lb
    zeta = float(rng.uniform(0.3, 2.0)) if zeta is None else zeta
    zetb = float(rng.uniform(0.3, 2.0)) if zetb is None else zetb
    npts = list(npts)
    cell = np.asarray(cell, dtype=np.float64)
    dh = np.diag(cell / np.asarray(npts))
    if shear:
        dh[1, 0] = shear * dh[0, 0]
    dh_inv = np.linalg.inv(dh)
    ngrid = list(npts) if window is None else list(window)
    ra = rng.uniform(0.0, 1.0, 3) @ (dh * np.asarray(npts)[:, None])
    rb = ra + rng.normal(0.0, 0.6, 3)
    rab = rb - ra
    if radius is None:
        radius = math.sqrt(math.log(1.0 / eps) / (zeta + zetb))
    maxco = ncoset(max(la_max, lb_max) + 1) + 2
    o1, o2 = 1, 2
    t = {
        "compute_tau": bool(tau), "use_ortho": bool(ortho),
        "la_max": la_max, "la_min": la_min, "lb_max": lb_max, "lb_min": lb_min,
        "zeta": zeta, "zetb": zetb, "rscale": float(rng.choice([1.0, 2.0])),
        "dh": dh, "dh_inv": dh_inv, "ra": ra, "rab": rab,
        "npts": npts, "ngrid": ngrid, "lb_grid": [0, 0, 0], "periodic": list(periodic),
        "lmax": max(la_max, lb_max) + 2, "radius": radius,
        "maxco": maxco + max(o1, o2), "o1": o1, "o2": o2,
    }
    if ortho:
        lbc, ubc, sb = oracle.sphere_bounds(radius, np.diag(dh))
        t.update(lb_cube=lbc, ub_cube=ubc, sphere_bounds=sb, nspheres=len(sb))
    pab = rng.uniform(-1.0, 1.0, (t["maxco"], t["maxco"]))
    return t, pab


def fixture(name):
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    return read_task(os.path.join(here, "golden", name + ".task"))


FIXTURES = ["collocate_ortho_density", "collocate_ortho_tau", "collocate_general_density",
            "collocate_general_tau"]


def fixture_plane(t):
    """Expected x=0 plane (z, y) from a fixture's grid entries."""
    n = t["ngrid"]
    ref = np.zeros((n[2], n[1], n[0]))
    for z, y, x, v in t["grid_entries"]:
        ref[z, y, x] = v
    return ref[:, :, 0]
