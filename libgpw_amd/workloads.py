"""Seeded synthetic workloads (pairs, coefficient blocks, potential grids).

Each generator is deterministic across platforms (numpy PCG64, seed = 1000 +
configuration index) and returns a dict:
    name, npts (x, y, z), dh (3x3 rows = step vectors), pairs (arrays),
    pab (flat float64, one packed block per pair), nblock, seed.
The cube radius is the minimal form r = sqrt(ln(1/eps) / (zeta + zetb)) with
eps = 1e-10, the same value for the CPU oracle and the GPU kernels; the grid
cube and sphere bounds follow from it by the CP2K cube rule.
"""
from __future__ import annotations

import math

import numpy as np

EPS = 1e-10


def ncoset(l):
    return (l + 1) * (l + 2) * (l + 3) // 6


def radius_from_eps(zetp, eps=EPS):
    return np.sqrt(math.log(1.0 / eps) / np.asarray(zetp, dtype=np.float64))


def _min_image(d, cell):
    return d - cell * np.round(d / cell)


def _blocks(la_max, lb_max):
    sizes = np.array([ncoset(int(b)) * ncoset(int(a)) for a, b in zip(la_max, lb_max)], dtype=np.int64)
    off = np.zeros(len(sizes), dtype=np.int64)
    if len(sizes) > 1:
        off[1:] = np.cumsum(sizes)[:-1]
    return off, int(sizes.sum())


def make_pairs(rng, n, lmax_a, lmax_b, zmin, zmax, cell, coef="uniform", la=None, lb=None):
    cell = np.asarray(cell, dtype=np.float64)
    la_max = rng.integers(0, lmax_a + 1, n) if la is None else np.full(n, la)
    lb_max = rng.integers(0, lmax_b + 1, n) if lb is None else np.full(n, lb)
    zeta = np.exp(rng.uniform(math.log(zmin), math.log(zmax), n))
    zetb = np.exp(rng.uniform(math.log(zmin), math.log(zmax), n))
    ra = rng.uniform(0.0, 1.0, (n, 3)) * cell
    rb = rng.uniform(0.0, 1.0, (n, 3)) * cell
    rab = _min_image(rb - ra, cell)
    off, nblock = _blocks(la_max, lb_max)
    if coef == "uniform":
        pab = rng.uniform(-1.0, 1.0, nblock)
    else:
        pab = rng.standard_normal(nblock)
    pairs = {
        "la_max": la_max.astype(np.int32), "la_min": la_max.astype(np.int32),
        "lb_max": lb_max.astype(np.int32), "lb_min": lb_max.astype(np.int32),
        "zeta": zeta, "zetb": zetb, "rscale": np.ones(n), "ra": ra, "rab": rab,
        "radius": radius_from_eps(zeta + zetb), "block_offset": off,
    }
    return pairs, pab, nblock


def config(index: int, npairs: int | None = None) -> dict:
    """Workloads of the benchmark suite (sizes as quoted in their names).

    1: one pair, la=lb=1, zeta 0.5 / 1.2, 10 bohr cubic cell, 48^3 grid
    2: 20,000 pairs, la,lb in 0..3, exponents log-U[0.1, 50], 15 bohr, 128^3
    3: 4,096 pairs, la,lb in 0..2, log-U[0.2, 20], 20x20x24 bohr, 96x96x112
    4: 100,000 pairs, la,lb in 0..4, log-U[0.05, 100], 25 bohr, 200^3, N(0,1)
    npairs overrides the pair count (for reduced test sizes).
    """
    seed = 1000 + index
    rng = np.random.default_rng(seed)
    if index == 1:
        cell = np.array([10.0, 10.0, 10.0])
        npts = (48, 48, 48)
        pairs, pab, nblock = make_pairs(rng, 1, 1, 1, 1.0, 1.0, cell, la=1, lb=1)
        pairs["zeta"][:] = 0.5
        pairs["zetb"][:] = 1.2
        pairs["radius"][:] = radius_from_eps(1.7)
        name = "plain-1pair-la1lb1-48^3"
    elif index == 2:
        cell = np.array([15.0, 15.0, 15.0])
        npts = (128, 128, 128)
        pairs, pab, nblock = make_pairs(rng, npairs or 20000, 3, 3, 0.1, 50.0, cell)
        name = f"batched-uneven-{len(pairs['la_max'])}pairs-l0..3-128^3"
    elif index == 3:
        cell = np.array([20.0, 20.0, 24.0])
        npts = (96, 96, 112)
        pairs, pab, nblock = make_pairs(rng, npairs or 4096, 2, 2, 0.2, 20.0, cell)
        name = f"differentiation-{len(pairs['la_max'])}pairs-l0..2-96x96x112"
    elif index == 4:
        cell = np.array([25.0, 25.0, 25.0])
        npts = (200, 200, 200)
        pairs, pab, nblock = make_pairs(rng, npairs or 100000, 4, 4, 0.05, 100.0, cell, coef="normal")
        name = f"mixed-{len(pairs['la_max'])}pairs-l0..4-200^3"
    else:
        raise ValueError(f"unknown configuration {index}")
    dh = np.diag(cell / np.asarray(npts, dtype=np.float64))
    return {"name": name, "index": index, "npts": npts, "dh": dh, "cell": cell, "pairs": pairs,
            "pab": pab, "nblock": nblock, "seed": seed}


def potential(cfg: dict, seed_offset: int = 7) -> np.ndarray:
    """Uniform[-1, 1] potential grid (z, y, x) for the integrate step."""
    rng = np.random.default_rng(cfg["seed"] + seed_offset)
    n = cfg["npts"]
    return rng.uniform(-1.0, 1.0, (n[2], n[1], n[0]))


def subset(cfg: dict, idx) -> dict:
    """A configuration restricted to the pairs idx (blocks re-packed)."""
    idx = np.asarray(idx)
    p = cfg["pairs"]
    q = {k: np.asarray(v)[idx].copy() for k, v in p.items()}
    off, nblock = _blocks(q["la_max"], q["lb_max"])
    pab = np.zeros(nblock)
    for new_i, old_i in enumerate(idx):
        sz = ncoset(int(p["lb_max"][old_i])) * ncoset(int(p["la_max"][old_i]))
        pab[off[new_i]:off[new_i] + sz] = cfg["pab"][p["block_offset"][old_i]:p["block_offset"][old_i] + sz]
    q["block_offset"] = off
    out = dict(cfg)
    out.update(pairs=q, pab=pab, nblock=nblock, name=cfg["name"] + f"-subset{len(idx)}")
    return out
