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
    """n random pairs; cell: three lengths (orthorhombic) or a 3x3 matrix
    whose rows are the cell vectors (centres uniform in fractional
    coordinates, rab the minimum image in fractional coordinates)."""
    cell = np.asarray(cell, dtype=np.float64)
    la_max = rng.integers(0, lmax_a + 1, n) if la is None else np.full(n, la)
    lb_max = rng.integers(0, lmax_b + 1, n) if lb is None else np.full(n, lb)
    zeta = np.exp(rng.uniform(math.log(zmin), math.log(zmax), n))
    zetb = np.exp(rng.uniform(math.log(zmin), math.log(zmax), n))
    if cell.ndim == 1:
        ra = rng.uniform(0.0, 1.0, (n, 3)) * cell
        rb = rng.uniform(0.0, 1.0, (n, 3)) * cell
        rab = _min_image(rb - ra, cell)
    else:
        fa = rng.uniform(0.0, 1.0, (n, 3))
        fb = rng.uniform(0.0, 1.0, (n, 3))
        ra = fa @ cell
        rab = _min_image(fb - fa, 1.0) @ cell
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
    5: 500,000 pairs, la,lb in 0..5, log-U[0.03, 500], triclinic cell
       a = b = 30, c = 32 bohr, gamma = 80 deg, four grid levels
       256^3 / 128^3 / 64^3 / 32^3, pairs assigned to levels by exponent
       (see ``levels``); the returned dict holds every pair, ``level(cfg, l)``
       one level's task list.
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
    elif index == 5:
        return _multigrid(rng, seed, npairs or 500000)
    else:
        raise ValueError(f"unknown configuration {index}")
    dh = np.diag(cell / np.asarray(npts, dtype=np.float64))
    return {"name": name, "index": index, "npts": npts, "dh": dh, "cell": cell, "pairs": pairs,
            "pab": pab, "nblock": nblock, "seed": seed}


# CP2K's multigrid mapping: a product Gaussian of exponent zetp goes to the
# coarsest level whose plane-wave cutoff pi^2 / (2 h^2) (Hartree, h the
# largest step length) is at least REL_CUTOFF * zetp, or to the finest level
# when none is (REL_CUTOFF = 20 Hartree = CP2K's default 40 Ry).
REL_CUTOFF = 20.0
LEVEL_NPTS = (256, 128, 64, 32)


def triclinic_cell(a, b, c, alpha_deg, beta_deg, gamma_deg) -> np.ndarray:
    """Cell vectors (rows) from lengths and angles, a along x, b in the xy plane."""
    al, be, ga = (math.radians(v) for v in (alpha_deg, beta_deg, gamma_deg))
    va = np.array([a, 0.0, 0.0])
    vb = np.array([b * math.cos(ga), b * math.sin(ga), 0.0])
    cx = c * math.cos(be)
    cy = c * (math.cos(al) - math.cos(be) * math.cos(ga)) / math.sin(ga)
    vc = np.array([cx, cy, math.sqrt(c * c - cx * cx - cy * cy)])
    return np.stack([va, vb, vc])


def level_of(zetp, level_dh) -> np.ndarray:
    """Grid level (0 = finest) of each exponent by the relative-cutoff rule."""
    zetp = np.asarray(zetp, dtype=np.float64)
    cut = [math.pi ** 2 / (2.0 * max(np.linalg.norm(r) for r in dh) ** 2) for dh in level_dh]
    lev = np.zeros(zetp.shape, dtype=np.int64)
    for l in range(1, len(level_dh)):
        lev[cut[l] >= REL_CUTOFF * zetp] = l
    return lev


def _multigrid(rng, seed, n) -> dict:
    hmat = triclinic_cell(30.0, 30.0, 32.0, 90.0, 90.0, 80.0)
    pairs, pab, nblock = make_pairs(rng, n, 5, 5, 0.03, 500.0, hmat)
    level_dh = [hmat / float(m) for m in LEVEL_NPTS]
    lev = level_of(pairs["zeta"] + pairs["zetb"], level_dh)
    levels = [{"npts": (m, m, m), "dh": level_dh[l], "index": np.flatnonzero(lev == l)}
              for l, m in enumerate(LEVEL_NPTS)]
    return {"name": f"multigrid-{n}pairs-l0..5-triclinic-256/128/64/32^3", "index": 5,
            "npts": levels[0]["npts"], "dh": level_dh[0], "cell": hmat, "pairs": pairs,
            "pab": pab, "nblock": nblock, "seed": seed, "levels": levels}


def level(cfg: dict, l: int) -> dict:
    """One grid level of a multigrid configuration as a task-list config."""
    lv = cfg["levels"][l]
    out = subset(cfg, lv["index"])
    out.update(npts=lv["npts"], dh=lv["dh"], name=cfg["name"] + f"-level{l}", level=l)
    out.pop("levels", None)
    return out


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


def contracted_shells(n: int, seed: int = 77, cell=(12.0, 12.0, 12.0), lmax: int = 2,
                      nprim=(3, 2)):
    """n random contracted shell pairs (for tests of the contracted front end):
    la, lb in 0..lmax, nprim primitives per side with exponents log-uniform in
    [0.3, 30] and contraction coefficients uniform [0.2, 1]; returns (shells,
    cell matrix)."""
    rng = np.random.default_rng(seed)
    hmat = np.diag(np.asarray(cell, dtype=np.float64))
    shells = []
    for _ in range(n):
        na, nb = int(nprim[0]), int(nprim[1])
        shells.append({
            "ra": rng.uniform(0.0, 1.0, 3) @ hmat,
            "rb": rng.uniform(0.0, 1.0, 3) @ hmat,
            "la": int(rng.integers(0, lmax + 1)), "lb": int(rng.integers(0, lmax + 1)),
            "exps_a": np.exp(rng.uniform(math.log(0.3), math.log(30.0), na)),
            "coefs_a": rng.uniform(0.2, 1.0, na),
            "exps_b": np.exp(rng.uniform(math.log(0.3), math.log(30.0), nb)),
            "coefs_b": rng.uniform(0.2, 1.0, nb),
        })
    return shells, hmat
