"""Reader and writer for the CP2K ``#Grid collocate task v1`` text format.

Format as written by the reference's task dump (src/lib/grid_base_ref_c.cc:1104-1180)
and read by its replay (src/lib/grid_base_ref_replay.cc:26-215). The dump only
ever wrote entries of the x = 0 plane (its inner loop advances ``j`` instead of
``k``, grid_base_ref_c.cc:1160), so ``grid`` entries here are sparse
(z, y, x) -> value records, and ``write_task`` writes every non-zero point.
"""
from __future__ import annotations

import numpy as np

HEADER = "#Grid collocate task v1"
FOOTER = "#THE_END"


class TaskFormatError(ValueError):
    pass


def read_task(path: str) -> dict:
    with open(path) as fh:
        lines = [ln.rstrip("\n") for ln in fh]
    if not lines or lines[0] != HEADER:
        raise TaskFormatError(f"{path}: missing '{HEADER}' header")
    pos = 1

    def take(key, n, conv):
        nonlocal pos
        if pos >= len(lines):
            raise TaskFormatError(f"{path}: unexpected end of file, wanted {key}")
        parts = lines[pos].split()
        if not parts or parts[0] != key or len(parts) != n + 1:
            raise TaskFormatError(f"{path}:{pos + 1}: expected '{key}' with {n} values")
        pos += 1
        vals = [conv(v) for v in parts[1:]]
        return vals[0] if n == 1 else vals

    t: dict = {}
    t["compute_tau"] = bool(take("compute_tau", 1, int))
    t["use_ortho"] = bool(take("use_ortho", 1, int))
    for k in ("la_max", "la_min", "lb_max", "lb_min"):
        t[k] = take(k, 1, int)
    for k in ("zeta", "zetb", "rscale"):
        t[k] = take(k, 1, float)
    for name in ("dh", "dh_inv"):
        rows = []
        for i in range(3):
            vals = take(name, 4, float)
            if int(vals[0]) != i:
                raise TaskFormatError(f"{path}: {name} row {i} out of order")
            rows.append(vals[1:])
        t[name] = np.array(rows)
    t["ra"] = np.array(take("ra", 3, float))
    t["rab"] = np.array(take("rab", 3, float))
    for k in ("npts", "ngrid", "lb_grid", "periodic"):
        t[k] = take(k, 3, int)
    t["periodic"] = [bool(v) for v in t["periodic"]]
    t["lmax"] = take("lmax", 1, int)
    t["radius"] = take("radius", 1, float)
    if t["use_ortho"]:
        t["lb_cube"] = take("lb_cube", 3, int)
        t["ub_cube"] = take("ub_cube", 3, int)
    t["nspheres"] = take("nspheres", 1, int)
    nnz = take("nspheres_nonzero", 1, int)
    sb = np.zeros(max(t["nspheres"], 0), dtype=np.int32)
    for _ in range(nnz):
        idx, val = take("sphere_bounds", 2, int)
        sb[idx] = val
    t["sphere_bounds"] = sb
    for k in ("maxco", "o1", "o2"):
        t[k] = take(k, 1, int)
    pab = np.zeros((t["maxco"], t["maxco"]))
    for i in range(t["maxco"]):
        for j in range(t["maxco"]):
            vals = take("pab", 3, float)
            if int(vals[0]) != i or int(vals[1]) != j:
                raise TaskFormatError(f"{path}: pab entry out of order")
            pab[i, j] = vals[2]
    t["pab"] = pab
    n = take("ngrid_nonzero", 1, int)
    entries = []
    for _ in range(n):
        parts = lines[pos].split()
        if parts[0] != "grid" or len(parts) != 5:
            raise TaskFormatError(f"{path}:{pos + 1}: expected grid entry")
        entries.append((int(parts[1]), int(parts[2]), int(parts[3]), float(parts[4])))
        pos += 1
    t["grid_entries"] = entries
    if pos >= len(lines) or lines[pos] != FOOTER:
        raise TaskFormatError(f"{path}: missing '{FOOTER}'")
    return t


def write_task(path: str, t: dict, grid: np.ndarray) -> None:
    """Write a task with every non-zero point of ``grid`` (z, y, x)."""
    D = 21
    with open(path, "w") as fh:
        w = fh.write
        w(HEADER + "\n")
        w(f"compute_tau {int(t['compute_tau'])}\n")
        w(f"use_ortho {int(t['use_ortho'])}\n")
        for k in ("la_max", "la_min", "lb_max", "lb_min"):
            w(f"{k} {int(t[k])}\n")
        for k in ("zeta", "zetb", "rscale"):
            w(f"{k} {float(t[k]):.{D}e}\n")
        for name in ("dh", "dh_inv"):
            for i in range(3):
                r = t[name][i]
                w(f"{name} {i} {r[0]:.{D}e} {r[1]:.{D}e} {r[2]:.{D}e}\n")
        w("ra " + " ".join(f"{v:.{D}e}" for v in t["ra"]) + "\n")
        w("rab " + " ".join(f"{v:.{D}e}" for v in t["rab"]) + "\n")
        for k in ("npts", "ngrid", "lb_grid"):
            w(f"{k} " + " ".join(str(int(v)) for v in t[k]) + "\n")
        w("periodic " + " ".join(str(int(bool(v))) for v in t["periodic"]) + "\n")
        w(f"lmax {int(t['lmax'])}\n")
        w(f"radius {float(t['radius']):.{D}e}\n")
        if t["use_ortho"]:
            w("lb_cube " + " ".join(str(int(v)) for v in t["lb_cube"]) + "\n")
            w("ub_cube " + " ".join(str(int(v)) for v in t["ub_cube"]) + "\n")
        sb = np.asarray(t.get("sphere_bounds", []), dtype=np.int64)
        w(f"nspheres {len(sb)}\n")
        nz = np.nonzero(sb)[0]
        w(f"nspheres_nonzero {len(nz)}\n")
        for i in nz:
            w(f"sphere_bounds {int(i)} {int(sb[i])}\n")
        w(f"maxco {int(t['maxco'])}\n")
        w(f"o1 {int(t['o1'])}\n")
        w(f"o2 {int(t['o2'])}\n")
        pab = np.asarray(t["pab"])
        for i in range(t["maxco"]):
            for j in range(t["maxco"]):
                w(f"pab {i} {j} {pab[i, j]:.{D}e}\n")
        idx = np.argwhere(grid != 0.0)
        w(f"ngrid_nonzero {len(idx)}\n")
        for z, y, x in idx:
            w(f"grid {z} {y} {x} {grid[z, y, x]:.{D}e}\n")
        w(FOOTER + "\n")


def ncoset(l: int) -> int:
    return 0 if l < 0 else (l + 1) * (l + 2) * (l + 3) // 6


def batch(tasks: list) -> dict:
    """Inputs of one task list for tasks that share a grid and compute_tau:
    grid (npts, ngrid, lb_grid, periodic, dh), pairs (arrays, one per task)
    and the packed pab blocks (each task's ncoset(lb_max) x ncoset(la_max)
    block at row o2, column o1 of its maxco x maxco pab). The task list
    derives each pair's cube from its radius; for orthorhombic tasks that is
    the file's lb_cube / ub_cube / sphere_bounds when they follow the CP2K
    cube rule (libgpw_amd.pair_geometry tells)."""
    if not tasks:
        raise ValueError("no tasks")
    t0 = tasks[0]
    keys = ("npts", "ngrid", "lb_grid", "periodic")
    for t in tasks[1:]:
        if any(list(t[k]) != list(t0[k]) for k in keys) or not np.array_equal(t["dh"], t0["dh"]) \
                or bool(t["compute_tau"]) != bool(t0["compute_tau"]):
            raise ValueError("tasks of one batch need one grid and one compute_tau")
    pairs = {k: [] for k in ("la_max", "la_min", "lb_max", "lb_min", "zeta", "zetb", "rscale", "ra",
                             "rab", "radius", "block_offset")}
    blocks, off = [], 0
    for t in tasks:
        na, nb = ncoset(t["la_max"]), ncoset(t["lb_max"])
        for k in ("la_max", "la_min", "lb_max", "lb_min", "zeta", "zetb", "rscale", "radius"):
            pairs[k].append(t[k])
        pairs["ra"].append(np.asarray(t["ra"]))
        pairs["rab"].append(np.asarray(t["rab"]))
        pairs["block_offset"].append(off)
        blocks.append(np.asarray(t["pab"])[t["o2"]:t["o2"] + nb, t["o1"]:t["o1"] + na].ravel())
        off += na * nb
    out = {k: np.asarray(v) for k, v in pairs.items()}
    for k in ("la_max", "la_min", "lb_max", "lb_min"):
        out[k] = out[k].astype(np.int32)
    return {"npts": list(t0["npts"]), "ngrid": list(t0["ngrid"]), "lb_grid": list(t0["lb_grid"]),
            "periodic": [int(bool(v)) for v in t0["periodic"]], "dh": np.asarray(t0["dh"]),
            "compute_tau": bool(t0["compute_tau"]), "pairs": out,
            "pab": np.concatenate(blocks), "nblock": off}


def expected_grid(t: dict) -> np.ndarray:
    """The task's listed grid values as a dense (z, y, x) array."""
    n = t["ngrid"]
    g = np.zeros((n[2], n[1], n[0]))
    for z, y, x, v in t["grid_entries"]:
        g[z, y, x] = v
    return g
