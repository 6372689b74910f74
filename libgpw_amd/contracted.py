"""Contracted-basis front end.

A contracted shell pair -- atoms A and B, Cartesian shells la and lb,
primitive exponents and contraction coefficients on each side -- is a sum
over its primitive pairs:

    phi_a phi_b = sum_{i,j} c_a[i] c_b[j] g(zeta_i, A) g(zeta_j, B).

Every primitive pair of a shell pair becomes one task-list pair that reads the
shell pair's own coefficient block (the contracted density-matrix block) with
rscale = c_a[i] * c_b[j]. Collocate therefore needs no expanded coefficients,
and integrate returns the contracted block directly: the task list gives the
primitive pairs of one block to one workgroup of its back-transform, which
adds them into the block in order (race-free and deterministic).

This is the batching idea of the reference's drafts: a container of Gaussian
pairs filled per contracted pair (src/lib/gaussian.hpp:1-114, add_gaussian at
:103), and the per-primitive-pair set-up of an overlap object -- minimum image
of B (src/lib/Overlap.cc:3-30), exponent sums and a radius per primitive pair
(:32-117) -- with the contraction coefficients applied to the primitive
results (:260). Those drafts do not compile; this is written from their intent.
"""
from __future__ import annotations

import math

import numpy as np

from . import TaskList

EPS = 1e-10


def ncoset(l):
    return (l + 1) * (l + 2) * (l + 3) // 6


class ShellPairs:
    """Primitive expansion of contracted shell pairs.

    shells: list of dicts with ra, rb (3), la, lb (ints), exps_a, coefs_a,
    exps_b, coefs_b (1D, same length per side); cell: 3x3 rows = cell vectors
    (B is moved to its minimum image of A). block_offset of shell pair s is
    the sum of the block sizes ncoset(lb) * ncoset(la) of the shell pairs
    before it, unless the dicts give "block_offset".
    """

    def __init__(self, shells, cell, eps: float = EPS):
        cell = np.asarray(cell, dtype=np.float64)
        inv = np.linalg.inv(cell)
        cols = {k: [] for k in ("la_max", "la_min", "lb_max", "lb_min", "zeta", "zetb", "rscale",
                                "ra", "rab", "radius", "block_offset")}
        self.shell_of = []
        self.block_offset = []
        off = 0
        for s, sh in enumerate(shells):
            la, lb = int(sh["la"]), int(sh["lb"])
            bo = int(sh.get("block_offset", off))
            self.block_offset.append(bo)
            off = max(off, bo + ncoset(lb) * ncoset(la))
            ra = np.asarray(sh["ra"], dtype=np.float64)
            d = np.asarray(sh["rb"], dtype=np.float64) - ra
            frac = d @ inv
            rab = (frac - np.round(frac)) @ cell
            ea, ca = np.asarray(sh["exps_a"], float), np.asarray(sh["coefs_a"], float)
            eb, cb = np.asarray(sh["exps_b"], float), np.asarray(sh["coefs_b"], float)
            for i in range(len(ea)):
                for j in range(len(eb)):
                    zp = ea[i] + eb[j]
                    cols["la_max"].append(la); cols["la_min"].append(la)
                    cols["lb_max"].append(lb); cols["lb_min"].append(lb)
                    cols["zeta"].append(ea[i]); cols["zetb"].append(eb[j])
                    cols["rscale"].append(ca[i] * cb[j])
                    cols["ra"].append(ra); cols["rab"].append(rab)
                    cols["radius"].append(math.sqrt(math.log(1.0 / eps) / zp))
                    cols["block_offset"].append(bo)
                    self.shell_of.append(s)
        self.nblock = off
        self.nshells = len(shells)
        self.shell_of = np.asarray(self.shell_of, dtype=np.int64)
        ints = ("la_max", "la_min", "lb_max", "lb_min")
        self.pairs = {k: (np.asarray(v, dtype=np.int32) if k in ints else
                          np.asarray(v, dtype=np.int64) if k == "block_offset" else
                          np.asarray(v, dtype=np.float64)) for k, v in cols.items()}

    @property
    def nprimitive(self) -> int:
        return len(self.shell_of)

    def tasklist(self, grid_desc, compute_tau: bool = False, splits: int = 1) -> TaskList:
        return TaskList(grid_desc, self.pairs, compute_tau=compute_tau, splits=splits)

    def forces_by_shell(self, forces) -> np.ndarray:
        """Per-primitive forces (nprimitive, 6) summed per shell pair: (nshells, 2, 3)."""
        f = np.asarray(forces.cpu().numpy() if hasattr(forces, "cpu") else forces).reshape(-1, 6)
        out = np.zeros((self.nshells, 6))
        np.add.at(out, self.shell_of, f)
        return out.reshape(-1, 2, 3)
