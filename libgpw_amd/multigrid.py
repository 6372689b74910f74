"""Multigrid collocate / integrate: one task list per grid level.

CP2K keeps a hierarchy of grids of decreasing resolution and maps every
product Gaussian to one level by its exponent (the relative-cutoff rule,
``workloads.level_of``). Each level here is an independent task list on its
own grid; the pairs of all levels share one packed coefficient buffer (pab)
and one hab buffer, addressed by each pair's block_offset, so the levels
write disjoint hab blocks. Forces come back per level (pairs in level order).

``MultiGrid.capture`` records collocate + integrate (+ forces) of every level
in one HIP graph (torch.cuda.CUDAGraph); task-list calls allocate nothing and
clear memory with fill kernels, so they are legal inside capture.
"""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import TaskList, grid_desc


class MultiGrid:
    def __init__(self, cfg: dict, compute_tau: bool = False, splits: int = 1):
        """cfg: a multigrid configuration (``workloads.config(5)``): ``pairs``
        (every pair, global block offsets) and ``levels`` (npts, dh, index)."""
        self.levels = []
        p = cfg["pairs"]
        jobs = []
        for lv in cfg["levels"]:
            idx = np.asarray(lv["index"])
            if idx.size == 0:
                continue
            sub = {k: np.asarray(v)[idx] for k, v in p.items()}
            jobs.append((lv, idx, sub, grid_desc(lv["npts"], lv["dh"])))
        # The levels are planned concurrently, one thread each: the planner
        # runs in the library with the GIL released, so the levels' host
        # geometry, uploads and device work lists overlap. Each thread plans
        # on the caller's current device.
        device = torch.cuda.current_device()

        def plan(job):
            torch.cuda.set_device(device)
            return TaskList(job[3], job[2], compute_tau=compute_tau, splits=splits)

        with ThreadPoolExecutor(max_workers=max(1, len(jobs))) as ex:
            tls = list(ex.map(plan, jobs))
        for (lv, idx, _, g), tl in zip(jobs, tls):
            self.levels.append({"npts": tuple(lv["npts"]), "index": idx, "grid_desc": g, "tl": tl})
        self.npairs = int(sum(len(lv["index"]) for lv in self.levels))

    def autotune(self, max_splits: int = 0, reps: int = 3):
        return [lv["tl"].autotune(max_splits=max_splits, reps=reps) for lv in self.levels]

    def grids(self, device="cuda"):
        """Zeroed level grids (z, y, x)."""
        return [torch.zeros((n[2], n[1], n[0]), dtype=torch.float64, device=device)
                for n in (lv["npts"] for lv in self.levels)]

    def forces_buffers(self, device="cuda"):
        return [torch.zeros((len(lv["index"]), 6), dtype=torch.float64, device=device)
                for lv in self.levels]

    def collocate(self, pab, grids):
        for lv, g in zip(self.levels, grids):
            lv["tl"].collocate(pab, g)
        return grids

    def integrate(self, pots, hab, pab=None, forces=None):
        for k, (lv, v) in enumerate(zip(self.levels, pots)):
            lv["tl"].integrate(v, hab, pab=pab, forces=None if forces is None else forces[k])
        return hab

    def _level_streams(self):
        if getattr(self, "_streams", None) is None:
            self._streams = [torch.cuda.Stream() for _ in self.levels]
        return self._streams

    def step(self, pab, grids, pots, hab, forces=None, concurrent=True):
        """One collocate + integrate (+ forces) of every level.

        The levels share nothing they write (own grid, own task-list scratch,
        disjoint hab blocks, own force buffer), so with ``concurrent`` each
        level runs on a stream of its own, forked from and joined back to the
        current stream; the coarse levels, too small to fill the GPU, then run
        beside the fine one. Results are the same either way."""
        for g in grids:
            g.zero_()
        hab.zero_()
        if forces is not None:
            for f in forces:
                f.zero_()
        pab_f = pab if forces is not None else None
        if not concurrent:
            self.collocate(pab, grids)
            self.integrate(pots, hab, pab=pab_f, forces=forces)
            return
        cur = torch.cuda.current_stream()
        streams = self._level_streams()
        for k, (lv, s) in enumerate(zip(self.levels, streams)):
            s.wait_stream(cur)
            with torch.cuda.stream(s):
                lv["tl"].collocate(pab, grids[k])
                lv["tl"].integrate(pots[k], hab, pab=pab_f,
                                   forces=None if forces is None else forces[k])
        for s in streams:
            cur.wait_stream(s)

    def capture(self, pab, grids, pots, hab, forces=None):
        """HIP graph of ``step``; replaying it recomputes grids, hab and forces
        from the current contents of pab and pots. The graph is recorded on
        one stream, levels one after the other (a capture that forked to
        other streams and joined them again ended in a crash of the capture's
        end), so the eager concurrent ``step`` is the faster form."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):  # warm-up outside the graph
            self.step(pab, grids, pots, hab, forces, concurrent=False)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self.step(pab, grids, pots, hab, forces, concurrent=False)
        return graph

    def forces_by_pair(self, forces) -> np.ndarray:
        """Level force buffers back in global pair order: (npairs, 2, 3)."""
        out = np.zeros((self.npairs, 6))
        for lv, f in zip(self.levels, forces):
            out[lv["index"]] = f.cpu().numpy()
        return out.reshape(-1, 2, 3)

    def close(self):
        for lv in self.levels:
            lv["tl"].close()
