"""Differentiable collocation for PyTorch.

``collocate(pab, ra, rb, pairs, grid)`` returns the collocated density grid
(z, y, x) of all pairs. Its backward is the library's integrate path:

* dL/dpab = integrate(dL/dgrid)   (integrate is the exact adjoint of collocate)
* dL/dRa, dL/dRb = the centre derivatives of <dL/dgrid, collocate(pab; Ra, Rb)>
  computed by the same integrate launch (forces), at the grid index set of
  the forward pass.

``Collocator`` keeps one task list across calls for a fixed set of pairs
whose centres move, re-planning it in place instead of building a new one.

The grid index set (cube centres, spheres) is fixed by the forward positions;
derivatives are those of the smooth function on that set, which is what a
central finite difference measures as long as no cube centre crosses a grid
plane inside the stencil.
"""
from __future__ import annotations

import numpy as np
import torch

from . import TaskList, grid_desc as _grid_desc


class _Collocate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pab, ra, rb, meta, gdesc, shape, compute_tau, splits, owner):
        pairs = dict(meta)
        ra_h = ra.detach().double().cpu().numpy()
        rb_h = rb.detach().double().cpu().numpy()
        pairs["ra"] = ra_h
        pairs["rab"] = rb_h - ra_h
        tl = owner.tasklist(pairs) if owner is not None else \
            TaskList(gdesc, pairs, compute_tau=compute_tau, splits=splits)
        grid = torch.zeros(shape, dtype=torch.float64, device=pab.device)
        pab_c = pab.detach().contiguous()
        tl.collocate(pab_c, grid)
        ctx.tl = tl
        ctx.npairs = len(meta["la_max"])
        ctx.save_for_backward(pab_c)
        ctx.pos_dtype = ra.dtype
        ctx.pos_device = ra.device
        return grid

    @staticmethod
    def backward(ctx, grad_grid):
        (pab,) = ctx.saved_tensors
        tl = ctx.tl
        need_pos = ctx.needs_input_grad[1] or ctx.needs_input_grad[2]
        hab = torch.zeros_like(pab)
        forces = (torch.zeros((ctx.npairs, 6), dtype=torch.float64, device=pab.device)
                  if need_pos else None)
        tl.integrate(grad_grid.contiguous().double(), hab, pab=pab if need_pos else None,
                     forces=forces)
        g_ra = g_rb = None
        if need_pos:
            g_ra = forces[:, 0:3].to(device=ctx.pos_device, dtype=ctx.pos_dtype)
            g_rb = forces[:, 3:6].to(device=ctx.pos_device, dtype=ctx.pos_dtype)
        return hab, g_ra, g_rb, None, None, None, None, None, None


class Collocator:
    """Differentiable collocation of one fixed set of pairs on one grid,
    called repeatedly as the centres move (the force / geometry-optimisation
    loop). The task list is planned on the first call and re-planned in
    place on later ones (gpw_tasklist_update: same handle, autotuned splits
    kept); a call with unchanged centres reuses the plan as it is. The
    backward of a call uses the plan of that call, so call backward before
    the next forward.

        col = Collocator(pairs, npts, dh)
        rho = col(pab, ra, rb)            # differentiable in pab, ra, rb
    """

    def __init__(self, pairs: dict, npts, dh, compute_tau: bool = False, splits: int = 1,
                 autotune: bool = False):
        self.meta = {k: np.asarray(v) for k, v in pairs.items() if k not in ("ra", "rab")}
        self.gdesc = _grid_desc(npts, dh)
        self.shape = (int(npts[2]), int(npts[1]), int(npts[0]))
        self.compute_tau = compute_tau
        self.splits = splits
        self.autotune = autotune
        self._tl = None
        self._key = None
        self.replans = 0

    def tasklist(self, pairs: dict) -> TaskList:
        key = (np.asarray(pairs["ra"]).tobytes(), np.asarray(pairs["rab"]).tobytes())
        if self._tl is None:
            self._tl = TaskList(self.gdesc, pairs, compute_tau=self.compute_tau, splits=self.splits)
            if self.autotune:
                self._tl.autotune()
        elif key != self._key:
            self._tl.update(pairs)
            self.replans += 1
        self._key = key
        return self._tl

    def __call__(self, pab, ra, rb):
        return _Collocate.apply(pab, ra, rb, self.meta, self.gdesc, self.shape, self.compute_tau,
                                self.splits, self)

    def close(self):
        if self._tl is not None:
            self._tl.close()
            self._tl = None


def collocate(pab: torch.Tensor, ra: torch.Tensor, rb: torch.Tensor, pairs: dict, npts, dh,
              compute_tau: bool = False, splits: int = 1) -> torch.Tensor:
    """Differentiable batched collocation.

    pab: packed coefficient blocks (float64, device), layout of
         ``pairs["block_offset"]``; ra, rb: (npairs, 3) centres (any device);
    pairs: dict with la_max, la_min, lb_max, lb_min, zeta, zetb, rscale,
         radius, block_offset (numpy); npts: (x, y, z); dh: 3x3 step vectors.
    Returns the grid (z, y, x) on pab's device.
    """
    meta = {k: np.asarray(v) for k, v in pairs.items() if k not in ("ra", "rab")}
    gdesc = _grid_desc(npts, dh)
    shape = (int(npts[2]), int(npts[1]), int(npts[0]))
    return _Collocate.apply(pab, ra, rb, meta, gdesc, shape, compute_tau, splits, None)


def capture_step(tl: TaskList, pab, grid, pot, hab, forces=None):
    """Capture collocate + integrate of a task list in a HIP graph.

    Returns a torch.cuda.CUDAGraph; replaying it zeroes ``grid`` and ``hab``
    and recomputes both from the current contents of ``pab`` and ``pot``.
    Task-list calls allocate nothing and clear memory with fill kernels, so
    they are legal inside capture.
    """
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):  # warm up outside the graph (library and torch kernels load)
        grid.zero_(); hab.zero_()
        tl.collocate(pab, grid)
        tl.integrate(pot, hab, pab=pab if forces is not None else None, forces=forces)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        grid.zero_()
        hab.zero_()
        if forces is not None:
            forces.zero_()
        tl.collocate(pab, grid)
        tl.integrate(pot, hab, pab=pab if forces is not None else None, forces=forces)
    return g
