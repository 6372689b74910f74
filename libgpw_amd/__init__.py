"""libgpw_amd: MI355X-native collocate / integrate of primitive Gaussian pairs.

The compute path is the HIP library ``_lib/libgpw_amd.so`` (kernels in
``csrc/``), reached through its C ABI (``include/libgpw_amd.h``). This module
is the Python host side: ctypes bindings, a task-list wrapper over torch
device tensors, the single-task drop-in, and the autograd Function. There is
no CPU fallback: every compute entry point raises if the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libgpw_amd.so")

GPW_LMAX = 8
STATUS = {0: "GPW_OK", 1: "GPW_ERR_ARG", 2: "GPW_ERR_LMAX", 3: "GPW_ERR_BOUNDS",
          4: "GPW_ERR_SPHERE", 5: "GPW_ERR_HIP", 6: "GPW_ERR_ALLOC", 7: "GPW_ERR_UNSUPPORTED"}

# every symbol declared in include/libgpw_amd.h
EXPORTS = (
    "gpw_collocate_pgf_product_rspace", "gpw_integrate_pgf_product_rspace",
    "gpw_collocate_pgf_product_rspace_dev", "gpw_integrate_pgf_product_rspace_dev",
    "gpw_integrate_pgf_product_rspace_noalloc",
    "gpw_tasklist_create", "gpw_tasklist_destroy", "gpw_tasklist_update", "gpw_tasklist_stats_get",
    "gpw_tasklist_plan_stats", "gpw_pair_geometry",
    "gpw_tasklist_collocate", "gpw_tasklist_integrate",
    "gpw_tasklist_collocate_plain", "gpw_tasklist_integrate_plain",
    "gpw_tasklist_set_timing", "gpw_tasklist_timings",
    "gpw_collocate_core", "gpw_integrate_core", "gpw_collocate_core_host",
    "gpw_integrate_core_host", "gpw_tasklist_autotune", "gpw_last_error",
    "gpw_fold_tables", "gpw_collocate_folded", "gpw_integrate_folded", "gpw_fold_tables_host",
    "gpw_collocate_folded_host", "gpw_integrate_folded_host", "gpw_nonzero_values",
)
PHASES = ("coef", "collocate", "reduce", "integrate", "moments", "hab")


class GpwError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class GridDesc(C.Structure):
    _fields_ = [("npts", C.c_int * 3), ("ngrid", C.c_int * 3), ("lb_grid", C.c_int * 3),
                ("periodic", C.c_int * 3), ("dh", (C.c_double * 3) * 3),
                ("dh_inv", (C.c_double * 3) * 3)]


class PairDesc(C.Structure):
    _fields_ = [("la_max", C.c_int), ("la_min", C.c_int), ("lb_max", C.c_int), ("lb_min", C.c_int),
                ("zeta", C.c_double), ("zetb", C.c_double), ("rscale", C.c_double),
                ("ra", C.c_double * 3), ("rab", C.c_double * 3), ("radius", C.c_double),
                ("block_offset", C.c_longlong)]


class TaskListStats(C.Structure):
    _fields_ = [("npairs", C.c_longlong), ("nsub", C.c_longlong), ("ntiles", C.c_longlong),
                ("nentries", C.c_longlong), ("cube_points", C.c_longlong),
                ("tile_points", C.c_longlong), ("lp_max", C.c_int), ("tile", C.c_int * 3),
                ("splits", C.c_int), ("csplits", C.c_int * 4), ("isplits", (C.c_int * 4) * 2),
                ("flops_collocate", C.c_double), ("flops_integrate", C.c_double * 2),
                ("plan_ms", C.c_double * 4), ("csmall_pairs", C.c_longlong),
                ("csmall_points", C.c_longlong), ("collocate_tile_points", C.c_longlong)]


_lib = None


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libgpw_amd: HIP library not built ({LIB_PATH}); run __graft_entry__.build() "
                "or python libgpw_amd/build.py")
        # torch links the HIP runtime by its unversioned name and would load a
        # second copy next to ours (SONAME libamdhip64.so.7) if ours came
        # first; loading torch first makes both use the one runtime.
        import torch  # noqa: F401
        L = C.CDLL(LIB_PATH)
        dp, ip, vp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p
        single = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                  C.c_double, dp, dp, dp, dp, ip, ip, ip, ip, C.c_int, C.c_double, ip, ip, C.c_int,
                  ip, C.c_int, C.c_int, C.c_int]
        L.gpw_collocate_pgf_product_rspace.argtypes = single + [dp, dp]
        L.gpw_integrate_pgf_product_rspace.argtypes = single + [dp, dp, dp, dp]
        L.gpw_collocate_pgf_product_rspace_dev.argtypes = single + [vp, vp, vp]
        L.gpw_integrate_pgf_product_rspace_dev.argtypes = single + [vp, vp, vp, vp, vp]
        L.gpw_integrate_pgf_product_rspace_noalloc.argtypes = single + [vp, vp, vp, vp, vp]
        L.gpw_tasklist_create.argtypes = [C.POINTER(GridDesc), C.c_int, C.POINTER(PairDesc),
                                          C.c_int, ip, C.c_int, C.POINTER(vp)]
        L.gpw_tasklist_destroy.argtypes = [vp]
        L.gpw_tasklist_update.argtypes = [vp, C.c_int, C.POINTER(PairDesc)]
        L.gpw_tasklist_stats_get.argtypes = [vp, C.POINTER(TaskListStats)]
        L.gpw_tasklist_plan_stats.argtypes = [C.POINTER(GridDesc), C.c_int, C.POINTER(PairDesc),
                                              C.c_int, C.c_int, C.POINTER(TaskListStats)]
        L.gpw_pair_geometry.argtypes = [C.POINTER(GridDesc), C.POINTER(PairDesc), C.c_int, ip, dp,
                                        ip, ip, ip, C.c_int, ip]
        L.gpw_tasklist_collocate.argtypes = [vp, vp, vp, vp]
        L.gpw_tasklist_integrate.argtypes = [vp, vp, vp, vp, vp, vp]
        L.gpw_tasklist_collocate_plain.argtypes = [vp, vp, vp, vp]
        L.gpw_tasklist_integrate_plain.argtypes = [vp, vp, vp, vp, vp, vp]
        L.gpw_tasklist_set_timing.argtypes = [vp, C.c_int]
        L.gpw_tasklist_timings.argtypes = [vp, dp, C.c_int, ip]
        L.gpw_collocate_core.argtypes = [ip, C.c_int, vp, vp, C.c_int, vp, vp]
        L.gpw_integrate_core.argtypes = [ip, C.c_int, vp, C.c_int, vp, vp, vp]
        L.gpw_collocate_core_host.argtypes = [ip, C.c_int, dp, dp, C.c_int, dp]
        L.gpw_integrate_core_host.argtypes = [ip, C.c_int, dp, C.c_int, dp, dp]
        L.gpw_tasklist_autotune.argtypes = [vp, C.c_int, C.c_int, vp, dp, dp]
        L.gpw_fold_tables.argtypes = [C.c_int, ip, ip, ip, vp, C.c_int, C.c_int, vp, C.c_int, vp,
                                      C.c_int, vp]
        L.gpw_collocate_folded.argtypes = [C.c_int, vp, vp, C.c_int, vp, C.c_int, ip, ip, vp, vp]
        L.gpw_integrate_folded.argtypes = [C.c_int, vp, C.c_int, vp, C.c_int, ip, ip, vp, vp, vp]
        L.gpw_fold_tables_host.argtypes = [C.c_int, ip, ip, ip, dp, C.c_int, C.c_int, dp, C.c_int,
                                           ip, C.c_int]
        L.gpw_collocate_folded_host.argtypes = [C.c_int, dp, dp, C.c_int, ip, C.c_int, ip, ip, dp]
        L.gpw_integrate_folded_host.argtypes = [C.c_int, dp, C.c_int, ip, C.c_int, ip, ip, dp, dp]
        L.gpw_nonzero_values.argtypes = [C.c_int, C.c_int, C.c_int, ip, C.c_int]
        L.gpw_last_error.argtypes = []
        L.gpw_last_error.restype = C.c_char_p
        for n in EXPORTS:
            if n != "gpw_last_error":
                getattr(L, n).restype = C.c_int
        _lib = L
    return _lib


def _check(code):
    if code != 0:
        msg = lib().gpw_last_error()
        raise GpwError(code, msg.decode() if msg else "")


def ncoset(l: int) -> int:
    return 0 if l < 0 else (l + 1) * (l + 2) * (l + 3) // 6


def grid_desc(npts, dh, ngrid=None, lb_grid=(0, 0, 0), periodic=(1, 1, 1)) -> GridDesc:
    g = GridDesc()
    dh = np.asarray(dh, dtype=np.float64)
    inv = np.linalg.inv(dh)
    ngrid = npts if ngrid is None else ngrid
    for i in range(3):
        g.npts[i] = int(npts[i]); g.ngrid[i] = int(ngrid[i])
        g.lb_grid[i] = int(lb_grid[i]); g.periodic[i] = int(bool(periodic[i]))
        for j in range(3):
            g.dh[i][j] = float(dh[i, j]); g.dh_inv[i][j] = float(inv[i, j])
    return g


# numpy view of gpw_pair_desc (same field offsets as the ctypes structure)
_PAIR_DTYPE = np.dtype({
    "names": [f for f, _ in PairDesc._fields_],
    "formats": [np.int32] * 4 + [np.float64] * 3 + [(np.float64, 3), (np.float64, 3), np.float64,
                                                    np.int64],
    "offsets": [getattr(PairDesc, f).offset for f, _ in PairDesc._fields_],
    "itemsize": C.sizeof(PairDesc)})


def pair_array(p: dict) -> np.ndarray:
    """gpw_pair_desc array (numpy, the C layout) of a dict of per-pair
    arrays, filled column-wise; pass it to the C ABI with _pairs_ptr."""
    n = len(p["la_max"])
    v = np.empty(max(n, 1), dtype=_PAIR_DTYPE)
    if n == 0:
        return v[:0]
    for k in ("la_max", "la_min", "lb_max", "lb_min", "zeta", "zetb", "rscale", "radius",
              "block_offset"):
        v[k] = np.asarray(p[k])
    v["ra"] = np.asarray(p["ra"], dtype=np.float64).reshape(n, 3)
    v["rab"] = np.asarray(p["rab"], dtype=np.float64).reshape(n, 3)
    return v


def _pairs_ptr(arr: np.ndarray):
    return arr.ctypes.data_as(C.POINTER(PairDesc))


def _stats_dict(s: TaskListStats) -> dict:
    out = {k: getattr(s, k) for k, _ in s._fields_}
    out["tile"] = list(s.tile)
    out["csplits"] = list(s.csplits)
    out["isplits"] = [list(r) for r in s.isplits]
    out["flops_integrate"] = list(s.flops_integrate)
    out["plan_ms"] = dict(zip(("host", "upload", "device_items", "buffers"), s.plan_ms))
    return out


def plan_stats(grid: GridDesc, pairs: dict, compute_tau: bool = False, splits: int = 1) -> dict:
    """Host-only planning statistics (no device needed)."""
    arr = pair_array(pairs)
    s = TaskListStats()
    _check(lib().gpw_tasklist_plan_stats(C.byref(grid), len(arr), _pairs_ptr(arr), int(compute_tau),
                                         int(splits), C.byref(s)))
    return _stats_dict(s)


def pair_geometry(grid: GridDesc, pairs: dict, i: int, compute_tau: bool = False) -> dict:
    """Host-only geometry of pair i: cube centre, roffset, cube and sphere bounds."""
    arr = pair_array({k: np.asarray(v)[i:i + 1] for k, v in pairs.items()})
    cc, lb, ub = (C.c_int * 3)(), (C.c_int * 3)(), (C.c_int * 3)()
    roff = (C.c_double * 3)()
    cap = 1 << 18
    sb = np.zeros(cap, dtype=np.int32)
    n = C.c_int(0)
    _check(lib().gpw_pair_geometry(C.byref(grid), _pairs_ptr(arr), int(compute_tau), cc, roff, lb, ub,
                                   sb.ctypes.data_as(C.POINTER(C.c_int)), cap, C.byref(n)))
    return {"cube_centre": list(cc), "roffset": list(roff), "lb_cube": list(lb),
            "ub_cube": list(ub), "sphere_bounds": sb[:n.value].copy()}


def _ptr(t, numel=None, name="tensor", exact=False):
    """Device pointer of a contiguous float64 tensor; with numel, checks that
    it holds at least (exact: exactly) that many elements, so that a short
    buffer is an error here and not an out-of-bounds access on the device."""
    if t is None:
        return None
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise ValueError(f"{name}: need a contiguous float64 device tensor "
                         f"(got {t.dtype} on {t.device}, contiguous={t.is_contiguous()})")
    if numel is not None:
        n = t.numel()
        if (exact and n != numel) or n < numel:
            raise ValueError(f"{name}: {n} elements, need {'exactly ' if exact else 'at least '}{numel}")
    return C.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    return C.c_void_p(int(stream))


class TaskList:
    """Batched collocate / integrate of many pairs on one grid (device resident).

    Mirrors a CP2K-style task list: geometry (cube centres, sphere bounds, tile
    work lists) is built once here; collocate/integrate run stream-ordered on
    torch device tensors and allocate nothing.
    """

    def __init__(self, grid: GridDesc, pairs: dict, compute_tau: bool = False, splits: int = 1):
        self.grid = grid
        self.npairs = len(pairs["la_max"])
        self._pairs = pair_array(pairs)
        # sizes the device buffers of collocate / integrate must have
        self.ngrid_points = int(grid.ngrid[0]) * int(grid.ngrid[1]) * int(grid.ngrid[2])
        self.nblock = 0
        if self.npairs:
            la = np.asarray(pairs["la_max"], dtype=np.int64)
            lb = np.asarray(pairs["lb_max"], dtype=np.int64)
            nc = lambda l: (l + 1) * (l + 2) * (l + 3) // 6
            ends = np.asarray(pairs["block_offset"], dtype=np.int64) + nc(la) * nc(lb)
            self.nblock = int(ends.max())
        h = C.c_void_p()
        _check(lib().gpw_tasklist_create(C.byref(grid), self.npairs, _pairs_ptr(self._pairs),
                                         int(compute_tau), None, int(splits), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gpw_tasklist_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update(self, pairs: dict):
        """Re-plan in place for new pair data (same number of pairs, e.g.
        moved centres), keeping the autotuned splits."""
        arr = pair_array(pairs)
        if len(arr) != self.npairs:
            raise ValueError(f"update needs {self.npairs} pairs, got {len(arr)}")
        _check(lib().gpw_tasklist_update(self._h, self.npairs, _pairs_ptr(arr)))
        self._pairs = arr
        la = np.asarray(pairs["la_max"], dtype=np.int64)
        lb = np.asarray(pairs["lb_max"], dtype=np.int64)
        nc = lambda l: (l + 1) * (l + 2) * (l + 3) // 6
        if self.npairs:
            self.nblock = int((np.asarray(pairs["block_offset"], dtype=np.int64) + nc(la) * nc(lb)).max())
        return self

    def stats(self) -> dict:
        s = TaskListStats()
        _check(lib().gpw_tasklist_stats_get(self._h, C.byref(s)))
        return _stats_dict(s)

    def set_timing(self, enable: bool = True):
        _check(lib().gpw_tasklist_set_timing(self._h, int(enable)))

    def timings(self) -> dict:
        """{phase: (total ms, calls)} since the previous query (waits for the GPU)."""
        ms = (C.c_double * len(PHASES))()
        calls = (C.c_int * len(PHASES))()
        _check(lib().gpw_tasklist_timings(self._h, ms, len(PHASES), calls))
        return {p: (ms[i], calls[i]) for i, p in enumerate(PHASES)}

    def autotune(self, max_splits: int = 0, reps: int = 3, stream=None) -> tuple:
        """Measure and keep the best workgroups-per-tile per kernel class
        (max_splits 0: automatic range); returns (ms before, ms after) of the
        collocate + integrate kernels."""
        before, after = C.c_double(0.0), C.c_double(0.0)
        _check(lib().gpw_tasklist_autotune(self._h, int(max_splits), int(reps), _stream(stream),
                                           C.byref(before), C.byref(after)))
        return before.value, after.value

    def collocate(self, pab, grid, stream=None, plain=False):
        f = lib().gpw_tasklist_collocate_plain if plain else lib().gpw_tasklist_collocate
        _check(f(self._h, _ptr(pab, self.nblock, "pab"),
                 _ptr(grid, self.ngrid_points, "grid", exact=True), _stream(stream)))
        return grid

    def integrate(self, grid, hab, pab=None, forces=None, stream=None, plain=False):
        f = lib().gpw_tasklist_integrate_plain if plain else lib().gpw_tasklist_integrate
        if forces is not None and pab is None:
            raise ValueError("forces need pab")
        _check(f(self._h, _ptr(grid, self.ngrid_points, "grid", exact=True),
                 _ptr(pab, self.nblock, "pab"), _ptr(hab, self.nblock, "hab"),
                 _ptr(forces, 6 * self.npairs, "forces"), _stream(stream)))
        return hab


def _ci(v):
    return (C.c_int * len(v))(*[int(x) for x in v])


def _cd(v):
    v = np.ascontiguousarray(v, dtype=np.float64).ravel()
    return (C.c_double * len(v))(*v)


def _single_args(t: dict):
    sb = np.asarray(t.get("sphere_bounds", []), dtype=np.int32)
    if sb.size == 0:
        sb = np.zeros(1, dtype=np.int32)
    return [int(t["compute_tau"]), int(t["use_ortho"]), int(t["la_max"]), int(t["la_min"]),
            int(t["lb_max"]), int(t["lb_min"]), float(t["zeta"]), float(t["zetb"]),
            float(t["rscale"]), _cd(t["dh"]), _cd(t["dh_inv"]), _cd(t["ra"]), _cd(t["rab"]),
            _ci(t["npts"]), _ci(t["ngrid"]), _ci(t["lb_grid"]), _ci(t["periodic"]),
            int(t["lmax"]), float(t["radius"]), _ci(t.get("lb_cube", (0, 0, 0))),
            _ci(t.get("ub_cube", (0, 0, 0))), int(t.get("nspheres", len(sb))), _ci(sb),
            int(t["maxco"]), int(t["o1"]), int(t["o2"])]


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check_single_sizes(t: dict, grid, pab=None, hab=None):
    """The host arrays of a single task must cover what the library reads or
    writes: grid ngrid[2] x ngrid[1] x ngrid[0], pab / hab maxco x maxco."""
    ng = int(t["ngrid"][0]) * int(t["ngrid"][1]) * int(t["ngrid"][2])
    if grid.size != ng:
        raise ValueError(f"grid: {grid.size} elements, need ngrid product {ng}")
    mc = int(t["maxco"]) ** 2
    for name, a in (("pab", pab), ("hab", hab)):
        if a is not None and a.size < mc:
            raise ValueError(f"{name}: {a.size} elements, need maxco^2 = {mc}")


def collocate_pgf_product_rspace(t: dict, pab: np.ndarray, grid: np.ndarray) -> np.ndarray:
    """Drop-in for grid_collocate_pgf_product_rspace on host arrays (grid += ...)."""
    pab = np.ascontiguousarray(pab, dtype=np.float64)
    if not (isinstance(grid, np.ndarray) and grid.dtype == np.float64 and grid.flags["C_CONTIGUOUS"]):
        raise ValueError("grid: need a C-contiguous float64 numpy array (it is accumulated in place)")
    _check_single_sizes(t, grid, pab=pab)
    _check(lib().gpw_collocate_pgf_product_rspace(*_single_args(t), _dptr(pab), _dptr(grid)))
    return grid


def integrate_pgf_product_rspace(t: dict, grid: np.ndarray, pab: np.ndarray | None = None,
                                 hab: np.ndarray | None = None, forces: bool = False):
    """Adjoint of collocate_pgf_product_rspace; returns (hab, forces or None)."""
    grid = np.ascontiguousarray(grid, dtype=np.float64)
    if hab is None:
        hab = np.zeros((t["maxco"], t["maxco"]))
    if not (hab.dtype == np.float64 and hab.flags["C_CONTIGUOUS"]):
        raise ValueError("hab: need a C-contiguous float64 numpy array (it is accumulated in place)")
    if forces and pab is None:
        raise ValueError("forces need pab")
    _check_single_sizes(t, grid, pab=pab, hab=hab)
    f = np.zeros(6) if forces else None
    pab_c = None if pab is None else np.ascontiguousarray(pab, dtype=np.float64)
    _check(lib().gpw_integrate_pgf_product_rspace(
        *_single_args(t), _dptr(grid), None if pab_c is None else _dptr(pab_c), _dptr(hab),
        None if f is None else _dptr(f)))
    return hab, (None if f is None else f.reshape(2, 3))


def _single_dev_sizes(t: dict, grid, pab=None, hab=None, forces=None):
    ng = int(t["ngrid"][0]) * int(t["ngrid"][1]) * int(t["ngrid"][2])
    mc = int(t["maxco"]) ** 2
    return (_ptr(grid, ng, "grid", exact=True), _ptr(pab, mc, "pab"), _ptr(hab, mc, "hab"),
            _ptr(forces, 6, "forces"))


def collocate_pgf_product_rspace_dev(t: dict, pab, grid, stream=None):
    """The single-task collocate on device tensors (pab maxco x maxco, grid
    ngrid[2] x ngrid[1] x ngrid[0], accumulated), stream ordered."""
    g, p, _, _ = _single_dev_sizes(t, grid, pab=pab)
    _check(lib().gpw_collocate_pgf_product_rspace_dev(*_single_args(t), p, g, _stream(stream)))
    return grid


def integrate_pgf_product_rspace_dev(t: dict, grid, hab, pab=None, forces=None, stream=None):
    """The single-task integrate on device tensors: hab (maxco x maxco) and
    forces (6, optional; needs pab) accumulated."""
    if forces is not None and pab is None:
        raise ValueError("forces need pab")
    g, p, h, f = _single_dev_sizes(t, grid, pab=pab, hab=hab, forces=forces)
    _check(lib().gpw_integrate_pgf_product_rspace_dev(*_single_args(t), g, p, h, f, _stream(stream)))
    return hab


def collocate_core(length, co, P, stream=None):
    """libgpw collocate_core on device tensors: returns V (length[0..2])."""
    import torch
    L = co.shape[0]
    if tuple(co.shape) != (L, L, L) or P.dim() != 3 or P.shape[0] != 3 or P.shape[1] != L:
        raise ValueError(f"collocate_core: co must be L^3 and P 3 x L x ldp (co {tuple(co.shape)}, "
                         f"P {tuple(P.shape)})")
    V = torch.empty(tuple(int(x) for x in length), dtype=torch.float64, device=co.device)
    _check(lib().gpw_collocate_core(_ci(length), L, _ptr(co, L ** 3, "co"), _ptr(P, P.numel(), "P"),
                                    int(P.shape[2]), _ptr(V), _stream(stream)))
    return V


def integrate_core(length, P, V, stream=None):
    import torch
    L = P.shape[1]
    if P.dim() != 3 or P.shape[0] != 3:
        raise ValueError(f"integrate_core: P must be 3 x L x ldp (got {tuple(P.shape)})")
    nv = int(length[0]) * int(length[1]) * int(length[2])
    out = torch.empty((L, L, L), dtype=torch.float64, device=V.device)
    _check(lib().gpw_integrate_core(_ci(length), L, _ptr(P), int(P.shape[2]),
                                    _ptr(V, nv, "V", exact=True), _ptr(out), _stream(stream)))
    return out


# ---------------------------------------------------------------------------
# libgpw folded-table path (device tensors)

def _iptr(t, numel=None, name="mask"):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.int32 and t.is_contiguous()):
        raise ValueError(f"{name}: need a contiguous int32 device tensor")
    if numel is not None and t.numel() < numel:
        raise ValueError(f"{name}: {t.numel()} elements, need at least {numel}")
    return C.c_void_p(t.data_ptr())


def fold_tables(unfolded, length, center, period, ldp, with_mask=True, stream=None):
    """Periodic folds (3, L, ldp) and masks (3, ldp) of device tables
    unfolded (3, L, ldu) (compute_folded_polynomials_all_powers +
    compute_nonzero_values)."""
    import torch
    L, ldu = int(unfolded.shape[1]), int(unfolded.shape[2])
    folded = torch.empty((3, L, ldp), dtype=torch.float64, device=unfolded.device)
    mask = torch.empty((3, ldp), dtype=torch.int32, device=unfolded.device) if with_mask else None
    _check(lib().gpw_fold_tables(L, _ci(length), _ci(center), _ci(period),
                                 _ptr(unfolded, 3 * L * ldu, "unfolded"), L, ldu, _ptr(folded), ldp,
                                 None if mask is None else C.c_void_p(mask.data_ptr()), ldp,
                                 _stream(stream)))
    return folded, mask


def collocate_folded(co, P, mask, grid_size, lower_corner, density, stream=None):
    """collocate<T>: density (grid_size, the local window) += the masked
    contraction of co (L^3) with the folded tables P (3, L, ldp)."""
    L, ldp = int(P.shape[1]), int(P.shape[2])
    n = int(grid_size[0]) * int(grid_size[1]) * int(grid_size[2])
    _check(lib().gpw_collocate_folded(L, _ptr(co, L ** 3, "co"), _ptr(P, 3 * L * ldp, "P"), ldp,
                                      _iptr(mask, 3 * int(mask.shape[1])), int(mask.shape[1]),
                                      _ci(grid_size), _ci(lower_corner),
                                      _ptr(density, n, "density", exact=True), _stream(stream)))
    return density


def integrate_folded(P, mask, grid_size, lower_corner, V, stream=None):
    """integrate<T>: co[g][b][a] of the potential V on the masked window."""
    import torch
    L, ldp = int(P.shape[1]), int(P.shape[2])
    n = int(grid_size[0]) * int(grid_size[1]) * int(grid_size[2])
    co = torch.empty((L, L, L), dtype=torch.float64, device=P.device)
    _check(lib().gpw_integrate_folded(L, _ptr(P, 3 * L * ldp, "P"), ldp,
                                      _iptr(mask, 3 * int(mask.shape[1])), int(mask.shape[1]),
                                      _ci(grid_size), _ci(lower_corner), _ptr(V, n, "V", exact=True),
                                      _ptr(co), _stream(stream)))
    return co
