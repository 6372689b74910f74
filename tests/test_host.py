"""CPU-only checks of the product library: the C ABI loads and exports every
declared symbol, the host planner's index sets equal the oracle's, the
work accounting is right, and argument errors come back as status codes
before any device work."""
import os
import re
import subprocess

import numpy as np
import pytest

from helpers import FIXTURES, fixture, random_task

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "libgpw_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w[\w\s\*]*?\b(gpw_\w+)\s*\(", txt, re.M)))


def test_library_exports_header_symbols(gpw):
    syms = header_symbols()
    assert len(syms) >= 12
    out = subprocess.run(["nm", "-D", "--defined-only", gpw.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(\w+)$", out, re.M))
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    assert set(gpw.EXPORTS) == set(syms)


def test_planner_index_sets_match_oracle(gpw, oracle):
    from libgpw_amd import workloads as W
    for idx in (2, 3):
        cfg = W.config(idx, npairs=200)
        g = gpw.grid_desc(cfg["npts"], cfg["dh"])
        dh_inv = np.linalg.inv(cfg["dh"])
        p = cfg["pairs"]
        for i in range(0, 200, 7):
            geo = gpw.pair_geometry(g, p, i)
            lb, ub, sb = oracle.sphere_bounds(p["radius"][i], np.diag(cfg["dh"]))
            assert geo["lb_cube"] == lb and geo["ub_cube"] == ub
            assert np.array_equal(geo["sphere_bounds"], sb)
            f = p["zetb"][i] / (p["zeta"][i] + p["zetb"][i])
            rp = p["ra"][i] + f * p["rab"][i]
            assert geo["cube_centre"] == [int(np.floor(rp[d] * dh_inv[d, d])) for d in range(3)]


@pytest.mark.parametrize("name", ["collocate_ortho_density", "collocate_ortho_tau"])
def test_planner_reproduces_fixture_bounds(gpw, name):
    t = fixture(name)
    g = gpw.grid_desc(t["npts"], t["dh"])
    pairs = {"la_max": [t["la_max"]], "la_min": [t["la_min"]], "lb_max": [t["lb_max"]],
             "lb_min": [t["lb_min"]], "zeta": [t["zeta"]], "zetb": [t["zetb"]],
             "rscale": [t["rscale"]], "ra": [t["ra"]], "rab": [t["rab"]], "radius": [t["radius"]],
             "block_offset": [0]}
    geo = gpw.pair_geometry(g, pairs, 0, compute_tau=t["compute_tau"])
    assert geo["lb_cube"] == t["lb_cube"] and geo["ub_cube"] == t["ub_cube"]
    n = len(geo["sphere_bounds"])
    assert np.array_equal(geo["sphere_bounds"], t["sphere_bounds"][:n])
    assert not np.asarray(t["sphere_bounds"][n:]).any()


def sphere_count(sb, lb, ub):
    """points of the sphere_bounds walk (grid_base_ref_c.cc:413-474), 8 per block"""
    n, sci = 0, 0
    kgmin = sb[sci]; sci += 1
    for kg in range(kgmin, 1):
        jgmin = sb[sci]; sci += 1
        for jg in range(jgmin, 1):
            igmin = sb[sci]; sci += 1
            n += 8 * (1 - igmin)
    return n


def test_plan_stats_counts_sphere_points(gpw, oracle):
    from libgpw_amd import workloads as W
    cfg = W.config(2, npairs=300)
    g = gpw.grid_desc(cfg["npts"], cfg["dh"])
    st = gpw.plan_stats(g, cfg["pairs"])
    p = cfg["pairs"]
    total = 0
    for i in range(300):
        lb, ub, sb = oracle.sphere_bounds(p["radius"][i], np.diag(cfg["dh"]))
        total += sphere_count(sb, lb, ub)
    assert st["cube_points"] == total
    assert st["npairs"] == 300 and st["nsub"] >= 300 and st["nentries"] >= st["nsub"]
    assert st["tile_points"] == st["nentries"] * 32 * 16 * 16


def test_plan_stats_general_cell_counts_oracle_points(gpw, oracle):
    # triclinic cell: the planner's per-row quadratic bounds select the same
    # points as the oracle's general path (count of points it writes)
    from libgpw_amd import workloads as W
    cfg = W.level(W.config(5, npairs=1200), 0)
    p = cfg["pairs"]
    # pairs whose Gaussian-product prefactor does not underflow (every point
    # of the sphere then holds a non-zero value)
    mu = p["zeta"] * p["zetb"] / (p["zeta"] + p["zetb"])
    keep = np.flatnonzero(mu * (p["rab"] ** 2).sum(1) < 200.0)
    idx = keep[::3]
    sub = W.subset(cfg, idx)
    g = gpw.grid_desc(sub["npts"], sub["dh"])
    st = gpw.plan_stats(g, sub["pairs"])
    og = oracle.make_grid(sub["npts"], sub["dh"])
    total = 0
    for i in range(len(idx)):
        one = W.subset(sub, [i])
        grid = oracle.batch_collocate(og, oracle.make_pairs(one["pairs"]), one["pab"])
        total += np.count_nonzero(grid)
    assert len(idx) >= 25 and total > 0
    assert st["cube_points"] == total


def test_pair_geometry_general_cell_box(gpw):
    # lattice box of a general cell holds every lattice point of the sphere
    from libgpw_amd import workloads as W
    cfg = W.level(W.config(5, npairs=200), 1)
    g = gpw.grid_desc(cfg["npts"], cfg["dh"])
    p = cfg["pairs"]
    dh = np.asarray(cfg["dh"])
    for i in range(0, len(p["la_max"]), 5):
        geo = gpw.pair_geometry(g, p, i)
        assert geo["sphere_bounds"].size == 0
        f = p["zetb"][i] / (p["zeta"][i] + p["zetb"][i])
        rp = p["ra"][i] + f * p["rab"][i]
        gp = rp @ np.linalg.inv(dh)
        assert np.allclose(geo["roffset"], gp, rtol=0, atol=1e-12)
        lb, ub = np.array(geo["lb_cube"]), np.array(geo["ub_cube"])
        # brute force: lattice points within the radius, searched in a wide box
        c = np.floor(gp).astype(int)
        span = int(np.ceil(p["radius"][i] / np.min(np.abs(np.linalg.eigvalsh(dh @ dh.T))) ** 0.5)) + 3
        ax = [np.arange(c[d] - span, c[d] + span + 1) for d in range(3)]
        K, J, I = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        pts = np.stack([I, J, K], -1).reshape(-1, 3)
        r2 = (((pts - gp) @ dh) ** 2).sum(-1)
        inside = pts[r2 <= p["radius"][i] ** 2]
        assert len(inside) > 0
        assert (inside.min(0) >= lb).all() and (inside.max(0) <= ub).all()
        assert (ub - lb).max() <= 2 * span


def test_plan_splits_wide_cubes(gpw):
    # a cube wider than the period is cut into sub-boxes that each fit it
    from libgpw_amd import workloads as W
    cfg = W.config(2, npairs=50)
    p = cfg["pairs"]
    p["zeta"][:] = 0.05
    p["zetb"][:] = 0.05
    p["radius"][:] = W.radius_from_eps(0.1)  # ~15 bohr: cube ~ 2 periods of 15 bohr
    g = gpw.grid_desc(cfg["npts"], cfg["dh"])
    st = gpw.plan_stats(g, p)
    assert st["nsub"] >= 50 * 8
    # every grid tile is touched by every pair
    assert st["nentries"] >= 50 * st["ntiles"]


def _single(t, pab, gpw):
    grid = np.zeros((t["ngrid"][2], t["ngrid"][1], t["ngrid"][0]))
    gpw.collocate_pgf_product_rspace(t, pab, grid)


def test_dropin_argument_errors(gpw, oracle):
    rng = np.random.default_rng(0)
    t, pab = random_task(rng, oracle, la=(0, 2), lb=(0, 2))
    cases = [
        (dict(t, lmax=1), 2),                                   # lmax below la_max
        (dict(t, la_min=3), 1),                               # la_min > la_max
        (dict(t, ngrid=[t["ngrid"][0] - 1] + t["ngrid"][1:]), 1),  # periodic, ngrid != npts
        (dict(t, o1=t["maxco"]), 1),                          # block outside pab
    ]
    bad_sb = np.array(t["sphere_bounds"]).copy()
    bad_sb[0] = t["lb_cube"][2] - 1
    cases.append((dict(t, sphere_bounds=bad_sb), 4))
    cases.append((dict(t, sphere_bounds=np.array(t["sphere_bounds"][:3])), 4))
    cases.append((dict(t, use_ortho=False, radius=0.0), 1))   # general path needs a radius
    for tt, code in cases:
        with pytest.raises(gpw.GpwError) as e:
            _single(tt, pab, gpw)
        assert e.value.code == code, (e.value, code)


def test_tasklist_rejects_overlapping_blocks(gpw):
    # blocks may be shared whole (contracted shell pairs), never partially
    from libgpw_amd import workloads as W
    cfg = W.config(2, npairs=3)
    g = gpw.grid_desc(cfg["npts"], cfg["dh"])
    p = {k: np.array(v, copy=True) for k, v in cfg["pairs"].items()}
    p["la_max"][:] = p["la_min"][:] = 1
    p["lb_max"][:] = p["lb_min"][:] = 1
    p["block_offset"][:] = [0, 5, 20]          # 9-entry blocks: 0..8 and 5..13 overlap
    with pytest.raises(gpw.GpwError) as e:
        gpw.TaskList(g, p)
    assert e.value.code == 1 and "overlap" in str(e.value)
    p["block_offset"][:] = [0, 0, 20]
    p["la_max"][1] = p["la_min"][1] = 2        # shared block, different shell
    with pytest.raises(gpw.GpwError) as e:
        gpw.TaskList(g, p)
    assert e.value.code == 1 and "sharing" in str(e.value)


def test_contracted_expansion(gpw):
    # primitive pairs of a shell pair: all exponent combinations, rscale =
    # c_a c_b, one shared block, minimum-image rab
    from libgpw_amd import workloads as W
    from libgpw_amd.contracted import ShellPairs
    shells, cell = W.contracted_shells(4, nprim=(3, 2))
    sp = ShellPairs(shells, cell)
    assert sp.nprimitive == 4 * 6
    for s, sh in enumerate(shells):
        idx = np.flatnonzero(sp.shell_of == s)
        assert len(set(sp.pairs["block_offset"][idx])) == 1
        got = sorted(zip(sp.pairs["zeta"][idx], sp.pairs["zetb"][idx], sp.pairs["rscale"][idx]))
        want = sorted((a, b, ca * cb) for a, ca in zip(sh["exps_a"], sh["coefs_a"])
                      for b, cb in zip(sh["exps_b"], sh["coefs_b"]))
        assert np.allclose(got, want, rtol=1e-15, atol=0)
        rab = sp.pairs["rab"][idx[0]]
        assert np.all(np.abs(rab) <= np.diag(cell) / 2 + 1e-12)
        d = sh["rb"] - sh["ra"] - rab
        assert np.allclose(d / np.diag(cell), np.round(d / np.diag(cell)), atol=1e-12)


def test_tasklist_argument_errors(gpw):
    from libgpw_amd import workloads as W
    cfg = W.config(2, npairs=4)
    g = gpw.grid_desc(cfg["npts"], cfg["dh"])
    p = {k: np.array(v, copy=True) for k, v in cfg["pairs"].items()}
    p["la_max"][0] = 9
    with pytest.raises(gpw.GpwError) as e:
        gpw.plan_stats(g, p)
    assert e.value.code == 1
    p = {k: np.array(v, copy=True) for k, v in cfg["pairs"].items()}
    p["radius"][1] = 0.0
    with pytest.raises(gpw.GpwError):
        gpw.plan_stats(g, p)


@pytest.mark.parametrize("name", FIXTURES)
def test_taskfile_roundtrip(tmp_path, name):
    from libgpw_amd.taskfile import read_task, write_task
    t = fixture(name)
    grid = np.zeros((t["ngrid"][2], t["ngrid"][1], t["ngrid"][0]))
    for z, y, x, v in t["grid_entries"]:
        grid[z, y, x] = v
    path = tmp_path / "t.task"
    write_task(str(path), t, grid)
    u = read_task(str(path))
    for k in ("la_max", "la_min", "lb_max", "lb_min", "maxco", "o1", "o2", "lmax", "npts", "ngrid"):
        assert u[k] == t[k]
    for k in ("zeta", "zetb", "rscale", "radius"):
        assert u[k] == t[k]
    assert np.array_equal(u["pab"], t["pab"]) and np.array_equal(u["dh"], t["dh"])
    assert u["grid_entries"] == t["grid_entries"]


def test_replay_driver_cli(gpw, tmp_path):
    exe = os.path.join(os.path.dirname(gpw.LIB_PATH), "gpw_replay")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stdout
    bad = tmp_path / "bad.task"
    bad.write_text("#Grid collocate task v1\ncompute_tau x\n")
    r = subprocess.run([exe, str(bad)], capture_output=True, text=True)
    assert r.returncode == 2 and "compute_tau" in r.stderr


def test_cxx_header_compiles(gpw, tmp_path):
    # the drop-in header accepts a reference-style row-major array type
    src = tmp_path / "use.cpp"
    src.write_text(r'''
#include <vector>
#include "libgpw_amd.hpp"
template <int N> struct A { std::vector<double> d; int n[N];
  double &operator()(int i, int j) { return d[i * n[1] + j]; }
  const double &operator()(int i, int j) const { return d[i * n[1] + j]; }
  double &operator()(int i, int j, int k) { return d[(i * n[1] + j) * n[2] + k]; }
  const double &operator()(int i, int j, int k) const { return d[(i * n[1] + j) * n[2] + k]; }
  size_t size(int i) const { return n[i]; } };
int main() {
  A<2> pab{std::vector<double>(16), {4, 4}}; A<3> grid{std::vector<double>(8), {2, 2, 2}};
  double dh[3][3] = {{1,0,0},{0,1,0},{0,0,1}}, ra[3] = {0,0,0}, rab[3] = {0,0,0};
  int n[3] = {2, 2, 2}, lb[3] = {0, 0, 0}, lbc[3] = {-1,-1,-1}, ubc[3] = {2,2,2};
  bool per[3] = {true, true, true};
  std::vector<int> sb = {0};
  try {  // lmax = 0 < la_max = 1: must come back as GPW_ERR_LMAX
    grid_collocate_pgf_product_rspace(false, true, 1, 0, 0, 0, 1.0, 1.0, 1.0, dh, dh, ra, rab, n, n,
                                      lb, per, 0, 1.0, lbc, ubc, 1, sb, 4, 0, 0, pab, grid);
  } catch (const gpw::error &e) { return e.code == GPW_ERR_LMAX ? 0 : 5; }
  return 6;
}
''')
    exe = tmp_path / "use"
    inc = os.path.join(ROOT, "include")
    libdir = os.path.dirname(gpw.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-I" + inc, str(src), "-o", str(exe), "-L" + libdir,
                    "-lgpw_amd", "-Wl,-rpath," + libdir], check=True)
    assert subprocess.run([str(exe)]).returncode == 0


def test_nonzero_values_matches_oracle_mask(gpw, oracle):
    """gpw_nonzero_values (host-only) is the oracle's support mask for
    windows shorter and longer than the period, centres far outside it."""
    import ctypes as C
    rng = np.random.default_rng(3)
    for _ in range(60):
        period = int(rng.integers(1, 40))
        length = int(rng.integers(0, 90))
        center = int(rng.integers(-200, 200))
        ldm = period + int(rng.integers(0, 5))
        m = np.full(ldm, 7, dtype=np.int32)
        assert gpw.lib().gpw_nonzero_values(center, period, length,
                                            m.ctypes.data_as(C.POINTER(C.c_int)), ldm) == 0
        U = np.ones((3, 1, max(length, 1)))
        _, ref = oracle.fold_tables(U, [length] * 3, [center] * 3, [period] * 3, ldm)
        assert np.array_equal(m, ref[0]), (period, length, center)
    m = np.zeros(4, dtype=np.int32)
    assert gpw.lib().gpw_nonzero_values(0, 5, 3, m.ctypes.data_as(C.POINTER(C.c_int)), 4) == 1


def test_cpp_templates_compile_and_link(gpw, tmp_path):
    """The C++ drop-in header instantiates every template with an mdarray-like
    row-major type (operator(), size(i) in declaration order) and links
    against the library (no GPU needed: nothing is run)."""
    src = tmp_path / "use_hpp.cpp"
    src.write_text(r'''
#include <vector>
#include "libgpw_amd.hpp"
template <class T, int N> struct arr {
    std::vector<T> v; int d[N];
    template <class... I> arr(I... n) : d{int(n)...} { size_t s = 1; for (int x : d) s *= x; v.resize(s); }
    size_t size(int i) const { return d[i]; }
    template <class... I> T &operator()(I... i) { return v[off(i...)]; }
    template <class... I> const T &operator()(I... i) const { return v[off(i...)]; }
    template <class... I> size_t off(I... i) const { int ix[N] = {int(i)...}; size_t o = 0;
        for (int k = 0; k < N; k++) o = o * d[k] + ix[k]; return o; }
};
int main() {
    const int c[3] = {1, 2, 3}, p[3] = {8, 8, 8}, len[3] = {12, 12, 12}, gs[3] = {8, 8, 8}, lc[3] = {0, 0, 0};
    arr<double, 3> pa(3, 4, 12), pf(3, 3, 8), co(3, 3, 3), V(8, 8, 8), d(8, 8, 8), vt(8, 8, 8), P(3, 3, 8);
    arr<int, 2> m(3, 8);
    arr<double, 2> pab(4, 4);
    compute_folded_polynomials_all_powers(c, p, len, pa, pf);
    for (int dir = 0; dir < 3; dir++) compute_nonzero_values(c[dir], p[dir], len[dir], 0, 8, dir, m);
    collocate(co, pf, m, gs, lc, vt, d);
    integrate(pf, m, gs, lc, V, co);
    collocate_core(gs, co, P, d);
    integrate_core(gs, P, V, co);
    double dh[3][3] = {}, dhi[3][3] = {}, ra[3] = {}, rab[3] = {};
    const int npts[3] = {8, 8, 8}, lbg[3] = {0, 0, 0}, lb[3] = {-1, -1, -1}, ub[3] = {2, 2, 2};
    const bool per[3] = {true, true, true};
    std::vector<int> sb;
    grid_collocate_pgf_product_rspace(false, true, 0, 0, 0, 0, 1.0, 1.0, 1.0, dh, dhi, ra, rab, npts,
                                      npts, lbg, per, 1, 1.0, lb, ub, 0, sb, 4, 0, 0, pab, d);
    grid_integrate_pgf_product_rspace(false, true, 0, 0, 0, 0, 1.0, 1.0, 1.0, dh, dhi, ra, rab, npts,
                                      npts, lbg, per, 1, 1.0, lb, ub, 0, sb, 4, 0, 0, d, pab, pab);
    return 0;
}
''')
    exe = tmp_path / "use_hpp"
    lib_dir = os.path.dirname(gpw.LIB_PATH)
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"), str(src),
                        "-o", str(exe), "-L" + lib_dir, "-lgpw_amd", "-Wl,-rpath," + lib_dir],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
