// gpw_single.cpp -- the reference's one-pair interface without a task list.
//
// grid_collocate_pgf_product_rspace (src/lib/grid_base_ref_c.h:40-67) is
// called once per pair by its callers, so each call here does only what one
// pair needs: the host geometry of the pair (cube centre and sub-boxes from
// the caller's lb_cube / ub_cube / sphere_bounds, grid_base_ref_c.cc:534-575,
// or the lattice box of the general path, :700-721), one upload of it into a
// device region, the pair transform (k_coef / k_hab) and one plain-kernel
// launch per sub-box.
//
// Three ways to call:
//   * host arrays (the reference's own signature): only the grid points of the
//     pair's cube travel. Collocate writes the cube into compact per-sub-box
//     boxes that the host adds into the grid through the periodic map;
//     integrate gathers the boxes from the grid. When the boxes hold more
//     points than the grid (cubes wider than the period), the whole grid
//     travels instead.
//   * device arrays (the _dev entry points): grid, pab, hab and forces stay on
//     the device; pab and hab are read and written in place as maxco x maxco
//     arrays (block at row o2, column o1); stream ordered.
// The device region of both is a per-thread, per-device workspace that grows
// on demand and is kept until the process ends.
//
// Built with -ffp-contract=off (host geometry rounds as written).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "gpw_host.hpp"

using namespace gpw;

namespace {

constexpr size_t ALIGN = 256;
size_t align_up(size_t n) { return (n + ALIGN - 1) & ~(ALIGN - 1); }

// validated arguments of one call
struct Task {
    gpw_grid_desc g{};
    gpw_pair_desc p{};
    int tau = 0, use_ortho = 0;
    int maxco = 0, o1 = 0, o2 = 0;
};

// The reference's preconditions as status codes (it asserts or runs out of
// bounds): l ranges, lmax (alpha sized by lmax, grid_base_ref_c.cc:946;
// coef_xyt by 2 lmax, :246-248), the pab block inside maxco, cube arguments.
gpw_status task_args(int compute_tau, int use_ortho, int la_max, int la_min, int lb_max, int lb_min,
                     double zeta, double zetb, double rscale, const double *dh, const double *dh_inv,
                     const double *ra, const double *rab, const int *npts, const int *ngrid,
                     const int *lb_grid, const int *periodic, int lmax, double radius,
                     const int *lb_cube, const int *ub_cube, int maxco, int o1, int o2, Task &t) {
    if (!dh || !dh_inv || !ra || !rab || !npts || !ngrid || !lb_grid || !periodic)
        return fail(GPW_ERR_ARG, "NULL argument");
    const int tau = compute_tau ? 1 : 0;
    if (la_min < 0 || lb_min < 0 || la_max < la_min || lb_max < lb_min)
        return fail(GPW_ERR_ARG, "bad angular momentum range");
    if (lmax < la_max + tau || lmax < lb_max + tau || 2 * lmax < la_max + lb_max + 2 * tau)
        return fail(GPW_ERR_LMAX, "lmax below the local angular momentum");
    if (o1 < 0 || o2 < 0 || o1 + ncoset(la_max) > maxco || o2 + ncoset(lb_max) > maxco)
        return fail(GPW_ERR_ARG, "pab block outside maxco");
    if (use_ortho && (!lb_cube || !ub_cube)) return fail(GPW_ERR_ARG, "ortho path needs lb_cube/ub_cube");
    if (!use_ortho && !(radius > 0.0)) return fail(GPW_ERR_ARG, "general path needs radius > 0");
    for (int i = 0; i < 3; i++) {
        t.g.npts[i] = npts[i];
        t.g.ngrid[i] = ngrid[i];
        t.g.lb_grid[i] = lb_grid[i];
        t.g.periodic[i] = periodic[i] ? 1 : 0;
        for (int j = 0; j < 3; j++) {
            t.g.dh[i][j] = dh[3 * i + j];
            t.g.dh_inv[i][j] = dh_inv[3 * i + j];
        }
    }
    gpw_status st = check_grid(&t.g);
    if (st) return st;
    t.p.la_max = la_max;
    t.p.la_min = la_min;
    t.p.lb_max = lb_max;
    t.p.lb_min = lb_min;
    t.p.zeta = zeta;
    t.p.zetb = zetb;
    t.p.rscale = rscale;
    for (int i = 0; i < 3; i++) {
        t.p.ra[i] = ra[i];
        t.p.rab[i] = rab[i];
    }
    t.p.radius = radius;
    t.p.block_offset = 0;
    if ((st = check_pair(t.p, tau))) return st;
    t.tau = tau;
    t.use_ortho = use_ortho ? 1 : 0;
    t.maxco = maxco;
    t.o1 = o1;
    t.o2 = o2;
    return GPW_OK;
}

// where the grid values of a call live
enum Mode {
    DEV,          // caller's device grid, indexed through the periodic map
    HOST_WINDOW,  // compact boxes of the sub-boxes (host adds / gathers)
    HOST_FULL,    // a device copy of the whole grid
};

// host geometry of one call and the layout of its device region
struct Plan {
    Task t;
    PairDev pd{};
    Profile prof;
    std::vector<SubDev> subs;     // sub-boxes, grid mapping
    std::vector<SubDev> bsubs;    // the same, mapped onto their compact boxes
    std::vector<long long> boff;  // offset of each box (doubles)
    long long nbox = 0;           // doubles of all boxes
    long long ngrid_pts = 0;
    TileGrid tg{};
    LatticeHost lat;  // tables (general cells that are not stepped)
    int lat_stepped = 0;
    double lat_h[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int general = 0, lp = 0, mstride = 0;
    long long ws_doubles = 0;
    Mode mode = DEV;
    // byte offsets in the device region; [0, head) is uploaded from the host
    size_t o_pair = 0, o_subs = 0, o_bsubs = 0, o_prof = 0, o_lcol = 0, o_lval = 0, o_lblk = 0,
           o_pab = 0, head = 0;
    size_t o_grid = 0, o_coef = 0, o_mbuf = 0, o_ws = 0, o_hab = 0, o_forces = 0, total = 0;
};

gpw_status make_plan(const Task &t, int nspheres, const int *sphere_bounds, const int *lb_cube,
                     const int *ub_cube, bool host_arrays, Plan &P) {
    P.t = t;
    const gpw_grid_desc &g = t.g;
    P.pd = make_pair_dev(t.p, t.tau);
    P.pd.coef_off = 0;
    P.lp = P.pd.lp;
    PairGeom geo{};
    gpw_status st;
    if (t.use_ortho) {
        // orthorhombic path on dh[i][i] as the reference takes it on the
        // caller's word (grid_base_ref_c.cc:981, :546), whatever dh holds
        if ((st = profile_from_bounds(lb_cube, ub_cube, nspheres, sphere_bounds, P.prof))) return st;
        cube_centre(g, P.pd.rp, geo.cc, geo.roff);
        for (int d = 0; d < 3; d++) {
            geo.lb[d] = lb_cube[d];
            geo.ub[d] = ub_cube[d];
        }
        geo.prof = 0;
        if ((st = make_subs(g, P.pd, 0, geo, 0, P.prof.n[1] + 1, P.subs))) return st;
    } else {
        general_geometry(g, P.pd.rp, t.p.radius, geo);
        if ((st = make_subs(g, P.pd, 0, geo, -1, 0, P.subs))) return st;
    }
    P.general = t.use_ortho ? 0 : 1;
    TileGrid &tg = P.tg;
    tg.nx = g.ngrid[0];
    tg.ny = g.ngrid[1];
    tg.nz = g.ngrid[2];
    tg.ntx = tg.nty = tg.ntz = 1;
    tg.general = P.general;
    for (int d = 0; d < 3; d++) {
        tg.dr[d] = g.dh[d][d];
        tg.period[d] = g.periodic[d] ? g.npts[d] : BIG;
        for (int e = 0; e < 3; e++) tg.dh[d][e] = g.dh[d][e];
    }
    P.ngrid_pts = (long long)g.ngrid[0] * g.ngrid[1] * g.ngrid[2];
    // compact boxes
    P.bsubs = P.subs;
    P.boff.assign(P.subs.size(), 0);
    P.nbox = 0;
    for (size_t i = 0; i < P.subs.size(); i++) {
        SubDev &b = P.bsubs[i];
        long long v = 1;
        for (int d = 0; d < 3; d++) {
            b.base[d] = 0;
            b.period[d] = BIG;
            v *= (long long)(b.ub[d] - b.lb[d] + 1);
        }
        P.boff[i] = P.nbox;
        P.nbox += v;
    }
    P.mode = !host_arrays ? DEV : P.nbox < P.ngrid_pts ? HOST_WINDOW : HOST_FULL;
    if (P.general && stepped_cell(g.dh, P.lat_h)) {
        P.lat_stepped = 1;
    } else if (P.general) {
        // lattice transforms of the two degrees the kernels read: lp (k_coef)
        // and lp + 2 (k_hab with the force extension)
        const int lps[2] = {P.lp, std::min(P.lp + 2, LPMAX)};
        if ((st = lattice_tables_some(g.dh, lps, 2, P.lat))) return st;
    }
    P.mstride = ncoset(P.lp + 2);
    const int lam = t.p.la_max + t.tau, lbm = t.p.lb_max + t.tau;
    long long ws = 0;
    if (coef_needs_workspace(lam, lbm)) ws = (long long)coef_ws_doubles(lam, lbm);
    ws = std::max(ws, (long long)hab_ws_need(t.p.la_max, t.p.lb_max, t.tau,
                                              hab_cross_entries(t.p.la_min, t.p.la_max, t.p.lb_min,
                                                                t.p.lb_max)));
    P.ws_doubles = ws;
    // device pair: in DEV mode pab / hab are the caller's maxco x maxco
    // arrays, read and written in place; otherwise a packed block
    const int na = ncoset(t.p.la_max), nb = ncoset(t.p.lb_max);
    if (P.mode == DEV) {
        P.pd.block_off = (long long)t.o2 * t.maxco + t.o1;
        P.pd.ld = t.maxco;
    } else {
        P.pd.block_off = 0;
        P.pd.ld = na;
    }
    // layout
    size_t o = 0;
    auto put = [&](size_t bytes) {
        const size_t at = o;
        o = align_up(o + bytes);
        return at;
    };
    P.o_pair = put(sizeof(PairDev));
    P.o_subs = put(P.subs.size() * sizeof(SubDev));
    P.o_bsubs = put(P.mode == HOST_WINDOW ? P.bsubs.size() * sizeof(SubDev) : 0);
    P.o_prof = put(P.prof.T.size() * sizeof(int));
    P.o_lcol = put(P.lat.col.size() * sizeof(int));
    P.o_lval = put(P.lat.val.size() * sizeof(double));
    P.o_lblk = put(P.lat.blk.size() * sizeof(LatBlk));
    P.o_pab = put(P.mode == DEV ? 0 : (size_t)na * nb * sizeof(double));
    P.head = o;
    const long long gpts = P.mode == HOST_WINDOW ? P.nbox : P.mode == HOST_FULL ? P.ngrid_pts : 0;
    P.o_grid = put((size_t)gpts * sizeof(double));
    P.o_coef = put((size_t)ncoset(P.lp) * sizeof(double));
    P.o_mbuf = put((size_t)P.mstride * sizeof(double));
    P.o_ws = put((size_t)P.ws_doubles * sizeof(double));
    P.o_hab = put(P.mode == DEV ? 0 : (size_t)na * nb * sizeof(double));
    P.o_forces = put(P.mode == DEV ? 0 : 6 * sizeof(double));
    P.total = o;
    return GPW_OK;
}

// host image of [0, head) of the region (plus, for integrate in window mode,
// the gathered boxes right behind it)
void stage_head(const Plan &P, const double *pab_host, std::vector<char> &stage) {
    stage.assign(P.head, 0);
    std::memcpy(stage.data() + P.o_pair, &P.pd, sizeof(PairDev));
    if (!P.subs.empty()) std::memcpy(stage.data() + P.o_subs, P.subs.data(), P.subs.size() * sizeof(SubDev));
    if (P.mode == HOST_WINDOW)
        std::memcpy(stage.data() + P.o_bsubs, P.bsubs.data(), P.bsubs.size() * sizeof(SubDev));
    if (!P.prof.T.empty()) std::memcpy(stage.data() + P.o_prof, P.prof.T.data(), P.prof.T.size() * sizeof(int));
    if (P.general && !P.lat_stepped) {
        if (!P.lat.col.empty()) {
            std::memcpy(stage.data() + P.o_lcol, P.lat.col.data(), P.lat.col.size() * sizeof(int));
            std::memcpy(stage.data() + P.o_lval, P.lat.val.data(), P.lat.val.size() * sizeof(double));
        }
        std::memcpy(stage.data() + P.o_lblk, P.lat.blk.data(), P.lat.blk.size() * sizeof(LatBlk));
    }
    if (P.mode != DEV) {
        // packed block ncoset(lb_max) x ncoset(la_max) of the caller's pab
        const Task &t = P.t;
        const int na = ncoset(t.p.la_max), nb = ncoset(t.p.lb_max);
        double *blk = reinterpret_cast<double *>(stage.data() + P.o_pab);
        for (int j = 0; j < nb; j++)
            for (int i = 0; i < na; i++)
                blk[(size_t)j * na + i] =
                    pab_host ? pab_host[(size_t)(t.o2 + j) * t.maxco + t.o1 + i] : 0.0;
    }
}

// Visits the grid points of every sub-box row in box order: f(box offset,
// grid offset, run length) for runs contiguous in both.
template <class F>
void for_box_rows(const Plan &P, F &&f) {
    const int nx = P.t.g.ngrid[0], ny = P.t.g.ngrid[1];
    for (size_t si = 0; si < P.subs.size(); si++) {
        const SubDev &s = P.subs[si];
        const int bx = s.ub[0] - s.lb[0] + 1, by = s.ub[1] - s.lb[1] + 1, bz = s.ub[2] - s.lb[2] + 1;
        long long bo = P.boff[si];
        for (int k = 0; k < bz; k++) {
            const int gz = s.period[2] == BIG ? s.base[2] + k : imod_h(s.base[2] + k, s.period[2]);
            for (int j = 0; j < by; j++, bo += bx) {
                const int gy = s.period[1] == BIG ? s.base[1] + j : imod_h(s.base[1] + j, s.period[1]);
                const long long row = ((long long)gz * ny + gy) * nx;
                // x run, split once where it wraps around the period
                const int x0 = s.period[0] == BIG ? s.base[0] : imod_h(s.base[0], s.period[0]);
                const int first = s.period[0] == BIG ? bx : std::min(bx, s.period[0] - x0);
                f(bo, row + x0, first);
                if (first < bx) f(bo + first, row, bx - first);
            }
        }
    }
}

LatticeTab lattice_view(const Plan &P, char *R) {
    LatticeTab t{};
    if (!P.general) return t;
    if (P.lat_stepped) {
        t.stepped = 1;
        for (int i = 0; i < 6; i++) t.h[i] = P.lat_h[i];
        return t;
    }
    t.col = reinterpret_cast<int *>(R + P.o_lcol);
    t.val = reinterpret_cast<double *>(R + P.o_lval);
    t.blk = reinterpret_cast<LatBlk *>(R + P.o_lblk);
    return t;
}

// the TileGrid a sub-box's launch indexes: the grid, or its compact box
TileGrid launch_grid(const Plan &P, size_t si) {
    TileGrid tg = P.tg;
    if (P.mode == HOST_WINDOW) {
        const SubDev &s = P.subs[si];
        tg.nx = s.ub[0] - s.lb[0] + 1;
        tg.ny = s.ub[1] - s.lb[1] + 1;
        tg.nz = s.ub[2] - s.lb[2] + 1;
    }
    return tg;
}

// device part of collocate: coef, then one plain launch per sub-box into
// `grid` (the caller's grid in DEV mode, else the region's boxes / copy)
gpw_status run_collocate(const Plan &P, char *R, const double *pab_dev, double *grid,
                         hipStream_t st) {
    const PairDev *dpair = reinterpret_cast<const PairDev *>(R + P.o_pair);
    double *coef = reinterpret_cast<double *>(R + P.o_coef);
    double *ws = P.ws_doubles ? reinterpret_cast<double *>(R + P.o_ws) : nullptr;
    const double *pab = P.mode == DEV ? pab_dev : reinterpret_cast<const double *>(R + P.o_pab);
    HIP_TRY(launch_coef(dpair, nullptr, 0, 1, P.t.tau, P.t.p.la_max, P.t.p.lb_max,
                        block_shell_entries(P.t.p.la_min, P.t.p.la_max, P.t.p.lb_min, P.t.p.lb_max), pab, coef,
                        lattice_view(P, R), ws, P.ws_doubles, 1, st));
    const SubDev *subs = reinterpret_cast<const SubDev *>(R + (P.mode == HOST_WINDOW ? P.o_bsubs : P.o_subs));
    const int *prof = reinterpret_cast<const int *>(R + P.o_prof);
    for (size_t si = 0; si < P.subs.size(); si++) {
        const SubDev &s = P.subs[si];
        double *target = P.mode == HOST_WINDOW ? grid + P.boff[si] : grid;
        HIP_TRY(launch_collocate_plain(launch_grid(P, si), subs, (int)si, s.ub[2] - s.lb[2] + 1,
                                       s.ub[1] - s.lb[1] + 1, coef, prof, target, st));
    }
    return GPW_OK;
}

// device part of integrate: moments of every sub-box, then hab / forces
gpw_status run_integrate(const Plan &P, char *R, const double *grid, const double *pab,
                         double *hab, double *forces, hipStream_t st) {
    const PairDev *dpair = reinterpret_cast<const PairDev *>(R + P.o_pair);
    double *Mbuf = reinterpret_cast<double *>(R + P.o_mbuf);
    double *ws = P.ws_doubles ? reinterpret_cast<double *>(R + P.o_ws) : nullptr;
    const int wf = forces ? 1 : 0;
    HIP_TRY(launch_fill(Mbuf, P.mstride, 0.0, st));
    const SubDev *subs = reinterpret_cast<const SubDev *>(R + (P.mode == HOST_WINDOW ? P.o_bsubs : P.o_subs));
    const int *prof = reinterpret_cast<const int *>(R + P.o_prof);
    for (size_t si = 0; si < P.subs.size(); si++) {
        const SubDev &s = P.subs[si];
        const double *src = P.mode == HOST_WINDOW ? grid + P.boff[si] : grid;
        HIP_TRY(launch_moments_plain(launch_grid(P, si), subs, (int)si, s.ub[2] - s.lb[2] + 1,
                                     s.ub[1] - s.lb[1] + 1, 2 * wf, prof, src, P.mstride, Mbuf, st));
    }
    HIP_TRY(launch_hab(dpair, nullptr, nullptr, 0, 1, P.t.tau, wf, P.t.p.la_max, P.t.p.lb_max,
                       hab_cross_entries(P.t.p.la_min, P.t.p.la_max, P.t.p.lb_min, P.t.p.lb_max), Mbuf,
                       P.mstride, pab, hab, forces, lattice_view(P, R), ws, P.ws_doubles, 1, st));
    return GPW_OK;
}

// per-thread, per-device workspace (grow only; kept until the process ends)
struct CachedRegion {
    int device = -1;
    char *ptr = nullptr;
    size_t cap = 0;
};
thread_local std::vector<CachedRegion> t_regions;
thread_local std::vector<char> t_stage;
thread_local std::vector<double> t_host;

}  // namespace

gpw_status gpw::thread_workspace(size_t need, hipStream_t st, char **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    CachedRegion *r = nullptr;
    for (auto &c : t_regions)
        if (c.device == dev) r = &c;
    if (!r) {
        t_regions.push_back(CachedRegion{dev, nullptr, 0});
        r = &t_regions.back();
    }
    if (r->cap < need) {
        if (r->ptr) {
            // earlier calls of this thread (on any stream) may still read it
            (void)st;
            HIP_TRY(hipDeviceSynchronize());
            HIP_TRY(hipFree(r->ptr));
            r->ptr = nullptr;
            r->cap = 0;
        }
        const size_t cap = std::max(need, (size_t)1 << 20);
        HIP_TRY(hipMalloc((void **)&r->ptr, cap));
        r->cap = cap;
    }
    *out = r->ptr;
    return GPW_OK;
}

namespace {

// Scratch of the allocation-free entry point: one array per GPU, loaded
// with the library's code object. Host code never passes `scratch` itself to
// a kernel or a copy (on the host it is only a stand-in symbol): it uses the
// device address that hipGetSymbolAddress returns for the current GPU.
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[GPW_NOALLOC_SCRATCH_BYTES];
std::mutex g_scratch_mu;
char *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// device address of the current GPU's scratch, looked up on the first call
// on that GPU (with it current) and kept
gpw_status scratch_region(char **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(GPW_ERR_UNSUPPORTED, "device ordinal too large");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(GPW_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<char *>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return GPW_OK;
}

}  // namespace

extern "C" {

gpw_status gpw_collocate_pgf_product_rspace(
    int compute_tau, int use_ortho, int la_max, int la_min, int lb_max, int lb_min, double zeta,
    double zetb, double rscale, const double *dh, const double *dh_inv, const double *ra,
    const double *rab, const int *npts, const int *ngrid, const int *lb_grid, const int *periodic,
    int lmax, double radius, const int *lb_cube, const int *ub_cube, int nspheres,
    const int *sphere_bounds, int maxco, int o1, int o2, const double *pab, double *grid) {
    if (!pab || !grid) return fail(GPW_ERR_ARG, "NULL argument");
    Task t;
    gpw_status st = task_args(compute_tau, use_ortho, la_max, la_min, lb_max, lb_min, zeta, zetb, rscale,
                              dh, dh_inv, ra, rab, npts, ngrid, lb_grid, periodic, lmax, radius,
                              lb_cube, ub_cube, maxco, o1, o2, t);
    if (st) return st;
    Plan P;
    if ((st = make_plan(t, nspheres, sphere_bounds, lb_cube, ub_cube, true, P))) return st;
    char *R = nullptr;
    if ((st = thread_workspace(P.total, nullptr, &R))) return st;
    stage_head(P, pab, t_stage);
    HIP_TRY(hipMemcpy(R, t_stage.data(), P.head, hipMemcpyHostToDevice));
    const long long gpts = P.mode == HOST_WINDOW ? P.nbox : P.ngrid_pts;
    double *dgrid = reinterpret_cast<double *>(R + P.o_grid);
    HIP_TRY(launch_fill(dgrid, gpts, 0.0, nullptr));
    if ((st = run_collocate(P, R, nullptr, dgrid, nullptr))) return st;
    t_host.resize((size_t)gpts);
    HIP_TRY(hipMemcpy(t_host.data(), dgrid, (size_t)gpts * sizeof(double), hipMemcpyDeviceToHost));
    // grid += the pair's values (points outside the sphere hold +0)
    if (P.mode == HOST_WINDOW) {
        const double *v = t_host.data();
        for_box_rows(P, [&](long long bo, long long go, int n) {
            for (int i = 0; i < n; i++) grid[go + i] += v[bo + i];
        });
    } else {
        for (long long i = 0; i < gpts; i++) grid[i] += t_host[(size_t)i];
    }
    return GPW_OK;
}

gpw_status gpw_integrate_pgf_product_rspace(
    int compute_tau, int use_ortho, int la_max, int la_min, int lb_max, int lb_min, double zeta,
    double zetb, double rscale, const double *dh, const double *dh_inv, const double *ra,
    const double *rab, const int *npts, const int *ngrid, const int *lb_grid, const int *periodic,
    int lmax, double radius, const int *lb_cube, const int *ub_cube, int nspheres,
    const int *sphere_bounds, int maxco, int o1, int o2, const double *grid, const double *pab,
    double *hab, double *forces) {
    if (!grid || !hab || (forces && !pab)) return fail(GPW_ERR_ARG, "NULL argument");
    Task t;
    gpw_status st = task_args(compute_tau, use_ortho, la_max, la_min, lb_max, lb_min, zeta, zetb, rscale,
                              dh, dh_inv, ra, rab, npts, ngrid, lb_grid, periodic, lmax, radius,
                              lb_cube, ub_cube, maxco, o1, o2, t);
    if (st) return st;
    Plan P;
    if ((st = make_plan(t, nspheres, sphere_bounds, lb_cube, ub_cube, true, P))) return st;
    char *R = nullptr;
    if ((st = thread_workspace(P.total, nullptr, &R))) return st;
    stage_head(P, pab, t_stage);
    if (P.mode == HOST_WINDOW) {
        // head and the gathered boxes in one upload (o_grid follows the head)
        t_stage.resize(P.o_grid + (size_t)P.nbox * sizeof(double));
        double *box = reinterpret_cast<double *>(t_stage.data() + P.o_grid);
        for_box_rows(P, [&](long long bo, long long go, int n) {
            std::memcpy(box + bo, grid + go, (size_t)n * sizeof(double));
        });
        HIP_TRY(hipMemcpy(R, t_stage.data(), t_stage.size(), hipMemcpyHostToDevice));
    } else {
        HIP_TRY(hipMemcpy(R, t_stage.data(), P.head, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(R + P.o_grid, grid, (size_t)P.ngrid_pts * sizeof(double), hipMemcpyHostToDevice));
    }
    const int na = ncoset(la_max), nb = ncoset(lb_max);
    double *dhab = reinterpret_cast<double *>(R + P.o_hab);
    double *dforces = reinterpret_cast<double *>(R + P.o_forces);
    HIP_TRY(launch_fill(dhab, (long long)na * nb, 0.0, nullptr));
    HIP_TRY(launch_fill(dforces, 6, 0.0, nullptr));
    if ((st = run_integrate(P, R, reinterpret_cast<const double *>(R + P.o_grid),
                            reinterpret_cast<const double *>(R + P.o_pab), dhab,
                            forces ? dforces : nullptr, nullptr)))
        return st;
    t_host.resize((size_t)na * nb + 6);
    HIP_TRY(hipMemcpy(t_host.data(), dhab, (size_t)na * nb * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(t_host.data() + (size_t)na * nb, dforces, 6 * sizeof(double), hipMemcpyDeviceToHost));
    for (int j = 0; j < nb; j++)
        for (int i = 0; i < na; i++) hab[(size_t)(o2 + j) * maxco + o1 + i] += t_host[(size_t)j * na + i];
    if (forces)
        for (int k = 0; k < 6; k++) forces[k] += t_host[(size_t)na * nb + k];
    return GPW_OK;
}

gpw_status gpw_collocate_pgf_product_rspace_dev(
    int compute_tau, int use_ortho, int la_max, int la_min, int lb_max, int lb_min, double zeta,
    double zetb, double rscale, const double *dh, const double *dh_inv, const double *ra,
    const double *rab, const int *npts, const int *ngrid, const int *lb_grid, const int *periodic,
    int lmax, double radius, const int *lb_cube, const int *ub_cube, int nspheres,
    const int *sphere_bounds, int maxco, int o1, int o2, const double *pab_dev, double *grid_dev,
    void *stream) {
    if (!pab_dev || !grid_dev) return fail(GPW_ERR_ARG, "NULL argument");
    Task t;
    gpw_status st = task_args(compute_tau, use_ortho, la_max, la_min, lb_max, lb_min, zeta, zetb, rscale,
                              dh, dh_inv, ra, rab, npts, ngrid, lb_grid, periodic, lmax, radius,
                              lb_cube, ub_cube, maxco, o1, o2, t);
    if (st) return st;
    Plan P;
    if ((st = make_plan(t, nspheres, sphere_bounds, lb_cube, ub_cube, false, P))) return st;
    hipStream_t hs = (hipStream_t)stream;
    char *R = nullptr;
    if ((st = thread_workspace(P.total, hs, &R))) return st;
    stage_head(P, nullptr, t_stage);
    HIP_TRY(hipMemcpyAsync(R, t_stage.data(), P.head, hipMemcpyHostToDevice, hs));
    return run_collocate(P, R, pab_dev, grid_dev, hs);
}

gpw_status gpw_integrate_pgf_product_rspace_dev(
    int compute_tau, int use_ortho, int la_max, int la_min, int lb_max, int lb_min, double zeta,
    double zetb, double rscale, const double *dh, const double *dh_inv, const double *ra,
    const double *rab, const int *npts, const int *ngrid, const int *lb_grid, const int *periodic,
    int lmax, double radius, const int *lb_cube, const int *ub_cube, int nspheres,
    const int *sphere_bounds, int maxco, int o1, int o2, const double *grid_dev, const double *pab_dev,
    double *hab_dev, double *forces_dev, void *stream) {
    if (!grid_dev || !hab_dev || (forces_dev && !pab_dev)) return fail(GPW_ERR_ARG, "NULL argument");
    Task t;
    gpw_status st = task_args(compute_tau, use_ortho, la_max, la_min, lb_max, lb_min, zeta, zetb, rscale,
                              dh, dh_inv, ra, rab, npts, ngrid, lb_grid, periodic, lmax, radius,
                              lb_cube, ub_cube, maxco, o1, o2, t);
    if (st) return st;
    Plan P;
    if ((st = make_plan(t, nspheres, sphere_bounds, lb_cube, ub_cube, false, P))) return st;
    hipStream_t hs = (hipStream_t)stream;
    char *R = nullptr;
    if ((st = thread_workspace(P.total, hs, &R))) return st;
    stage_head(P, nullptr, t_stage);
    HIP_TRY(hipMemcpyAsync(R, t_stage.data(), P.head, hipMemcpyHostToDevice, hs));
    return run_integrate(P, R, grid_dev, pab_dev, hab_dev, forces_dev, hs);
}

gpw_status gpw_integrate_pgf_product_rspace_noalloc(
    int compute_tau, int use_ortho, int la_max, int la_min, int lb_max, int lb_min, double zeta,
    double zetb, double rscale, const double *dh, const double *dh_inv, const double *ra,
    const double *rab, const int *npts, const int *ngrid, const int *lb_grid, const int *periodic,
    int lmax, double radius, const int *lb_cube, const int *ub_cube, int nspheres,
    const int *sphere_bounds, int maxco, int o1, int o2, const double *grid_dev, const double *pab_dev,
    double *hab_dev, double *forces_dev, void *stream) {
    if (!grid_dev || !hab_dev || (forces_dev && !pab_dev)) return fail(GPW_ERR_ARG, "NULL argument");
    Task t;
    gpw_status st = task_args(compute_tau, use_ortho, la_max, la_min, lb_max, lb_min, zeta, zetb, rscale,
                              dh, dh_inv, ra, rab, npts, ngrid, lb_grid, periodic, lmax, radius,
                              lb_cube, ub_cube, maxco, o1, o2, t);
    if (st) return st;
    Plan P;
    if ((st = make_plan(t, nspheres, sphere_bounds, lb_cube, ub_cube, false, P))) return st;
    if (P.total > (size_t)GPW_NOALLOC_SCRATCH_BYTES)
        return fail(GPW_ERR_ARG, "task needs " + std::to_string(P.total) +
                                     " bytes of device scratch, more than GPW_NOALLOC_SCRATCH_BYTES (" +
                                     std::to_string((size_t)GPW_NOALLOC_SCRATCH_BYTES) + ")");
    char *R = nullptr;
    if ((st = scratch_region(&R))) return st;
    hipStream_t hs = (hipStream_t)stream;
    // the whole region is rewritten or cleared before it is read: the head
    // by this upload, the moments by a fill kernel, the workspace by k_hab
    stage_head(P, nullptr, t_stage);
    HIP_TRY(hipMemcpyAsync(R, t_stage.data(), P.head, hipMemcpyHostToDevice, hs));
    return run_integrate(P, R, grid_dev, pab_dev, hab_dev, forces_dev, hs);
}

}  // extern "C"
