// gpw_plan.cpp -- host side of the library: per-pair geometry, sub-box split,
// tile work lists, device buffers, and the C ABI of include/libgpw_amd.h.
//
// Geometry is computed here in FP64 with contraction disabled (-ffp-contract=off)
// so that cube centres and sphere bounds are bit-identical to a straightforward
// CPU evaluation of the reference formulas:
//   pair set-up ....... src/lib/grid_base_ref_c.cc:878-886
//   cube centre ....... src/lib/grid_base_ref_c.cc:534-547
//   periodic map ...... src/lib/grid_base_ref_c.cc:318-351
//   sphere bounds ..... consumed as at src/lib/grid_base_ref_c.cc:413-470; for
//                       task lists they are generated from the pair radius by
//                       the rule of CP2K's cube_utils (radius rounded up to a
//                       multiple of the smallest spacing), which reproduces the
//                       two orthorhombic fixtures of the reference exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "gpw_host.hpp"


using namespace gpw;

static thread_local std::string g_last_error;

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// f(begin, end) over [0, n) in contiguous chunks on up to 16 host threads
// (one chunk below 4,096 items); chunks are numbered in order, so results
// gathered per chunk concatenate in item order
template <class F>
static void parallel_chunks(size_t n, F &&f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    nt = (unsigned)std::min<size_t>(nt, std::max<size_t>(1, n / 2048));
    if (nt <= 1) {
        f(0, n, 0);
        return;
    }
    const size_t chunk = (n + nt - 1) / nt;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) {
        const size_t b = t * chunk, e = std::min(n, b + chunk);
        if (b < e) th.emplace_back([&f, b, e, t] { f(b, e, t); });
    }
    for (auto &x : th) x.join();
}

gpw_status gpw::fail(gpw_status code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int gpw::imod_h(int a, int m) { return (a % m + m) % m; }

// ---------------------------------------------------------------------------
// sphere profiles: T[kk][jj] = -igmin (>= 0) or -1 for an empty row

static int sphere_extent(double base, double h, double R2) {
    int n = 0;
    for (;;) {
        const double x = (double)(n + 1) * h;
        if (base + x * x <= R2)
            n++;
        else
            return n;
    }
}

// sphere_extent(base, h, R2) when it is known to be at most `upper`
static int sphere_extent_below(double base, double h, double R2, int upper) {
    int n = upper;
    while (n > 0) {
        const double x = (double)n * h;
        if (base + x * x <= R2) return n;
        n--;
    }
    return 0;
}

// Profile for a radius on an orthorhombic grid with spacings dr (the rule
// that regenerates the reference's fixture sphere_bounds).
void gpw::profile_from_radius(double radius, const double dr[3], Profile &P, int lb[3],
                                int ub[3]) {
    double drmin = std::min(dr[0], std::min(dr[1], dr[2]));
    int imr = (int)std::ceil(radius / drmin);
    if (imr < 1) imr = 1;
    const double R = imr * drmin;
    const double R2 = R * R;
    P.R2 = R2;
    for (int d = 0; d < 3; d++) {
        const int n = sphere_extent(0.0, dr[d], R2);
        P.n[d] = n;
        lb[d] = -n;
        ub[d] = 1 + n;
    }
    P.T.assign((size_t)(P.n[2] + 1) * (P.n[1] + 1), -1);
    // the predicate base + (n h)^2 <= R2 is monotone in n and in base, so
    // each extent is the previous row's (or plane's) searched downwards
    int jn = P.n[1];
    for (int kk = 0; kk <= P.n[2]; kk++) {
        const double zk = -kk * dr[2];
        const double base = zk * zk;
        jn = sphere_extent_below(base, dr[1], R2, jn);
        int ext = P.n[0];
        for (int jj = 0; jj <= jn; jj++) {
            const double yj = -jj * dr[1];
            const double b2 = base + yj * yj;
            ext = sphere_extent_below(b2, dr[0], R2, ext);
            P.T[(size_t)kk * (P.n[1] + 1) + jj] = ext;
        }
    }
}

// Profile from a caller's sphere_bounds (grid_base_ref_c.cc:413-470 walk).
gpw_status gpw::profile_from_bounds(const int lb[3], const int ub[3], int nspheres,
                                      const int *sb, Profile &P) {
    if (nspheres < 1 || !sb) return fail(GPW_ERR_SPHERE, "sphere_bounds empty");
    for (int d = 0; d < 3; d++) P.n[d] = std::max(std::max(-lb[d], ub[d] - 1), 0);
    P.T.assign((size_t)(P.n[2] + 1) * (P.n[1] + 1), -1);
    int sci = 0;
    const int kgmin = sb[sci++];
    if (kgmin > 0 || kgmin < lb[2] || 1 - kgmin > ub[2])
        return fail(GPW_ERR_SPHERE, "sphere_bounds: kgmin outside the cube");
    for (int kg = kgmin; kg <= 0; kg++) {
        if (sci >= nspheres) return fail(GPW_ERR_SPHERE, "sphere_bounds too short");
        const int jgmin = sb[sci++];
        if (jgmin > 0 || jgmin < lb[1] || 1 - jgmin > ub[1])
            return fail(GPW_ERR_SPHERE, "sphere_bounds: jgmin outside the cube");
        for (int jg = jgmin; jg <= 0; jg++) {
            if (sci >= nspheres) return fail(GPW_ERR_SPHERE, "sphere_bounds too short");
            const int igmin = sb[sci++];
            if (igmin > 0 || igmin < lb[0] || 1 - igmin > ub[0])
                return fail(GPW_ERR_SPHERE, "sphere_bounds: igmin outside the cube");
            P.T[(size_t)(-kg) * (P.n[1] + 1) + (-jg)] = -igmin;
        }
    }
    return GPW_OK;
}

// ---------------------------------------------------------------------------

// work-item classes by polynomial length L = lp + 1 (one kernel instantiation
// each, so registers are sized for the class, not for the largest L)

enum { PH_COEF, PH_CTILES, PH_CREDUCE, PH_ITILES, PH_ISUM, PH_HAB, NPH };
constexpr int NSCLS = 4;
// tile-list set of integrate without (wf = 0) / with (wf = 1) forces
static inline int iset(int wf) { return wf ? 1 : 2; }
constexpr int NEVENT = 2048;

struct gpw_tasklist {
    gpw_grid_desc grid{};
    int compute_tau = 0;
    int npairs = 0;
    int lp_max = 0;
    int la_max = 0, lb_max = 0;  // largest given la_max / lb_max
    TileGrid tg{};
    int ntiles = 0, splits = 1;
    // workgroups per tile for each kernel class: collocate, integrate
    // (without / with the force extension); set from `splits`, then by autotune
    int csplit[NCLS] = {1, 1, 1, 1};
    int isplit[2][NCLS] = {{1, 1, 1, 1}, {1, 1, 1, 1}};
    int cpart_cap = 0;             // partial grids allocated in d_cpartial
    long long nentries = 0;
    long long cube_points = 0, tile_points = 0;
    double flops_collocate = 0.0, flops_integrate[2] = {0.0, 0.0};
    double plan_ms[4] = {0.0, 0.0, 0.0, 0.0};  // see gpw_tasklist_stats
    // tile lists: set 0 = collocate (every item), set 1 = integrate with the
    // force extension (lp_add 2), set 2 = integrate without; the integrate
    // sets leave out the pairs of the small-sphere kernel (small_*)
    long long class_entries[NSET][NCLS] = {};
    long long list_off[NSET][NCLS] = {};    // segment of the class in the tile lists
    // sub-box ids of the pairs integrated by k_moments_small, per [force
    // extension][class of L]: d_small[small_off[w][c] .. + small_cnt[w][c])
    long long small_off[2][NSCLS] = {}, small_cnt[2][NSCLS] = {};
    // sub-box ids of the pairs collocated by k_collocate_small, per class of
    // L: d_csmall[csmall_off[c] .. + csmall_cnt[c]); their boxes in d_sval at
    // d_box_off[sub] (box-major doubles, sval_total in all)
    long long csmall_off[NSCLS] = {}, csmall_cnt[NSCLS] = {};
    long long sval_total = 0;

    std::vector<PairDev> pairs_h;
    std::vector<SubDev> subs_h;
    std::vector<int> pair_sub_start;  // CSR pair -> subs (subs are contiguous per pair)

    PairDev *d_pairs = nullptr;
    SubDev *d_subs = nullptr;
    int *d_prof = nullptr;
    int *d_tile_start = nullptr;  // [NSET][NCLS][ntiles + 1]
    int *d_small = nullptr;
    unsigned char *d_small_mask = nullptr;  // [3][npairs]: integrated (w/o, with forces) / collocated per pair
    int *d_csmall = nullptr;
    long long *d_box_off = nullptr;          // per sub-box: its small-collocate box in d_sval (-1: none)
    double *d_sval = nullptr;
    EntryRec *d_recs = nullptr;   // item records in tile-list order, per (lp_add, class)
    int *d_pe_start = nullptr;
    long long *d_ent_off0 = nullptr, *d_ent_off2 = nullptr;
    double *d_coef = nullptr;
    double *d_partial = nullptr;   // integrate partial moments
    double *d_cpartial = nullptr;  // collocate partial grids (splits > 1)
    double *d_Mbuf = nullptr;
    int mstride = 0;
    // general cells: lattice transforms (CSR) per lp and direction
    int *d_mt_col = nullptr;
    double *d_mt_val = nullptr;
    LatBlk *d_mt_blk = nullptr;
    int lat_stepped = 0;
    double lat_h[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    LatticeTab lat() const {
        LatticeTab t{d_mt_col, d_mt_val, d_mt_blk};
        t.stepped = lat_stepped;
        for (int i = 0; i < 6; i++) t.h[i] = lat_h[i];
        return t;
    }
    long long coef_total = 0;        // doubles of d_coef
    // pairs grouped by block size for the per-pair transform kernels (one
    // launch per group, LDS sized for the group): d_order[start .. start+count)
    // within a group, pairs sharing a coefficient block (contracted shell
    // pairs: one block, one primitive pair per exponent pair) are adjacent and
    // form one segment, walked by one workgroup of k_hab
    struct PairGroup {
        int start, count, la_max, lb_max, seg0, nseg, nab;
        int nhx;  // most k_hab cross entries of a pair (hab_cross_entries)
    };
    int nab_max = 0;  // most block-shell entries of a pair (block_shell_entries)
    int nhx_max = 0;  // most k_hab cross entries of a pair (hab_cross_entries)
    std::vector<PairGroup> groups;
    int *d_order = nullptr;
    int *d_segs = nullptr;
    double *d_hab_ws = nullptr;      // k_coef / k_hab workspace for blocks too large for LDS
    long long hab_ws_stride = 0;
    int hab_ws_blocks = 0;

    // per-phase device timing (events recorded on the launch stream)
    bool timing = false;
    std::vector<hipEvent_t> events;
    std::vector<int> event_phase;
    int events_used = 0;

    // frees the planned buffers (not the collocate partial grids or the
    // timing events, which a re-plan of the same grid keeps)
    void release_plan() {
        void **ptrs[] = {(void **)&d_pairs, (void **)&d_subs, (void **)&d_prof, (void **)&d_tile_start,
                         (void **)&d_small, (void **)&d_small_mask, (void **)&d_csmall, (void **)&d_box_off,
                         (void **)&d_sval, (void **)&d_recs, (void **)&d_pe_start,
                         (void **)&d_ent_off0, (void **)&d_ent_off2, (void **)&d_coef, (void **)&d_partial,
                         (void **)&d_Mbuf, (void **)&d_mt_col, (void **)&d_mt_val,
                         (void **)&d_mt_blk, (void **)&d_hab_ws, (void **)&d_order, (void **)&d_segs};
        for (void **p : ptrs) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
        }
    }
    ~gpw_tasklist() {
        release_plan();
        if (d_cpartial) (void)hipFree(d_cpartial);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
    }
};

// brackets one phase with events when timing is on
struct PhaseTimer {
    gpw_tasklist *tl;
    hipStream_t st;
    int slot = -1;
    PhaseTimer(gpw_tasklist *t, hipStream_t s, int phase) : tl(t), st(s) {
        if (!tl->timing || tl->events_used + 2 > (int)tl->events.size()) return;
        slot = tl->events_used;
        tl->events_used += 2;
        tl->event_phase[slot / 2] = phase;
        (void)hipEventRecord(tl->events[slot], st);
    }
    ~PhaseTimer() {
        if (slot >= 0) (void)hipEventRecord(tl->events[slot + 1], st);
    }
};

template <class T>
static gpw_status upload(T **dst, const std::vector<T> &src) {
    const size_t n = std::max<size_t>(src.size(), 1);
    HIP_TRY(hipMalloc((void **)dst, n * sizeof(T)));
    if (!src.empty()) HIP_TRY(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return GPW_OK;
}

bool gpw::is_ortho(const gpw_grid_desc &g) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            if (i != j && g.dh[i][j] != 0.0) return false;
    return true;
}

gpw_status gpw::check_pair(const gpw_pair_desc &p, int tau) {
    if (p.la_min < 0 || p.lb_min < 0 || p.la_max < p.la_min || p.lb_max < p.lb_min)
        return fail(GPW_ERR_ARG, "bad angular momentum range");
    // +1 headroom for the force extension
    if (p.la_max + tau + 1 > LMAX || p.lb_max + tau + 1 > LMAX)
        return fail(GPW_ERR_ARG, "angular momentum above GPW_LMAX");
    if (!(p.zeta > 0.0) || !(p.zetb > 0.0)) return fail(GPW_ERR_ARG, "exponents must be > 0");
    return GPW_OK;
}

// FP64 pair set-up exactly as grid_collocate_internal (grid_base_ref_c.cc:878-886)
PairDev gpw::make_pair_dev(const gpw_pair_desc &p, int tau) {
    PairDev d{};
    d.zeta = p.zeta;
    d.zetb = p.zetb;
    d.zetp = p.zeta + p.zetb;
    const double f = p.zetb / d.zetp;
    const double rab2 = p.rab[0] * p.rab[0] + p.rab[1] * p.rab[1] + p.rab[2] * p.rab[2];
    d.prefactor = p.rscale * std::exp(-p.zeta * f * rab2);
    for (int i = 0; i < 3; i++) {
        d.ra[i] = p.ra[i];
        d.rp[i] = p.ra[i] + f * p.rab[i];
        d.rb[i] = p.ra[i] + p.rab[i];
    }
    d.la_max = p.la_max;
    d.la_min = p.la_min;
    d.lb_max = p.lb_max;
    d.lb_min = p.lb_min;
    d.lp = p.la_max + p.lb_max + 2 * tau;
    d.block_off = p.block_offset;
    d.ld = ncoset(p.la_max);
    return d;
}

// cube centre and offset (grid_base_ref_c.cc:534-547)
void gpw::cube_centre(const gpw_grid_desc &g, const double rp[3], int cc[3], double roff[3]) {
    for (int i = 0; i < 3; i++) {
        double v = 0.0;
        for (int j = 0; j < 3; j++) v += g.dh_inv[j][i] * rp[j];
        cc[i] = (int)std::floor(v);
        roff[i] = rp[i] - ((double)cc[i]) * g.dh[i][i];
    }
}

// General cell (grid_base_ref_c.cc:636-644, 700-721): lattice centre gp and
// the lattice box around the sphere of the pair's radius, from the 27 points
// rp + {-1,0,1}^3 * radius.
void gpw::general_geometry(const gpw_grid_desc &g, const double rp[3], double radius,
                             PairGeom &geo) {
    for (int i = 0; i < 3; i++) {
        geo.cc[i] = 0;
        double v = 0.0;
        for (int j = 0; j < 3; j++) v += g.dh_inv[j][i] * rp[j];
        geo.roff[i] = v;
        geo.lb[i] = 2147483647;
        geo.ub[i] = -2147483647 - 1;
    }
    for (int i = -1; i <= 1; i++)
        for (int j = -1; j <= 1; j++)
            for (int k = -1; k <= 1; k++) {
                const double x = rp[0] + i * radius;
                const double y = rp[1] + j * radius;
                const double z = rp[2] + k * radius;
                for (int d = 0; d < 3; d++) {
                    const double r = g.dh_inv[0][d] * x + g.dh_inv[1][d] * y + g.dh_inv[2][d] * z;
                    geo.lb[d] = std::min(geo.lb[d], (int)std::floor(r));
                    geo.ub[d] = std::max(geo.ub[d], (int)std::ceil(r));
                }
            }
    geo.prof = -1;
    geo.R2 = radius * radius;
}

// i-range of the lattice row (j, k) inside the sphere (grid_base_ref_c.cc:790-811)
static bool general_row_host(const gpw_grid_desc &g, double gp0, double dj, double dk, double R2,
                             int &lo, int &hi) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (int i = 0; i < 3; i++) {
        const double pointk = g.dh[2][i] * dk;
        const double pointj = pointk + g.dh[1][i] * dj;
        const double v = pointj - gp0 * g.dh[0][i];
        a += g.dh[0][i] * g.dh[0][i];
        b += 2.0 * v * g.dh[0][i];
        c += v * v;
    }
    double d = b * b - 4 * a * (c - R2);
    if (d < 0.0) return false;
    d = std::sqrt(d);
    lo = (int)std::ceil((-b - d) / (2.0 * a));
    hi = (int)std::floor((-b + d) / (2.0 * a));
    return true;
}

// Split the pair's cube into sub-boxes that fit the period, append SubDevs.
gpw_status gpw::make_subs(const gpw_grid_desc &g, const PairDev &pd, int pair_index,
                         const PairGeom &geo, int prof_off, int prof_stride,
                         std::vector<SubDev> &out) {
    int rlo[3][64], rhi[3][64], base_of[3][64], nr[3];
    for (int d = 0; d < 3; d++) {
        const int lb = geo.lb[d], ub = geo.ub[d];
        nr[d] = 0;
        if (g.periodic[d]) {
            const int N = g.npts[d];
            for (long long s = lb; s <= ub; s += N) {
                if (nr[d] >= 64) return fail(GPW_ERR_ARG, "cube spans more than 64 periods");
                rlo[d][nr[d]] = (int)s;
                rhi[d][nr[d]] = (int)std::min<long long>(ub, s + N - 1);
                base_of[d][nr[d]++] = imod_h(geo.cc[d] + (int)s, N);
            }
        } else {
            const int gshift = imod_h(geo.cc[d] + lb + g.lb_grid[d], g.npts[d]) - lb;
            if (lb + gshift < 0 || ub + gshift > g.ngrid[d] - 1)
                return fail(GPW_ERR_BOUNDS, "non-periodic cube outside the local grid");
            rlo[d][0] = lb;
            rhi[d][0] = ub;
            base_of[d][0] = lb + gshift;
            nr[d] = 1;
        }
    }
    for (int iz = 0; iz < nr[2]; iz++)
        for (int iy = 0; iy < nr[1]; iy++)
            for (int ix = 0; ix < nr[0]; ix++) {
                SubDev s{};
                const int idx[3] = {ix, iy, iz};
                for (int d = 0; d < 3; d++) {
                    s.roff[d] = geo.roff[d];
                    s.lb[d] = rlo[d][idx[d]];
                    s.ub[d] = rhi[d][idx[d]];
                    s.base[d] = base_of[d][idx[d]];
                    s.period[d] = g.periodic[d] ? g.npts[d] : BIG;
                }
                s.zetp = pd.zetp;
                s.lp = pd.lp;
                s.pair = pair_index;
                s.coef_off = pd.coef_off;
                s.R2 = geo.R2;
                s.prof_off = prof_off;
                s.prof_stride = prof_stride;
                out.push_back(s);
            }
    return GPW_OK;
}

// tiles along one axis touched by the grid indices of a sub-box
static void axis_tiles(const SubDev &s, int d, int n, int T, int ntile, std::vector<int> &out) {
    out.clear();
    const int len = s.ub[d] - s.lb[d] + 1;
    std::vector<char> hit(ntile, 0);
    if (s.period[d] != BIG) {
        const int start = s.base[d];
        const int stop = start + len - 1;  // may exceed n: wraps
        if (stop < n) {
            for (int t = start / T; t <= stop / T; t++) hit[t] = 1;
        } else {
            for (int t = start / T; t < ntile; t++) hit[t] = 1;
            for (int t = 0; t <= (stop - n) / T; t++) hit[t] = 1;
        }
    } else {
        for (int t = s.base[d] / T; t <= (s.base[d] + len - 1) / T; t++) hit[t] = 1;
    }
    for (int t = 0; t < ntile; t++)
        if (hit[t]) out.push_back(t);
}

// Sphere points of a sub-box, with the rows (z, y) and z-planes that hold
// any: the algorithmic work of the contraction order z, y, x
// (grid_base_ref_c.cc:431-494) is C3(lp) FMAs per plane, C2(lp) per row and
// lp + 1 per point.
struct SubWork {
    long long pts = 0, rows = 0, planes = 0;
};

static SubWork sub_points_general(const gpw_grid_desc &g, const SubDev &s) {
    SubWork w;
    for (int k = s.lb[2]; k <= s.ub[2]; k++) {
        const long long rows0 = w.rows;
        for (int j = s.lb[1]; j <= s.ub[1]; j++) {
            int lo, hi;
            if (!general_row_host(g, s.roff[0], j - s.roff[1], k - s.roff[2], s.R2, lo, hi)) continue;
            lo = std::max(lo, s.lb[0]);
            hi = std::min(hi, s.ub[0]);
            if (hi >= lo) {
                w.pts += hi - lo + 1;
                w.rows++;
            }
        }
        if (w.rows > rows0) w.planes++;
    }
    return w;
}

// fold offset of the sphere profiles: |i| for i <= 0, i - 1 above
static int fold_off_h(int i) { return i <= 0 ? -i : i - 1; }

static SubWork sub_points(const SubDev &s, const std::vector<int> &prof) {
    auto fo = fold_off_h;
    SubWork w;
    for (int k = s.lb[2]; k <= s.ub[2]; k++) {
        const long long rows0 = w.rows;
        for (int j = s.lb[1]; j <= s.ub[1]; j++) {
            int T;
            if (s.prof_off < 0)
                T = BIG;
            else
                T = prof[s.prof_off + fo(k) * s.prof_stride + fo(j)];
            if (T < 0) continue;
            // offsets i in [lb, ub] with fo(i) <= T  <=>  i in [-T, T+1]
            const int lo = std::max(s.lb[0], -T), hi = std::min(s.ub[0], T + 1);
            if (hi >= lo) {
                w.pts += hi - lo + 1;
                w.rows++;
            }
        }
        if (w.rows > rows0) w.planes++;
    }
    return w;
}

// FP64 flops of one contraction (collocate or integrate) of a sub-box at
// product angular momentum lp: 2 per FMA
static double contraction_flops(const SubWork &w, int lp) {
    return 2.0 * ((double)(lp + 1) * w.pts + (double)ncoef2(lp) * w.rows + (double)ncoset(lp) * w.planes);
}

// Lattice transform of a general cell (grid_base_ref_c.cc:647-698): for
// (x - xp)^lx (y - yp)^ly (z - zp)^lz expanded in lattice displacements
// (i - ip)^il (j - jp)^jl (k - kp)^kl, T[ijk][xyz] with both indices packed
// as (z-like, y-like, x-like) triples of total degree <= lp.
void gpw::multinomial_matrix(const double dh[3][3], int lp, double *T) {
    const int n = ncoset(lp);
    std::vector<double> hp((size_t)(lp + 1) * 9);  // hp[k][j][i] = dh[j][i]^k
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            hp[(0 * 3 + j) * 3 + i] = 1.0;
            for (int k = 1; k <= lp; k++) hp[(k * 3 + j) * 3 + i] = hp[((k - 1) * 3 + j) * 3 + i] * dh[j][i];
        }
    auto H = [&](int k, int j, int i) { return hp[(k * 3 + j) * 3 + i]; };
    std::vector<double> fac(lp + 1, 1.0);
    for (int k = 1; k <= lp; k++) fac[k] = fac[k - 1] * k;
    for (size_t q = 0; q < (size_t)n * n; q++) T[q] = 0.0;
    for (int klx = 0; klx <= lp; klx++)
        for (int jlx = 0; jlx <= lp - klx; jlx++)
            for (int ilx = 0; ilx <= lp - klx - jlx; ilx++) {
                const int lx = ilx + jlx + klx;
                for (int kly = 0; kly <= lp - lx; kly++)
                    for (int jly = 0; jly <= lp - lx - kly; jly++)
                        for (int ily = 0; ily <= lp - lx - kly - jly; ily++) {
                            const int ly = ily + jly + kly;
                            for (int klz = 0; klz <= lp - lx - ly; klz++)
                                for (int jlz = 0; jlz <= lp - lx - ly - klz; jlz++)
                                    for (int ilz = 0; ilz <= lp - lx - ly - klz - jlz; ilz++) {
                                        const int lz = ilz + jlz + klz;
                                        const double w =
                                            H(ilx, 0, 0) * H(jlx, 1, 0) * H(klx, 2, 0) *
                                            H(ily, 0, 1) * H(jly, 1, 1) * H(kly, 2, 1) *
                                            H(ilz, 0, 2) * H(jlz, 1, 2) * H(klz, 2, 2) *
                                            fac[lx] * fac[ly] * fac[lz] /
                                            (fac[ilx] * fac[ily] * fac[ilz] * fac[jlx] * fac[jly] *
                                             fac[jlz] * fac[klx] * fac[kly] * fac[klz]);
                                        const int ijk = cidx(lp, klx + kly + klz, jlx + jly + jlz,
                                                             ilx + ily + ilz);
                                        const int xyz = cidx(lp, lz, ly, lx);
                                        T[(size_t)ijk * n + xyz] += w;
                                    }
                        }
            }
}

// Lattice transforms T(lp) of a general cell, in both directions (dir 0: T,
// dir 1: T^T), exact zeros dropped, each (lp, dir) a block of rows padded to
// its longest row and stored entry-major (LatticeTab).
static gpw_status lattice_tables_of(const double dh[3][3], const std::vector<int> &lps, LatticeHost &out) {
    out.col.clear();
    out.val.clear();
    out.blk.assign(2 * (LPMAX + 1), LatBlk{});
    std::vector<double> T;
    for (int lp : lps) {
        const int n = ncoset(lp);
        T.assign((size_t)n * n, 0.0);
        multinomial_matrix(dh, lp, T.data());
        const double *Tl = T.data();
        for (int dir = 0; dir < 2; dir++) {
            auto at = [&](int r, int c) { return dir == 0 ? Tl[(size_t)r * n + c] : Tl[(size_t)c * n + r]; };
            int width = 0;
            for (int r = 0; r < n; r++) {
                int nz = 0;
                for (int c = 0; c < n; c++) nz += at(r, c) != 0.0;
                width = std::max(width, nz);
            }
            LatBlk &b = out.blk[2 * lp + dir];
            b.base = (long long)out.val.size();
            b.width = width;
            b.n = n;
            out.col.resize(out.col.size() + (size_t)width * n, 0);
            out.val.resize(out.val.size() + (size_t)width * n, 0.0);
            for (int r = 0; r < n; r++) {
                int k = 0;
                for (int c = 0; c < n; c++) {
                    const double v = at(r, c);
                    if (v == 0.0) continue;
                    out.col[b.base + (size_t)k * n + r] = c;
                    out.val[b.base + (size_t)k * n + r] = v;
                    k++;
                }
            }
        }
    }
    if (out.val.size() > 0x7fffffffULL) return fail(GPW_ERR_ARG, "lattice transform too large");
    return GPW_OK;
}

bool gpw::stepped_cell(const double dh[3][3], double h[6]) {
    if (dh[0][1] != 0.0 || dh[0][2] != 0.0 || dh[1][2] != 0.0) return false;
    h[0] = dh[0][0];
    h[1] = dh[1][0];
    h[2] = dh[2][0];
    h[3] = dh[1][1];
    h[4] = dh[2][1];
    h[5] = dh[2][2];
    return true;
}

gpw_status gpw::lattice_tables(const double dh[3][3], int lpm, LatticeHost &out) {
    std::vector<int> lps;
    for (int lp = 0; lp <= lpm; lp++) lps.push_back(lp);
    return lattice_tables_of(dh, lps, out);
}

// only the degrees lps[0 .. n) (rows of other degrees are absent)
gpw_status gpw::lattice_tables_some(const double dh[3][3], const int *lps, int n, LatticeHost &out) {
    std::vector<int> v;
    for (int i = 0; i < n; i++)
        if (lps[i] >= 0 && lps[i] <= LPMAX && std::find(v.begin(), v.end(), lps[i]) == v.end())
            v.push_back(lps[i]);
    return lattice_tables_of(dh, v, out);
}

// Host part of a task list: profiles, pairs, sub-boxes, tile work lists.
// Work items ("entries") are (sub-box, tile) pairs numbered pair-major, so a
// pair's partial moments are contiguous; per (lp_add, class) a CSR over tiles
// lists the entry ids the tile kernels sweep.
struct HostPlan {
    std::vector<int> prof_flat;
    std::vector<int> ent_sub, pe_start;
    std::vector<long long> off0, off2;
    std::vector<int> tile_start;  // [NSET][NCLS][ntiles + 1]
    std::vector<EntryRec> recs;   // tile-list order, segments per (set, class)
    std::vector<int> small;       // sub-box ids for k_moments_small
    std::vector<unsigned char> small_mask;  // [3][npairs]: integrate w/o, with forces; collocate
    std::vector<int> csmall;      // sub-box ids for k_collocate_small
    std::vector<long long> box_off;  // per sub-box: offset of its small-collocate box, -1 none
    long long coef_total = 0, partial_total = 0;
};

// Host part shared by both planners: profiles, pairs, sub-boxes (with the
// squared sphere radius of their pair), the tile grid, and the pairs of the
// small-sphere integrate kernel. O(pairs + sub-boxes).
static gpw_status plan_common(gpw_tasklist *tl, const std::vector<gpw_pair_desc> &pairs,
                              const std::vector<PairGeom> &geos, const std::vector<Profile> &profiles,
                              int splits, HostPlan &hp) {
    const gpw_grid_desc &g = tl->grid;
    std::vector<int> &prof_flat = hp.prof_flat;
    std::vector<int> prof_off(profiles.size()), prof_stride(profiles.size());
    for (size_t i = 0; i < profiles.size(); i++) {
        prof_off[i] = (int)prof_flat.size();
        prof_stride[i] = profiles[i].n[1] + 1;
        prof_flat.insert(prof_flat.end(), profiles[i].T.begin(), profiles[i].T.end());
    }
    // pairs + coefficient offsets
    const size_t np = pairs.size();
    tl->pairs_h.resize(np);
    parallel_chunks(np, [&](size_t b, size_t e, unsigned) {
        for (size_t i = b; i < e; i++) tl->pairs_h[i] = make_pair_dev(pairs[i], tl->compute_tau);
    });
    long long coef_total = 0;
    for (size_t i = 0; i < np; i++) {
        PairDev &pd = tl->pairs_h[i];
        pd.coef_off = (int)coef_total;
        coef_total += ncoset(pd.lp);
        if (coef_total > 0x7fffffffLL) return fail(GPW_ERR_ARG, "too many coefficients");
        tl->lp_max = std::max(tl->lp_max, pd.lp);
        tl->la_max = std::max(tl->la_max, pairs[i].la_max);
        tl->lb_max = std::max(tl->lb_max, pairs[i].lb_max);
    }
    hp.coef_total = coef_total;
    // sub-boxes (contiguous per pair), built per chunk of pairs on threads
    // and concatenated in pair order; orthorhombic sub-boxes carry the
    // squared radius of their profile (general ones their pair's radius^2)
    std::vector<std::vector<SubDev>> part(16);
    std::vector<std::vector<int>> part_start(16);
    std::vector<size_t> part_begin(16, 0);
    std::vector<int> bad(16, 0);
    std::vector<std::string> why(16);
    parallel_chunks(np, [&](size_t b, size_t e, unsigned t) {
        part_begin[t] = b;
        part[t].reserve((e - b) + (e - b) / 4);
        for (size_t i = b; i < e; i++) {
            part_start[t].push_back((int)part[t].size());
            const int pi = geos[i].prof;
            PairGeom geo = geos[i];
            if (pi >= 0) geo.R2 = profiles[pi].R2;
            gpw_status st = make_subs(tl->grid, tl->pairs_h[i], (int)i, geo, pi >= 0 ? prof_off[pi] : -1,
                                      pi >= 0 ? prof_stride[pi] : 0, part[t]);
            if (st) {
                bad[t] = st;
                why[t] = gpw_last_error();
                return;
            }
        }
    });
    for (int t = 0; t < 16; t++)
        if (bad[t]) return fail(bad[t], why[t]);
    size_t nsub = 0;
    for (int t = 0; t < 16; t++) nsub += part[t].size();
    if (nsub > 0x3fffffffULL) return fail(GPW_ERR_ARG, "too many sub-boxes");
    tl->subs_h.resize(nsub);
    tl->pair_sub_start.assign(np + 1, 0);
    size_t at = 0;
    for (int t = 0; t < 16; t++) {
        if (part[t].empty()) continue;
        std::copy(part[t].begin(), part[t].end(), tl->subs_h.begin() + at);
        for (size_t k = 0; k < part_start[t].size(); k++)
            tl->pair_sub_start[part_begin[t] + k] = (int)at + part_start[t][k];
        at += part[t].size();
    }
    tl->pair_sub_start[np] = (int)nsub;
    // tiles
    TileGrid &tg = tl->tg;
    tg.nx = g.ngrid[0];
    tg.ny = g.ngrid[1];
    tg.nz = g.ngrid[2];
    tg.ntx = (tg.nx + TX - 1) / TX;
    tg.nty = (tg.ny + TY - 1) / TY;
    tg.ntz = (tg.nz + TZ - 1) / TZ;
    tg.general = !is_ortho(g) ? 1 : 0;
    for (int d = 0; d < 3; d++) {
        tg.dr[d] = g.dh[d][d];
        tg.period[d] = g.periodic[d] ? g.npts[d] : BIG;
        for (int e = 0; e < 3; e++) tg.dh[d][e] = g.dh[d][e];
    }
    tl->ntiles = tg.ntx * tg.nty * tg.ntz;
    tl->splits = std::max(1, splits);
    for (int c = 0; c < NCLS; c++) {
        tl->csplit[c] = tl->splits;
        tl->isplit[0][c] = tl->isplit[1][c] = tl->splits;
    }
    // pairs for the small-sphere integrate kernel, per force extension
    hp.small.clear();
    hp.small_mask.assign(3 * pairs.size(), 0);
    for (int w = 0; w < 2; w++) {
        std::vector<int> by_cls[NSCLS];
        for (size_t p = 0; p < pairs.size(); p++) {
            if (tl->pair_sub_start[p + 1] - tl->pair_sub_start[p] != 1) continue;
            const int si = tl->pair_sub_start[p];
            const SubDev &sd = tl->subs_h[si];
            const int L = sd.lp + 2 * w + 1;
            const int nyz = tl->tg.general ? SMALL_NY : SMALL_NYZ_ORTHO;
            if (L > SMALL_LMAX || sd.ub[0] - sd.lb[0] + 1 > SMALL_NX ||
                sd.ub[1] - sd.lb[1] + 1 > std::min(nyz, SMALL_NY) ||
                sd.ub[2] - sd.lb[2] + 1 > std::min(nyz, SMALL_NZ))
                continue;
            hp.small_mask[w * pairs.size() + p] = 1;
            by_cls[std::min((L - 1) / 4, NSCLS - 1)].push_back(si);
        }
        for (int c = 0; c < NSCLS; c++) {
            tl->small_off[w][c] = (long long)hp.small.size();
            tl->small_cnt[w][c] = (long long)by_cls[c].size();
            hp.small.insert(hp.small.end(), by_cls[c].begin(), by_cls[c].end());
        }
    }
    // pairs for the small-sphere collocate, their boxes in sub-box order
    hp.csmall.clear();
    hp.box_off.assign(tl->subs_h.size(), -1);
    {
        const int lim = tl->tg.general ? CSMALL_N : CSMALL_N_ORTHO;
        std::vector<int> by_cls[NSCLS];
        long long tot = 0;
        for (size_t p = 0; p < pairs.size() && lim > 0; p++) {
            if (tl->pair_sub_start[p + 1] - tl->pair_sub_start[p] != 1) continue;
            const int si = tl->pair_sub_start[p];
            const SubDev &sd = tl->subs_h[si];
            const int L = sd.lp + 1;
            long long vol = 1;
            bool fits = L <= SMALL_LMAX;
            for (int d = 0; d < 3; d++) {
                fits = fits && sd.ub[d] - sd.lb[d] + 1 <= lim;
                vol *= sd.ub[d] - sd.lb[d] + 1;
            }
            if (!fits) continue;
            hp.small_mask[2 * pairs.size() + p] = 1;
            hp.box_off[si] = tot;
            tot += vol;
            by_cls[std::min((L - 1) / 4, NSCLS - 1)].push_back(si);
        }
        for (int c = 0; c < NSCLS; c++) {
            tl->csmall_off[c] = (long long)hp.csmall.size();
            tl->csmall_cnt[c] = (long long)by_cls[c].size();
            hp.csmall.insert(hp.csmall.end(), by_cls[c].begin(), by_cls[c].end());
        }
        tl->sval_total = tot;
    }
    tl->mstride = ncoset(tl->lp_max + 2);
    return GPW_OK;
}

// Host planner of the work items (gpw_tasklist_plan_stats; the task lists
// themselves are planned on the device, gpw_devplan). Items ("entries") are
// (sub-box, tile) pairs numbered pair-major, so a pair's partial moments are
// contiguous; per (set, class) a CSR over tiles lists the item records in
// the order the tile kernels sweep them (sub-box order within a tile).
static gpw_status plan_entries_host(gpw_tasklist *tl, HostPlan &hp) {
    const gpw_grid_desc &g = tl->grid;
    const TileGrid &tg = tl->tg;
    const size_t npairs = tl->pairs_h.size();
    std::vector<int> ent_tile;
    std::vector<int> tx, ty, tz;
    hp.pe_start.assign(npairs + 1, 0);
    tl->cube_points = 0;
    tl->flops_collocate = tl->flops_integrate[0] = tl->flops_integrate[1] = 0.0;
    for (size_t si = 0; si < tl->subs_h.size(); si++) {
        const SubDev &s = tl->subs_h[si];
        axis_tiles(s, 0, tg.nx, TX, tg.ntx, tx);
        axis_tiles(s, 1, tg.ny, TY, tg.nty, ty);
        axis_tiles(s, 2, tg.nz, TZ, tg.ntz, tz);
        // the tiles that hold a point of the sub-box's sphere (the device
        // planner's test, with the host's row bounds)
        auto row = [&](int j, int k, int &lo, int &hi) {
            if (tg.general) return general_row_host(g, s.roff[0], j - s.roff[1], k - s.roff[2], s.R2, lo, hi) && lo <= hi;
            const int T = s.prof_off < 0 ? BIG : hp.prof_flat[s.prof_off + fold_off_h(k) * s.prof_stride + fold_off_h(j)];
            lo = -T;
            hi = T + 1;
            return T >= 0;
        };
        for (int a : tz)
            for (int b : ty)
                for (int c : tx) {
                    if (!item_has_points(s, tg, c, b, a, row)) continue;
                    ent_tile.push_back((a * tg.nty + b) * tg.ntx + c);
                    hp.ent_sub.push_back((int)si);
                    hp.pe_start[s.pair + 1]++;
                }
        const SubWork w = tg.general ? sub_points_general(g, s) : sub_points(s, hp.prof_flat);
        tl->cube_points += w.pts;
        tl->flops_collocate += contraction_flops(w, s.lp);
        tl->flops_integrate[0] += contraction_flops(w, s.lp);
        tl->flops_integrate[1] += contraction_flops(w, s.lp + 2);
    }
    for (size_t p = 0; p < npairs; p++) hp.pe_start[p + 1] += hp.pe_start[p];
    const size_t ne = ent_tile.size();
    if (ne > 0x7fffffffULL) return fail(GPW_ERR_ARG, "too many work items");
    tl->nentries = (long long)ne;
    tl->tile_points = tl->nentries * (long long)(TX * TY * TZ);
    // partial-moment offsets, pair-major
    hp.off0.assign(ne, 0);
    hp.off2.assign(ne, 0);
    long long acc0 = 0, acc2 = 0;
    for (size_t e = 0; e < ne; e++) {
        const int lp = tl->subs_h[hp.ent_sub[e]].lp;
        hp.off0[e] = acc0;
        hp.off2[e] = acc2;
        acc0 += ncoset(lp);
        acc2 += ncoset(lp + 2);
    }
    hp.partial_total = acc2;
    // per (set, class): CSR over tiles of the item records
    const int nt1 = tl->ntiles + 1;
    hp.tile_start.assign((size_t)NSET * NCLS * nt1, 0);
    hp.recs.clear();
    hp.recs.reserve(NSET * ne);
    for (int a = 0; a < NSET; a++) {
        const int lp_add = a == 1 ? 2 : 0;
        std::vector<int> cls(ne);
        for (size_t e = 0; e < ne; e++) {
            const SubDev &sd = tl->subs_h[hp.ent_sub[e]];
            const bool csm = hp.small_mask[2 * npairs + sd.pair] != 0;
            const bool skip = a == 0 ? csm
                            : a == 1 ? hp.small_mask[npairs + sd.pair] != 0
                            : a == 2 ? hp.small_mask[sd.pair] != 0 : !csm;
            cls[e] = skip ? -1 : a == 3 ? 0 : class_of(sd.lp + lp_add);
        }
        for (int c = 0; c < NCLS; c++) {
            int *start = hp.tile_start.data() + (size_t)(a * NCLS + c) * nt1;
            for (size_t e = 0; e < ne; e++)
                if (cls[e] == c) start[ent_tile[e] + 1]++;
            for (int t = 0; t < tl->ntiles; t++) start[t + 1] += start[t];
            tl->list_off[a][c] = (long long)hp.recs.size();
            tl->class_entries[a][c] = start[tl->ntiles];
            hp.recs.resize(hp.recs.size() + start[tl->ntiles]);
            EntryRec *list = hp.recs.data() + tl->list_off[a][c];
            std::vector<int> fillp(start, start + tl->ntiles);
            for (size_t e = 0; e < ne; e++) {
                if (cls[e] != c) continue;
                const SubDev &sd = tl->subs_h[hp.ent_sub[e]];
                EntryRec r{};
                for (int d = 0; d < 3; d++) {
                    r.roff[d] = sd.roff[d];
                    r.base[d] = sd.base[d];
                    r.lb[d] = sd.lb[d];
                    r.ub[d] = sd.ub[d];
                }
                r.zetp = sd.zetp;
                r.R2 = sd.R2;
                r.ent_off = a == 3 ? hp.box_off[hp.ent_sub[e]] : a == 1 ? hp.off2[e] : hp.off0[e];
                r.lp = sd.lp;
                r.coef_off = sd.coef_off;
                r.prof_off = sd.prof_off;
                r.prof_stride = sd.prof_stride;
                list[fillp[ent_tile[e]]++] = r;
            }
        }
    }
    return GPW_OK;
}

static gpw_status plan_tasklist(gpw_tasklist *tl, const std::vector<gpw_pair_desc> &pairs,
                                const std::vector<PairGeom> &geos,
                                const std::vector<Profile> &profiles, int splits, HostPlan &hp) {
    gpw_status st = plan_common(tl, pairs, geos, profiles, splits, hp);
    if (st) return st;
    return plan_entries_host(tl, hp);
}

// Groups of pairs by la_max + lb_max (+ 2 tau), which sizes the LDS and the
// registers of their transforms, so that small blocks do not run at the
// occupancy of the largest one; inside a group pairs are sorted by block offset and pairs of
// one block form a segment. Blocks may be shared only whole (same offset,
// same la_max and lb_max); any other overlap is an argument error.
static gpw_status group_pairs(gpw_tasklist *tl, std::vector<int> &order, std::vector<int> &segs) {
    const int t = tl->compute_tau;
    // overlap check over all blocks (pairs usually come in block order:
    // then no sort is needed)
    bool sorted = true;
    for (int i = 1; i < tl->npairs && sorted; i++)
        sorted = tl->pairs_h[i - 1].block_off <= tl->pairs_h[i].block_off;
    {
        std::vector<int> idx(tl->npairs);
        for (int i = 0; i < tl->npairs; i++) idx[i] = i;
        if (!sorted)
            std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
                return tl->pairs_h[a].block_off < tl->pairs_h[b].block_off;
            });
        long long end = -1;
        const PairDev *prev = nullptr;
        for (int i : idx) {
            const PairDev &p = tl->pairs_h[i];
            const long long sz = (long long)ncoset(p.lb_max) * ncoset(p.la_max);
            if (prev && p.block_off == prev->block_off) {
                if (p.la_max != prev->la_max || p.lb_max != prev->lb_max)
                    return fail(GPW_ERR_ARG, "pairs sharing a block need equal la_max and lb_max");
                continue;
            }
            if (p.block_off < end) return fail(GPW_ERR_ARG, "coefficient blocks overlap");
            end = p.block_off + sz;
            prev = &p;
        }
    }
    std::vector<std::vector<int>> byc;
    for (int i = 0; i < tl->npairs; i++) {
        const PairDev &p = tl->pairs_h[i];
        const int c = p.la_max + p.lb_max + 2 * t;
        if ((int)byc.size() <= c) byc.resize(c + 1);
        byc[c].push_back(i);
    }
    order.clear();
    segs.clear();
    tl->groups.clear();
    tl->nab_max = 0;
    tl->nhx_max = 0;
    for (auto &v : byc) {
        if (v.empty()) continue;
        if (!sorted)
            std::stable_sort(v.begin(), v.end(), [&](int a, int b) {
                return tl->pairs_h[a].block_off < tl->pairs_h[b].block_off;
            });
        gpw_tasklist::PairGroup g{(int)order.size(), (int)v.size(), 0, 0, (int)segs.size(), 0, 0, 0};
        long long last = -1;
        for (int i : v) {
            if (tl->pairs_h[i].block_off != last) {
                segs.push_back((int)order.size());
                g.nseg++;
                last = tl->pairs_h[i].block_off;
            }
            order.push_back(i);
            const PairDev &pd = tl->pairs_h[i];
            g.la_max = std::max(g.la_max, pd.la_max);
            g.lb_max = std::max(g.lb_max, pd.lb_max);
            g.nab = std::max(g.nab, block_shell_entries(pd.la_min, pd.la_max, pd.lb_min, pd.lb_max));
            g.nhx = std::max(g.nhx, hab_cross_entries(pd.la_min, pd.la_max, pd.lb_min, pd.lb_max));
        }
        tl->groups.push_back(g);
        tl->nab_max = std::max(tl->nab_max, g.nab);
        tl->nhx_max = std::max(tl->nhx_max, g.nhx);
    }
    segs.push_back((int)order.size());
    return GPW_OK;
}

// device memory freed when the scope ends (the planner's temporaries)
struct DevTemp {
    std::vector<void *> ptrs;
    ~DevTemp() {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T>
    hipError_t alloc(T **p, size_t n) {
        *p = nullptr;
        hipError_t e = hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

// Device planner of the work items, from the uploaded sub-boxes: the same
// items, partial-moment offsets, per-(set, class) tile CSR and records as
// plan_entries_host, with the same order of items within each tile.
//  1. k_plan_count: items per sub-box (tiles holding a point of its sphere)
//     and the algorithmic work (stats);
//  2. host: prefix sums over sub-boxes (item starts, partial offsets);
//  3. k_plan_entries: each item's sub-box, tile, offsets and sort keys
//     (class * ntiles + tile per set), with per-key counts -> tile CSR;
//  4. per set, a stable radix sort of the keys (items of one key keep their
//     sub-box order) and k_plan_records writes the records in that order.
static gpw_status plan_device(gpw_tasklist *tl, HostPlan &hp, hipStream_t st) {
    const int nsub = (int)tl->subs_h.size();
    const int npairs = tl->npairs;
    const int ntiles = tl->ntiles;
    DevTemp tmp;
    int *d_counts = nullptr;
    unsigned long long *d_work = nullptr;
    const int nwork = 3 * (LPMAX + 1);
    HIP_TRY(tmp.alloc(&d_counts, nsub));
    HIP_TRY(tmp.alloc(&d_work, nwork));
    HIP_TRY(hipMemsetAsync(d_work, 0, nwork * sizeof(unsigned long long), st));
    HIP_TRY(launch_plan_count(tl->d_subs, nsub, tl->d_prof, tl->tg, d_counts, d_work, st));
    std::vector<int> counts(nsub);
    std::vector<unsigned long long> work(nwork);
    if (nsub) HIP_TRY(hipMemcpyAsync(counts.data(), d_counts, nsub * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(work.data(), d_work, nwork * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    // stats
    tl->cube_points = 0;
    tl->flops_collocate = tl->flops_integrate[0] = tl->flops_integrate[1] = 0.0;
    for (int lp = 0; lp <= LPMAX; lp++) {
        SubWork w;
        w.pts = (long long)work[3 * lp];
        w.rows = (long long)work[3 * lp + 1];
        w.planes = (long long)work[3 * lp + 2];
        tl->cube_points += w.pts;
        tl->flops_collocate += contraction_flops(w, lp);
        tl->flops_integrate[0] += contraction_flops(w, lp);
        tl->flops_integrate[1] += contraction_flops(w, lp + 2);
    }
    // item starts (sub-box-major; sub-boxes are pair-major) and partial offsets
    std::vector<long long> ent_start(nsub + 1), base0(nsub), base2(nsub);
    long long ne = 0, acc0 = 0, acc2 = 0;
    hp.pe_start.assign(npairs + 1, 0);
    for (int s = 0; s < nsub; s++) {
        const SubDev &sd = tl->subs_h[s];
        ent_start[s] = ne;
        base0[s] = acc0;
        base2[s] = acc2;
        ne += counts[s];
        acc0 += (long long)counts[s] * ncoset(sd.lp);
        acc2 += (long long)counts[s] * ncoset(sd.lp + 2);
    }
    ent_start[nsub] = ne;
    for (int p = 0; p < npairs; p++) hp.pe_start[p] = (int)ent_start[tl->pair_sub_start[p]];
    hp.pe_start[npairs] = (int)ne;
    if (ne > 0x7fffffffLL) return fail(GPW_ERR_ARG, "too many work items");
    tl->nentries = ne;
    tl->tile_points = ne * (long long)(TX * TY * TZ);
    hp.partial_total = acc2;
    gpw_status stt;
    if ((stt = upload(&tl->d_pe_start, hp.pe_start))) return stt;
    // items, keys, per-key counts
    const int n = (int)ne;
    const size_t nkey = (size_t)NCLS * ntiles;
    long long *d_es = nullptr, *d_b0 = nullptr, *d_b2 = nullptr;
    int *d_ent_sub = nullptr, *d_vals = nullptr, *d_sorted = nullptr, *d_hist = nullptr;
    unsigned *d_keys = nullptr, *d_keys_out = nullptr;
    HIP_TRY(tmp.alloc(&d_es, nsub + 1));
    HIP_TRY(tmp.alloc(&d_b0, nsub));
    HIP_TRY(tmp.alloc(&d_b2, nsub));
    HIP_TRY(tmp.alloc(&d_ent_sub, n));
    HIP_TRY(tmp.alloc(&d_vals, n));
    HIP_TRY(tmp.alloc(&d_sorted, n));
    HIP_TRY(tmp.alloc(&d_keys, (size_t)NSET * n));
    HIP_TRY(tmp.alloc(&d_keys_out, n));
    HIP_TRY(tmp.alloc(&d_hist, NSET * nkey));
    HIP_TRY(hipMalloc((void **)&tl->d_ent_off0, sizeof(long long) * std::max(n, 1)));
    HIP_TRY(hipMalloc((void **)&tl->d_ent_off2, sizeof(long long) * std::max(n, 1)));
    HIP_TRY(hipMemcpyAsync(d_es, ent_start.data(), (nsub + 1) * sizeof(long long), hipMemcpyHostToDevice, st));
    if (nsub) {
        HIP_TRY(hipMemcpyAsync(d_b0, base0.data(), nsub * sizeof(long long), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_b2, base2.data(), nsub * sizeof(long long), hipMemcpyHostToDevice, st));
    }
    HIP_TRY(hipMemsetAsync(d_hist, 0, NSET * nkey * sizeof(int), st));
    HIP_TRY(launch_plan_entries(tl->d_subs, nsub, d_es, d_b0, d_b2, tl->d_prof, tl->tg, ntiles, tl->d_small_mask, npairs, n,
                                d_ent_sub, tl->d_ent_off0, tl->d_ent_off2, d_keys, d_vals, d_hist, st));
    std::vector<int> hist(NSET * nkey);
    HIP_TRY(hipMemcpyAsync(hist.data(), d_hist, NSET * nkey * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    // tile CSR per (set, class), record segments in set-major, class-major order
    const int nt1 = ntiles + 1;
    hp.tile_start.assign((size_t)NSET * NCLS * nt1, 0);
    long long nrec = 0, set_base[NSET];
    for (int a = 0; a < NSET; a++) {
        set_base[a] = nrec;
        for (int c = 0; c < NCLS; c++) {
            int *start = hp.tile_start.data() + (size_t)(a * NCLS + c) * nt1;
            for (int t = 0; t < ntiles; t++) start[t + 1] = start[t] + hist[(size_t)a * nkey + (size_t)c * ntiles + t];
            tl->list_off[a][c] = nrec;
            tl->class_entries[a][c] = start[ntiles];
            nrec += start[ntiles];
        }
    }
    if ((stt = upload(&tl->d_tile_start, hp.tile_start))) return stt;
    HIP_TRY(hipMalloc((void **)&tl->d_recs, sizeof(EntryRec) * std::max<long long>(nrec, 1)));
    // sort each set's keys (values: item ids) and write its records
    int end_bit = 1;
    while ((1ULL << end_bit) <= (unsigned long long)nkey) end_bit++;
    size_t tbytes = 0;
    HIP_TRY(sort_pairs_u32(nullptr, &tbytes, d_keys, d_keys_out, d_vals, d_sorted, n, end_bit, st));
    void *d_temp = nullptr;
    HIP_TRY(tmp.alloc((char **)&d_temp, tbytes));
    for (int a = 0; a < NSET; a++) {
        const long long nvalid = (a + 1 < NSET ? set_base[a + 1] : nrec) - set_base[a];
        if (nvalid == 0) continue;
        size_t tb = tbytes;
        HIP_TRY(sort_pairs_u32(d_temp, &tb, d_keys + (size_t)a * n, d_keys_out, d_vals, d_sorted, n, end_bit, st));
        HIP_TRY(launch_plan_records(tl->d_subs, d_ent_sub, a == 1 ? tl->d_ent_off2 : tl->d_ent_off0,
                                    a == 3 ? tl->d_box_off : nullptr, d_sorted,
                                    (int)nvalid, tl->d_recs + set_base[a], st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    return GPW_OK;
}

// Uploads the host plan and, unless the host planned the work items
// (hp.recs filled), plans them on the device; then allocates the work
// buffers.
static gpw_status upload_tasklist(gpw_tasklist *tl, HostPlan &hp, bool device_items) {
    gpw_status st;
    std::vector<int> order, segs;
    if ((st = group_pairs(tl, order, segs))) return st;  // argument checks before device work
    if ((st = upload(&tl->d_pairs, tl->pairs_h))) return st;
    if ((st = upload(&tl->d_order, order))) return st;
    if ((st = upload(&tl->d_segs, segs))) return st;
    if ((st = upload(&tl->d_subs, tl->subs_h))) return st;
    if ((st = upload(&tl->d_prof, hp.prof_flat))) return st;
    tl->tg.prof = tl->d_prof;
    if ((st = upload(&tl->d_small, hp.small))) return st;
    if ((st = upload(&tl->d_small_mask, hp.small_mask))) return st;
    if ((st = upload(&tl->d_csmall, hp.csmall))) return st;
    if ((st = upload(&tl->d_box_off, hp.box_off))) return st;
    if (tl->sval_total > 0) HIP_TRY(hipMalloc((void **)&tl->d_sval, sizeof(double) * tl->sval_total));
    if (device_items) {
        const double t0 = now_ms();
        tl->plan_ms[1] += t0 - tl->plan_ms[3];  // uploads since the host plan
        if ((st = plan_device(tl, hp, nullptr))) return st;
        tl->plan_ms[2] = now_ms() - t0;
        tl->plan_ms[3] = now_ms();
    } else {
        if ((st = upload(&tl->d_tile_start, hp.tile_start))) return st;
        if ((st = upload(&tl->d_recs, hp.recs))) return st;
        if ((st = upload(&tl->d_pe_start, hp.pe_start))) return st;
        if ((st = upload(&tl->d_ent_off0, hp.off0))) return st;
        if ((st = upload(&tl->d_ent_off2, hp.off2))) return st;
    }
    HIP_TRY(hipMalloc((void **)&tl->d_coef, sizeof(double) * std::max<long long>(hp.coef_total, 1)));
    tl->coef_total = hp.coef_total;
    HIP_TRY(hipMalloc((void **)&tl->d_partial, sizeof(double) * std::max<long long>(hp.partial_total, 1)));
    HIP_TRY(hipMalloc((void **)&tl->d_Mbuf,
                      sizeof(double) * std::max<long long>((long long)tl->npairs * tl->mstride, 1)));
    if (tl->splits > 1 && !tl->d_cpartial) {  // (a re-plan keeps its partial grids)
        const long long ng = (long long)tl->tg.nx * tl->tg.ny * tl->tg.nz;
        HIP_TRY(hipMalloc((void **)&tl->d_cpartial, sizeof(double) * ng * tl->splits));
        tl->cpart_cap = tl->splits;
    }
    {
        // workspace of k_coef / k_hab (stream-ordered, never used by both at
        // once), sized for the largest block and the force extension
        const int t = tl->compute_tau;
        const int lam = tl->la_max + t, lbm = tl->lb_max + t;
        long long stride = 0;
        if (coef_needs_workspace(lam, lbm)) stride = (long long)coef_ws_doubles(lam, lbm);
        stride = std::max(stride, (long long)hab_ws_need(tl->la_max, tl->lb_max, t, tl->nhx_max));
        if (stride > 0) {
            tl->hab_ws_stride = stride;
            tl->hab_ws_blocks = std::min(tl->npairs, HAB_WS_BLOCKS);
            HIP_TRY(hipMalloc((void **)&tl->d_hab_ws, sizeof(double) * stride * tl->hab_ws_blocks));
        }
    }
    if (tl->tg.general && stepped_cell(tl->grid.dh, tl->lat_h)) {
        tl->lat_stepped = 1;
    } else if (tl->tg.general) {
        LatticeHost lh;
        if ((st = lattice_tables(tl->grid.dh, std::min(tl->lp_max + 2, LPMAX), lh))) return st;
        if ((st = upload(&tl->d_mt_col, lh.col))) return st;
        if ((st = upload(&tl->d_mt_val, lh.val))) return st;
        if ((st = upload(&tl->d_mt_blk, lh.blk))) return st;
    }
    return GPW_OK;
}

static gpw_status finish_tasklist(gpw_tasklist *tl, const std::vector<gpw_pair_desc> &pairs,
                                  const std::vector<PairGeom> &geos,
                                  const std::vector<Profile> &profiles, int splits) {
    HostPlan hp;
    const double t0 = now_ms();
    gpw_status st = plan_common(tl, pairs, geos, profiles, splits, hp);
    if (st) return st;
    tl->plan_ms[0] += now_ms() - t0;
    tl->plan_ms[3] = now_ms();  // start of the uploads (upload_tasklist books them)
    if ((st = upload_tasklist(tl, hp, true))) return st;
    tl->plan_ms[3] = now_ms() - tl->plan_ms[3];  // buffers after the device items
    return GPW_OK;
}

gpw_status gpw::check_grid(const gpw_grid_desc *g) {
    if (!g) return fail(GPW_ERR_ARG, "grid descriptor is NULL");
    for (int d = 0; d < 3; d++) {
        if (g->npts[d] <= 0 || g->ngrid[d] <= 0) return fail(GPW_ERR_ARG, "grid sizes must be > 0");
        if (g->periodic[d] && g->ngrid[d] != g->npts[d])
            return fail(GPW_ERR_ARG, "periodic axes need ngrid == npts");
    }
    return GPW_OK;
}

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

const char *gpw_last_error(void) { return g_last_error.c_str(); }

// geometry of every pair from its radius (cube rule of CP2K's cube_utils)
static gpw_status radius_geometry(const gpw_grid_desc *grid, int npairs, const gpw_pair_desc *pairs,
                                  int tau, std::vector<PairGeom> &geos,
                                  std::vector<Profile> &profiles) {
    geos.assign(npairs, PairGeom{});
    const double dr[3] = {grid->dh[0][0], grid->dh[1][1], grid->dh[2][2]};
    const double drmin = std::min(dr[0], std::min(dr[1], dr[2]));
    const bool ortho = is_ortho(*grid);
    // per pair (threads): checks, cube centre / lattice box, profile key imr
    std::vector<long long> imr(ortho ? npairs : 0);
    std::vector<int> bad(16, -1);
    std::vector<std::string> why(16);
    std::vector<gpw_status> code(16, GPW_OK);
    parallel_chunks((size_t)npairs, [&](size_t b, size_t e, unsigned t) {
        for (size_t i = b; i < e; i++) {
            gpw_status st = check_pair(pairs[i], tau);
            if (!st && !(pairs[i].radius > 0.0)) st = fail(GPW_ERR_ARG, "radius must be > 0");
            if (st) {
                bad[t] = (int)i;
                code[t] = st;
                why[t] = gpw_last_error();
                return;
            }
            const PairDev pd = make_pair_dev(pairs[i], tau);
            PairGeom &geo = geos[i];
            if (!ortho) {
                general_geometry(*grid, pd.rp, pairs[i].radius, geo);
                continue;
            }
            cube_centre(*grid, pd.rp, geo.cc, geo.roff);
            long long m = (long long)std::ceil(pairs[i].radius / drmin);
            imr[i] = m < 1 ? 1 : m;
        }
    });
    for (int t = 0; t < 16; t++)  // the first failing chunk in pair order
        if (bad[t] >= 0) return fail(code[t], why[t]);
    if (!ortho) return GPW_OK;
    // one profile per imr, numbered in order of first appearance
    std::map<long long, int> prof_cache;
    std::vector<int> prof_of(npairs);
    for (int i = 0; i < npairs; i++) {
        auto it = prof_cache.find(imr[i]);
        if (it == prof_cache.end()) {
            Profile P;
            int lb[3], ub[3];
            profile_from_radius(pairs[i].radius, dr, P, lb, ub);
            profiles.push_back(std::move(P));
            it = prof_cache.emplace(imr[i], (int)profiles.size() - 1).first;
        }
        prof_of[i] = it->second;
    }
    parallel_chunks((size_t)npairs, [&](size_t b, size_t e, unsigned) {
        for (size_t i = b; i < e; i++) {
            PairGeom &geo = geos[i];
            geo.prof = prof_of[i];
            for (int d = 0; d < 3; d++) {
                geo.lb[d] = -profiles[geo.prof].n[d];
                geo.ub[d] = 1 + profiles[geo.prof].n[d];
            }
        }
    });
    return GPW_OK;
}

gpw_status gpw_tasklist_create(const gpw_grid_desc *grid, int npairs, const gpw_pair_desc *pairs,
                               int compute_tau, const int *tile, int splits, gpw_tasklist **out) {
    if (!out || npairs < 0 || (npairs > 0 && !pairs)) return fail(GPW_ERR_ARG, "bad arguments");
    *out = nullptr;
    gpw_status st = check_grid(grid);
    if (st) return st;
    if (tile && (tile[0] != TX || tile[1] != TY || tile[2] != TZ))
        return fail(GPW_ERR_UNSUPPORTED, "only the built-in tile shape is compiled");
    auto *tl = new (std::nothrow) gpw_tasklist();
    if (!tl) return fail(GPW_ERR_ALLOC, "host allocation failed");
    tl->grid = *grid;
    tl->compute_tau = compute_tau ? 1 : 0;
    tl->npairs = npairs;
    std::vector<gpw_pair_desc> pv(pairs, pairs + npairs);
    std::vector<PairGeom> geos;
    std::vector<Profile> profiles;
    const double t0 = now_ms();
    st = radius_geometry(grid, npairs, pairs, tl->compute_tau, geos, profiles);
    tl->plan_ms[0] = now_ms() - t0;
    if (st || (st = finish_tasklist(tl, pv, geos, profiles, splits))) {
        delete tl;
        return st;
    }
    *out = tl;
    return GPW_OK;
}

gpw_status gpw_tasklist_update(gpw_tasklist *tl, int npairs, const gpw_pair_desc *pairs) {
    if (!tl || npairs != tl->npairs || (npairs > 0 && !pairs))
        return fail(GPW_ERR_ARG, "update needs the task list's number of pairs");
    // the new plan is built on the host first; device buffers are replaced
    // only once it is valid
    std::vector<PairGeom> geos;
    std::vector<Profile> profiles;
    const double t0 = now_ms();
    gpw_status st = radius_geometry(&tl->grid, npairs, pairs, tl->compute_tau, geos, profiles);
    if (st) return st;
    int csplit[NCLS], isplit[2][NCLS];
    std::copy(tl->csplit, tl->csplit + NCLS, csplit);
    std::copy(&tl->isplit[0][0], &tl->isplit[0][0] + 2 * NCLS, &isplit[0][0]);
    HIP_TRY(hipDeviceSynchronize());  // earlier calls may still read the old plan
    tl->release_plan();
    tl->lp_max = tl->la_max = tl->lb_max = 0;
    for (int a = 0; a < NSET; a++)
        for (int c = 0; c < NCLS; c++) tl->class_entries[a][c] = tl->list_off[a][c] = 0;
    tl->plan_ms[0] = now_ms() - t0;
    tl->plan_ms[1] = tl->plan_ms[2] = tl->plan_ms[3] = 0.0;
    std::vector<gpw_pair_desc> pv(pairs, pairs + npairs);
    if ((st = finish_tasklist(tl, pv, geos, profiles, tl->splits))) {
        tl->release_plan();
        tl->npairs = 0;  // unusable until a successful update
        return st;
    }
    // keep the measured splits (the partial grids were kept for them)
    std::copy(csplit, csplit + NCLS, tl->csplit);
    std::copy(&isplit[0][0], &isplit[0][0] + 2 * NCLS, &tl->isplit[0][0]);
    return GPW_OK;
}

gpw_status gpw_tasklist_plan_stats(const gpw_grid_desc *grid, int npairs, const gpw_pair_desc *pairs,
                                   int compute_tau, int splits, gpw_tasklist_stats *out) {
    if (!out || npairs < 0 || (npairs > 0 && !pairs)) return fail(GPW_ERR_ARG, "bad arguments");
    gpw_status st = check_grid(grid);
    if (st) return st;
    gpw_tasklist tl;
    tl.grid = *grid;
    tl.compute_tau = compute_tau ? 1 : 0;
    tl.npairs = npairs;
    std::vector<gpw_pair_desc> pv(pairs, pairs + npairs);
    std::vector<PairGeom> geos;
    std::vector<Profile> profiles;
    HostPlan hp;
    if ((st = radius_geometry(grid, npairs, pairs, tl.compute_tau, geos, profiles)) ||
        (st = plan_tasklist(&tl, pv, geos, profiles, splits, hp)))
        return st;
    return gpw_tasklist_stats_get(&tl, out);
}

gpw_status gpw_pair_geometry(const gpw_grid_desc *grid, const gpw_pair_desc *pair, int compute_tau,
                             int cube_centre_out[3], double roffset_out[3], int lb_cube_out[3],
                             int ub_cube_out[3], int *sphere_bounds, int cap, int *nspheres) {
    if (!grid || !pair || !cube_centre_out || !roffset_out || !lb_cube_out || !ub_cube_out || !nspheres)
        return fail(GPW_ERR_ARG, "NULL argument");
    gpw_status st = check_grid(grid);
    if (st) return st;
    std::vector<PairGeom> geos;
    std::vector<Profile> profiles;
    if ((st = radius_geometry(grid, 1, pair, compute_tau ? 1 : 0, geos, profiles))) return st;
    const PairGeom &g = geos[0];
    for (int d = 0; d < 3; d++) {
        cube_centre_out[d] = g.cc[d];
        roffset_out[d] = g.roff[d];
        lb_cube_out[d] = g.lb[d];
        ub_cube_out[d] = g.ub[d];
    }
    if (g.prof < 0) {  // general cell: lattice box, no sphere_bounds
        *nspheres = 0;
        return GPW_OK;
    }
    const Profile &P = profiles[g.prof];
    // re-encode the profile in the reference's sphere_bounds layout
    int n = 0;
    auto put = [&](int v) {
        if (sphere_bounds && n < cap) sphere_bounds[n] = v;
        n++;
    };
    const int kgmin = g.lb[2];
    put(kgmin);
    for (int kg = kgmin; kg <= 0; kg++) {
        int jn = -1;
        for (int jj = 0; jj <= P.n[1]; jj++)
            if (P.T[(size_t)(-kg) * (P.n[1] + 1) + jj] >= 0) jn = jj;
        put(-jn);
        for (int jg = -jn; jg <= 0; jg++) put(-P.T[(size_t)(-kg) * (P.n[1] + 1) + (-jg)]);
    }
    *nspheres = n;
    return n <= cap ? GPW_OK : fail(GPW_ERR_ARG, "sphere_bounds buffer too small");
}

gpw_status gpw_tasklist_destroy(gpw_tasklist *tl) {
    delete tl;
    return GPW_OK;
}

gpw_status gpw_tasklist_stats_get(const gpw_tasklist *tl, gpw_tasklist_stats *o) {
    if (!tl || !o) return fail(GPW_ERR_ARG, "NULL argument");
    o->npairs = tl->npairs;
    o->nsub = (long long)tl->subs_h.size();
    o->ntiles = tl->ntiles;
    o->nentries = tl->nentries;
    o->cube_points = tl->cube_points;
    o->tile_points = tl->tile_points;
    o->lp_max = tl->lp_max;
    o->tile[0] = TX;
    o->tile[1] = TY;
    o->tile[2] = TZ;
    o->splits = tl->splits;
    o->flops_collocate = tl->flops_collocate;
    o->flops_integrate[0] = tl->flops_integrate[0];
    o->flops_integrate[1] = tl->flops_integrate[1];
    for (int i = 0; i < 4; i++) o->plan_ms[i] = tl->plan_ms[i];
    o->csmall_pairs = tl->csmall_cnt[0] + tl->csmall_cnt[1] + tl->csmall_cnt[2] + tl->csmall_cnt[3];
    o->csmall_points = tl->sval_total;
    long long ce = 0;
    for (int c = 0; c < NCLS; c++) ce += tl->class_entries[0][c];
    o->collocate_tile_points = ce * (long long)(TX * TY * TZ);
    for (int c = 0; c < NCLS; c++) {
        o->csplits[c] = tl->csplit[c];
        o->isplits[0][c] = tl->isplit[0][c];
        o->isplits[1][c] = tl->isplit[1][c];
    }
    return GPW_OK;
}

// largest collocate split over the classes that have work
static int collocate_splits_max(const gpw_tasklist *tl) {
    int smax = 1;
    for (int c = 0; c < NCLS; c++)
        if (tl->class_entries[0][c]) smax = std::max(smax, tl->csplit[c]);
    return smax;
}

// Orthorhombic lists whose work is all in classes 0 and 1, both split: the
// two classes run in one launch, each into partial grids of its own
static bool collocate_fused(const gpw_tasklist *tl) {
    return !tl->tg.general && tl->class_entries[0][0] && tl->class_entries[0][1] &&
           !tl->class_entries[0][2] && !tl->class_entries[0][3] && tl->csplit[0] > 1 && tl->csplit[1] > 1;
}

// partial grids the collocate launches write (0: they add into the grid)
static int collocate_partials(const gpw_tasklist *tl) {
    if (collocate_fused(tl)) return tl->csplit[0] + tl->csplit[1];
    const int smax = collocate_splits_max(tl);
    return smax > 1 ? smax : 0;
}

// Collocate tile launches of every class. With smax == 1 the kernels add into
// the grid; otherwise into smax partial grids (reduced afterwards): the class
// with the most splits runs first and writes all smax partials (workgroups
// without items write zeros), the others add into theirs. Fused classes 0
// and 1 (collocate_fused, when the partial grids are there) write theirs.
static hipError_t collocate_classes(gpw_tasklist *tl, double *grid, int smax, bool fused, hipStream_t st) {
    const int nt1 = tl->ntiles + 1;
    if (fused)
        return launch_collocate_tiles_84(tl->tg, tl->d_recs + tl->list_off[0][1], tl->d_tile_start + (size_t)nt1,
                                         tl->csplit[1], tl->d_recs + tl->list_off[0][0], tl->d_tile_start,
                                         tl->csplit[0], tl->d_coef, tl->ntiles, tl->d_cpartial, st);
    int order[NCLS] = {0, 1, 2, 3};
    std::stable_sort(order, order + NCLS, [&](int a, int b) { return tl->csplit[a] > tl->csplit[b]; });
    bool first = true;
    for (int k = 0; k < NCLS; k++) {
        const int c = order[k];
        if (tl->class_entries[0][c] == 0) continue;
        const int *start = tl->d_tile_start + (size_t)c * nt1;
        const EntryRec *recs = tl->d_recs + tl->list_off[0][c];
        hipError_t e;
        if (smax == 1) {
            e = launch_collocate_tiles(c, tl->tg, recs, start, tl->d_coef, tl->ntiles, 1, grid, 0, st);
        } else {
            e = launch_collocate_tiles(c, tl->tg, recs, start, tl->d_coef, tl->ntiles, tl->csplit[c],
                                       tl->d_cpartial, first ? 1 : 2, st);
            first = false;
        }
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// autotuner: workgroups per tile for every (kernel, class), timed on the
// device on scratch inputs (run times depend on the index sets, not on values)

gpw_status gpw_tasklist_autotune(gpw_tasklist *tl, int max_splits, int reps, void *stream,
                                 double *ms_before, double *ms_after) {
    if (!tl) return fail(GPW_ERR_ARG, "NULL argument");
    if (tl->npairs == 0) return GPW_OK;
    const long long ng = (long long)tl->tg.nx * tl->tg.ny * tl->tg.nz;
    if (max_splits <= 0)  // automatic: 8, more on grids of few tiles (>= 4096 workgroups)
        max_splits = std::max(8, std::min(128, 4096 / std::max(tl->ntiles, 1)));
    max_splits = std::min(max_splits, 128);
    // partial grids of collocate stay within 2 GiB
    max_splits = (int)std::max<long long>(1, std::min<long long>(max_splits, (2LL << 30) / (8 * ng)));
    reps = std::min(std::max(reps, 1), 20);
    hipStream_t st = (hipStream_t)stream;
    std::vector<int> cand;
    for (int s : {1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128})
        if (s <= max_splits) cand.push_back(s);
    const int cmax = cand.back();
    const int nt1 = tl->ntiles + 1;
    double *grid = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (grid) (void)hipFree(grid);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
#define AT_TRY(x)                                                          \
    do {                                                                   \
        hipError_t err_ = (x);                                             \
        if (err_ != hipSuccess) {                                          \
            cleanup();                                                     \
            return fail(GPW_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(err_)); \
        }                                                                  \
    } while (0)
    AT_TRY(hipMalloc((void **)&grid, sizeof(double) * ng));
    if (tl->cpart_cap < cmax) {
        if (tl->d_cpartial) AT_TRY(hipFree(tl->d_cpartial));
        tl->d_cpartial = nullptr;
        tl->cpart_cap = 0;
        AT_TRY(hipMalloc((void **)&tl->d_cpartial, sizeof(double) * ng * cmax));
        tl->cpart_cap = cmax;
    }
    AT_TRY(hipEventCreate(&e0));
    AT_TRY(hipEventCreate(&e1));
    AT_TRY(launch_fill(grid, ng, 0.0, st));
    AT_TRY(launch_fill(tl->d_coef, std::max<long long>(tl->coef_total, 1), 0.0, st));
    // best-of-reps time of one launch sequence, after one warm-up
    auto timed = [&](auto &&launch, float &ms) -> hipError_t {
        hipError_t e = launch();
        if (e != hipSuccess) return e;
        ms = 1e30f;
        for (int r = 0; r < reps; r++) {
            if ((e = hipEventRecord(e0, st)) != hipSuccess) return e;
            if ((e = launch()) != hipSuccess) return e;
            if ((e = hipEventRecord(e1, st)) != hipSuccess) return e;
            if ((e = hipEventSynchronize(e1)) != hipSuccess) return e;
            float t = 0.0f;
            if ((e = hipEventElapsedTime(&t, e0, e1)) != hipSuccess) return e;
            ms = std::min(ms, t);
        }
        return hipSuccess;
    };
    const int nc = (int)cand.size();
    // collocate: per class and split, plus the partial-grid reduction per split
    std::vector<float> tc((size_t)NCLS * nc, 0.0f), tr(nc, 0.0f);
    for (int c = 0; c < NCLS; c++) {
        if (tl->class_entries[0][c] == 0) continue;
        const int *start = tl->d_tile_start + (size_t)c * nt1;
        const EntryRec *recs = tl->d_recs + tl->list_off[0][c];
        for (int i = 0; i < nc; i++) {
            const int sp = cand[i];
            AT_TRY(timed([&] {
                return sp == 1 ? launch_collocate_tiles(c, tl->tg, recs, start, tl->d_coef, tl->ntiles, 1,
                                                        grid, 0, st)
                               : launch_collocate_tiles(c, tl->tg, recs, start, tl->d_coef, tl->ntiles,
                                                        sp, tl->d_cpartial, 1, st);
            }, tc[(size_t)c * nc + i]));
        }
    }
    for (int i = 1; i < nc; i++)
        AT_TRY(timed([&] { return launch_reduce_partials(tl->d_cpartial, cand[i], ng, grid, st); }, tr[i]));
    // choose the largest split smax, then each class's best split <= smax
    double best = 1e300;
    int bestc[NCLS] = {1, 1, 1, 1};
    auto cost_of = [&](int imax, int *pick) {
        double tot = imax > 0 ? tr[imax] : 0.0;
        for (int c = 0; c < NCLS; c++) {
            pick[c] = cand[imax];
            if (tl->class_entries[0][c] == 0) continue;
            int bi = 0;
            for (int i = 1; i <= imax; i++)
                if (tc[(size_t)c * nc + i] < tc[(size_t)c * nc + bi]) bi = i;
            pick[c] = cand[bi];
            tot += tc[(size_t)c * nc + bi];
        }
        return tot;
    };
    for (int imax = 0; imax < nc; imax++) {
        int pick[NCLS];
        const double tot = cost_of(imax, pick);
        if (tot < best) {
            best = tot;
            std::copy(pick, pick + NCLS, bestc);
        }
    }
    // classes 0 and 1 fused (collocate_fused): each into partial grids of its
    // own, so the reduction sums splits0 + splits1 of them (time taken as
    // proportional to the grids read plus the one written, from the largest
    // measured reduction)
    if (!tl->tg.general && tl->class_entries[0][0] && tl->class_entries[0][1] &&
        !tl->class_entries[0][2] && !tl->class_entries[0][3] && nc > 1) {
        const double tred1 = tr[nc - 1] / (cand[nc - 1] + 1);
        for (int ia = 1; ia < nc; ia++)
            for (int ib = 1; ib < nc; ib++) {
                const double tot = tc[ia] + tc[(size_t)nc + ib] + tred1 * (cand[ia] + cand[ib] + 1);
                if (tot < best) {
                    best = tot;
                    bestc[0] = cand[ia];
                    bestc[1] = cand[ib];
                    bestc[2] = bestc[3] = std::max(cand[ia], cand[ib]);
                }
            }
    }
    // cost of the splits in use before tuning (collocate + integrate)
    double before = 0.0, after = best;
    {
        const int smax0 = collocate_splits_max(tl);
        int i0 = 0;
        while (i0 + 1 < nc && cand[i0] < smax0) i0++;
        before += (smax0 > 1 && cand[i0] == smax0) ? tr[i0] : 0.0;
        for (int c = 0; c < NCLS; c++) {
            if (tl->class_entries[0][c] == 0) continue;
            for (int i = 0; i < nc; i++)
                if (cand[i] == tl->csplit[c]) before += tc[(size_t)c * nc + i];
        }
    }
    std::copy(bestc, bestc + NCLS, tl->csplit);
    // integrate, without and with the force extension
    for (int wf = 0; wf < 2; wf++) {
        const int is = iset(wf);
        for (int c = 0; c < NCLS; c++) {
            if (tl->class_entries[is][c] == 0) continue;
            const int *start = tl->d_tile_start + (size_t)(is * NCLS + c) * nt1;
            const EntryRec *recs = tl->d_recs + tl->list_off[is][c];
            float bt = 1e30f, t0 = 0.0f;
            int bs = 1;
            for (int i = 0; i < nc; i++) {
                float t;
                AT_TRY(timed([&] {
                    return launch_integrate_tiles(c, tl->tg, recs, start, grid, 2 * wf, tl->ntiles, cand[i],
                                                  tl->d_partial, st);
                }, t));
                if (cand[i] == tl->isplit[wf][c]) t0 = t;
                if (t < bt) {
                    bt = t;
                    bs = cand[i];
                }
            }
            if (wf == 0) {
                before += t0;
                after += bt;
            }
            tl->isplit[wf][c] = bs;
        }
    }
#undef AT_TRY
    // keep only the partial grids the chosen collocate splits use
    const int npart = collocate_partials(tl);
    if (tl->cpart_cap != npart) {
        if (tl->d_cpartial) (void)hipFree(tl->d_cpartial);
        tl->d_cpartial = nullptr;
        tl->cpart_cap = 0;
        if (npart > 0) {
            hipError_t e = hipMalloc((void **)&tl->d_cpartial, sizeof(double) * ng * npart);
            if (e != hipSuccess) {
                cleanup();
                return fail(GPW_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
            }
            tl->cpart_cap = npart;
        }
    }
    cleanup();
    if (ms_before) *ms_before = before;
    if (ms_after) *ms_after = after;
    return GPW_OK;
}

static gpw_status run_coef(gpw_tasklist *tl, const double *pab, int pair0, int n, hipStream_t st) {
    if (n <= 0) return GPW_OK;
    if (pair0 == 0 && n == tl->npairs) {  // the whole list: one launch per size group
        for (const auto &g : tl->groups)
            HIP_TRY(launch_coef(tl->d_pairs, tl->d_order, g.start, g.count, tl->compute_tau, g.la_max,
                                g.lb_max, g.nab, pab, tl->d_coef, tl->lat(), tl->d_hab_ws,
                                tl->hab_ws_stride, tl->hab_ws_blocks, st));
        return GPW_OK;
    }
    HIP_TRY(launch_coef(tl->d_pairs, nullptr, pair0, n, tl->compute_tau, tl->la_max, tl->lb_max, tl->nab_max, pab,
                        tl->d_coef, tl->lat(), tl->d_hab_ws, tl->hab_ws_stride,
                        tl->hab_ws_blocks, st));
    return GPW_OK;
}

gpw_status gpw_tasklist_collocate(gpw_tasklist *tl, const double *pab, double *grid, void *stream) {
    if (!tl || !pab || !grid) return fail(GPW_ERR_ARG, "NULL argument");
    if (tl->npairs == 0) return GPW_OK;
    hipStream_t st = (hipStream_t)stream;
    {
        PhaseTimer ph(tl, st, PH_COEF);
        gpw_status s = run_coef(tl, pab, 0, tl->npairs, st);
        if (s) return s;
    }
    const int smax = collocate_splits_max(tl);
    // fused classes when their partial grids are allocated (after autotune)
    const bool fused = collocate_fused(tl) && tl->cpart_cap >= collocate_partials(tl);
    const int nred = fused ? collocate_partials(tl) : (smax > 1 ? smax : 0);
    {
        PhaseTimer ph(tl, st, PH_CTILES);
        HIP_TRY(collocate_classes(tl, grid, smax, fused, st));
        // small spheres: each pair's box into the scratch buffer
        for (int c = 0; c < NSCLS; c++)
            HIP_TRY(launch_collocate_small(c, tl->tg, tl->d_subs, tl->d_csmall + tl->csmall_off[c],
                                           (int)tl->csmall_cnt[c], tl->d_coef, tl->d_prof, tl->d_box_off,
                                           tl->d_sval, st));
    }
    const bool gather = tl->class_entries[3][0] > 0;
    if (nred > 0 || gather) {
        PhaseTimer ph(tl, st, PH_CREDUCE);
        const long long ng = (long long)tl->tg.nx * tl->tg.ny * tl->tg.nz;
        if (nred > 0) HIP_TRY(launch_reduce_partials(tl->d_cpartial, nred, ng, grid, st));
        // then the small boxes, tile by tile in list order
        if (gather)
            HIP_TRY(launch_gather_small(tl->tg, tl->d_recs + tl->list_off[3][0],
                                        tl->d_tile_start + (size_t)3 * NCLS * (tl->ntiles + 1), tl->ntiles,
                                        tl->d_sval, grid, st));
    }
    return GPW_OK;
}

gpw_status gpw_tasklist_integrate(gpw_tasklist *tl, const double *grid, const double *pab,
                                  double *hab, double *forces, void *stream) {
    if (!tl || !grid || !hab) return fail(GPW_ERR_ARG, "NULL argument");
    if (forces && !pab) return fail(GPW_ERR_ARG, "forces need pab");
    if (tl->npairs == 0) return GPW_OK;
    hipStream_t st = (hipStream_t)stream;
    const int wf = forces ? 1 : 0;
    const int lp_add = 2 * wf;
    const long long *off = wf ? tl->d_ent_off2 : tl->d_ent_off0;
    const int nt1 = tl->ntiles + 1;
    const int is = iset(wf);
    {
        PhaseTimer ph(tl, st, PH_ITILES);
        const bool only01 = tl->class_entries[is][2] == 0 && tl->class_entries[is][3] == 0;
        if (!tl->tg.general && only01 && tl->class_entries[is][0] && tl->class_entries[is][1]) {
            // the two light classes together in one launch (class 1 first)
            HIP_TRY(launch_integrate_tiles_84(
                tl->tg, tl->d_recs + tl->list_off[is][1], tl->d_tile_start + (size_t)(is * NCLS + 1) * nt1,
                tl->isplit[wf][1], tl->d_recs + tl->list_off[is][0],
                tl->d_tile_start + (size_t)(is * NCLS + 0) * nt1, tl->isplit[wf][0], grid, lp_add,
                tl->ntiles, tl->d_partial, st));
        } else
        for (int c = 0; c < NCLS; c++) {
            if (tl->class_entries[is][c] == 0) continue;
            const int *start = tl->d_tile_start + (size_t)(is * NCLS + c) * nt1;
            const EntryRec *recs = tl->d_recs + tl->list_off[is][c];
            HIP_TRY(launch_integrate_tiles(c, tl->tg, recs, start, grid, lp_add, tl->ntiles,
                                           tl->isplit[wf][c], tl->d_partial, st));
        }
        // small spheres: one workgroup per pair, moments straight into Mbuf
        for (int c = 0; c < NSCLS; c++)
            HIP_TRY(launch_moments_small(c, tl->tg, tl->d_subs, tl->d_small + tl->small_off[wf][c],
                                         (int)tl->small_cnt[wf][c], tl->d_prof, grid, lp_add,
                                         tl->mstride, tl->d_Mbuf, st));
    }
    {
        PhaseTimer ph(tl, st, PH_ISUM);
        HIP_TRY(launch_sum_entries(tl->d_pairs, tl->npairs, tl->compute_tau, lp_add, tl->d_pe_start,
                                   off, tl->d_partial, tl->mstride,
                                   tl->d_small_mask + (size_t)wf * tl->npairs, tl->d_Mbuf, st));
    }
    {
        PhaseTimer ph(tl, st, PH_HAB);
        for (const auto &g : tl->groups)
            HIP_TRY(launch_hab(tl->d_pairs, tl->d_order, tl->d_segs, g.seg0, g.nseg, tl->compute_tau,
                               wf, g.la_max, g.lb_max, g.nhx, tl->d_Mbuf, tl->mstride, pab, hab, forces,
                               tl->lat(), tl->d_hab_ws, tl->hab_ws_stride,
                               tl->hab_ws_blocks, st));
    }
    return GPW_OK;
}

gpw_status gpw_tasklist_set_timing(gpw_tasklist *tl, int enable) {
    if (!tl) return fail(GPW_ERR_ARG, "NULL argument");
    if (enable && tl->events.empty()) {
        tl->events.resize(NEVENT);
        tl->event_phase.assign(NEVENT / 2, 0);
        for (int i = 0; i < NEVENT; i++) HIP_TRY(hipEventCreate(&tl->events[i]));
    }
    tl->timing = enable != 0;
    tl->events_used = 0;
    return GPW_OK;
}

gpw_status gpw_tasklist_timings(gpw_tasklist *tl, double *ms, int n, int *calls) {
    if (!tl || !ms || n < NPH) return fail(GPW_ERR_ARG, "need room for GPW_NPHASES values");
    for (int i = 0; i < n; i++) ms[i] = 0.0;
    if (calls)
        for (int i = 0; i < NPH; i++) calls[i] = 0;
    if (tl->events_used > 0) HIP_TRY(hipEventSynchronize(tl->events[tl->events_used - 1]));
    for (int k = 0; k < tl->events_used / 2; k++) {
        float t = 0.0f;
        HIP_TRY(hipEventElapsedTime(&t, tl->events[2 * k], tl->events[2 * k + 1]));
        ms[tl->event_phase[k]] += t;
        if (calls) calls[tl->event_phase[k]]++;
    }
    tl->events_used = 0;
    return GPW_OK;
}

gpw_status gpw_tasklist_collocate_plain(gpw_tasklist *tl, const double *pab, double *grid,
                                        void *stream) {
    if (!tl || !pab || !grid) return fail(GPW_ERR_ARG, "NULL argument");
    hipStream_t st = (hipStream_t)stream;
    for (int p = 0; p < tl->npairs; p++) {
        gpw_status s = run_coef(tl, pab, p, 1, st);
        if (s) return s;
        for (int si = tl->pair_sub_start[p]; si < tl->pair_sub_start[p + 1]; si++) {
            const SubDev &sd = tl->subs_h[si];
            HIP_TRY(launch_collocate_plain(tl->tg, tl->d_subs, si, sd.ub[2] - sd.lb[2] + 1,
                                           sd.ub[1] - sd.lb[1] + 1, tl->d_coef, tl->d_prof, grid, st));
        }
    }
    return GPW_OK;
}

gpw_status gpw_tasklist_integrate_plain(gpw_tasklist *tl, const double *grid, const double *pab,
                                        double *hab, double *forces, void *stream) {
    if (!tl || !grid || !hab) return fail(GPW_ERR_ARG, "NULL argument");
    if (forces && !pab) return fail(GPW_ERR_ARG, "forces need pab");
    hipStream_t st = (hipStream_t)stream;
    const int wf = forces ? 1 : 0;
    for (int p = 0; p < tl->npairs; p++) {
        HIP_TRY(launch_fill(tl->d_Mbuf + (long long)p * tl->mstride, tl->mstride, 0.0, st));
        for (int si = tl->pair_sub_start[p]; si < tl->pair_sub_start[p + 1]; si++) {
            const SubDev &sd = tl->subs_h[si];
            HIP_TRY(launch_moments_plain(tl->tg, tl->d_subs, si, sd.ub[2] - sd.lb[2] + 1,
                                         sd.ub[1] - sd.lb[1] + 1, 2 * wf, tl->d_prof, grid,
                                         tl->mstride, tl->d_Mbuf, st));
        }
        const PairDev &pd = tl->pairs_h[p];
        HIP_TRY(launch_hab(tl->d_pairs, nullptr, nullptr, p, 1, tl->compute_tau, wf, pd.la_max, pd.lb_max,
                           hab_cross_entries(pd.la_min, pd.la_max, pd.lb_min, pd.lb_max), tl->d_Mbuf,
                           tl->mstride, pab, hab, forces,
                           tl->lat(), tl->d_hab_ws, tl->hab_ws_stride,
                           tl->hab_ws_blocks, st));
    }
    return GPW_OK;
}

gpw_status gpw_collocate_core(const int *length, int L, const double *co, const double *P, int ldp,
                              double *V, void *stream) {
    if (!length || !co || !P || !V || L < 1 || L > LPMAX + 1 || ldp < std::max(length[0], std::max(length[1], length[2])))
        return fail(GPW_ERR_ARG, "bad arguments");
    HIP_TRY(launch_collocate_core(length, L, co, P, ldp, V, (hipStream_t)stream));
    return GPW_OK;
}

gpw_status gpw_integrate_core(const int *length, int L, const double *P, int ldp, const double *V,
                              double *Velems, void *stream) {
    if (!length || !P || !V || !Velems || L < 1 || L > LPMAX + 1 ||
        ldp < std::max(length[0], std::max(length[1], length[2])))
        return fail(GPW_ERR_ARG, "bad arguments");
    HIP_TRY(launch_integrate_core(length, L, P, ldp, V, Velems, (hipStream_t)stream));
    return GPW_OK;
}

// host-array variants: stage through device memory
static gpw_status core_host(bool collocate, const int *length, int L, const double *co,
                            const double *P, int ldp, const double *V, double *out) {
    if (!length || !P || !out || (collocate ? !co : !V) || L < 1 || L > LPMAX + 1 ||
        length[0] < 0 || length[1] < 0 || length[2] < 0 ||
        ldp < std::max(length[0], std::max(length[1], length[2])))
        return fail(GPW_ERR_ARG, "bad arguments");
    const size_t nv = (size_t)length[0] * length[1] * length[2];
    const size_t np = (size_t)3 * L * ldp, nc = (size_t)L * L * L;
    double *d_P = nullptr, *d_a = nullptr, *d_b = nullptr;
    gpw_status rc = GPW_OK;
    hipError_t e;
    const size_t na = collocate ? nc : nv, nb = collocate ? nv : nc;
    if ((e = hipMalloc((void **)&d_P, np * sizeof(double))) != hipSuccess ||
        (e = hipMalloc((void **)&d_a, std::max<size_t>(na, 1) * sizeof(double))) != hipSuccess ||
        (e = hipMalloc((void **)&d_b, std::max<size_t>(nb, 1) * sizeof(double))) != hipSuccess) {
        rc = fail(GPW_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    } else if ((e = hipMemcpy(d_P, P, np * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess ||
               (e = hipMemcpy(d_a, collocate ? co : V, na * sizeof(double), hipMemcpyHostToDevice)) !=
                   hipSuccess) {
        rc = fail(GPW_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    } else {
        rc = collocate ? gpw_collocate_core(length, L, d_a, d_P, ldp, d_b, nullptr)
                       : gpw_integrate_core(length, L, d_P, ldp, d_a, d_b, nullptr);
        if (rc == GPW_OK &&
            (e = hipMemcpy(out, d_b, nb * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
            rc = fail(GPW_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    }
    for (double *p : {d_P, d_a, d_b})
        if (p) (void)hipFree(p);
    return rc;
}

gpw_status gpw_collocate_core_host(const int *length, int L, const double *co, const double *P,
                                   int ldp, double *V) {
    return core_host(true, length, L, co, P, ldp, nullptr, V);
}

gpw_status gpw_integrate_core_host(const int *length, int L, const double *P, int ldp,
                                   const double *V, double *Velems) {
    return core_host(false, length, L, nullptr, P, ldp, V, Velems);
}

}  // extern "C"
