// gpw_internal.hpp -- structures shared by the host planner and the HIP kernels.
#pragma once

#include <cstdint>

#include "libgpw_amd.h"

namespace gpw {

constexpr int LMAX = GPW_LMAX;            // per centre, after tau / force shifts
constexpr int LPMAX = 2 * LMAX;           // product angular momentum
constexpr int NCOSET_LMAX = (LMAX + 1) * (LMAX + 2) * (LMAX + 3) / 6;
constexpr int BIG = 0x3fffffff;           // "outside" marker for offsets

// Tile of the grid owned by one workgroup in the tile kernels: TX points
// along x per row, TY x TZ rows, one row per thread.
constexpr int TX = 32, TY = 16, TZ = 16;
constexpr int NROW = TY * TZ;             // 256 threads per workgroup
// k_coef / k_hab keep their per-pair blocks in LDS up to these sizes, else in
// a global workspace of HAB_WS_BLOCKS slices (one per resident workgroup);
// TR_NT: their largest workgroup
constexpr size_t COEF_LDS_MAX = 24 * 1024;
constexpr size_t HAB_LDS_MAX = 28 * 1024;
constexpr int HAB_WS_BLOCKS = 4096;
constexpr int TR_NT = 256;
// the one-pair-per-wave transforms (tau = 0) take blocks up to this LDS size
constexpr size_t WAVE_LDS_MAX = 64 * 1024;
// pairs whose cube is one sub-box of at most SMALL_NX x SMALL_NY x SMALL_NZ
// points (general cells; SMALL_NX x SMALL_NYZ_ORTHO^2 on orthorhombic ones,
// where the tile kernels are faster) are integrated by k_moments_small (one
// workgroup per pair) instead of the tiles; limits chosen by measurement
constexpr int SMALL_NX = 64, SMALL_NY = 48, SMALL_NZ = 48, SMALL_NYZ_ORTHO = 32, SMALL_LMAX = 17;

// collocate of small spheres: pairs whose cube is one sub-box of at most
// CSMALL_N points per axis (general cells; CSMALL_N_ORTHO on orthorhombic
// ones, 0: none) with lp + 1 <= SMALL_LMAX are computed per pair into a
// scratch box and added to the grid tile by tile (k_collocate_small,
// k_gather_small) instead of the collocate tile kernels; limits measured
constexpr int CSMALL_N = 16, CSMALL_N_ORTHO = 0;

// work-item classes by polynomial length L = lp + 1 (one tile-kernel
// instantiation each, so registers are sized for the class, not for the
// largest L), and the four tile-list sets: 0 collocate, 1 integrate with
// the force extension (lp + 2), 2 integrate without, 3 the small-sphere
// collocate boxes per tile (class 0 only)
constexpr int NCLS = 4, NSET = 4;
__host__ __device__ constexpr int class_of(int lp) {
    return lp + 1 <= 4 ? 0 : lp + 1 <= 8 ? 1 : lp + 1 <= 12 ? 2 : 3;
}

__host__ __device__ constexpr int ncoset(int l) {
    return l < 0 ? 0 : ((l + 1) * (l + 2) * (l + 3)) / 6;
}
__host__ __device__ constexpr int ncoef2(int l) { return ((l + 1) * (l + 2)) / 2; }
// sum of ncoset(k) over k <= l (0 for l < 0)
__host__ __device__ constexpr int c4(int l) {
    return l < 0 ? 0 : ((l + 1) * (l + 2) * (l + 3) * (l + 4)) / 24;
}

// heff entries k_hab keeps with forces and without tau: the block's own b
// shells with the a shells la_min - 1 .. la_max + 1, then the b shells
// lb_min - 1 and lb_max + 1 with the block's own a shells
__host__ __device__ constexpr int hab_cross_entries(int la_min, int la_max, int lb_min, int lb_max) {
    const int nax = ncoset(la_max + 1) - ncoset((la_min > 0 ? la_min - 1 : 0) - 1);
    const int nbx = ncoset(lb_max + 1) - ncoset((lb_min > 0 ? lb_min - 1 : 0) - 1);
    const int nam = ncoset(la_max) - ncoset(la_min - 1), nbm = ncoset(lb_max) - ncoset(lb_min - 1);
    return nbm * nax + (nbx - nbm) * nam;
}

// CP2K Cartesian order (reference src/lib/grid_base_ref_c.cc:39-46)
__host__ __device__ constexpr int coset(int lx, int ly, int lz) {
    const int l = lx + ly + lz;
    return l == 0 ? 0 : ncoset(l - 1) + ((l - lx) * (l - lx + 1)) / 2 + lz;
}

// packed (lzp, lyp, lxp), lzp outermost, lxp fastest, lxp+lyp+lzp <= lp
// (reference src/lib/grid_base_ref_c.cc:250-257)
__host__ __device__ constexpr int cidx(int lp, int lzp, int lyp, int lxp) {
    // sum_{z<lzp} C2(lp-z) = C3(lp) - C3(lp-lzp)
    const int m = lp - lzp;
    return ncoset(lp) - ncoset(m) + lyp * (m + 1) - (lyp * (lyp - 1)) / 2 + lxp;
}
// packed (lyp, lxp) triangle over lp: lyp outer
__host__ __device__ constexpr int c2idx(int lp, int lyp, int lxp) {
    return lyp * (lp + 1) - (lyp * (lyp - 1)) / 2 + lxp;
}

// Lattice transforms of a general cell for every lp <= LPMAX and both
// directions (dir 0: coef_ijk = T coef_xyz, dir 1: M_xyz = T^T M_ijk). T is
// block-diagonal by degree and, for cells with axis-aligned steps, sparse
// within the blocks. Each (lp, dir) is a block of n = ncoset(lp) rows padded
// to a common width: entry k of row r is col / val[base + k * n + r], the
// non-zeros of a row in ascending column order, then (column 0, value 0)
// fill. Consecutive rows are consecutive in memory, so the lanes of a wave
// that own consecutive rows load each entry together, and a row's entries
// are independent loads.
struct LatBlk {
    long long base = 0;
    int width = 0, n = 0;
};
// Cells whose steps are triangular, dh[0] = (a0, 0, 0), dh[1] = (b0, b1, 0),
// dh[2] = (c0, c1, c2) (a along x, b in the xy plane: the usual cell
// orientation), need no tables: the transform runs in three passes from
// h = {a0, b0, c0, b1, c1, c2} (lattice_stepped).
struct LatticeTab {
    const int *col = nullptr;
    const double *val = nullptr;
    const LatBlk *blk = nullptr;  // [2 * lp + dir]; null: stepped or orthorhombic cell
    int stepped = 0;
    double h[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    __host__ __device__ bool general() const { return blk != nullptr || stepped != 0; }
};

// Per-pair parameters on the device (one per primitive pair).
struct PairDev {
    double zeta, zetb, prefactor, zetp;
    double ra[3], rb[3], rp[3];
    int la_max, la_min, lb_max, lb_min;   // as given (before tau shift)
    int lp;                               // product l of the collocated block
    int coef_off;                         // offset of coef (doubles)
    long long block_off;                  // offset of pab/hab block (doubles)
    int ld;                               // ncoset(la_max): block leading dimension
    int pad;
};

// One sub-box of a pair's cube that does not overlap itself under the
// periodic map. Offsets i along axis d lie in [lb[d], ub[d]]; grid index g
// maps back to i = lb + mod(g - base, period). On orthorhombic grids i is the
// offset from the cube centre and the displacement is i*dr - roff; on general
// grids i is the absolute lattice index and roff holds the centre in lattice
// coordinates (displacement i - roff).
struct SubDev {
    double roff[3];   // ortho: displacement X = i*dr - roff
    double zetp;
    int base[3];
    int period[3];
    int lb[3], ub[3];
    int lp;
    int pair;
    int coef_off;
    int prof_off;     // offset of the sphere profile (ints), -1: full box
    int prof_stride;  // row stride of the profile (n_y + 1)
    int pad;
    double R2;        // general cells: squared radius of the row bounds
};

// One (sub-box, tile) work item as the tile kernels read it: stored in the
// order of the per-tile work lists so a workgroup streams its records.
struct EntryRec {
    double roff[3];      // displacement of offset i along d: i*dr[d] - roff[d]
    double zetp;
    double R2;           // ortho: (imr * min spacing)^2; general: radius^2
    long long ent_off;   // offset of the item's partial moments (integrate)
    int base[3];         // grid index g -> offset i = lb + mod(g - base, period)
                         // (general cells: roff is the lattice centre gp)
    int lb[3], ub[3];
    int lp;              // product l (without the force extension)
    int coef_off;
    int prof_off;        // orthorhombic: the sub-box's sphere profile (see SubDev)
    int prof_stride;
    int pad;
};
static_assert(sizeof(EntryRec) == 104, "EntryRec layout");

// Tile-kernel launch description.
struct TileGrid {
    int nx, ny, nz;          // local grid dims
    int ntx, nty, ntz;       // tiles per axis
    int period[3];           // npts on periodic axes, BIG otherwise
    int general;             // non-orthorhombic cell (lattice-coordinate path)
    double dr[3];
    double dh[3][3];         // step vectors (rows), used by the general path
    const int *prof;         // sphere profiles of the sub-boxes (device; orthorhombic)
};

// a mod m in [0, m) (m > 0)
__host__ __device__ inline int imod_hd(int a, int m) {
    const int r = a % m;
    return r < 0 ? r + m : r;
}

// Offsets of sub-box s along axis d covered by the grid indices g0..g1 of one
// tile: at most two ascending runs a[r]..b[r] (empty when a[r] > b[r]), the
// second one after a periodic wrap, each clipped to [lb, ub].
__host__ __device__ inline void tile_offset_runs(const SubDev &s, int d, int g0, int g1, int (&a)[2],
                                                 int (&b)[2]) {
    const int P = s.period[d];
    const int s0 = imod_hd(g0 - s.base[d], P);
    const int last = s0 + (g1 - g0);  // < P + TX: no overflow (P <= BIG)
    a[0] = s.lb[d] + s0;
    b[0] = s.lb[d] + (last < P ? last : P - 1);
    a[1] = s.lb[d];
    b[1] = last >= P ? s.lb[d] + (last - P) : s.lb[d] - 1;
    for (int r = 0; r < 2; r++) b[r] = b[r] < s.ub[d] ? b[r] : s.ub[d];
}

// Whether the sphere of sub-box s has a point in tile (tx, ty, tz): a (sub-box,
// tile) pair without one is no work item. row(j, k, lo, hi) gives the x
// offsets lo..hi of row (j, k) inside the sphere (false: none) by the same
// rule the tile kernels apply point by point, so host and device planners
// and the kernels agree on every item.
template <class Row>
__host__ __device__ inline bool item_has_points(const SubDev &s, const TileGrid &tg, int tx, int ty, int tz,
                                                Row row) {
    int xa[2], xb[2];
    const int gx1 = (tx + 1) * TX < tg.nx ? (tx + 1) * TX : tg.nx;
    tile_offset_runs(s, 0, tx * TX, gx1 - 1, xa, xb);
    if (xa[0] > xb[0] && xa[1] > xb[1]) return false;
    const int gz1 = (tz + 1) * TZ < tg.nz ? (tz + 1) * TZ : tg.nz;
    const int gy1 = (ty + 1) * TY < tg.ny ? (ty + 1) * TY : tg.ny;
    for (int gz = tz * TZ; gz < gz1; gz++) {
        const int k = s.lb[2] + imod_hd(gz - s.base[2], s.period[2]);
        if (k > s.ub[2]) continue;
        for (int gy = ty * TY; gy < gy1; gy++) {
            const int j = s.lb[1] + imod_hd(gy - s.base[1], s.period[1]);
            if (j > s.ub[1]) continue;
            int lo, hi;
            if (!row(j, k, lo, hi)) continue;
            for (int r = 0; r < 2; r++)
                if ((lo > xa[r] ? lo : xa[r]) <= (hi < xb[r] ? hi : xb[r])) return true;
        }
    }
    return false;
}

}  // namespace gpw
