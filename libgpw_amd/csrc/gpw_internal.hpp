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
constexpr size_t HAB_LDS_MAX = 64 * 1024;
constexpr int HAB_WS_BLOCKS = 4096;
constexpr int TR_NT = 256;
// pairs whose cube is one sub-box of at most SMALL_NX x SMALL_NY x SMALL_NZ
// points (general cells; SMALL_NX x SMALL_NYZ_ORTHO^2 on orthorhombic ones,
// where the tile kernels are faster) are integrated by k_moments_small (one
// workgroup per pair) instead of the tiles; limits chosen by measurement
constexpr int SMALL_NX = 64, SMALL_NY = 48, SMALL_NZ = 48, SMALL_NYZ_ORTHO = 32, SMALL_LMAX = 12;

// work-item classes by polynomial length L = lp + 1 (one tile-kernel
// instantiation each, so registers are sized for the class, not for the
// largest L), and the three tile-list sets: 0 collocate, 1 integrate with
// the force extension (lp + 2), 2 integrate without
constexpr int NCLS = 4, NSET = 3;
__host__ __device__ constexpr int class_of(int lp) {
    return lp + 1 <= 4 ? 0 : lp + 1 <= 8 ? 1 : lp + 1 <= 12 ? 2 : 3;
}

__host__ __device__ constexpr int ncoset(int l) {
    return l < 0 ? 0 : ((l + 1) * (l + 2) * (l + 3)) / 6;
}
__host__ __device__ constexpr int ncoef2(int l) { return ((l + 1) * (l + 2)) / 2; }

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

// Lattice transforms of a general cell in CSR form, for every lp <= LPMAX and
// both directions (dir 0: coef_ijk = T coef_xyz, dir 1: M_xyz = T^T M_ijk).
// Rows of (lp, dir) start at rp + rp_off[2 * lp + dir]; their entries are
// absolute indices into col / val. T is block-diagonal by degree and, for
// cells with axis-aligned steps, sparse within the blocks.
struct LatticeCSR {
    const int *rp = nullptr;
    const int *col = nullptr;
    const double *val = nullptr;
    const long long *rp_off = nullptr;
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
    int pad;
};
static_assert(sizeof(EntryRec) == 96, "EntryRec layout");

// Tile-kernel launch description.
struct TileGrid {
    int nx, ny, nz;          // local grid dims
    int ntx, nty, ntz;       // tiles per axis
    int period[3];           // npts on periodic axes, BIG otherwise
    int general;             // non-orthorhombic cell (lattice-coordinate path)
    double dr[3];
    double dh[3][3];         // step vectors (rows), used by the general path
};

}  // namespace gpw
