// gpw_host.hpp -- host-side helpers shared by the task-list planner
// (gpw_plan.cpp) and the single-task engine (gpw_single.cpp): pair set-up,
// cube geometry, sub-box split, lattice transforms, error reporting, and the
// kernel launchers of gpw_kernels.hip. Both host files are compiled with
// -ffp-contract=off, so the geometry rounds as written.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "gpw_internal.hpp"

namespace gpw {

// records the message for gpw_last_error() and returns code
gpw_status fail(gpw_status code, const std::string &msg);

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            return ::gpw::fail(GPW_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

int imod_h(int a, int m);

// sphere profile: T[kk][jj] = -igmin (>= 0) or -1 for an empty row
struct Profile {
    int n[3];              // table covers fold offsets 0..n[d]
    std::vector<int> T;    // (n[2]+1) x (n[1]+1)
    double R2 = 0.0;       // squared sphere radius when built from a radius
};

// cube of one pair: centre, bounds and its sphere profile (or lattice box)
struct PairGeom {
    int cc[3];
    double roff[3];  // general cells: lattice centre gp
    int lb[3], ub[3];
    int prof;        // index into profile list, -1 none
    double R2 = 0.0; // general cells: radius^2
};

// lattice transforms of a general cell in the padded row layout of
// LatticeTab (see lattice_tables)
struct LatticeHost {
    std::vector<int> col;
    std::vector<double> val;
    std::vector<LatBlk> blk;
};

void profile_from_radius(double radius, const double dr[3], Profile &P, int lb[3], int ub[3]);
gpw_status profile_from_bounds(const int lb[3], const int ub[3], int nspheres, const int *sb,
                               Profile &P);
bool is_ortho(const gpw_grid_desc &g);
gpw_status check_pair(const gpw_pair_desc &p, int tau);
gpw_status check_grid(const gpw_grid_desc *g);
PairDev make_pair_dev(const gpw_pair_desc &p, int tau);
void cube_centre(const gpw_grid_desc &g, const double rp[3], int cc[3], double roff[3]);
void general_geometry(const gpw_grid_desc &g, const double rp[3], double radius, PairGeom &geo);
gpw_status make_subs(const gpw_grid_desc &g, const PairDev &pd, int pair_index, const PairGeom &geo,
                     int prof_off, int prof_stride, std::vector<SubDev> &out);
void multinomial_matrix(const double dh[3][3], int lp, double *T);
// lattice transforms of a general cell: all degrees lp <= lpm, or only the
// listed ones
// triangular steps (LatticeTab::stepped): h = {a0, b0, c0, b1, c1, c2}
bool stepped_cell(const double dh[3][3], double h[6]);
gpw_status lattice_tables(const double dh[3][3], int lpm, LatticeHost &out);
gpw_status lattice_tables_some(const double dh[3][3], const int *lps, int n, LatticeHost &out);

// tile-kernel launchers per cell kind (GEN: general cell), defined in
// gpw_tiles_colloc.hpp / gpw_tiles_integ.hpp and instantiated in the
// gpw_tiles_{colloc,integ}_{ortho,general}.hip translation units
template <bool GEN>
hipError_t launch_collocate_tiles_t(int cls, const TileGrid &tg, const EntryRec *recs,
                                    const int *tile_start, const double *coef, int ntiles, int splits,
                                    double *out, int out_mode, hipStream_t st);
template <bool GEN>
hipError_t launch_integrate_tiles_t(int cls, const TileGrid &tg, const EntryRec *recs,
                                    const int *tile_start, const double *grid, int lp_add,
                                    int ntiles, int splits, double *partial, hipStream_t st);

// kernel launchers (gpw_kernels.hip, gpw_tiles_integ_ortho.hip)
hipError_t launch_collocate_tiles(int cls, const TileGrid &tg, const EntryRec *recs,
                                  const int *tile_start, const double *coef, int ntiles, int splits,
                                  double *out, int out_mode, hipStream_t st);
hipError_t launch_reduce_partials(const double *partial, int splits, long long ng, double *grid,
                                  hipStream_t st);
hipError_t launch_integrate_tiles(int cls, const TileGrid &tg, const EntryRec *recs,
                                  const int *tile_start, const double *grid, int lp_add,
                                  int ntiles, int splits, double *partial, hipStream_t st);
hipError_t launch_collocate_tiles_84(const TileGrid &tg, const EntryRec *recs8, const int *start8,
                                     int splits8, const EntryRec *recs4, const int *start4, int splits4,
                                     const double *coef, int ntiles, double *partial, hipStream_t st);
hipError_t launch_integrate_tiles_84(const TileGrid &tg, const EntryRec *recs8, const int *start8,
                                     int splits8, const EntryRec *recs4, const int *start4,
                                     int splits4, const double *grid, int lp_add, int ntiles,
                                     double *partial, hipStream_t st);
// nab: the most (a, b) entries of the block shells la_min..la_max x
// lb_min..lb_max among the launched pairs (LDS of the one-pair-per-wave path)
int block_shell_entries(int la_min, int la_max, int lb_min, int lb_max);
hipError_t launch_coef(const PairDev *pairs, const int *order, int pair0, int npairs, int tau,
                       int lam_max, int lbm_max, int nab, const double *pab, double *coef,
                       const LatticeTab &lt, double *ws,
                       long long ws_stride, int ws_blocks, hipStream_t st);
size_t coef_smem_bytes(int lam, int lbm, bool lds_blocks);
size_t coef_ws_doubles(int lam, int lbm);
bool coef_needs_workspace(int lam, int lbm);
// nhx: the most hab_cross_entries among the launched pairs (forces without tau)
hipError_t launch_hab(const PairDev *pairs, const int *order, const int *segs, int seg0, int nseg,
                      int tau, int with_forces, int la_max, int lb_max, int nhx, const double *Mbuf,
                      int mstride, const double *pab, double *hab, double *forces,
                      const LatticeTab &lt, double *ws,
                      long long ws_stride, int ws_blocks, hipStream_t st);
size_t hab_ws_need(int la_max, int lb_max, int tau, int nhx);
hipError_t launch_moments_small(int cls, const TileGrid &tg, const SubDev *subs, const int *list,
                                int n, const int *prof, const double *grid, int lp_add, int mstride,
                                double *Mbuf, hipStream_t st);
hipError_t launch_sum_entries(const PairDev *pairs, int npairs, int tau, int lp_add,
                              const int *pe_start, const long long *ent_off,
                              const double *partial, int mstride, const unsigned char *skip,
                              double *Mbuf, hipStream_t st);
hipError_t launch_fill(double *p, long long n, double v, hipStream_t st);
hipError_t launch_collocate_plain(const TileGrid &tg, const SubDev *subs, int si, int nz_box,
                                  int ny_box, const double *coef, const int *prof, double *grid,
                                  hipStream_t st);
hipError_t launch_moments_plain(const TileGrid &tg, const SubDev *subs, int si, int nz_box,
                                int ny_box, int lp_add, const int *prof, const double *grid,
                                int mstride, double *Mbuf, hipStream_t st);
hipError_t launch_collocate_core(const int *length, int L, const double *co, const double *P,
                                 int ldp, double *V, hipStream_t st);
hipError_t launch_integrate_core(const int *length, int L, const double *P, int ldp,
                                 const double *V, double *Velems, hipStream_t st);
hipError_t launch_fold_tables(int L, const int *length, const int *center, const int *period,
                              const double *unfolded, int lu, int ldu, double *folded, int ldp,
                              int *mask, int ldm, hipStream_t st);
hipError_t launch_collocate_folded(int L, const double *co, const double *P, int ldp, const int *mask,
                                   int ldm, const int *gs, const int *lc, double *density,
                                   hipStream_t st);
hipError_t launch_integrate_folded(int L, const double *P, int ldp, const int *mask, int ldm,
                                   const int *gs, const int *lc, const double *V, double *co,
                                   hipStream_t st);

// device planner of the tile work lists (gpw_kernels.hip)
hipError_t launch_plan_count(const SubDev *subs, int nsub, const int *prof, const TileGrid &tg, int *counts,
                             unsigned long long *work, hipStream_t st);
hipError_t launch_plan_entries(const SubDev *subs, int nsub, const long long *ent_start, const long long *base0,
                               const long long *base2, const int *prof, const TileGrid &tg, int ntiles,
                               const unsigned char *small_mask, int npairs, int ne, int *ent_sub,
                               long long *off0, long long *off2, unsigned *keys, int *vals, int *hist,
                               hipStream_t st);
hipError_t launch_plan_records(const SubDev *subs, const int *ent_sub, const long long *off,
                               const long long *sub_off, const int *sorted_e, int n, EntryRec *recs,
                               hipStream_t st);
hipError_t launch_collocate_small(int cls, const TileGrid &tg, const SubDev *subs, const int *list, int n,
                                  const double *coef, const int *prof, const long long *box_off, double *sval,
                                  hipStream_t st);
hipError_t launch_gather_small(const TileGrid &tg, const EntryRec *recs, const int *tile_start, int ntiles,
                               const double *sval, double *grid, hipStream_t st);
hipError_t sort_pairs_u32(void *temp, size_t *temp_bytes, const unsigned *keys_in, unsigned *keys_out,
                          const int *vals_in, int *vals_out, int n, int end_bit, hipStream_t st);

// the calling thread's device workspace on the current device (grow only,
// kept until the process ends; see gpw_single.cpp)
gpw_status thread_workspace(size_t need, hipStream_t st, char **out);

}  // namespace gpw
