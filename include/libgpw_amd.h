/*
 * libgpw_amd.h -- C ABI of the MI355X collocate/integrate library.
 *
 * Plain C types only (no STL, no torch). Every entry point returns a
 * gpw_status; GPW_OK is 0. Device pointers are HIP device memory on the
 * current device; `stream` is a hipStream_t passed as void* (NULL = default
 * stream). Grids are row-major (z, y, x), x fastest, FP64, and are always
 * ACCUMULATED into (+=), as in the reference.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   gpw_collocate_pgf_product_rspace  <- grid_collocate_pgf_product_rspace,
 *                                         src/lib/grid_base_ref_c.h:40-67
 *   gpw_integrate_pgf_product_rspace  <- no counterpart in the reference: the
 *                                         adjoint of the above with the same
 *                                         argument list (integrate wiring
 *                                         sketched at src/lib/collocate.cc:963-1113)
 *   gpw_tasklist_*                    <- batched form of the above; batching
 *                                         drafts at src/lib/gaussian.hpp:1-114 and
 *                                         src/lib/collocate_gpu.cc:1-124
 *   gpw_collocate_core / gpw_integrate_core
 *                                     <- collocate_core<T> / integrate_core<T>,
 *                                         src/lib/collocate.cc:543 / :438
 *   gpw_collocate_folded / gpw_integrate_folded
 *                                     <- collocate<T> / integrate<T>,
 *                                         src/lib/collocate.cc:681-745 / :626-679
 *   gpw_fold_tables / gpw_nonzero_values
 *                                     <- compute_folded_polynomials_all_powers /
 *                                         compute_nonzero_values,
 *                                         src/lib/collocate.cc:308-331 / :32-57
 */
#ifndef LIBGPW_AMD_H
#define LIBGPW_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef int gpw_status;

#define GPW_OK 0
#define GPW_ERR_ARG 1        /* bad angular momentum range, sizes or pointers   */
#define GPW_ERR_LMAX 2       /* lmax below the local angular momentum           */
#define GPW_ERR_BOUNDS 3     /* non-periodic cube outside the local grid window  */
#define GPW_ERR_SPHERE 4     /* sphere_bounds inconsistent with lb/ub_cube       */
#define GPW_ERR_HIP 5        /* a HIP runtime call failed                        */
#define GPW_ERR_ALLOC 6      /* device or host allocation failed                 */
#define GPW_ERR_UNSUPPORTED 7

/* largest angular momentum per centre after tau / force shifts */
#define GPW_LMAX 8

/* Grid description shared by every pair of a task list. */
typedef struct gpw_grid_desc {
    int npts[3];        /* global points per axis (x, y, z)                     */
    int ngrid[3];       /* local points per axis; == npts on periodic axes       */
    int lb_grid[3];     /* lower bounds of the local window                      */
    int periodic[3];    /* 1 = periodic axis                                     */
    double dh[3][3];    /* dh[i] = grid step vector of axis i (rows)             */
    double dh_inv[3][3];
} gpw_grid_desc;

/* One primitive Gaussian pair. */
typedef struct gpw_pair_desc {
    int la_max, la_min, lb_max, lb_min;
    double zeta, zetb, rscale;
    double ra[3], rab[3];
    double radius;            /* cube radius; the cube and sphere bounds follow it  */
    long long block_offset;   /* offset (doubles) of the pair's pab / hab block;
                                 the block is ncoset(lb_max) rows (b functions) x
                                 ncoset(la_max) columns (a functions), row-major   */
} gpw_pair_desc;

typedef struct gpw_tasklist gpw_tasklist;

/* Statistics of a task list (algorithmic work for roofline accounting). */
typedef struct gpw_tasklist_stats {
    long long npairs;
    long long nsub;           /* sub-boxes after splitting cubes larger than the period */
    long long ntiles;
    long long nentries;       /* (sub-box, tile) work items                             */
    long long cube_points;    /* grid points inside all pairs' spheres (folded count)   */
    long long tile_points;    /* points swept by the tile kernels                       */
    int lp_max;
    int tile[3];              /* tile shape (x, y, z)                                   */
    int splits;               /* work-list splits per tile given at creation            */
    int csplits[4];           /* workgroups per tile of the collocate kernel, per class */
    int isplits[2][4];        /* same for integrate, without / with forces              */
    /* algorithmic FP64 flops of one pass (2 per FMA) in the contraction order
     * z, y, x: C3(lp) FMAs per z-plane, C2(lp) per (z, y) row and lp + 1 per
     * point of each pair's sphere; integrate with forces runs at lp + 2       */
    double flops_collocate;
    double flops_integrate[2];  /* without / with forces                          */
    /* wall time of gpw_tasklist_create (ms): host geometry and plan of pairs and
     * sub-boxes, their upload, the work items planned on the device, the
     * work buffers (0 for gpw_tasklist_plan_stats)                           */
    double plan_ms[4];
    /* pairs collocated by the small-sphere path (one box per pair, added to
     * the grid tile by tile in a fixed order) and the points of their boxes;
     * the collocate tile kernels sweep (nentries of the other pairs) tiles */
    long long csmall_pairs;
    long long csmall_points;
    long long collocate_tile_points;  /* points swept by the collocate tile kernels */
} gpw_tasklist_stats;

/* ---- single task, host memory (drop-in for the reference's collocate) ---- */

/* grid_collocate_pgf_product_rspace (src/lib/grid_base_ref_c.h:40-67).
 * dh, dh_inv: 3x3 row-major; pab: maxco x maxco row-major host array, the
 * sub-block starts at pab[o2][o1]; grid: ngrid[2] x ngrid[1] x ngrid[0] host
 * array, accumulated. lb_cube/ub_cube/sphere_bounds are only read when
 * use_ortho != 0; use_ortho != 0 takes the orthorhombic path on dh[i][i]
 * whatever the off-diagonal steps (as the reference does). Runs on the GPU
 * (one pair, plain kernels); only the grid points of the pair's cube cross
 * PCIe (the whole grid when the cube holds more points than the grid). The
 * library keeps a per-thread device workspace that grows on demand and is
 * held until the process ends. */
gpw_status gpw_collocate_pgf_product_rspace(
    int compute_tau, int use_ortho, int la_max, int la_min, int lb_max, int lb_min,
    double zeta, double zetb, double rscale, const double *dh, const double *dh_inv,
    const double *ra, const double *rab, const int *npts, const int *ngrid,
    const int *lb_grid, const int *periodic, int lmax, double radius, const int *lb_cube,
    const int *ub_cube, int nspheres, const int *sphere_bounds, int maxco, int o1, int o2,
    const double *pab, double *grid);

/* Adjoint of the above (no counterpart in the reference):
 *   hab[o2+jco][o1+ico] += d <grid, collocate(pab)> / d pab[o2+jco][o1+ico]
 * for every (ico, jco) the collocation reads. If forces != NULL (6 doubles),
 * forces[0..2] += dE/dRa and forces[3..5] += dE/dRb with E = <grid,
 * collocate(pab)> at a fixed grid index set; pab is then read as well. */
gpw_status gpw_integrate_pgf_product_rspace(
    int compute_tau, int use_ortho, int la_max, int la_min, int lb_max, int lb_min,
    double zeta, double zetb, double rscale, const double *dh, const double *dh_inv,
    const double *ra, const double *rab, const int *npts, const int *ngrid,
    const int *lb_grid, const int *periodic, int lmax, double radius, const int *lb_cube,
    const int *ub_cube, int nspheres, const int *sphere_bounds, int maxco, int o1, int o2,
    const double *grid, const double *pab, double *hab, double *forces);

/* The same two calls on device memory (no counterpart in the reference):
 * pab / hab are maxco x maxco device arrays (block at row o2, column o1),
 * read and written in place; grid is the ngrid[2] x ngrid[1] x ngrid[0]
 * device grid (accumulated); forces_dev: NULL or 6 device doubles
 * (accumulated). Stream ordered on `stream`; the pair's geometry is uploaded
 * into the calling thread's device workspace (grown with hipMalloc when a
 * larger cube needs it), so calls of one thread must stay on one stream. */
gpw_status gpw_collocate_pgf_product_rspace_dev(
    int compute_tau, int use_ortho, int la_max, int la_min, int lb_max, int lb_min,
    double zeta, double zetb, double rscale, const double *dh, const double *dh_inv,
    const double *ra, const double *rab, const int *npts, const int *ngrid,
    const int *lb_grid, const int *periodic, int lmax, double radius, const int *lb_cube,
    const int *ub_cube, int nspheres, const int *sphere_bounds, int maxco, int o1, int o2,
    const double *pab_dev, double *grid_dev, void *stream);
gpw_status gpw_integrate_pgf_product_rspace_dev(
    int compute_tau, int use_ortho, int la_max, int la_min, int lb_max, int lb_min,
    double zeta, double zetb, double rscale, const double *dh, const double *dh_inv,
    const double *ra, const double *rab, const int *npts, const int *ngrid,
    const int *lb_grid, const int *periodic, int lmax, double radius, const int *lb_cube,
    const int *ub_cube, int nspheres, const int *sphere_bounds, int maxco, int o1, int o2,
    const double *grid_dev, const double *pab_dev, double *hab_dev, double *forces_dev,
    void *stream);

/* Allocation-free variant of gpw_integrate_pgf_product_rspace_dev (no
 * counterpart in the reference): for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc. Same arguments and results; it calls no hipMalloc,
 * hipFree or other device-allocating routine. Its device scratch (the pair's
 * geometry, moments and transform workspace) is a fixed array of
 * GPW_NOALLOC_SCRATCH_BYTES per GPU, held by the library's code object until
 * the process ends; a task needing more returns GPW_ERR_ARG before any
 * device work. Every call on one GPU uses that GPU's one array: calls that
 * overlap on a GPU (other host threads, or other streams not ordered with
 * this one) race, so order them on one stream. */
#define GPW_NOALLOC_SCRATCH_BYTES (8u << 20)
gpw_status gpw_integrate_pgf_product_rspace_noalloc(
    int compute_tau, int use_ortho, int la_max, int la_min, int lb_max, int lb_min,
    double zeta, double zetb, double rscale, const double *dh, const double *dh_inv,
    const double *ra, const double *rab, const int *npts, const int *ngrid,
    const int *lb_grid, const int *periodic, int lmax, double radius, const int *lb_cube,
    const int *ub_cube, int nspheres, const int *sphere_bounds, int maxco, int o1, int o2,
    const double *grid_dev, const double *pab_dev, double *hab_dev, double *forces_dev,
    void *stream);

/* ---- batched task lists (device memory, stream ordered) ------------------ */

/* Builds the per-pair geometry (cube centre, cube and sphere bounds from each
 * pair's radius, splitting of cubes larger than the period) and the tile
 * work lists on the host, and uploads them. compute_tau applies to all pairs.
 * tile: NULL or {32, 16, 16} (the only tile shape: it is fixed by the
 * kernels' 4 x 8 register blocks); splits: workgroups per tile for every
 * kernel class (gpw_tasklist_autotune replaces them by measured choices). */
gpw_status gpw_tasklist_create(const gpw_grid_desc *grid, int npairs,
                               const gpw_pair_desc *pairs, int compute_tau,
                               const int *tile, int splits, gpw_tasklist **out);
gpw_status gpw_tasklist_destroy(gpw_tasklist *tl);
/* Re-plans a task list in place for new pair data (same number of pairs,
 * typically moved centres): geometry and work items are rebuilt, the
 * autotuned splits and the collocate partial grids are kept. Waits for the
 * device first; allocates, so not inside captured or timed regions. On
 * failure the task list has no pairs until a successful update. */
gpw_status gpw_tasklist_update(gpw_tasklist *tl, int npairs, const gpw_pair_desc *pairs);
gpw_status gpw_tasklist_stats_get(const gpw_tasklist *tl, gpw_tasklist_stats *out);

/* Host-only planning (no device needed): the statistics gpw_tasklist_create
 * would report for these pairs. */
gpw_status gpw_tasklist_plan_stats(const gpw_grid_desc *grid, int npairs,
                                   const gpw_pair_desc *pairs, int compute_tau, int splits,
                                   gpw_tasklist_stats *out);
/* Host-only: the geometry the task list derives for one pair -- cube centre
 * (FLOOR(dh_inv rp), src/lib/grid_base_ref_c.cc:534-542), the offset roffset
 * (:544-547), lb_cube / ub_cube and sphere_bounds in the layout consumed at
 * src/lib/grid_base_ref_c.cc:413-470 (what CP2K's cube_utils supplies). */
gpw_status gpw_pair_geometry(const gpw_grid_desc *grid, const gpw_pair_desc *pair,
                             int compute_tau, int cube_centre[3], double roffset[3],
                             int lb_cube[3], int ub_cube[3], int *sphere_bounds, int cap,
                             int *nspheres);

/* Autotuner (no counterpart in the reference): times the tile kernels of
 * every class with 1, 2, 3, 4, 6, 8, 12, 16, 24, ..., 128 (<= max_splits)
 * workgroups per tile on scratch inputs, best of `reps` runs after a warm-up
 * (max_splits <= 0: 8, raised on grids of few tiles so that up to 4096
 * workgroups are tried; collocate partial grids stay within 2 GiB), and keeps the
 * fastest choice per (collocate / integrate without / with forces, class);
 * collocate also counts the reduction of its partial grids. Allocates a
 * scratch grid and the partial grids, so call it at set-up, not inside
 * captured or timed regions. ms_before / ms_after (may be NULL): summed
 * kernel times of collocate + integrate with the previous and the chosen
 * splits. */
gpw_status gpw_tasklist_autotune(gpw_tasklist *tl, int max_splits, int reps, void *stream,
                                 double *ms_before, double *ms_after);

/* grid_dev += collocation of every pair (pab_dev: packed blocks). */
gpw_status gpw_tasklist_collocate(gpw_tasklist *tl, const double *pab_dev, double *grid_dev,
                                  void *stream);
/* hab_dev (packed blocks) += integrate; forces_dev (npairs x 6) += dE/dRa,
 * dE/dRb when non-NULL (pab_dev then read). */
gpw_status gpw_tasklist_integrate(gpw_tasklist *tl, const double *grid_dev,
                                  const double *pab_dev, double *hab_dev, double *forces_dev,
                                  void *stream);
/* The same two operations, one kernel launch sequence per pair (plain form). */
gpw_status gpw_tasklist_collocate_plain(gpw_tasklist *tl, const double *pab_dev,
                                        double *grid_dev, void *stream);
gpw_status gpw_tasklist_integrate_plain(gpw_tasklist *tl, const double *grid_dev,
                                        const double *pab_dev, double *hab_dev,
                                        double *forces_dev, void *stream);

/* Per-phase device timing (hipEvents recorded on the launch stream).
 * gpw_tasklist_set_timing(tl, 1) creates the events (do it outside timed or
 * captured regions); gpw_tasklist_timings waits for the recorded work and
 * returns the summed milliseconds per phase since the previous query in
 * ms[0..GPW_NPHASES-1] (and the number of timed calls per phase in calls,
 * which may be NULL), then resets. Phases: */
#define GPW_PHASE_COEF 0          /* pab -> product-centre coefficients  */
#define GPW_PHASE_COLLOCATE 1     /* collocate tile kernels              */
#define GPW_PHASE_REDUCE 2        /* sum of collocate split buffers      */
#define GPW_PHASE_INTEGRATE 3     /* integrate tile kernels              */
#define GPW_PHASE_MOMENTS 4       /* per-pair sums of partial moments    */
#define GPW_PHASE_HAB 5           /* back-transform to hab and forces    */
#define GPW_NPHASES 6
gpw_status gpw_tasklist_set_timing(gpw_tasklist *tl, int enable);
gpw_status gpw_tasklist_timings(gpw_tasklist *tl, double *ms, int n, int *calls);

/* ---- libgpw tensor-path contraction (device memory) ---------------------- */
/* collocate_core (src/lib/collocate.cc:543):
 *   V[z][y][x] = sum co[a][g][b] * P[0][g][z] * P[1][b][y] * P[2][a][x]
 * co: L x L x L, P: 3 x L x ldp, V: length[0] x length[1] x length[2]
 * (overwritten, as the reference's Vtmp). */
gpw_status gpw_collocate_core(const int *length, int L, const double *co_dev,
                              const double *P_dev, int ldp, double *V_dev, void *stream);
/* integrate_core (src/lib/collocate.cc:438):
 *   Velems[g][b][a] = sum V[z][y][x] * P[0][g][z] * P[1][b][y] * P[2][a][x] */
gpw_status gpw_integrate_core(const int *length, int L, const double *P_dev, int ldp,
                              const double *V_dev, double *Velems_dev, void *stream);

/* The same two contractions on host arrays (the library stages them through
 * device memory): drop-in for the reference's CPU templates. */
gpw_status gpw_collocate_core_host(const int *length, int L, const double *co, const double *P,
                                   int ldp, double *V);
gpw_status gpw_integrate_core_host(const int *length, int L, const double *P, int ldp,
                                   const double *V, double *Velems);

/* ---- libgpw folded-table path --------------------------------------------- */
/* Periodic folding of 1D tables, compute_folded_polynomials_all_powers
 * (src/lib/collocate.cc:308-331) with the support masks of
 * compute_nonzero_values (:32-57), for the three directions d = 0 (z), 1 (y),
 * 2 (x): entry i of an unfolded table is grid coordinate
 * center[d] - (length[d] - 1) / 2 + i, and
 *   folded[d][a][g] = sum of the entries whose coordinate is g mod period[d]
 *   mask[d][g]      = 1 where some entry lands (every g < period[d] when
 *                     length[d] >= period[d]), else 0
 * for g < ldp (ldm); both are overwritten. unfolded: (3, lu, ldu), rows
 * a < L read; folded: (3, L, ldp); mask: (3, ldm) ints or NULL. Device
 * memory, stream ordered. */
gpw_status gpw_fold_tables(int L, const int *length, const int *center, const int *period,
                           const double *unfolded, int lu, int ldu, double *folded, int ldp,
                           int *mask, int ldm, void *stream);
/* collocate<T> (src/lib/collocate.cc:681-745) on a local grid window of
 * grid_size = (nz, ny, nx) points whose (0, 0, 0) is global coordinate
 * lower_corner (z, y, x): for every point whose three masks are set,
 *   density[z][y][x] += sum_{a,g,b} co[a][g][b] P[0][g][c0 + z] P[1][b][c1 + y]
 *                       P[2][a][c2 + x]
 * (co: L^3, a the x power, g the z power, b the y power; P: (3, L, ldp)
 * folded tables; mask: (3, ldm)). Points with a clear mask are untouched. */
gpw_status gpw_collocate_folded(int L, const double *co, const double *P, int ldp,
                                const int *mask, int ldm, const int *grid_size,
                                const int *lower_corner, double *density, void *stream);
/* integrate<T> (src/lib/collocate.cc:626-679), the adjoint on the same set:
 *   co[g][b][a] = sum V[z][y][x] P[0][g][c0 + z] P[1][b][c1 + y] P[2][a][c2 + x]
 * (co overwritten; zero when no point is set). */
gpw_status gpw_integrate_folded(int L, const double *P, int ldp, const int *mask, int ldm,
                                const int *grid_size, const int *lower_corner, const double *V,
                                double *co, void *stream);
/* The same three on host arrays (staged through the calling thread's device
 * workspace; drop-ins for the reference's CPU templates). */
gpw_status gpw_fold_tables_host(int L, const int *length, const int *center, const int *period,
                                const double *unfolded, int lu, int ldu, double *folded, int ldp,
                                int *mask, int ldm);
gpw_status gpw_collocate_folded_host(int L, const double *co, const double *P, int ldp,
                                     const int *mask, int ldm, const int *grid_size,
                                     const int *lower_corner, double *density);
gpw_status gpw_integrate_folded_host(int L, const double *P, int ldp, const int *mask, int ldm,
                                     const int *grid_size, const int *lower_corner,
                                     const double *V, double *co);
/* compute_nonzero_values (src/lib/collocate.cc:32-57) for one direction, on
 * the host: mask[0 .. ldm) as described for gpw_fold_tables. */
gpw_status gpw_nonzero_values(int center, int period, int length, int *mask, int ldm);

/* last HIP error string recorded by the library (thread-local) */
const char *gpw_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
