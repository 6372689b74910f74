/*
 * gpw_oracle.h -- CPU FP64 restatement of the reference collocate/integrate
 * path. TEST INFRASTRUCTURE ONLY: it is the checker that tests, the smoke
 * check and the CPU-baseline leg of the benchmark compare against. Nothing in
 * the libgpw_amd product links, loads or calls it.
 *
 * Conventions follow src/lib/grid_base_ref_c.cc of the reference:
 *   - grids are row-major (z, y, x), x fastest          (grid_base_ref_c.cc:496-503)
 *   - pab is row-major (maxco, maxco), pab[(o2+jco)*ld + o1+ico], row = b, col = a
 *                                                        (grid_base_ref_c.cc:101,185)
 *   - Cartesian function order is CP2K's coset           (grid_base_ref_c.cc:39-46)
 */
#ifndef GPW_ORACLE_H
#define GPW_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_LMAX 8 /* largest angular momentum per centre after tau/force shifts */

enum {
    ORC_OK = 0,
    ORC_ERR_ARG = 1,     /* bad angular momentum range or grid description */
    ORC_ERR_LMAX = 2,    /* lmax smaller than the local angular momentum     */
    ORC_ERR_BOUNDS = 3,  /* non-periodic cube outside the local grid window  */
    ORC_ERR_SPHERE = 4,  /* sphere_bounds inconsistent with lb_cube/ub_cube  */
    ORC_ERR_ALLOC = 5
};

/* One collocation task: the argument list of grid_collocate_pgf_product_rspace
 * (reference grid_base_ref_c.h:40-67), gathered into a struct. */
typedef struct orc_task {
    int compute_tau, use_ortho;
    int la_max, la_min, lb_max, lb_min;
    double zeta, zetb, rscale;
    double dh[3][3], dh_inv[3][3];
    double ra[3], rab[3];
    int npts[3], ngrid[3], lb_grid[3], periodic[3];
    int lmax;
    double radius;
    int lb_cube[3], ub_cube[3];
    int nspheres;
    const int *sphere_bounds;
    int maxco, o1, o2;
    int pab_ld; /* leading dimension of pab/hab; 0 means maxco */
} orc_task;

int orc_ncoset(int l);
int orc_coset(int lx, int ly, int lz);

/* grid += collocation of the task (reference grid_collocate_internal). */
int orc_collocate(const orc_task *t, const double *pab, double *grid);

/* Adjoint of orc_collocate with respect to pab:
 *   hab[(o2+jco)*ld + o1+ico] += d<grid, collocate(pab)>/d pab[...]
 * and, when forces != NULL, forces[0..2] += dE/dRa, forces[3..5] += dE/dRb for
 * E = <grid, collocate(pab)>, at fixed grid index set. pab is only read when
 * forces != NULL. */
int orc_integrate(const orc_task *t, const double *grid, const double *pab,
                  double *hab, double *forces);

/* Direct evaluation rscale * sum_ab pab_ab phi_a(r) phi_b(r) (or
 * 0.5 grad phi_a . grad phi_b for tau) on the same grid point set, with no
 * coefficient transform and no 1D tables. */
int orc_collocate_brute(const orc_task *t, const double *pab, double *grid);

/* Cartesian product-centre coefficients coef_xyz (packed lzp, lyp, lxp) of
 * the task; returns lp via *lp_out. coef must hold (lp+1)(lp+2)(lp+3)/6. */
int orc_coef_xyz(const orc_task *t, const double *pab, double *coef, int *lp_out);

/* Cube and sphere bounds for an orthorhombic grid of spacings dr[3] and a
 * radius, in the layout grid_collocate_core consumes
 * (grid_base_ref_c.cc:413-470). Returns the number of sphere_bounds entries,
 * or -needed if cap is too small. */
int orc_sphere_bounds(double radius, const double dr[3], int lb_cube[3],
                      int ub_cube[3], int *sphere_bounds, int cap);

/* libgpw tensor-path contractions (reference collocate.cc:438-623 and the
 * naive loops of collocate_test.cc:11-33, 77-99):
 *   V[z][y][x] (+)= sum co[a][g][b] * P[0][g][z] * P[1][b][y] * P[2][a][x]
 *   Velems[g][b][a] = sum V[z][y][x] * P[0][g][z] * P[1][b][y] * P[2][a][x]
 * co is (L, L, L), P is (3, L, ldp), V is (length[0], length[1], length[2]). */
void orc_collocate_core(const int length[3], int L, const double *co,
                        const double *P, int ldp, double *V);
void orc_integrate_core(const int length[3], int L, const double *P, int ldp,
                        const double *V, double *Velems);

/* libgpw folded-table path (reference collocate.cc:32-57, 276-331, 626-745,
 * polynomials.hpp:5-86, utils.hpp:20-181), defects fixed:
 *   fold:      folded[d][a][g] = sum_i unfolded[d][a][i] over the i whose grid
 *              coordinate center[d] - (length[d]-1)/2 + i is g mod period[d];
 *              mask[d][g] = 1 on those g (all g < period when length >= period).
 *              unfolded (3, L, ldu), folded (3, L, ldp), mask (3, ldm) or NULL.
 *   collocate: density[z][y][x] += sum co[a][g][b] P[0][g][c0+z] P[1][b][c1+y]
 *              P[2][a][c2+x] on the local window grid_size (z, y, x) at
 *              lower_corner c where all three masks are set;
 *   integrate: co[g][b][a] = sum V[z][y][x] P[0][g][c0+z] ... on the same set. */
void orc_fold_tables(int L, const int length[3], const int center[3], const int period[3],
                     const double *unfolded, int ldu, double *folded, int ldp, int *mask,
                     int ldm);
void orc_collocate_folded(int L, const double *co, const double *P, int ldp, const int *mask,
                          int ldm, const int grid_size[3], const int lower_corner[3],
                          double *density);
void orc_integrate_folded(int L, const double *P, int ldp, const int *mask, int ldm,
                          const int grid_size[3], const int lower_corner[3], const double *V,
                          double *co);

/* ---- batches of primitive pairs on one grid (CPU baseline and test driver) */
typedef struct orc_pair {
    int la_max, la_min, lb_max, lb_min;
    double zeta, zetb, rscale;
    double ra[3], rab[3];
    double radius;
    long long block_offset; /* offset (doubles) of the pair's pab/hab block */
} orc_pair;

typedef struct orc_grid {
    int npts[3];
    double dh[3][3], dh_inv[3][3];
    int orthorhombic;
} orc_grid;

/* Every pair uses a packed block ncoset(lb_max) x ncoset(la_max) (row b,
 * col a) at block_offset. Periodic grid, ngrid == npts. nthreads <= 0 uses
 * the OpenMP default. compute_tau applies to every pair. */
int orc_batch_collocate(const orc_grid *g, int npairs, const orc_pair *pairs,
                        int compute_tau, const double *pab, double *grid,
                        int nthreads);
int orc_batch_integrate(const orc_grid *g, int npairs, const orc_pair *pairs,
                        int compute_tau, const double *grid, const double *pab,
                        double *hab, double *forces, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
