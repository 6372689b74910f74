/*
 * gpw_oracle.c -- CPU FP64 restatement of the reference collocate/integrate
 * path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: tests, the smoke check
 * and the CPU-baseline leg of the benchmark compare the HIP product against
 * it. The product (libgpw_amd) never links or calls it.
 *
 * What follows which reference lines (paths relative to the reference root):
 *   coset / ncoset ............ src/lib/grid_base_ref_c.cc:18-46
 *   pab extraction (rho, tau) . src/lib/grid_base_ref_c.cc:65-192
 *   alpha (binomial shift) .... src/lib/grid_base_ref_c.cc:195-231
 *   coef_xyz (cab -> cxyz) .... src/lib/grid_base_ref_c.cc:234-315 (evaluated
 *                               as the equivalent direct triple product)
 *   periodic / windowed map ... src/lib/grid_base_ref_c.cc:318-351
 *   1D tables ................. src/lib/grid_base_ref_c.cc:355-400 (direct exp
 *                               per point instead of the 3-exp recursion)
 *   ortho contraction ......... src/lib/grid_base_ref_c.cc:403-589
 *   general (non-ortho) path .. src/lib/grid_base_ref_c.cc:593-846
 *   pair setup ................ src/lib/grid_base_ref_c.cc:878-986
 *   libgpw tensor contraction . src/lib/collocate_test.cc:11-33, 77-99
 *   libgpw fold / mask ........ src/lib/collocate.cc:32-57, 276-331
 *   libgpw collocate<T> /
 *   integrate<T> (pack,
 *   unpack) ................... src/lib/collocate.cc:626-745,
 *                               src/lib/polynomials.hpp:5-86, src/lib/utils.hpp:20-181
 * Integrate and forces have no runnable reference: they are defined here as
 * the exact adjoint of collocate (math in src/lib/doc.tex:363-415) and as the
 * analytic derivative with respect to the two centres at a fixed index set.
 *
 * Parity pins (see tests/test_oracle.py): the four CP2K .task fixtures, a
 * direct brute-force evaluator (orc_collocate_brute), ortho == general on an
 * orthorhombic cell, adjointness <collocate(P), W> == <P, integrate(W)>,
 * central finite differences for forces, and the libgpw naive contractions;
 * the folded-table path by a direct periodic image sum of separable Gaussian
 * tables, adjointness, and windows that are slices of the full period.
 * The cube/sphere-bounds generator reproduces both ortho fixtures exactly.
 *
 * Build with -ffp-contract=off so index arithmetic is reproducible.
 */
#include "gpw_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LPMAX (2 * ORC_LMAX)
#define NALPHA (2 * ORC_LMAX + 1)

/* ------------------------------------------------------------------------ */
/* basic index helpers                                                       */

int orc_ncoset(int l) { return l < 0 ? 0 : ((l + 1) * (l + 2) * (l + 3)) / 6; }

int orc_coset(int lx, int ly, int lz) {
    const int l = lx + ly + lz;
    if (l == 0) return 0;
    return orc_ncoset(l - 1) + ((l - lx) * (l - lx + 1)) / 2 + lz;
}

static int imod(int a, int m) { return (a % m + m) % m; }
static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }

/* packed (lzp, lyp, lxp) index with lxp+lyp+lzp <= lp, lzp outermost,
 * lxp fastest (grid_base_ref_c.cc:250-257). */
static int cidx(int lp, int lzp, int lyp, int lxp) {
    int o = 0;
    for (int z = 0; z < lzp; z++) o += ((lp - z + 1) * (lp - z + 2)) / 2;
    const int m = lp - lzp;
    o += lyp * (m + 1) - (lyp * (lyp - 1)) / 2;
    return o + lxp;
}

/* ------------------------------------------------------------------------ */
/* pair set-up                                                               */

typedef struct {
    int la_max, la_min, lb_max, lb_min; /* local ranges after tau shift */
    int la_max_x, la_min_x, lb_max_x, lb_min_x; /* + force extension */
    double zetp, prefactor, rp[3], rb[3];
    int lp, lp_x;
} pair_setup;

static int setup_pair(const orc_task *t, int forces, pair_setup *s) {
    if (t->la_min < 0 || t->lb_min < 0 || t->la_max < t->la_min ||
        t->lb_max < t->lb_min)
        return ORC_ERR_ARG;
    const int tau = t->compute_tau ? 1 : 0;
    s->la_max = t->la_max + tau;
    s->lb_max = t->lb_max + tau;
    s->la_min = imax(t->la_min - tau, 0);
    s->lb_min = imax(t->lb_min - tau, 0);
    const int fx = forces ? 1 : 0;
    s->la_max_x = s->la_max + fx;
    s->lb_max_x = s->lb_max + fx;
    s->la_min_x = imax(s->la_min - fx, 0);
    s->lb_min_x = imax(s->lb_min - fx, 0);
    if (s->la_max_x > ORC_LMAX || s->lb_max_x > ORC_LMAX) return ORC_ERR_ARG;
    /* reference preconditions: alpha sized by lmax (grid_base_ref_c.cc:946),
     * coef_xyt/coef_xtt sized by 2*lmax (:246-248, :739, :768) */
    if (t->lmax < s->la_max || t->lmax < s->lb_max) return ORC_ERR_LMAX;
    s->lp = s->la_max + s->lb_max;
    s->lp_x = s->la_max_x + s->lb_max_x;
    if (2 * t->lmax < s->lp) return ORC_ERR_LMAX;

    s->zetp = t->zeta + t->zetb;
    const double f = t->zetb / s->zetp;
    const double rab2 =
        t->rab[0] * t->rab[0] + t->rab[1] * t->rab[1] + t->rab[2] * t->rab[2];
    s->prefactor = t->rscale * exp(-t->zeta * f * rab2);
    for (int i = 0; i < 3; i++) {
        s->rp[i] = t->ra[i] + f * t->rab[i];
        s->rb[i] = t->ra[i] + t->rab[i];
    }
    return ORC_OK;
}

/* alpha[d][lb][la][k]: (x-a)^la (x-b)^lb = sum_k alpha (x-p)^k */
typedef double alpha_t[3][ORC_LMAX + 1][ORC_LMAX + 1][NALPHA];

static void fill_alpha(const double ra[3], const double rb[3], const double rp[3],
                       int la_max, int lb_max, alpha_t alpha) {
    memset(alpha, 0, sizeof(alpha_t));
    for (int d = 0; d < 3; d++) {
        const double pa = rp[d] - ra[d];
        const double pb = rp[d] - rb[d];
        for (int la = 0; la <= la_max; la++) {
            for (int lb = 0; lb <= lb_max; lb++) {
                double ca = 1.0, powa = 1.0;
                for (int i = 0; i <= la; i++) {
                    double cb = 1.0, powb = 1.0;
                    for (int j = 0; j <= lb; j++) {
                        alpha[d][lb][la][(la - i) + (lb - j)] += ca * cb * powa * powb;
                        cb *= (double)(lb - j) / (double)(j + 1);
                        powb *= pb;
                    }
                    ca *= (double)(la - i) / (double)(i + 1);
                    powa *= pa;
                }
            }
        }
    }
}

/* effective block pab_eff[jco][ico] (dims ncoset(lb_max) x ncoset(la_max) of
 * the local ranges) -- rho copy or the tau combination. */
static void prepare_pab(const orc_task *t, const pair_setup *s, const double *pab,
                        double *eff, int ld_eff) {
    const int ld = t->pab_ld > 0 ? t->pab_ld : t->maxco;
    const int nla = orc_ncoset(s->la_max), nlb = orc_ncoset(s->lb_max);
    for (int j = 0; j < nlb; j++)
        for (int i = 0; i < nla; i++) eff[j * ld_eff + i] = 0.0;
    for (int la = t->la_min; la <= t->la_max; la++)
        for (int lxa = 0; lxa <= la; lxa++)
            for (int lya = 0; lya <= la - lxa; lya++) {
                const int lza = la - lxa - lya;
                const int ico = orc_coset(lxa, lya, lza);
                for (int lb = t->lb_min; lb <= t->lb_max; lb++)
                    for (int lxb = 0; lxb <= lb; lxb++)
                        for (int lyb = 0; lyb <= lb - lxb; lyb++) {
                            const int lzb = lb - lxb - lyb;
                            const int jco = orc_coset(lxb, lyb, lzb);
                            const double p = pab[(t->o2 + jco) * ld + t->o1 + ico];
                            if (!t->compute_tau) {
                                eff[jco * ld_eff + ico] = p;
                                continue;
                            }
                            const int a[3] = {lxa, lya, lza}, b[3] = {lxb, lyb, lzb};
                            for (int d = 0; d < 3; d++) {
                                int am[3] = {a[0], a[1], a[2]}, ap[3] = {a[0], a[1], a[2]};
                                int bm[3] = {b[0], b[1], b[2]}, bp[3] = {b[0], b[1], b[2]};
                                am[d] = a[d] > 0 ? a[d] - 1 : 0;
                                bm[d] = b[d] > 0 ? b[d] - 1 : 0;
                                ap[d] = a[d] + 1;
                                bp[d] = b[d] + 1;
                                const int iam = orc_coset(am[0], am[1], am[2]);
                                const int iap = orc_coset(ap[0], ap[1], ap[2]);
                                const int jbm = orc_coset(bm[0], bm[1], bm[2]);
                                const int jbp = orc_coset(bp[0], bp[1], bp[2]);
                                eff[jbm * ld_eff + iam] += 0.5 * a[d] * b[d] * p;
                                eff[jbp * ld_eff + iam] += 0.5 * (-2.0 * a[d] * t->zetb) * p;
                                eff[jbm * ld_eff + iap] += 0.5 * (-2.0 * t->zeta * b[d]) * p;
                                eff[jbp * ld_eff + iap] += 0.5 * (4.0 * t->zeta * t->zetb) * p;
                            }
                        }
            }
}

/* adjoint of prepare_pab: hab += d(sum eff*heff)/d pab */
static void prepare_pab_adjoint(const orc_task *t, const double *heff, int ld_eff,
                                double *hab) {
    const int ld = t->pab_ld > 0 ? t->pab_ld : t->maxco;
    for (int la = t->la_min; la <= t->la_max; la++)
        for (int lxa = 0; lxa <= la; lxa++)
            for (int lya = 0; lya <= la - lxa; lya++) {
                const int lza = la - lxa - lya;
                const int ico = orc_coset(lxa, lya, lza);
                for (int lb = t->lb_min; lb <= t->lb_max; lb++)
                    for (int lxb = 0; lxb <= lb; lxb++)
                        for (int lyb = 0; lyb <= lb - lxb; lyb++) {
                            const int lzb = lb - lxb - lyb;
                            const int jco = orc_coset(lxb, lyb, lzb);
                            double acc;
                            if (!t->compute_tau) {
                                acc = heff[jco * ld_eff + ico];
                            } else {
                                const int a[3] = {lxa, lya, lza}, b[3] = {lxb, lyb, lzb};
                                acc = 0.0;
                                for (int d = 0; d < 3; d++) {
                                    int am[3] = {a[0], a[1], a[2]}, ap[3] = {a[0], a[1], a[2]};
                                    int bm[3] = {b[0], b[1], b[2]}, bp[3] = {b[0], b[1], b[2]};
                                    am[d] = a[d] > 0 ? a[d] - 1 : 0;
                                    bm[d] = b[d] > 0 ? b[d] - 1 : 0;
                                    ap[d] = a[d] + 1;
                                    bp[d] = b[d] + 1;
                                    const int iam = orc_coset(am[0], am[1], am[2]);
                                    const int iap = orc_coset(ap[0], ap[1], ap[2]);
                                    const int jbm = orc_coset(bm[0], bm[1], bm[2]);
                                    const int jbp = orc_coset(bp[0], bp[1], bp[2]);
                                    acc += 0.5 * a[d] * b[d] * heff[jbm * ld_eff + iam];
                                    acc += 0.5 * (-2.0 * a[d] * t->zetb) * heff[jbp * ld_eff + iam];
                                    acc += 0.5 * (-2.0 * t->zeta * b[d]) * heff[jbm * ld_eff + iap];
                                    acc += 0.5 * (4.0 * t->zeta * t->zetb) * heff[jbp * ld_eff + iap];
                                }
                            }
                            hab[(t->o2 + jco) * ld + t->o1 + ico] += acc;
                        }
            }
}

/* coef[cidx(lp,...)] = prefactor * sum_{a,b} eff[b][a] prod_d alpha */
static void coef_from_eff(const pair_setup *s, const alpha_t alpha, const double *eff,
                          int ld_eff, double *coef) {
    const int lp = s->lp;
    for (int i = 0; i < orc_ncoset(lp); i++) coef[i] = 0.0;
    for (int la = s->la_min; la <= s->la_max; la++)
        for (int lxa = 0; lxa <= la; lxa++)
            for (int lya = 0; lya <= la - lxa; lya++) {
                const int lza = la - lxa - lya;
                const int ico = orc_coset(lxa, lya, lza);
                for (int lb = s->lb_min; lb <= s->lb_max; lb++)
                    for (int lxb = 0; lxb <= lb; lxb++)
                        for (int lyb = 0; lyb <= lb - lxb; lyb++) {
                            const int lzb = lb - lxb - lyb;
                            const int jco = orc_coset(lxb, lyb, lzb);
                            const double p = s->prefactor * eff[jco * ld_eff + ico];
                            if (p == 0.0) continue;
                            for (int lzp = 0; lzp <= lza + lzb; lzp++) {
                                const double pz = p * alpha[2][lzb][lza][lzp];
                                for (int lyp = 0; lyp <= lya + lyb; lyp++) {
                                    const double pyz = pz * alpha[1][lyb][lya][lyp];
                                    for (int lxp = 0; lxp <= lxa + lxb; lxp++)
                                        coef[cidx(lp, lzp, lyp, lxp)] +=
                                            pyz * alpha[0][lxb][lxa][lxp];
                                }
                            }
                        }
            }
}

/* heff[b][a] = prefactor * sum_l M[l] prod_d alpha, over the extended range */
static void heff_from_moments(const pair_setup *s, const alpha_t alpha, const double *M,
                              double *heff, int ld_eff) {
    const int lp = s->lp_x;
    for (int lb = 0; lb <= s->lb_max_x; lb++)
        for (int la = 0; la <= s->la_max_x; la++) {
            for (int lxa = 0; lxa <= la; lxa++)
                for (int lya = 0; lya <= la - lxa; lya++) {
                    const int lza = la - lxa - lya;
                    const int ico = orc_coset(lxa, lya, lza);
                    for (int lxb = 0; lxb <= lb; lxb++)
                        for (int lyb = 0; lyb <= lb - lxb; lyb++) {
                            const int lzb = lb - lxb - lyb;
                            const int jco = orc_coset(lxb, lyb, lzb);
                            double acc = 0.0;
                            if (la >= s->la_min_x && lb >= s->lb_min_x) {
                                for (int lzp = 0; lzp <= lza + lzb; lzp++)
                                    for (int lyp = 0; lyp <= lya + lyb; lyp++)
                                        for (int lxp = 0; lxp <= lxa + lxb; lxp++)
                                            acc += M[cidx(lp, lzp, lyp, lxp)] *
                                                   alpha[2][lzb][lza][lzp] *
                                                   alpha[1][lyb][lya][lyp] *
                                                   alpha[0][lxb][lxa][lxp];
                            }
                            heff[jco * ld_eff + ico] = s->prefactor * acc;
                        }
                }
        }
}

/* dE/dRa, dE/dRb from heff (extended block) and eff (local block):
 *   d phi_a / d A_d = 2 zeta phi_{a+1_d} - a_d phi_{a-1_d}                    */
static void forces_from_heff(const orc_task *t, const pair_setup *s, const double *eff,
                             int ld_eff, const double *heff, int ldh, double *forces) {
    double fa[3] = {0, 0, 0}, fb[3] = {0, 0, 0};
    for (int la = s->la_min; la <= s->la_max; la++)
        for (int lxa = 0; lxa <= la; lxa++)
            for (int lya = 0; lya <= la - lxa; lya++) {
                const int lza = la - lxa - lya;
                const int a[3] = {lxa, lya, lza};
                const int ico = orc_coset(lxa, lya, lza);
                for (int lb = s->lb_min; lb <= s->lb_max; lb++)
                    for (int lxb = 0; lxb <= lb; lxb++)
                        for (int lyb = 0; lyb <= lb - lxb; lyb++) {
                            const int lzb = lb - lxb - lyb;
                            const int b[3] = {lxb, lyb, lzb};
                            const int jco = orc_coset(lxb, lyb, lzb);
                            const double p = eff[jco * ld_eff + ico];
                            if (p == 0.0) continue;
                            for (int d = 0; d < 3; d++) {
                                int ap[3] = {a[0], a[1], a[2]}, am[3] = {a[0], a[1], a[2]};
                                int bp[3] = {b[0], b[1], b[2]}, bm[3] = {b[0], b[1], b[2]};
                                ap[d]++;
                                bp[d]++;
                                double ga = 2.0 * t->zeta *
                                            heff[jco * ldh + orc_coset(ap[0], ap[1], ap[2])];
                                double gb = 2.0 * t->zetb *
                                            heff[orc_coset(bp[0], bp[1], bp[2]) * ldh + ico];
                                if (a[d] > 0) {
                                    am[d]--;
                                    ga -= a[d] * heff[jco * ldh + orc_coset(am[0], am[1], am[2])];
                                }
                                if (b[d] > 0) {
                                    bm[d]--;
                                    gb -= b[d] * heff[orc_coset(bm[0], bm[1], bm[2]) * ldh + ico];
                                }
                                fa[d] += p * ga;
                                fb[d] += p * gb;
                            }
                        }
            }
    for (int d = 0; d < 3; d++) {
        forces[d] += fa[d];
        forces[3 + d] += fb[d];
    }
}

/* ------------------------------------------------------------------------ */
/* orthorhombic point set                                                    */

typedef struct {
    int cc[3];
    double roff[3], dr[3];
    int gshift[3]; /* non-periodic: grid index = offset + gshift */
} ortho_geom;

static int ortho_setup(const orc_task *t, const double rp[3], ortho_geom *g) {
    for (int i = 0; i < 3; i++) {
        double v = 0.0;
        for (int j = 0; j < 3; j++) v += t->dh_inv[j][i] * rp[j];
        g->cc[i] = (int)floor(v);
        g->dr[i] = t->dh[i][i];
        g->roff[i] = rp[i] - ((double)g->cc[i]) * t->dh[i][i];
    }
    for (int d = 0; d < 3; d++) {
        if (t->npts[d] <= 0 || t->ngrid[d] <= 0) return ORC_ERR_ARG;
        if (t->periodic[d]) {
            if (t->ngrid[d] != t->npts[d]) return ORC_ERR_ARG;
            g->gshift[d] = 0;
        } else {
            g->gshift[d] = imod(g->cc[d] + t->lb_cube[d] + t->lb_grid[d], t->npts[d]) -
                           t->lb_cube[d];
            if (t->lb_cube[d] + g->gshift[d] < 0 ||
                t->ub_cube[d] + g->gshift[d] > t->ngrid[d] - 1)
                return ORC_ERR_BOUNDS;
        }
    }
    return ORC_OK;
}

static inline int ortho_map(const orc_task *t, const ortho_geom *g, int d, int off) {
    return t->periodic[d] ? imod(g->cc[d] + off, t->npts[d]) : off + g->gshift[d];
}

/* check the whole sphere_bounds walk against lb_cube/ub_cube first */
static int check_sphere(const orc_task *t) {
    const int *sb = t->sphere_bounds;
    int sci = 0;
    if (t->nspheres < 1 || sb == NULL) return ORC_ERR_SPHERE;
    const int kgmin = sb[sci++];
    if (kgmin > 0 || kgmin < t->lb_cube[2] || 1 - kgmin > t->ub_cube[2]) return ORC_ERR_SPHERE;
    for (int kg = kgmin; kg <= 0; kg++) {
        if (sci >= t->nspheres) return ORC_ERR_SPHERE;
        const int jgmin = sb[sci++];
        if (jgmin > 0 || jgmin < t->lb_cube[1] || 1 - jgmin > t->ub_cube[1])
            return ORC_ERR_SPHERE;
        for (int jg = jgmin; jg <= 0; jg++) {
            if (sci >= t->nspheres) return ORC_ERR_SPHERE;
            const int igmin = sb[sci++];
            if (igmin > 0 || igmin < t->lb_cube[0] || 1 - igmin > t->ub_cube[0])
                return ORC_ERR_SPHERE;
        }
    }
    return ORC_OK;
}

/* 1D tables over cube offsets [lb, ub]: pol[(off-lb)*(lp+1)+k] */
static double *ortho_tables(const orc_task *t, const ortho_geom *g, int lp, double zetp) {
    size_t n = 0;
    for (int d = 0; d < 3; d++) n += (size_t)(t->ub_cube[d] - t->lb_cube[d] + 1) * (lp + 1);
    double *pol = (double *)malloc(n * sizeof(double));
    if (!pol) return NULL;
    double *p = pol;
    for (int d = 0; d < 3; d++) {
        for (int off = t->lb_cube[d]; off <= t->ub_cube[d]; off++) {
            const double r = off * g->dr[d] - g->roff[d];
            double v = exp(-zetp * r * r);
            for (int k = 0; k <= lp; k++) {
                *p++ = v;
                v *= r;
            }
        }
    }
    return pol;
}

static int collocate_ortho(const orc_task *t, const pair_setup *s, const double *coef,
                           double *grid) {
    ortho_geom g;
    int err = ortho_setup(t, s->rp, &g);
    if (err) return err;
    if ((err = check_sphere(t))) return err;
    const int lp = s->lp, L = lp + 1;
    double *pol = ortho_tables(t, &g, lp, s->zetp);
    if (!pol) return ORC_ERR_ALLOC;
    const double *px = pol;
    const double *py = px + (size_t)(t->ub_cube[0] - t->lb_cube[0] + 1) * L;
    const double *pz = py + (size_t)(t->ub_cube[1] - t->lb_cube[1] + 1) * L;
    const size_t nx = t->ngrid[0], ny = t->ngrid[1];
    const int *sb = t->sphere_bounds;
    int sci = 0;
    const int kgmin = sb[sci++];
    for (int kg = kgmin; kg <= 0; kg++) {
        const int jgmin = sb[sci++];
        const int *igmins = sb + sci;
        sci += 1 - jgmin;
        for (int sk = 0; sk < 2; sk++) {
            const int koff = sk ? 1 - kg : kg;
            const int gk = ortho_map(t, &g, 2, koff);
            const double *tz = pz + (size_t)(koff - t->lb_cube[2]) * L;
            double cxy[LPMAX + 1][LPMAX + 1];
            for (int lyp = 0; lyp <= lp; lyp++)
                for (int lxp = 0; lxp <= lp - lyp; lxp++) {
                    double acc = 0.0;
                    for (int lzp = 0; lzp <= lp - lyp - lxp; lzp++)
                        acc += coef[cidx(lp, lzp, lyp, lxp)] * tz[lzp];
                    cxy[lyp][lxp] = acc;
                }
            for (int jg = jgmin; jg <= 0; jg++) {
                const int igmin = igmins[jg - jgmin];
                for (int sj = 0; sj < 2; sj++) {
                    const int joff = sj ? 1 - jg : jg;
                    const int gj = ortho_map(t, &g, 1, joff);
                    const double *ty = py + (size_t)(joff - t->lb_cube[1]) * L;
                    double cx[LPMAX + 1];
                    for (int lxp = 0; lxp <= lp; lxp++) {
                        double acc = 0.0;
                        for (int lyp = 0; lyp <= lp - lxp; lyp++) acc += cxy[lyp][lxp] * ty[lyp];
                        cx[lxp] = acc;
                    }
                    double *row = grid + ((size_t)gk * ny + gj) * nx;
                    for (int ig = igmin; ig <= 0; ig++) {
                        for (int si = 0; si < 2; si++) {
                            const int ioff = si ? 1 - ig : ig;
                            const double *tx = px + (size_t)(ioff - t->lb_cube[0]) * L;
                            double v = 0.0;
                            for (int lxp = 0; lxp <= lp; lxp++) v += cx[lxp] * tx[lxp];
                            row[ortho_map(t, &g, 0, ioff)] += v;
                        }
                    }
                }
            }
        }
    }
    free(pol);
    return ORC_OK;
}

/* moments M[cidx(lp,..)] = sum_points V * X^lxp Y^lyp Z^lzp exp(-zetp r^2) */
static int moments_ortho(const orc_task *t, const pair_setup *s, int lp, const double *grid,
                         double *M) {
    ortho_geom g;
    int err = ortho_setup(t, s->rp, &g);
    if (err) return err;
    if ((err = check_sphere(t))) return err;
    const int L = lp + 1;
    double *pol = ortho_tables(t, &g, lp, s->zetp);
    if (!pol) return ORC_ERR_ALLOC;
    const double *px = pol;
    const double *py = px + (size_t)(t->ub_cube[0] - t->lb_cube[0] + 1) * L;
    const double *pz = py + (size_t)(t->ub_cube[1] - t->lb_cube[1] + 1) * L;
    const size_t nx = t->ngrid[0], ny = t->ngrid[1];
    for (int i = 0; i < orc_ncoset(lp); i++) M[i] = 0.0;
    const int *sb = t->sphere_bounds;
    int sci = 0;
    const int kgmin = sb[sci++];
    for (int kg = kgmin; kg <= 0; kg++) {
        const int jgmin = sb[sci++];
        const int *igmins = sb + sci;
        sci += 1 - jgmin;
        for (int sk = 0; sk < 2; sk++) {
            const int koff = sk ? 1 - kg : kg;
            const int gk = ortho_map(t, &g, 2, koff);
            const double *tz = pz + (size_t)(koff - t->lb_cube[2]) * L;
            double txy[LPMAX + 1][LPMAX + 1];
            memset(txy, 0, sizeof(txy));
            for (int jg = jgmin; jg <= 0; jg++) {
                const int igmin = igmins[jg - jgmin];
                for (int sj = 0; sj < 2; sj++) {
                    const int joff = sj ? 1 - jg : jg;
                    const int gj = ortho_map(t, &g, 1, joff);
                    const double *ty = py + (size_t)(joff - t->lb_cube[1]) * L;
                    const double *row = grid + ((size_t)gk * ny + gj) * nx;
                    double tx[LPMAX + 1];
                    for (int lxp = 0; lxp <= lp; lxp++) tx[lxp] = 0.0;
                    for (int ig = igmin; ig <= 0; ig++) {
                        for (int si = 0; si < 2; si++) {
                            const int ioff = si ? 1 - ig : ig;
                            const double *pxr = px + (size_t)(ioff - t->lb_cube[0]) * L;
                            const double v = row[ortho_map(t, &g, 0, ioff)];
                            for (int lxp = 0; lxp <= lp; lxp++) tx[lxp] += v * pxr[lxp];
                        }
                    }
                    for (int lyp = 0; lyp <= lp; lyp++)
                        for (int lxp = 0; lxp <= lp - lyp; lxp++) txy[lyp][lxp] += tx[lxp] * ty[lyp];
                }
            }
            for (int lzp = 0; lzp <= lp; lzp++)
                for (int lyp = 0; lyp <= lp - lzp; lyp++)
                    for (int lxp = 0; lxp <= lp - lzp - lyp; lxp++)
                        M[cidx(lp, lzp, lyp, lxp)] += txy[lyp][lxp] * tz[lzp];
        }
    }
    free(pol);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* general (non-orthorhombic) point set                                      */

static const double factorial[LPMAX + 1] = {
    1.0, 1.0, 2.0, 6.0, 24.0, 120.0, 720.0, 5040.0, 40320.0, 362880.0, 3628800.0,
    39916800.0, 479001600.0, 6227020800.0, 87178291200.0, 1307674368000.0,
    20922789888000.0};

/* coef_ijk = T coef_xyz (forward) or M_xyz = T^T M_ijk (adjoint); both packed
 * with cidx(lp, kl/lz, jl/ly, il/lx). */
static void multinomial(const orc_task *t, int lp, const double *src, double *dst, int adjoint) {
    double hp[LPMAX + 1][3][3]; /* hp[k][j][i] = dh[j][i]^k */
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            hp[0][j][i] = 1.0;
            for (int k = 1; k <= lp; k++) hp[k][j][i] = hp[k - 1][j][i] * t->dh[j][i];
        }
    for (int i = 0; i < orc_ncoset(lp); i++) dst[i] = 0.0;
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
                                        const int il = ilx + ily + ilz;
                                        const int jl = jlx + jly + jlz;
                                        const int kl = klx + kly + klz;
                                        const double w =
                                            hp[ilx][0][0] * hp[jlx][1][0] * hp[klx][2][0] *
                                            hp[ily][0][1] * hp[jly][1][1] * hp[kly][2][1] *
                                            hp[ilz][0][2] * hp[jlz][1][2] * hp[klz][2][2] *
                                            factorial[lx] * factorial[ly] * factorial[lz] /
                                            (factorial[ilx] * factorial[ily] * factorial[ilz] *
                                             factorial[jlx] * factorial[jly] * factorial[jlz] *
                                             factorial[klx] * factorial[kly] * factorial[klz]);
                                        const int ijk = cidx(lp, kl, jl, il);
                                        const int xyz = cidx(lp, lz, ly, lx);
                                        if (adjoint)
                                            dst[xyz] += w * src[ijk];
                                        else
                                            dst[ijk] += w * src[xyz];
                                    }
                        }
            }
}

typedef struct {
    double gp[3];
    int index_min[3], index_max[3], offset[3];
} general_geom;

static int general_setup(const orc_task *t, const double rp[3], general_geom *g) {
    for (int i = 0; i < 3; i++) {
        g->gp[i] = 0.0;
        for (int j = 0; j < 3; j++) g->gp[i] += t->dh_inv[j][i] * rp[j];
        g->index_min[i] = 2147483647;
        g->index_max[i] = -2147483647 - 1;
    }
    for (int i = -1; i <= 1; i++)
        for (int j = -1; j <= 1; j++)
            for (int k = -1; k <= 1; k++) {
                const double x = rp[0] + i * t->radius;
                const double y = rp[1] + j * t->radius;
                const double z = rp[2] + k * t->radius;
                for (int d = 0; d < 3; d++) {
                    const double r = t->dh_inv[0][d] * x + t->dh_inv[1][d] * y + t->dh_inv[2][d] * z;
                    g->index_min[d] = imin(g->index_min[d], (int)floor(r));
                    g->index_max[d] = imax(g->index_max[d], (int)ceil(r));
                }
            }
    for (int d = 0; d < 3; d++) {
        if (t->npts[d] <= 0 || t->ngrid[d] <= 0) return ORC_ERR_ARG;
        if (t->periodic[d]) {
            if (t->ngrid[d] != t->npts[d]) return ORC_ERR_ARG;
        } else {
            g->offset[d] = imod(g->index_min[d] + t->lb_grid[d], t->npts[d]);
            if (g->index_max[d] - g->index_min[d] + g->offset[d] > t->ngrid[d] - 1)
                return ORC_ERR_BOUNDS;
        }
    }
    return ORC_OK;
}

static inline int general_map(const orc_task *t, const general_geom *g, int d, int i) {
    return t->periodic[d] ? imod(i, t->npts[d]) : i - g->index_min[d] + g->offset[d];
}

/* i-range of the row (j, k) inside the sphere (grid_base_ref_c.cc:790-811);
 * returns 0 for an empty row, else fills the quadratic coefficients already
 * scaled by -zetp. */
static int general_row(const orc_task *t, const general_geom *g, double zetp, double dj,
                       double dk, int *ismin, int *ismax, double *qa, double *qb, double *qc) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (int i = 0; i < 3; i++) {
        const double pointk = t->dh[2][i] * dk;
        const double pointj = pointk + t->dh[1][i] * dj;
        const double v = pointj - g->gp[0] * t->dh[0][i];
        a += t->dh[0][i] * t->dh[0][i];
        b += 2.0 * v * t->dh[0][i];
        c += v * v;
    }
    double d = b * b - 4 * a * (c - t->radius * t->radius);
    if (d < 0.0) return 0;
    d = sqrt(d);
    *ismin = (int)ceil((-b - d) / (2.0 * a));
    *ismax = (int)floor((-b + d) / (2.0 * a));
    *qa = a * -zetp;
    *qb = b * -zetp;
    *qc = c * -zetp;
    return 1;
}

static int collocate_general(const orc_task *t, const pair_setup *s, const double *coef,
                             double *grid) {
    general_geom g;
    int err = general_setup(t, s->rp, &g);
    if (err) return err;
    const int lp = s->lp;
    double cijk[1024];
    multinomial(t, lp, coef, cijk, 0);
    const size_t nx = t->ngrid[0], ny = t->ngrid[1];
    for (int k = g.index_min[2]; k <= g.index_max[2]; k++) {
        const double dk = k - g.gp[2];
        const int gk = general_map(t, &g, 2, k);
        double cxy[LPMAX + 1][LPMAX + 1];
        memset(cxy, 0, sizeof(cxy));
        double dkp = 1.0;
        for (int kl = 0; kl <= lp; kl++) {
            for (int jl = 0; jl <= lp - kl; jl++)
                for (int il = 0; il <= lp - kl - jl; il++)
                    cxy[jl][il] += cijk[cidx(lp, kl, jl, il)] * dkp;
            dkp *= dk;
        }
        for (int j = g.index_min[1]; j <= g.index_max[1]; j++) {
            const double dj = j - g.gp[1];
            const int gj = general_map(t, &g, 1, j);
            double cx[LPMAX + 1];
            for (int il = 0; il <= lp; il++) cx[il] = 0.0;
            double djp = 1.0;
            for (int jl = 0; jl <= lp; jl++) {
                for (int il = 0; il <= lp - jl; il++) cx[il] += cxy[jl][il] * djp;
                djp *= dj;
            }
            int ismin, ismax;
            double qa, qb, qc;
            if (!general_row(t, &g, s->zetp, dj, dk, &ismin, &ismax, &qa, &qb, &qc)) continue;
            double *row = grid + ((size_t)gk * ny + gj) * nx;
            for (int i = ismin; i <= ismax; i++) {
                const double di = i - g.gp[0];
                double res = 0.0, dip = 1.0;
                for (int il = 0; il <= lp; il++) {
                    res += cx[il] * dip;
                    dip *= di;
                }
                res *= exp((qa * i + qb) * i + qc);
                row[general_map(t, &g, 0, i)] += res;
            }
        }
    }
    return ORC_OK;
}

static int moments_general(const orc_task *t, const pair_setup *s, int lp, const double *grid,
                           double *M) {
    general_geom g;
    int err = general_setup(t, s->rp, &g);
    if (err) return err;
    double mijk[1024];
    for (int i = 0; i < orc_ncoset(lp); i++) mijk[i] = 0.0;
    const size_t nx = t->ngrid[0], ny = t->ngrid[1];
    for (int k = g.index_min[2]; k <= g.index_max[2]; k++) {
        const double dk = k - g.gp[2];
        const int gk = general_map(t, &g, 2, k);
        double txy[LPMAX + 1][LPMAX + 1];
        memset(txy, 0, sizeof(txy));
        for (int j = g.index_min[1]; j <= g.index_max[1]; j++) {
            const double dj = j - g.gp[1];
            const int gj = general_map(t, &g, 1, j);
            int ismin, ismax;
            double qa, qb, qc;
            if (!general_row(t, &g, s->zetp, dj, dk, &ismin, &ismax, &qa, &qb, &qc)) continue;
            const double *row = grid + ((size_t)gk * ny + gj) * nx;
            double tx[LPMAX + 1];
            for (int il = 0; il <= lp; il++) tx[il] = 0.0;
            for (int i = ismin; i <= ismax; i++) {
                const double di = i - g.gp[0];
                const double v = row[general_map(t, &g, 0, i)] * exp((qa * i + qb) * i + qc);
                double dip = 1.0;
                for (int il = 0; il <= lp; il++) {
                    tx[il] += v * dip;
                    dip *= di;
                }
            }
            double djp = 1.0;
            for (int jl = 0; jl <= lp; jl++) {
                for (int il = 0; il <= lp - jl; il++) txy[jl][il] += tx[il] * djp;
                djp *= dj;
            }
        }
        double dkp = 1.0;
        for (int kl = 0; kl <= lp; kl++) {
            for (int jl = 0; jl <= lp - kl; jl++)
                for (int il = 0; il <= lp - kl - jl; il++) mijk[cidx(lp, kl, jl, il)] += txy[jl][il] * dkp;
            dkp *= dk;
        }
    }
    multinomial(t, lp, mijk, M, 1);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* public single-task entry points                                           */

int orc_coef_xyz(const orc_task *t, const double *pab, double *coef, int *lp_out) {
    pair_setup s;
    int err = setup_pair(t, 0, &s);
    if (err) return err;
    alpha_t alpha;
    fill_alpha(t->ra, s.rb, s.rp, s.la_max, s.lb_max, alpha);
    const int ld_eff = orc_ncoset(ORC_LMAX);
    double *eff = (double *)calloc((size_t)ld_eff * ld_eff, sizeof(double));
    if (!eff) return ORC_ERR_ALLOC;
    prepare_pab(t, &s, pab, eff, ld_eff);
    coef_from_eff(&s, alpha, eff, ld_eff, coef);
    free(eff);
    if (lp_out) *lp_out = s.lp;
    return ORC_OK;
}

int orc_collocate(const orc_task *t, const double *pab, double *grid) {
    pair_setup s;
    int err = setup_pair(t, 0, &s);
    if (err) return err;
    double coef[1024];
    if ((err = orc_coef_xyz(t, pab, coef, NULL))) return err;
    return t->use_ortho ? collocate_ortho(t, &s, coef, grid) : collocate_general(t, &s, coef, grid);
}

int orc_integrate(const orc_task *t, const double *grid, const double *pab, double *hab,
                  double *forces) {
    pair_setup s;
    const int want_f = forces != NULL;
    int err = setup_pair(t, want_f, &s);
    if (err) return err;
    double M[1024];
    err = t->use_ortho ? moments_ortho(t, &s, s.lp_x, grid, M)
                       : moments_general(t, &s, s.lp_x, grid, M);
    if (err) return err;
    alpha_t alpha;
    fill_alpha(t->ra, s.rb, s.rp, s.la_max_x, s.lb_max_x, alpha);
    const int ld = orc_ncoset(ORC_LMAX);
    double *heff = (double *)calloc((size_t)ld * ld, sizeof(double));
    double *eff = (double *)calloc((size_t)ld * ld, sizeof(double));
    if (!heff || !eff) {
        free(heff);
        free(eff);
        return ORC_ERR_ALLOC;
    }
    heff_from_moments(&s, alpha, M, heff, ld);
    if (hab) prepare_pab_adjoint(t, heff, ld, hab);
    if (want_f) {
        prepare_pab(t, &s, pab, eff, ld);
        forces_from_heff(t, &s, eff, ld, heff, ld, forces);
    }
    free(heff);
    free(eff);
    return ORC_OK;
}

/* phi_a and its gradient at displacement r from the centre */
static void cart_gauss(const int a[3], double zeta, const double r[3], double *phi,
                       double grad[3]) {
    double pw[3], pwm[3], pwp[3];
    for (int d = 0; d < 3; d++) {
        pw[d] = 1.0;
        for (int k = 0; k < a[d]; k++) pw[d] *= r[d];
        pwm[d] = 1.0;
        for (int k = 0; k < a[d] - 1; k++) pwm[d] *= r[d];
        pwp[d] = pw[d] * r[d];
    }
    const double e = exp(-zeta * (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]));
    *phi = pw[0] * pw[1] * pw[2] * e;
    for (int d = 0; d < 3; d++) {
        const double dd = (a[d] > 0 ? a[d] * pwm[d] : 0.0) - 2.0 * zeta * pwp[d];
        grad[d] = dd * pw[(d + 1) % 3] * pw[(d + 2) % 3] * e;
    }
}

static double brute_value(const orc_task *t, const pair_setup *s, const double *pab,
                          const double rel[3]) {
    /* rel = r - rp */
    const int ld = t->pab_ld > 0 ? t->pab_ld : t->maxco;
    double ra_rel[3], rb_rel[3];
    for (int d = 0; d < 3; d++) {
        ra_rel[d] = rel[d] + (s->rp[d] - t->ra[d]);
        rb_rel[d] = rel[d] + (s->rp[d] - s->rb[d]);
    }
    double acc = 0.0;
    for (int la = t->la_min; la <= t->la_max; la++)
        for (int lxa = 0; lxa <= la; lxa++)
            for (int lya = 0; lya <= la - lxa; lya++) {
                const int a[3] = {lxa, lya, la - lxa - lya};
                double pa, ga[3];
                cart_gauss(a, t->zeta, ra_rel, &pa, ga);
                const int ico = orc_coset(a[0], a[1], a[2]);
                for (int lb = t->lb_min; lb <= t->lb_max; lb++)
                    for (int lxb = 0; lxb <= lb; lxb++)
                        for (int lyb = 0; lyb <= lb - lxb; lyb++) {
                            const int b[3] = {lxb, lyb, lb - lxb - lyb};
                            double pb, gb[3];
                            cart_gauss(b, t->zetb, rb_rel, &pb, gb);
                            const int jco = orc_coset(b[0], b[1], b[2]);
                            const double p = pab[(t->o2 + jco) * ld + t->o1 + ico];
                            if (t->compute_tau)
                                acc += p * 0.5 * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
                            else
                                acc += p * pa * pb;
                        }
            }
    return t->rscale * acc;
}

int orc_collocate_brute(const orc_task *t, const double *pab, double *grid) {
    pair_setup s;
    int err = setup_pair(t, 0, &s);
    if (err) return err;
    const size_t nx = t->ngrid[0], ny = t->ngrid[1];
    if (t->use_ortho) {
        ortho_geom g;
        if ((err = ortho_setup(t, s.rp, &g))) return err;
        if ((err = check_sphere(t))) return err;
        const int *sb = t->sphere_bounds;
        int sci = 0;
        const int kgmin = sb[sci++];
        for (int kg = kgmin; kg <= 0; kg++) {
            const int jgmin = sb[sci++];
            const int *igmins = sb + sci;
            sci += 1 - jgmin;
            for (int jg = jgmin; jg <= 0; jg++) {
                const int igmin = igmins[jg - jgmin];
                for (int ig = igmin; ig <= 0; ig++)
                    for (int sk = 0; sk < 2; sk++)
                        for (int sj = 0; sj < 2; sj++)
                            for (int si = 0; si < 2; si++) {
                                const int off[3] = {si ? 1 - ig : ig, sj ? 1 - jg : jg,
                                                    sk ? 1 - kg : kg};
                                double rel[3];
                                for (int d = 0; d < 3; d++) rel[d] = off[d] * g.dr[d] - g.roff[d];
                                const double v = brute_value(t, &s, pab, rel);
                                grid[((size_t)ortho_map(t, &g, 2, off[2]) * ny +
                                      ortho_map(t, &g, 1, off[1])) * nx +
                                     ortho_map(t, &g, 0, off[0])] += v;
                            }
            }
        }
        return ORC_OK;
    }
    general_geom g;
    if ((err = general_setup(t, s.rp, &g))) return err;
    for (int k = g.index_min[2]; k <= g.index_max[2]; k++)
        for (int j = g.index_min[1]; j <= g.index_max[1]; j++) {
            const double dk = k - g.gp[2], dj = j - g.gp[1];
            int ismin, ismax;
            double qa, qb, qc;
            if (!general_row(t, &g, s.zetp, dj, dk, &ismin, &ismax, &qa, &qb, &qc)) continue;
            for (int i = ismin; i <= ismax; i++) {
                const double di = i - g.gp[0];
                double rel[3];
                for (int d = 0; d < 3; d++)
                    rel[d] = di * t->dh[0][d] + dj * t->dh[1][d] + dk * t->dh[2][d];
                grid[((size_t)general_map(t, &g, 2, k) * ny + general_map(t, &g, 1, j)) * nx +
                     general_map(t, &g, 0, i)] += brute_value(t, &s, pab, rel);
            }
        }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* cube / sphere bounds                                                      */

/* largest n >= 0 with base + (n*h)^2 <= R2 */
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

int orc_sphere_bounds(double radius, const double dr[3], int lb_cube[3], int ub_cube[3],
                      int *sb, int cap) {
    double drmin = dr[0];
    if (dr[1] < drmin) drmin = dr[1];
    if (dr[2] < drmin) drmin = dr[2];
    int imr = (int)ceil(radius / drmin);
    if (imr < 1) imr = 1;
    const double R = imr * drmin;
    const double R2 = R * R;
    for (int d = 0; d < 3; d++) {
        const int n = sphere_extent(0.0, dr[d], R2);
        lb_cube[d] = -n;
        ub_cube[d] = 1 + n;
    }
    int count = 0;
    const int kgmin = lb_cube[2];
    if (count < cap) sb[count] = kgmin;
    count++;
    for (int kg = kgmin; kg <= 0; kg++) {
        const double zk = kg * dr[2];
        const double base = zk * zk;
        const int jgmin = -sphere_extent(base, dr[1], R2);
        if (count < cap) sb[count] = jgmin;
        count++;
        for (int jg = jgmin; jg <= 0; jg++) {
            const double yj = jg * dr[1];
            const double b2 = base + yj * yj;
            if (count < cap) sb[count] = -sphere_extent(b2, dr[0], R2);
            count++;
        }
    }
    return count <= cap ? count : -count;
}

/* ------------------------------------------------------------------------ */
/* libgpw tensor-path contractions                                           */

void orc_collocate_core(const int length[3], int L, const double *co, const double *P, int ldp,
                        double *V) {
    const double *Pz = P, *Py = P + (size_t)L * ldp, *Px = P + (size_t)2 * L * ldp;
    for (int a = 0; a < L; a++)
        for (int g = 0; g < L; g++)
            for (int b = 0; b < L; b++) {
                const double c = co[((size_t)a * L + g) * L + b];
                for (int z = 0; z < length[0]; z++) {
                    const double c1 = c * Pz[(size_t)g * ldp + z];
                    for (int y = 0; y < length[1]; y++) {
                        const double c2 = c1 * Py[(size_t)b * ldp + y];
                        double *row = V + ((size_t)z * length[1] + y) * length[2];
                        for (int x = 0; x < length[2]; x++) row[x] += c2 * Px[(size_t)a * ldp + x];
                    }
                }
            }
}

void orc_integrate_core(const int length[3], int L, const double *P, int ldp, const double *V,
                        double *Velems) {
    const double *Pz = P, *Py = P + (size_t)L * ldp, *Px = P + (size_t)2 * L * ldp;
    for (int g = 0; g < L; g++)
        for (int b = 0; b < L; b++)
            for (int a = 0; a < L; a++) {
                double res = 0.0;
                for (int z = 0; z < length[0]; z++)
                    for (int y = 0; y < length[1]; y++) {
                        const double c1 = Pz[(size_t)g * ldp + z] * Py[(size_t)b * ldp + y];
                        const double *row = V + ((size_t)z * length[1] + y) * length[2];
                        for (int x = 0; x < length[2]; x++) res += c1 * Px[(size_t)a * ldp + x] * row[x];
                    }
                Velems[((size_t)g * L + b) * L + a] = res;
            }
}

/* ------------------------------------------------------------------------ */
/* libgpw folded-table path                                                  */
/*
 * Periodic folding of 1D tables, compute_folded_polynomials (5-argument form,
 * collocate.cc:276-301) called per direction and power by
 * compute_folded_polynomials_all_powers (collocate.cc:308-331): entry i of an
 * unfolded table of `length` points is the grid coordinate
 * center - (length - 1) / 2 + i, and the folded table is the image sum
 * F[g] = sum over i with (center - (length - 1) / 2 + i) mod period == g.
 * The support mask is compute_nonzero_values (collocate.cc:32-57): all ones
 * when length >= period, else ones on those grid coordinates. Defects of the
 * reference not carried over: the (length + 1) / 2 centring of the other
 * overload (collocate.cc:140,162), the negative remainder of
 * (i + p0 + period) % period when p0 < -period, the integer `16 * period` /
 * `32 * period` offsets that bound how far a table may reach.
 */
static int pmod(long long a, int m) {
    long long r = a % m;
    return (int)(r < 0 ? r + m : r);
}

void orc_fold_tables(int L, const int length[3], const int center[3], const int period[3],
                     const double *unfolded, int ldu, double *folded, int ldp, int *mask,
                     int ldm) {
    for (int d = 0; d < 3; d++) {
        const long long p0 = (long long)center[d] - (length[d] - 1) / 2;
        for (int a = 0; a < L; a++) {
            double *f = folded + ((size_t)d * L + a) * ldp;
            const double *u = unfolded + ((size_t)d * L + a) * ldu;
            for (int g = 0; g < ldp; g++) f[g] = 0.0;
            for (int i = 0; i < length[d]; i++) f[pmod(p0 + i, period[d])] += u[i];
        }
        if (mask) {
            int *m = mask + (size_t)d * ldm;
            for (int g = 0; g < ldm; g++) m[g] = 0;
            if (length[d] >= period[d]) {
                for (int g = 0; g < period[d] && g < ldm; g++) m[g] = 1;
            } else {
                for (int i = 0; i < length[d]; i++) m[pmod(p0 + i, period[d])] = 1;
            }
        }
    }
}

/*
 * collocate<T> (collocate.cc:681-745): the coefficient cube contracted with
 * the folded tables over the points of the local grid window whose three
 * masks are set -- pack the support (compute_polynomials_non_zero_elems,
 * polynomials.hpp:5-86), collocate_core on it, add it back
 * (unpack_elems_and_add, utils.hpp:102-181):
 *   density[z][y][x] += sum co[a][g][b] P[0][g][c0+z] P[1][b][c1+y] P[2][a][c2+x]
 * for 0 <= z < grid_size[0] etc. and mask[d][c_d + .] != 0, c = lower_corner.
 * Defects not carried over: the lower corner counted twice
 * (polynomials.hpp:21-22), the loop bounds of pack / unpack on the wrong axes
 * (utils.hpp:60,81,127,140,161). Vtmp (the packed scratch) is not needed.
 */
void orc_collocate_folded(int L, const double *co, const double *P, int ldp, const int *mask,
                          int ldm, const int grid_size[3], const int lower_corner[3],
                          double *density) {
    const double *Pz = P, *Py = P + (size_t)L * ldp, *Px = P + (size_t)2 * L * ldp;
    const int *mz = mask, *my = mask + ldm, *mx = mask + 2 * (size_t)ldm;
    const int c0 = lower_corner[0], c1 = lower_corner[1], c2 = lower_corner[2];
    for (int z = 0; z < grid_size[0]; z++) {
        if (!mz[c0 + z]) continue;
        for (int y = 0; y < grid_size[1]; y++) {
            if (!my[c1 + y]) continue;
            double *row = density + ((size_t)z * grid_size[1] + y) * grid_size[2];
            for (int x = 0; x < grid_size[2]; x++) {
                if (!mx[c2 + x]) continue;
                double v = 0.0;
                for (int a = 0; a < L; a++)
                    for (int g = 0; g < L; g++)
                        for (int b = 0; b < L; b++)
                            v += co[((size_t)a * L + g) * L + b] * Pz[(size_t)g * ldp + c0 + z] *
                                 Py[(size_t)b * ldp + c1 + y] * Px[(size_t)a * ldp + c2 + x];
                row[x] += v;
            }
        }
    }
}

/*
 * integrate<T> (collocate.cc:626-679; its call of pack_elems does not
 * compile, collocate.cc:674): the potential on the masked support of the
 * local window packed (pack_elems, utils.hpp:20-99) and contracted by
 * integrate_core:
 *   co[g][b][a] = sum V[z][y][x] P[0][g][c0+z] P[1][b][c1+y] P[2][a][c2+x]
 * over the same points as orc_collocate_folded; co is overwritten (zero when
 * no point is set, where the reference would leave it untouched).
 */
void orc_integrate_folded(int L, const double *P, int ldp, const int *mask, int ldm,
                          const int grid_size[3], const int lower_corner[3], const double *V,
                          double *co) {
    const double *Pz = P, *Py = P + (size_t)L * ldp, *Px = P + (size_t)2 * L * ldp;
    const int *mz = mask, *my = mask + ldm, *mx = mask + 2 * (size_t)ldm;
    const int c0 = lower_corner[0], c1 = lower_corner[1], c2 = lower_corner[2];
    for (size_t i = 0; i < (size_t)L * L * L; i++) co[i] = 0.0;
    for (int z = 0; z < grid_size[0]; z++) {
        if (!mz[c0 + z]) continue;
        for (int y = 0; y < grid_size[1]; y++) {
            if (!my[c1 + y]) continue;
            const double *row = V + ((size_t)z * grid_size[1] + y) * grid_size[2];
            for (int x = 0; x < grid_size[2]; x++) {
                if (!mx[c2 + x]) continue;
                for (int g = 0; g < L; g++)
                    for (int b = 0; b < L; b++)
                        for (int a = 0; a < L; a++)
                            co[((size_t)g * L + b) * L + a] += row[x] * Pz[(size_t)g * ldp + c0 + z] *
                                                              Py[(size_t)b * ldp + c1 + y] *
                                                              Px[(size_t)a * ldp + c2 + x];
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* batches                                                                   */

static int make_task(const orc_grid *g, const orc_pair *p, int compute_tau, orc_task *t,
                     int **sb_buf, int *sb_cap) {
    memset(t, 0, sizeof(*t));
    t->compute_tau = compute_tau;
    t->use_ortho = g->orthorhombic;
    t->la_max = p->la_max;
    t->la_min = p->la_min;
    t->lb_max = p->lb_max;
    t->lb_min = p->lb_min;
    t->zeta = p->zeta;
    t->zetb = p->zetb;
    t->rscale = p->rscale;
    memcpy(t->dh, g->dh, sizeof(t->dh));
    memcpy(t->dh_inv, g->dh_inv, sizeof(t->dh_inv));
    for (int d = 0; d < 3; d++) {
        t->ra[d] = p->ra[d];
        t->rab[d] = p->rab[d];
        t->npts[d] = g->npts[d];
        t->ngrid[d] = g->npts[d];
        t->periodic[d] = 1;
    }
    t->lmax = ORC_LMAX;
    t->radius = p->radius;
    t->maxco = orc_ncoset(p->la_max);
    t->pab_ld = orc_ncoset(p->la_max);
    if (g->orthorhombic) {
        const double dr[3] = {g->dh[0][0], g->dh[1][1], g->dh[2][2]};
        int n = orc_sphere_bounds(p->radius, dr, t->lb_cube, t->ub_cube, *sb_buf, *sb_cap);
        if (n < 0) {
            free(*sb_buf);
            *sb_cap = -n;
            *sb_buf = (int *)malloc(sizeof(int) * (size_t)(*sb_cap));
            if (!*sb_buf) return ORC_ERR_ALLOC;
            n = orc_sphere_bounds(p->radius, dr, t->lb_cube, t->ub_cube, *sb_buf, *sb_cap);
        }
        t->nspheres = n;
        t->sphere_bounds = *sb_buf;
    }
    return ORC_OK;
}

int orc_batch_collocate(const orc_grid *g, int npairs, const orc_pair *pairs, int compute_tau,
                        const double *pab, double *grid, int nthreads) {
    const size_t ng = (size_t)g->npts[0] * g->npts[1] * g->npts[2];
    int err = ORC_OK;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    double **priv = (double **)calloc((size_t)nthreads, sizeof(double *));
    if (!priv) return ORC_ERR_ALLOC;
#pragma omp parallel num_threads(nthreads)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *mine = (double *)calloc(ng, sizeof(double));
        priv[tid] = mine;
        int cap = 4096;
        int *sb = (int *)malloc(sizeof(int) * cap);
        if (!mine || !sb) {
#pragma omp atomic write
            err = ORC_ERR_ALLOC;
        }
#pragma omp for schedule(dynamic, 8)
        for (int i = 0; i < npairs; i++) {
            if (!mine || !sb) continue;
            orc_task t;
            int e = make_task(g, &pairs[i], compute_tau, &t, &sb, &cap);
            if (!e) e = orc_collocate(&t, pab + pairs[i].block_offset, mine);
            if (e) {
#pragma omp atomic write
                err = e;
            }
        }
        free(sb);
#pragma omp for schedule(static)
        for (long long q = 0; q < (long long)ng; q++) {
            double acc = 0.0;
            for (int th = 0; th < nthreads; th++)
                if (priv[th]) acc += priv[th][q];
            grid[q] += acc;
        }
    }
    for (int th = 0; th < nthreads; th++) free(priv[th]);
    free(priv);
    return err;
}

int orc_batch_integrate(const orc_grid *g, int npairs, const orc_pair *pairs, int compute_tau,
                        const double *grid, const double *pab, double *hab, double *forces,
                        int nthreads) {
    int err = ORC_OK;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
#pragma omp parallel num_threads(nthreads)
    {
        int cap = 4096;
        int *sb = (int *)malloc(sizeof(int) * cap);
#pragma omp for schedule(dynamic, 8)
        for (int i = 0; i < npairs; i++) {
            if (!sb) continue;
            orc_task t;
            int e = make_task(g, &pairs[i], compute_tau, &t, &sb, &cap);
            if (!e)
                e = orc_integrate(&t, grid, pab ? pab + pairs[i].block_offset : NULL,
                                  hab + pairs[i].block_offset, forces ? forces + 6 * (size_t)i : NULL);
            if (e) {
#pragma omp atomic write
                err = e;
            }
        }
        if (!sb) {
#pragma omp atomic write
            err = ORC_ERR_ALLOC;
        }
        free(sb);
    }
    return err;
}
