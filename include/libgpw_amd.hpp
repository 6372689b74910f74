// libgpw_amd.hpp -- C++ drop-in wrappers with the reference's signatures.
//
// A caller of the reference includes this header instead of grid_base_ref_c.h /
// collocate.hpp and links libgpw_amd.so. The array arguments are templates so
// the reference's own mdarray<T, N, CblasRowMajor> (row-major, contiguous,
// operator() returning T&, size(i) in declaration order) is accepted as is,
// without this header depending on mdarray.hpp or a BLAS header.
//
// Signatures replaced (paths relative to the reference root):
//   grid_collocate_pgf_product_rspace  src/lib/grid_base_ref_c.h:40-67
//   collocate_core<T>                  src/lib/collocate.cc:543 (decl. src/lib/collocate.hpp:4-7)
//   integrate_core<T>                  src/lib/collocate.cc:438
//   collocate<T> / integrate<T>        src/lib/collocate.cc:681-745 / :626-679
//   compute_folded_polynomials_all_powers, compute_nonzero_values
//                                      src/lib/collocate.cc:308-331 / :32-57
// and adds grid_integrate_pgf_product_rspace, the adjoint with the same
// argument list (no counterpart in the reference).
//
// Errors: the reference asserts or aborts; these wrappers throw
// gpw::error carrying the library's status code.
#ifndef LIBGPW_AMD_HPP
#define LIBGPW_AMD_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "libgpw_amd.h"

namespace gpw {

struct error : std::runtime_error {
    int code;
    error(int c, const char *what) : std::runtime_error(what), code(c) {}
};

inline void check(gpw_status s) {
    if (s != GPW_OK) throw error(s, gpw_last_error());
}

}  // namespace gpw

template <class Pab, class Grid>
void grid_collocate_pgf_product_rspace(
    const bool compute_tau, const bool use_ortho, const int la_max, const int la_min,
    const int lb_max, const int lb_min, const double zeta, const double zetb, const double rscale,
    const double dh[3][3], const double dh_inv[3][3], const double ra[3], const double rab[3],
    const int npts[3], const int ngrid[3], const int lb_grid[3], const bool periodic[3],
    const int lmax, const double radius, const int lb_cube[3], const int ub_cube[3],
    const int nspheres, const std::vector<int> &sphere_bounds, const int maxco, const int o1,
    const int o2, const Pab &pab, Grid &grid) {
    const int per[3] = {periodic[0], periodic[1], periodic[2]};
    gpw::check(gpw_collocate_pgf_product_rspace(
        compute_tau, use_ortho, la_max, la_min, lb_max, lb_min, zeta, zetb, rscale, &dh[0][0],
        &dh_inv[0][0], ra, rab, npts, ngrid, lb_grid, per, lmax, radius, lb_cube, ub_cube,
        nspheres, sphere_bounds.empty() ? nullptr : sphere_bounds.data(), maxco, o1, o2,
        &pab(0, 0), &grid(0, 0, 0)));
}

// hab(o2+jco, o1+ico) += d<grid, collocate(pab)>/d pab(o2+jco, o1+ico);
// forces (if non-null, 6 doubles) += dE/dRa, dE/dRb.
template <class Grid, class Pab, class Hab>
void grid_integrate_pgf_product_rspace(
    const bool compute_tau, const bool use_ortho, const int la_max, const int la_min,
    const int lb_max, const int lb_min, const double zeta, const double zetb, const double rscale,
    const double dh[3][3], const double dh_inv[3][3], const double ra[3], const double rab[3],
    const int npts[3], const int ngrid[3], const int lb_grid[3], const bool periodic[3],
    const int lmax, const double radius, const int lb_cube[3], const int ub_cube[3],
    const int nspheres, const std::vector<int> &sphere_bounds, const int maxco, const int o1,
    const int o2, const Grid &grid, const Pab &pab, Hab &hab, double *forces = nullptr) {
    const int per[3] = {periodic[0], periodic[1], periodic[2]};
    gpw::check(gpw_integrate_pgf_product_rspace(
        compute_tau, use_ortho, la_max, la_min, lb_max, lb_min, zeta, zetb, rscale, &dh[0][0],
        &dh_inv[0][0], ra, rab, npts, ngrid, lb_grid, per, lmax, radius, lb_cube, ub_cube,
        nspheres, sphere_bounds.empty() ? nullptr : sphere_bounds.data(), maxco, o1, o2,
        &grid(0, 0, 0), &pab(0, 0), &hab(0, 0), forces));
}

// V(z, y, x) = sum co(a, g, b) P(0, g, z) P(1, b, y) P(2, a, x)  (Vtmp overwritten)
template <class Co, class Poly, class V>
void collocate_core(const int length[3], const Co &co, const Poly &p_alpha_beta_reduced, V &Vtmp) {
    const int L = (int)co.size(0);
    gpw::check(gpw_collocate_core_host(length, L, &co(0, 0, 0), &p_alpha_beta_reduced(0, 0, 0),
                                       (int)p_alpha_beta_reduced.size(2), &Vtmp(0, 0, 0)));
}

// Velems(g, b, a) = sum V(z, y, x) P(0, g, z) P(1, b, y) P(2, a, x)
template <class Poly, class V, class E>
void integrate_core(const int *length, const Poly &polynomials, const V &Vgrid, E &Velems) {
    const int L = (int)polynomials.size(1);
    gpw::check(gpw_integrate_core_host(length, L, &polynomials(0, 0, 0),
                                       (int)polynomials.size(2), &Vgrid(0, 0, 0),
                                       &Velems(0, 0, 0)));
}

// compute_folded_polynomials_all_powers (src/lib/collocate.cc:308-331):
// p_alpha (3, L + 1, length) unfolded tables (rows a < L are folded; the
// reference keeps x - x_i in the last row), p_alpha_beta_folded (3, L, ldp)
// overwritten with the periodic folds.
template <class Pa, class Pf>
void compute_folded_polynomials_all_powers(const int center[3], const int period[3],
                                           const int length[3], const Pa &p_alpha,
                                           Pf &p_alpha_beta_folded) {
    gpw::check(gpw_fold_tables_host((int)p_alpha_beta_folded.size(1), length, center, period,
                                    &p_alpha(0, 0, 0), (int)p_alpha.size(1), (int)p_alpha.size(2),
                                    &p_alpha_beta_folded(0, 0, 0), (int)p_alpha_beta_folded.size(2),
                                    nullptr, 0));
}

// compute_nonzero_values (src/lib/collocate.cc:32-57): row `dir` of the
// support masks non_zeroes_ (3, ldm)
template <class Mask>
void compute_nonzero_values(const int center, const int period, const int length,
                            const int lower_corner, const int length_dir, const int dir,
                            Mask &non_zeroes_) {
    (void)lower_corner;
    (void)length_dir;
    gpw::check(gpw_nonzero_values(center, period, length, &non_zeroes_(dir, 0),
                                  (int)non_zeroes_.size(1)));
}

// collocate<T> (src/lib/collocate.cc:681-745): density (grid_size[0] x
// grid_size[1] x grid_size[2], the local window at grid_lower_corner) += the
// coefficient cube co (L^3) contracted with the folded tables on the points
// whose masks (3, ldm) are set. Vtmp, the reference's packed scratch, is not
// used.
template <class Co, class Pf, class Mask, class Vt, class Dens>
void collocate(const Co &co, const Pf &p_alpha_beta_folded, const Mask &yz_non_vanishing_elements,
               const int grid_size[3], const int grid_lower_corner[3], Vt &Vtmp, Dens &density) {
    (void)Vtmp;
    gpw::check(gpw_collocate_folded_host(
        (int)co.size(0), &co(0, 0, 0), &p_alpha_beta_folded(0, 0, 0),
        (int)p_alpha_beta_folded.size(2), &yz_non_vanishing_elements(0, 0),
        (int)yz_non_vanishing_elements.size(1), grid_size, grid_lower_corner, &density(0, 0, 0)));
}

// integrate<T> (src/lib/collocate.cc:626-679): co(g, b, a) = the potential V
// (the local window) contracted with the folded tables on the same points
template <class Pf, class Mask, class V, class Co>
void integrate(const Pf &p_alpha_beta_folded, const Mask &yz_non_vanishing_elements,
               const int *grid_size, const int grid_lower_corner[3], const V &Vgrid, Co &co) {
    gpw::check(gpw_integrate_folded_host(
        (int)p_alpha_beta_folded.size(1), &p_alpha_beta_folded(0, 0, 0),
        (int)p_alpha_beta_folded.size(2), &yz_non_vanishing_elements(0, 0),
        (int)yz_non_vanishing_elements.size(1), grid_size, grid_lower_corner, &Vgrid(0, 0, 0),
        &co(0, 0, 0)));
}

#endif
