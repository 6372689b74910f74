// gpw_folded.cpp -- C ABI of the libgpw folded-table path: periodic folding
// of 1D tables with their support masks (compute_folded_polynomials_all_powers
// and compute_nonzero_values, src/lib/collocate.cc:308-331, :32-57) and the
// collocate<T> / integrate<T> wrappers over the masked support of a local grid
// window (src/lib/collocate.cc:681-745, :626-679), on device arrays and on
// host arrays (staged through the calling thread's device workspace).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "gpw_host.hpp"

using namespace gpw;

namespace {

size_t up256(size_t n) { return (n + 255) & ~(size_t)255; }

gpw_status check_L(int L) {
    if (L < 1 || L > LPMAX + 1) return fail(GPW_ERR_ARG, "L outside 1 .. 2 GPW_LMAX + 1");
    return GPW_OK;
}

gpw_status check_fold(int L, const int *length, const int *center, const int *period, int lu, int ldu,
                      int ldp, bool with_mask, int ldm) {
    if (!length || !center || !period) return fail(GPW_ERR_ARG, "NULL argument");
    gpw_status st = check_L(L);
    if (st) return st;
    if (lu < L) return fail(GPW_ERR_ARG, "fewer unfolded rows per direction than L");
    for (int d = 0; d < 3; d++) {
        if (period[d] < 1 || length[d] < 0) return fail(GPW_ERR_ARG, "period must be > 0, length >= 0");
        if (length[d] > ldu) return fail(GPW_ERR_ARG, "unfolded table longer than its row");
        if (period[d] > ldp) return fail(GPW_ERR_ARG, "folded row shorter than the period");
        if (with_mask && period[d] > ldm) return fail(GPW_ERR_ARG, "mask row shorter than the period");
    }
    return GPW_OK;
}

gpw_status check_window(int L, int ldp, int ldm, const int *gs, const int *lc) {
    if (!gs || !lc) return fail(GPW_ERR_ARG, "NULL argument");
    gpw_status st = check_L(L);
    if (st) return st;
    for (int d = 0; d < 3; d++) {
        if (gs[d] < 0 || lc[d] < 0) return fail(GPW_ERR_ARG, "grid_size and lower corner must be >= 0");
        if ((long long)lc[d] + gs[d] > ldp || (long long)lc[d] + gs[d] > ldm)
            return fail(GPW_ERR_ARG, "local window beyond the folded tables or masks");
    }
    return GPW_OK;
}

}  // namespace

extern "C" {

gpw_status gpw_fold_tables(int L, const int *length, const int *center, const int *period,
                           const double *unfolded, int lu, int ldu, double *folded, int ldp, int *mask,
                           int ldm, void *stream) {
    if (!unfolded || !folded) return fail(GPW_ERR_ARG, "NULL argument");
    gpw_status st = check_fold(L, length, center, period, lu, ldu, ldp, mask != nullptr, ldm);
    if (st) return st;
    HIP_TRY(launch_fold_tables(L, length, center, period, unfolded, lu, ldu, folded, ldp, mask, ldm,
                               (hipStream_t)stream));
    return GPW_OK;
}

gpw_status gpw_collocate_folded(int L, const double *co, const double *P, int ldp, const int *mask,
                                int ldm, const int *grid_size, const int *lower_corner, double *density,
                                void *stream) {
    if (!co || !P || !mask || !density) return fail(GPW_ERR_ARG, "NULL argument");
    gpw_status st = check_window(L, ldp, ldm, grid_size, lower_corner);
    if (st) return st;
    HIP_TRY(launch_collocate_folded(L, co, P, ldp, mask, ldm, grid_size, lower_corner, density,
                                    (hipStream_t)stream));
    return GPW_OK;
}

gpw_status gpw_integrate_folded(int L, const double *P, int ldp, const int *mask, int ldm,
                                const int *grid_size, const int *lower_corner, const double *V, double *co,
                                void *stream) {
    if (!co || !P || !mask || !V) return fail(GPW_ERR_ARG, "NULL argument");
    gpw_status st = check_window(L, ldp, ldm, grid_size, lower_corner);
    if (st) return st;
    HIP_TRY(launch_integrate_folded(L, P, ldp, mask, ldm, grid_size, lower_corner, V, co,
                                    (hipStream_t)stream));
    return GPW_OK;
}

gpw_status gpw_fold_tables_host(int L, const int *length, const int *center, const int *period,
                                const double *unfolded, int lu, int ldu, double *folded, int ldp,
                                int *mask, int ldm) {
    if (!unfolded || !folded) return fail(GPW_ERR_ARG, "NULL argument");
    gpw_status st = check_fold(L, length, center, period, lu, ldu, ldp, mask != nullptr, ldm);
    if (st) return st;
    const size_t nu = (size_t)3 * lu * ldu, nf = (size_t)3 * L * ldp, nm = mask ? (size_t)3 * ldm : 0;
    const size_t ou = 0, of = up256(nu * sizeof(double)), om = of + up256(nf * sizeof(double));
    char *R = nullptr;
    if ((st = thread_workspace(om + nm * sizeof(int) + 256, nullptr, &R))) return st;
    HIP_TRY(hipMemcpy(R + ou, unfolded, nu * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(launch_fold_tables(L, length, center, period, (const double *)(R + ou), lu, ldu,
                               (double *)(R + of), ldp, mask ? (int *)(R + om) : nullptr, ldm, nullptr));
    HIP_TRY(hipMemcpy(folded, R + of, nf * sizeof(double), hipMemcpyDeviceToHost));
    if (mask) HIP_TRY(hipMemcpy(mask, R + om, nm * sizeof(int), hipMemcpyDeviceToHost));
    return GPW_OK;
}

gpw_status gpw_collocate_folded_host(int L, const double *co, const double *P, int ldp, const int *mask,
                                     int ldm, const int *grid_size, const int *lower_corner,
                                     double *density) {
    if (!co || !P || !mask || !density) return fail(GPW_ERR_ARG, "NULL argument");
    gpw_status st = check_window(L, ldp, ldm, grid_size, lower_corner);
    if (st) return st;
    const size_t nc = (size_t)L * L * L, np = (size_t)3 * L * ldp, nm = (size_t)3 * ldm;
    const size_t nd = (size_t)grid_size[0] * grid_size[1] * grid_size[2];
    const size_t oc = 0, op = up256(nc * 8), om = op + up256(np * 8), od = om + up256(nm * 4);
    char *R = nullptr;
    if ((st = thread_workspace(od + nd * 8 + 256, nullptr, &R))) return st;
    HIP_TRY(hipMemcpy(R + oc, co, nc * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(R + op, P, np * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(R + om, mask, nm * 4, hipMemcpyHostToDevice));
    // the device adds into zeros; the host adds the result into density
    HIP_TRY(launch_fill((double *)(R + od), (long long)nd, 0.0, nullptr));
    HIP_TRY(launch_collocate_folded(L, (const double *)(R + oc), (const double *)(R + op), ldp,
                                    (const int *)(R + om), ldm, grid_size, lower_corner,
                                    (double *)(R + od), nullptr));
    std::vector<double> v(nd);
    HIP_TRY(hipMemcpy(v.data(), R + od, nd * 8, hipMemcpyDeviceToHost));
    const int *mz = mask, *my = mask + ldm, *mx = mask + 2 * (size_t)ldm;
    for (int z = 0; z < grid_size[0]; z++) {
        if (!mz[lower_corner[0] + z]) continue;
        for (int y = 0; y < grid_size[1]; y++) {
            if (!my[lower_corner[1] + y]) continue;
            const size_t row = ((size_t)z * grid_size[1] + y) * grid_size[2];
            for (int x = 0; x < grid_size[2]; x++)
                if (mx[lower_corner[2] + x]) density[row + x] += v[row + x];
        }
    }
    return GPW_OK;
}

gpw_status gpw_integrate_folded_host(int L, const double *P, int ldp, const int *mask, int ldm,
                                     const int *grid_size, const int *lower_corner, const double *V,
                                     double *co) {
    if (!co || !P || !mask || !V) return fail(GPW_ERR_ARG, "NULL argument");
    gpw_status st = check_window(L, ldp, ldm, grid_size, lower_corner);
    if (st) return st;
    const size_t nc = (size_t)L * L * L, np = (size_t)3 * L * ldp, nm = (size_t)3 * ldm;
    const size_t nv = (size_t)grid_size[0] * grid_size[1] * grid_size[2];
    const size_t oc = 0, op = up256(nc * 8), om = op + up256(np * 8), ov = om + up256(nm * 4);
    char *R = nullptr;
    if ((st = thread_workspace(ov + nv * 8 + 256, nullptr, &R))) return st;
    HIP_TRY(hipMemcpy(R + op, P, np * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(R + om, mask, nm * 4, hipMemcpyHostToDevice));
    if (nv) HIP_TRY(hipMemcpy(R + ov, V, nv * 8, hipMemcpyHostToDevice));
    HIP_TRY(launch_integrate_folded(L, (const double *)(R + op), ldp, (const int *)(R + om), ldm,
                                    grid_size, lower_corner, (const double *)(R + ov), (double *)(R + oc),
                                    nullptr));
    HIP_TRY(hipMemcpy(co, R + oc, nc * 8, hipMemcpyDeviceToHost));
    return GPW_OK;
}

// Support mask of one direction (compute_nonzero_values, collocate.cc:32-57):
// mask[g] = 1 for the grid coordinates center - (length - 1) / 2 + i mod
// period, i < length (every g < period when length >= period), 0 elsewhere
// up to ldm. Host-only index arithmetic.
gpw_status gpw_nonzero_values(int center, int period, int length, int *mask, int ldm) {
    if (!mask || period < 1 || length < 0 || period > ldm)
        return fail(GPW_ERR_ARG, "bad arguments (period must be 1 .. ldm)");
    for (int g = 0; g < ldm; g++) mask[g] = 0;
    if (length >= period) {
        for (int g = 0; g < period; g++) mask[g] = 1;
        return GPW_OK;
    }
    const long long p0 = (long long)center - (length - 1) / 2;
    for (int i = 0; i < length; i++) {
        long long r = (p0 + i) % period;
        mask[r < 0 ? r + period : r] = 1;
    }
    return GPW_OK;
}

}  // extern "C"
