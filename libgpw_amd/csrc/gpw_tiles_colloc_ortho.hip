// gpw_tiles_colloc_ortho.hip -- the collocate tile kernels of orthorhombic
// cells, compiled in a translation unit of their own.
#include "gpw_host.hpp"
#include "gpw_tiles_colloc.hpp"

namespace gpw {
template hipError_t launch_collocate_tiles_t<false>(int, const TileGrid &, const EntryRec *,
                                                      const int *, const double *, int, int, double *,
                                                      int, hipStream_t);

// classes 1 (LM 8) and 0 (LM 4) of orthorhombic lists in one launch, into
// split buffers 0 .. splits8 + splits4 - 1
hipError_t launch_collocate_tiles_84(const TileGrid &tg, const EntryRec *recs8, const int *start8,
                                     int splits8, const EntryRec *recs4, const int *start4, int splits4,
                                     const double *coef, int ntiles, double *partial, hipStream_t st) {
    if (tg.general) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_collocate_tiles2<8, 4, false>), dim3(ntiles * (splits8 + splits4)), dim3(NROW),
                       0, st, tg, recs8, start8, splits8, recs4, start4, splits4, coef, ntiles, partial);
    return hipGetLastError();
}
}  // namespace gpw
