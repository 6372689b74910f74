// gpw_tiles_colloc_ortho.hip -- the collocate tile kernels of orthorhombic
// cells, compiled in a translation unit of their own.
#include "gpw_host.hpp"
#include "gpw_tiles_colloc.hpp"

namespace gpw {
template hipError_t launch_collocate_tiles_t<false>(int, const TileGrid &, const EntryRec *,
                                                      const int *, const double *, int, int, double *,
                                                      int, hipStream_t);
}  // namespace gpw
