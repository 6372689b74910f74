// gpw_tiles_integ_general.hip -- the integrate tile kernels of general
// (non-orthorhombic) cells, compiled in a translation unit of their own.
#include "gpw_host.hpp"
#include "gpw_tiles_integ.hpp"

namespace gpw {
template hipError_t launch_integrate_tiles_t<true>(int, const TileGrid &, const EntryRec *,
                                                   const int *, const double *, int, int, int,
                                                   double *, hipStream_t);
}  // namespace gpw
