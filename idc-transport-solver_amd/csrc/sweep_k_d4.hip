// Staggered sweep instantiated for double precision, DG degree 3 (P = 4).
#include "sweep_kernel.hpp"

namespace tsw {
TSW_SWEEP_INSTANTIATE(double, 4)
}  // namespace tsw
