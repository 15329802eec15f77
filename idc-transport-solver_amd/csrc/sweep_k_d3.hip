// Staggered sweep instantiated for double precision, DG degree 2 (P = 3).
#include "sweep_kernel.hpp"

namespace tsw {
TSW_SWEEP_INSTANTIATE(double, 3)
}  // namespace tsw
