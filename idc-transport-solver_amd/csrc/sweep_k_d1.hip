// Staggered sweep instantiated for double precision, DG degree 0 (P = 1).
#include "sweep_kernel.hpp"

namespace tsw {
TSW_SWEEP_INSTANTIATE(double, 1)
}  // namespace tsw
