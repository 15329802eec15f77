// Staggered sweep instantiated for double precision, DG degree 1 (P = 2).
#include "sweep_kernel.hpp"

namespace tsw {
TSW_SWEEP_INSTANTIATE(double, 2)
}  // namespace tsw
