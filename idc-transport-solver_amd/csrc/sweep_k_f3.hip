// Staggered sweep instantiated for float precision, DG degree 2 (P = 3).
#include "sweep_kernel.hpp"

namespace tsw {
TSW_SWEEP_INSTANTIATE(float, 3)
}  // namespace tsw
