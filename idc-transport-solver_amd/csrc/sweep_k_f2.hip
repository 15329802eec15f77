// Staggered sweep instantiated for float precision, DG degree 1 (P = 2).
#include "sweep_kernel.hpp"

namespace tsw {
TSW_SWEEP_INSTANTIATE(float, 2)
}  // namespace tsw
