// Staggered sweep instantiated for float precision, DG degree 0 (P = 1).
#include "sweep_kernel.hpp"

namespace tsw {
TSW_SWEEP_INSTANTIATE(float, 1)
}  // namespace tsw
