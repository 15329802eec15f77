// Staggered sweep instantiated for float precision, DG degree 3 (P = 4).
#include "sweep_kernel.hpp"

namespace tsw {
TSW_SWEEP_INSTANTIATE(float, 4)
}  // namespace tsw
