// Ordinate-lane sweep instantiated for float precision, DG degrees 0-2 (P = 1..3).
#include "sweep_lanes.hpp"

namespace tsw {
TSW_LANES_INSTANTIATE(float, 1)
TSW_LANES_INSTANTIATE(float, 2)
TSW_LANES_INSTANTIATE(float, 3)
}  // namespace tsw
