// Ordinate-lane sweep instantiated for double precision, DG degrees 0-2 (P = 1..3).
#include "sweep_lanes.hpp"

namespace tsw {
TSW_LANES_INSTANTIATE(double, 1)
TSW_LANES_INSTANTIATE(double, 2)
TSW_LANES_INSTANTIATE(double, 3)
}  // namespace tsw
