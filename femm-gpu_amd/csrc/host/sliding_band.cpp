// Rotor angle of a sliding-band model (see sliding_band.hpp).

#include "sliding_band.hpp"

#include <cmath>
#include <cstdint>

namespace femm
{
namespace host
{

bool RotateAirGaps(const femm_air_gap_t *in, int32_t count, double posed_deg, double angle_deg,
                   std::vector<femm_air_gap_t> *out, std::vector<femm_age_point_t> *points)
{
  const double delta_deg = angle_deg - posed_deg;
  if (!std::isfinite(delta_deg) || count < 0 || (count > 0 && in == nullptr))
    return false;
  size_t total = 0;
  for (int32_t g = 0; g < count; ++g) {
    if (in[g].cells <= 0 || in[g].points == nullptr || !(in[g].arc_length_deg > 0.0))
      return false;
    total += static_cast<size_t>(in[g].cells) + 1;
  }
  out->assign(in, in + count);
  points->clear();
  points->reserve(total);
  for (int32_t g = 0; g < count; ++g) {
    const femm_air_gap_t &src = in[g];
    const int64_t cells = src.cells;
    // FEMM rebuilds the records for every rotor angle; turning the rotor by whole cells
    // moves the inner ring's records by as many positions, the fraction is the shift.
    const double shifted = src.inner_shift + delta_deg * static_cast<double>(cells) /
                                                 src.arc_length_deg;
    const int64_t whole = static_cast<int64_t>(std::floor(shifted));
    (*out)[g].inner_shift = shifted - static_cast<double>(whole);
    const size_t first = points->size();
    points->insert(points->end(), src.points, src.points + cells + 1);
    femm_age_point_t *rec = points->data() + first;
    for (int64_t k = 0; k < cells; ++k) {
      int64_t from = (k - whole) % cells;
      if (from < 0)
        from += cells;
      for (int q = 0; q < 2; ++q) {
        rec[k].node[q] = src.points[from].node[q];
        rec[k].weight[q] = src.points[from].weight[q];
      }
    }
    for (int q = 0; q < 2; ++q) {
      rec[cells].node[q] = rec[0].node[q];
      rec[cells].weight[q] = rec[0].weight[q];
    }
  }
  size_t at = 0;
  for (femm_air_gap_t &age : *out) {
    age.points = points->data() + at;
    at += static_cast<size_t>(age.cells) + 1;
  }
  return true;
}

}  // namespace host
}  // namespace femm
