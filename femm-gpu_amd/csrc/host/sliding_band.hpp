// Rotor angle of a sliding-band model: the air-gap elements' inner ring records turned by
// whole cells and the fractional inner shift (reference gpu_solver/src/io/sample_request.cpp:
// 448-491, ApplySlidingBandAngle).
#pragma once

#include <vector>

#include "femm_model.h"

namespace femm
{
namespace host
{

// The air-gap elements of a model posed at `posed_deg` turned to `angle_deg`: for each
// element the inner shift advances by (angle - posed) x cells / arc, the whole cells of it
// rotate the inner-side (node 0, node 1) pairs of the first `cells` records cyclically, and
// the last record's inner side repeats the first's; outer sides are kept.  `points` receives
// every element's records (out[g].points points into it).  False for a non-finite angle
// difference or an element without cells or arc.
bool RotateAirGaps(const femm_air_gap_t *in, int32_t count, double posed_deg, double angle_deg,
                   std::vector<femm_air_gap_t> *out, std::vector<femm_age_point_t> *points);

}  // namespace host
}  // namespace femm
