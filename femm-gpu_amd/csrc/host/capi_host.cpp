// C ABI of the wire contract's host side: artifact parsing, SHA-256 and the command line.

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>

#include "api_guard.hpp"
#include "artifact.hpp"
#include "femm_wire.h"
#include "json.hpp"
#include "sliding_band.hpp"

struct femm_artifact {
  femm::host::Artifact artifact;
};

namespace
{

void CopyText(const std::string &text, char *out, size_t capacity)
{
  if (out == nullptr || capacity == 0)
    return;
  const size_t n = std::min(text.size(), capacity - 1);
  std::memcpy(out, text.data(), n);
  out[n] = '\0';
}

}  // namespace

extern "C" {

int femm_air_gaps_rotate(int32_t count, const femm_air_gap_t *gaps, double posed_deg,
                         double angle_deg, femm_air_gap_t *out, femm_age_point_t *out_points,
                         size_t point_capacity)
{
  FEMM_API_BEGIN
  if (count < 0 || (count > 0 && (gaps == nullptr || out == nullptr)))
    return FEMM_INVALID_ARGUMENT;
  std::vector<femm_air_gap_t> rotated;
  std::vector<femm_age_point_t> points;
  if (!femm::host::RotateAirGaps(gaps, count, posed_deg, angle_deg, &rotated, &points))
    return FEMM_INVALID_ARGUMENT;
  if (points.size() > point_capacity || (!points.empty() && out_points == nullptr))
    return FEMM_INVALID_ARGUMENT;
  std::copy(points.begin(), points.end(), out_points);
  size_t at = 0;
  for (int32_t g = 0; g < count; ++g) {
    out[g] = rotated[g];
    out[g].points = out_points + at;
    at += static_cast<size_t>(rotated[g].cells) + 1;
  }
  return FEMM_OK;
  FEMM_API_END
}

int femm_artifact_parse(const char *text, size_t length, femm_artifact **out, char *error,
                        size_t error_capacity)
{
  FEMM_API_BEGIN
  if (out == nullptr || (text == nullptr && length > 0))
    return FEMM_INVALID_ARGUMENT;
  *out = nullptr;
  std::unique_ptr<femm_artifact> a(new femm_artifact);
  std::string message;
  if (!femm::host::ParseArtifact(std::string(text ? text : "", length), &a->artifact, &message)) {
    CopyText(message, error, error_capacity);
    return FEMM_INVALID_ARGUMENT;
  }
  CopyText("", error, error_capacity);
  *out = a.release();
  return FEMM_OK;
  FEMM_API_END
}

int femm_artifact_model(const femm_artifact *artifact, femm_model_t *model)
{
  FEMM_API_BEGIN
  if (artifact == nullptr || model == nullptr)
    return FEMM_INVALID_ARGUMENT;
  *model = artifact->artifact.View();
  return FEMM_OK;
  FEMM_API_END
}

int femm_artifact_currents(const femm_artifact *artifact, double *currents)
{
  FEMM_API_BEGIN
  if (artifact == nullptr || (currents == nullptr && !artifact->artifact.currents.empty()))
    return FEMM_INVALID_ARGUMENT;
  std::copy(artifact->artifact.currents.begin(), artifact->artifact.currents.end(), currents);
  return FEMM_OK;
  FEMM_API_END
}

void femm_artifact_destroy(femm_artifact *artifact)
{
  FEMM_API_BEGIN
  delete artifact;
  FEMM_API_END_VOID
}

int femm_sha256_hex(const void *bytes, size_t length, char hex[65])
{
  FEMM_API_BEGIN
  if (hex == nullptr || (bytes == nullptr && length > 0))
    return FEMM_INVALID_ARGUMENT;
  const std::string digest =
      femm::host::Sha256Hex(std::string(static_cast<const char *>(bytes ? bytes : ""), length));
  CopyText(digest, hex, 65);
  return FEMM_OK;
  FEMM_API_END
}

}  // extern "C"
