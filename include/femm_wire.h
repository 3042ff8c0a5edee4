/*
 * femm_gpu_amd wire contract (C ABI): the JSON request / response protocol of the solve path
 * and the mesh-artifact parser, over the solver entry points of femm_gpu_amd.h.
 *
 * Plain pointers and sizes; status codes of femm_model.h; nothing throws.  The command
 * line (femm_cli_main) is what the reference's solver executable exposes; the artifact
 * functions let a host embed the parser without the command line.
 */
#ifndef FEMM_WIRE_H
#define FEMM_WIRE_H

#include <stddef.h>
#include <stdint.h>

#include "femm_model.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The solver command line.  Replaces the reference's executable entry (RunCli,
 * gpu_solver/src/app/cli.cpp:20-181) for this path:
 *   --solve REQUEST RESPONSE              gpu_femm_planar_dc_sample_v1 /
 *                                         gpu_femm_magnetostatic_sample_v1
 *                                         (PlanarDcSingleSampleAdapter, app/sample_adapter.cpp:330-489)
 *   --motor-single-sample REQUEST RESPONSE gpu_femm_motor_sample_v1 (MotorSingleSampleAdapter,
 *                                         app/sample_adapter.cpp:135-179)
 *   --motor-batch REQUEST RESPONSE        gpu_femm_motor_batch_v1 (MotorBatchAdapter,
 *                                         app/motor_batch.cpp:32-491)
 *   --mesh-artifact ARTIFACT              parse and summarise an artifact
 *   --capabilities                        the protocols and schemas of this build
 * Returns the process exit status: 0 when every response says PASS, 1 when a response says
 * FAIL or none could be published, 2 for an unknown command line.  argv[0] names the
 * executable (a response may not overwrite it). */
int femm_cli_main(int argc, const char *const *argv);

/* Parsed mesh artifact (GpuFemmMeshArtifact, io/mesh_artifact.h; parser
 * ParseGpuFemmMeshArtifactJson, io/mesh_artifact.cpp:14-440): schemas
 * gpu_femm_planar_dc_mesh_v1, gpu_femm_magnetostatic_mesh_v1 and gpu_femm_mesh_v1.
 * On failure returns FEMM_INVALID_ARGUMENT, no handle, and the reason in error
 * (NUL-terminated, truncated to error_capacity). */
typedef struct femm_artifact femm_artifact;
int femm_artifact_parse(const char *text, size_t length, femm_artifact **out, char *error,
                        size_t error_capacity);
/* The artifact's model (SI units, per-region material labels); the arrays belong to the
 * artifact and stay valid until femm_artifact_destroy. */
int femm_artifact_model(const femm_artifact *artifact, femm_model_t *model);
/* currents[model.circuit_count]: the circuits' current_A. */
int femm_artifact_currents(const femm_artifact *artifact, double *currents);
void femm_artifact_destroy(femm_artifact *artifact);

/* The rotor angle of a sliding-band model (ApplySlidingBandAngle,
 * gpu_solver/src/io/sample_request.cpp:448-491): the air-gap elements of a model posed at
 * posed_deg, turned to angle_deg.  Each element's inner shift advances by
 * (angle_deg - posed_deg) x cells / arc_length_deg; its whole cells rotate the inner-side
 * (node 0, node 1) pairs of the first `cells` records cyclically and the last record's inner
 * side repeats the first's; outer sides stay.  out[count] gets the elements, out_points (at
 * least the sum of cells + 1, point_capacity) their records.  FEMM_INVALID_ARGUMENT for a
 * non-finite angle difference, an element without cells or arc, or too small a buffer. */
int femm_air_gaps_rotate(int32_t count, const femm_air_gap_t *gaps, double posed_deg,
                         double angle_deg, femm_air_gap_t *out, femm_age_point_t *out_points,
                         size_t point_capacity);

/* SHA-256 of length bytes as 64 lower-case hex digits and a NUL (Sha256Hex,
 * io/strict_json.cpp:284-400): artifact identities bind the exact bytes on disk. */
int femm_sha256_hex(const void *bytes, size_t length, char hex[65]);

#ifdef __cplusplus
}
#endif

#endif
