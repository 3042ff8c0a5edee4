/*
 * Flat, C-compatible description of a P1 magnetostatic model.
 *
 * This is the data contract shared by the MI355X solver library
 * (femm_gpu_amd.h) and by the CPU oracle used in tests.  It carries exactly the
 * fields of the reference's NonlinearModel (gpu_solver/src/physics/nonlinear_model.h:74-97)
 * in plain arrays: SI units (metres, A/m, T), counter-clockwise triangles,
 * one material per triangle, raw B-H knots (the spline is built by the callee),
 * Dirichlet nodes and signed periodic / anti-periodic node pairs.
 * No pointer is retained by a callee after it returns.
 */
#ifndef FEMM_MODEL_H
#define FEMM_MODEL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: identical integers and meaning to the reference's Status enum
 * (gpu_solver/src/core/types.h:11-28). */
enum {
  FEMM_OK = 0,
  FEMM_INPUT_IO = 1,
  FEMM_INVALID_ARGUMENT = 2,
  FEMM_UNSUPPORTED_FEATURE = 3,
  FEMM_MESH_INVALID = 4,
  FEMM_BOUNDARY_INVALID = 5,
  FEMM_GPU_UNAVAILABLE = 6,
  FEMM_GPU_ALLOCATION_FAILED = 7,
  FEMM_ASSEMBLY_FAILED = 8,
  FEMM_LINEAR_SOLVE_NOT_CONVERGED = 9,
  FEMM_LINEAR_SOLVE_BREAKDOWN = 10,
  FEMM_NUMERICAL_NONFINITE = 11,
  FEMM_OUTPUT_IO = 12,
  FEMM_INTERNAL_ERROR = 13,
  FEMM_NONLINEAR_SOLVE_NOT_CONVERGED = 14,
  FEMM_INVALID_MATERIAL = 15
};

enum { FEMM_PLANAR = 0, FEMM_AXISYMMETRIC = 1 };

/* One material label (reference NonlinearMaterial, nonlinear_model.h:38-52).
 * nu(B) = reluctivity_zero * (1 + alpha |B|^2) unless bh_curve >= 0. */
typedef struct femm_material {
  double reluctivity_zero; /* m/H */
  double alpha;            /* 1/T^2 */
  double source_j_per_a;   /* turns / region area, A/m^2 per A */
  double hc;               /* PM coercivity, A/m */
  double magnetization_deg;
  int32_t bh_curve;        /* index into the curve table or -1 */
  int32_t circuit;         /* series circuit index or -1 */
  int32_t group;           /* FEMM block-label group or -1 */
  int32_t reserved;
} femm_material_t;

/* A(node_b) = sign * A(node_a), sign = +1 periodic, -1 anti-periodic
 * (reference SignedNodeConstraint, signed_node_constraints.h:10-15). */
typedef struct femm_node_constraint {
  int32_t node_a;
  int32_t node_b;
  int32_t sign;
} femm_node_constraint_t;

/* One cell boundary of an air-gap element (reference AirGapQuadPoint,
 * nonlinear_model.h:63-69): (node[0], weight[0]) and (node[1], weight[1]) are the inner-ring
 * nodes at the start and end of the cell, (node[2], weight[2]) and (node[3], weight[3]) the
 * outer-ring ones; weights are +1 or -1 (an anti-periodic image). */
typedef struct femm_age_point {
  int32_t node[4];
  double weight[4];
} femm_age_point_t;

/* FEMM's native air-gap element (reference AirGapElement, nonlinear_model.h:71-81): the
 * annulus between an inner and an outer ring, meshed independently, cut into `cells` equal
 * angular cells; no triangles are built inside it.  inner_shift / outer_shift place each
 * ring's nodes within a cell (fraction of a cell), so the rotor can turn without remeshing.
 * Planar models only. */
typedef struct femm_air_gap {
  int32_t antiperiodic;            /* 1: the arc closes with a sign change */
  int32_t cells;                   /* equal angular cells, >= 2 */
  double center_xy[2];             /* m */
  double inner_radius_m;
  double outer_radius_m;
  double arc_length_deg;           /* angular extent of the element */
  double inner_shift;
  double outer_shift;
  const femm_age_point_t *points;  /* [cells + 1] */
} femm_air_gap_t;

typedef struct femm_model {
  int32_t problem_type;            /* FEMM_PLANAR or FEMM_AXISYMMETRIC (x = r, y = z) */
  int32_t node_count;
  const double *node_xy;           /* [node_count][2] */
  int32_t triangle_count;
  const int32_t *triangle_nodes;   /* [triangle_count][3], counter-clockwise */
  const int32_t *triangle_material;/* [triangle_count] */
  int32_t material_count;
  const femm_material_t *materials;
  int32_t curve_count;
  const int32_t *curve_offsets;    /* [curve_count + 1] into curve_b / curve_h */
  const double *curve_b;           /* raw knots, B[0] = H[0] = 0, B strictly increasing */
  const double *curve_h;
  int32_t dirichlet_count;
  const int32_t *dirichlet_nodes;
  const double *dirichlet_values;
  int32_t constraint_count;
  const femm_node_constraint_t *constraints;
  int32_t circuit_count;
  double depth_m;                  /* planar depth; ignored for r-z */
  int32_t air_gap_count;           /* native air-gap elements (may be 0) */
  const femm_air_gap_t *air_gaps;
} femm_model_t;

/* Newton policy (reference NonlinearOptions, nonlinear_model.h:99-106, and
 * ProductionSampleOptions, nonlinear_model.cpp:195-205). */
typedef struct femm_newton_options {
  double relative_tolerance;        /* step gate is 100x this; equation residual gate is 1x */
  double linear_relative_tolerance; /* final PCG ||b-Ax|| / ||b|| */
  int32_t max_newton_iterations;
  int32_t max_linear_iterations;
  int32_t recover_nonconvergence;   /* bounded Armijo recovery for exhausted items */
  int32_t adaptive_linear_tolerance;
  int32_t require_equation_residual; /* 1: accept an item only when its undamped step is
                                        below 100x relative_tolerance AND its equation residual
                                        is <= relative_tolerance (the two gates of the
                                        reference's SolveSafeguarded, nonlinear_solver.cu:623-627);
                                        0: the step gate alone (NewtonStepConverged) */
  int32_t reserved;
} femm_newton_options_t;

#ifdef __cplusplus
}
#endif

#endif
