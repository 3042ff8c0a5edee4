// FEMM's native air-gap element (AGE) on the host: validation, the 10-node stencil of a cell,
// and the cell's element matrix (reference gpu_solver/src/native_airgap_matrix.cpp:5-22, whose
// matrix text comes from fkn/prob1big.cpp:158-213, and physics/nonlinear_model.cpp:569-622).
//
// The element, as derived here from what the reference computes with it: a cell is a
// rectangle in (arc length s, radius r) of width w = R dt (R the mean radius, dt the cell
// angle) and height h = ro - ri, with local coordinates u = s / w, v = (r - ri) / h.  The
// potential is bilinear between four virtual corners; each corner is the Catmull-Rom
// interpolant of the ring nodes around it, so a ring whose nodes sit a fraction c of a cell
// away from the corners (its shift) couples through five nodes: j = 0..4 at u = j - 2 + c.
// The stiffness of the rectangle is K A (x) M + K^-1 M (x) A, with K = h / w, A the 1D
// stiffness [[1, -1], [-1, 1]] and M the mean of the consistent and the lumped 1D mass
// matrices, [[5, 1], [1, 5]] / 12 (the bilinear element with its twist mode stiffened
// twice).  The 10 x 10 matrix is T^T S T with T the 4 x 10 corner interpolation.
#pragma once

#include <cstdint>

#include "femm_model.h"

namespace femm
{

// Catmull-Rom weights of the ring nodes at -1, 0, 1, 2 for the point t in [0, 1] between the
// middle two.
void CatmullRomWeights(double t, double w[4]);

// The 10 x 10 cell matrix for aspect K (h / w), its inverse Ki, and ring shifts ci (inner)
// and co (outer).  False for non-positive or non-finite inputs (BuildNativeFemmAirGapMatrix,
// native_airgap_matrix.cpp:5-11).
bool AgeCellMatrix(double K, double Ki, double ci, double co, double matrix[10][10]);

// K, Ki and the shifts the assembly uses for an element: the cell angle dt, K = 2 (ro - ri) /
// (dt (ro + ri)), and the shifts made relative, one of them 0 or 1
// (nonlinear_model.cpp:573-587).
void AgeAssemblyParameters(const femm_air_gap_t &age, double *K, double *Ki, double *ci,
                           double *co);

// The ten nodes and weights of cell k (BuildAirGapStencil, postprocess/airgap_postprocess.cpp:
// 10-38): inner ring nodes of the previous, this, next and second-next boundary records,
// then the outer ring's; an anti-periodic element flips the wrapped first and last nodes.
bool AgeStencil(const femm_air_gap_t &age, int32_t k, int32_t node[10], double weight[10]);

// The AGE rules of ValidateNonlinearModel (nonlinear_model.cpp:281-297): planar model,
// >= 2 cells, 0 < ri < ro, arc > 0, finite centre and shifts, node indices in range, finite
// weights.  Returns a femm status code.
int ValidateAirGaps(const femm_model_t *model);

// SI scale of the AGE matrix (fkn works in relative permeability): 1 / mu0.
constexpr double kAgeReluctivity = 1.0 / (4.0e-7 * 3.141592653589793238462643383279502884);

}  // namespace femm
