// Shared-operator tangent solves (femm_tangent_multi_rhs): dA/dI along circuit-current
// directions at a converged state, one PCG right-hand side per direction on one Newton
// operator.  Compiled in a translation unit of its own, so the solve kernels' inlining is
// unaffected.

#include "solve_kernels.hpp"

namespace femm
{
namespace
{

// Block d solves J dA_d = sum_c dir[d][c] s_c, where J is the Newton operator K + C that
// AssembleKernel left in item 0's values and s_c is circuit c's per-ampere source load
// reduced to the free rows (SolveTangentMultiRhs, solvers/nonlinear_solver.cu:85-172, with
// the reference's Jacobi PCG from x0 = 0).  The derivative is expanded with the DOF signs
// and zero on prescribed nodes (ExpandAssemblyDerivative, physics/nonlinear_model.cpp:413-426).
__global__ __launch_bounds__(kBlockThreads) void TangentKernel(DevicePlan P, DeviceWork W,
                                                               const double *directions, double tol,
                                                               int maxit, double *da,
                                                               femm_item_result_t *results)
{
  __shared__ double lds[2 * kWaves * 4];
  BlockTeam R{BlockReducer{lds, 0}, WholeItem(P)};
  const int item = blockIdx.x;
  const ItemPointers op = ItemWork(P, W, 0, false);
  const ItemPointers w = ItemWork(P, W, item, false);
  const double *dir = directions + static_cast<size_t>(item) * P.C;
  for (int row = R.S.row0 + threadIdx.x; row < R.S.row1; row += kBlockThreads) {
    double s = 0.0;
    for (int c = 0; c < P.C; ++c)
      s += dir[c] * P.circuit_rhs[static_cast<size_t>(c) * P.n + row];
    w.rhs[row] = s;
  }
  R.Sync();
  const BlockPrec<double> point{nullptr, 1, VectorStride(P.n)};
  const LinearInfo li =
      Pcg64(P, op.vals, op.inv, point, w.rhs, w.x, w.r, w.p, w.ap, tol, maxit, false, R);
  R.Sync();
  double *d = da + static_cast<size_t>(item) * P.N;
  for (int v = threadIdx.x; v < P.N; v += kBlockThreads) {
    const int f = P.node_to_free[v];
    d[v] = (f >= 0 && li.status == FEMM_OK) ? static_cast<double>(P.node_sign[v]) * w.x[f] : 0.0;
  }
  if (threadIdx.x == 0) {
    femm_item_result_t r{};
    r.status = li.status;
    r.newton_iterations = 0;
    r.linear_iterations = li.iterations;
    r.fp64_iterations = li.iterations;
    r.refinements = 0;
    r.relative_step = 0.0;
    r.equation_residual = 0.0;
    r.last_linear_tolerance = tol;
    r.linear_residual = li.residual;
    results[item] = r;
  }
}

}  // namespace

int LaunchTangent(const DeviceBuffers &dev, int32_t directions, const double *direction_currents,
                  double tol, int32_t maxit, double *da, femm_item_result_t *results,
                  hipStream_t stream)
{
  if (directions <= 0 || directions > dev.work.capacity)
    return FEMM_INVALID_ARGUMENT;
  hipLaunchKernelGGL(TangentKernel, dim3(directions), dim3(kBlockThreads), 0, stream, dev.plan,
                     dev.work, direction_currents, tol, maxit, da, results);
  return hipGetLastError() == hipSuccess ? FEMM_OK : FEMM_INTERNAL_ERROR;
}

}  // namespace femm
