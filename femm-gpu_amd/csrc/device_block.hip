// Block-Jacobi solve forms (kFormBlock, FP64 and mixed precision) of the templates in
// solve_kernels.hpp, compiled apart from the point-Jacobi forms of device.hip.

#include "solve_kernels.hpp"

namespace femm
{
namespace
{

// PCG of the current values with the block-Jacobi preconditioner of `block_size` rows
// (DeterministicPcgKernel<..., kBlock = true>, solvers/pcg_kernels.cu:10-344, after
// BlockJacobiPlan::Prepare, block_jacobi_preconditioner.cu:188-258).
__global__ __launch_bounds__(kBlockThreads) void BlockPcgKernel(DevicePlan P, DeviceWork W,
                                                                double tol, int maxit,
                                                                int block_size, double *info)
{
  __shared__ double lds[2 * kWaves * 4];
  BlockTeam R{BlockReducer{lds, 0}, WholeItem(P)};
  const int item = blockIdx.x;
  const ItemPointers w = ItemWork(P, W, item, false);
  for (int row = R.S.row0 + threadIdx.x; row < R.S.row1; row += kBlockThreads) {
    const double d = w.vals[P.diag_sell[row]];
    w.inv[row] = (isfinite(d) && d > 0.0) ? 1.0 / d : 0.0;
  }
  R.Sync();
  BuildBlockJacobi(P.n, P.slice_ptr, P.slice_width, P.sell_cols, P.diag_sell, R.S.row0, R.S.row1,
                   w.vals, w.inv, block_size, w.bjac, nullptr);
  R.Sync();
  const BlockPrec<double> B{w.bjac, block_size, VectorStride(P.n)};
  const LinearInfo li =
      Pcg64<true, 1>(P, w.vals, w.inv, B, w.rhs, w.x, w.r, w.p, w.ap, tol, maxit, false, R);
  if (threadIdx.x == 0) {
    info[3 * item] = li.status;
    info[3 * item + 1] = li.iterations;
    info[3 * item + 2] = li.residual;
  }
}

}  // namespace

int LaunchBlockSolve(int form, int mode, const DeviceBuffers &dev, int32_t items,
                     int32_t partitions, const double *currents, const NewtonArgs &args,
                     const double *warm, double *a, double *flux, femm_item_result_t *results,
                     hipStream_t stream, const LaunchExtras &x)
{
  if (form & kFormMixed)
    return LaunchFormSolveMode<kFormMixed | kFormBlock>(mode, dev, items, partitions, currents,
                                                        args, warm, a, flux, results, stream, x);
  return LaunchFormSolveMode<kFormBlock>(mode, dev, items, partitions, currents, args, warm, a,
                                         flux, results, stream, x);
}

const void *BlockGroupKernel(int form, bool recovery)
{
  return (form & kFormMixed) ? FormGroupKernel<kFormMixed | kFormBlock>(recovery)
                             : FormGroupKernel<kFormBlock>(recovery);
}

int LaunchBlockPcg(const DeviceBuffers &dev, int32_t items, double tol, int32_t maxit,
                   int32_t block_size, double *info, hipStream_t stream)
{
  hipLaunchKernelGGL(BlockPcgKernel, dim3(items), dim3(kBlockThreads), 0, stream, dev.plan,
                     dev.work, tol, maxit, block_size, info);
  return hipGetLastError() == hipSuccess ? FEMM_OK : FEMM_INTERNAL_ERROR;
}

}  // namespace femm
