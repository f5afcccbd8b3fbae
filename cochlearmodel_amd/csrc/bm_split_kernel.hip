// BM/OHC velocity kernel with a time block spread over a workgroup of 2 or 4
// waves (2 or 1 sections per lane), a launch layout of the autotuner.  It
// runs the same device functions as bm_kernel (bm_device.h), exchanging the
// wave-boundary pressures and partial reductions through LDS; its results are
// identical to bm_kernel's.  A translation unit of its own, so that the
// one-wave kernels' code object is the same with or without it.
#include <hip/hip_runtime.h>

#include "bm_device.h"

namespace cm {


// One block per workgroup of 4 / N waves (N = 2 or 1 sections per lane).
template <int N, bool TOL, int WS, int U>
__global__ __launch_bounds__(256 / N) void bm_split_kernel(BmLaunch a) {
  constexpr int W = 4 / N;
  __shared__ float xa[2 * 4 * W];
  __shared__ float4 pold_lds[WS ? W * 65 : 1];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int idx = blockIdx.x;   // the whole workgroup takes one block
  Lay<N> L{lane, wave, xa, 0};
  if (idx < a.n_items)
    run_item<N, TOL, WS, U>(a, __builtin_amdgcn_readfirstlane(a.order ? a.order[idx] : idx), L,
                            pold_lds + (WS ? wave * 65 : 0));
}

}  // namespace cm

// Launcher used by cm_launch_bm for 2 or 1 sections per lane: one workgroup
// of 4 / N waves per block; Jacobi unroll 1, 2 or 4.
extern "C" int cm_launch_bm_split(const cm::BmLaunch *a, int sections_per_lane, int unroll, hipStream_t stream) {
  const bool tol = a->sc.jacobi_mode == 1;
  const int ws = tol ? a->sc.warm_start : 0;
  if (ws < 0 || ws > 1) return -1;
  const int n = sections_per_lane;
  const int u = unroll == 1 ? 0 : (unroll == 2 ? 1 : (unroll == 4 ? 2 : -1));
  if ((n != 2 && n != 1) || u < 0) return -1;
  const int mode = tol ? 1 + ws : 0;   // 0 fixed sweeps, 1 tolerance, 2 tolerance + extrapolation
  const dim3 grid(a->n_items), block(256 / n);
#define CM_SPLIT_CASE(N, U)                                                                 \
  case ((N) == 2 ? 0 : 1) * 3 + ((U) == 1 ? 0 : ((U) == 2 ? 1 : 2)):                        \
    if (mode == 0)                                                                          \
      hipLaunchKernelGGL((cm::bm_split_kernel<N, false, 0, U>), grid, block, 0, stream, *a); \
    else if (mode == 1)                                                                     \
      hipLaunchKernelGGL((cm::bm_split_kernel<N, true, 0, U>), grid, block, 0, stream, *a);  \
    else                                                                                    \
      hipLaunchKernelGGL((cm::bm_split_kernel<N, true, 1, U>), grid, block, 0, stream, *a);  \
    break;
  switch ((n == 2 ? 0 : 1) * 3 + u) {
    CM_SPLIT_CASE(2, 1)
    CM_SPLIT_CASE(2, 2)
    CM_SPLIT_CASE(2, 4)
    CM_SPLIT_CASE(1, 1)
    CM_SPLIT_CASE(1, 2)
    CM_SPLIT_CASE(1, 4)
    default: return -1;
  }
#undef CM_SPLIT_CASE
  return hipGetLastError() == hipSuccess ? 0 : -2;
}
