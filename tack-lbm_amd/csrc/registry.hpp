// Kernel registry shared by the instantiation units and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "lbm_step.hpp"

namespace tlbm {

constexpr int kObstacles = 5;  // TLBM_CYLINDER .. TLBM_HEATSINK

using StepLaunch = hipError_t (*)(const StepArgs&, dim3 grid, dim3 block, hipStream_t stream);
// Launchers of one (storage, collision) unit, indexed [obstacle][side][report].
using StepTable = StepLaunch[kObstacles][3][2];
// Launchers of the temperature variant of one collision (FP32), indexed as StepTable.
using HeatTable = StepLaunch[kObstacles][3][2];

// Quiet steps come in four builds: plain or non-temporal population stores (NT), chosen per
// launch by a.stream_stores (the host sets it for large lattices), each with per-lane or
// per-wave row addressing (ROWS, for rows that are whole waves: a.row_waves); the NT + ROWS
// build also streams the unshifted pulls with non-temporal loads.  Report steps use plain
// loads and stores.
template <int C, int O, int S, bool R, int ST, bool H = false>
hipError_t launch_step(const StepArgs& a, dim3 grid, dim3 block, hipStream_t stream) {
  if constexpr (!R) {
    if (a.stream_stores && a.row_waves)
      hipLaunchKernelGGL((lbm_step_kernel<C, O, S, R, ST, H, true, true>), grid, block, 0, stream, a);
    else if (a.stream_stores)
      hipLaunchKernelGGL((lbm_step_kernel<C, O, S, R, ST, H, true, false>), grid, block, 0, stream, a);
    else if (a.row_waves)
      hipLaunchKernelGGL((lbm_step_kernel<C, O, S, R, ST, H, false, true>), grid, block, 0, stream, a);
    else
      hipLaunchKernelGGL((lbm_step_kernel<C, O, S, R, ST, H, false, false>), grid, block, 0, stream, a);
    return hipGetLastError();
  }
  hipLaunchKernelGGL((lbm_step_kernel<C, O, S, R, ST, H, false, false>), grid, block, 0, stream, a);
  return hipGetLastError();
}

using BatchLaunch = hipError_t (*)(const BatchArgs&, bool nt, dim3 grid, dim3 block,
                                   hipStream_t stream);
using BatchTable = BatchLaunch[kObstacles][3][2];

// `nt`: the batch's population buffers are large enough for non-temporal stores (quiet steps).
template <int C, int O, int S, bool R, int ST>
hipError_t launch_batch(const BatchArgs& b, bool nt, dim3 grid, dim3 block, hipStream_t stream) {
  if constexpr (!R) {
    if (nt) {
      hipLaunchKernelGGL((lbm_batch_kernel<C, O, S, R, ST, true>), grid, block, 0, stream, b);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((lbm_batch_kernel<C, O, S, R, ST, false>), grid, block, 0, stream, b);
  return hipGetLastError();
}

template <int C, int ST>
void fill_table(StepTable& t, BatchTable& bt) {
  static_for<0, kObstacles>([&](auto o) {
    static_for<0, 3>([&](auto s) {
      constexpr int O = decltype(o)::value, S = decltype(s)::value;
      t[O][S][0] = &launch_step<C, O, S, false, ST>;
      t[O][S][1] = &launch_step<C, O, S, true, ST>;
      bt[O][S][0] = &launch_batch<C, O, S, false, ST>;
      bt[O][S][1] = &launch_batch<C, O, S, true, ST>;
    });
  });
}

template <int C>
void fill_heat_table(HeatTable& t) {
  static_for<0, kObstacles>([&](auto o) {
    static_for<0, 3>([&](auto s) {
      constexpr int O = decltype(o)::value, S = decltype(s)::value;
      t[O][S][0] = &launch_step<C, O, S, false, TLBM_F32, true>;
      t[O][S][1] = &launch_step<C, O, S, true, TLBM_F32, true>;
    });
  });
}

}  // namespace tlbm

// One per (storage, collision) unit; defined in step_inst.hip compiled with
// -DTLBM_STORAGE=<s> -DTLBM_COLL=<c>.
// (library-internal: hidden from the shared object's exports)
#define TLBM_DECLARE_UNIT(s, c)                       \
  __attribute__((visibility("hidden"))) void tlbm_register_##s##_##c(tlbm::StepTable& t, \
                                                                      tlbm::BatchTable& bt);
TLBM_DECLARE_UNIT(0, 0)
TLBM_DECLARE_UNIT(0, 1)
TLBM_DECLARE_UNIT(0, 2)
TLBM_DECLARE_UNIT(1, 0)
TLBM_DECLARE_UNIT(1, 1)
TLBM_DECLARE_UNIT(1, 2)
TLBM_DECLARE_UNIT(2, 0)
TLBM_DECLARE_UNIT(2, 1)
TLBM_DECLARE_UNIT(2, 2)

// One per collision; defined in heat_inst.hip compiled with -DTLBM_COLL=<c>.
#define TLBM_DECLARE_HEAT_UNIT(c) \
  __attribute__((visibility("hidden"))) void tlbm_register_heat_##c(tlbm::HeatTable& t);
TLBM_DECLARE_HEAT_UNIT(0)
TLBM_DECLARE_HEAT_UNIT(1)
TLBM_DECLARE_HEAT_UNIT(2)
