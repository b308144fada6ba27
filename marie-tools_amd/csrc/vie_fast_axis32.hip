// Compile-time-length axis passes in FP32, all lengths and tile widths.
#include "vie_fast_kernels.hpp"

namespace vie {

bool launch_axis_fast_fp32(int in_kind, int out_kind, bool inv, int P, int TX, int sched, const AxisArgs& a, dim3 grid,
                           size_t lds, hipStream_t s, const void* tw) {
  return axis_typed<float>(in_kind, out_kind, inv, P, TX, sched, a, grid, lds, s, tw);
}

}  // namespace vie
