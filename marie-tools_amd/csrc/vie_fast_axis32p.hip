// Compile-time-length FP32 axis passes that read or write packed vectors in
// FP32 (the first and last pass of the mixed-precision solver's operator),
// all lengths and tile widths; a unit of its own so it compiles in parallel
// with the other FP32 passes.
#include "vie_fast_kernels.hpp"

namespace vie {

bool launch_axis_fast_fp32_p32(int in_kind, int out_kind, bool inv, int P, int TX, int sched, const AxisArgs& a,
                               dim3 grid, size_t lds, hipStream_t s, const void* tw) {
  return axis_typed_p32<float>(in_kind, out_kind, inv, P, TX, sched, a, grid, lds, s, tw);
}

}  // namespace vie
