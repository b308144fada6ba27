// Compile-time-length axis passes in FP64, all lengths and tile widths, and
// the queries for the compiled lengths.
#include "vie_fast_kernels.hpp"

namespace vie {

bool launch_axis_fast_fp64(int in_kind, int out_kind, bool inv, int P, int TX, int sched, const AxisArgs& a, dim3 grid,
                           size_t lds, hipStream_t s, const void* tw) {
  return axis_typed<double>(in_kind, out_kind, inv, P, TX, sched, a, grid, lds, s, tw);
}

bool fast_axis_length(int P) { return dispatch_len(P, [](auto) {}, AxisLens{}); }
bool fast_x_length(int P) { return dispatch_len(P, [](auto) {}, XLens{}); }

}  // namespace vie
