// Launch entry points of the specialised kernels (the axis-pass
// instantiations live in vie_fast_axis64.hip / vie_fast_axis32.hip).
#include "vie_passes.hpp"

namespace vie {

bool launch_axis_fast(int in_kind, int out_kind, bool inv, bool fp64, int P, int TX, int sched, const AxisArgs& a,
                      dim3 grid, size_t lds, hipStream_t s, const void* tw) {
  if (in_packed32(in_kind) || out_packed32(out_kind)) {
    if (fp64) throw Error{VIE_ERR_ARG, "packed FP32 vectors need the FP32 operator"};
    return fast_axis_length(P) && launch_axis_fast_fp32_p32(in_kind, out_kind, inv, P, TX, sched, a, grid, lds, s, tw);
  }
  return fp64 ? launch_axis_fast_fp64(in_kind, out_kind, inv, P, TX, sched, a, grid, lds, s, tw)
              : launch_axis_fast_fp32(in_kind, out_kind, inv, P, TX, sched, a, grid, lds, s, tw);
}

bool launch_x_fast(bool fp64, bool pwl, bool isk, bool adj, int P, const void* args, int64_t blocks, size_t lds,
                   hipStream_t s, const void* tw) {
  return fp64 ? launch_x_fast_fp64(pwl, isk, adj, P, args, blocks, lds, s, tw)
              : launch_x_fast_fp32(pwl, isk, adj, P, args, blocks, lds, s, tw);
}

}  // namespace vie
