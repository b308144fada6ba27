// Fused x pass, float instantiations (a translation unit of its own so the
// specialised kernels compile in parallel).
#include "vie_fast_kernels.hpp"

namespace vie {

bool launch_x_fast_fp32(bool pwl, bool isk, bool adj, int P, const void* args, int64_t blocks, size_t lds,
                        hipStream_t s, const void* tw) {
  return x_typed<float>(pwl, isk, adj, P, args, blocks, lds, s, tw);
}

}  // namespace vie
