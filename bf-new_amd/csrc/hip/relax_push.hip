// relax_push.hip -- the relaxation kernel variants for batches of push queries only (dense chaotic and frontier push).
#include "relax_kernels.hpp"

#include "bfn_launch.hpp"

namespace bfn {

hipError_t relax_launch_push(std::uint32_t block, bool persistent, bool cooperative,
                        std::uint32_t grid, std::size_t lds_bytes, std::uint32_t max_v,
                        std::uint32_t work_words, const DevGraph& g, const DevBatch& b,
                        hipStream_t stream) {
  return relax_launch_engines<kEnginesPush>(block, persistent, cooperative, grid, lds_bytes, max_v,
                                   work_words, g, b, stream);
}

hipError_t relax_occupancy_push(std::uint32_t block, bool persistent, bool instrumented,
                           std::uint32_t min_waves, std::size_t lds_bytes, int* blocks_per_cu) {
  return relax_occupancy_engines<kEnginesPush>(block, persistent, instrumented, min_waves, lds_bytes,
                                      blocks_per_cu);
}

}  // namespace bfn
