// relax.hip -- batch preparation kernels and the dispatch of the relaxation
// kernels (relax_kernels.hpp) to their per-block-size translation units.
#include "bfn_internal.hpp"
#include "bfn_launch.hpp"

namespace bfn {

// ---------------------------------------------------------------------------
// Batch preparation.
// ---------------------------------------------------------------------------
__global__ void reset_control_kernel(DevControl* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    for (int i = 0; i < 3; ++i) {
      c->list_count[i] = 0;
      c->claim[i] = 0;
    }
    c->round = 0;
    c->done = 0;
    c->stop_reason = kStopNone;
    c->error_bits = 0;
    c->barrier_count = 0;
    c->barrier_generation = 0;
    c->rounds_executed = 0;
    c->items_processed = 0;
    c->edges_examined = 0;
    c->atomic_attempts = 0;
    c->atomic_successes = 0;
    c->max_list = 0;
    c->edges_relaxed = 0;
    c->algorithmic_bytes = 0;
    c->closure_sweeps = 0;
    for (int p = 0; p < kPhaseCount; ++p) c->phase_cycles[p] = 0;
  }
}

__global__ void fill_u32_kernel(std::uint32_t* p, unsigned long long n, std::uint32_t v) {
  const unsigned long long stride = static_cast<unsigned long long>(gridDim.x) * blockDim.x;
  for (unsigned long long i = static_cast<unsigned long long>(blockIdx.x) * blockDim.x +
                              threadIdx.x;
       i < n; i += stride)
    p[i] = v;
}

// Resolves each slot's in/out neighbour tiles to the query's slots.
__global__ void prep_slots_kernel(DevGraph g, DevBatch b) {
  const std::uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= b.slot_count) return;
  const std::uint32_t q = b.slot_query[s];
  const std::uint32_t T = b.tiles[s];
  const std::uint32_t in0 = g.in_nbr_offsets[T], in1 = g.in_nbr_offsets[T + 1];
  const std::uint32_t ib = b.slot_in_base[s];
  for (std::uint32_t k = in0; k < in1; ++k) b.slot_in[ib + (k - in0)] = find_slot(b, q, g.in_nbr_tiles[k]);
  const std::uint32_t o0 = g.out_nbr_offsets[T], o1 = g.out_nbr_offsets[T + 1];
  const std::uint32_t obs = b.slot_out_base[s];
  for (std::uint32_t k = o0; k < o1; ++k) b.slot_out[obs + (k - o0)] = find_slot(b, q, g.out_nbr_tiles[k]);
  fill_slot_desc(g, b, s);
  b.stamp[s] = 0;
  b.last_change[q] = 0;
}

// Zero the sources and queue round 0: Jacobi queues the source tiles and
// every tile they feed; push engines queue the source tiles.
__global__ void seed_kernel(DevGraph g, DevBatch b) {
  const std::uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= b.query_count) return;
  const bool jacobi = b.engine[q] == kJacobiPull;
  for (std::uint32_t i = b.source_offsets[q]; i < b.source_offsets[q + 1]; ++i) {
    const std::uint32_t u = b.sources[i];
    const std::uint32_t T = g.owner_tiles[u];
    const std::uint32_t s = find_slot(b, q, T);
    if (s == kNone) {
      atomicOr(&b.control->error_bits, 4u);  // device_failure: host validated this
      continue;
    }
    const std::uint32_t v0 = g.tile_vertex_offsets[T];
    const std::uint32_t npt = g.tile_vertex_offsets[T + 1] - v0;
    std::uint32_t* column = b.dist + b.slot_dist[s];
    column[u - v0] = 0u;
    if (jacobi) column[npt + (u - v0)] = 0u;
    if (jacobi) enqueue(b, s, 0); else enqueue_pushed(b, s, 0, 0u);
    if (jacobi) {
      const std::uint32_t ob = b.slot_out_base[s];
      const std::uint32_t KO = g.out_nbr_offsets[T + 1] - g.out_nbr_offsets[T];
      for (std::uint32_t k = 0; k < KO; ++k) {
        const std::uint32_t ns = b.slot_out[ob + k];
        if (ns != kNone) enqueue(b, ns, 0);
      }
    }
  }
}

hipError_t relax_launch(std::uint32_t block, bool persistent, bool cooperative,
                        std::uint32_t grid, std::size_t lds_bytes, std::uint32_t max_v,
                        std::uint32_t work_words, const DevGraph& g, const DevBatch& b,
                        hipStream_t stream) {
  switch (block) {
#define BFN_CASE(B) \
  case B: return relax_launch_b##B(persistent, cooperative, grid, lds_bytes, max_v, work_words, g, b, stream);
    BFN_CASE(64) BFN_CASE(128) BFN_CASE(256) BFN_CASE(512) BFN_CASE(1024)
#undef BFN_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t relax_occupancy(std::uint32_t block, bool persistent, bool instrumented,
                           std::uint32_t min_waves, std::uint32_t kind, std::size_t lds_bytes,
                           int* blocks_per_cu) {
  switch (block) {
#define BFN_CASE(B) \
  case B: return relax_occupancy_b##B(persistent, instrumented, min_waves, kind, lds_bytes, blocks_per_cu);
    BFN_CASE(64) BFN_CASE(128) BFN_CASE(256) BFN_CASE(512) BFN_CASE(1024)
#undef BFN_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t prep_launch(const DevGraph& g, const DevBatch& b, std::uint64_t pool_words,
                       hipStream_t stream) {
  hipLaunchKernelGGL(reset_control_kernel, dim3(1), dim3(64), 0, stream, b.control);
  const std::uint32_t fill_grid = 4096;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(fill_grid), dim3(256), 0, stream, b.dist,
                     static_cast<unsigned long long>(pool_words), kInfBits);
  if (b.slot_count)
    hipLaunchKernelGGL(prep_slots_kernel, dim3((b.slot_count + 255) / 256), dim3(256), 0,
                       stream, g, b);
  if (b.slot_min)
    hipLaunchKernelGGL(fill_u32_kernel, dim3(256), dim3(256), 0, stream, b.slot_min,
                       2ull * b.slot_count + 3ull * b.query_count, kInfBits);
  if (b.query_count)
    hipLaunchKernelGGL(seed_kernel, dim3((b.query_count + 127) / 128), dim3(128), 0, stream, g,
                       b);
  return hipGetLastError();
}

}  // namespace bfn

namespace bfn {
hipError_t fill_launch(std::uint32_t* p, unsigned long long n, std::uint32_t value,
                       hipStream_t stream) {
  if (n) hipLaunchKernelGGL(fill_u32_kernel, dim3(2048), dim3(256), 0, stream, p, n, value);
  return hipGetLastError();
}
}  // namespace bfn
