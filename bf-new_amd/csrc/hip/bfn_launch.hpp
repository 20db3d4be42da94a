// bfn_launch.hpp -- host-callable launch wrappers of the engine kernels.
#pragma once

#include "bfn_internal.hpp"
#include "bfnew_amd.h"

#include <string>

namespace bfn {

// Records the calling thread's last error (bfn_last_error) and returns code.
int32_t set_error(int32_t code, const std::string& message);
// Accessors of the handles defined in capi.hip, for the server.
int device_ordinal(bfn_device device);
std::size_t device_lds_limit(bfn_device device);
std::uint32_t device_cu_count(bfn_device device);
// Copies the graph's device view and the largest in/out neighbour-tile list.
void graph_view(bfn_graph graph, DevGraph* view, std::uint32_t* max_in, std::uint32_t* max_out);
// What a batch entry point outside capi.hip needs of a resident graph: its
// device view, host copies of the tile and neighbour-list offsets ([T+1]
// each), its stream and its device.
struct GraphHost {
  DevGraph view;
  const std::uint32_t* tile_offsets;
  const std::uint32_t* in_offsets;
  const std::uint32_t* out_offsets;
  hipStream_t stream;
  int ordinal;
  std::uint32_t cu_count;
  std::size_t lds_limit;
};
GraphHost graph_host(bfn_graph graph);
// Checks of bfn_batch_submit's arguments, shared by every batch entry point.
int32_t check_run_options(const bfn_run_options* options);
int32_t check_batch(bfn_graph graph, const bfn_batch_desc* batch, const bfn_run_options* options);

struct PathArgs;

hipError_t relax_launch(std::uint32_t block, bool persistent, bool cooperative,
                        std::uint32_t grid, std::size_t lds_bytes, std::uint32_t max_v,
                        std::uint32_t work_words, const DevGraph& g, const DevBatch& b,
                        hipStream_t stream);
hipError_t relax_occupancy(std::uint32_t engines, std::uint32_t block, bool persistent,
                           bool instrumented, std::uint32_t min_waves, std::size_t lds_bytes,
                           int* blocks_per_cu);
// The kernel variants of one engine set (relax_mixed.hip, relax_push.hip,
// relax_pull.hip); relax_launch and relax_occupancy pick by engine set.
#define BFN_RELAX_ENGINE_SET(name)                                                             \
  hipError_t relax_launch_##name(std::uint32_t block, bool persistent, bool cooperative,      \
                                 std::uint32_t grid, std::size_t lds_bytes, std::uint32_t max_v, \
                                 std::uint32_t work_words, const DevGraph& g, const DevBatch& b, \
                                 hipStream_t stream);                                          \
  hipError_t relax_occupancy_##name(std::uint32_t block, bool persistent, bool instrumented,  \
                                    std::uint32_t min_waves, std::size_t lds_bytes,            \
                                    int* blocks_per_cu);
BFN_RELAX_ENGINE_SET(mixed)
BFN_RELAX_ENGINE_SET(push)
BFN_RELAX_ENGINE_SET(pull)
#undef BFN_RELAX_ENGINE_SET
// Resets the control block, fills the labels and seeds round 0; with
// `structure` also resolves the slots' neighbour slots and descriptors (the
// first run of a batch; its reruns keep them).
hipError_t prep_launch(const DevGraph& g, const DevBatch& b, std::uint64_t pool_words,
                       hipStream_t stream, bool structure = true);
// Workgroups per CU (at most `promised`) whose persistent grid the device
// actually places at once, checked by a probe launch of the variant that `b`
// selects (b.engines, b.min_waves, b.instrumentation; b.control is reset and
// used); 0 when not even one per CU passes.
hipError_t coresident_per_cu(int device, int promised, std::uint32_t cus, std::size_t lds_bytes,
                             std::uint32_t max_v, std::uint32_t work_words, const DevGraph& g,
                             const DevBatch& b, hipStream_t stream, std::uint32_t block,
                             int* placed);
// Inputs and outputs of the tile-image derivation (graph.hip).
struct TileImageArgs {
  std::uint32_t tile_count;
  const std::uint32_t* tile_vertex_offsets;
  const std::uint32_t* owner_tiles;
  const std::uint32_t* row_offsets;
  const std::uint32_t* destinations;
  const std::uint32_t* row_runs;
  const std::uint32_t* run_dst_tiles;
  const float* weights;
  const std::uint32_t* column_offsets;
  const std::uint32_t* sources;
  const std::uint32_t* col_runs;
  const std::uint32_t* run_src_tiles;
  const float* in_weights;
  std::uint32_t* counts_scratch;
  std::uint32_t* in_offsets;
  std::uint32_t* out_offsets;
  std::uint32_t** in_tiles;   // allocated here
  std::uint32_t** in_stage;   // allocated here
  std::uint32_t** out_tiles;  // allocated here
  std::uint32_t** out_stage;  // allocated here
  std::uint32_t* csc_pack;
  std::uint32_t* csr_pack;
  uint2* csc_rec;
  uint2* csr_rec;
  std::uint32_t* csc_perm;
  std::uint32_t* csr_perm;
  std::uint32_t* csc_intra;
  std::uint32_t* csr_intra;
  std::uint32_t* csr_bnd;     // [V] first boundary record of each vertex
  std::uint32_t* stats;
  std::uint32_t* host_in_total;
  std::uint32_t* host_out_total;
};
hipError_t derive_tile_image(const TileImageArgs& args, hipStream_t stream);
// LDS words of the work area of a relaxation workgroup: Jacobi stages every
// in-neighbour slab; push engines keep the snapshot and, in direct mode, the
// boundary record starts and the list of pushing vertices; proposals mode
// adds the out-neighbour proposal slabs after the snapshot.
inline std::uint32_t relax_work_words(const DevGraph& g, bool any_jacobi, bool any_push,
                                      bool proposals) {
  std::uint32_t w = g.max_tile_vertices;
  if (any_jacobi) w = w > g.max_stage_floats ? w : g.max_stage_floats;
  if (any_push) {
    const std::uint32_t push = proposals ? g.max_tile_vertices + g.max_out_stage_floats
                                         : 3u * g.max_tile_vertices;
    w = w > push ? w : push;
  }
  return w;
}
// Re-reads the record weights from the orientation weight arrays.
hipError_t refresh_record_weights(const DevGraph& g, const std::uint32_t* csr_perm,
                                  const std::uint32_t* csc_perm, hipStream_t stream);
hipError_t paths_launch(const PathArgs& a, std::uint32_t query_count, hipStream_t stream);
hipError_t scan_u64_launch(const unsigned long long* in, unsigned long long* out, std::uint32_t n,
                           hipStream_t stream);
hipError_t predecessors_launch(const DevGraph& g, const DevBatch& b, std::uint32_t q,
                               std::uint32_t slots, std::uint32_t copy, std::uint32_t* out,
                               hipStream_t stream);
hipError_t scatter_launch(const DevGraph& g, const DevBatch& b, std::uint32_t q,
                          std::uint32_t slots, std::uint32_t copy, std::uint32_t* out,
                          hipStream_t stream);
hipError_t gather_region_launch(const DevGraph& g, const DevBatch& b, std::uint32_t q,
                                std::uint32_t slots, std::uint32_t copy,
                                const unsigned long long* offsets, std::uint32_t* out,
                                hipStream_t stream);
hipError_t finalize_launch(const DevGraph& g, const DevBatch& b,
                           const unsigned long long* relaxations, void* out, hipStream_t stream);
hipError_t relaxation_count_launch(const DevGraph& g, const DevBatch& b,
                                   unsigned long long* out, hipStream_t stream);
hipError_t occupancy_launch(const std::uint32_t* arena, unsigned long long n, int* occ,
                            hipStream_t stream);
hipError_t node_cost_launch(const DevGraph& g, const float* base, const int* occ, float* csr_w,
                            float* csc_w, hipStream_t stream);
hipError_t fill_launch(std::uint32_t* p, unsigned long long n, std::uint32_t value,
                       hipStream_t stream);

}  // namespace bfn
