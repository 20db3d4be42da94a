// capi.hip -- host side of the engine library: handles, memory, control loops.
//
// Control (GpuRunOptions::control_mode, gpu_api.hpp:20-24):
//   persistent_cooperative -- one hipLaunchCooperativeKernel runs every round
//                             with a grid barrier between rounds
//                             (batched_jacobi.hip.cpp:838-878 shape);
//   chunked / per_round    -- rounds_per_chunk (round kernel + one-thread
//                             advance) pairs, then one controller copy and
//                             stream sync (batched_jacobi.hip.cpp:1540-1587);
//   bfn_tuning.use_graph   -- graph_rounds round pairs and the controller
//                             copy captured once into a hipGraph and replayed
//                             after the preparation launches; rounds beyond
//                             convergence exit at their first instruction,
//                             and a host-polled tail finishes a batch that
//                             needs more rounds than were captured.
#include "bfnew_amd.h"

#include "bfn_launch.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace bfn;

namespace {

thread_local std::string g_error;

int32_t fail(int32_t code, const std::string& message) {
  g_error = message;
  return code;
}

#define BFN_HIP(expr)                                                                 \
  do {                                                                                \
    const hipError_t bfn_status_ = (expr);                                            \
    if (bfn_status_ != hipSuccess)                                                    \
      return fail(BFN_E_HIP, std::string(#expr) + ": " + hipGetErrorString(bfn_status_)); \
  } while (0)

struct DeviceBuffer {
  void* ptr = nullptr;
  std::size_t capacity = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() {
    if (ptr) (void)hipFree(ptr);
  }
  // Geometric growth, like DeviceBuffer::reserve (runtime.hip.cpp:222-240).
  hipError_t reserve(std::size_t bytes) {
    if (bytes <= capacity && ptr) return hipSuccess;
    const std::size_t grown = std::max({bytes, std::size_t{256}, capacity + capacity / 2});
    if (ptr) {
      (void)hipFree(ptr);
      ptr = nullptr;
      capacity = 0;
    }
    hipError_t e = hipMalloc(&ptr, grown);
    if (e == hipSuccess) capacity = grown;
    return e;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

struct PinnedBuffer {
  void* ptr = nullptr;
  std::size_t capacity = 0;
  PinnedBuffer() = default;
  PinnedBuffer(const PinnedBuffer&) = delete;
  PinnedBuffer& operator=(const PinnedBuffer&) = delete;
  ~PinnedBuffer() {
    if (ptr) (void)hipHostFree(ptr);
  }
  hipError_t reserve(std::size_t bytes) {
    if (bytes <= capacity && ptr) return hipSuccess;
    if (ptr) {
      (void)hipHostFree(ptr);
      ptr = nullptr;
      capacity = 0;
    }
    std::size_t grown = std::max<std::size_t>(bytes, 256);
    grown = std::max(grown, capacity + capacity / 2);
    hipError_t e = hipHostMalloc(&ptr, grown, hipHostMallocDefault);
    if (e == hipSuccess) capacity = grown;
    return e;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

constexpr std::size_t align_up(std::size_t x, std::size_t a) { return (x + a - 1) / a * a; }

// LDS words of one relaxation workgroup: own + orig labels, the work area
// (staged slabs, or snapshot + proposals), the table and the flags.
std::size_t lds_words(const DevGraph& g, std::uint32_t max_v, std::uint32_t work_words) {
  return relax_lds_words(max_v, work_words, g.max_neighbor_tiles);
}

bool valid_block(std::uint32_t b) {
  return b == 64 || b == 128 || b == 256 || b == 512 || b == 1024;
}

}  // namespace

struct bfn_device_s {
  int ordinal = 0;
  hipDeviceProp_t props{};
};

struct bfn_graph_s {
  bfn_device device = nullptr;
  hipStream_t stream = nullptr;
  DevGraph view{};
  std::uint32_t V = 0, E = 0, T = 0, Rcsr = 0, Rcsc = 0;
  DeviceBuffer owner, tile_off, row_off, dst, w, row_runs, run_edges, run_dst, col_off, src, cw,
      eid, col_runs, crun_edges, crun_src, in_off, out_off, csc_pack, csr_pack, csc_rec, csr_rec,
      csc_perm, csr_perm, csc_intra, csr_intra, csr_bnd, base_cost, occ_scratch, counts, stats;
  std::uint32_t* in_tiles = nullptr;
  std::uint32_t* in_stage = nullptr;
  std::uint32_t* out_tiles = nullptr;
  std::uint32_t* out_stage = nullptr;
  std::vector<std::uint32_t> h_tile_off, h_in_off, h_out_off;
  bfn_graph_info info{};
  ~bfn_graph_s() {
    if (in_tiles) (void)hipFree(in_tiles);
    if (in_stage) (void)hipFree(in_stage);
    if (out_tiles) (void)hipFree(out_tiles);
    if (out_stage) (void)hipFree(out_stage);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct GraphKey {
  std::uint32_t block = 0, grid = 0, rounds = 0;
  std::size_t lds = 0;
  std::uint32_t max_v = 0, work = 0;
  std::uint64_t pool_words = 0;
  DevGraph g{};
  DevBatch b{};
  bool operator==(const GraphKey& o) const {
    return block == o.block && grid == o.grid && rounds == o.rounds && lds == o.lds &&
           max_v == o.max_v && work == o.work && pool_words == o.pool_words &&
           std::memcmp(&g, &o.g, sizeof g) == 0 && std::memcmp(&b, &o.b, sizeof b) == 0;
  }
};

struct bfn_workspace_s {
  bfn_device device = nullptr;
  hipStream_t stream = nullptr;
  std::uint64_t budget = 0;
  // batch metadata in one device buffer + separately grown state
  DeviceBuffer meta, slot_in, slot_out, slot_out_off, slot_desc, stamp, prio, lists, pool, last_change, control,
      status, relax,
      scratch_v, paths_summary, paths_len, paths_off, frames, onpath, clean, arena_v, arena_l,
      arena_e, occupancy, region_off;
  PinnedBuffer staging, h_status, h_control, h_paths;
  // last batch
  bfn_graph graph = nullptr;
  DevBatch view{};
  std::uint32_t Q = 0, S = 0;
  std::uint64_t pool_words = 0, frame_words = 0;
  std::vector<std::uint32_t> h_target_off, h_tile_off, h_tiles;
  std::vector<std::uint32_t> h_engine;
  bfn_run_options options{};
  bfn_tuning tuning{};
  std::uint32_t block = 256, grid = 0, max_v = 0, work_words = 0;
  std::size_t lds_bytes = 0;
  const unsigned long long* frame_base_ptr = nullptr;  // in `meta`
  bool submitted = false, completed = false, have_paths = false;
  bool paths_counted = false;  // pass 1 of this completion's paths is on the host
  bool slots_resolved = false; // the batch's neighbour slots and descriptors are on the device
  std::uint64_t arena_vertices = 0;
  // Path-kernel marks in `onpath`: the next pass marks with stamps
  // [path_stamp + 1, path_stamp + targets]; the buffer was last cleared at
  // onpath_cleared for onpath_cleared_words words.
  std::uint32_t path_stamp = 0;
  const void* onpath_cleared = nullptr;
  std::uint64_t onpath_cleared_words = 0;
  std::uint64_t launches = 0, host_checks = 0, host_syncs = 0, controller_copies = 0;
  hipEvent_t ev_start = nullptr, ev_relax0 = nullptr, ev_relax1 = nullptr, ev_stop = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph_obj = nullptr;
  GraphKey graph_key{};
  ~bfn_workspace_s() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph_obj) (void)hipGraphDestroy(graph_obj);
    for (hipEvent_t e : {ev_start, ev_relax0, ev_relax1, ev_stop})
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// Entry points keep the C linkage of their declarations in bfnew_amd.h.
const char* bfn_last_error(void) { return g_error.c_str(); }
uint32_t bfn_abi_version(void) { return 1u; }

int32_t bfn_device_open(int32_t ordinal, bfn_device* out) {
  if (!out) return fail(BFN_E_INVALID_ARGUMENT, "null output handle");
  int count = 0;
  BFN_HIP(hipGetDeviceCount(&count));
  if (ordinal < 0 || ordinal >= count)
    return fail(BFN_E_INVALID_ARGUMENT, "device ordinal out of range");
  BFN_HIP(hipSetDevice(ordinal));
  auto* d = new bfn_device_s;
  d->ordinal = ordinal;
  if (hipGetDeviceProperties(&d->props, ordinal) != hipSuccess) {
    delete d;
    return fail(BFN_E_HIP, "hipGetDeviceProperties failed");
  }
  *out = d;
  return BFN_OK;
}

int32_t bfn_device_close(bfn_device device) {
  delete device;
  return BFN_OK;
}

int32_t bfn_device_query(bfn_device device, bfn_device_info* out) {
  if (!device || !out) return fail(BFN_E_INVALID_ARGUMENT, "null argument");
  std::memset(out, 0, sizeof *out);
  out->compute_units = static_cast<std::uint32_t>(device->props.multiProcessorCount);
  out->wavefront_size = static_cast<std::uint32_t>(device->props.warpSize);
  out->lds_bytes_per_cu = static_cast<std::uint32_t>(device->props.maxSharedMemoryPerMultiProcessor);
  out->global_memory_bytes = device->props.totalGlobalMem;
  std::snprintf(out->arch, sizeof out->arch, "%s", device->props.gcnArchName);
  const bool multi_die = std::strncmp(out->arch, "gfx94", 5) == 0 || std::strncmp(out->arch, "gfx95", 5) == 0;
  out->xcd_count = multi_die ? 8u : 1u;
  return BFN_OK;
}

int32_t bfn_device_synchronize(bfn_device device) {
  if (!device) return fail(BFN_E_INVALID_ARGUMENT, "null device");
  BFN_HIP(hipSetDevice(device->ordinal));
  BFN_HIP(hipDeviceSynchronize());
  return BFN_OK;
}

// ---------------------------------------------------------------------------
// Graph.
// ---------------------------------------------------------------------------
namespace {

bool offsets_valid(const std::uint32_t* off, std::uint32_t n, std::uint32_t total) {
  if (!off || off[0] != 0 || off[n] != total) return false;
  for (std::uint32_t i = 0; i < n; ++i)
    if (off[i + 1] < off[i]) return false;
  return true;
}

// Tile runs of one orientation (device_layout.cpp:280-322): per row, the runs
// cover the row's edges in order, are non-empty and maximal, name a tile in
// range, and every edge of a run has its other endpoint in the run's tile.
// The device derivation of the neighbour lists relies on all of it.
const char* runs_invalid(std::uint32_t V, std::uint32_t T, const std::uint32_t* owner,
                         const std::uint32_t* row_offsets, const std::uint32_t* endpoints,
                         const std::uint32_t* row_runs, const std::uint32_t* run_edges,
                         const std::uint32_t* run_tiles) {
  for (std::uint32_t v = 0; v < V; ++v) {
    std::uint32_t cursor = row_offsets[v];
    std::uint32_t previous = kNone;
    for (std::uint32_t r = row_runs[v]; r < row_runs[v + 1]; ++r) {
      const std::uint32_t b = run_edges[r], e = run_edges[r + 1];
      if (b != cursor) return "tile runs do not cover their row's edges in order";
      if (e <= b) return "empty tile run";
      if (e > row_offsets[v + 1]) return "tile run crosses its row";
      const std::uint32_t tile = run_tiles[r];
      if (tile >= T) return "tile run names a tile out of range";
      if (tile == previous) return "tile runs are not maximal";
      for (std::uint32_t x = b; x < e; ++x)
        if (owner[endpoints[x]] != tile) return "an edge's endpoint is not in its run's tile";
      previous = tile;
      cursor = e;
    }
    if (cursor != row_offsets[v + 1]) return "tile runs do not cover their row's edges in order";
  }
  return nullptr;
}

template <typename T>
hipError_t upload(DeviceBuffer& buf, const T* host, std::size_t count, hipStream_t stream) {
  hipError_t e = buf.reserve(sizeof(T) * std::max<std::size_t>(count, 1));
  if (e != hipSuccess || count == 0) return e;
  e = hipMemcpyAsync(buf.ptr, host, sizeof(T) * count, hipMemcpyHostToDevice, stream);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(stream);  // the caller may free host memory on return
}

}  // namespace

namespace {

// Checks of validate_device_graph_layout32 / validate_tile_run_layout
// (device_layout.cpp:253-393) that the engine image depends on; writes the
// largest tile size.
int32_t validate_layout(const bfn_graph_layout32* L, std::uint32_t* largest_tile) {
  if (!L) return fail(BFN_E_INVALID_ARGUMENT, "null layout");
  const std::uint32_t V = L->vertex_count, E = L->edge_count, T = L->tile_count;
  if (T == 0) return fail(BFN_E_INVALID_ARGUMENT, "graph has no tiles");
  if (!L->owner_tiles && V) return fail(BFN_E_INVALID_ARGUMENT, "missing owner tiles");
  if (!offsets_valid(L->tile_vertex_offsets, T, V))
    return fail(BFN_E_INVALID_ARGUMENT, "tile vertex offsets are not a partition of the vertices");
  if (!offsets_valid(L->csr_row_offsets, V, E) || !offsets_valid(L->csc_column_offsets, V, E))
    return fail(BFN_E_INVALID_ARGUMENT, "invalid CSR/CSC offsets");
  if (!offsets_valid(L->csr_row_run_offsets, V, L->csr_run_count) ||
      !offsets_valid(L->csc_column_run_offsets, V, L->csc_run_count) ||
      !offsets_valid(L->csr_run_edge_offsets, L->csr_run_count, E) ||
      !offsets_valid(L->csc_run_edge_offsets, L->csc_run_count, E))
    return fail(BFN_E_INVALID_ARGUMENT, "invalid tile-run offsets");
  std::uint32_t max_v = 0;
  for (std::uint32_t t = 0; t < T; ++t) {
    const std::uint32_t n = L->tile_vertex_offsets[t + 1] - L->tile_vertex_offsets[t];
    max_v = std::max(max_v, n);
    for (std::uint32_t v = L->tile_vertex_offsets[t]; v < L->tile_vertex_offsets[t + 1]; ++v)
      if (L->owner_tiles[v] != t)
        return fail(BFN_E_INVALID_ARGUMENT, "owner tiles disagree with the tile spans");
  }
  if (max_v > kMaxTileVertices)
    return fail(BFN_E_NOT_REPRESENTABLE, "a tile holds more than 4096 vertices");
  if (E && (!L->csr_destinations || !L->csr_weights || !L->csc_sources || !L->csc_weights ||
            !L->csc_edge_ids))
    return fail(BFN_E_INVALID_ARGUMENT, "missing edge arrays");
  for (std::uint32_t e = 0; e < E; ++e)
    if (!(L->csr_weights[e] >= 0.0f) || !(L->csc_weights[e] >= 0.0f))
      return fail(BFN_E_INVALID_ARGUMENT, "edge weights must be nonnegative");
  for (std::uint32_t e = 0; e < E; ++e)
    if (L->csr_destinations[e] >= V || L->csc_sources[e] >= V || L->csc_edge_ids[e] >= E)
      return fail(BFN_E_INVALID_ARGUMENT, "edge endpoint or edge id out of range");
  if (!L->csr_run_destination_tiles || !L->csc_run_source_tiles)
    return fail(BFN_E_INVALID_ARGUMENT, "missing tile-run tiles");
  if (const char* why = runs_invalid(V, T, L->owner_tiles, L->csr_row_offsets, L->csr_destinations,
                                     L->csr_row_run_offsets, L->csr_run_edge_offsets,
                                     L->csr_run_destination_tiles))
    return fail(BFN_E_INVALID_ARGUMENT, std::string("CSR: ") + why);
  if (const char* why = runs_invalid(V, T, L->owner_tiles, L->csc_column_offsets, L->csc_sources,
                                     L->csc_column_run_offsets, L->csc_run_edge_offsets,
                                     L->csc_run_source_tiles))
    return fail(BFN_E_INVALID_ARGUMENT, std::string("CSC: ") + why);
  *largest_tile = max_v;
  return BFN_OK;
}

}  // namespace

int32_t bfn_graph_validate(const bfn_graph_layout32* layout) {
  std::uint32_t max_v = 0;
  return validate_layout(layout, &max_v);
}

int32_t bfn_graph_upload(bfn_device device, const bfn_graph_layout32* L, bfn_graph* out) {
  if (!device || !L || !out) return fail(BFN_E_INVALID_ARGUMENT, "null argument");
  std::uint32_t max_v = 0;
  if (int32_t rc = validate_layout(L, &max_v)) return rc;
  const std::uint32_t V = L->vertex_count, E = L->edge_count, T = L->tile_count;

  BFN_HIP(hipSetDevice(device->ordinal));
  auto* g = new bfn_graph_s;
  g->device = device;
  g->V = V;
  g->E = E;
  g->T = T;
  g->Rcsr = L->csr_run_count;
  g->Rcsc = L->csc_run_count;
  auto cleanup = [&](int32_t code, const std::string& m) {
    delete g;
    return fail(code, m);
  };
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(BFN_E_HIP, "stream creation failed");
  hipStream_t s = g->stream;
  hipError_t e = hipSuccess;
#define BFN_UP(buf, ptr, n) \
  if (e == hipSuccess) e = upload(g->buf, L->ptr, n, s)
  BFN_UP(owner, owner_tiles, V);
  BFN_UP(tile_off, tile_vertex_offsets, T + 1);
  BFN_UP(row_off, csr_row_offsets, V + 1);
  BFN_UP(dst, csr_destinations, E);
  BFN_UP(w, csr_weights, E);
  BFN_UP(row_runs, csr_row_run_offsets, V + 1);
  BFN_UP(run_edges, csr_run_edge_offsets, g->Rcsr + 1);
  BFN_UP(run_dst, csr_run_destination_tiles, g->Rcsr);
  BFN_UP(col_off, csc_column_offsets, V + 1);
  BFN_UP(src, csc_sources, E);
  BFN_UP(cw, csc_weights, E);
  BFN_UP(eid, csc_edge_ids, E);
  BFN_UP(col_runs, csc_column_run_offsets, V + 1);
  BFN_UP(crun_edges, csc_run_edge_offsets, g->Rcsc + 1);
  BFN_UP(crun_src, csc_run_source_tiles, g->Rcsc);
#undef BFN_UP
  if (e == hipSuccess) e = g->in_off.reserve(sizeof(std::uint32_t) * (T + 1));
  if (e == hipSuccess) e = g->out_off.reserve(sizeof(std::uint32_t) * (T + 1));
  if (e == hipSuccess) e = g->counts.reserve(sizeof(std::uint32_t) * (T + 1));
  const std::size_t e_words = std::max(E, 1u);
  if (e == hipSuccess) e = g->csc_pack.reserve(sizeof(std::uint32_t) * e_words);
  if (e == hipSuccess) e = g->csr_pack.reserve(sizeof(std::uint32_t) * e_words);
  if (e == hipSuccess) e = g->csc_rec.reserve(sizeof(uint2) * e_words);
  // The push visits read the CSR stream two records per 16-byte load from
  // even positions, so it holds one record to spare.
  if (e == hipSuccess) e = g->csr_rec.reserve(sizeof(uint2) * (e_words + 1));
  if (e == hipSuccess) e = g->csc_perm.reserve(sizeof(std::uint32_t) * e_words);
  if (e == hipSuccess) e = g->csr_perm.reserve(sizeof(std::uint32_t) * e_words);
  if (e == hipSuccess) e = g->csc_intra.reserve(sizeof(std::uint32_t) * (T + 1));
  if (e == hipSuccess) e = g->csr_intra.reserve(sizeof(std::uint32_t) * (T + 1));
  if (e == hipSuccess) e = g->csr_bnd.reserve(sizeof(std::uint32_t) * std::max(V, 1u));
  if (e == hipSuccess) e = g->stats.reserve(sizeof(std::uint32_t) * 8);
  if (e == hipSuccess) e = hipMemsetAsync(g->stats.ptr, 0, sizeof(std::uint32_t) * 8, s);
  if (e != hipSuccess) return cleanup(BFN_E_HIP, std::string("graph upload: ") + hipGetErrorString(e));

  std::uint32_t in_total = 0, out_total = 0;
  TileImageArgs ta{};
  ta.tile_count = T;
  ta.tile_vertex_offsets = g->tile_off.as<std::uint32_t>();
  ta.owner_tiles = g->owner.as<std::uint32_t>();
  ta.row_offsets = g->row_off.as<std::uint32_t>();
  ta.destinations = g->dst.as<std::uint32_t>();
  ta.row_runs = g->row_runs.as<std::uint32_t>();
  ta.run_dst_tiles = g->run_dst.as<std::uint32_t>();
  ta.weights = g->w.as<float>();
  ta.column_offsets = g->col_off.as<std::uint32_t>();
  ta.sources = g->src.as<std::uint32_t>();
  ta.col_runs = g->col_runs.as<std::uint32_t>();
  ta.run_src_tiles = g->crun_src.as<std::uint32_t>();
  ta.in_weights = g->cw.as<float>();
  ta.counts_scratch = g->counts.as<std::uint32_t>();
  ta.in_offsets = g->in_off.as<std::uint32_t>();
  ta.out_offsets = g->out_off.as<std::uint32_t>();
  ta.in_tiles = &g->in_tiles;
  ta.in_stage = &g->in_stage;
  ta.out_tiles = &g->out_tiles;
  ta.out_stage = &g->out_stage;
  ta.csc_pack = g->csc_pack.as<std::uint32_t>();
  ta.csr_pack = g->csr_pack.as<std::uint32_t>();
  ta.csc_rec = g->csc_rec.as<uint2>();
  ta.csr_rec = g->csr_rec.as<uint2>();
  ta.csc_perm = g->csc_perm.as<std::uint32_t>();
  ta.csr_perm = g->csr_perm.as<std::uint32_t>();
  ta.csc_intra = g->csc_intra.as<std::uint32_t>();
  ta.csr_intra = g->csr_intra.as<std::uint32_t>();
  ta.csr_bnd = g->csr_bnd.as<std::uint32_t>();
  ta.stats = g->stats.as<std::uint32_t>();
  ta.host_in_total = &in_total;
  ta.host_out_total = &out_total;
  e = derive_tile_image(ta, s);
  if (e != hipSuccess) return cleanup(BFN_E_HIP, std::string("tile image: ") + hipGetErrorString(e));
  std::uint32_t stats[8] = {};
  g->h_tile_off.assign(L->tile_vertex_offsets, L->tile_vertex_offsets + T + 1);
  g->h_in_off.resize(T + 1);
  g->h_out_off.resize(T + 1);
  e = hipMemcpyAsync(stats, g->stats.ptr, sizeof stats, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(g->h_in_off.data(), g->in_off.ptr, sizeof(std::uint32_t) * (T + 1),
                       hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(g->h_out_off.data(), g->out_off.ptr, sizeof(std::uint32_t) * (T + 1),
                       hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return cleanup(BFN_E_HIP, std::string("tile image: ") + hipGetErrorString(e));
  if (stats[0]) return cleanup(BFN_E_NOT_REPRESENTABLE, "a tile has more than 255 neighbour tiles");

  DevGraph& v = g->view;
  v.vertex_count = V;
  v.edge_count = E;
  v.tile_count = T;
  v.max_tile_vertices = std::max(max_v, 1u);
  v.max_stage_floats = std::max(stats[1], 1u);
  v.max_in_edges = stats[2];
  v.max_out_edges = stats[5];
  v.max_out_stage_floats = stats[6];
  std::uint32_t max_nbr = 1;
  for (std::uint32_t t = 0; t < T; ++t)
    max_nbr = std::max({max_nbr, g->h_in_off[t + 1] - g->h_in_off[t], g->h_out_off[t + 1] - g->h_out_off[t]});
  v.max_neighbor_tiles = max_nbr;
  v.owner_tiles = g->owner.as<std::uint32_t>();
  v.tile_vertex_offsets = g->tile_off.as<std::uint32_t>();
  v.row_offsets = g->row_off.as<std::uint32_t>();
  v.destinations = g->dst.as<std::uint32_t>();
  v.weights = g->w.as<float>();
  v.column_offsets = g->col_off.as<std::uint32_t>();
  v.sources = g->src.as<std::uint32_t>();
  v.in_weights = g->cw.as<float>();
  v.in_edge_ids = g->eid.as<std::uint32_t>();
  v.in_nbr_offsets = g->in_off.as<std::uint32_t>();
  v.in_nbr_tiles = g->in_tiles;
  v.in_nbr_stage = g->in_stage;
  v.out_nbr_offsets = g->out_off.as<std::uint32_t>();
  v.out_nbr_tiles = g->out_tiles;
  v.out_nbr_stage = g->out_stage;
  v.csc_pack = g->csc_pack.as<std::uint32_t>();
  v.csr_pack = g->csr_pack.as<std::uint32_t>();
  v.csc_rec = g->csc_rec.as<uint2>();
  v.csr_rec = g->csr_rec.as<uint2>();
  v.csc_intra = g->csc_intra.as<std::uint32_t>();
  v.csr_intra = g->csr_intra.as<std::uint32_t>();
  v.csr_bnd = g->csr_bnd.as<std::uint32_t>();

  bfn_graph_info& info = g->info;
  info.max_tile_vertices = max_v;
  info.max_tile_in_edges = stats[2];
  info.max_tile_out_edges = stats[5];
  info.max_in_neighbor_tiles = 0;
  info.max_out_neighbor_tiles = 0;
  for (std::uint32_t t = 0; t < T; ++t) {
    info.max_in_neighbor_tiles = std::max(info.max_in_neighbor_tiles, g->h_in_off[t + 1] - g->h_in_off[t]);
    info.max_out_neighbor_tiles = std::max(info.max_out_neighbor_tiles, g->h_out_off[t + 1] - g->h_out_off[t]);
  }
  info.max_stage_floats = stats[1];
  info.has_zero_weight = stats[4];
  // reference image + derived tile lists + packs (4 B) + records (8 B) +
  // record permutations (4 B) per edge and orientation
  info.device_bytes = 4ull * (3ull * V + 2 + T + 1 + 6ull * E + 2ull * (g->Rcsr + g->Rcsc + 2) +
                              4ull * (T + 1) + 2ull * in_total + 2ull * out_total) +
                      2ull * E * (4 + 8 + 4);
  *out = g;
  return BFN_OK;
}

int32_t bfn_graph_free(bfn_graph graph) {
  if (graph) {
    (void)hipSetDevice(graph->device->ordinal);
    (void)hipDeviceSynchronize();
  }
  delete graph;
  return BFN_OK;
}

int32_t bfn_graph_query(bfn_graph graph, bfn_graph_info* out) {
  if (!graph || !out) return fail(BFN_E_INVALID_ARGUMENT, "null argument");
  *out = graph->info;
  return BFN_OK;
}

int32_t bfn_graph_update_weights(bfn_graph graph, const float* csr_weights,
                                 const float* csc_weights) {
  if (!graph || !csr_weights || !csc_weights) return fail(BFN_E_INVALID_ARGUMENT, "null argument");
  // The two orientations hold the same edges in different orders: push
  // engines read the CSR weights, pull and path reconstruction the CSC
  // weights.  Their multisets must agree: compared through order-free exact
  // sums of the weights' bit patterns and of their squares (mod 2^64) and
  // the zero counts -- checked in O(E) without matching the edges.
  std::uint64_t csr_sum = 0, csc_sum = 0, csr_sq = 0, csc_sq = 0, csr_zero = 0, csc_zero = 0;
  for (std::uint32_t e = 0; e < graph->E; ++e) {
    if (!(csr_weights[e] >= 0.0f) || !(csc_weights[e] >= 0.0f) || csr_weights[e] > 3.4e38f ||
        csc_weights[e] > 3.4e38f)
      return fail(BFN_E_INVALID_ARGUMENT, "weights must be finite and nonnegative");
    std::uint32_t a, b;
    std::memcpy(&a, &csr_weights[e], 4);
    std::memcpy(&b, &csc_weights[e], 4);
    csr_sum += a;
    csc_sum += b;
    csr_sq += static_cast<std::uint64_t>(a) * a;
    csc_sq += static_cast<std::uint64_t>(b) * b;
    csr_zero += a == 0u;
    csc_zero += b == 0u;
  }
  if (csr_sum != csc_sum || csr_sq != csc_sq || csr_zero != csc_zero)
    return fail(BFN_E_INVALID_ARGUMENT, "CSR and CSC weights do not hold the same edge weights");
  graph->info.has_zero_weight = csr_zero != 0;
  BFN_HIP(hipSetDevice(graph->device->ordinal));
  BFN_HIP(hipMemcpyAsync(graph->w.ptr, csr_weights, sizeof(float) * graph->E,
                         hipMemcpyHostToDevice, graph->stream));
  BFN_HIP(hipMemcpyAsync(graph->cw.ptr, csc_weights, sizeof(float) * graph->E,
                         hipMemcpyHostToDevice, graph->stream));
  BFN_HIP(refresh_record_weights(graph->view, graph->csr_perm.as<std::uint32_t>(),
                                 graph->csc_perm.as<std::uint32_t>(), graph->stream));
  BFN_HIP(hipStreamSynchronize(graph->stream));
  return BFN_OK;
}

namespace {
int32_t apply_node_costs(bfn_graph graph, const float* base_cost, const int* device_occ) {
  if (base_cost) {
    for (std::uint32_t v = 0; v < graph->V; ++v)
      if (!(base_cost[v] >= 0.0f) || base_cost[v] > 3.4e38f)
        return fail(BFN_E_INVALID_ARGUMENT, "node costs must be finite and nonnegative");
    BFN_HIP(graph->base_cost.reserve(sizeof(float) * std::max(graph->V, 1u)));
    BFN_HIP(hipMemcpyAsync(graph->base_cost.ptr, base_cost, sizeof(float) * graph->V,
                           hipMemcpyHostToDevice, graph->stream));
  } else if (!graph->base_cost.ptr) {
    return fail(BFN_E_STATE, "no base node costs uploaded yet");
  }
  BFN_HIP(node_cost_launch(graph->view, graph->base_cost.as<float>(), device_occ,
                           graph->w.as<float>(), graph->cw.as<float>(), graph->stream));
  BFN_HIP(refresh_record_weights(graph->view, graph->csr_perm.as<std::uint32_t>(),
                                 graph->csc_perm.as<std::uint32_t>(), graph->stream));
  BFN_HIP(hipStreamSynchronize(graph->stream));
  return BFN_OK;
}
}  // namespace

int32_t bfn_graph_update_node_costs(bfn_graph graph, const float* base_cost,
                                    const int32_t* occupancy) {
  if (!graph) return fail(BFN_E_INVALID_ARGUMENT, "null graph");
  BFN_HIP(hipSetDevice(graph->device->ordinal));
  const int* occ = nullptr;
  if (occupancy) {
    for (std::uint32_t v = 0; v < graph->V; ++v)
      if (occupancy[v] < 0) return fail(BFN_E_INVALID_ARGUMENT, "negative occupancy");
    BFN_HIP(graph->occ_scratch.reserve(sizeof(int) * std::max(graph->V, 1u)));
    BFN_HIP(hipMemcpyAsync(graph->occ_scratch.ptr, occupancy, sizeof(int) * graph->V,
                           hipMemcpyHostToDevice, graph->stream));
    occ = graph->occ_scratch.as<int>();
  }
  return apply_node_costs(graph, base_cost, occ);
}

int32_t bfn_graph_update_node_costs_from_workspace(bfn_graph graph, bfn_workspace ws,
                                                   const float* base_cost) {
  if (!graph || !ws) return fail(BFN_E_INVALID_ARGUMENT, "null argument");
  if (ws->graph != graph) return fail(BFN_E_STATE, "workspace ran on another graph");
  BFN_HIP(hipSetDevice(graph->device->ordinal));
  // Count the occupancy of the workspace's current compact paths on device.
  if (int32_t rc = bfn_batch_occupancy(ws, nullptr)) return rc;
  return apply_node_costs(graph, base_cost, ws->occupancy.as<int>());
}

// ---------------------------------------------------------------------------
// Workspace and batches.
// ---------------------------------------------------------------------------
int32_t bfn_workspace_create(bfn_device device, uint64_t budget_bytes, bfn_workspace* out) {
  if (!device || !out) return fail(BFN_E_INVALID_ARGUMENT, "null argument");
  BFN_HIP(hipSetDevice(device->ordinal));
  auto* ws = new bfn_workspace_s;
  ws->device = device;
  ws->budget = budget_bytes;
  if (hipStreamCreateWithFlags(&ws->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ws;
    return fail(BFN_E_HIP, "stream creation failed");
  }
  for (hipEvent_t* e : {&ws->ev_start, &ws->ev_relax0, &ws->ev_relax1, &ws->ev_stop}) {
    if (hipEventCreate(e) != hipSuccess) {
      delete ws;
      return fail(BFN_E_HIP, "event creation failed");
    }
  }
  *out = ws;
  return BFN_OK;
}

int32_t bfn_workspace_free(bfn_workspace ws) {
  if (ws) {
    (void)hipSetDevice(ws->device->ordinal);
    (void)hipStreamSynchronize(ws->stream);
  }
  delete ws;
  return BFN_OK;
}

void bfn_tuning_default(bfn_tuning* out) {
  if (!out) return;
  std::memset(out, 0, sizeof *out);
  out->tile_sweeps = 1;
  out->items_per_claim = 1;
  out->graph_rounds = 0;
  out->use_graph = 0;
  out->count_relaxations = 0;
}

int32_t bfn::check_run_options(const bfn_run_options* o) {
  if (!o) return fail(BFN_E_INVALID_ARGUMENT, "null run options");
  if (o->engine > 2) return fail(BFN_E_INVALID_ARGUMENT, "invalid engine");
  if (o->control_mode > 2) return fail(BFN_E_INVALID_ARGUMENT, "invalid control mode");
  if (o->control_mode == 1 && o->rounds_per_chunk == 0)
    return fail(BFN_E_INVALID_ARGUMENT, "invalid rounds per chunk");
  if (!valid_block(o->block_size)) return fail(BFN_E_INVALID_ARGUMENT, "invalid block size");
  if (o->grid_policy > 1) return fail(BFN_E_INVALID_ARGUMENT, "invalid grid policy");
  if (o->grid_policy == 1 && o->blocks_per_wgp == 0)
    return fail(BFN_E_INVALID_ARGUMENT, "invalid blocks per WGP");
  if (o->instrumentation > 2) return fail(BFN_E_INVALID_ARGUMENT, "invalid instrumentation");
  if (o->maximum_rounds == 0) return fail(BFN_E_INVALID_ARGUMENT, "invalid maximum rounds");
  if (o->enable_per_lane_convergence > 1) return fail(BFN_E_INVALID_ARGUMENT, "invalid per-lane flag");
  return BFN_OK;
}

namespace {

int32_t validate_options(const bfn_run_options* o) { return check_run_options(o); }

struct BatchPlanHost {
  std::uint32_t Q = 0, S = 0;
  std::uint64_t pool_words = 0, frame_words = 0, in_total = 0, out_total = 0;
  std::vector<std::uint32_t> slot_query, slot_tile, query_slot, slot_in_base, slot_out_base, engine;
  std::vector<unsigned long long> slot_dist, frame_base;
  bool any_jacobi = false;
  bool any_push = false;
};

}  // namespace

// Host checks of a batch description (shapes, canonical terminals inside
// their query's selected tiles, sorted unique tiles in range, engines).
int32_t bfn::check_batch(bfn_graph g, const bfn_batch_desc* d, const bfn_run_options* o) {
  if (!d) return fail(BFN_E_INVALID_ARGUMENT, "null batch");
  const std::uint32_t Q = d->query_count;
  if (Q == 0) return fail(BFN_E_INVALID_ARGUMENT, "empty batch");
  if (!d->source_offsets || !d->target_offsets || !d->tile_offsets || !d->sources ||
      !d->targets || !d->selected_tiles)
    return fail(BFN_E_INVALID_ARGUMENT, "missing batch arrays");
  for (std::uint32_t q = 0; q < Q; ++q) {
    const std::uint32_t s0 = d->source_offsets[q], s1 = d->source_offsets[q + 1];
    const std::uint32_t t0 = d->target_offsets[q], t1 = d->target_offsets[q + 1];
    const std::uint32_t k0 = d->tile_offsets[q], k1 = d->tile_offsets[q + 1];
    if (s1 <= s0 || t1 <= t0 || k1 <= k0)
      return fail(BFN_E_INVALID_ARGUMENT, "query " + std::to_string(q) +
                                              " needs sources, targets and selected tiles");
    for (std::uint32_t i = k0; i < k1; ++i) {
      if (d->selected_tiles[i] >= g->T || (i > k0 && d->selected_tiles[i] <= d->selected_tiles[i - 1]))
        return fail(BFN_E_INVALID_ARGUMENT, "selected tiles must be sorted, unique and in range");
    }
    auto selected = [&](std::uint32_t v) {
      std::uint32_t tile = 0;
      // owner tile by binary search of the tile spans
      const auto& off = g->h_tile_off;
      tile = static_cast<std::uint32_t>(std::upper_bound(off.begin(), off.end(), v) - off.begin()) - 1;
      return std::binary_search(d->selected_tiles + k0, d->selected_tiles + k1, tile);
    };
    for (std::uint32_t i = s0; i < s1; ++i) {
      if (d->sources[i] >= g->V || (i > s0 && d->sources[i] <= d->sources[i - 1]))
        return fail(BFN_E_INVALID_ARGUMENT, "sources must be canonical and in range");
      if (!selected(d->sources[i]))
        return fail(BFN_E_INVALID_ARGUMENT, "a source lies outside its query's selected tiles");
    }
    for (std::uint32_t i = t0; i < t1; ++i) {
      if (d->targets[i] >= g->V || (i > t0 && d->targets[i] <= d->targets[i - 1]))
        return fail(BFN_E_INVALID_ARGUMENT, "targets must be canonical and in range");
      if (!selected(d->targets[i]))
        return fail(BFN_E_INVALID_ARGUMENT, "a target lies outside its query's selected tiles");
    }
    const std::uint32_t eng = d->engine_by_query ? d->engine_by_query[q] : o->engine;
    if (eng > 2) return fail(BFN_E_INVALID_ARGUMENT, "invalid per-query engine");
  }
  if (d->source_offsets[0] != 0 || d->target_offsets[0] != 0 || d->tile_offsets[0] != 0)
    return fail(BFN_E_INVALID_ARGUMENT, "offsets must start at zero");
  return BFN_OK;
}

namespace {

int32_t plan_batch(bfn_graph g, const bfn_batch_desc* d, const bfn_run_options* o,
                   BatchPlanHost& p) {
  if (int32_t rc = check_batch(g, d, o)) return rc;
  const std::uint32_t Q = d->query_count;
  p.Q = Q;
  p.S = d->tile_offsets[Q];
  const std::uint32_t S = p.S;
  p.slot_query.resize(S);
  p.slot_tile.resize(S);
  p.query_slot.resize(S);
  p.slot_in_base.resize(S);
  p.slot_out_base.resize(S);
  p.slot_dist.resize(S);
  p.engine.resize(Q);
  p.frame_base.resize(Q);
  // Tile-major slot numbering (counting sort of the (tile, query) pairs).
  std::vector<std::uint32_t> cursor(g->T + 1, 0);
  for (std::uint32_t i = 0; i < S; ++i) ++cursor[d->selected_tiles[i] + 1];
  for (std::uint32_t t = 0; t < g->T; ++t) cursor[t + 1] += cursor[t];
  std::uint64_t frames = 0;
  for (std::uint32_t q = 0; q < Q; ++q) {
    p.engine[q] = d->engine_by_query ? d->engine_by_query[q] : o->engine;
    p.any_jacobi |= p.engine[q] == kJacobiPull;
    p.any_push |= p.engine[q] != kJacobiPull;
    p.frame_base[q] = frames;
    for (std::uint32_t pos = d->tile_offsets[q]; pos < d->tile_offsets[q + 1]; ++pos) {
      const std::uint32_t T = d->selected_tiles[pos];
      const std::uint32_t s = cursor[T]++;
      p.query_slot[pos] = s;
      p.slot_tile[s] = T;
      p.slot_query[s] = q;
      frames += g->h_tile_off[T + 1] - g->h_tile_off[T];
    }
  }
  // Labels and neighbour tables in slot order: a tile's queries sit together.
  std::uint64_t pool = 0, in_total = 0, out_total = 0;
  for (std::uint32_t s = 0; s < S; ++s) {
    const std::uint32_t T = p.slot_tile[s];
    p.slot_dist[s] = pool;
    pool += 2ull * (g->h_tile_off[T + 1] - g->h_tile_off[T]);
    p.slot_in_base[s] = static_cast<std::uint32_t>(in_total);
    p.slot_out_base[s] = static_cast<std::uint32_t>(out_total);
    in_total += g->h_in_off[T + 1] - g->h_in_off[T];
    out_total += g->h_out_off[T + 1] - g->h_out_off[T];
  }
  if (in_total >= 0xFFFFFFFFull || out_total >= 0xFFFFFFFFull)
    return fail(BFN_E_NOT_REPRESENTABLE, "batch neighbour tables exceed 32-bit indexing");
  p.pool_words = pool;
  p.frame_words = frames;
  p.in_total = in_total;
  p.out_total = out_total;
  return BFN_OK;
}

// Batch metadata, uploaded in one copy: 0 tile offsets, 1 query tiles, 2-5
// sources / targets, 6 engines, 7 slot query, 8 slot pool offset, 9-10
// neighbour bases, 11 query slot, 12 slot tile, 13 frame base (u64 per query,
// used by compact paths).  Returns the total bytes.
constexpr int kMetaSegments = 14;
std::size_t meta_layout(std::uint32_t Q, std::uint32_t S, std::uint32_t ns, std::uint32_t nt,
                        std::size_t* seg, std::size_t* sizes) {
  const std::size_t n[kMetaSegments] = {4ull * (Q + 1), 4ull * S, 4ull * (Q + 1), 4ull * ns,
                                        4ull * (Q + 1), 4ull * nt, 4ull * Q, 4ull * S, 8ull * S,
                                        4ull * S, 4ull * S, 4ull * S, 4ull * S, 8ull * Q};
  std::size_t off = 0;
  for (int i = 0; i < kMetaSegments; ++i) {
    seg[i] = off;
    sizes[i] = n[i];
    off = align_up(off + n[i], 256);
  }
  return off;
}

std::uint64_t batch_bytes(const BatchPlanHost& p, std::uint64_t meta_bytes) {
  return meta_bytes + 4ull * (p.in_total + p.out_total) + 8ull * p.out_total +
         (20ull + sizeof(SlotDesc)) * p.S +
         4ull * p.pool_words + 4ull * p.Q * 2 + sizeof(QueryStatusDev) * p.Q + 8ull * p.Q +
         sizeof(DevControl);
}

// Sets up the DevBatch view over the workspace buffers.
void bind_view(bfn_workspace ws, const BatchPlanHost& p, const std::size_t* seg,
               const bfn_run_options* o, const bfn_tuning* t) {
  char* m = ws->meta.as<char>();
  DevBatch& b = ws->view;
  b.query_count = p.Q;
  b.slot_count = p.S;
  b.maximum_rounds = o->maximum_rounds >= 0xFFFFFFFFull ? 0u : static_cast<std::uint32_t>(o->maximum_rounds);
  b.tile_sweeps = t->tile_sweeps;
  b.items_per_claim = std::max(1u, t->items_per_claim);
  b.instrumentation = o->instrumentation;
  b.boundary_mode = t->boundary_mode;
  b.min_waves = t->min_waves;
  b.ordered = t->claim_order;
  std::memcpy(&b.window, &t->push_window, sizeof b.window);
  b.tile_begin = reinterpret_cast<const std::uint32_t*>(m + seg[0]);
  b.tile_end = b.tile_begin + 1;
  b.query_tiles = reinterpret_cast<const std::uint32_t*>(m + seg[1]);
  b.query_slot = reinterpret_cast<const std::uint32_t*>(m + seg[11]);
  b.tiles = reinterpret_cast<const std::uint32_t*>(m + seg[12]);
  b.source_offsets = reinterpret_cast<const std::uint32_t*>(m + seg[2]);
  b.sources = reinterpret_cast<const std::uint32_t*>(m + seg[3]);
  b.target_offsets = reinterpret_cast<const std::uint32_t*>(m + seg[4]);
  b.targets = reinterpret_cast<const std::uint32_t*>(m + seg[5]);
  b.engine = reinterpret_cast<const std::uint32_t*>(m + seg[6]);
  b.slot_query = reinterpret_cast<const std::uint32_t*>(m + seg[7]);
  b.slot_dist = reinterpret_cast<const unsigned long long*>(m + seg[8]);
  b.slot_in_base = reinterpret_cast<const std::uint32_t*>(m + seg[9]);
  b.slot_out_base = reinterpret_cast<const std::uint32_t*>(m + seg[10]);
  b.slot_in = ws->slot_in.as<std::uint32_t>();
  b.slot_out = ws->slot_out.as<std::uint32_t>();
  b.slot_out_off = ws->slot_out_off.as<unsigned long long>();
  b.slot_desc = ws->slot_desc.as<SlotDesc>();
  b.stamp = ws->stamp.as<std::uint32_t>();
  b.active = b.stamp;
  b.prio = b.window ? ws->prio.as<std::uint32_t>() : nullptr;
  b.lists = ws->lists.as<std::uint32_t>();
  b.dist = ws->pool.as<std::uint32_t>();
  b.last_change = ws->last_change.as<std::uint32_t>();
  b.pending = nullptr;
  b.control = ws->control.as<DevControl>();
}

// Launches preparation, relaxation and finalisation for the bound batch.
int32_t run_resident_batch(bfn_workspace ws) {
  bfn_graph g = ws->graph;
  hipStream_t s = ws->stream;
  const bfn_run_options& o = ws->options;
  const bfn_tuning& t = ws->tuning;
  const DevBatch& b = ws->view;
  ws->completed = false;
  ws->have_paths = false;
  ws->paths_counted = false;
  BFN_HIP(hipEventRecord(ws->ev_start, s));
  const bool graph_mode = t.use_graph && o.control_mode != 0;
  unsigned long long* relax = t.count_relaxations ? ws->relax.as<unsigned long long>() : nullptr;
  auto finalize = [&]() -> hipError_t {
    hipError_t e = hipSuccess;
    if (relax) {
      e = fill_launch(reinterpret_cast<std::uint32_t*>(relax), 2ull * b.query_count, 0u, s);
      if (e == hipSuccess) e = relaxation_count_launch(g->view, b, relax, s);
    }
    if (e == hipSuccess) e = finalize_launch(g->view, b, relax, ws->status.ptr, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(ws->h_status.ptr, ws->status.ptr, sizeof(QueryStatusDev) * b.query_count,
                         hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(ws->h_control.ptr, ws->control.ptr, sizeof(DevControl),
                         hipMemcpyDeviceToHost, s);
    return e;
  };
  if (graph_mode) {
    // Preparation and finalisation run as ordinary launches; the graph holds
    // the relaxation rounds and the control read-back, so the relax events
    // bracket relaxation only.
    const std::uint32_t rounds = std::max(1u, t.graph_rounds);
    GraphKey key;
    key.block = ws->block;
    key.grid = ws->grid;
    key.rounds = rounds;
    key.lds = ws->lds_bytes;
    key.max_v = ws->max_v;
    key.work = ws->work_words;
    key.pool_words = ws->pool_words;
    key.g = g->view;
    key.b = b;
    if (!ws->graph_exec || !(key == ws->graph_key)) {
      if (ws->graph_exec) (void)hipGraphExecDestroy(ws->graph_exec);
      if (ws->graph_obj) (void)hipGraphDestroy(ws->graph_obj);
      ws->graph_exec = nullptr;
      ws->graph_obj = nullptr;
      BFN_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      hipError_t e = hipSuccess;
      for (std::uint32_t r = 0; r < rounds && e == hipSuccess; ++r)
        e = relax_launch(ws->block, false, false, ws->grid, ws->lds_bytes, ws->max_v,
                         ws->work_words, g->view, b, s);
      if (e == hipSuccess)
        e = hipMemcpyAsync(ws->h_control.ptr, ws->control.ptr, sizeof(DevControl),
                           hipMemcpyDeviceToHost, s);
      hipGraph_t captured = nullptr;
      const hipError_t end = hipStreamEndCapture(s, &captured);
      if (e != hipSuccess || end != hipSuccess)
        return fail(BFN_E_HIP, std::string("graph capture: ") +
                                   hipGetErrorString(e != hipSuccess ? e : end));
      ws->graph_obj = captured;
      BFN_HIP(hipGraphInstantiate(&ws->graph_exec, captured, nullptr, nullptr, 0));
      ws->graph_key = key;
    }
    BFN_HIP(prep_launch(g->view, b, ws->pool_words, s, !ws->slots_resolved));
    ws->slots_resolved = true;
    BFN_HIP(hipEventRecord(ws->ev_relax0, s));
    BFN_HIP(hipGraphLaunch(ws->graph_exec, s));
    ws->launches += 2ull * rounds;
    BFN_HIP(hipStreamSynchronize(s));
    ++ws->host_syncs;
    ++ws->controller_copies;
    DevControl* hc = ws->h_control.as<DevControl>();
    // Continue host-polled, rounds_per_chunk rounds at a time, when the
    // captured rounds did not suffice (push schedules vary by a round or two).
    const std::uint32_t chunk = std::max(1u, o.rounds_per_chunk);
    while (!hc->done) {
      for (std::uint32_t r = 0; r < chunk; ++r)
        BFN_HIP(relax_launch(ws->block, false, false, ws->grid, ws->lds_bytes, ws->max_v,
                             ws->work_words, g->view, b, s));
      ws->launches += 2ull * chunk;
      BFN_HIP(hipMemcpyAsync(ws->h_control.ptr, ws->control.ptr, sizeof(DevControl),
                             hipMemcpyDeviceToHost, s));
      BFN_HIP(hipStreamSynchronize(s));
      ++ws->host_checks;
      ++ws->controller_copies;
    }
    BFN_HIP(hipEventRecord(ws->ev_relax1, s));
    BFN_HIP(finalize());
    BFN_HIP(hipEventRecord(ws->ev_stop, s));
    return BFN_OK;
  }

  BFN_HIP(prep_launch(g->view, b, ws->pool_words, s, !ws->slots_resolved));
  ws->slots_resolved = true;
  BFN_HIP(hipEventRecord(ws->ev_relax0, s));
  if (o.control_mode == 0) {
    BFN_HIP(relax_launch(ws->block, true, true, ws->grid, ws->lds_bytes, ws->max_v,
                         ws->work_words, g->view, b, s));
    ++ws->launches;
  } else {
    const std::uint32_t chunk = o.control_mode == 2 ? 1u : o.rounds_per_chunk;
    DevControl* hc = ws->h_control.as<DevControl>();
    for (;;) {
      for (std::uint32_t r = 0; r < chunk; ++r)
        BFN_HIP(relax_launch(ws->block, false, false, ws->grid, ws->lds_bytes, ws->max_v,
                             ws->work_words, g->view, b, s));
      ws->launches += 2ull * chunk;
      BFN_HIP(hipMemcpyAsync(ws->h_control.ptr, ws->control.ptr, sizeof(DevControl),
                             hipMemcpyDeviceToHost, s));
      BFN_HIP(hipStreamSynchronize(s));
      ++ws->host_checks;
      ++ws->host_syncs;
      ++ws->controller_copies;
      if (hc->done) break;
    }
  }
  BFN_HIP(hipEventRecord(ws->ev_relax1, s));
  BFN_HIP(finalize());
  BFN_HIP(hipEventRecord(ws->ev_stop, s));
  return BFN_OK;
}

}  // namespace

int32_t bfn_batch_estimate(bfn_graph graph, const bfn_batch_desc* batch,
                           const bfn_run_options* options, uint64_t* bytes) {
  if (!graph || !bytes) return fail(BFN_E_INVALID_ARGUMENT, "null argument");
  if (int32_t rc = validate_options(options)) return rc;
  BatchPlanHost p;
  if (int32_t rc = plan_batch(graph, batch, options, p)) return rc;
  std::size_t seg[kMetaSegments], sizes[kMetaSegments];
  const std::size_t meta = meta_layout(p.Q, p.S, batch->source_offsets[p.Q],
                                       batch->target_offsets[p.Q], seg, sizes);
  *bytes = batch_bytes(p, meta);
  return BFN_OK;
}

int32_t bfn_batch_submit(bfn_workspace ws, bfn_graph graph, const bfn_batch_desc* d,
                         const bfn_run_options* o, const bfn_tuning* tuning) {
  if (!ws || !graph) return fail(BFN_E_INVALID_ARGUMENT, "null handle");
  // From here on the workspace holds no batch until this one is bound: a
  // refused submit never leaves the previous batch's results readable under
  // the new batch's description.
  ws->submitted = ws->completed = ws->have_paths = ws->paths_counted = false;
  ws->slots_resolved = false;
  if (ws->device != graph->device) return fail(BFN_E_INVALID_ARGUMENT, "graph is on another device");
  if (int32_t rc = validate_options(o)) return rc;
  bfn_tuning t;
  if (tuning) t = *tuning; else bfn_tuning_default(&t);
  if (t.tile_sweeps > 1 || t.boundary_mode > 1 || t.use_graph > 1 || t.min_waves > 8 ||
      t.claim_order > 1)
    return fail(BFN_E_INVALID_ARGUMENT,
                "invalid tuning (tile_sweeps, boundary_mode, use_graph, min_waves, claim_order)");
  if (!(t.push_window >= 0.0f) || !std::isfinite(t.push_window))
    return fail(BFN_E_INVALID_ARGUMENT, "push_window must be a finite non-negative number");
  if (t.claim_order && o->control_mode == 0)
    return fail(BFN_E_INVALID_ARGUMENT,
                "claim_order 1 compacts the queue between round launches: it needs per-round "
                "control (control_mode 1 or 2)");
  BatchPlanHost p;
  if (int32_t rc = plan_batch(graph, d, o, p)) return rc;
  BFN_HIP(hipSetDevice(ws->device->ordinal));
  // Nothing of a previous batch may still be running on this stream.
  BFN_HIP(hipStreamSynchronize(ws->stream));

  const std::uint32_t Q = p.Q, S = p.S;
  const std::uint32_t ns = d->source_offsets[Q], nt = d->target_offsets[Q];
  // Metadata segments (one H2D copy from pinned staging).
  std::size_t seg[kMetaSegments], sizes[kMetaSegments];
  const std::size_t meta_bytes = meta_layout(Q, S, ns, nt, seg, sizes);
  const std::uint64_t need = batch_bytes(p, meta_bytes);
  if (ws->budget && need > ws->budget)
    return fail(BFN_E_OUT_OF_MEMORY, "batch needs " + std::to_string(need) +
                                         " bytes of workspace, budget is " + std::to_string(ws->budget));

  // Launch shape, checked before the workspace changes: a refused batch
  // leaves the workspace without a batch (status, results and rerun then
  // report BFN_E_STATE), never a mix of the old and the new one.
  const DevGraph& gv = graph->view;
  const std::uint32_t max_v = gv.max_tile_vertices;
  const std::size_t lds_limit = ws->device->props.sharedMemPerBlock;
  // Push engines stage their boundary proposals after the snapshot when the
  // LDS holds them; otherwise they push with direct global atomics.
  const bool proposals = p.any_push && t.boundary_mode == 0 &&
                         lds_words(gv, max_v, gv.max_tile_vertices + gv.max_out_stage_floats) * 4ull <=
                             lds_limit;
  const std::uint32_t work_words = relax_work_words(gv, p.any_jacobi, p.any_push, proposals);
  const std::size_t lds_bytes = 4ull * lds_words(gv, max_v, work_words);
  if (lds_bytes > lds_limit)
    return fail(BFN_E_NOT_REPRESENTABLE, "tile neighbourhood exceeds the LDS budget (" +
                                             std::to_string(lds_bytes) + " > " +
                                             std::to_string(lds_limit) + " bytes)");
  int per_cu = 0;
  const std::uint32_t engines = !p.any_jacobi ? kEnginesPush : !p.any_push ? kEnginesPull : kEnginesMixed;
  BFN_HIP(relax_occupancy(engines, o->block_size, o->control_mode == 0, o->instrumentation != 0,
                          t.min_waves, lds_bytes, &per_cu));
  if (per_cu <= 0) return fail(BFN_E_NOT_REPRESENTABLE, "relaxation kernel cannot be resident");
  const std::uint32_t cus = static_cast<std::uint32_t>(ws->device->props.multiProcessorCount);
  std::uint32_t blocks = static_cast<std::uint32_t>(per_cu) * cus;
  if (o->grid_policy == 1) blocks = std::min(blocks, o->blocks_per_wgp * cus);

  BFN_HIP(ws->staging.reserve(meta_bytes));
  char* h = ws->staging.as<char>();
  std::memcpy(h + seg[0], d->tile_offsets, sizes[0]);
  std::memcpy(h + seg[1], d->selected_tiles, sizes[1]);
  std::memcpy(h + seg[2], d->source_offsets, sizes[2]);
  std::memcpy(h + seg[3], d->sources, sizes[3]);
  std::memcpy(h + seg[4], d->target_offsets, sizes[4]);
  std::memcpy(h + seg[5], d->targets, sizes[5]);
  std::memcpy(h + seg[6], p.engine.data(), sizes[6]);
  std::memcpy(h + seg[7], p.slot_query.data(), sizes[7]);
  std::memcpy(h + seg[8], p.slot_dist.data(), sizes[8]);
  std::memcpy(h + seg[9], p.slot_in_base.data(), sizes[9]);
  std::memcpy(h + seg[10], p.slot_out_base.data(), sizes[10]);
  std::memcpy(h + seg[11], p.query_slot.data(), sizes[11]);
  std::memcpy(h + seg[12], p.slot_tile.data(), sizes[12]);
  std::memcpy(h + seg[13], p.frame_base.data(), sizes[13]);

  BFN_HIP(ws->meta.reserve(meta_bytes));
  BFN_HIP(ws->slot_in.reserve(4ull * std::max<std::uint64_t>(p.in_total, 1)));
  BFN_HIP(ws->slot_out.reserve(4ull * std::max<std::uint64_t>(p.out_total, 1)));
  BFN_HIP(ws->slot_out_off.reserve(8ull * std::max<std::uint64_t>(p.out_total, 1)));
  BFN_HIP(ws->slot_desc.reserve(sizeof(SlotDesc) * std::max(S, 1u)));
  BFN_HIP(ws->stamp.reserve(4ull * std::max(S, 1u)));
  if (t.push_window > 0.0f) BFN_HIP(ws->prio.reserve(4ull * std::max(S, 1u)));
  BFN_HIP(ws->lists.reserve(12ull * std::max(S, 1u)));
  BFN_HIP(ws->pool.reserve(4ull * std::max<std::uint64_t>(p.pool_words, 1)));
  BFN_HIP(ws->last_change.reserve(4ull * Q));
  BFN_HIP(ws->control.reserve(sizeof(DevControl)));
  BFN_HIP(ws->status.reserve(sizeof(QueryStatusDev) * Q));
  BFN_HIP(ws->relax.reserve(8ull * Q));
  BFN_HIP(ws->h_status.reserve(sizeof(QueryStatusDev) * Q));
  BFN_HIP(ws->h_control.reserve(sizeof(DevControl)));
  BFN_HIP(hipMemcpyAsync(ws->meta.ptr, h, meta_bytes, hipMemcpyHostToDevice, ws->stream));

  ws->graph = graph;
  ws->Q = Q;
  ws->S = S;
  ws->pool_words = p.pool_words;
  ws->frame_words = p.frame_words;
  ws->options = *o;
  ws->tuning = t;
  ws->h_target_off.assign(d->target_offsets, d->target_offsets + Q + 1);
  ws->h_tile_off.assign(d->tile_offsets, d->tile_offsets + Q + 1);
  ws->h_tiles.assign(d->selected_tiles, d->selected_tiles + S);
  ws->h_engine = p.engine;
  ws->launches = ws->host_checks = ws->host_syncs = ws->controller_copies = 0;
  bind_view(ws, p, seg, o, &t);
  ws->frame_base_ptr = reinterpret_cast<const unsigned long long*>(ws->meta.as<char>() + seg[13]);
  if (!proposals) ws->view.boundary_mode = 1;
  ws->view.engines = engines;
  if (o->control_mode == 0) {
    // A persistent grid must be co-resident: checked on the device.
    int placed = 0;
    BFN_HIP(coresident_per_cu(ws->device->ordinal, per_cu, cus, lds_bytes, max_v, work_words,
                              graph->view, ws->view, ws->stream, o->block_size, &placed));
    if (placed <= 0)
      return fail(BFN_E_NOT_REPRESENTABLE, "no persistent grid of the relaxation kernel is co-resident");
    blocks = static_cast<std::uint32_t>(placed) * cus;
    if (o->grid_policy == 1) blocks = std::min(blocks, o->blocks_per_wgp * cus);
  }
  // A round holds each slot at most once, so workgroups beyond the slot
  // count never find work; a persistent grid of them would still wait at
  // every barrier (small batches: one query per launch).
  blocks = std::min(blocks, std::max(S, 1u));
  ws->block = o->block_size;
  ws->max_v = max_v;
  ws->work_words = work_words;
  ws->lds_bytes = lds_bytes;
  ws->grid = std::max(1u, blocks);
  ws->submitted = true;
  return run_resident_batch(ws);
}

int32_t bfn_batch_rerun(bfn_workspace ws) {
  if (!ws || !ws->submitted) return fail(BFN_E_STATE, "no batch has been submitted");
  BFN_HIP(hipSetDevice(ws->device->ordinal));
  BFN_HIP(hipStreamSynchronize(ws->stream));
  ws->launches = ws->host_checks = ws->host_syncs = ws->controller_copies = 0;
  return run_resident_batch(ws);
}

int32_t bfn_batch_wait(bfn_workspace ws) {
  if (!ws || !ws->submitted) return fail(BFN_E_STATE, "no batch has been submitted");
  BFN_HIP(hipSetDevice(ws->device->ordinal));
  BFN_HIP(hipStreamSynchronize(ws->stream));
  ++ws->host_syncs;
  ws->completed = true;
  return BFN_OK;
}

int32_t bfn_batch_status(bfn_workspace ws, bfn_run_status* status, bfn_query_status* per_query,
                         bfn_work_statistics* work) {
  if (!ws || !ws->completed) return fail(BFN_E_STATE, "batch has not completed");
  const DevControl* c = ws->h_control.as<DevControl>();
  const QueryStatusDev* qs = ws->h_status.as<QueryStatusDev>();
  if (per_query) {
    for (std::uint32_t q = 0; q < ws->Q; ++q) {
      per_query[q].rounds_completed = qs[q].rounds;
      per_query[q].relaxations = qs[q].relaxations;
      per_query[q].converged = qs[q].converged;
      per_query[q].reached_all_targets = qs[q].reached;
      per_query[q].bounding_box_miss = qs[q].miss;
      per_query[q].stop_reason = qs[q].stop_reason;
      per_query[q].error_bits = qs[q].error_bits;
      per_query[q].engine = qs[q].engine;
    }
  }
  if (status) {
    std::memset(status, 0, sizeof *status);
    bool all_converged = c->error_bits == 0;
    for (std::uint32_t q = 0; q < ws->Q; ++q) {
      all_converged &= qs[q].converged != 0;
      if (q < 32) {
        const std::uint32_t bit = 1u << q;
        status->valid_lane_mask |= bit;
        if (qs[q].reached) status->reached_target_mask |= bit;
        if (qs[q].miss) status->bounding_box_miss_mask |= bit;
        if (qs[q].converged) status->converged_lane_mask |= bit;
        else status->active_lane_mask |= bit;
      }
    }
    const bool jacobi = ws->options.engine == kJacobiPull;
    status->final_distance_slot = jacobi ? (c->rounds_executed & 1u) : 0u;
    status->converged = all_converged ? 1u : 0u;
    status->rounds_completed = c->rounds_executed;
    status->stop_reason = c->error_bits ? kStopDeviceFailure : c->stop_reason;
    status->error_bits = c->error_bits;
    if (!all_converged) {
      status->reached_target_mask = 0;
      status->bounding_box_miss_mask = 0;
    }
  }
  if (work) {
    std::memset(work, 0, sizeof *work);
    work->kernel_dispatches = ws->launches;
    work->host_checks = ws->host_checks;
    work->host_synchronizations = ws->host_syncs;
    work->controller_copies = ws->controller_copies;
    work->maximum_queue_size = c->max_list;
    work->atomic_attempts = c->atomic_attempts;
    work->successful_atomic_updates = c->atomic_successes;
    work->active_vertices = c->items_processed;
    std::uint64_t relax = 0;
    for (std::uint32_t q = 0; q < ws->Q; ++q) relax += qs[q].relaxations;
    work->edges_examined = relax;
  }
  return BFN_OK;
}

namespace {
int32_t ensure_scratch_v(bfn_workspace ws) {
  BFN_HIP(ws->scratch_v.reserve(4ull * std::max(ws->graph->V, 1u)));
  return BFN_OK;
}
}  // namespace

int32_t bfn_batch_distances(bfn_workspace ws, uint32_t query, float* out) {
  if (!ws || !ws->completed || !out) return fail(BFN_E_STATE, "batch has not completed");
  if (query >= ws->Q) return fail(BFN_E_INVALID_ARGUMENT, "query index out of range");
  if (int32_t rc = ensure_scratch_v(ws)) return rc;
  const DevControl* c = ws->h_control.as<DevControl>();
  const std::uint32_t copy = c->rounds_executed & 1u;
  const std::uint32_t slots = ws->h_tile_off[query + 1] - ws->h_tile_off[query];
  hipStream_t s = ws->stream;
  BFN_HIP(fill_launch(ws->scratch_v.as<std::uint32_t>(), ws->graph->V, kInfBits, s));
  BFN_HIP(scatter_launch(ws->graph->view, ws->view, query, slots, copy,
                         ws->scratch_v.as<std::uint32_t>(), s));
  BFN_HIP(hipMemcpyAsync(out, ws->scratch_v.ptr, 4ull * ws->graph->V, hipMemcpyDeviceToHost, s));
  BFN_HIP(hipStreamSynchronize(s));
  return BFN_OK;
}

int32_t bfn_batch_region_distances(bfn_workspace ws, uint32_t query, float* out,
                                   uint64_t capacity, uint64_t* count) {
  if (!ws || !ws->completed) return fail(BFN_E_STATE, "batch has not completed");
  if (query >= ws->Q) return fail(BFN_E_INVALID_ARGUMENT, "query index out of range");
  if (!count) return fail(BFN_E_INVALID_ARGUMENT, "null count");
  const std::uint32_t p0 = ws->h_tile_off[query], p1 = ws->h_tile_off[query + 1];
  const auto& toff = ws->graph->h_tile_off;
  std::vector<unsigned long long> offsets(p1 - p0);
  std::uint64_t n = 0;
  for (std::uint32_t i = p0; i < p1; ++i) {
    offsets[i - p0] = n;
    n += toff[ws->h_tiles[i] + 1] - toff[ws->h_tiles[i]];
  }
  *count = n;
  if (!out) return BFN_OK;
  if (capacity < n) return fail(BFN_E_INVALID_ARGUMENT, "output capacity too small");
  if (int32_t rc = ensure_scratch_v(ws)) return rc;
  BFN_HIP(ws->region_off.reserve(8ull * std::max<std::size_t>(offsets.size(), 1)));
  hipStream_t s = ws->stream;
  const DevControl* c = ws->h_control.as<DevControl>();
  BFN_HIP(hipMemcpyAsync(ws->region_off.ptr, offsets.data(), 8ull * offsets.size(),
                         hipMemcpyHostToDevice, s));
  BFN_HIP(gather_region_launch(ws->graph->view, ws->view, query, p1 - p0,
                               c->rounds_executed & 1u, ws->region_off.as<unsigned long long>(),
                               ws->scratch_v.as<std::uint32_t>(), s));
  if (n) BFN_HIP(hipMemcpyAsync(out, ws->scratch_v.ptr, 4ull * n, hipMemcpyDeviceToHost, s));
  BFN_HIP(hipStreamSynchronize(s));
  return BFN_OK;
}

int32_t bfn_batch_predecessors(bfn_workspace ws, uint32_t query, uint32_t* out) {
  if (!ws || !ws->completed || !out) return fail(BFN_E_STATE, "batch has not completed");
  if (query >= ws->Q) return fail(BFN_E_INVALID_ARGUMENT, "query index out of range");
  if (int32_t rc = ensure_scratch_v(ws)) return rc;
  const DevControl* c = ws->h_control.as<DevControl>();
  const std::uint32_t copy = c->rounds_executed & 1u;
  const std::uint32_t slots = ws->h_tile_off[query + 1] - ws->h_tile_off[query];
  hipStream_t s = ws->stream;
  BFN_HIP(fill_launch(ws->scratch_v.as<std::uint32_t>(), ws->graph->V, kNone, s));
  BFN_HIP(predecessors_launch(ws->graph->view, ws->view, query, slots, copy,
                              ws->scratch_v.as<std::uint32_t>(), s));
  BFN_HIP(hipMemcpyAsync(out, ws->scratch_v.ptr, 4ull * ws->graph->V, hipMemcpyDeviceToHost, s));
  BFN_HIP(hipStreamSynchronize(s));
  return BFN_OK;
}

int32_t bfn_batch_paths(bfn_workspace ws, bfn_target_summary* summaries, uint64_t* path_offsets,
                        uint32_t* vertices, float* labels, uint32_t* edge_ids,
                        uint64_t arena_capacity, uint64_t* vertex_total) {
  if (!ws || !ws->completed) return fail(BFN_E_STATE, "batch has not completed");
  if (!summaries || !path_offsets || !vertex_total)
    return fail(BFN_E_INVALID_ARGUMENT, "null summaries, offsets or total");
  hipStream_t s = ws->stream;
  const std::uint32_t NT = ws->h_target_off[ws->Q];
  const DevControl* c = ws->h_control.as<DevControl>();
  const QueryStatusDev* qs = ws->h_status.as<QueryStatusDev>();
  BFN_HIP(ws->paths_summary.reserve(28ull * NT));
  BFN_HIP(ws->paths_len.reserve(8ull * NT));
  BFN_HIP(ws->paths_off.reserve(8ull * (NT + 1)));
  BFN_HIP(ws->frames.reserve(12ull * std::max<std::uint64_t>(ws->frame_words, 1)));
  BFN_HIP(ws->onpath.reserve(4ull * std::max<std::uint64_t>(ws->pool_words, 1)));
  BFN_HIP(ws->clean.reserve(4ull * ws->Q));
  BFN_HIP(ws->h_paths.reserve(28ull * NT + 8ull * (NT + 1) + 4ull * ws->Q));
  char* hp = ws->h_paths.as<char>();
  PathArgs a{};
  a.g = ws->graph->view;
  a.b = ws->view;
  a.jacobi_copy = c->rounds_executed & 1u;
  a.pass = 1;
  a.target_total = NT;
  a.query_clean = ws->clean.as<std::uint32_t>();
  a.frame_base = ws->frame_base_ptr;
  a.frame_vertex = ws->frames.as<std::uint32_t>();
  a.frame_last = a.frame_vertex + ws->frame_words;
  a.frame_via = a.frame_last + ws->frame_words;
  a.onpath = ws->onpath.as<std::uint32_t>();
  a.summaries = ws->paths_summary.as<std::uint32_t>();
  a.lengths = ws->paths_len.as<unsigned long long>();
  a.offsets = ws->paths_off.as<unsigned long long>();
  // Every pass of the path kernel marks the vertices on its walks with its
  // own stamp range, so the marks that earlier passes (and earlier batches)
  // left on found paths never match a later stamp: the marks are cleared
  // only when the buffer is new or larger than what was cleared, or when the
  // 32-bit stamps would wrap.
  auto take_stamps = [&]() -> hipError_t {
    if (ws->onpath_cleared != ws->onpath.ptr || ws->onpath_cleared_words < ws->pool_words ||
        ws->path_stamp > 0xFFFFFFFFu - NT) {
      const hipError_t e = fill_launch(ws->onpath.as<std::uint32_t>(), ws->pool_words, 0u, s);
      if (e != hipSuccess) return e;
      ws->onpath_cleared = ws->onpath.ptr;
      ws->onpath_cleared_words = ws->pool_words;
      ws->path_stamp = 0;
    }
    a.stamp_base = ws->path_stamp;
    ws->path_stamp += NT;
    return hipSuccess;
  };
  // Pass 1 (summaries, lengths, offsets) runs once per completed batch: a
  // second call, the one that brings the arenas, reuses its results.
  if (!ws->paths_counted) {
    std::uint32_t* hclean = reinterpret_cast<std::uint32_t*>(hp + 28ull * NT + 8ull * (NT + 1));
    for (std::uint32_t q = 0; q < ws->Q; ++q)
      hclean[q] = (qs[q].converged && qs[q].error_bits == 0) ? 1u : 0u;
    BFN_HIP(hipMemcpyAsync(ws->clean.ptr, hclean, 4ull * ws->Q, hipMemcpyHostToDevice, s));
    BFN_HIP(take_stamps());
    BFN_HIP(paths_launch(a, ws->Q, s));
    BFN_HIP(scan_u64_launch(a.lengths, ws->paths_off.as<unsigned long long>(), NT, s));
    BFN_HIP(hipMemcpyAsync(hp, ws->paths_summary.ptr, 28ull * NT, hipMemcpyDeviceToHost, s));
    BFN_HIP(hipMemcpyAsync(hp + 28ull * NT, ws->paths_off.ptr, 8ull * (NT + 1), hipMemcpyDeviceToHost, s));
    BFN_HIP(hipStreamSynchronize(s));
    ws->paths_counted = true;
  }
  std::memcpy(summaries, hp, 28ull * NT);
  std::memcpy(path_offsets, hp + 28ull * NT, 8ull * (NT + 1));
  const std::uint64_t total = path_offsets[NT];
  *vertex_total = total;
  if (!vertices && !labels && !edge_ids) return BFN_OK;
  if (!vertices || !labels || !edge_ids) return fail(BFN_E_INVALID_ARGUMENT, "arenas must all be given");
  if (arena_capacity < total) return fail(BFN_E_INVALID_ARGUMENT, "arena capacity too small");
  BFN_HIP(ws->arena_v.reserve(4ull * std::max<std::uint64_t>(total, 1)));
  BFN_HIP(ws->arena_l.reserve(4ull * std::max<std::uint64_t>(total, 1)));
  BFN_HIP(ws->arena_e.reserve(4ull * std::max<std::uint64_t>(total, 1)));
  BFN_HIP(take_stamps());
  a.pass = 2;
  a.arena_vertices = ws->arena_v.as<std::uint32_t>();
  a.arena_labels = ws->arena_l.as<std::uint32_t>();
  a.arena_edges = ws->arena_e.as<std::uint32_t>();
  BFN_HIP(fill_launch(a.arena_edges, total, kNone, s));
  BFN_HIP(paths_launch(a, ws->Q, s));
  if (total) {
    BFN_HIP(hipMemcpyAsync(vertices, ws->arena_v.ptr, 4ull * total, hipMemcpyDeviceToHost, s));
    BFN_HIP(hipMemcpyAsync(labels, ws->arena_l.ptr, 4ull * total, hipMemcpyDeviceToHost, s));
    BFN_HIP(hipMemcpyAsync(edge_ids, ws->arena_e.ptr, 4ull * total, hipMemcpyDeviceToHost, s));
  }
  BFN_HIP(hipStreamSynchronize(s));
  ws->have_paths = true;
  ws->arena_vertices = total;
  return BFN_OK;
}

int32_t bfn_batch_occupancy(bfn_workspace ws, int32_t* out) {
  if (!ws || !ws->have_paths) return fail(BFN_E_STATE, "no compact paths in this workspace");
  hipStream_t s = ws->stream;
  const std::uint32_t V = ws->graph->V;
  BFN_HIP(ws->occupancy.reserve(4ull * std::max(V, 1u)));
  BFN_HIP(fill_launch(ws->occupancy.as<std::uint32_t>(), V, 0u, s));
  BFN_HIP(occupancy_launch(ws->arena_v.as<std::uint32_t>(), ws->arena_vertices,
                           ws->occupancy.as<int>(), s));
  if (out) BFN_HIP(hipMemcpyAsync(out, ws->occupancy.ptr, 4ull * V, hipMemcpyDeviceToHost, s));
  BFN_HIP(hipStreamSynchronize(s));
  return BFN_OK;
}

int32_t bfn_batch_counters(bfn_workspace ws, bfn_engine_counters* out) {
  if (!ws || !ws->completed || !out) return fail(BFN_E_STATE, "batch has not completed");
  const DevControl* c = ws->h_control.as<DevControl>();
  std::memset(out, 0, sizeof *out);
  out->tile_visits = c->items_processed;
  out->edges_relaxed = c->edges_relaxed;
  out->algorithmic_bytes = c->algorithmic_bytes;
  out->atomic_attempts = c->atomic_attempts;
  out->atomic_successes = c->atomic_successes;
  out->maximum_list = c->max_list;
  out->rounds_executed = c->rounds_executed;
  out->closure_sweeps = c->closure_sweeps;
  for (int p = 0; p < kPhaseCount; ++p) out->phase_cycles[p] = c->phase_cycles[p];
  out->candidate_evaluations = c->candidates;
  return BFN_OK;
}

int32_t bfn_batch_timing(bfn_workspace ws, float* relax_ms, float* total_ms, uint64_t* launches) {
  if (!ws || !ws->completed) return fail(BFN_E_STATE, "batch has not completed");
  if (relax_ms) BFN_HIP(hipEventElapsedTime(relax_ms, ws->ev_relax0, ws->ev_relax1));
  if (total_ms) BFN_HIP(hipEventElapsedTime(total_ms, ws->ev_start, ws->ev_stop));
  if (launches) *launches = ws->launches;
  return BFN_OK;
}


namespace bfn {
int32_t set_error(int32_t code, const std::string& message) { return fail(code, message); }
int device_ordinal(bfn_device device) { return device->ordinal; }
std::size_t device_lds_limit(bfn_device device) { return device->props.sharedMemPerBlock; }
std::uint32_t device_cu_count(bfn_device device) {
  return static_cast<std::uint32_t>(device->props.multiProcessorCount);
}
void graph_view(bfn_graph graph, DevGraph* view, std::uint32_t* max_in, std::uint32_t* max_out) {
  *view = graph->view;
  *max_in = graph->info.max_in_neighbor_tiles;
  *max_out = graph->info.max_out_neighbor_tiles;
}
GraphHost graph_host(bfn_graph graph) {
  GraphHost h;
  h.view = graph->view;
  h.tile_offsets = graph->h_tile_off.data();
  h.in_offsets = graph->h_in_off.data();
  h.out_offsets = graph->h_out_off.data();
  h.stream = graph->stream;
  h.ordinal = graph->device->ordinal;
  h.cu_count = static_cast<std::uint32_t>(graph->device->props.multiProcessorCount);
  h.lds_limit = graph->device->props.sharedMemPerBlock;
  return h;
}
}  // namespace bfn
