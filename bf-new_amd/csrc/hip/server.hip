// server.hip -- continuous-batching route server over a resident graph.
//
// One host thread owns one stream and a fixed-capacity batch: query records
// indexed by a query slot, tile slots in one contiguous arena (a query's
// selected tiles occupy one contiguous slot range), label and path scratch
// sized by the arena.  The loop repeats:
//   admit    -- requests waiting in the inbox take a free query slot and a
//               free slot range while the arena has room (admission by
//               workspace memory); one kernel writes their slot records,
//               resolves neighbour slots, resets labels and queues their
//               source tiles into the round that runs next;
//   relax    -- rounds_per_chunk rounds of the shared round kernel over every
//               admitted query: queries admitted now run alongside queries
//               that are half way;
//   harvest  -- a query whose queued-work counter reached zero has converged;
//               its target label (and path length) is copied out, its slots
//               return to the arena and a response is posted.
// Requests therefore join the running batch between chunks, and their work
// overlaps on the device with everyone else's.
#include "bfnew_amd.h"

#include "bfn_launch.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

using namespace bfn;

namespace bfn {

struct AdmitRecord {
  std::uint32_t query;
  std::uint32_t slot_begin;
  std::uint32_t slot_count;
  std::uint32_t source;
  std::uint32_t target;
  std::uint32_t engine;
  std::uint32_t tiles_offset;
  std::uint32_t pad;
};

// One block per admitted query.
__global__ void admit_kernel(DevGraph g, DevBatch b, const AdmitRecord* records,
                             const std::uint32_t* tiles, std::uint32_t* onpath,
                             unsigned long long* frame_base, std::uint32_t* active,
                             std::uint32_t slab_words, std::uint32_t kin, std::uint32_t kout) {
  const AdmitRecord r = records[blockIdx.x];
  const std::uint32_t tid = threadIdx.x;
  const std::uint32_t q = r.query;
  for (std::uint32_t i = tid; i < r.slot_count; i += blockDim.x) {
    const std::uint32_t s = r.slot_begin + i;
    const_cast<std::uint32_t*>(b.tiles)[s] = tiles[r.tiles_offset + i];
    const_cast<std::uint32_t*>(b.slot_query)[s] = q;
    b.stamp[s] = 0;
  }
  if (tid == 0) {
    const_cast<std::uint32_t*>(b.tile_begin)[q] = r.slot_begin;
    const_cast<std::uint32_t*>(b.tile_end)[q] = r.slot_begin + r.slot_count;
    const_cast<std::uint32_t*>(b.sources)[q] = r.source;
    const_cast<std::uint32_t*>(b.targets)[q] = r.target;
    const_cast<std::uint32_t*>(b.engine)[q] = r.engine;
    b.last_change[q] = 0;
    b.pending[q] = 0;
    frame_base[q] = static_cast<unsigned long long>(r.slot_begin) * (slab_words / 2);
    active[q] = 1;
  }
  __syncthreads();
  const std::uint32_t qb = r.slot_begin, qe = r.slot_begin + r.slot_count;
  for (std::uint32_t i = tid; i < r.slot_count; i += blockDim.x) {
    const std::uint32_t s = r.slot_begin + i;
    const std::uint32_t T = b.tiles[s];
    const std::uint32_t in0 = g.in_nbr_offsets[T], in1 = g.in_nbr_offsets[T + 1];
    for (std::uint32_t k = in0; k < in1; ++k)
      b.slot_in[static_cast<std::size_t>(s) * kin + (k - in0)] = find_slot(b, q, g.in_nbr_tiles[k]);
    const std::uint32_t o0 = g.out_nbr_offsets[T], o1 = g.out_nbr_offsets[T + 1];
    for (std::uint32_t k = o0; k < o1; ++k)
      b.slot_out[static_cast<std::size_t>(s) * kout + (k - o0)] = find_slot(b, q, g.out_nbr_tiles[k]);
  }
  const unsigned long long w0 = static_cast<unsigned long long>(qb) * slab_words;
  const unsigned long long w1 = static_cast<unsigned long long>(qe) * slab_words;
  for (unsigned long long w = w0 + tid; w < w1; w += blockDim.x) {
    b.dist[w] = kInfBits;
    onpath[w] = 0;
  }
  __syncthreads();
  __threadfence();
  if (tid == 0) {
    const std::uint32_t round = load_agent(&b.control->round);
    const std::uint32_t u = r.source;
    const std::uint32_t T = g.owner_tiles[u];
    const std::uint32_t s = find_slot(b, q, T);
    if (s == kNone) {
      atomicOr(&b.control->error_bits, 4u);
      return;
    }
    const std::uint32_t v0 = g.tile_vertex_offsets[T];
    const std::uint32_t npt = g.tile_vertex_offsets[T + 1] - v0;
    std::uint32_t* column = b.dist + b.slot_dist[s];
    column[u - v0] = 0u;
    const bool jacobi = r.engine == kJacobiPull;
    if (jacobi) column[npt + (u - v0)] = 0u;
    enqueue(b, s, round);
    if (jacobi) {
      const std::uint32_t KO = g.out_nbr_offsets[T + 1] - g.out_nbr_offsets[T];
      for (std::uint32_t k = 0; k < KO; ++k) {
        const std::uint32_t ns = b.slot_out[static_cast<std::size_t>(s) * kout + k];
        if (ns != kNone) enqueue(b, ns, round);
      }
    }
    store_agent(&b.control->done, 0u);
  }
}

// Reports converged queries: pending work reached zero.
__global__ void harvest_kernel(DevGraph g, DevBatch b, std::uint32_t capacity,
                               std::uint32_t* active, std::uint32_t* done_list,
                               std::uint32_t* result_bits) {
  const std::uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= capacity || !active[q]) return;
  if (load_agent(&b.pending[q]) != 0) return;
  active[q] = 0;
  const std::uint32_t pos = atomicAdd(&done_list[0], 1u);
  done_list[1 + pos] = q;
  const std::uint32_t t = b.targets[q];
  const std::uint32_t T = g.owner_tiles[t];
  const std::uint32_t s = find_slot(b, q, T);
  std::uint32_t label = kInfBits;
  if (s != kNone) {
    const std::uint32_t v0 = g.tile_vertex_offsets[T];
    const std::uint32_t npt = g.tile_vertex_offsets[T + 1] - v0;
    const std::uint32_t col = b.engine[q] == kJacobiPull ? (b.control->round & 1u) : 0u;
    label = b.dist[b.slot_dist[s] + static_cast<unsigned long long>(col) * npt + (t - v0)];
  }
  result_bits[q] = label;
}

__global__ void server_init_kernel(DevBatch b, std::uint32_t capacity, std::uint32_t slots,
                                   std::uint32_t slab_words, std::uint32_t kin, std::uint32_t kout,
                                   std::uint32_t* active, std::uint32_t* src_off,
                                   std::uint32_t* tgt_off, std::uint32_t* clean) {
  const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= capacity) {
    src_off[i] = i;
    tgt_off[i] = i;
  }
  if (i < capacity) {
    active[i] = 0;
    clean[i] = 1;
  }
  if (i < slots) {
    unsigned long long* sd = const_cast<unsigned long long*>(b.slot_dist);
    sd[i] = static_cast<unsigned long long>(i) * slab_words;
    const_cast<std::uint32_t*>(b.slot_in_base)[i] = i * kin;
    const_cast<std::uint32_t*>(b.slot_out_base)[i] = i * kout;
    b.stamp[i] = 0;
  }
  if (i == 0) {
    DevControl* c = b.control;
    for (int k = 0; k < 3; ++k) {
      c->list_count[k] = 0;
      c->claim[k] = 0;
    }
    c->round = 0;
    c->done = 1;
    c->stop_reason = 0;
    c->error_bits = 0;
    c->rounds_executed = 0;
  }
}

}  // namespace bfn

namespace {

int32_t fail(int32_t code, const std::string& m) { return bfn::set_error(code, m); }

double now_ns() {
  return std::chrono::duration<double, std::nano>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct Pending {
  bfn_request request;
  std::vector<std::uint32_t> tiles;
  double submit_ns;
};

struct Active {
  std::uint64_t request_id;
  std::uint32_t query;
  std::uint32_t slot_begin, slot_count;
  double submit_ns;
};

// A response with its compact path (source first; empty when unreached).
struct Delivered {
  bfn_response response;
  std::vector<std::uint32_t> vertices;
  std::vector<std::uint32_t> edges;
};

}  // namespace

struct bfn_server_s {
  bfn_device device = nullptr;
  bfn_graph graph = nullptr;
  bfn_server_options options{};
  hipStream_t stream = nullptr;
  DevGraph gview{};
  DevBatch view{};
  std::uint32_t Qcap = 0, Scap = 0, slab_words = 0, kin = 0, kout = 0, block = 256;
  std::uint32_t grid = 0;
  std::size_t lds_bytes = 0;
  std::uint32_t max_v = 0, work_words = 0;
  std::vector<void*> allocations;
  std::uint32_t *d_tiles = nullptr, *d_slot_query = nullptr, *d_in_base = nullptr,
                *d_out_base = nullptr, *d_slot_in = nullptr, *d_slot_out = nullptr,
                *d_stamp = nullptr, *d_lists = nullptr, *d_pool = nullptr, *d_onpath = nullptr,
                *d_frames = nullptr, *d_begin = nullptr, *d_end = nullptr, *d_src_off = nullptr,
                *d_src = nullptr, *d_tgt_off = nullptr, *d_tgt = nullptr, *d_engine = nullptr,
                *d_last = nullptr, *d_pending = nullptr, *d_active = nullptr, *d_done = nullptr,
                *d_result = nullptr, *d_clean = nullptr, *d_summary = nullptr,
                *d_upload_tiles = nullptr;
  unsigned long long *d_slot_dist = nullptr, *d_frame_base = nullptr, *d_lengths = nullptr,
                     *d_path_offsets = nullptr;
  // Compact-path arenas of one step's converged queries (grown on demand).
  std::uint32_t *d_arena_v = nullptr, *d_arena_l = nullptr, *d_arena_e = nullptr;
  std::uint32_t *h_arena_v = nullptr, *h_arena_e = nullptr;
  unsigned long long* h_path_offsets = nullptr;
  std::uint64_t arena_capacity = 0;
  DevControl* d_control = nullptr;
  AdmitRecord* d_records = nullptr;
  // pinned staging
  AdmitRecord* h_records = nullptr;
  std::uint32_t* h_tiles = nullptr;
  std::uint32_t* h_done = nullptr;
  std::uint32_t* h_result = nullptr;
  std::uint32_t* h_summary = nullptr;
  std::uint32_t upload_tile_capacity = 0;
  // host state
  std::map<std::uint32_t, std::uint32_t> free_ranges;  // begin -> length
  std::vector<std::uint32_t> free_queries;
  std::vector<Active> active;
  std::deque<Pending> inbox;
  std::deque<Delivered> outbox;
  std::mutex mutex;
  std::condition_variable wake, delivered;
  bool stopping = false;
  std::string fatal;
  std::thread worker;
  std::uint64_t launches = 0, admitted = 0, completed = 0;

  bool take_range(std::uint32_t count, std::uint32_t* begin) {
    for (auto it = free_ranges.begin(); it != free_ranges.end(); ++it) {
      if (it->second >= count) {
        *begin = it->first;
        const std::uint32_t rest = it->second - count;
        const std::uint32_t next = it->first + count;
        free_ranges.erase(it);
        if (rest) free_ranges[next] = rest;
        return true;
      }
    }
    return false;
  }
  void give_range(std::uint32_t begin, std::uint32_t count) {
    auto it = free_ranges.emplace(begin, count).first;
    auto next = std::next(it);
    if (next != free_ranges.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_ranges.erase(next);
    }
    if (it != free_ranges.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_ranges.erase(it);
      }
    }
  }
  void run();
  hipError_t step(std::uint32_t admit_count, std::uint32_t tile_count);
  hipError_t reserve_arena(std::uint64_t vertices);
  void free_arena();
};

void bfn_server_s::free_arena() {
  for (std::uint32_t* p : {d_arena_v, d_arena_l, d_arena_e})
    if (p) (void)hipFree(p);
  for (std::uint32_t* p : {h_arena_v, h_arena_e})
    if (p) (void)hipHostFree(p);
  d_arena_v = d_arena_l = d_arena_e = h_arena_v = h_arena_e = nullptr;
  arena_capacity = 0;
}

hipError_t bfn_server_s::reserve_arena(std::uint64_t vertices) {
  if (vertices <= arena_capacity) return hipSuccess;
  free_arena();
  const std::uint64_t n = std::max<std::uint64_t>(vertices, 2 * static_cast<std::uint64_t>(max_v));
  hipError_t e = hipSuccess;
  for (std::uint32_t** p : {&d_arena_v, &d_arena_l, &d_arena_e})
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(p), 4 * n);
  for (std::uint32_t** p : {&h_arena_v, &h_arena_e})
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void**>(p), 4 * n, hipHostMallocDefault);
  if (e != hipSuccess) {
    free_arena();
    return e;
  }
  arena_capacity = n;
  return hipSuccess;
}

hipError_t bfn_server_s::step(std::uint32_t admit_count, std::uint32_t tile_count) {
  hipStream_t s = stream;
  hipError_t e = hipSuccess;
  if (admit_count) {
    if ((e = hipMemcpyAsync(d_records, h_records, sizeof(AdmitRecord) * admit_count,
                            hipMemcpyHostToDevice, s)) != hipSuccess)
      return e;
    if ((e = hipMemcpyAsync(d_upload_tiles, h_tiles, 4ull * tile_count, hipMemcpyHostToDevice,
                            s)) != hipSuccess)
      return e;
    hipLaunchKernelGGL(admit_kernel, dim3(admit_count), dim3(256), 0, s, gview, view, d_records,
                       d_upload_tiles, d_onpath, d_frame_base, d_active, slab_words, kin, kout);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  // No more workgroups than live slots (a round visits each at most once);
  // `active` belongs to the worker thread that runs this step.
  std::uint64_t live = 0;
  for (const Active& a : active) live += a.slot_count;
  const std::uint32_t chunk_grid =
      static_cast<std::uint32_t>(std::min<std::uint64_t>(grid, std::max<std::uint64_t>(live, 1)));
  for (std::uint32_t r = 0; r < options.rounds_per_chunk; ++r) {
    if ((e = relax_launch(block, false, false, chunk_grid, lds_bytes, max_v, work_words, gview, view,
                          s)) != hipSuccess)
      return e;
    launches += 2;
  }
  if ((e = fill_launch(d_done, 1, 0u, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(harvest_kernel, dim3((Qcap + 255) / 256), dim3(256), 0, s, gview, view, Qcap,
                     d_active, d_done, d_result);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(h_done, d_done, 4ull * (Qcap + 1), hipMemcpyDeviceToHost, s)) != hipSuccess)
    return e;
  if ((e = hipMemcpyAsync(h_result, d_result, 4ull * Qcap, hipMemcpyDeviceToHost, s)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  const std::uint32_t ndone = h_done[0];
  if (ndone && options.want_paths) {
    // compact paths of the converged queries (their labels stay put until
    // their slots are handed out again)
    PathArgs a{};
    a.g = gview;
    a.b = view;
    a.jacobi_copy = 0;  // resolved per query below through the engine column
    a.pass = 1;
    a.target_total = Qcap;
    a.stamp_base = 0;  // pass 2 marks with [Qcap + 1, 2 Qcap]
    a.query_clean = d_clean;
    a.query_list = d_done + 1;
    a.frame_base = d_frame_base;
    a.frame_vertex = d_frames;
    a.frame_last = d_frames + static_cast<std::size_t>(Scap) * max_v;
    a.frame_via = a.frame_last + static_cast<std::size_t>(Scap) * max_v;
    a.onpath = d_onpath;
    a.summaries = d_summary;
    a.lengths = d_lengths;
    a.jacobi_copy_from_control = 1;
    if ((e = paths_launch(a, ndone, s)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(h_summary, d_summary, 28ull * Qcap, hipMemcpyDeviceToHost, s)) !=
        hipSuccess)
      return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    // Second pass: the complete paths themselves, packed at host-computed
    // offsets, then copied back with the summaries' step.
    std::uint64_t total = 0;
    for (std::uint32_t i = 0; i < ndone; ++i) {
      const std::uint32_t q = h_done[1 + i];
      const bool complete = h_summary[7ull * q + 5] == 0u && h_summary[7ull * q + 6] == 1u;
      h_path_offsets[q] = total;
      if (complete) total += h_summary[7ull * q + 3] + 1ull;
    }
    if (total) {
      if ((e = reserve_arena(total)) != hipSuccess) return e;
      if ((e = hipMemcpyAsync(d_path_offsets, h_path_offsets, 8ull * Qcap, hipMemcpyHostToDevice,
                              s)) != hipSuccess)
        return e;
      a.pass = 2;
      a.stamp_base = Qcap;
      a.offsets = d_path_offsets;
      a.arena_vertices = d_arena_v;
      a.arena_labels = d_arena_l;
      a.arena_edges = d_arena_e;
      if ((e = paths_launch(a, ndone, s)) != hipSuccess) return e;
      if ((e = hipMemcpyAsync(h_arena_v, d_arena_v, 4 * total, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return e;
      if ((e = hipMemcpyAsync(h_arena_e, d_arena_e, 4 * total, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return e;
      if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

void bfn_server_s::run() {
  (void)hipSetDevice(device_ordinal(device));
  for (;;) {
    std::uint32_t admit_count = 0, tile_count = 0;
    {
      std::unique_lock<std::mutex> lock(mutex);
      if (active.empty()) wake.wait(lock, [&] { return stopping || !inbox.empty(); });
      if (stopping) return;
      // Admission by free query records and free contiguous slot ranges.
      while (!inbox.empty() && !free_queries.empty()) {
        Pending& p = inbox.front();
        const auto count = static_cast<std::uint32_t>(p.tiles.size());
        if (tile_count + count > upload_tile_capacity) break;
        std::uint32_t begin = 0;
        if (!take_range(count, &begin)) break;
        const std::uint32_t q = free_queries.back();
        free_queries.pop_back();
        active.push_back(Active{p.request.request_id, q, begin, count, p.submit_ns});
        AdmitRecord& rec = h_records[admit_count++];
        rec = AdmitRecord{q, begin, count, p.request.source, p.request.target, options.engine,
                          tile_count, 0};
        std::memcpy(h_tiles + tile_count, p.tiles.data(), 4ull * count);
        tile_count += count;
        inbox.pop_front();
      }
    }
    const hipError_t e = step(admit_count, tile_count);
    if (e != hipSuccess) {
      std::lock_guard<std::mutex> lock(mutex);
      fatal = std::string("server step: ") + hipGetErrorString(e);
      stopping = true;
      delivered.notify_all();
      return;
    }
    const std::uint32_t ndone = h_done[0];
    std::lock_guard<std::mutex> lock(mutex);
    admitted += admit_count;
    if (!ndone) continue;
    const double t = now_ns();
    for (std::uint32_t i = 0; i < ndone; ++i) {
      const std::uint32_t q = h_done[1 + i];
      auto it = std::find_if(active.begin(), active.end(),
                             [&](const Active& a) { return a.query == q; });
      if (it == active.end()) continue;
      Delivered d;
      bfn_response& r = d.response;
      r = bfn_response{};
      r.request_id = it->request_id;
      std::memcpy(&r.distance, &h_result[q], 4);
      r.reached = h_result[q] != kInfBits;
      r.path_length = options.want_paths ? h_summary[7ull * q + 3] : 0;
      r.submit_ns = it->submit_ns;
      r.complete_ns = t;
      if (options.want_paths && h_summary[7ull * q + 5] == 0u && h_summary[7ull * q + 6] == 1u) {
        const std::uint64_t off = h_path_offsets[q];
        const std::uint32_t len = h_summary[7ull * q + 3];
        d.vertices.assign(h_arena_v + off, h_arena_v + off + len + 1);
        d.edges.assign(h_arena_e + off, h_arena_e + off + len);
      }
      outbox.push_back(std::move(d));
      give_range(it->slot_begin, it->slot_count);
      free_queries.push_back(q);
      active.erase(it);
      ++completed;
    }
    delivered.notify_all();
  }
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------







int32_t bfn_server_create(bfn_device device, bfn_graph graph, const bfn_server_options* o,
                          bfn_server* out) {
  if (!device || !graph || !o || !out) return fail(BFN_E_INVALID_ARGUMENT, "null argument");
  if (o->engine > 2 || o->rounds_per_chunk == 0 || o->max_active_queries == 0)
    return fail(BFN_E_INVALID_ARGUMENT, "invalid server options");
  if (o->block_size != 64 && o->block_size != 128 && o->block_size != 256 &&
      o->block_size != 512 && o->block_size != 1024)
    return fail(BFN_E_INVALID_ARGUMENT, "invalid block size");
  auto* sv = new bfn_server_s;
  sv->device = device;
  sv->graph = graph;
  sv->options = *o;
  auto bail = [&](int32_t code, const std::string& m) {
    for (void* p : sv->allocations) (void)hipFree(p);
    if (sv->h_records) (void)hipHostFree(sv->h_records);
    if (sv->h_tiles) (void)hipHostFree(sv->h_tiles);
    if (sv->h_done) (void)hipHostFree(sv->h_done);
    if (sv->h_result) (void)hipHostFree(sv->h_result);
    if (sv->h_summary) (void)hipHostFree(sv->h_summary);
    if (sv->h_path_offsets) (void)hipHostFree(sv->h_path_offsets);
    if (sv->stream) (void)hipStreamDestroy(sv->stream);
    delete sv;
    return fail(code, m);
  };
  (void)hipSetDevice(device_ordinal(device));
  graph_view(graph, &sv->gview, &sv->kin, &sv->kout);
  if (sv->kin == 0 || sv->kout == 0)
    return bail(BFN_E_INVALID_ARGUMENT, "graph view unavailable");
  sv->max_v = sv->gview.max_tile_vertices;
  sv->slab_words = 2 * sv->max_v;
  // Push engines push boundary edges with direct atomics here (no proposal
  // area), which keeps more workgroups resident for small requests.
  sv->work_words = relax_work_words(sv->gview, o->engine == kJacobiPull, o->engine != kJacobiPull, false);
  sv->lds_bytes = 4ull * relax_lds_words(sv->max_v, sv->work_words, sv->gview.max_neighbor_tiles);
  if (sv->lds_bytes > device_lds_limit(device))
    return bail(BFN_E_NOT_REPRESENTABLE, "tile neighbourhood exceeds the LDS budget");
  sv->block = o->block_size;
  sv->Qcap = o->max_active_queries;
  // bytes per tile slot: two label columns, path stamps, three frame words
  // per vertex, neighbour tables, slot records and three list entries
  const std::uint64_t per_slot = 4ull * sv->slab_words * 2 + 12ull * sv->max_v +
                                 4ull * (sv->kin + sv->kout) + 40;
  const std::uint64_t per_query = 4ull * 16 + 28;
  const std::uint64_t budget = o->budget_bytes ? o->budget_bytes : (1ull << 30);
  if (budget <= per_query * sv->Qcap + per_slot)
    return bail(BFN_E_OUT_OF_MEMORY, "server budget holds no tile slots");
  const std::uint64_t slots = (budget - per_query * sv->Qcap) / per_slot;
  sv->Scap = static_cast<std::uint32_t>(std::min<std::uint64_t>(slots, 0x7FFFFFFFull / sv->slab_words));
  sv->upload_tile_capacity = std::min<std::uint32_t>(sv->Scap, 1u << 22);
  if (hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(BFN_E_HIP, "stream creation failed");
  auto alloc = [&](auto** ptr, std::uint64_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<std::uint64_t>(bytes, 256)) != hipSuccess) return false;
    sv->allocations.push_back(p);
    *ptr = static_cast<std::remove_reference_t<decltype(*ptr)>>(p);
    return true;
  };
  const std::uint64_t S = sv->Scap, Q = sv->Qcap;
  bool ok = alloc(&sv->d_tiles, 4 * S) && alloc(&sv->d_slot_query, 4 * S) &&
            alloc(&sv->d_in_base, 4 * S) && alloc(&sv->d_out_base, 4 * S) &&
            alloc(&sv->d_slot_in, 4 * S * sv->kin) && alloc(&sv->d_slot_out, 4 * S * sv->kout) &&
            alloc(&sv->d_stamp, 4 * S) && alloc(&sv->d_lists, 12 * S) &&
            alloc(&sv->d_pool, 4 * S * sv->slab_words) && alloc(&sv->d_onpath, 4 * S * sv->slab_words) &&
            alloc(&sv->d_frames, 12 * S * sv->max_v) && alloc(&sv->d_slot_dist, 8 * S) &&
            alloc(&sv->d_begin, 4 * Q) && alloc(&sv->d_end, 4 * Q) && alloc(&sv->d_src_off, 4 * (Q + 1)) &&
            alloc(&sv->d_src, 4 * Q) && alloc(&sv->d_tgt_off, 4 * (Q + 1)) && alloc(&sv->d_tgt, 4 * Q) &&
            alloc(&sv->d_engine, 4 * Q) && alloc(&sv->d_last, 4 * Q) && alloc(&sv->d_pending, 4 * Q) &&
            alloc(&sv->d_active, 4 * Q) && alloc(&sv->d_done, 4 * (Q + 1)) &&
            alloc(&sv->d_result, 4 * Q) && alloc(&sv->d_clean, 4 * Q) &&
            alloc(&sv->d_summary, 28 * Q) && alloc(&sv->d_lengths, 8 * Q) &&
            alloc(&sv->d_frame_base, 8 * Q) && alloc(&sv->d_path_offsets, 8 * Q) &&
            alloc(&sv->d_control, sizeof(DevControl)) &&
            alloc(&sv->d_records, sizeof(AdmitRecord) * Q) &&
            alloc(&sv->d_upload_tiles, 4ull * sv->upload_tile_capacity);
  if (!ok) return bail(BFN_E_OUT_OF_MEMORY, "server allocation failed");
  if (hipHostMalloc(reinterpret_cast<void**>(&sv->h_records), sizeof(AdmitRecord) * Q) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&sv->h_tiles), 4ull * sv->upload_tile_capacity) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&sv->h_done), 4 * (Q + 1)) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&sv->h_result), 4 * Q) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&sv->h_summary), 28 * Q) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&sv->h_path_offsets), 8 * Q) != hipSuccess)
    return bail(BFN_E_OUT_OF_MEMORY, "pinned staging allocation failed");
  DevBatch& b = sv->view;
  b.query_count = sv->Qcap;
  b.slot_count = sv->Scap;
  b.maximum_rounds = 0;
  b.tile_sweeps = o->tile_sweeps;
  b.items_per_claim = 1;
  b.instrumentation = 0;
  b.boundary_mode = 1;
  b.min_waves = 0;
  b.ordered = 0;  // arrival-order lists: admitted work joins the running round
  b.active = nullptr;
  b.window = 0;
  b.prio = nullptr;
  b.engines = o->engine == kJacobiPull ? kEnginesPull : kEnginesPush;  // one engine per server
  // A request's slots form one contiguous range of the arena, in tile
  // order, so positions and slots coincide.
  b.tile_begin = sv->d_begin;
  b.tile_end = sv->d_end;
  b.query_tiles = sv->d_tiles;
  b.query_slot = nullptr;
  b.tiles = sv->d_tiles;
  b.source_offsets = sv->d_src_off;
  b.sources = sv->d_src;
  b.target_offsets = sv->d_tgt_off;
  b.targets = sv->d_tgt;
  b.engine = sv->d_engine;
  b.slot_query = sv->d_slot_query;
  b.slot_dist = sv->d_slot_dist;
  b.slot_in_base = sv->d_in_base;
  b.slot_out_base = sv->d_out_base;
  b.slot_in = sv->d_slot_in;
  b.slot_out = sv->d_slot_out;
  b.stamp = sv->d_stamp;
  b.lists = sv->d_lists;
  b.dist = sv->d_pool;
  b.last_change = sv->d_last;
  b.pending = sv->d_pending;
  b.control = sv->d_control;
  const std::uint32_t n = std::max(sv->Scap, sv->Qcap + 1);
  hipLaunchKernelGGL(server_init_kernel, dim3((n + 255) / 256), dim3(256), 0, sv->stream, b, sv->Qcap,
                     sv->Scap, sv->slab_words, sv->kin, sv->kout, sv->d_active, sv->d_src_off,
                     sv->d_tgt_off, sv->d_clean);
  int per_cu = 0;
  if (relax_occupancy(b.engines, sv->block, false, false, 0, sv->lds_bytes, &per_cu) != hipSuccess ||
      per_cu <= 0)
    return bail(BFN_E_NOT_REPRESENTABLE, "relaxation kernel cannot be resident");
  sv->grid = static_cast<std::uint32_t>(per_cu) * device_cu_count(device);
  if (hipStreamSynchronize(sv->stream) != hipSuccess) return bail(BFN_E_HIP, "server init failed");
  sv->free_ranges[0] = sv->Scap;
  for (std::uint32_t q = sv->Qcap; q-- > 0;) sv->free_queries.push_back(q);
  sv->worker = std::thread([sv] { sv->run(); });
  *out = sv;
  return BFN_OK;
}

int32_t bfn_server_submit(bfn_server sv, const bfn_request* r) {
  if (!sv || !r || (r->tile_count && !r->selected_tiles))
    return fail(BFN_E_INVALID_ARGUMENT, "null argument");
  const DevGraph& g = sv->gview;
  if (r->source >= g.vertex_count || r->target >= g.vertex_count || r->tile_count == 0)
    return fail(BFN_E_INVALID_ARGUMENT, "request terminals or tiles out of range");
  if (r->tile_count > sv->Scap || r->tile_count > sv->upload_tile_capacity)
    return fail(BFN_E_OUT_OF_MEMORY, "request region exceeds the server's slot arena");
  Pending p;
  p.request = *r;
  p.request.selected_tiles = nullptr;
  p.tiles.assign(r->selected_tiles, r->selected_tiles + r->tile_count);
  for (std::size_t i = 0; i < p.tiles.size(); ++i)
    if (p.tiles[i] >= g.tile_count || (i && p.tiles[i] <= p.tiles[i - 1]))
      return fail(BFN_E_INVALID_ARGUMENT, "selected tiles must be sorted, unique and in range");
  p.submit_ns = now_ns();
  {
    std::lock_guard<std::mutex> lock(sv->mutex);
    if (!sv->fatal.empty()) return fail(BFN_E_HIP, sv->fatal);
    sv->inbox.push_back(std::move(p));
  }
  sv->wake.notify_one();
  return BFN_OK;
}

int32_t bfn_server_poll(bfn_server sv, bfn_response* responses, uint32_t capacity, uint32_t* count,
                        uint32_t timeout_ms) {
  if (!sv || !count || (capacity && !responses)) return fail(BFN_E_INVALID_ARGUMENT, "null argument");
  std::unique_lock<std::mutex> lock(sv->mutex);
  sv->delivered.wait_for(lock, std::chrono::milliseconds(timeout_ms),
                         [&] { return !sv->outbox.empty() || sv->stopping; });
  std::uint32_t n = 0;
  while (n < capacity && !sv->outbox.empty()) {
    responses[n++] = sv->outbox.front().response;
    sv->outbox.pop_front();
  }
  *count = n;
  if (!sv->fatal.empty()) return fail(BFN_E_HIP, sv->fatal);
  return BFN_OK;
}

int32_t bfn_server_poll_paths(bfn_server sv, bfn_response* responses, uint32_t capacity,
                              uint32_t* count, uint32_t timeout_ms, uint32_t* vertices,
                              uint32_t* edge_ids, uint64_t vertex_capacity, uint64_t* path_offsets) {
  if (!sv || !count || !path_offsets || (capacity && (!responses || !vertices || !edge_ids)))
    return fail(BFN_E_INVALID_ARGUMENT, "null argument");
  std::unique_lock<std::mutex> lock(sv->mutex);
  sv->delivered.wait_for(lock, std::chrono::milliseconds(timeout_ms),
                         [&] { return !sv->outbox.empty() || sv->stopping; });
  std::uint32_t n = 0;
  std::uint64_t used = 0;
  path_offsets[0] = 0;
  while (n < capacity && !sv->outbox.empty()) {
    const Delivered& d = sv->outbox.front();
    if (used + d.vertices.size() > vertex_capacity) {
      if (n == 0) {
        *count = 0;
        return fail(BFN_E_INVALID_ARGUMENT, "vertex capacity below the next path's " +
                                                std::to_string(d.vertices.size()) + " vertices");
      }
      break;
    }
    responses[n] = d.response;
    std::copy(d.vertices.begin(), d.vertices.end(), vertices + used);
    std::copy(d.edges.begin(), d.edges.end(), edge_ids + used);
    used += d.vertices.size();
    path_offsets[++n] = used;
    sv->outbox.pop_front();
  }
  *count = n;
  if (!sv->fatal.empty()) return fail(BFN_E_HIP, sv->fatal);
  return BFN_OK;
}

int32_t bfn_server_stats(bfn_server sv, uint64_t* launches, uint64_t* admitted, uint64_t* completed) {
  if (!sv) return fail(BFN_E_INVALID_ARGUMENT, "null server");
  std::lock_guard<std::mutex> lock(sv->mutex);
  if (launches) *launches = sv->launches;
  if (admitted) *admitted = sv->admitted;
  if (completed) *completed = sv->completed;
  return BFN_OK;
}

int32_t bfn_server_destroy(bfn_server sv) {
  if (!sv) return BFN_OK;
  {
    std::lock_guard<std::mutex> lock(sv->mutex);
    sv->stopping = true;
  }
  sv->wake.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  (void)hipStreamSynchronize(sv->stream);
  for (void* p : sv->allocations) (void)hipFree(p);
  (void)hipHostFree(sv->h_records);
  (void)hipHostFree(sv->h_tiles);
  (void)hipHostFree(sv->h_done);
  (void)hipHostFree(sv->h_result);
  (void)hipHostFree(sv->h_summary);
  (void)hipHostFree(sv->h_path_offsets);
  sv->free_arena();
  (void)hipStreamDestroy(sv->stream);
  delete sv;
  return BFN_OK;
}
