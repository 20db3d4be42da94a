// noalloc.hip -- bfn_batch_run_static: a batch run that allocates nothing.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc): the same
// engines as bfn_batch_submit, with every piece of per-batch device state
// (control block, batch description, slot tables, work lists, label pool,
// per-query status, target labels) carved out of one file-scope static
// device array, `scratch`, of kStaticScratchBytes per GPU.  The caller
// passes the batch (host arrays, copied in) and the host arrays the results
// are copied out to.
//
// Slots are numbered by position (query-major: query q's selected tiles are
// slots tile_offsets[q] .. tile_offsets[q+1]), so the slot tables follow
// from the description by one scan on the device and no host-side plan is
// built.  Every call on a GPU uses that GPU's one array; a call clears what
// must start cleared (control block, labels, stamps, list counters) with
// kernels before it runs, and the array keeps the last call's contents.
// Calls that overlap on one GPU (from several host threads) race on it.
#include "bfnew_amd.h"

#include "bfn_launch.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

namespace bfn {
namespace {

constexpr std::size_t kStaticScratchBytes = 64ull << 20;  // 64 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) std::uint32_t scratch[kStaticScratchBytes / 4];

// Device address of each GPU's `scratch`, looked up once per device with
// that device current (host code never uses the array's name as an address).
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

hipError_t scratch_address(int ordinal, void** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[ordinal]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[ordinal] = p;
  }
  *out = g_scratch_ptr[ordinal];
  return hipSuccess;
}

constexpr std::size_t align256(std::size_t x) { return (x + 255) / 256 * 256; }

// Per-slot tables from the description: slot s is position s, its query the
// q with tile_offsets[q] <= s < tile_offsets[q+1]; pool offsets and
// neighbour-table bases are exclusive prefix sums over the slots (two label
// copies of npt words; the tile's in- and out-neighbour counts).  One
// workgroup; a carry runs across 1024-slot chunks.
__global__ __launch_bounds__(1024) void static_slots_kernel(DevGraph g, DevBatch b) {
  __shared__ unsigned long long pool[1024];
  __shared__ std::uint32_t kin[1024], kout[1024];
  const std::uint32_t tid = threadIdx.x;
  unsigned long long pool_carry = 0;
  std::uint32_t in_carry = 0, out_carry = 0;
  for (std::uint32_t base = 0; base < b.slot_count; base += 1024) {
    const std::uint32_t s = base + tid;
    unsigned long long words = 0;
    std::uint32_t ni = 0, no = 0;
    if (s < b.slot_count) {
      const std::uint32_t T = b.query_tiles[s];
      words = 2ull * (g.tile_vertex_offsets[T + 1] - g.tile_vertex_offsets[T]);
      ni = g.in_nbr_offsets[T + 1] - g.in_nbr_offsets[T];
      no = g.out_nbr_offsets[T + 1] - g.out_nbr_offsets[T];
    }
    pool[tid] = words;
    kin[tid] = ni;
    kout[tid] = no;
    __syncthreads();
    for (std::uint32_t off = 1; off < 1024; off <<= 1) {
      const unsigned long long pw = tid >= off ? pool[tid - off] : 0ull;
      const std::uint32_t iw = tid >= off ? kin[tid - off] : 0u;
      const std::uint32_t ow = tid >= off ? kout[tid - off] : 0u;
      __syncthreads();
      pool[tid] += pw;
      kin[tid] += iw;
      kout[tid] += ow;
      __syncthreads();
    }
    if (s < b.slot_count) {
      // query of slot s: the last q with tile_begin[q] <= s
      std::uint32_t lo = 0, hi = b.query_count;
      while (hi - lo > 1) {
        const std::uint32_t mid = (lo + hi) >> 1;
        if (b.tile_begin[mid] <= s) lo = mid; else hi = mid;
      }
      const_cast<std::uint32_t*>(b.slot_query)[s] = lo;
      const_cast<unsigned long long*>(b.slot_dist)[s] = pool_carry + pool[tid] - words;
      const_cast<std::uint32_t*>(b.slot_in_base)[s] = in_carry + kin[tid] - ni;
      const_cast<std::uint32_t*>(b.slot_out_base)[s] = out_carry + kout[tid] - no;
    }
    pool_carry += pool[1023];
    in_carry += kin[1023];
    out_carry += kout[1023];
    __syncthreads();
  }
}

// Label of each target of each query (copy 0, or the Jacobi column the last
// round wrote), as float bits.
__global__ void static_targets_kernel(DevGraph g, DevBatch b, std::uint32_t* out) {
  const std::uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= b.query_count) return;
  const std::uint32_t jacobi_copy = b.control->rounds_executed & 1u;
  for (std::uint32_t i = b.target_offsets[q]; i < b.target_offsets[q + 1]; ++i) {
    const std::uint32_t v = b.targets[i];
    const std::uint32_t T = g.owner_tiles[v];
    const std::uint32_t s = find_slot(b, q, T);
    std::uint32_t label = kInfBits;
    if (s != kNone) {
      const std::uint32_t v0 = g.tile_vertex_offsets[T];
      const std::uint32_t npt = g.tile_vertex_offsets[T + 1] - v0;
      const std::uint32_t col = b.engine[q] == kJacobiPull ? jacobi_copy : 0u;
      label = b.dist[b.slot_dist[s] + static_cast<unsigned long long>(col) * npt + (v - v0)];
    }
    out[i] = label;
  }
}

__global__ void static_fill_kernel(std::uint32_t* p, std::uint32_t n, std::uint32_t value) {
  const std::uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

int32_t fail(int32_t code, const std::string& message) { return set_error(code, message); }

#define BFN_HIP_STATIC(expr)                                                                  \
  do {                                                                                        \
    const hipError_t bfn_status_ = (expr);                                                    \
    if (bfn_status_ != hipSuccess)                                                            \
      return fail(BFN_E_HIP, std::string(#expr) + ": " + hipGetErrorString(bfn_status_));     \
  } while (0)

}  // namespace
}  // namespace bfn

using namespace bfn;

uint64_t bfn_static_scratch_bytes(void) { return kStaticScratchBytes; }

int32_t bfn_batch_run_static(bfn_graph graph, const bfn_batch_desc* d, const bfn_run_options* o,
                             const bfn_tuning* tuning, bfn_query_status* per_query,
                             float* target_distances) {
  if (!graph || !per_query || !target_distances) return fail(BFN_E_INVALID_ARGUMENT, "null argument");
  if (int32_t rc = check_run_options(o)) return rc;
  bfn_tuning t;
  if (tuning) t = *tuning; else bfn_tuning_default(&t);
  if (t.tile_sweeps > 1 || t.boundary_mode > 1 || t.min_waves > 8 || t.claim_order > 1)
    return fail(BFN_E_INVALID_ARGUMENT,
                "invalid tuning (tile_sweeps, boundary_mode, min_waves, claim_order)");
  if (!(t.push_window >= 0.0f) || !std::isfinite(t.push_window))
    return fail(BFN_E_INVALID_ARGUMENT, "push_window must be a finite non-negative number");
  if (t.claim_order && o->control_mode == 0)
    return fail(BFN_E_INVALID_ARGUMENT, "claim_order 1 needs per-round control (control_mode 1 or 2)");
  if (t.use_graph || t.count_relaxations)
    return fail(BFN_E_INVALID_ARGUMENT,
                "graph replay and relaxation counts are not available without allocation");
  if (int32_t rc = check_batch(graph, d, o)) return rc;
  const GraphHost gh = graph_host(graph);
  if (gh.ordinal < 0 || gh.ordinal >= kMaxDevices) return fail(BFN_E_INVALID_ARGUMENT, "device ordinal");
  const DevGraph& gv = gh.view;

  // Sizes, on the host, without building a plan.
  const std::uint32_t Q = d->query_count, S = d->tile_offsets[Q];
  const std::uint32_t ns = d->source_offsets[Q], nt = d->target_offsets[Q];
  std::uint64_t pool_words = 0, in_total = 0, out_total = 0;
  bool any_jacobi = false, any_push = false;
  for (std::uint32_t q = 0; q < Q; ++q) {
    const std::uint32_t e = d->engine_by_query ? d->engine_by_query[q] : o->engine;
    any_jacobi |= e == kJacobiPull;
    any_push |= e != kJacobiPull;
  }
  for (std::uint32_t i = 0; i < S; ++i) {
    const std::uint32_t T = d->selected_tiles[i];
    pool_words += 2ull * (gh.tile_offsets[T + 1] - gh.tile_offsets[T]);
    in_total += gh.in_offsets[T + 1] - gh.in_offsets[T];
    out_total += gh.out_offsets[T + 1] - gh.out_offsets[T];
  }
  // Scratch carve-up, 256-byte aligned segments.
  std::size_t at = 0;
  auto take = [&](std::size_t bytes) {
    const std::size_t here = at;
    at = align256(at + std::max<std::size_t>(bytes, 4));
    return here;
  };
  const std::size_t o_control = take(sizeof(DevControl));
  const std::size_t o_status = take(sizeof(QueryStatusDev) * Q);
  const std::size_t o_targets = take(4ull * nt);
  const std::size_t o_tile_off = take(4ull * (Q + 1));
  const std::size_t o_tiles = take(4ull * S);
  const std::size_t o_src_off = take(4ull * (Q + 1));
  const std::size_t o_src = take(4ull * ns);
  const std::size_t o_tgt_off = take(4ull * (Q + 1));
  const std::size_t o_tgt = take(4ull * nt);
  const std::size_t o_engine = take(4ull * Q);
  const std::size_t o_slot_query = take(4ull * S);
  const std::size_t o_slot_dist = take(8ull * S);
  const std::size_t o_in_base = take(4ull * S);
  const std::size_t o_out_base = take(4ull * S);
  const std::size_t o_desc = take(sizeof(SlotDesc) * S);
  const std::size_t o_slot_in = take(4ull * in_total);
  const std::size_t o_slot_out = take(4ull * out_total);
  const std::size_t o_stamp = take(4ull * S);
  const std::size_t o_prio = take(t.push_window > 0.0f ? 4ull * S : 0ull);
  const std::size_t o_lists = take(12ull * S);
  const std::size_t o_last = take(4ull * Q);
  const std::size_t o_pool = take(4ull * pool_words);
  if (at > kStaticScratchBytes)
    return fail(BFN_E_OUT_OF_MEMORY, "batch needs " + std::to_string(at) +
                                         " bytes of device scratch; the static scratch holds " +
                                         std::to_string(kStaticScratchBytes));
  if (pool_words > 0xFFFFFFFFull || in_total >= 0xFFFFFFFFull || out_total >= 0xFFFFFFFFull)
    return fail(BFN_E_NOT_REPRESENTABLE, "batch exceeds 32-bit slot indexing");

  // Launch shape (direct boundary pushes: no proposal slabs).
  const std::uint32_t max_v = gv.max_tile_vertices;
  const std::uint32_t work_words = relax_work_words(gv, any_jacobi, any_push, false);
  const std::size_t lds_bytes = 4ull * relax_lds_words(max_v, work_words, gv.max_neighbor_tiles);
  if (lds_bytes > gh.lds_limit)
    return fail(BFN_E_NOT_REPRESENTABLE, "tile neighbourhood exceeds the LDS budget");
  const bool persistent = o->control_mode == 0;
  int per_cu = 0;
  BFN_HIP_STATIC(hipSetDevice(gh.ordinal));
  const std::uint32_t engines = !any_jacobi ? kEnginesPush : !any_push ? kEnginesPull : kEnginesMixed;
  BFN_HIP_STATIC(relax_occupancy(engines, o->block_size, persistent, o->instrumentation != 0, t.min_waves,
                                 lds_bytes, &per_cu));
  if (per_cu <= 0) return fail(BFN_E_NOT_REPRESENTABLE, "relaxation kernel cannot be resident");
  std::uint32_t grid = static_cast<std::uint32_t>(per_cu) * gh.cu_count;
  if (o->grid_policy == 1) grid = std::min(grid, o->blocks_per_wgp * gh.cu_count);
  grid = std::max(grid, 1u);

  void* base_ptr = nullptr;
  BFN_HIP_STATIC(scratch_address(gh.ordinal, &base_ptr));
  char* m = static_cast<char*>(base_ptr);
  auto u32 = [&](std::size_t off) { return reinterpret_cast<std::uint32_t*>(m + off); };
  DevBatch b{};
  b.query_count = Q;
  b.slot_count = S;
  b.maximum_rounds = o->maximum_rounds >= 0xFFFFFFFFull ? 0u : static_cast<std::uint32_t>(o->maximum_rounds);
  b.tile_sweeps = t.tile_sweeps;
  b.items_per_claim = std::max(1u, t.items_per_claim);
  b.instrumentation = o->instrumentation;
  b.boundary_mode = 1;
  b.min_waves = t.min_waves;
  b.ordered = t.claim_order;
  b.engines = engines;
  std::memcpy(&b.window, &t.push_window, sizeof b.window);
  b.tile_begin = u32(o_tile_off);
  b.tile_end = u32(o_tile_off) + 1;
  b.query_tiles = u32(o_tiles);
  b.query_slot = nullptr;  // slot = position
  b.tiles = u32(o_tiles);
  b.source_offsets = u32(o_src_off);
  b.sources = u32(o_src);
  b.target_offsets = u32(o_tgt_off);
  b.targets = u32(o_tgt);
  b.engine = u32(o_engine);
  b.slot_query = u32(o_slot_query);
  b.slot_dist = reinterpret_cast<const unsigned long long*>(m + o_slot_dist);
  b.slot_in_base = u32(o_in_base);
  b.slot_out_base = u32(o_out_base);
  b.slot_in = u32(o_slot_in);
  b.slot_out = u32(o_slot_out);
  b.slot_out_off = nullptr;
  b.slot_desc = reinterpret_cast<const SlotDesc*>(m + o_desc);
  b.stamp = u32(o_stamp);
  b.active = u32(o_stamp);
  b.prio = b.window ? u32(o_prio) : nullptr;
  b.lists = u32(o_lists);
  b.dist = u32(o_pool);
  b.last_change = u32(o_last);
  b.pending = nullptr;
  b.control = reinterpret_cast<DevControl*>(m + o_control);

  const hipStream_t s = gh.stream;
  // Nothing of an earlier call may still read the scratch.
  BFN_HIP_STATIC(hipStreamSynchronize(s));
  if (persistent) {
    // A persistent grid must be co-resident: checked on the device.
    int placed = 0;
    BFN_HIP_STATIC(coresident_per_cu(gh.ordinal, per_cu, gh.cu_count, lds_bytes, max_v, work_words,
                                     gv, b, s, o->block_size, &placed));
    if (placed <= 0)
      return fail(BFN_E_NOT_REPRESENTABLE, "no persistent grid of the relaxation kernel is co-resident");
    grid = static_cast<std::uint32_t>(placed) * gh.cu_count;
    if (o->grid_policy == 1) grid = std::min(grid, o->blocks_per_wgp * gh.cu_count);
  }
  grid = std::min(grid, std::max(S, 1u));  // no more workgroups than slots
  auto h2d = [&](std::size_t off, const void* src, std::size_t bytes) {
    return bytes ? hipMemcpyAsync(m + off, src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
  };
  BFN_HIP_STATIC(h2d(o_tile_off, d->tile_offsets, 4ull * (Q + 1)));
  BFN_HIP_STATIC(h2d(o_tiles, d->selected_tiles, 4ull * S));
  BFN_HIP_STATIC(h2d(o_src_off, d->source_offsets, 4ull * (Q + 1)));
  BFN_HIP_STATIC(h2d(o_src, d->sources, 4ull * ns));
  BFN_HIP_STATIC(h2d(o_tgt_off, d->target_offsets, 4ull * (Q + 1)));
  BFN_HIP_STATIC(h2d(o_tgt, d->targets, 4ull * nt));
  if (d->engine_by_query) {
    BFN_HIP_STATIC(h2d(o_engine, d->engine_by_query, 4ull * Q));
  } else {
    hipLaunchKernelGGL(static_fill_kernel, dim3((Q + 255) / 256), dim3(256), 0, s, u32(o_engine), Q,
                       o->engine);
    BFN_HIP_STATIC(hipGetLastError());
  }
  hipLaunchKernelGGL(static_slots_kernel, dim3(1), dim3(1024), 0, s, gv, b);
  BFN_HIP_STATIC(hipGetLastError());
  // Control reset, label fill, neighbour slots, stamps, seeds.
  BFN_HIP_STATIC(prep_launch(gv, b, pool_words, s));

  if (persistent) {
    BFN_HIP_STATIC(relax_launch(o->block_size, true, true, grid, lds_bytes, max_v, work_words, gv, b, s));
  } else {
    const std::uint32_t chunk = o->control_mode == 2 ? 1u : o->rounds_per_chunk;
    for (;;) {
      for (std::uint32_t r = 0; r < chunk; ++r)
        BFN_HIP_STATIC(relax_launch(o->block_size, false, false, grid, lds_bytes, max_v, work_words, gv, b, s));
      std::uint32_t done = 0;  // pageable host word: a synchronous read-back
      BFN_HIP_STATIC(hipMemcpyAsync(&done, &b.control->done, sizeof done, hipMemcpyDeviceToHost, s));
      BFN_HIP_STATIC(hipStreamSynchronize(s));
      if (done) break;
    }
  }
  BFN_HIP_STATIC(finalize_launch(gv, b, nullptr, m + o_status, s));
  hipLaunchKernelGGL(static_targets_kernel, dim3((Q + 127) / 128), dim3(128), 0, s, gv, b, u32(o_targets));
  BFN_HIP_STATIC(hipGetLastError());
  static_assert(sizeof(bfn_query_status) == sizeof(QueryStatusDev), "per-query status layout");
  BFN_HIP_STATIC(hipMemcpyAsync(per_query, m + o_status, sizeof(QueryStatusDev) * Q,
                                hipMemcpyDeviceToHost, s));
  BFN_HIP_STATIC(hipMemcpyAsync(target_distances, m + o_targets, 4ull * nt, hipMemcpyDeviceToHost, s));
  BFN_HIP_STATIC(hipStreamSynchronize(s));
  return BFN_OK;
}
