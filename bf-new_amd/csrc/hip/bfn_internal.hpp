// bfn_internal.hpp -- device-side records of the MI355X engine.
//
// Layout in HBM (all tile-major; vertex ids are the reference's reordered ids,
// so every tile owns one contiguous vertex range, reorder.cpp:266-273):
//
//   graph (immutable, shared by every query)
//     reference image : owner_tiles, CSR/CSC offsets, endpoints, weights,
//                       CSC edge ids, runs            (DeviceGraphLayout32)
//     tile directory  : tile_vertex_offsets[T+1]
//                       in_nbr_offsets[T+1] / in_nbr_tiles[]   (source tiles of
//                       each tile's incoming edges, the tile itself first)
//                       out_nbr_offsets[T+1] / out_nbr_tiles[] (same, outgoing)
//                       in_nbr_stage[]  LDS float offset of each in-neighbour slab
//                       out_nbr_stage[] LDS float offset of each out-neighbour
//                       proposal slab (k >= 1; the tile itself has none)
//     packed edges    : csc_pack[E] = k << 24 | src_local << 12 | dst_local
//                       (k indexes the destination tile's in-neighbour list)
//                       csr_pack[E] = k << 24 | src_local << 12 | dst_local
//                       (k indexes the source tile's out-neighbour list)
//     relax streams   : csc_rec[E] / csr_rec[E] = (pack, weight bits), each
//                       tile's block reordered [internal | boundary]; the
//                       engines keep the internal part in registers for the
//                       tile closure and stream the boundary part once
//
//   batch (one workspace; any number of queries)
//     slot s = (query q, selected tile T).  A batch numbers its slots
//     tile-major (all slots of tile 0, then tile 1, ...; by query within a
//     tile), so the labels of one tile's queries sit together in the pool.
//     query_tiles[tile_begin[q] .. tile_end[q]) are q's selected tiles, sorted;
//     query_slot[] holds the slot of each of those positions.
//     slot_dist[s]   float offset of the slot's two npt-long copies in the pool:
//                    Jacobi uses them as the two distance columns (read r&1,
//                    write (r+1)&1); push engines keep distances in copy 0 and
//                    the last-pushed snapshot in copy 1.
//     slot_in[]/slot_out[]  slot of each in/out-neighbour tile for q, or NONE
//     stamp[s]       1 + last round s was queued for (dedups the work lists)
//     lists[3][S]    round-robin work lists; list r%3 is consumed in round r,
//                    list (r+1)%3 is filled, list (r+2)%3 is reset.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bfn {

constexpr std::uint32_t kNone = 0xFFFFFFFFu;
constexpr std::uint32_t kInfBits = 0x7F800000u;  // +inf; non-negative floats order as u32
constexpr std::uint32_t kMaxTileVertices = 4096;  // 12-bit local ids
constexpr std::uint32_t kMaxNeighborTiles = 255;  // 8-bit neighbour index
constexpr std::uint32_t kPackLocalMask = 0xFFFu;
constexpr int kPhaseCount = 5;  // visit phases timed by instrumented runs
// LDS table of a visit, sized by the graph's largest neighbour-tile list K
// (rounded up to 4): neighbour slots [0, K), pushed-into marks [K, 2K),
// per-neighbour offsets [2K, 4K) (two words each: a pool offset, or one
// word: a staged slab offset); then the flag words.
constexpr std::uint32_t kLdsFlagWords = 16;

// Relaxation kernel variants by the engines of a batch: a kernel that only
// holds the push visit (or only the pull visit) needs fewer registers per
// lane than one that holds both, so more waves are resident.
constexpr std::uint32_t kEnginesMixed = 0;
constexpr std::uint32_t kEnginesPush = 1;
constexpr std::uint32_t kEnginesPull = 2;
// Workgroups are dispatched round-robin over the 8 XCDs (blockIdx % 8):
// per-XCD counters (claims, window minima) are indexed by that group.  The
// grid barrier gathers its arrivals on kArrivalGroups counters (a multiple
// of 8, so each counter's workgroups share an XCD) before the top counter.
constexpr std::uint32_t kBarrierGroups = 8;
constexpr std::uint32_t kMaxArrivalGroups = 64;  // arrival counters in DevControl
constexpr std::uint32_t kArrivalGroups = 32;     // used by the relaxation kernels
constexpr std::uint32_t kBarrierStride = 32;  // words: one 128-byte line per group

enum Engine : std::uint32_t { kJacobiPull = 0, kDensePush = 1, kFrontierPush = 2 };

enum StopReason : std::uint32_t {
  kStopNone = 0,
  kStopConverged = 1,
  kStopMaximumRounds = 2,
  kStopQueueOverflow = 3,
  kStopInvalidController = 4,
  kStopDeviceFailure = 5,
};

struct DevGraph {
  std::uint32_t vertex_count;
  std::uint32_t edge_count;
  std::uint32_t tile_count;
  std::uint32_t max_tile_vertices;
  std::uint32_t max_stage_floats;
  std::uint32_t max_in_edges;
  std::uint32_t max_out_edges;
  std::uint32_t max_out_stage_floats;  // largest out-neighbour proposal area (k >= 1)
  std::uint32_t max_neighbor_tiles;    // largest in- or out-neighbour tile list (<= 255)
  std::uint32_t pad0;
  const std::uint32_t* owner_tiles;
  const std::uint32_t* tile_vertex_offsets;
  const std::uint32_t* row_offsets;
  const std::uint32_t* destinations;
  const float* weights;
  const std::uint32_t* column_offsets;
  const std::uint32_t* sources;
  const float* in_weights;
  const std::uint32_t* in_edge_ids;
  const std::uint32_t* in_nbr_offsets;
  const std::uint32_t* in_nbr_tiles;
  const std::uint32_t* in_nbr_stage;
  const std::uint32_t* out_nbr_offsets;
  const std::uint32_t* out_nbr_tiles;
  const std::uint32_t* out_nbr_stage;   // LDS offset of each out-neighbour proposal slab
  const std::uint32_t* csc_pack;
  const std::uint32_t* csr_pack;
  // Relaxation streams: each tile's block as [internal | boundary] records
  // (pack, weight bits); *_intra[T] = internal records of tile T.
  const uint2* csc_rec;
  const uint2* csr_rec;
  const std::uint32_t* csc_intra;
  const std::uint32_t* csr_intra;
  // csr_bnd[v]: position in csr_rec of v's first boundary record; v's
  // boundary records run to csr_bnd[v + 1], or to the end of the tile's CSR
  // block for the tile's last vertex.
  const std::uint32_t* csr_bnd;
};

// Control block; lives in device memory, one per workspace.
// Words that workgroups hit with atomics or polls during a round sit on
// 128-byte lines of their own: agent-scope atomics and coherent loads of one
// line are served one after another at the memory side, so enqueues
// (list_count), claims (claim) and the barrier polls of finished workgroups
// (barrier_generation) would otherwise queue behind each other.
struct DevControl {
  std::uint32_t round;            // next round to execute
  std::uint32_t done;
  std::uint32_t stop_reason;
  std::uint32_t error_bits;
  std::uint32_t rounds_executed;
  alignas(128) std::uint32_t list_count[3];
  alignas(128) std::uint32_t claim[3];
  alignas(128) std::uint32_t barrier_count;
  alignas(128) std::uint32_t barrier_generation;
  alignas(128) unsigned long long items_processed;
  unsigned long long edges_examined;
  unsigned long long atomic_attempts;
  unsigned long long atomic_successes;
  unsigned long long max_list;
  unsigned long long edges_relaxed;      // edge records evaluated by visits
  unsigned long long algorithmic_bytes;  // modelled HBM bytes of the visits
  unsigned long long closure_sweeps;     // tile-closure sweeps (one per block pass)
  unsigned long long candidates;         // candidate labels computed by the visits
  unsigned long long phase_cycles[kPhaseCount];  // summed over workgroups
  // Slot-order claims: the compacted list of the round is cut into one
  // segment per XCD group, [seg_begin[x], seg_begin[x+1]) (read once per
  // round launch); seg_claim[x * kBarrierStride] is the next unclaimed
  // position of segment x, each counter on its own 128-byte line.
  std::uint32_t seg_begin[kBarrierGroups + 1];
  alignas(128) std::uint32_t seg_claim[kBarrierGroups * kBarrierStride];
  // Push window: rounds visit the queued slots whose priority (smallest
  // label pushed in since their last visit) is at most `window_limit`.
  // Round r gathers the smallest priority it leaves queued for round r + 1
  // in next_min[(r+1) % 3] (one counter per XCD group, each on its own
  // line); the limit of round r + 1 is that minimum plus the window.
  std::uint32_t window_limit;
  alignas(128) std::uint32_t next_min[3 * kBarrierGroups * kBarrierStride];
  // Arrival counters of the two-level grid barrier, one per group of
  // workgroups (blockIdx % G), each on its own 128-byte line.
  alignas(128) std::uint32_t group_arrivals[kMaxArrivalGroups * kBarrierStride];
};

// What a visit of slot s needs before it can load any label: written once
// per batch by the preparation kernel, so a visit starts with one 64-byte
// read instead of a chain of dependent table reads.  The neighbour list and
// edge block are those of the slot's engine: incoming tiles and the CSC
// block for Jacobi pull, outgoing tiles and the CSR block for the push
// engines.
struct SlotDesc {
  std::uint32_t tile, query, engine, v0;
  std::uint32_t npt, nbr0, nbr_count, nbr_base;  // g.*_nbr_offsets[tile], list length, b.slot_*_base[s]
  std::uint32_t e0, ei, e1, dist_lo;             // edge block [e0, ei) internal, [ei, e1) boundary
  std::uint32_t dist_hi;                         // pool offset of the slot's labels (64 bits)
  std::uint32_t out_nbr0, out_count, out_base;   // out-neighbour list and slots (whom to queue)
};
static_assert(sizeof(SlotDesc) == 64, "slot descriptor is one 64-byte record");

struct DevBatch {
  std::uint32_t query_count;
  std::uint32_t slot_count;
  std::uint32_t maximum_rounds;   // 0 = unlimited
  std::uint32_t tile_sweeps;      // Jacobi: 1 point, 0 closure
  std::uint32_t items_per_claim;
  std::uint32_t instrumentation;
  std::uint32_t boundary_mode;    // push: 0 LDS proposals when they fit, 1 direct atomics
  std::uint32_t min_waves;        // kernel variant bounded to >= this many waves per SIMD
  std::uint32_t ordered;          // slot-order claims: queue as bits in `active`
  std::uint32_t window;           // push window (float bits), 0 = off
  std::uint32_t engines;          // kEnginesMixed / kEnginesPush / kEnginesPull: the
                                  // relaxation kernel variant, by the engines the batch holds
  std::uint32_t probe;            // persistent kernel: pass one grid barrier and exit
  const std::uint32_t* tile_begin;      // [Q]   first position of q in query_tiles
  const std::uint32_t* tile_end;        // [Q]   one past q's last position
  const std::uint32_t* query_tiles;     // [.]   q's selected tiles, sorted
  const std::uint32_t* query_slot;      // [.]   slot of each position, or null: the same
  const std::uint32_t* tiles;           // [S]   tile of slot
  const std::uint32_t* source_offsets;  // [Q+1]
  const std::uint32_t* sources;
  const std::uint32_t* target_offsets;  // [Q+1]
  const std::uint32_t* targets;
  const std::uint32_t* engine;          // [Q]
  const std::uint32_t* slot_query;      // [S]
  const unsigned long long* slot_dist;  // [S]
  const std::uint32_t* slot_in_base;    // [S]
  const std::uint32_t* slot_out_base;   // [S]
  const SlotDesc* slot_desc;            // [S] visit records, or null: derived per visit
  std::uint32_t* slot_in;
  std::uint32_t* slot_out;
  // Pool offset of each out-neighbour slot (slot_out order), ~0 for none;
  // null when visits look the offsets up through slot_dist.
  unsigned long long* slot_out_off;
  std::uint32_t* stamp;                 // [S]
  // Slot-order claims: bit s of `active` queues slot s for the next round
  // (ceil(S/32) words, sharing the stamp buffer, which that mode leaves
  // unused); the compacted list of a round is lists[0 .. n).
  std::uint32_t* active;
  std::uint32_t* prio;                  // [S] push window: priority of each slot, or null
  std::uint32_t* lists;                 // [3*S]
  std::uint32_t* dist;                  // float bits pool
  std::uint32_t* last_change;           // [Q] 1 + last round with a decrease
  std::uint32_t* pending;               // [Q] queued work items of q, or null
  DevControl* control;
};

// Per-query outcome: rounds, convergence, all-targets reach.
struct QueryStatusDev {
  unsigned long long rounds;
  unsigned long long relaxations;
  std::uint32_t converged;
  std::uint32_t reached;
  std::uint32_t miss;
  std::uint32_t stop_reason;
  std::uint32_t error_bits;
  std::uint32_t engine;
};

// Arguments of the compact-path passes.
struct PathArgs {
  DevGraph g;
  DevBatch b;
  std::uint32_t jacobi_copy;            // column holding final Jacobi labels
  std::uint32_t pass;                   // 1: summaries + lengths, 2: emit arenas
  std::uint32_t target_total;
  std::uint32_t stamp_base;             // target t's walk marks onpath with stamp_base + t + 1
  const std::uint32_t* query_clean;     // [Q] converged without error
  const unsigned long long* frame_base; // [Q] frame scratch offset
  std::uint32_t* frame_vertex;
  std::uint32_t* frame_last;
  std::uint32_t* frame_via;
  std::uint32_t* onpath;                // pool-sized stamps
  std::uint32_t* summaries;             // [targets][7]
  unsigned long long* lengths;          // [targets] vertex count of each path
  const unsigned long long* offsets;    // [targets+1] (pass 2)
  std::uint32_t* arena_vertices;
  std::uint32_t* arena_labels;
  std::uint32_t* arena_edges;
  const std::uint32_t* query_list;      // queries to reconstruct (null: all)
  std::uint32_t jacobi_copy_from_control;  // take the column from round parity
  std::uint32_t pad2;
};

// ---- memory helpers -------------------------------------------------------
// Loads of values other workgroups may be lowering concurrently go through
// agent-scope atomic loads (coherent across XCDs); monotone relaxation makes a
// stale value safe, freshness only speeds convergence.
__device__ __forceinline__ std::uint32_t load_agent(const std::uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_agent(std::uint32_t* p, std::uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ std::uint32_t atomic_min_agent(std::uint32_t* p, std::uint32_t v) {
  return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ std::uint32_t atomic_max_agent(std::uint32_t* p, std::uint32_t v) {
  return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ std::uint32_t atomic_add_agent(std::uint32_t* p, std::uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ std::uint32_t atomic_exch_agent(std::uint32_t* p, std::uint32_t v) {
  return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_or_agent(std::uint32_t* p, std::uint32_t v) {
  __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ float as_float(std::uint32_t b) { return __uint_as_float(b); }
__device__ __forceinline__ std::uint32_t as_bits(float f) { return __float_as_uint(f); }

// Queues slot for `round` once: the stamp dedups, the list append claims a
// position, and the query's pending work (when tracked) counts the item.
__device__ __forceinline__ void enqueue(const DevBatch& b, std::uint32_t slot,
                                        std::uint32_t round) {
  if (b.ordered) {
    atomic_or_agent(&b.active[slot >> 5], 1u << (slot & 31u));
    return;
  }
  const std::uint32_t tag = round + 1;
  if (atomic_max_agent(&b.stamp[slot], tag) < tag) {
    const std::uint32_t li = round % 3;
    const std::uint32_t pos = atomic_add_agent(&b.control->list_count[li], 1u);
    b.lists[static_cast<std::size_t>(li) * b.slot_count + pos] = slot;
    if (b.pending) atomic_add_agent(&b.pending[b.slot_query[slot]], 1u);
  }
}

// Wave-wide enqueue: every lane of the wave calls it in uniform control
// flow; lanes with `want` queue `slot` for `round`.  The stamps dedup per
// slot as in enqueue(), and the lanes that queue a slot first share one
// list-position claim per wave (one atomic on the list counter instead of
// one per slot).
__device__ __forceinline__ void enqueue_wave(const DevBatch& b, std::uint32_t slot, bool want,
                                             std::uint32_t round) {
  if (b.ordered) {
    // A bit per slot: setting it twice queues the slot once, and no lane
    // waits for a returned value.
    if (want) atomic_or_agent(&b.active[slot >> 5], 1u << (slot & 31u));
    return;
  }
  const std::uint32_t tag = round + 1;
  const std::uint32_t lane = threadIdx.x & 63u;
  const bool first = want && atomic_max_agent(&b.stamp[slot], tag) < tag;
  const unsigned long long m = __ballot(first);
  if (m == 0) return;
  const std::uint32_t li = round % 3;
  std::uint32_t base = 0;
  if (lane == 0) base = atomic_add_agent(&b.control->list_count[li], static_cast<std::uint32_t>(__popcll(m)));
  base = __shfl(base, 0);
  if (first) {
    const std::uint32_t rank = static_cast<std::uint32_t>(__popcll(m & ((1ull << lane) - 1ull)));
    b.lists[static_cast<std::size_t>(li) * b.slot_count + base + rank] = slot;
    if (b.pending) atomic_add_agent(&b.pending[b.slot_query[slot]], 1u);
  }
}

// Neighbour entries of a visit's LDS table, and the words of LDS one
// relaxation workgroup uses: own + orig labels, the work area, the table and
// the flags.
__host__ __device__ constexpr std::uint32_t lds_table_entries(std::uint32_t max_neighbor_tiles) {
  return (max_neighbor_tiles + 3u) & ~3u;
}
__host__ __device__ constexpr std::size_t relax_lds_words(std::uint32_t max_v, std::uint32_t work_words,
                                                          std::uint32_t max_neighbor_tiles) {
  return 2ull * max_v + work_words + 4ull * lds_table_entries(max_neighbor_tiles) + kLdsFlagWords;
}

// Grid barriers of the persistent kernels (every workgroup of the grid
// calls them the same number of times; the grid is co-resident).  A wait is
// bounded: after kBarrierSpinLimit polls (seconds, far beyond any round) the
// workgroup records kErrorBarrierTimeout and the barrier returns false, so a
// grid that is not co-resident ends the run with a device failure instead
// of holding the GPU.
constexpr std::uint32_t kBarrierSpinLimit = 1u << 23;
constexpr std::uint32_t kProbeSpinLimit = 1u << 16;  // co-residency probe launches
constexpr std::uint32_t kErrorBarrierTimeout = 8u;

__device__ __forceinline__ bool barrier_wait(DevControl* c, std::uint32_t gen, std::uint32_t limit) {
  for (std::uint32_t spins = 0; load_agent(&c->barrier_generation) == gen; ++spins) {
    if (spins == limit) {
      __hip_atomic_fetch_or(&c->error_bits, kErrorBarrierTimeout, __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

// Flat: every workgroup arrives on one counter; the last arriver resets it
// and advances the generation that the others wait on.
__device__ __forceinline__ bool grid_barrier_flat(DevControl* c, std::uint32_t blocks,
                                                  std::uint32_t limit = kBarrierSpinLimit) {
  __syncthreads();
  bool ok = true;
  if (threadIdx.x == 0) {
    const std::uint32_t gen = load_agent(&c->barrier_generation);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const std::uint32_t arrived = atomic_add_agent(&c->barrier_count, 1u);
    if (arrived == blocks - 1) {
      store_agent(&c->barrier_count, 0u);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      store_agent(&c->barrier_generation, gen + 1);
    } else {
      ok = barrier_wait(c, gen, limit);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  return __syncthreads_or(!ok) == 0;
}

// Two-level: a workgroup arrives on its group's counter (group =
// blockIdx % G; G is a multiple of 8, so every group's workgroups sit on the
// XCD that the dispatcher's round-robin placed them on); the last arriver of
// a group resets that counter and arrives on the global one; the last group
// to arrive resets the global counter and advances the generation.  Every
// counter is reset before the generation moves, so the next barrier's
// arrivals always find zeros.  On MI355X one barrier of 1280 workgroups
// costs 21.4 / 20.6 / 20.4 / 20.6 us with G = 8 / 16 / 32 / 64 (31.7 us
// with one counter): the arrivals are not the part that grows with the
// grid, and G = 32 is the measured best.
template <std::uint32_t G = kArrivalGroups>
__device__ __forceinline__ bool grid_barrier_grouped(DevControl* c, std::uint32_t blocks,
                                                     std::uint32_t limit = kBarrierSpinLimit) {
  static_assert(G % kBarrierGroups == 0 && G <= kMaxArrivalGroups, "arrival groups");
  __syncthreads();
  bool ok = true;
  if (threadIdx.x == 0) {
    const std::uint32_t gen = load_agent(&c->barrier_generation);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const std::uint32_t g = blockIdx.x % G;
    const std::uint32_t members = (blocks + G - 1u - g) / G;
    const std::uint32_t groups = blocks < G ? blocks : G;
    std::uint32_t* mine = &c->group_arrivals[g * kBarrierStride];
    bool last = false;
    if (atomic_add_agent(mine, 1u) == members - 1) {
      store_agent(mine, 0u);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      last = atomic_add_agent(&c->barrier_count, 1u) == groups - 1;
    }
    if (last) {
      store_agent(&c->barrier_count, 0u);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      store_agent(&c->barrier_generation, gen + 1);
    } else {
      ok = barrier_wait(c, gen, limit);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  return __syncthreads_or(!ok) == 0;
}

// Position of key in the sorted range values[begin, end), or kNone.
__device__ __forceinline__ std::uint32_t find_sorted(const std::uint32_t* values,
                                                     std::uint32_t begin, std::uint32_t end,
                                                     std::uint32_t key) {
  std::uint32_t lo = begin, hi = end;
  while (lo < hi) {
    std::uint32_t mid = (lo + hi) >> 1;
    if (values[mid] < key) lo = mid + 1; else hi = mid;
  }
  return (lo < end && values[lo] == key) ? lo : kNone;
}

// Slot at position pos of a query's selected tiles.
__device__ __forceinline__ std::uint32_t slot_at(const DevBatch& b, std::uint32_t pos) {
  return b.query_slot ? b.query_slot[pos] : pos;
}

// The visit record of slot s, derived from the batch and graph tables.
__device__ __forceinline__ SlotDesc make_slot_desc(const DevGraph& g, const DevBatch& b,
                                                   std::uint32_t s) {
  SlotDesc d;
  d.tile = b.tiles[s];
  d.query = b.slot_query[s];
  d.engine = b.engine[d.query];
  d.v0 = g.tile_vertex_offsets[d.tile];
  d.npt = g.tile_vertex_offsets[d.tile + 1] - d.v0;
  const bool pull = d.engine == kJacobiPull;
  const std::uint32_t* nbr = pull ? g.in_nbr_offsets : g.out_nbr_offsets;
  d.nbr0 = nbr[d.tile];
  d.nbr_count = nbr[d.tile + 1] - d.nbr0;
  d.nbr_base = pull ? b.slot_in_base[s] : b.slot_out_base[s];
  const std::uint32_t* off = pull ? g.column_offsets : g.row_offsets;
  d.e0 = off[d.v0];
  d.e1 = off[d.v0 + d.npt];
  d.ei = d.e0 + (pull ? g.csc_intra : g.csr_intra)[d.tile];
  const unsigned long long dist = b.slot_dist[s];
  d.dist_lo = static_cast<std::uint32_t>(dist);
  d.dist_hi = static_cast<std::uint32_t>(dist >> 32);
  d.out_nbr0 = g.out_nbr_offsets[d.tile];
  d.out_count = g.out_nbr_offsets[d.tile + 1] - d.out_nbr0;
  d.out_base = b.slot_out_base[s];
  return d;
}

__device__ __forceinline__ SlotDesc load_slot_desc(const DevGraph& g, const DevBatch& b,
                                                   std::uint32_t s) {
  return b.slot_desc ? b.slot_desc[s] : make_slot_desc(g, b, s);
}

// Slot of query q on tile, or kNone when q did not select the tile.
__device__ __forceinline__ std::uint32_t find_slot(const DevBatch& b, std::uint32_t q,
                                                   std::uint32_t tile) {
  const std::uint32_t pos = find_sorted(b.query_tiles, b.tile_begin[q], b.tile_end[q], tile);
  return pos == kNone ? kNone : slot_at(b, pos);
}

// Push window bookkeeping between rounds.  The limit of round r + 1 is the
// smallest priority round r left queued plus the window (never below the
// limit of round r), so the slot holding that priority is always visited
// in round r + 1 and every queued slot is visited eventually.
__device__ __forceinline__ std::uint32_t next_window_limit(DevControl* c, const DevBatch& b,
                                                           std::uint32_t r, std::uint32_t limit) {
  std::uint32_t m = kInfBits;
  for (std::uint32_t x = 0; x < kBarrierGroups; ++x)
    m = min(m, load_agent(&c->next_min[(((r + 1) % 3) * kBarrierGroups + x) * kBarrierStride]));
  if (m == kInfBits) return limit;
  return max(limit, as_bits(as_float(m) + as_float(b.window)));
}
__device__ __forceinline__ void reset_next_min(DevControl* c, std::uint32_t buffer) {
  for (std::uint32_t x = 0; x < kBarrierGroups; ++x)
    store_agent(&c->next_min[(buffer * kBarrierGroups + x) * kBarrierStride], kInfBits);
}

}  // namespace bfn
