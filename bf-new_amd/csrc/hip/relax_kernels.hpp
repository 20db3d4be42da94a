// relax_kernels.hpp -- the relaxation visits, the round loop and the
// relaxation kernels (persistent and one round per launch) as templates over
// block size, records per thread, instrumentation, wave bound and engine
// set; each relax_<engines>.hip instantiates one engine set, so the variants
// compile in parallel.
//
// Work unit: a slot s = (query q, selected tile T).  One workgroup visits one
// slot at a time.  The tile's edge records (pack, weight) are read once per
// visit, coalesced: the internal records (both ends in T) into registers,
// where they stay for the tile closure (thread t holds records t, t+B, ...),
// the boundary records streamed in chunks.  The distances the visit touches
// are staged in LDS and relaxed there with ds_min_u32 on the order-preserving
// u32 image of the fp32 distance (all distances are +0, finite positive or
// +inf, so unsigned order equals float order).
//
// Semantics (reference engines):
//   Jacobi pull   batched_jacobi.hip.cpp:466-694 -- two columns, round r reads
//                 column r&1 and writes column (r+1)&1:
//                 d'[v] = min(d[v], min over admitted in-edges fl(d[u] + w)).
//                 tile_sweeps = 1 is exactly that (same rounds as the
//                 reference); tile_sweeps = 0 additionally closes the tile's
//                 internal edges in LDS each round (block Jacobi: the same
//                 fixed point, fewer rounds).
//   dense push    batched_dense_chaotic_push.hip.cpp:392-627 -- one column,
//                 in place, atomicMin(d[dst], fl(d[src] + w)) over every
//                 admitted out-edge of a dirty tile.
//   frontier push batched_frontier_push.hip.cpp:710-906 -- as dense, but
//                 only from vertices whose label dropped below the label last
//                 pushed (column 1 holds that snapshot).
// Every schedule reaches the same bits: fl(x + w) is monotone in x, so the
// relaxation operator has one greatest fixed point.
//
// Admission: a neighbour tile that the query did not select is staged as
// +inf (pull) or has no slot (push), so edges from or to it never relax.
#pragma once

#include "bfn_internal.hpp"

namespace bfn {

// LDS carve-up (u32 words): [own | orig | work | table | flags], with
// `work` = staged neighbour slabs (pull), or the pushed snapshot followed by
// the boundary proposal slabs (push), and `table` = per-neighbour LDS offsets
// (pull) or neighbour slots, "pushed into" marks and proposal slab offsets
// (push).
struct Lds {
  std::uint32_t* own;
  std::uint32_t* orig;
  std::uint32_t* work;
  std::uint32_t* prop;   // push, proposals: the boundary proposal slabs
  std::uint32_t* bnd;    // push, direct: boundary record start per vertex
  std::uint32_t* list;   // push, direct: the visit's pushing vertices
  std::uint32_t* slots;  // per neighbour: its slot (push) / staged slab offset (pull)
  std::uint32_t* marks;  // per neighbour: pushed into during this visit
  std::uint32_t* offs;   // per neighbour: pool offset (two words) / proposal slab offset
  std::uint32_t* flags;
};

__device__ __forceinline__ Lds carve(std::uint32_t* base, std::uint32_t max_v,
                                     std::uint32_t work_words, std::uint32_t max_neighbor_tiles) {
  const std::uint32_t k = lds_table_entries(max_neighbor_tiles);
  Lds l;
  l.own = base;
  l.orig = base + max_v;
  l.work = base + 2 * max_v;
  l.prop = l.work + max_v;
  l.bnd = l.work + max_v;
  l.list = l.work + 2 * max_v;
  l.slots = l.work + work_words;
  l.marks = l.slots + k;
  l.offs = l.marks + k;
  l.flags = l.offs + 2 * k;
  return l;
}

// Boundary records per pushing vertex loaded together (one load, then one
// filter read, in flight per record).  Six covers a vertex of the tiled
// routing graphs (four intra-tile and six inter-tile edges per node) in one
// chunk, and at six the 256-thread push-only kernel fits 79 VGPRs: six
// workgroups per CU, where eight records per chunk needed 85 (five).
constexpr int kPushChunk = 6;
static_assert(kPushChunk % 2 == 0, "boundary records are loaded in pairs");

// Edge records in registers: thread t holds records t, t + B, t + 2B, ... of
// the range it loaded.
template <int B, int N>
struct RecordRegs {
  std::uint32_t pack[N];
  float w[N];
};

// Loads records [base, base + N*B) clipped to `end` (kNone past the end).
template <int B, int N>
__device__ __forceinline__ void load_records(RecordRegs<B, N>& r, const uint2* rec,
                                             std::uint32_t base, std::uint32_t end) {
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const std::uint32_t e = base + j * B + threadIdx.x;
    if (e < end) {
      const uint2 x = rec[e];
      r.pack[j] = x.x;
      r.w[j] = __uint_as_float(x.y);
    } else {
      r.pack[j] = kNone;
      r.w[j] = 0.0f;
    }
  }
}

// Boundary records per thread per streamed chunk (loads in flight per lane):
// more for the large blocks, fewer where registers bound the occupancy.
constexpr int stream_records(int block) { return block >= 512 ? 8 : 4; }

// Per-workgroup counters of the visits, compiled in for instrumented runs
// only (I = true; otherwise every counter is dead code).  Phase clocks are
// shader-clock cycles read by thread 0 at phase boundaries: 0 setup (tables,
// label staging), 1 internal records (pull: first pass), 2 tile closure,
// 3 boundary records, 4 write-back and enqueue.
template <bool I>
struct VisitStats {
  unsigned long long attempts = 0, successes = 0, bytes = 0, edges = 0, items = 0, sweeps = 0,
                     evals = 0;
  unsigned long long phase[kPhaseCount] = {};
  long long t = 0;
  __device__ __forceinline__ void start() {
    if constexpr (I) {
      if (threadIdx.x == 0) t = clock64();
    }
  }
  __device__ __forceinline__ void mark(int p) {
    if constexpr (I) {
      if (threadIdx.x == 0) {
        const long long now = clock64();
        phase[p] += static_cast<unsigned long long>(now - t);
        t = now;
      }
    }
  }
};

__device__ __forceinline__ std::uint32_t pk_k(std::uint32_t p) { return p >> 24; }
__device__ __forceinline__ std::uint32_t pk_src(std::uint32_t p) { return (p >> 12) & kPackLocalMask; }
__device__ __forceinline__ std::uint32_t pk_dst(std::uint32_t p) { return p & kPackLocalMask; }

// ---------------------------------------------------------------------------
// Jacobi pull visit.
// ---------------------------------------------------------------------------
template <int B, int EPT, bool I>
__device__ __forceinline__ void visit_pull(const DevGraph& g, const DevBatch& b, const SlotDesc& d,
                           std::uint32_t r, const Lds& l, VisitStats<I>& st) {
  const std::uint32_t tid = threadIdx.x;
  st.start();
  const std::uint32_t q = d.query;
  const std::uint32_t npt = d.npt;
  const std::uint32_t nb0 = d.nbr0;
  const std::uint32_t K = d.nbr_count;
  const std::uint32_t rp = r & 1u;
  const std::uint32_t in_base = d.nbr_base;
  const unsigned long long dist = (static_cast<unsigned long long>(d.dist_hi) << 32) | d.dist_lo;

  for (std::uint32_t k = tid; k < K; k += B) l.slots[k] = g.in_nbr_stage[nb0 + k];
  __syncthreads();
  // Stage the read column of every in-neighbour slab (k = 0 is T itself).
  for (std::uint32_t k = 0; k < K; ++k) {
    const std::uint32_t nt = g.in_nbr_tiles[nb0 + k];
    const std::uint32_t n = g.tile_vertex_offsets[nt + 1] - g.tile_vertex_offsets[nt];
    const std::uint32_t ns = b.slot_in[in_base + k];
    std::uint32_t* dst = l.work + l.slots[k];
    if (ns == kNone) {
      for (std::uint32_t i = tid; i < n; i += B) dst[i] = kInfBits;
    } else {
      if (tid == 0) st.bytes += 4ull * n;
      const std::uint32_t* src = b.dist + b.slot_dist[ns] + static_cast<std::size_t>(rp) * n;
      for (std::uint32_t i = tid; i < n; i += B) dst[i] = src[i];
    }
  }
  __syncthreads();
  for (std::uint32_t i = tid; i < npt; i += B) l.own[i] = l.work[i];
  __syncthreads();
  st.mark(0);

  // CSC block of T: internal records [e0, ei), boundary records [ei, e1).
  const std::uint32_t e0 = d.e0;
  const std::uint32_t e1 = d.e1;
  const std::uint32_t ei = d.ei;
  const bool resident = ei - e0 <= static_cast<std::uint32_t>(B * EPT);
  if (tid == 0) {
    // edge records (pack + weight) once, the written column once
    st.bytes += 8ull * (e1 - e0) + 4ull * npt;
    st.edges += e1 - e0;
  }
  // Pass 1: every admitted in-edge reads the staged read column.
  auto relax_staged = [&](std::uint32_t p, float w) {
    if constexpr (I) ++st.evals;
    const std::uint32_t cb = as_bits(as_float(l.work[l.slots[pk_k(p)] + pk_src(p)]) + w);
    if (cb < l.own[pk_dst(p)]) atomicMin(&l.own[pk_dst(p)], cb);
  };
  // The internal records: loaded once here, and kept in registers through
  // the tile closure when they fit (resident); not kept across visits, so
  // they hold no registers outside this visit.
  RecordRegs<B, EPT> internal;
  for (std::uint32_t base = e0; base < ei; base += B * EPT) {
    load_records<B, EPT>(internal, g.csc_rec, base, ei);
#pragma unroll
    for (int j = 0; j < EPT; ++j)
      if (internal.pack[j] != kNone) relax_staged(internal.pack[j], internal.w[j]);
  }
  st.mark(1);
  for (std::uint32_t base = ei; base < e1; base += B * stream_records(B)) {
    RecordRegs<B, stream_records(B)> chunk;
    load_records<B, stream_records(B)>(chunk, g.csc_rec, base, e1);
#pragma unroll
    for (int j = 0; j < stream_records(B); ++j)
      if (chunk.pack[j] != kNone) relax_staged(chunk.pack[j], chunk.w[j]);
  }
  __syncthreads();
  st.mark(3);
  if (b.tile_sweeps == 0) {
    // Close the tile's internal edges in place; the internal records stay in
    // registers when they fit.
    for (;;) {
      bool changed = false;
      for (std::uint32_t base = e0; base < ei; base += B * EPT) {
        if (!resident) load_records<B, EPT>(internal, g.csc_rec, base, ei);
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
          const std::uint32_t p = internal.pack[j];
          if (p == kNone) continue;
          if constexpr (I) ++st.evals;
          const std::uint32_t cb = as_bits(as_float(l.own[pk_src(p)]) + internal.w[j]);
          if (cb < l.own[pk_dst(p)]) {
            atomicMin(&l.own[pk_dst(p)], cb);
            changed = true;
          }
        }
      }
      ++st.sweeps;
      if (!__syncthreads_or(changed)) break;
    }
  }
  st.mark(2);

  std::uint32_t* out = b.dist + dist + static_cast<std::size_t>(rp ^ 1u) * npt;
  bool changed = false;
  for (std::uint32_t i = tid; i < npt; i += B) {
    const std::uint32_t v = l.own[i];
    out[i] = v;
    changed |= (v != l.work[i]);
  }
  if (__syncthreads_or(changed)) {
    if (tid == 0) atomic_max_agent(&b.last_change[q], r + 1);
    // every selected out-neighbour (the tile itself first) reads the new
    // column next round
    for (std::uint32_t j0 = 0; j0 < d.out_count; j0 += B) {
      const std::uint32_t j = j0 + tid;
      const std::uint32_t ns = j < d.out_count ? b.slot_out[d.out_base + j] : kNone;
      enqueue_wave(b, ns == kNone ? 0u : ns, ns != kNone, r + 1);
    }
  }
  st.mark(4);
}

// ---------------------------------------------------------------------------
// Dense / frontier push visit.
// ---------------------------------------------------------------------------
template <int B, int EPT, bool I>
__device__ __forceinline__ void visit_push(const DevGraph& g, const DevBatch& b, const SlotDesc& d,
                           std::uint32_t r, bool frontier, const Lds& l,
                           VisitStats<I>& st, std::uint32_t& queued_min) {
  const std::uint32_t tid = threadIdx.x;
  st.start();
  const std::uint32_t q = d.query;
  const std::uint32_t v0 = d.v0;
  const std::uint32_t npt = d.npt;
  const std::uint32_t ob0 = d.nbr0;
  const std::uint32_t KO = d.nbr_count;
  const std::uint32_t out_base = d.nbr_base;
  std::uint32_t* column = b.dist + ((static_cast<unsigned long long>(d.dist_hi) << 32) | d.dist_lo);
  std::uint32_t* snapshot = column + npt;
  // CSR block of T: internal records [e0, ei), boundary records [ei, e1).
  const std::uint32_t e0 = d.e0;
  const std::uint32_t e1 = d.e1;
  const std::uint32_t ei = d.ei;
  const bool resident = ei - e0 <= static_cast<std::uint32_t>(B * EPT);
  // The internal records are requested first, so that they travel together
  // with the label loads below (one round trip for both).  They live only
  // through the tile closure: kept in registers across visits of the same
  // tile they would stay live through the boundary phase too, and those 16
  // VGPRs cost the 256-thread push-only kernel two of its eight workgroups
  // per CU.
  RecordRegs<B, EPT> internal;
  if (resident) load_records<B, EPT>(internal, g.csr_rec, e0, ei);

  const bool proposals = b.boundary_mode == 0;
  for (std::uint32_t k = tid; k < KO; k += B) {
    const std::uint32_t ns = b.slot_out[out_base + k];
    l.slots[k] = ns;
    l.marks[k] = kInfBits;
    if (proposals) {
      l.offs[k] = g.out_nbr_stage[ob0 + k];
    } else {
      // pool offset of the neighbour's labels, as two words
      const unsigned long long off =
          b.slot_out_off ? b.slot_out_off[out_base + k] : (ns != kNone ? b.slot_dist[ns] : ~0ull);
      l.offs[2 * k] = static_cast<std::uint32_t>(off);
      l.offs[2 * k + 1] = static_cast<std::uint32_t>(off >> 32);
    }
  }
  // Dense pushes every finite vertex: an infinite snapshot admits them all.
  if (((npt | v0) & 1u) == 0) {
    // Even tile size and first vertex: labels, snapshot and record starts
    // two per 8-byte load (pool offsets are sums of 2 npt, so all even).
    for (std::uint32_t i = 2 * tid; i < npt; i += 2 * B) {
      const unsigned long long x = __hip_atomic_load(
          reinterpret_cast<const unsigned long long*>(column + i), __ATOMIC_RELAXED,
          __HIP_MEMORY_SCOPE_AGENT);
      l.own[i] = l.orig[i] = static_cast<std::uint32_t>(x);
      l.own[i + 1] = l.orig[i + 1] = static_cast<std::uint32_t>(x >> 32);
      const uint2 sn = frontier ? *reinterpret_cast<const uint2*>(snapshot + i)
                                : make_uint2(kInfBits, kInfBits);
      l.work[i] = sn.x;
      l.work[i + 1] = sn.y;
      if (!proposals) {
        const uint2 bd = *reinterpret_cast<const uint2*>(g.csr_bnd + v0 + i);
        l.bnd[i] = bd.x;
        l.bnd[i + 1] = bd.y;
      }
    }
  } else {
    for (std::uint32_t i = tid; i < npt; i += B) {
      const std::uint32_t x = load_agent(column + i);
      l.own[i] = x;
      l.orig[i] = x;
      l.work[i] = frontier ? snapshot[i] : kInfBits;
      if (!proposals) l.bnd[i] = g.csr_bnd[v0 + i];
    }
  }
  if (tid == 0) l.flags[1] = 0;
  if (proposals) {
    // Clear the proposal slab of every selected out-neighbour (k >= 1).
    for (std::uint32_t k = 1; k < KO; ++k) {
      if (b.slot_out[out_base + k] == kNone) continue;
      const std::uint32_t nt = g.out_nbr_tiles[ob0 + k];
      const std::uint32_t n = g.tile_vertex_offsets[nt + 1] - g.tile_vertex_offsets[nt];
      std::uint32_t* slab = l.prop + g.out_nbr_stage[ob0 + k];
      for (std::uint32_t i = tid; i < n; i += B) slab[i] = kInfBits;
    }
  }
  __syncthreads();
  st.mark(0);

  if (tid == 0) {
    // internal records once, the column (and the snapshot for frontier)
    // once; proposals stream every boundary record, direct mode reads the
    // record starts and then only the pushing vertices' records
    st.bytes += 8ull * (ei - e0) + 4ull * npt * (frontier ? 2 : 1) +
                (proposals ? 8ull * (e1 - ei) : 4ull * npt);
    st.edges += (ei - e0) + (proposals ? e1 - ei : 0u);
  }
  st.mark(1);
  // Internal edges to the tile's local fixed point, from frontier sources.
  for (;;) {
    bool changed = false;
    for (std::uint32_t base = e0; base < ei; base += B * EPT) {
      if (!resident) load_records<B, EPT>(internal, g.csr_rec, base, ei);
#pragma unroll
      for (int j = 0; j < EPT; ++j) {
        const std::uint32_t p = internal.pack[j];
        if (p == kNone) continue;
        const std::uint32_t sv = l.own[pk_src(p)];
        if (!(sv < l.work[pk_src(p)])) continue;
        if constexpr (I) ++st.evals;
        const std::uint32_t cb = as_bits(as_float(sv) + internal.w[j]);
        if (cb < l.own[pk_dst(p)]) {
          atomicMin(&l.own[pk_dst(p)], cb);
          changed = true;
        }
      }
    }
    ++st.sweeps;
    if (!__syncthreads_or(changed)) break;
  }
  st.mark(2);
  // Boundary records, streamed once.  Labels only fall during a run, so an
  // ordinary (cacheable, possibly stale) load of a neighbour label is never
  // below the current label: comparing against it never drops a decrease, at
  // worst it issues an atomic or marks a neighbour that did not need it.
  unsigned long long att = 0, suc = 0;
  if (proposals) {
    // Gather the best candidate per destination in LDS, then compare each
    // neighbour slab once, coalesced, and issue one atomic per improvement.
    for (std::uint32_t base = ei; base < e1; base += B * stream_records(B)) {
      RecordRegs<B, stream_records(B)> chunk;
      load_records<B, stream_records(B)>(chunk, g.csr_rec, base, e1);
#pragma unroll
      for (int j = 0; j < stream_records(B); ++j) {
        const std::uint32_t p = chunk.pack[j];
        if (p == kNone) continue;
        const std::uint32_t sv = l.own[pk_src(p)];
        if (!(sv < l.work[pk_src(p)])) continue;
        if (l.slots[pk_k(p)] == kNone) continue;
        const std::uint32_t cb = as_bits(as_float(sv) + chunk.w[j]);
        std::uint32_t* slot = &l.prop[l.offs[pk_k(p)] + pk_dst(p)];
        ++att;
        if (cb < *slot) atomicMin(slot, cb);
      }
    }
    __syncthreads();
    for (std::uint32_t k = 1; k < KO; ++k) {
      const std::uint32_t ns = l.slots[k];
      if (ns == kNone) continue;
      const std::uint32_t nt = g.out_nbr_tiles[ob0 + k];
      const std::uint32_t n = g.tile_vertex_offsets[nt + 1] - g.tile_vertex_offsets[nt];
      const std::uint32_t* slab = l.prop + l.offs[k];
      std::uint32_t* target = b.dist + b.slot_dist[ns];
      if (tid == 0) st.bytes += 4ull * n;  // the neighbour slab, read once
      std::uint32_t low = kInfBits;
      for (std::uint32_t i = tid; i < n; i += B) {
        const std::uint32_t pv = slab[i];
        if (pv != kInfBits && pv < target[i]) {
          atomic_min_agent(target + i, pv);
          ++suc;
          low = min(low, pv);
        }
      }
      if (low != kInfBits) atomicMin(&l.marks[k], low);
    }
  } else {
    // Direct mode: only vertices that push (label below the snapshot;
    // dense: finite) read their boundary records.  They are listed in LDS
    // (one claim per wave) and taken one vertex per thread; a vertex's
    // boundary records are consecutive, loaded kPushChunk at a time, then
    // their filter reads are issued together, then one no-return atomic per
    // candidate that beats the label read.  The filter reads are ordinary
    // (cacheable, possibly stale) loads: labels only fall, so a stale label
    // is never below the true one and the filter never drops a decrease; a
    // neighbour is marked when its read label was above a candidate, which
    // covers every actual decrease.  (A returning atomic per candidate in
    // place of the read measured 1.6x slower on 128x128x512 tiles: device-
    // scope atomics are performed past the L2, at the memory side.)
    for (std::uint32_t base = 0; base < npt; base += B) {
      const std::uint32_t i = base + tid;
      const bool push = i < npt && l.own[i] < l.work[i];
      const unsigned long long m = __ballot(push);
      std::uint32_t at = 0;
      if ((tid & 63u) == 0 && m) at = atomicAdd(&l.flags[1], static_cast<std::uint32_t>(__popcll(m)));
      at = __shfl(at, 0);
      if (push) l.list[at + __popcll(m & ((1ull << (tid & 63u)) - 1ull))] = i;
    }
    __syncthreads();
    const std::uint32_t n_push = l.flags[1];
    unsigned long long loaded = 0;
    for (std::uint32_t j = tid; j < n_push; j += B) {
      const std::uint32_t i = l.list[j];
      const float sf = as_float(l.own[i]);
      const std::uint32_t rb = l.bnd[i];
      const std::uint32_t re = i + 1 < npt ? l.bnd[i + 1] : e1;
      loaded += re - rb;
      // Chunks start at even positions, so a lane's records arrive two per
      // 16-byte load (the stream holds one record to spare past its end).
      for (std::uint32_t r0 = rb & ~1u; r0 < re; r0 += kPushChunk) {
        std::uint32_t pk[kPushChunk];
        float w[kPushChunk];
#pragma unroll
        for (int c = 0; c < kPushChunk; c += 2) {
          uint4 x = make_uint4(kNone, 0u, kNone, 0u);
          if (r0 + c < re) x = *reinterpret_cast<const uint4*>(g.csr_rec + r0 + c);
          const bool in0 = r0 + c >= rb && r0 + c < re;
          const bool in1 = r0 + c + 1 >= rb && r0 + c + 1 < re;
          pk[c] = in0 ? x.x : kNone;
          w[c] = in0 ? as_float(x.y) : 0.0f;
          pk[c + 1] = in1 ? x.z : kNone;
          w[c + 1] = in1 ? as_float(x.w) : 0.0f;
        }
        auto target = [&](std::uint32_t p) -> std::uint32_t* {
          const unsigned long long off = (static_cast<unsigned long long>(l.offs[2 * pk_k(p) + 1]) << 32) |
                                         l.offs[2 * pk_k(p)];
          return b.dist + off + pk_dst(p);
        };
        std::uint32_t seen[kPushChunk];
#pragma unroll
        for (int c = 0; c < kPushChunk; ++c) {
          if (pk[c] != kNone && l.slots[pk_k(pk[c])] == kNone) pk[c] = kNone;  // not selected
          seen[c] = pk[c] != kNone ? *target(pk[c]) : 0u;
        }
#pragma unroll
        for (int c = 0; c < kPushChunk; ++c) {
          if (pk[c] == kNone) continue;
          ++att;
          const std::uint32_t cb = as_bits(sf + w[c]);
          if (cb < seen[c]) {
            atomic_min_agent(target(pk[c]), cb);
            ++suc;
            // read first: most candidates into one neighbour lose to its mark
            if (cb < l.marks[pk_k(pk[c])]) atomicMin(&l.marks[pk_k(pk[c])], cb);
          }
        }
      }
    }
    if (loaded) {
      st.edges += loaded;  // per-thread counters, summed at the end of the round
      st.bytes += 8ull * loaded;
    }
  }
  st.attempts += att;
  st.successes += suc;
  st.evals += att;
  // direct: a remote label read per attempt; both: read-modify-write per
  // atomic issued (proposal slabs read once, counted above)
  st.bytes += (proposals ? 0ull : 4ull * att) + 8ull * suc;
  st.mark(3);
  bool changed = false;
  for (std::uint32_t i = tid; i < npt; i += B) {
    const std::uint32_t v = l.own[i];
    if (v < l.orig[i]) {
      atomic_min_agent(column + i, v);
      changed = true;
    }
    if (frontier && v < l.work[i]) snapshot[i] = v;
  }
  const bool pushed_any = __syncthreads_or(changed);
  // A marked neighbour holds the smallest candidate pushed into it; with a
  // push window that candidate lowers the neighbour's priority.
  bool marked = false;
  for (std::uint32_t k0 = 1; k0 < KO; k0 += B) {
    const std::uint32_t k = k0 + tid;
    const std::uint32_t low = k < KO ? l.marks[k] : kInfBits;
    const bool want = low != kInfBits;
    marked |= want;
    if (want && b.prio) {
      atomic_min_agent(&b.prio[l.slots[k]], low);
      queued_min = min(queued_min, low);
    }
    enqueue_wave(b, want ? l.slots[k] : 0u, want, r + 1);
  }
  if (__syncthreads_or(marked) || pushed_any) {
    if (tid == 0) atomic_max_agent(&b.last_change[q], r + 1);
  }
  st.mark(4);
}

// ---------------------------------------------------------------------------
// Round processing shared by the persistent and host-polled kernels.
// ---------------------------------------------------------------------------
// Round r's share of one workgroup.  With a push window (b.prio set), a
// push slot whose priority is above `limit` is not visited: it keeps its
// priority and is queued again for round r + 1.  The smallest priority the
// workgroup leaves queued (deferred or pushed into) goes to the round's
// next_min counter of its XCD group.
template <int B, int EPT, bool I, int M>
__device__ __forceinline__ void process_round(const DevGraph& g, const DevBatch& b, std::uint32_t r,
                              std::uint32_t n, std::uint32_t limit, const Lds& l) {
  const std::uint32_t chunk = b.items_per_claim ? b.items_per_claim : 1u;
  VisitStats<I> st;
  std::uint32_t queued_min = kInfBits;
  if (threadIdx.x == 0) l.flags[4] = kInfBits;
  auto visit = [&](std::uint32_t s) {
    const SlotDesc d = load_slot_desc(g, b, s);
    bool run = true;
    // A variant without the slot's engine cannot visit it; prep_slots_kernel
    // flags such a batch (the host picks the variant from the batch's
    // engines, so it does not happen).  Nothing is done here: any extra
    // global operation on this path costs the variant 18 VGPRs.
    if (d.engine == kJacobiPull) {
      if constexpr (M != kEnginesPush) visit_pull<B, EPT, I>(g, b, d, r, l, st);
    } else if constexpr (M != kEnginesPull) {
      if (b.prio) {
        if (threadIdx.x == 0) {
          const std::uint32_t p = atomic_exch_agent(&b.prio[s], kInfBits);
          const bool defer = p != kInfBits && p > limit;
          if (defer) {
            atomic_min_agent(&b.prio[s], p);
            enqueue(b, s, r + 1);
            queued_min = min(queued_min, p);
          }
          l.flags[3] = defer ? 1u : 0u;
        }
        __syncthreads();
        run = l.flags[3] == 0u;
      }
      if (run)
        visit_push<B, EPT, I>(g, b, d, r, d.engine == kFrontierPush, l, st, queued_min);
    }
    if (run) ++st.items;
    __syncthreads();
    // the visit's own enqueues are issued before its item is retired
    if (b.pending && threadIdx.x == 0) atomic_add_agent(&b.pending[d.query], 0xFFFFFFFFu);
  };
  if (b.ordered) {
    // Slot order: the round's list is cut into one segment per XCD group;
    // a workgroup claims from its own group's segment (blockIdx % 8: the
    // dispatcher places workgroups on the XCDs round-robin), then from the
    // others in turn.  Consecutive slots share a tile, so the slots of one
    // tile are visited by one XCD, close together in time, and a chunk of
    // them by one workgroup (the tile's records stay in that XCD's L2).
    // Thread 0 keeps the next claim in flight while a chunk is visited.
    DevControl* c = b.control;
    const std::uint32_t home = blockIdx.x % kBarrierGroups;
    std::uint32_t tried = 0;  // thread 0: segments exhausted so far
    std::uint32_t ahead = 0;  // thread 0: claim in flight on segment home + tried
    // Segment ends, read once into LDS (flags 8 .. 15): the claim counters
    // sit on lines of their own, which no plain load shares.
    std::uint32_t* seg_end = l.flags + 8;
    if (threadIdx.x == 0) {
      for (std::uint32_t x = 0; x < kBarrierGroups; ++x) seg_end[x] = c->seg_begin[x + 1];
      ahead = atomic_add_agent(&c->seg_claim[home * kBarrierStride], chunk);
    }
    for (;;) {
      if (threadIdx.x == 0) {
        std::uint32_t start = kNone, end = 0;
        while (tried < kBarrierGroups) {
          const std::uint32_t y = (home + tried) % kBarrierGroups;
          const std::uint32_t e = seg_end[y];
          if (ahead < e) {
            start = ahead;
            end = min(ahead + chunk, e);
            ahead = atomic_add_agent(&c->seg_claim[y * kBarrierStride], chunk);
            break;
          }
          if (++tried < kBarrierGroups)
            ahead = atomic_add_agent(&c->seg_claim[((home + tried) % kBarrierGroups) * kBarrierStride],
                                     chunk);
        }
        l.flags[0] = start;
        l.flags[2] = end;
      }
      __syncthreads();
      const std::uint32_t start = l.flags[0], end = l.flags[2];
      __syncthreads();
      if (start == kNone) break;
      for (std::uint32_t i = start; i < end; ++i) visit(b.lists[i]);
    }
  } else {
    const std::uint32_t li = r % 3;
    // The first ceil(n / chunk) workgroups claim until the list is exhausted;
    // the others would only find it exhausted, and skip the claim counter
    // (in a small round most of the grid would otherwise queue atomics on
    // its one line).
    const bool claims = static_cast<unsigned long long>(blockIdx.x) * chunk < n;
    for (; claims;) {
      if (threadIdx.x == 0) l.flags[0] = atomic_add_agent(&b.control->claim[li], chunk);
      __syncthreads();
      const std::uint32_t start = l.flags[0];
      __syncthreads();
      if (start >= n) break;
      const std::uint32_t end = min(start + chunk, n);
      for (std::uint32_t i = start; i < end; ++i)
        visit(b.lists[static_cast<std::size_t>(li) * b.slot_count + i]);
    }
  }
  if (b.prio) {
    if (queued_min != kInfBits) atomicMin(&l.flags[4], queued_min);
    __syncthreads();
    if (threadIdx.x == 0 && l.flags[4] != kInfBits)
      atomic_min_agent(&b.control->next_min[(((r + 1) % 3) * kBarrierGroups + blockIdx.x % kBarrierGroups) *
                                            kBarrierStride],
                       l.flags[4]);
  }
  if constexpr (I) {
    // One add per thread keeps the counters off the hot path's critical lanes.
    DevControl* c = b.control;
    // Byte and record counts are per thread: tile-level terms are added by
    // thread 0 of the visit, per-record terms by the thread that did them.
    if (st.attempts) atomicAdd(&c->atomic_attempts, st.attempts);
    if (st.successes) atomicAdd(&c->atomic_successes, st.successes);
    if (st.bytes) atomicAdd(&c->algorithmic_bytes, st.bytes);
    if (st.edges) atomicAdd(&c->edges_relaxed, st.edges);
    if (st.evals) atomicAdd(&c->candidates, st.evals);
    if (threadIdx.x == 0) {
      if (st.items) atomicAdd(&c->items_processed, st.items);
      if (st.sweeps) atomicAdd(&c->closure_sweeps, st.sweeps);
      for (int p = 0; p < kPhaseCount; ++p)
        if (st.phase[p]) atomicAdd(&c->phase_cycles[p], st.phase[p]);
    }
  }
}

// One cooperative launch runs every round: process list r%3, grid barrier,
// repeat while the next list is non-empty and the round limit allows.
template <int B, int EPT, bool I, int MW, int M>
__global__ __launch_bounds__(B, MW) void relax_persistent_kernel(DevGraph g, DevBatch b,
                                                             std::uint32_t max_v,
                                                             std::uint32_t work_words) {
  extern __shared__ std::uint32_t lds_words[];
  const Lds l = carve(lds_words, max_v, work_words, g.max_neighbor_tiles);
  DevControl* c = b.control;
  if (b.probe) {  // co-residency probe: one short-bounded grid barrier
    grid_barrier_grouped(c, gridDim.x, kProbeSpinLimit);
    return;
  }
  std::uint32_t r = load_agent(&c->round);
  if (load_agent(&c->done)) return;
  std::uint32_t n = 0;
  std::uint32_t limit = b.prio ? load_agent(&c->window_limit) : 0u;
  for (;;) {
    n = load_agent(&c->list_count[r % 3]);
    if (n == 0 || (b.maximum_rounds && r >= b.maximum_rounds)) break;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      store_agent(&c->list_count[(r + 2) % 3], 0u);
      store_agent(&c->claim[(r + 2) % 3], 0u);
      // read after the barrier of round r - 2, written in round r + 1
      if (b.prio) reset_next_min(c, (r + 2) % 3);
      if (b.instrumentation) atomicMax(&c->max_list, static_cast<unsigned long long>(n));
    }
    process_round<B, EPT, I, M>(g, b, r, n, limit, l);
    if (!grid_barrier_grouped(c, gridDim.x)) return;  // error bits set; the host reports a device failure
    if (b.prio) limit = next_window_limit(c, b, r, limit);
    ++r;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    c->round = r;
    c->rounds_executed = r;
    c->done = 1;
    c->stop_reason = n == 0 ? kStopConverged : kStopMaximumRounds;
  }
}

// Host-polled control: one round per launch (ordinary grid), then the
// advance kernel (jacobi_pull.hpp:153-220 shape: count the round, reset
// the list that the next round fills, stop on an empty list or the limit).
template <int B, int EPT, bool I, int MW, int M>
__global__ __launch_bounds__(B, MW) void relax_round_kernel(DevGraph g, DevBatch b,
                                                        std::uint32_t max_v,
                                                        std::uint32_t work_words) {
  extern __shared__ std::uint32_t lds_words[];
  const Lds l = carve(lds_words, max_v, work_words, g.max_neighbor_tiles);
  DevControl* c = b.control;
  if (load_agent(&c->done)) return;
  const std::uint32_t r = load_agent(&c->round);
  const std::uint32_t n = b.ordered ? 0u : load_agent(&c->list_count[r % 3]);
  const std::uint32_t limit = b.prio ? load_agent(&c->window_limit) : 0u;
  process_round<B, EPT, I, M>(g, b, r, n, limit, l);
}

// ---------------------------------------------------------------------------
// Launch wrappers (host side; the advance kernel between round launches is
// launched by relax.hip).
// ---------------------------------------------------------------------------
template <int B, int EPT, bool I, int MW, int M>
hipError_t launch_relax(bool persistent, bool cooperative, std::uint32_t grid,
                        std::size_t lds_bytes, std::uint32_t max_v, std::uint32_t work_words,
                        const DevGraph& g, const DevBatch& b, hipStream_t stream) {
  if (persistent) {
    void* args[] = {const_cast<DevGraph*>(&g), const_cast<DevBatch*>(&b), &max_v, &work_words};
    const void* kernel = reinterpret_cast<const void*>(&relax_persistent_kernel<B, EPT, I, MW, M>);
    if (cooperative)
      return hipLaunchCooperativeKernel(kernel, dim3(grid), dim3(B), args,
                                        static_cast<unsigned>(lds_bytes), stream);
    return hipLaunchKernel(kernel, dim3(grid), dim3(B), args, lds_bytes, stream);
  }
  hipLaunchKernelGGL((relax_round_kernel<B, EPT, I, MW, M>), dim3(grid), dim3(B), lds_bytes, stream,
                     g, b, max_v, work_words);
  return hipGetLastError();
}

template <int B, int EPT, bool I, int MW, int M>
hipError_t occupancy_relax(bool persistent, std::size_t lds_bytes, int* blocks_per_cu) {
  const void* kernel =
      persistent ? reinterpret_cast<const void*>(&relax_persistent_kernel<B, EPT, I, MW, M>)
                 : reinterpret_cast<const void*>(&relax_round_kernel<B, EPT, I, MW, M>);
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, kernel, B, lds_bytes);
}

// Instantiated kernel variants.  Block sizes come with the registers per
// thread that hold a tile's internal records through its closure (B * EPT =
// 2048: 512-vertex tiles with four internal out-edges per vertex fit).
// The 256- and 512-thread blocks also come bounded to at least 6, 7 or 8 waves
// per SIMD (fewer registers per lane, more visits in flight); instrumented
// runs use the unbounded variant.
template <int B, int EPT, typename F>
hipError_t with_waves(bool instrumented, std::uint32_t min_waves, F& f) {
  if (instrumented) return f.template operator()<B, EPT, true, 1>();
  if constexpr (B == 256 || B == 512) {
    if (min_waves >= 8) return f.template operator()<B, EPT, false, 8>();
    if (min_waves >= 7) return f.template operator()<B, EPT, false, 7>();
    if (min_waves >= 6) return f.template operator()<B, EPT, false, 6>();
  }
  return f.template operator()<B, EPT, false, 1>();
}

template <typename F>
hipError_t with_shape(std::uint32_t block, bool instrumented, std::uint32_t min_waves, F&& f) {
  switch (block) {
    case 64: return with_waves<64, 32>(instrumented, min_waves, f);
    case 128: return with_waves<128, 16>(instrumented, min_waves, f);
    case 256: return with_waves<256, 8>(instrumented, min_waves, f);
    case 512: return with_waves<512, 4>(instrumented, min_waves, f);
    case 1024: return with_waves<1024, 2>(instrumented, min_waves, f);
    default: return hipErrorInvalidValue;
  }
}

// The variants of one engine set, for every block size and wave bound.
template <int M>
hipError_t relax_launch_engines(std::uint32_t block, bool persistent, bool cooperative,
                                std::uint32_t grid, std::size_t lds_bytes, std::uint32_t max_v,
                                std::uint32_t work_words, const DevGraph& g, const DevBatch& b,
                                hipStream_t stream) {
  return with_shape(block, b.instrumentation != 0, b.min_waves,
                    [&]<int B, int EPT, bool I, int MW>() {
                      return launch_relax<B, EPT, I, MW, M>(persistent, cooperative, grid, lds_bytes,
                                                            max_v, work_words, g, b, stream);
                    });
}

template <int M>
hipError_t relax_occupancy_engines(std::uint32_t block, bool persistent, bool instrumented,
                                   std::uint32_t min_waves, std::size_t lds_bytes,
                                   int* blocks_per_cu) {
  return with_shape(block, instrumented, min_waves, [&]<int B, int EPT, bool I, int MW>() {
    return occupancy_relax<B, EPT, I, MW, M>(persistent, lds_bytes, blocks_per_cu);
  });
}

}  // namespace bfn
