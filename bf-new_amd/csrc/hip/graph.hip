// graph.hip -- derives the tile-major engine image from the reference layout.
//
// Input: DeviceGraphLayout32 (device_layout.hpp:72-93) already in HBM, with
// the tile runs of device_layout.cpp:186-251 and tile vertex offsets.
// Output, per tile T:
//   in_nbr_tiles  -- T, then the other source tiles of T's CSC runs, ascending
//   out_nbr_tiles -- T, then the other destination tiles of T's CSR runs
//   in_nbr_stage  -- LDS float offset of each in-neighbour slab (prefix of
//                    tile sizes in list order, so T's own slab is at 0)
//   csc_pack / csr_pack -- one u32 per edge at its orientation position:
//                    neighbour index << 24 | source-local id << 12 |
//                    destination-local id
//   csc_rec / csr_rec -- the relaxation stream: per tile, the tile's edge
//                    block reordered as [internal edges | boundary edges]
//                    (each part in orientation order), one (pack, weight
//                    bits) record per edge; *_perm[p] is the orientation
//                    position of record p and *_intra[T] the internal count.
// One workgroup per tile; the distinct tiles of the tile's runs are collected
// in an LDS hash set, ranked in LDS, and the packs are written per vertex.
#include "bfn_internal.hpp"
#include "bfn_launch.hpp"

namespace bfn {

namespace {

constexpr std::uint32_t kHashSize = 1024;

struct DeriveArgs {
  std::uint32_t tile_count;
  const std::uint32_t* tile_vertex_offsets;
  const std::uint32_t* owner_tiles;
  // orientation: CSC (in) or CSR (out)
  const std::uint32_t* offsets;        // [V+1]
  const std::uint32_t* endpoints;      // sources (CSC) or destinations (CSR)
  const std::uint32_t* run_offsets;    // [V+1] runs per vertex
  const std::uint32_t* run_tiles;      // remote tile per run
  std::uint32_t* counts;               // [T] distinct neighbour tiles (incl. T)
  const std::uint32_t* list_offsets;   // [T+1] (fill/pack passes)
  std::uint32_t* list_tiles;
  std::uint32_t* list_stage;           // LDS slab offsets in list order
  std::uint32_t* pack;                 // [E]
  std::uint32_t* stats;                // [0] error, [1] max stage floats, [2] max CSC edges
                                       // per tile, [3] max neighbours, [4] zero weight,
                                       // [5] max CSR edges per tile, [6] max proposal
                                       // floats (CSR stage)
  const float* weights;
  bool csc;
};

__device__ void insert(std::uint32_t* set, std::uint32_t key, std::uint32_t* overflow) {
  std::uint32_t h = (key * 2654435761u) >> 22;
  for (std::uint32_t probe = 0; probe < kHashSize; ++probe) {
    const std::uint32_t slot = (h + probe) & (kHashSize - 1);
    const std::uint32_t prev = atomicCAS(&set[slot], kNone, key);
    if (prev == kNone || prev == key) return;
  }
  *overflow = 1;
}

// Collects the distinct tiles into `set`; returns the compacted, ranked list
// in `list` (T first) and its length.
__device__ std::uint32_t collect(const DeriveArgs& a, std::uint32_t T, std::uint32_t* set,
                                 std::uint32_t* list, std::uint32_t* scratch) {
  const std::uint32_t tid = threadIdx.x;
  for (std::uint32_t i = tid; i < kHashSize; i += blockDim.x) set[i] = kNone;
  if (tid == 0) {
    scratch[0] = 0;  // overflow
    scratch[1] = 0;  // count
  }
  __syncthreads();
  if (tid == 0) insert(set, T, &scratch[0]);
  const std::uint32_t v0 = a.tile_vertex_offsets[T], v1 = a.tile_vertex_offsets[T + 1];
  const std::uint32_t r0 = a.run_offsets[v0], r1 = a.run_offsets[v1];
  for (std::uint32_t r = r0 + tid; r < r1; r += blockDim.x) insert(set, a.run_tiles[r], &scratch[0]);
  __syncthreads();
  // compact
  for (std::uint32_t i = tid; i < kHashSize; i += blockDim.x) {
    if (set[i] != kNone) {
      const std::uint32_t pos = atomicAdd(&scratch[1], 1u);
      scratch[2 + pos] = set[i];
    }
  }
  __syncthreads();
  const std::uint32_t n = scratch[1];
  // rank: T first, the rest ascending
  for (std::uint32_t i = tid; i < n; i += blockDim.x) {
    const std::uint32_t key = scratch[2 + i];
    std::uint32_t rank = 0;
    if (key != T) {
      rank = 1;
      for (std::uint32_t j = 0; j < n; ++j) {
        const std::uint32_t other = scratch[2 + j];
        if (other != T && other < key) ++rank;
      }
    }
    list[rank] = key;
  }
  __syncthreads();
  if (tid == 0 && (scratch[0] || n > kMaxNeighborTiles)) atomicOr(&a.stats[0], 1u);
  return n;
}

__global__ void count_kernel(DeriveArgs a) {
  __shared__ std::uint32_t set[kHashSize];
  __shared__ std::uint32_t list[kHashSize];
  __shared__ std::uint32_t scratch[kHashSize + 2];
  const std::uint32_t T = blockIdx.x;
  const std::uint32_t n = collect(a, T, set, list, scratch);
  if (threadIdx.x == 0) {
    a.counts[T] = n;
    atomicMax(&a.stats[3], n);
    const std::uint32_t v0 = a.tile_vertex_offsets[T], v1 = a.tile_vertex_offsets[T + 1];
    atomicMax(&a.stats[a.csc ? 2 : 5], a.offsets[v1] - a.offsets[v0]);
  }
}

__global__ void fill_pack_kernel(DeriveArgs a) {
  __shared__ std::uint32_t set[kHashSize];
  __shared__ std::uint32_t list[kHashSize];
  __shared__ std::uint32_t scratch[kHashSize + 2];
  const std::uint32_t T = blockIdx.x;
  const std::uint32_t n = collect(a, T, set, list, scratch);
  const std::uint32_t tid = threadIdx.x;
  const std::uint32_t base = a.list_offsets[T];
  if (n > kMaxNeighborTiles || a.list_offsets[T + 1] - base != n) return;
  for (std::uint32_t i = tid; i < n; i += blockDim.x) a.list_tiles[base + i] = list[i];
  if (a.list_stage && tid == 0) {
    // CSC: staged read slabs of every in-neighbour, T's own first (offset 0).
    // CSR: proposal slabs of the other out-neighbours only (k >= 1).
    std::uint32_t off = 0;
    for (std::uint32_t i = 0; i < n; ++i) {
      a.list_stage[base + i] = off;
      if (a.csc || i > 0) off += a.tile_vertex_offsets[list[i] + 1] - a.tile_vertex_offsets[list[i]];
    }
    atomicMax(&a.stats[a.csc ? 1 : 6], off);
  }
  const std::uint32_t v0 = a.tile_vertex_offsets[T], v1 = a.tile_vertex_offsets[T + 1];
  for (std::uint32_t v = v0 + tid; v < v1; v += blockDim.x) {
    for (std::uint32_t e = a.offsets[v]; e < a.offsets[v + 1]; ++e) {
      const std::uint32_t other = a.endpoints[e];
      const std::uint32_t ot = a.owner_tiles[other];
      std::uint32_t k = 0;
      if (ot != T) {
        std::uint32_t lo = 1, hi = n;
        while (lo < hi) {
          const std::uint32_t mid = (lo + hi) >> 1;
          if (list[mid] < ot) lo = mid + 1; else hi = mid;
        }
        k = lo;
      }
      const std::uint32_t other_local = other - a.tile_vertex_offsets[ot];
      const std::uint32_t self_local = v - v0;
      // CSC: the edge runs other -> v; CSR: v -> other.
      const std::uint32_t src_local = a.csc ? other_local : self_local;
      const std::uint32_t dst_local = a.csc ? self_local : other_local;
      a.pack[e] = (k << 24) | (src_local << 12) | dst_local;
      if (a.weights[e] == 0.0f) a.stats[4] = 1;
    }
  }
}

struct SplitArgs {
  const std::uint32_t* tile_vertex_offsets;
  const std::uint32_t* offsets;  // [V+1] orientation offsets
  const std::uint32_t* pack;     // [E] at orientation positions
  const float* weights;          // [E] at orientation positions
  uint2* rec;
  std::uint32_t* perm;
  std::uint32_t* intra;          // [T]
};

// Stable partition of each tile's block into internal and boundary records.
// A 64-lane ballot ranks the internal edges of a chunk inside each wave, the
// wave totals are scanned in LDS, and a carry runs across chunks.
__global__ __launch_bounds__(256) void split_records_kernel(SplitArgs a) {
  __shared__ std::uint32_t wave_total[16];
  __shared__ std::uint32_t internal_total;
  const std::uint32_t T = blockIdx.x;
  const std::uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const std::uint32_t e0 = a.offsets[a.tile_vertex_offsets[T]];
  const std::uint32_t e1 = a.offsets[a.tile_vertex_offsets[T + 1]];
  if (tid == 0) internal_total = 0;
  __syncthreads();
  std::uint32_t mine = 0;
  for (std::uint32_t e = e0 + tid; e < e1; e += blockDim.x) mine += (a.pack[e] >> 24) == 0u;
  if (mine) atomicAdd(&internal_total, mine);
  __syncthreads();
  const std::uint32_t n_int = internal_total;
  std::uint32_t carry = 0;  // internal edges before the current chunk
  for (std::uint32_t base = e0; base < e1; base += blockDim.x) {
    const std::uint32_t e = base + tid;
    const bool in_range = e < e1;
    const std::uint32_t p = in_range ? a.pack[e] : kNone;
    const bool internal = in_range && (p >> 24) == 0u;
    const unsigned long long m = __ballot(internal);
    const std::uint32_t below = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wave_total[wave] = __popcll(m);
    __syncthreads();
    std::uint32_t before = carry;
    for (std::uint32_t w = 0; w < wave; ++w) before += wave_total[w];
    std::uint32_t chunk = 0;
    for (std::uint32_t w = 0; w < (blockDim.x >> 6); ++w) chunk += wave_total[w];
    if (in_range) {
      const std::uint32_t rank_int = before + below;               // internal edges before e
      const std::uint32_t rank_bnd = (e - e0) - rank_int;          // boundary edges before e
      const std::uint32_t dst = internal ? e0 + rank_int : e0 + n_int + rank_bnd;
      a.rec[dst] = make_uint2(p, __float_as_uint(a.weights[e]));
      a.perm[dst] = e;
    }
    carry += chunk;
    __syncthreads();
  }
  if (tid == 0) a.intra[T] = n_int;
}

// Per vertex, the position of its first boundary record in the tile's
// [internal | boundary] block: the boundary part keeps orientation order, so
// a vertex's boundary records follow those of the tile's earlier vertices.
// One workgroup per tile; a carry runs across 256-vertex chunks.
__global__ __launch_bounds__(256) void boundary_starts_kernel(const std::uint32_t* tile_vertex_offsets,
                                                              const std::uint32_t* offsets,
                                                              const std::uint32_t* pack,
                                                              const std::uint32_t* intra,
                                                              std::uint32_t* starts) {
  __shared__ std::uint32_t buf[256];
  const std::uint32_t T = blockIdx.x, tid = threadIdx.x;
  const std::uint32_t v0 = tile_vertex_offsets[T], v1 = tile_vertex_offsets[T + 1];
  std::uint32_t carry = offsets[v0] + intra[T];
  for (std::uint32_t base = v0; base < v1; base += 256) {
    const std::uint32_t v = base + tid;
    std::uint32_t count = 0;
    if (v < v1)
      for (std::uint32_t e = offsets[v]; e < offsets[v + 1]; ++e) count += (pack[e] >> 24) != 0u;
    buf[tid] = count;
    __syncthreads();
    for (std::uint32_t off = 1; off < 256; off <<= 1) {
      const std::uint32_t y = tid >= off ? buf[tid - off] : 0u;
      __syncthreads();
      buf[tid] += y;
      __syncthreads();
    }
    if (v < v1) starts[v] = carry + buf[tid] - count;
    carry += buf[255];
    __syncthreads();
  }
}

__global__ void refresh_weights_kernel(uint2* rec, const std::uint32_t* perm, const float* w,
                                       std::uint32_t n) {
  const std::uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) rec[p].y = __float_as_uint(w[perm[p]]);
}

__global__ void exclusive_scan_kernel(const std::uint32_t* in, std::uint32_t* out,
                                      std::uint32_t n) {
  // Single workgroup: chunked Hillis-Steele in LDS, carry across chunks.
  __shared__ std::uint32_t buf[1024];
  __shared__ std::uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (std::uint32_t base = 0; base < n; base += 1024) {
    const std::uint32_t i = base + threadIdx.x;
    std::uint32_t x = i < n ? in[i] : 0;
    buf[threadIdx.x] = x;
    __syncthreads();
    for (std::uint32_t off = 1; off < 1024; off <<= 1) {
      std::uint32_t y = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += y;
      __syncthreads();
    }
    if (i < n) out[i] = carry + buf[threadIdx.x] - x;
    __syncthreads();
    if (threadIdx.x == 1023) carry += buf[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = carry;
}

}  // namespace

// Runs both orientations.  stats[0] != 0 means a tile has more than 255
// neighbour tiles (not representable in the 8-bit neighbour index).
hipError_t derive_tile_image(const TileImageArgs& t, hipStream_t stream) {
  DeriveArgs in{t.tile_count, t.tile_vertex_offsets, t.owner_tiles, t.column_offsets, t.sources,
                t.col_runs, t.run_src_tiles, t.counts_scratch, t.in_offsets, nullptr, nullptr,
                t.csc_pack, t.stats, t.in_weights, true};
  DeriveArgs out{t.tile_count, t.tile_vertex_offsets, t.owner_tiles, t.row_offsets,
                 t.destinations, t.row_runs, t.run_dst_tiles, t.counts_scratch, t.out_offsets,
                 nullptr, nullptr, t.csr_pack, t.stats, t.weights, false};
  hipError_t e;
  // In-neighbour lists.
  hipLaunchKernelGGL(count_kernel, dim3(t.tile_count), dim3(256), 0, stream, in);
  hipLaunchKernelGGL(exclusive_scan_kernel, dim3(1), dim3(1024), 0, stream, t.counts_scratch,
                     t.in_offsets, t.tile_count);
  if ((e = hipMemcpyAsync(t.host_in_total, t.in_offsets + t.tile_count, sizeof(std::uint32_t),
                          hipMemcpyDeviceToHost, stream)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
  if ((e = hipMalloc(t.in_tiles, sizeof(std::uint32_t) * (*t.host_in_total + 1))) != hipSuccess)
    return e;
  if ((e = hipMalloc(t.in_stage, sizeof(std::uint32_t) * (*t.host_in_total + 1))) != hipSuccess)
    return e;
  in.list_tiles = *t.in_tiles;
  in.list_stage = *t.in_stage;
  hipLaunchKernelGGL(fill_pack_kernel, dim3(t.tile_count), dim3(256), 0, stream, in);
  // Out-neighbour lists.
  hipLaunchKernelGGL(count_kernel, dim3(t.tile_count), dim3(256), 0, stream, out);
  hipLaunchKernelGGL(exclusive_scan_kernel, dim3(1), dim3(1024), 0, stream, t.counts_scratch,
                     t.out_offsets, t.tile_count);
  if ((e = hipMemcpyAsync(t.host_out_total, t.out_offsets + t.tile_count, sizeof(std::uint32_t),
                          hipMemcpyDeviceToHost, stream)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
  if ((e = hipMalloc(t.out_tiles, sizeof(std::uint32_t) * (*t.host_out_total + 1))) != hipSuccess)
    return e;
  if ((e = hipMalloc(t.out_stage, sizeof(std::uint32_t) * (*t.host_out_total + 1))) != hipSuccess)
    return e;
  out.list_tiles = *t.out_tiles;
  out.list_stage = *t.out_stage;
  hipLaunchKernelGGL(fill_pack_kernel, dim3(t.tile_count), dim3(256), 0, stream, out);
  // Relaxation streams.
  const SplitArgs sin{t.tile_vertex_offsets, t.column_offsets, t.csc_pack, t.in_weights,
                      t.csc_rec, t.csc_perm, t.csc_intra};
  const SplitArgs sout{t.tile_vertex_offsets, t.row_offsets, t.csr_pack, t.weights,
                       t.csr_rec, t.csr_perm, t.csr_intra};
  hipLaunchKernelGGL(split_records_kernel, dim3(t.tile_count), dim3(256), 0, stream, sin);
  hipLaunchKernelGGL(split_records_kernel, dim3(t.tile_count), dim3(256), 0, stream, sout);
  hipLaunchKernelGGL(boundary_starts_kernel, dim3(t.tile_count), dim3(256), 0, stream,
                     t.tile_vertex_offsets, t.row_offsets, t.csr_pack, t.csr_intra, t.csr_bnd);
  return hipGetLastError();
}

hipError_t refresh_record_weights(const DevGraph& g, const std::uint32_t* csr_perm,
                                  const std::uint32_t* csc_perm, hipStream_t stream) {
  const std::uint32_t n = g.edge_count;
  if (n == 0) return hipSuccess;
  const dim3 grid((n + 255) / 256);
  hipLaunchKernelGGL(refresh_weights_kernel, grid, dim3(256), 0, stream,
                     const_cast<uint2*>(g.csr_rec), csr_perm, g.weights, n);
  hipLaunchKernelGGL(refresh_weights_kernel, grid, dim3(256), 0, stream,
                     const_cast<uint2*>(g.csc_rec), csc_perm, g.in_weights, n);
  return hipGetLastError();
}

}  // namespace bfn
