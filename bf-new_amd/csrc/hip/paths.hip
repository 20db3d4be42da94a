// paths.hip -- result extraction: compact paths, predecessors, distances,
// per-query finalisation, occupancy and negotiated-cost weight updates.
//
// Compact paths follow compact_paths.cpp:117-206 (and the device version
// compact_path_results.hip.cpp:169-249): from each target, walk tight
// admitted incoming edges (d[u] + w == d[v], u selected and finite) in
// stable-EdgeId order, skipping vertices already on the path, backtracking
// when a vertex has no untried candidate, until a source is reached.  Here
// one wavefront owns one query: its 64 lanes scan a vertex's incoming edges
// together and a wave min-reduction picks the smallest untried EdgeId, so a
// step costs one pass over the in-edges instead of a sort.  A frame stores
// its vertex and the last EdgeId tried from it; resuming a frame asks for the
// smallest candidate above that EdgeId, which visits candidates in exactly
// the reference's cursor order.
#include "bfn_internal.hpp"

namespace bfn {


__device__ __forceinline__ std::uint32_t column_of(const DevBatch& b, std::uint32_t q,
                                                   std::uint32_t jacobi_copy) {
  return b.engine[q] == kJacobiPull ? jacobi_copy : 0u;
}

// Label of global vertex v for query q (+inf when its tile is not selected).
__device__ std::uint32_t label_of(const DevGraph& g, const DevBatch& b, std::uint32_t q,
                                  std::uint32_t v, std::uint32_t jacobi_copy,
                                  unsigned long long* pool_index) {
  const std::uint32_t T = g.owner_tiles[v];
  const std::uint32_t s = find_slot(b, q, T);
  if (s == kNone) {
    if (pool_index) *pool_index = ~0ull;
    return kInfBits;
  }
  const std::uint32_t v0 = g.tile_vertex_offsets[T];
  const std::uint32_t npt = g.tile_vertex_offsets[T + 1] - v0;
  const unsigned long long base = b.slot_dist[s] + (v - v0);
  if (pool_index) *pool_index = base;
  return b.dist[base + static_cast<unsigned long long>(column_of(b, q, jacobi_copy)) * npt];
}

__device__ __forceinline__ bool is_source(const DevBatch& b, std::uint32_t q, std::uint32_t v) {
  return find_sorted(b.sources, b.source_offsets[q], b.source_offsets[q + 1], v) != kNone;
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const std::uint32_t lo = __shfl_xor(static_cast<std::uint32_t>(x), off, 64);
    const std::uint32_t hi = __shfl_xor(static_cast<std::uint32_t>(x >> 32), off, 64);
    const unsigned long long y = (static_cast<unsigned long long>(hi) << 32) | lo;
    x = y < x ? y : x;
  }
  return x;
}

// One 64-lane block per query.
__global__ __launch_bounds__(64) void compact_paths_kernel(PathArgs a) {
  const std::uint32_t q = a.query_list ? a.query_list[blockIdx.x] : blockIdx.x;
  const DevGraph& g = a.g;
  const DevBatch& b = a.b;
  const std::uint32_t lane = threadIdx.x;
  const std::uint32_t copy =
      a.jacobi_copy_from_control ? (b.control->round & 1u) : a.jacobi_copy;
  const unsigned long long fb = a.frame_base[q];
  for (std::uint32_t t = b.target_offsets[q]; t < b.target_offsets[q + 1]; ++t) {
    const std::uint32_t target = b.targets[t];
    unsigned long long target_pool;
    const std::uint32_t dt = label_of(g, b, q, target, copy, &target_pool);
    std::uint32_t* sum = a.summaries + static_cast<std::size_t>(t) * 7;
    if (!a.query_clean[q] || dt == kInfBits) {
      if (a.pass == 1 && lane == 0) {
        sum[0] = target;
        sum[1] = 0;
        sum[2] = kInfBits;
        sum[3] = 0;
        sum[4] = 0;                                  // not_reached
        sum[5] = a.query_clean[q] ? 1u : 2u;         // unreachable / terminal failure
        sum[6] = 0;
        a.lengths[t] = 0;
      }
      continue;
    }
    const std::uint32_t stamp = a.stamp_base + t + 1;
    std::int32_t top = 0;
    std::uint32_t v = target;
    std::uint32_t last = kNone;  // EdgeId last tried from the top frame
    if (lane == 0) {
      a.frame_vertex[fb] = target;
      a.frame_via[fb] = kNone;
      a.onpath[target_pool] = stamp;
    }
    bool found_source = false;
    for (;;) {
      if (is_source(b, q, v)) {
        found_source = true;
        break;
      }
      const std::uint32_t T = g.owner_tiles[v];
      const std::uint32_t s = find_slot(b, q, T);
      const std::uint32_t dv = label_of(g, b, q, v, copy, nullptr);
      const std::uint32_t in0 = g.in_nbr_offsets[T];
      unsigned long long best = ~0ull;
      for (std::uint32_t e = g.column_offsets[v] + lane; e < g.column_offsets[v + 1]; e += 64) {
        const std::uint32_t eid = g.in_edge_ids[e];
        if (last != kNone && eid <= last) continue;
        const std::uint32_t p = g.csc_pack[e];
        const std::uint32_t k = p >> 24;
        const std::uint32_t ns = b.slot_in[b.slot_in_base[s] + k];
        if (ns == kNone) continue;
        const std::uint32_t nt = g.in_nbr_tiles[in0 + k];
        const std::uint32_t nn = g.tile_vertex_offsets[nt + 1] - g.tile_vertex_offsets[nt];
        const unsigned long long pool = b.slot_dist[ns] + ((p >> 12) & kPackLocalMask);
        const std::uint32_t du =
            b.dist[pool + static_cast<unsigned long long>(column_of(b, q, copy)) * nn];
        if (du == kInfBits) continue;
        if (as_bits(as_float(du) + g.in_weights[e]) != dv) continue;
        if (a.onpath[pool] == stamp) continue;
        const unsigned long long key =
            (static_cast<unsigned long long>(eid) << 32) | g.sources[e];
        best = key < best ? key : best;
      }
      best = wave_min_u64(best);
      __syncthreads();
      if (best != ~0ull) {
        const std::uint32_t eid = static_cast<std::uint32_t>(best >> 32);
        const std::uint32_t u = static_cast<std::uint32_t>(best);
        unsigned long long upool;
        label_of(g, b, q, u, copy, &upool);
        if (lane == 0) {
          a.frame_last[fb + top] = eid;
          a.frame_vertex[fb + top + 1] = u;
          a.frame_via[fb + top + 1] = eid;
          a.onpath[upool] = stamp;
        }
        ++top;
        v = u;
        last = kNone;
      } else {
        unsigned long long vpool;
        label_of(g, b, q, v, copy, &vpool);
        if (lane == 0) a.onpath[vpool] = 0;
        --top;
        if (top < 0) break;
        __threadfence_block();
        __syncthreads();
        v = a.frame_vertex[fb + top];
        last = a.frame_last[fb + top];
      }
      __threadfence_block();
      __syncthreads();
    }
    if (a.pass == 1) {
      if (lane == 0) {
        sum[0] = target;
        sum[2] = dt;
        sum[4] = 1;  // reached
        if (found_source) {
          sum[1] = a.frame_vertex[fb + top];
          sum[3] = static_cast<std::uint32_t>(top);
          sum[5] = 0;  // complete
          sum[6] = 1;
          a.lengths[t] = static_cast<unsigned long long>(top) + 1;
        } else {
          sum[1] = 0;
          sum[3] = 0;
          sum[5] = 3;  // no_tight_path
          sum[6] = 0;
          a.lengths[t] = 0;
        }
      }
    } else if (found_source) {
      // Edges share the vertex offsets; a path of L edges leaves slot L unused.
      const unsigned long long out = a.offsets[t];
      for (std::int32_t i = lane; i <= top; i += 64) {
        const std::uint32_t vert = a.frame_vertex[fb + (top - i)];
        a.arena_vertices[out + i] = vert;
        a.arena_labels[out + i] = label_of(g, b, q, vert, copy, nullptr);
        if (i < top) a.arena_edges[out + i] = a.frame_via[fb + (top - i)];
      }
    }
    __syncthreads();
  }
}

// Smallest-EdgeId tight admitted predecessor of each vertex of query q.
__global__ void predecessors_kernel(DevGraph g, DevBatch b, std::uint32_t q,
                                    std::uint32_t jacobi_copy, std::uint32_t* out) {
  const std::uint32_t pos = b.tile_begin[q] + blockIdx.x;
  if (pos >= b.tile_end[q]) return;
  const std::uint32_t s = slot_at(b, pos);
  const std::uint32_t T = b.tiles[s];
  const std::uint32_t v0 = g.tile_vertex_offsets[T];
  const std::uint32_t npt = g.tile_vertex_offsets[T + 1] - v0;
  const std::uint32_t in0 = g.in_nbr_offsets[T];
  const std::uint32_t col = column_of(b, q, jacobi_copy);
  for (std::uint32_t i = threadIdx.x; i < npt; i += blockDim.x) {
    const std::uint32_t v = v0 + i;
    std::uint32_t best_eid = kNone, best_u = kNone;
    const std::uint32_t dv = b.dist[b.slot_dist[s] + static_cast<unsigned long long>(col) * npt + i];
    if (dv != kInfBits && !is_source(b, q, v)) {
      for (std::uint32_t e = g.column_offsets[v]; e < g.column_offsets[v + 1]; ++e) {
        const std::uint32_t p = g.csc_pack[e];
        const std::uint32_t k = p >> 24;
        const std::uint32_t ns = b.slot_in[b.slot_in_base[s] + k];
        if (ns == kNone) continue;
        const std::uint32_t nt = g.in_nbr_tiles[in0 + k];
        const std::uint32_t nn = g.tile_vertex_offsets[nt + 1] - g.tile_vertex_offsets[nt];
        const std::uint32_t du = b.dist[b.slot_dist[ns] + static_cast<unsigned long long>(col) * nn +
                                        ((p >> 12) & kPackLocalMask)];
        if (du == kInfBits || as_bits(as_float(du) + g.in_weights[e]) != dv) continue;
        const std::uint32_t eid = g.in_edge_ids[e];
        if (eid < best_eid) {
          best_eid = eid;
          best_u = g.sources[e];
        }
      }
    }
    out[v] = best_u;
  }
}

// Scatters query q's final labels to global ids (out pre-filled with +inf).
__global__ void scatter_distances_kernel(DevGraph g, DevBatch b, std::uint32_t q,
                                         std::uint32_t jacobi_copy, std::uint32_t* out) {
  const std::uint32_t pos = b.tile_begin[q] + blockIdx.x;
  if (pos >= b.tile_end[q]) return;
  const std::uint32_t s = slot_at(b, pos);
  const std::uint32_t T = b.tiles[s];
  const std::uint32_t v0 = g.tile_vertex_offsets[T];
  const std::uint32_t npt = g.tile_vertex_offsets[T + 1] - v0;
  const std::uint32_t* src =
      b.dist + b.slot_dist[s] + static_cast<unsigned long long>(column_of(b, q, jacobi_copy)) * npt;
  for (std::uint32_t i = threadIdx.x; i < npt; i += blockDim.x) out[v0 + i] = src[i];
}


__global__ void finalize_kernel(DevGraph g, DevBatch b, const unsigned long long* relaxations,
                                QueryStatusDev* out) {
  const std::uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= b.query_count) return;
  const DevControl* c = b.control;
  // Round r writes Jacobi column (r+1)&1; unvisited tiles hold equal columns.
  const std::uint32_t jacobi_copy = c->rounds_executed & 1u;
  const std::uint32_t executed = c->rounds_executed;
  const std::uint32_t lc = b.last_change[q];
  const bool clean = c->error_bits == 0;
  // A query converged once a round after its last decrease found no change.
  const bool converged = clean && (lc < executed || c->stop_reason == kStopConverged);
  bool reached = true;
  for (std::uint32_t i = b.target_offsets[q]; i < b.target_offsets[q + 1]; ++i)
    reached &= label_of(g, b, q, b.targets[i], jacobi_copy, nullptr) != kInfBits;
  QueryStatusDev st;
  st.rounds = converged ? static_cast<unsigned long long>(lc) + 1 : executed;
  st.relaxations = relaxations ? relaxations[q] : 0;
  st.converged = converged;
  st.reached = converged && reached;
  st.miss = converged && !reached;
  st.stop_reason = !clean ? kStopDeviceFailure : (converged ? kStopConverged : c->stop_reason);
  st.error_bits = c->error_bits;
  st.engine = b.engine[q];
  out[q] = st;
}

// Work-normalised relaxation count: admitted out-edges of reached vertices.
__global__ void relaxation_count_kernel(DevGraph g, DevBatch b, unsigned long long* out) {
  const std::uint32_t jacobi_copy = b.control->rounds_executed & 1u;
  const std::uint32_t s = blockIdx.x;
  if (s >= b.slot_count) return;
  const std::uint32_t q = b.slot_query[s];
  const std::uint32_t T = b.tiles[s];
  const std::uint32_t v0 = g.tile_vertex_offsets[T];
  const std::uint32_t npt = g.tile_vertex_offsets[T + 1] - v0;
  const std::uint32_t col = column_of(b, q, jacobi_copy);
  unsigned long long count = 0;
  for (std::uint32_t i = threadIdx.x; i < npt; i += blockDim.x) {
    if (b.dist[b.slot_dist[s] + static_cast<unsigned long long>(col) * npt + i] == kInfBits) continue;
    const std::uint32_t v = v0 + i;
    for (std::uint32_t e = g.row_offsets[v]; e < g.row_offsets[v + 1]; ++e)
      count += b.slot_out[b.slot_out_base[s] + (g.csr_pack[e] >> 24)] != kNone;
  }
  if (count) atomicAdd(&out[q], count);
}

__global__ void occupancy_kernel(const std::uint32_t* arena_vertices, unsigned long long n,
                                 int* occupancy) {
  const unsigned long long stride = static_cast<unsigned long long>(gridDim.x) * blockDim.x;
  for (unsigned long long i = static_cast<unsigned long long>(blockIdx.x) * blockDim.x +
                              threadIdx.x;
       i < n; i += stride)
    atomicAdd(&occupancy[arena_vertices[i]], 1);
}

// w(u -> v) = base[v] * (1 + 0.5 * occ[v]) in this fp32 order, both orientations.
__global__ void node_cost_kernel(DevGraph g, const float* base, const int* occupancy,
                                 float* csr_weights, float* csc_weights) {
  const std::uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= g.vertex_count) return;
  const float occ = occupancy ? static_cast<float>(occupancy[v]) : 0.0f;
  const float scale = 1.0f + 0.5f * occ;
  const float w = base[v] * scale;
  for (std::uint32_t e = g.column_offsets[v]; e < g.column_offsets[v + 1]; ++e) csc_weights[e] = w;
  for (std::uint32_t e = g.row_offsets[v]; e < g.row_offsets[v + 1]; ++e) {
    const std::uint32_t d = g.destinations[e];
    const float od = occupancy ? static_cast<float>(occupancy[d]) : 0.0f;
    const float sd = 1.0f + 0.5f * od;
    csr_weights[e] = base[d] * sd;
  }
}

__global__ void scan_u64_kernel(const unsigned long long* in, unsigned long long* out,
                                std::uint32_t n) {
  __shared__ unsigned long long buf[1024];
  __shared__ unsigned long long carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (std::uint32_t base = 0; base < n; base += 1024) {
    const std::uint32_t i = base + threadIdx.x;
    const unsigned long long x = i < n ? in[i] : 0;
    buf[threadIdx.x] = x;
    __syncthreads();
    for (std::uint32_t off = 1; off < 1024; off <<= 1) {
      const unsigned long long y = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
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

// ---- launch wrappers ------------------------------------------------------
hipError_t paths_launch(const PathArgs& a, std::uint32_t query_count, hipStream_t stream) {
  hipLaunchKernelGGL(compact_paths_kernel, dim3(query_count), dim3(64), 0, stream, a);
  return hipGetLastError();
}

hipError_t scan_u64_launch(const unsigned long long* in, unsigned long long* out, std::uint32_t n,
                           hipStream_t stream) {
  hipLaunchKernelGGL(scan_u64_kernel, dim3(1), dim3(1024), 0, stream, in, out, n);
  return hipGetLastError();
}

hipError_t predecessors_launch(const DevGraph& g, const DevBatch& b, std::uint32_t q,
                               std::uint32_t slots, std::uint32_t copy, std::uint32_t* out,
                               hipStream_t stream) {
  if (slots)
    hipLaunchKernelGGL(predecessors_kernel, dim3(slots), dim3(256), 0, stream, g, b, q, copy, out);
  return hipGetLastError();
}

hipError_t scatter_launch(const DevGraph& g, const DevBatch& b, std::uint32_t q,
                          std::uint32_t slots, std::uint32_t copy, std::uint32_t* out,
                          hipStream_t stream) {
  if (slots)
    hipLaunchKernelGGL(scatter_distances_kernel, dim3(slots), dim3(256), 0, stream, g, b, q, copy,
                       out);
  return hipGetLastError();
}

// Copies query q's final labels tile by tile, in the order of its sorted
// selected tiles, to out[offsets[i] ...] (i = position within the query).
__global__ void gather_region_kernel(DevGraph g, DevBatch b, std::uint32_t q,
                                     std::uint32_t jacobi_copy, const unsigned long long* offsets,
                                     std::uint32_t* out) {
  const std::uint32_t i = blockIdx.x;
  const std::uint32_t pos = b.tile_begin[q] + i;
  if (pos >= b.tile_end[q]) return;
  const std::uint32_t s = slot_at(b, pos);
  const std::uint32_t T = b.tiles[s];
  const std::uint32_t npt = g.tile_vertex_offsets[T + 1] - g.tile_vertex_offsets[T];
  const std::uint32_t* src =
      b.dist + b.slot_dist[s] + static_cast<unsigned long long>(column_of(b, q, jacobi_copy)) * npt;
  std::uint32_t* dst = out + offsets[i];
  for (std::uint32_t k = threadIdx.x; k < npt; k += blockDim.x) dst[k] = src[k];
}

hipError_t gather_region_launch(const DevGraph& g, const DevBatch& b, std::uint32_t q,
                                std::uint32_t slots, std::uint32_t copy,
                                const unsigned long long* offsets, std::uint32_t* out,
                                hipStream_t stream) {
  if (slots)
    hipLaunchKernelGGL(gather_region_kernel, dim3(slots), dim3(256), 0, stream, g, b, q, copy,
                       offsets, out);
  return hipGetLastError();
}

hipError_t finalize_launch(const DevGraph& g, const DevBatch& b,
                           const unsigned long long* relaxations, void* out, hipStream_t stream) {
  hipLaunchKernelGGL(finalize_kernel, dim3((b.query_count + 127) / 128), dim3(128), 0, stream, g,
                     b, relaxations, static_cast<QueryStatusDev*>(out));
  return hipGetLastError();
}

hipError_t relaxation_count_launch(const DevGraph& g, const DevBatch& b,
                                   unsigned long long* out, hipStream_t stream) {
  if (b.slot_count)
    hipLaunchKernelGGL(relaxation_count_kernel, dim3(b.slot_count), dim3(256), 0, stream, g, b,
                       out);
  return hipGetLastError();
}

hipError_t occupancy_launch(const std::uint32_t* arena, unsigned long long n, int* occ,
                            hipStream_t stream) {
  if (n) hipLaunchKernelGGL(occupancy_kernel, dim3(1024), dim3(256), 0, stream, arena, n, occ);
  return hipGetLastError();
}

hipError_t node_cost_launch(const DevGraph& g, const float* base, const int* occ, float* csr_w,
                            float* csc_w, hipStream_t stream) {
  hipLaunchKernelGGL(node_cost_kernel, dim3((g.vertex_count + 255) / 256), dim3(256), 0, stream, g,
                     base, occ, csr_w, csc_w);
  return hipGetLastError();
}

}  // namespace bfn
