// relax.hip -- control of the relaxation engines: the advance kernels
// between round launches, batch preparation (control reset, label fill,
// slot tables, seeds) and the launch dispatch to the kernel variant of a
// batch's engine set (relax_kernels.hpp, instantiated by relax_mixed.hip,
// relax_push.hip and relax_pull.hip).
#include "bfn_launch.hpp"

#include <mutex>

namespace bfn {

__global__ void relax_advance_kernel(DevBatch b) {
  DevControl* c = b.control;
  if (c->done) return;
  const std::uint32_t r = c->round;
  const std::uint32_t n = c->list_count[r % 3];
  if (b.instrumentation && n > c->max_list) c->max_list = n;
  c->list_count[(r + 2) % 3] = 0;
  c->claim[(r + 2) % 3] = 0;
  if (b.prio) {
    c->window_limit = next_window_limit(c, b, r, c->window_limit);
    reset_next_min(c, (r + 2) % 3);
  }
  c->round = r + 1;
  c->rounds_executed = r + 1;
  if (c->list_count[(r + 1) % 3] == 0) {
    c->done = 1;
    c->stop_reason = kStopConverged;
  } else if (b.maximum_rounds && r + 1 >= b.maximum_rounds) {
    c->done = 1;
    c->stop_reason = kStopMaximumRounds;
  }
}

// Between rounds in slot order (one workgroup): count the round as
// relax_advance_kernel does, then compact the queued bits into lists[0, n)
// and clear them.  Slots are taken in groups of 64 (two bit words); group k
// belongs to segment k % 8, and the list holds segment 0's groups in order,
// then segment 1's, ..., so each segment is in slot order.  first = 1 (after
// seeding) compacts only.
constexpr std::uint32_t kCompactGroupsPerThread = 8;

__global__ __launch_bounds__(1024) void ordered_advance_kernel(DevBatch b, std::uint32_t first) {
  __shared__ std::uint32_t scan[1024];
  __shared__ std::uint32_t seg_start[kBarrierGroups];
  DevControl* c = b.control;
  if (c->done) return;
  const std::uint32_t tid = threadIdx.x;
  const std::uint32_t W = (b.slot_count + 31u) / 32u;
  const std::uint32_t NG = (W + 1u) / 2u;
  // Groups of segment x: x, x + 8, ...; its first position in the virtual
  // order (segment 0's groups first) is vstart[x].
  std::uint32_t vstart[kBarrierGroups + 1];
  vstart[0] = 0;
#pragma unroll
  for (std::uint32_t x = 0; x < kBarrierGroups; ++x)
    vstart[x + 1] = vstart[x] + (NG > x ? (NG - 1u - x) / kBarrierGroups + 1u : 0u);
  if (tid < kBarrierGroups) seg_start[tid] = kNone;
  std::uint32_t carry = 0;
  for (std::uint32_t base = 0; base < NG; base += 1024u * kCompactGroupsPerThread) {
    const std::uint32_t v0 = base + tid * kCompactGroupsPerThread;
    std::uint32_t w0[kCompactGroupsPerThread], w1[kCompactGroupsPerThread],
        grp[kCompactGroupsPerThread], seg[kCompactGroupsPerThread];
    std::uint32_t total = 0;
#pragma unroll
    for (std::uint32_t j = 0; j < kCompactGroupsPerThread; ++j) {
      const std::uint32_t v = v0 + j;
      w0[j] = w1[j] = 0;
      grp[j] = kNone;
      seg[j] = 0;
      if (v < NG) {
        std::uint32_t x = 0;
        while (v >= vstart[x + 1]) ++x;
        seg[j] = x;
        grp[j] = x + kBarrierGroups * (v - vstart[x]);
        w0[j] = b.active[2 * grp[j]];
        if (2 * grp[j] + 1 < W) w1[j] = b.active[2 * grp[j] + 1];
      }
    }
#pragma unroll
    for (std::uint32_t j = 0; j < kCompactGroupsPerThread; ++j) total += __popc(w0[j]) + __popc(w1[j]);
    scan[tid] = total;
    __syncthreads();
    for (std::uint32_t off = 1; off < 1024; off <<= 1) {
      const std::uint32_t add = tid >= off ? scan[tid - off] : 0u;
      __syncthreads();
      scan[tid] += add;
      __syncthreads();
    }
    std::uint32_t pos = carry + scan[tid] - total;
    const std::uint32_t chunk_total = scan[1023];
#pragma unroll
    for (std::uint32_t j = 0; j < kCompactGroupsPerThread; ++j) {
      if (grp[j] == kNone) continue;
      if (v0 + j == vstart[seg[j]]) seg_start[seg[j]] = pos;
      const std::uint32_t s0 = 64u * grp[j];
      for (std::uint32_t m = w0[j]; m; m &= m - 1u) b.lists[pos++] = s0 + __ffs(m) - 1u;
      for (std::uint32_t m = w1[j]; m; m &= m - 1u) b.lists[pos++] = s0 + 32u + __ffs(m) - 1u;
      if (w0[j]) b.active[2 * grp[j]] = 0;
      if (w1[j]) b.active[2 * grp[j] + 1] = 0;
    }
    carry += chunk_total;
    __syncthreads();
  }
  __syncthreads();
  if (tid == 0) {
    const std::uint32_t n = carry;
    // Segments without groups (NG < 8) are empty and sit at the end.
    for (std::uint32_t x = 0; x < kBarrierGroups; ++x) {
      const std::uint32_t at = seg_start[x] == kNone ? n : seg_start[x];
      c->seg_begin[x] = at;
      c->seg_claim[x * kBarrierStride] = at;
    }
    c->seg_begin[kBarrierGroups] = n;
    c->list_count[0] = n;
    if (b.instrumentation && n > c->max_list) c->max_list = n;
    if (!first) {
      const std::uint32_t r = c->round;
      if (b.prio) {
        c->window_limit = next_window_limit(c, b, r, c->window_limit);
        reset_next_min(c, (r + 2) % 3);
      }
      c->round = r + 1;
      c->rounds_executed = r + 1;
      if (n == 0) {
        c->done = 1;
        c->stop_reason = kStopConverged;
      } else if (b.maximum_rounds && r + 1 >= b.maximum_rounds) {
        c->done = 1;
        c->stop_reason = kStopMaximumRounds;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Batch preparation.
// ---------------------------------------------------------------------------
__global__ void reset_control_kernel(DevControl* c, std::uint32_t window) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    c->window_limit = window;  // round 0 visits the sources' slots (priority 0)
    for (std::uint32_t i = 0; i < 3 * kBarrierGroups * kBarrierStride; ++i) c->next_min[i] = kInfBits;
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
    c->candidates = 0;
    for (int p = 0; p < kPhaseCount; ++p) c->phase_cycles[p] = 0;
    for (std::uint32_t i = 0; i < kMaxArrivalGroups * kBarrierStride; ++i) c->group_arrivals[i] = 0;
  }
}

__global__ void fill_u32_kernel(std::uint32_t* p, unsigned long long n, std::uint32_t v) {
  const unsigned long long stride = static_cast<unsigned long long>(gridDim.x) * blockDim.x;
  for (unsigned long long i = static_cast<unsigned long long>(blockIdx.x) * blockDim.x +
                              threadIdx.x;
       i < n; i += stride)
    p[i] = v;
}

// Per run: each slot's queue stamp, window priority and its query's last
// change, and the engine-set check.  With `structure` (the first run of a
// batch) also the slot's in/out neighbour tiles resolved to the query's
// slots and its descriptor: they depend only on the batch and the graph's
// tile directory, so the reruns of a batch keep them.
__global__ void prep_slots_kernel(DevGraph g, DevBatch b, std::uint32_t structure) {
  const std::uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= b.slot_count) return;
  const std::uint32_t q = b.slot_query[s];
  if (structure) {
    const std::uint32_t T = b.tiles[s];
    const std::uint32_t in0 = g.in_nbr_offsets[T], in1 = g.in_nbr_offsets[T + 1];
    const std::uint32_t ib = b.slot_in_base[s];
    for (std::uint32_t k = in0; k < in1; ++k)
      b.slot_in[ib + (k - in0)] = find_slot(b, q, g.in_nbr_tiles[k]);
    const std::uint32_t o0 = g.out_nbr_offsets[T], o1 = g.out_nbr_offsets[T + 1];
    const std::uint32_t obs = b.slot_out_base[s];
    for (std::uint32_t k = o0; k < o1; ++k) {
      const std::uint32_t ns = find_slot(b, q, g.out_nbr_tiles[k]);
      b.slot_out[obs + (k - o0)] = ns;
      if (b.slot_out_off) b.slot_out_off[obs + (k - o0)] = ns == kNone ? ~0ull : b.slot_dist[ns];
    }
  }
  b.stamp[s] = 0;  // also every queue-bit word of slot-order claims (ceil(S/32) <= S)
  if (b.prio) b.prio[s] = kInfBits;
  // the relaxation kernel variant must hold the engine of every slot
  const bool pull = b.engine[q] == kJacobiPull;
  if ((b.engines == kEnginesPush && pull) || (b.engines == kEnginesPull && !pull))
    atomicOr(&b.control->error_bits, 4u);
  b.last_change[q] = 0;
  if (structure && b.slot_desc) const_cast<SlotDesc*>(b.slot_desc)[s] = make_slot_desc(g, b, s);
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
    if (b.prio && !jacobi) atomic_min_agent(&b.prio[s], 0u);
    enqueue(b, s, 0);
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

// Relaxation launch: the kernel variant of the batch's engine set, then
// (one round per launch) the advance kernel.
hipError_t relax_launch(std::uint32_t block, bool persistent, bool cooperative,
                        std::uint32_t grid, std::size_t lds_bytes, std::uint32_t max_v,
                        std::uint32_t work_words, const DevGraph& g, const DevBatch& b,
                        hipStream_t stream) {
  hipError_t e;
  switch (b.engines) {
    case kEnginesPush:
      e = relax_launch_push(block, persistent, cooperative, grid, lds_bytes, max_v, work_words, g, b, stream);
      break;
    case kEnginesPull:
      e = relax_launch_pull(block, persistent, cooperative, grid, lds_bytes, max_v, work_words, g, b, stream);
      break;
    default:
      e = relax_launch_mixed(block, persistent, cooperative, grid, lds_bytes, max_v, work_words, g, b, stream);
  }
  if (e != hipSuccess || persistent) return e;
  if (b.ordered)
    hipLaunchKernelGGL(ordered_advance_kernel, dim3(1), dim3(1024), 0, stream, b, 0u);
  else
    hipLaunchKernelGGL(relax_advance_kernel, dim3(1), dim3(1), 0, stream, b);
  return hipGetLastError();
}

hipError_t relax_occupancy(std::uint32_t engines, std::uint32_t block, bool persistent,
                           bool instrumented, std::uint32_t min_waves, std::size_t lds_bytes,
                           int* blocks_per_cu) {
  switch (engines) {
    case kEnginesPush:
      return relax_occupancy_push(block, persistent, instrumented, min_waves, lds_bytes, blocks_per_cu);
    case kEnginesPull:
      return relax_occupancy_pull(block, persistent, instrumented, min_waves, lds_bytes, blocks_per_cu);
    default:
      return relax_occupancy_mixed(block, persistent, instrumented, min_waves, lds_bytes, blocks_per_cu);
  }
}

// Workgroups per CU of a persistent grid that the device places at once.
// The grid's workgroups wait for each other at every grid barrier, so all of
// them must be resident together; the occupancy query has promised more than
// the device places (on MI355X: 7 per CU for the 256-thread Jacobi-only
// kernel, whose grid then never passed its first barrier).  So the count is
// checked by a probe launch of the same kernel variant with the same LDS:
// every workgroup passes one grid barrier with a short bounded wait and
// exits; a timeout lowers the count by one.  Results are kept per device,
// variant and LDS size in a fixed table (no allocation).
namespace {
struct ProbeResult {
  int device;
  std::uint32_t engines, block, min_waves, instrumented;
  std::size_t lds_bytes;
  int promised, placed;
};
constexpr int kProbeTable = 64;
ProbeResult g_probe[kProbeTable];
int g_probe_count = 0;
std::mutex g_probe_mutex;
}  // namespace

hipError_t coresident_per_cu(int device, int promised, std::uint32_t cus, std::size_t lds_bytes,
                             std::uint32_t max_v, std::uint32_t work_words, const DevGraph& g,
                             const DevBatch& b, hipStream_t stream, std::uint32_t block,
                             int* placed) {
  std::lock_guard<std::mutex> lock(g_probe_mutex);
  for (int i = 0; i < g_probe_count; ++i) {
    const ProbeResult& p = g_probe[i];
    if (p.device == device && p.engines == b.engines && p.block == block &&
        p.min_waves == b.min_waves && p.instrumented == (b.instrumentation != 0) &&
        p.lds_bytes == lds_bytes && p.promised == promised) {
      *placed = p.placed;
      return hipSuccess;
    }
  }
  DevBatch probe{};
  probe.engines = b.engines;
  probe.min_waves = b.min_waves;
  probe.instrumentation = b.instrumentation;
  probe.control = b.control;
  probe.probe = 1;
  int per_cu = promised;
  for (; per_cu >= 1; --per_cu) {
    hipLaunchKernelGGL(reset_control_kernel, dim3(1), dim3(64), 0, stream, b.control, 0u);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = relax_launch(block, true, true, static_cast<std::uint32_t>(per_cu) * cus, lds_bytes, max_v,
                       work_words, g, probe, stream);
    std::uint32_t bits = 0;
    if (e == hipSuccess)
      e = hipMemcpyAsync(&bits, &b.control->error_bits, sizeof bits, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return e;
    if ((bits & kErrorBarrierTimeout) == 0) break;
  }
  if (g_probe_count < kProbeTable)
    g_probe[g_probe_count++] = ProbeResult{device, b.engines, block, b.min_waves,
                                           b.instrumentation != 0u, lds_bytes, promised, per_cu};
  *placed = per_cu;
  return hipSuccess;
}

hipError_t prep_launch(const DevGraph& g, const DevBatch& b, std::uint64_t pool_words,
                       hipStream_t stream, bool structure) {
  hipLaunchKernelGGL(reset_control_kernel, dim3(1), dim3(64), 0, stream, b.control, b.window);
  const std::uint32_t fill_grid = 4096;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(fill_grid), dim3(256), 0, stream, b.dist,
                     static_cast<unsigned long long>(pool_words), kInfBits);
  if (b.slot_count)
    hipLaunchKernelGGL(prep_slots_kernel, dim3((b.slot_count + 255) / 256), dim3(256), 0,
                       stream, g, b, structure ? 1u : 0u);
  if (b.query_count)
    hipLaunchKernelGGL(seed_kernel, dim3((b.query_count + 127) / 128), dim3(128), 0, stream, g,
                       b);
  if (b.ordered) hipLaunchKernelGGL(ordered_advance_kernel, dim3(1), dim3(1024), 0, stream, b, 1u);
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
