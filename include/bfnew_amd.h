/*
 * bfnew_amd.h -- C ABI of the MI355X batched shortest-path engine.
 *
 * The reference has no C ABI: its batched relaxation path sits behind C++
 * engine classes (include/bfnew/hip/batched_jacobi.hpp:137-200,
 * batched_dense_chaotic_push.hpp:105-165, batched_frontier_push.hpp:91-144),
 * the single-query GpuSsspEngine::run (include/bfnew/gpu_api.hpp:203-211) and
 * the resident graph / workspace owners (include/bfnew/hip/runtime.hpp:238-327).
 * Each entry point below replaces one step of those surfaces; the citation on
 * each says which.  Everything is plain pointers, sizes and POD records; the
 * records that the reference copies across its host/device boundary keep the
 * reference's byte layout (gpu_api.hpp:224-251, compact_path_types.hpp:44-47).
 *
 * Conventions: every function returns an int32 status (0 = success, see
 * BFN_E_*); bfn_last_error() returns a thread-local message for the last
 * failure on the calling thread.  Handles are single-stream and are not
 * thread-safe, like the reference's engines (runtime.hpp:315-320); a server
 * handle owns its own thread.
 */
#ifndef BFNEW_AMD_H
#define BFNEW_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define BFN_OK 0
#define BFN_E_INVALID_ARGUMENT 1   /* std::invalid_argument in the reference */
#define BFN_E_HIP 2                /* HipRuntimeError (runtime.hpp:22-35) */
#define BFN_E_OUT_OF_MEMORY 3      /* admission/budget refusal */
#define BFN_E_NOT_REPRESENTABLE 4  /* graph shape exceeds the tile layout */
#define BFN_E_STATE 5              /* call out of order (no batch, not done) */
#define BFN_E_DEVICE_RESULT 6      /* std::runtime_error: inconsistent result */

/* ---- records shared with the reference ABI ------------------------------ */
/* GpuRunOptions, gpu_api.hpp:57-67 (48 bytes). */
typedef struct {
  uint32_t engine;               /* 0 jacobi_pull, 1 dense_chaotic_push, 2 frontier_push */
  uint32_t control_mode;         /* 0 persistent_cooperative, 1 chunked, 2 per_round */
  uint32_t rounds_per_chunk;
  uint32_t block_size;
  uint32_t grid_policy;          /* 0 occupancy_derived, 1 fixed_blocks_per_wgp */
  uint32_t blocks_per_wgp;
  uint32_t instrumentation;      /* 0 none, 1 light, 2 debug */
  uint32_t pad0;
  uint64_t maximum_rounds;
  uint32_t enable_per_lane_convergence;
  uint32_t pad1;
} bfn_run_options;

/* DeviceRunStatus, gpu_api.hpp:164-175 (48 bytes).  Lane masks cover the
 * first 32 queries of a batch, as in the reference's u32 LaneMask. */
typedef struct {
  uint32_t final_distance_slot;
  uint32_t converged;
  uint64_t rounds_completed;
  uint32_t reached_target_mask;
  uint32_t bounding_box_miss_mask;
  uint32_t valid_lane_mask;
  uint32_t active_lane_mask;
  uint32_t converged_lane_mask;
  uint32_t stop_reason;          /* DeviceStopReason, gpu_api.hpp:37-44 */
  uint32_t error_bits;           /* device_error bits, gpu_api.hpp:46-55 */
  uint32_t pad;
} bfn_run_status;

/* DeviceWorkStatistics, gpu_api.hpp:134-162 (168 bytes). */
typedef struct {
  uint64_t edges_examined;
  uint64_t successful_decreases;
  uint64_t active_vertices;
  uint64_t active_lane_rounds;
  uint64_t maximum_queue_size;
  uint64_t host_checks;
  uint64_t host_synchronizations;
  uint64_t controller_copies;
  uint64_t kernel_dispatches;
  uint64_t expansion_count;
  uint64_t atomic_attempts;
  uint64_t successful_atomic_updates;
  uint64_t queue_claims;
  uint64_t duplicate_suppressions;
  uint64_t mask_operations;
  uint64_t overflow_events;
  uint64_t high_contention_destinations;
  uint64_t changed_flag_updates;
  uint64_t full_edge_rounds;
  uint64_t empty_frontier_rounds;
  uint64_t small_frontier_rounds;
} bfn_work_statistics;

/* CompactTargetSummary, compact_path_types.hpp:31-47 (28 bytes). */
typedef struct {
  uint32_t target;
  uint32_t selected_source;
  float distance;
  uint32_t path_length;          /* edges */
  uint32_t reached;              /* CompactTargetReachStatus */
  uint32_t reconstruction;       /* CompactPathStatus: 0 complete, 1 unreachable,
                                    2 query_terminal_failure, 3 no_tight_path,
                                    4 path_length_overflow */
  uint32_t has_selected_source;
} bfn_target_summary;

/* ---- MI355X-side records (no reference counterpart) -------------------- */
/* Per-query outcome of a batch (the reference keeps these as bits of the
 * batch's LaneMask fields; a launch here holds hundreds of queries). */
typedef struct {
  uint64_t rounds_completed;     /* Jacobi: changing rounds + final no-change round */
  uint64_t relaxations;          /* admitted edges with a reachable source */
  uint32_t converged;
  uint32_t reached_all_targets;
  uint32_t bounding_box_miss;
  uint32_t stop_reason;
  uint32_t error_bits;
  uint32_t engine;
} bfn_query_status;

/* Engine tuning knobs chosen by the autotuner (bfn_tuning_default gives the
 * untuned defaults). */
typedef struct {
  uint32_t tile_sweeps;          /* Jacobi: 1 = point Jacobi (reference round
                                    semantics), 0 = sweep each tile to its local
                                    fixed point every round (block Jacobi) */
  uint32_t items_per_claim;      /* work items taken per queue claim */
  uint32_t graph_rounds;         /* rounds captured per hipGraph replay */
  uint32_t use_graph;            /* replay the batch step through a hipGraph */
  uint32_t count_relaxations;    /* compute per-query relaxation counts */
  uint32_t boundary_mode;        /* push engines: 0 = aggregate boundary pushes
                                    per destination in LDS, then one coalesced
                                    compare and an atomic per improved label
                                    (when the LDS holds the proposals);
                                    1 = per boundary edge, a read of the
                                    neighbour's label, then a global atomic
                                    when the candidate is below it */
  uint32_t min_waves;            /* 0 = compiler default; 6, 7 or 8 = the 256-
                                    and 512-thread kernels compiled for at least
                                    that many waves per SIMD */
  uint32_t claim_order;          /* per-round launches only: 0 = work items in
                                    arrival order (one list per round); 1 = work
                                    items in slot order (tile-major), split into
                                    eight segments that the workgroups of each
                                    XCD claim first (queued as bits, compacted
                                    between rounds) */
  float push_window;             /* push engines: 0 = a round visits every queued
                                    slot; w > 0 = a round visits only the queued
                                    slots whose smallest label pushed into them
                                    is at most m + w, m the smallest such label
                                    over the queue (the others stay queued for a
                                    later round): fewer repeated visits, more
                                    rounds.  Jacobi pull ignores it. */
} bfn_tuning;

/* The reference DeviceGraphLayout32 (device_layout.hpp:72-93) plus the tile
 * vertex offsets its WeightedGraph carries (graph.hpp:226-228). */
typedef struct {
  uint32_t vertex_count;
  uint32_t edge_count;
  uint32_t tile_count;
  uint32_t csr_run_count;
  uint32_t csc_run_count;
  uint32_t pad;
  const uint32_t *owner_tiles;             /* [V] */
  const uint32_t *tile_vertex_offsets;     /* [tile_count+1] */
  const uint32_t *csr_row_offsets;         /* [V+1] */
  const uint32_t *csr_destinations;        /* [E] */
  const float *csr_weights;                /* [E] */
  const uint32_t *csr_row_run_offsets;     /* [V+1] */
  const uint32_t *csr_run_edge_offsets;    /* [Rcsr+1] */
  const uint32_t *csr_run_destination_tiles; /* [Rcsr] */
  const uint32_t *csc_column_offsets;      /* [V+1] */
  const uint32_t *csc_sources;             /* [E] */
  const float *csc_weights;                /* [E] */
  const uint32_t *csc_edge_ids;            /* [E] */
  const uint32_t *csc_column_run_offsets;  /* [V+1] */
  const uint32_t *csc_run_edge_offsets;    /* [Rcsc+1] */
  const uint32_t *csc_run_source_tiles;    /* [Rcsc] */
} bfn_graph_layout32;

/* A batch: the POD projection of BatchDeviceDescription
 * (batch_layout.hpp:73-120) generalised from <=32 lanes to any number of
 * queries.  Sources and targets are canonical (sorted, deduplicated) global
 * vertex ids; selected tiles are sorted per query. */
typedef struct {
  uint32_t query_count;
  uint32_t pad;
  const uint32_t *query_ids;        /* [Q] */
  const uint32_t *source_offsets;   /* [Q+1] */
  const uint32_t *sources;
  const uint32_t *target_offsets;   /* [Q+1] */
  const uint32_t *targets;
  const uint32_t *tile_offsets;     /* [Q+1] */
  const uint32_t *selected_tiles;
  const uint32_t *engine_by_query;  /* [Q] or NULL: per-query engine choice */
} bfn_batch_desc;

typedef struct {
  uint64_t device_bytes;            /* resident graph bytes */
  uint32_t max_tile_vertices;
  uint32_t max_tile_in_edges;
  uint32_t max_tile_out_edges;
  uint32_t max_in_neighbor_tiles;
  uint32_t max_out_neighbor_tiles;
  uint32_t max_stage_floats;
  uint32_t has_zero_weight;
  uint32_t pad;
} bfn_graph_info;

typedef struct {
  uint32_t compute_units;
  uint32_t xcd_count;
  uint32_t wavefront_size;
  uint32_t lds_bytes_per_cu;
  uint64_t global_memory_bytes;
  char arch[32];
} bfn_device_info;

typedef struct bfn_device_s *bfn_device;
typedef struct bfn_graph_s *bfn_graph;
typedef struct bfn_workspace_s *bfn_workspace;

const char *bfn_last_error(void);
uint32_t bfn_abi_version(void);

/* Device and stream (HipStream, runtime.hpp:37-54). */
int32_t bfn_device_open(int32_t ordinal, bfn_device *out);
int32_t bfn_device_close(bfn_device device);
int32_t bfn_device_query(bfn_device device, bfn_device_info *out);
int32_t bfn_device_synchronize(bfn_device device);

/* ResidentDeviceGraph::upload_once_async + synchronize_upload
 * (runtime.hpp:238-270, resident_graph.hip.cpp:315).  The caller's host
 * arrays may be released when the call returns.  Derives the tile-major
 * engine image (tile neighbour lists, packed tile-local CSC/CSR) on device. */
int32_t bfn_graph_upload(bfn_device device, const bfn_graph_layout32 *layout,
                         bfn_graph *out);
/* The host checks bfn_graph_upload runs before it touches the device
 * (validate_device_graph_layout32 / validate_tile_run_layout,
 * device_layout.cpp:253-393): offsets, owner tiles, endpoints, nonnegative
 * weights and the tile runs of both orientations (in range, non-empty,
 * maximal, covering each row in order, each edge's endpoint in its run's
 * tile).  BFN_E_INVALID_ARGUMENT or BFN_E_NOT_REPRESENTABLE on failure;
 * needs no GPU. */
int32_t bfn_graph_validate(const bfn_graph_layout32 *layout);
int32_t bfn_graph_free(bfn_graph graph);
int32_t bfn_graph_query(bfn_graph graph, bfn_graph_info *out);
/* Replaces every edge weight; both orientations in the layout's orders. */
int32_t bfn_graph_update_weights(bfn_graph graph, const float *csr_weights,
                                 const float *csc_weights);
/* Negotiated congestion: on device, every edge u->v gets
 * w = base_cost[v] * (1.0f + 0.5f * (float)occupancy[v]) with that exact fp32
 * evaluation order.  occupancy may be NULL (all zero) or the batch's own
 * occupancy counts (bfn_batch_occupancy writes them on device). */
int32_t bfn_graph_update_node_costs(bfn_graph graph, const float *base_cost,
                                    const int32_t *occupancy);
/* Device-resident variant used between negotiated-cost rounds: counts the
 * occupancy of the workspace's last compact paths (bfn_batch_paths) on device
 * and applies it (BFN_E_STATE when the workspace holds no paths). */
int32_t bfn_graph_update_node_costs_from_workspace(bfn_graph graph,
                                                   bfn_workspace workspace,
                                                   const float *base_cost);

/* Workspace (ReusableBatched*Workspace, batched_workspace.hpp; admission by
 * budget as estimate_batch_workspace, batch_workspace.cpp:163). */
int32_t bfn_workspace_create(bfn_device device, uint64_t budget_bytes, bfn_workspace *out);
int32_t bfn_workspace_free(bfn_workspace workspace);
void bfn_tuning_default(bfn_tuning *out);
int32_t bfn_batch_estimate(bfn_graph graph, const bfn_batch_desc *batch,
                           const bfn_run_options *options, uint64_t *bytes);

/* Batched*Engine::run_* (batched_jacobi.hpp:137-200): prepares the batch
 * (upload + device-side slot mapping) and runs the relaxation on the
 * workspace stream.  Asynchronous; bfn_batch_wait completes it. */
int32_t bfn_batch_submit(bfn_workspace workspace, bfn_graph graph, const bfn_batch_desc *batch,
                         const bfn_run_options *options, const bfn_tuning *tuning);
int32_t bfn_batch_wait(bfn_workspace workspace);
/* Re-runs the batch already resident in the workspace (same queries, same
 * options) from its initial labels: the benchmark's step. */
int32_t bfn_batch_rerun(bfn_workspace workspace);
/* Batch status as DeviceRunStatus plus one record per query. */
int32_t bfn_batch_status(bfn_workspace workspace, bfn_run_status *status,
                         bfn_query_status *per_query, bfn_work_statistics *work);
/* run_with_distances: one query's labels scattered to global ids (+inf
 * outside its selected tiles). */
int32_t bfn_batch_distances(bfn_workspace workspace, uint32_t query, float *out);
/* One query's final labels over its selected tiles only, tile by tile in
 * the order of its sorted selected tiles (each tile's vertices in id order):
 * the compact form of run_with_distances that DeviceCompactDistanceMatrix
 * (include/bfnew/hip/compact_path_results.hpp:23-31) keeps per lane.  With
 * out == NULL only *count (the label count) is written. */
int32_t bfn_batch_region_distances(bfn_workspace workspace, uint32_t query, float *out,
                                   uint64_t capacity, uint64_t *count);
/* Smallest-stable-EdgeId tight admitted predecessor per vertex (0xFFFFFFFF
 * for sources, unreached and unselected vertices). */
int32_t bfn_batch_predecessors(bfn_workspace workspace, uint32_t query, uint32_t *out);
/* run_compact_paths / ReusableCompactPathWorkspace::reconstruct
 * (compact_path_results.hpp:102-134).  Reconstructs every target of every
 * query on device.  summaries has one entry per batch target (query order,
 * canonical target order); path_offsets[i]..path_offsets[i+1] index the
 * vertex and label arenas; target i's path_length edges start at
 * path_offsets[i] in the edge arena (which has the vertex arena's size).
 * Pass NULL arenas to learn *vertex_total first; the summaries and offsets
 * are computed once per completed batch, so the second call, with the
 * arenas, only emits the paths. */
int32_t bfn_batch_paths(bfn_workspace workspace, bfn_target_summary *summaries,
                        uint64_t *path_offsets, uint32_t *vertices, float *labels,
                        uint32_t *edge_ids, uint64_t arena_capacity, uint64_t *vertex_total);
/* Per-vertex count of completed paths through each vertex (int32[V]),
 * accumulated on device from the last bfn_batch_paths call. */
int32_t bfn_batch_occupancy(bfn_workspace workspace, int32_t *out);
/* Device time (ms) of the last batch's relaxation, measured with HIP events
 * on the workspace stream; and the number of relaxation kernel launches. */
int32_t bfn_batch_timing(bfn_workspace workspace, float *relax_ms, float *total_ms,
                         uint64_t *launches);

/* Engine counters of the last batch, collected when the run options ask for
 * instrumentation (light or debug); zero otherwise. */
typedef struct {
  uint64_t tile_visits;          /* (query, tile) work items processed */
  uint64_t edges_relaxed;        /* edge records evaluated by those visits */
  uint64_t algorithmic_bytes;    /* modelled HBM bytes: edge records, staged and
                                    written label slabs, remote label reads and
                                    atomics */
  uint64_t atomic_attempts;
  uint64_t atomic_successes;
  uint64_t maximum_list;
  uint64_t rounds_executed;
  uint64_t closure_sweeps;       /* tile-closure sweeps over internal records */
  uint64_t phase_cycles[5];      /* shader-clock cycles per visit phase summed
                                    over workgroups: setup and label staging,
                                    internal records, tile closure, boundary
                                    records, write-back and enqueue */
  uint64_t candidate_evaluations;  /* candidate labels fl(d[u] + w) computed:
                                    internal edges from a frontier source per
                                    closure sweep (every edge per pull pass),
                                    plus boundary candidates */
} bfn_engine_counters;
int32_t bfn_batch_counters(bfn_workspace workspace, bfn_engine_counters *out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the batch of bfn_batch_submit run to completion on the
 * graph's stream with no allocation at all.  Every piece of per-batch device
 * state lives in one static device array per GPU of
 * bfn_static_scratch_bytes() bytes (64 MiB), held until the process ends;
 * a batch that needs more returns BFN_E_OUT_OF_MEMORY before touching it.
 * The batch is copied in from the caller's host arrays; per_query
 * ([query_count]) and target_distances ([target_offsets[query_count]],
 * +inf when unreached) are copied out before the call returns.  Tuning as
 * bfn_batch_submit, except that use_graph and count_relaxations are refused
 * and push engines always push boundary edges with direct atomics.  Every
 * call on one GPU shares that GPU's array and keeps the last call's
 * contents; calls that overlap on one GPU race on it. */
int32_t bfn_batch_run_static(bfn_graph graph, const bfn_batch_desc *batch,
                             const bfn_run_options *options, const bfn_tuning *tuning,
                             bfn_query_status *per_query, float *target_distances);
uint64_t bfn_static_scratch_bytes(void);

/* Serving loop (continuous batching, north of the reference's single-batch
 * engines): a server owns a thread and a stream over a resident graph.
 * Requests are admitted by free workspace memory and join the running
 * persistent batch between chunks. */
typedef struct bfn_server_s *bfn_server;
typedef struct {
  uint64_t budget_bytes;
  uint32_t max_active_queries;
  uint32_t rounds_per_chunk;
  uint32_t engine;
  uint32_t tile_sweeps;
  uint32_t block_size;
  uint32_t want_paths;
} bfn_server_options;
typedef struct {
  uint64_t request_id;
  uint32_t source;
  uint32_t target;
  uint32_t tile_count;
  uint32_t pad;
  const uint32_t *selected_tiles;  /* sorted */
} bfn_request;
typedef struct {
  uint64_t request_id;
  float distance;
  uint32_t path_length;
  uint32_t reached;
  uint32_t pad;
  double submit_ns;                /* steady-clock stamps, host */
  double complete_ns;
} bfn_response;
int32_t bfn_server_create(bfn_device device, bfn_graph graph, const bfn_server_options *options,
                          bfn_server *out);
int32_t bfn_server_submit(bfn_server server, const bfn_request *request);
/* Blocks for up to timeout_ms for responses; returns the count in *count. */
int32_t bfn_server_poll(bfn_server server, bfn_response *responses, uint32_t capacity,
                        uint32_t *count, uint32_t timeout_ms);
/* bfn_server_poll with each response's compact path (servers created with
 * want_paths): response i's path vertices, source first, are
 * vertices[path_offsets[i] .. path_offsets[i+1]) and its path_length edge ids
 * start at edge_ids[path_offsets[i]] (an unreached target has an empty
 * path).  path_offsets has capacity + 1 entries; responses are returned while
 * their paths fit in vertex_capacity (BFN_E_INVALID_ARGUMENT when the next
 * path alone does not). */
int32_t bfn_server_poll_paths(bfn_server server, bfn_response *responses, uint32_t capacity,
                              uint32_t *count, uint32_t timeout_ms, uint32_t *vertices,
                              uint32_t *edge_ids, uint64_t vertex_capacity,
                              uint64_t *path_offsets);
int32_t bfn_server_stats(bfn_server server, uint64_t *launches, uint64_t *admitted,
                         uint64_t *completed);
int32_t bfn_server_destroy(bfn_server server);

#ifdef __cplusplus
}
#endif

#endif
