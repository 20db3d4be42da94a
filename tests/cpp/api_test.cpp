// api_test.cpp -- tests of the reference-named C++ API (gpu_api records and
// rules, batch descriptions, resident graph, batched and single-query
// engines) over the C ABI.
//
//   api_test cpu           rules and host layouts only; no device call
//   api_test planner G.tsv batch planner on the four-query fixture of the
//                          reference's tests/batch_plan_test.cpp:80-160; G.tsv
//                          holds that test's literal serialised plan
//   api_test gpu OUT.bin   runs every engine on the 16x16x64 synthetic graph
//                          (seed 1), checks the engines against each other,
//                          and writes the lanes and labels to OUT.bin for the
//                          oracle comparison done by tests/test_cpp_api.py
//
// The reference states the same rules in tests/gpu_api_test.cpp and its
// batched engine tests; the cases below are this library's own.
#include "bfnew/batch_layout.hpp"
#include "bfnew/batch_plan.hpp"
#include "bfnew/compact_paths.hpp"
#include "bfnew/gpu_api.hpp"
#include "bfnew/hip/batched_expansion.hpp"
#include "bfnew/hip/engines.hpp"
#include "bfnew/hip/runtime.hpp"
#include "bfnew/query.hpp"
#include "bfnew/synthetic.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

using namespace bfnew;

namespace {

int failures = 0;

void check(bool ok, const char* what, int line) {
  if (!ok) {
    std::fprintf(stderr, "FAIL line %d: %s\n", line, what);
    ++failures;
  }
}
#define CHECK(x) check((x), #x, __LINE__)

template <typename E, typename F>
bool throws(F&& f) {
  try {
    f();
  } catch (const E&) {
    return true;
  } catch (...) {
    return false;
  }
  return false;
}

GpuRunOptions options_for(EngineKind kind) {
  GpuRunOptions o;
  o.engine = kind;
  o.control_mode = ControlMode::persistent_cooperative;
  o.maximum_rounds = 1U << 20;
  return o;
}

// ---------------------------------------------------------------------------
void cpu_rules() {
  // Run options.
  GpuRunOptions o;
  CHECK(validate_gpu_run_options(o) == GpuRunOptionsError::none);
  GpuRunOptions bad = o;
  bad.engine = static_cast<EngineKind>(7);
  CHECK(validate_gpu_run_options(bad) == GpuRunOptionsError::invalid_engine);
  bad = o;
  bad.rounds_per_chunk = 0;
  CHECK(validate_gpu_run_options(bad) == GpuRunOptionsError::invalid_rounds_per_chunk);
  bad = o;
  bad.grid_policy = GridPolicy::fixed_blocks_per_wgp;
  CHECK(validate_gpu_run_options(bad) == GpuRunOptionsError::invalid_blocks_per_wgp);
  bad.blocks_per_wgp = 2;
  CHECK(validate_gpu_run_options(bad) == GpuRunOptionsError::none);
  bad = o;
  bad.blocks_per_wgp = 2;  // a derived grid takes no count
  CHECK(validate_gpu_run_options(bad) == GpuRunOptionsError::invalid_blocks_per_wgp);
  bad = o;
  bad.maximum_rounds = 0;
  CHECK(validate_gpu_run_options(bad) == GpuRunOptionsError::invalid_maximum_rounds);
  bad = o;
  bad.enable_per_lane_convergence = 2;
  CHECK(validate_gpu_run_options(bad) == GpuRunOptionsError::invalid_per_lane_flag);

  // Controllers: initial states are valid; each broken invariant is named.
  for (EngineKind k : {EngineKind::jacobi_pull, EngineKind::dense_chaotic_push,
                       EngineKind::frontier_push}) {
    GpuRunOptions ko = o;
    ko.engine = k;
    const DeviceController c = initialize_device_controller(ko, 0x5U, 3U);
    CHECK(validate_device_controller(c) == DeviceControllerError::none);
    CHECK(c.distance_write_slot == (k == EngineKind::jacobi_pull ? 1U : 0U));
  }
  GpuRunOptions fo = o;
  fo.engine = EngineKind::frontier_push;
  CHECK(throws<std::invalid_argument>([&] { (void)initialize_device_controller(fo, 1U, 0U); }));
  CHECK(throws<std::invalid_argument>([&] { (void)initialize_device_controller(o, 0U); }));
  DeviceController c = initialize_device_controller(o, 0x3U);
  DeviceController d = c;
  d.active_lane_mask = 0x7U;
  CHECK(validate_device_controller(d) == DeviceControllerError::mask_outside_valid_lanes);
  d = c;
  d.converged_lane_mask = 0x1U;
  CHECK(validate_device_controller(d) == DeviceControllerError::active_converged_overlap);
  d = c;
  d.distance_write_slot = 0U;
  CHECK(validate_device_controller(d) == DeviceControllerError::invalid_distance_slots);
  d = c;  // a converged stop: no active lane, every lane converged, no error
  d.done = 1U;
  d.execute_lane_mask = 0U;
  d.active_lane_mask = 0U;
  d.converged_lane_mask = 0x3U;
  d.stop_reason = static_cast<std::uint32_t>(DeviceStopReason::converged);
  CHECK(validate_device_controller(d) == DeviceControllerError::none);
  d.error_bits = device_error::queue_overflow;
  CHECK(validate_device_controller(d) == DeviceControllerError::inconsistent_error_state);
  d.error_bits = 0x80U;
  CHECK(validate_device_controller(d) == DeviceControllerError::unknown_error_bits);
  d = c;
  d.done = 1U;
  d.execute_lane_mask = 0U;
  d.stop_reason = static_cast<std::uint32_t>(DeviceStopReason::maximum_rounds);
  CHECK(validate_device_controller(d) == DeviceControllerError::inconsistent_terminal_state);
  d.rounds_completed = d.maximum_rounds;
  CHECK(validate_device_controller(d) == DeviceControllerError::none);

  // Run status.
  DeviceRunStatus s;
  s.valid_lane_mask = 0x3U;
  s.converged = 1U;
  s.converged_lane_mask = 0x3U;
  s.reached_target_mask = 0x1U;
  s.bounding_box_miss_mask = 0x2U;
  s.stop_reason = static_cast<std::uint32_t>(DeviceStopReason::converged);
  CHECK(validate_device_run_status(s) == DeviceRunStatusError::none);
  DeviceRunStatus t = s;
  t.reached_target_mask = 0x3U;
  CHECK(validate_device_run_status(t) == DeviceRunStatusError::reached_miss_overlap);
  t = s;
  t.converged = 0U;
  CHECK(validate_device_run_status(t) == DeviceRunStatusError::inconsistent_convergence_state);
  t = s;
  t.converged = 0U;
  t.stop_reason = static_cast<std::uint32_t>(DeviceStopReason::device_failure);
  t.error_bits = device_error::device_failure;
  CHECK(validate_device_run_status(t) == DeviceRunStatusError::error_reported_as_bounding_box_miss);
}

void cpu_layouts() {
  SyntheticTileGridConfig cfg;
  cfg.tiles_x = 8;
  cfg.tiles_y = 8;
  cfg.nodes_per_tile = 16;
  cfg.seed = 3;
  const SyntheticTileGraph gen = SyntheticTileGraph::generate(cfg, 1);
  const WeightedGraph g = gen.weighted_graph();
  const hip::ResidentGraphPlan plan = hip::make_resident_graph_plan(gen.layout());
  const auto offs = gen.tile_vertex_offsets();
  CHECK(plan.tile_vertex_offsets.size() == offs.size());
  CHECK(std::equal(offs.begin(), offs.end(), plan.tile_vertex_offsets.begin()));
  CHECK(plan.fingerprint == fingerprint_device_graph_layout32(gen.layout()));
  DeviceGraphLayout32 scrambled = gen.layout();
  std::swap(scrambled.owner_tiles.front(), scrambled.owner_tiles.back());
  CHECK(throws<std::invalid_argument>([&] { (void)hip::make_resident_graph_plan(scrambled); }));

  std::vector<RouteQuery> qs;
  const std::vector<VertexId> s0{VertexId{5}}, t0{VertexId{200}};
  const std::vector<VertexId> s1{VertexId{700}, VertexId{3}}, t1{VertexId{900}};
  qs.push_back(make_route_query(QueryId{10}, g, s0, t0, 1));
  qs.push_back(make_route_query(QueryId{11}, g, s1, t1, 0));
  BatchDeviceDescription d;
  prepare_batch_device_description(g, qs, 4, d);
  CHECK(validate_batch_device_description(g, d).ok());
  CHECK(d.valid_lane_mask == 0x3U);
  CHECK(d.source_offsets == (std::vector<std::uint32_t>{0, 1, 3, 3, 3}));
  CHECK(d.sources == (std::vector<std::uint32_t>{5, 3, 700}));
  std::size_t lane0 = 0;
  for (std::uint32_t t : d.union_tiles) lane0 += d.tile_lane_masks[t] & 1U;
  CHECK(lane0 == qs[0].selected_tiles.size());
  BatchDeviceDescription e = d;
  e.union_tiles.pop_back();
  CHECK(validate_batch_device_description(g, e).code ==
        BatchLayoutValidationErrorCode::invalid_union_tiles);
  e = d;
  std::swap(e.sources[1], e.sources[2]);  // lane 1 sources out of order
  CHECK(validate_batch_device_description(g, e).code ==
        BatchLayoutValidationErrorCode::invalid_terminal_payload);
  e = d;
  e.tile_lane_masks[d.union_tiles[0]] |= 0x8U;  // a bit of an invalid lane
  CHECK(validate_batch_device_description(g, e).code ==
        BatchLayoutValidationErrorCode::invalid_tile_lane_masks);
  CHECK(throws<std::invalid_argument>([&] { prepare_batch_device_description(g, qs, 1, d); }));
}

// ---------------------------------------------------------------------------
struct Writer {
  std::FILE* f;
  void u32(std::uint32_t v) { std::fwrite(&v, 4, 1, f); }
  void f32s(const std::vector<float>& v) { std::fwrite(v.data(), 4, v.size(), f); }
};

int gpu(const char* out_path) {
  SyntheticTileGridConfig cfg;  // 16 x 16 tiles x 64 nodes, seed 1
  const SyntheticTileGraph gen = SyntheticTileGraph::generate(cfg, 0);
  const WeightedGraph g = gen.weighted_graph();
  const TileRunLayout64 runs = build_tile_run_layout(g);
  hip::HipStream stream;
  hip::ResidentDeviceGraph resident;
  resident.upload_once_async(hip::make_resident_graph_plan(gen.layout()), stream);
  resident.wait_until_ready(stream);
  CHECK(throws<std::logic_error>(
      [&] { resident.upload_once_async(hip::make_resident_graph_plan(gen.layout()), stream); }));

  SyntheticQueryConfig qc;
  qc.count = 6;
  qc.seed = 17;
  qc.side_min = 3;
  qc.side_max = 10;
  std::vector<RouteQuery> qs;
  for (const SyntheticQuery& sq : generate_synthetic_queries(cfg, qc)) {
    const std::vector<VertexId> s{VertexId{sq.source}}, t{VertexId{sq.target}};
    qs.push_back(make_route_query(QueryId{sq.query_id}, g, s, t, 2));
  }
  const std::uint32_t W = 8;  // six valid lanes, two padded
  BatchDeviceDescription batch;
  prepare_batch_device_description(g, qs, W, batch);
  const std::uint32_t V = g.vertex_count();

  hip::ReusableBatchedWorkspace ws(stream);
  hip::BatchedJacobiPullEngine jacobi(g, runs, resident, ws, stream);
  hip::BatchedDenseChaoticPushEngine dense(g, runs, resident, ws, stream);
  hip::BatchedFrontierPushEngine frontier(g, runs, resident, ws, stream);

  const hip::BatchedRunOutput ref = jacobi.run_with_distances(batch, options_for(EngineKind::jacobi_pull));
  CHECK(ref.result.status.converged == 1U);
  CHECK(ref.result.status.valid_lane_mask == 0x3FU);
  CHECK(validate_device_run_status(ref.result.status) == DeviceRunStatusError::none);
  CHECK(ref.distances.size() == static_cast<std::size_t>(V) * W);
  for (std::uint32_t l = 0; l < 6; ++l) CHECK(ref.lane_convergence_rounds[l] > 0U);
  CHECK(ref.lane_convergence_rounds[6] == 0U && ref.lane_convergence_rounds[7] == 0U);

  auto same_bits = [](const std::vector<float>& a, const std::vector<float>& b) {
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), 4 * a.size()) == 0;
  };
  const auto d_out = dense.run_with_distances(batch, options_for(EngineKind::dense_chaotic_push));
  const auto f_out = frontier.run_with_distances(batch, options_for(EngineKind::frontier_push));
  CHECK(same_bits(d_out.distances, ref.distances));
  CHECK(same_bits(f_out.distances, ref.distances));
  GpuRunOptions chunked = options_for(EngineKind::jacobi_pull);
  chunked.control_mode = ControlMode::chunked_host_poll;
  chunked.instrumentation = InstrumentationLevel::light;
  const auto j2 = jacobi.run_with_distances(batch, chunked);
  CHECK(same_bits(j2.distances, ref.distances));
  CHECK(j2.result.status.rounds_completed == ref.result.status.rounds_completed);
  CHECK(j2.metrics.tile_visits > 0U);
  const DeviceRunStatus compact = frontier.run_compact_status(batch, options_for(EngineKind::frontier_push));
  CHECK(compact.converged == 1U && compact.valid_lane_mask == 0x3FU);
  CHECK(compact.reached_target_mask == ref.result.status.reached_target_mask);

  // Wrong engine for the class, and a round limit that stops early.
  CHECK(throws<std::invalid_argument>(
      [&] { (void)jacobi.run_status_only(batch, options_for(EngineKind::frontier_push)); }));
  GpuRunOptions one_round = options_for(EngineKind::jacobi_pull);
  one_round.maximum_rounds = 1U;
  const auto early = jacobi.run_status_only(batch, one_round);
  CHECK(early.result.status.converged == 0U);
  CHECK(early.result.status.stop_reason == static_cast<std::uint32_t>(DeviceStopReason::maximum_rounds));
  CHECK(validate_device_run_status(early.result.status) == DeviceRunStatusError::none);

  hip::ReusableCompactPathWorkspace pws;
  const hip::CompactPathBatchOutput paths =
      dense.run_compact_paths(batch, options_for(EngineKind::dense_chaotic_push), pws);
  CHECK(paths.targets.size() == batch.targets.size());

  // The expansion runner boundary: the planner's lane batch through
  // BatchedExpansionExecutor, in each transfer mode, equals the engine run
  // directly on the same planned description.
  {
    std::vector<RouteQuery> by_id = qs;
    std::sort(by_id.begin(), by_id.end(),
              [](const RouteQuery& x, const RouteQuery& y) { return x.query_id < y.query_id; });
    const SelectedRegionIndex regions{g, runs};
    const auto features = make_batch_query_features(g, regions, by_id);
    BatchPlannerPolicy policy;
    policy.lane_width = 8U;
    policy.minimum_jaccard_numerator = 0U;  // one batch of all six
    policy.maximum_union_inflation_numerator = 64U;
    const BatchPlan plan = make_overlapping_batch_plan(regions, features, policy);
    CHECK(plan.batches.size() == 1U && validate_batch_plan(regions, features, plan).ok());
    const BatchPlanEntry& entry = plan.batches[0];
    BatchDeviceDescription planned;
    prepare_batch_device_description(g, by_id, features, entry, planned);
    const ExpansionBatchContext ctx{};
    for (EngineKind kind : {EngineKind::jacobi_pull, EngineKind::dense_chaotic_push,
                            EngineKind::frontier_push}) {
      const GpuRunOptions o = options_for(kind);
      hip::BatchedExpansionExecutor status_exec(g, runs, resident, ws, stream, o);
      CHECK(status_exec.kind() == kind);
      const ExpansionBatchExecution st = status_exec(by_id, features, entry, ctx);
      const DeviceRunStatus direct =
          kind == EngineKind::jacobi_pull ? jacobi.run_compact_status(planned, o)
          : kind == EngineKind::dense_chaotic_push ? dense.run_compact_status(planned, o)
                                                    : frontier.run_compact_status(planned, o);
      CHECK(st.result.status.converged == 1U && st.result.status.valid_lane_mask == 0x3FU);
      CHECK(st.result.status.reached_target_mask == direct.reached_target_mask);
      CHECK(st.work_evidence == ExpansionWorkEvidence::unavailable && st.compact_paths.empty());

      hip::BatchedExpansionExecutor work_exec(g, runs, resident, ws, stream, o,
                                              hip::BatchedExpansionTransferMode::status_and_work_evidence);
      const ExpansionBatchExecution wk = work_exec(by_id, features, entry, ctx);
      CHECK(wk.work_evidence == ExpansionWorkEvidence::measured);
      CHECK(wk.shared_edge_work > 0U && wk.logical_lane_edge_work > 0U);
      CHECK(wk.result.status.reached_target_mask == direct.reached_target_mask);

      hip::BatchedExpansionExecutor path_exec(g, runs, resident, ws, stream, o,
                                              hip::BatchedExpansionTransferMode::compact_paths);
      const ExpansionBatchExecution pe = path_exec(by_id, features, entry, ctx);
      const hip::CompactPathBatchOutput dp =
          kind == EngineKind::jacobi_pull ? jacobi.run_compact_paths(planned, o, pws)
          : kind == EngineKind::dense_chaotic_push ? dense.run_compact_paths(planned, o, pws)
                                                    : frontier.run_compact_paths(planned, o, pws);
      CHECK(pe.compact_paths.size() == 6U);
      CHECK(pe.compact_transfer == dp.transfer);
      CHECK(pe.compact_transfer == measure_compact_transfer(pe.compact_paths));
      CHECK(pe.compact_execution.host_timing == CompactStageTimingEvidence::measured);
      CHECK(pe.compact_total_device_to_host_bytes ==
            pe.compact_transfer.total_bytes + sizeof(DeviceRunStatus) + pe.compact_error_bytes);
      for (const CompactPathPayload& p : pe.compact_paths)
        CHECK(validate_compact_path_payload(g, *std::find_if(by_id.begin(), by_id.end(), [&](const RouteQuery& q) {
                return q.query_id == p.query_id;
              }), p).ok());
      for (std::size_t i = 1; i < pe.compact_paths.size(); ++i)
        CHECK(pe.compact_paths[i - 1].query_id < pe.compact_paths[i].query_id);
      for (std::uint32_t lane = 0; lane < 6; ++lane) {
        const QueryId id = entry.query_ids_by_lane[lane];
        const auto it = std::find_if(pe.compact_paths.begin(), pe.compact_paths.end(),
                                     [&](const CompactPathPayload& p) { return p.query_id == id; });
        CHECK(it != pe.compact_paths.end());
        if (it == pe.compact_paths.end()) continue;
        const auto b0 = dp.targets.begin() + planned.target_offsets[lane];
        const auto b1 = dp.targets.begin() + planned.target_offsets[lane + 1];
        CHECK(std::equal(it->targets.begin(), it->targets.end(), b0, b1));
        CHECK(it->target_terminal_to_target == by_id[entry.query_indices_by_lane[lane]].target_terminal_to_target);
      }
    }
    GpuRunOptions traced = options_for(EngineKind::frontier_push);
    traced.instrumentation = InstrumentationLevel::light;
    CHECK(throws<std::invalid_argument>([&] {
      hip::BatchedExpansionExecutor bad(g, runs, resident, ws, stream, traced,
                                        hip::BatchedExpansionTransferMode::compact_paths);
    }));
  }

  // Single-query engines reproduce each lane.
  hip::ReusableDeviceWorkspace single_ws(stream);
  hip::JacobiPullEngine sj(g, runs, resident, single_ws, stream);
  hip::DenseChaoticPushEngine sd(g, runs, resident, single_ws, stream);
  hip::FrontierPushEngine sf(g, runs, resident, single_ws, stream);
  GpuSsspEngine* singles[] = {&sj, &sd, &sf};
  for (std::uint32_t l = 0; l < qs.size(); ++l) {
    std::vector<float> lane(V);
    for (std::uint32_t v = 0; v < V; ++v) lane[v] = ref.distances[static_cast<std::size_t>(v) * W + l];
    for (GpuSsspEngine* e : singles) {
      auto* base = static_cast<hip::SingleQueryEngineBase*>(e);
      const auto out = base->run_with_distances(qs[l], options_for(e->kind()));
      CHECK(out.result.status.converged == 1U);
      CHECK(same_bits(out.distances, lane));
      const GpuSsspResult r = e->run(qs[l], options_for(e->kind()));
      CHECK(r.status.converged == 1U && r.engine_kind == static_cast<std::uint32_t>(e->kind()));
    }
    const auto sel = sj.run_with_selected_distances(qs[l], options_for(EngineKind::jacobi_pull));
    std::size_t expect = 0;
    for (TileId t : qs[l].selected_tiles)
      expect += g.tile_vertex_offsets()[t.value() + 1] - g.tile_vertex_offsets()[t.value()];
    CHECK(sel.distances.size() == expect);
  }

  // Lanes, labels and paths for the oracle comparison.
  std::FILE* f = std::fopen(out_path, "wb");
  if (!f) {
    std::fprintf(stderr, "cannot write %s\n", out_path);
    return 2;
  }
  Writer w{f};
  w.u32(V);
  w.u32(W);
  w.u32(static_cast<std::uint32_t>(qs.size()));
  for (const RouteQuery& q : qs) {
    w.u32(q.sources[0].value());
    w.u32(q.targets[0].value());
    w.u32(static_cast<std::uint32_t>(q.selected_tiles.size()));
    for (TileId t : q.selected_tiles) w.u32(t.value());
  }
  w.f32s(ref.distances);
  for (const CompactTargetPath& p : paths.targets) {
    w.u32(static_cast<std::uint32_t>(p.summary.reconstruction));
    w.u32(static_cast<std::uint32_t>(p.vertices.size()));
    for (VertexId v : p.vertices) w.u32(v.value());
    for (EdgeId e : p.edge_ids) w.u32(static_cast<std::uint32_t>(e.value()));
    w.f32s(p.distance_labels);
  }
  std::fclose(f);
  return 0;
}



// ---------------------------------------------------------------------------
// Batch planner (reference tests/batch_plan_test.cpp:80-411 states the same
// fixture and expectations).

struct PlannerFixture {
  PartitionedGraph partitioned;
  TileRunLayout64 tile_runs;
  std::vector<RouteQuery> queries;
};

PlannerFixture planner_fixture() {
  const ResourceClassId rc{1U};
  std::vector<VertexMetadata> vertices{
      VertexMetadata::located(0, 0, rc),  VertexMetadata::located(1, 0, rc),
      VertexMetadata::located(10, 0, rc), VertexMetadata::located(30, 0, rc),
      VertexMetadata::unlocated(rc),      VertexMetadata::located(11, 0, rc),
      VertexMetadata::located(2, 0, rc)};
  const std::uint32_t ends[8][2] = {{0, 1}, {0, 2}, {0, 3}, {0, 3}, {0, 4}, {4, 0}, {2, 0}, {1, 2}};
  std::vector<EdgeInputRecord> edges;
  for (std::uint32_t i = 0; i < 8; ++i) {
    EdgeInputRecord e{};
    e.source = VertexId{ends[i][0]};
    e.destination = VertexId{ends[i][1]};
    e.weight = static_cast<float>(i + 1);
    e.provenance = PhysicalProvenance{i};
    edges.push_back(e);
  }
  const UniformGridPartitioner grid{SpatialOrderConfig{0, 0, 10U, 10U}};
  PlannerFixture f{grid.partition(InputGraph{std::move(vertices), std::move(edges)}), {}, {}};
  const WeightedGraph& g = f.partitioned.graph;
  const auto to_new = g.old_to_new();
  const auto tile_of = [&](std::uint32_t old) { return g.owner_tiles()[to_new[old].value()]; };
  const auto query = [&](std::uint32_t id, std::vector<std::uint32_t> src, std::vector<std::uint32_t> tgt,
                         std::uint32_t gen, std::vector<TileId> tiles) {
    std::vector<VertexId> s, t;
    for (auto v : src) s.push_back(to_new[v]);
    for (auto v : tgt) t.push_back(to_new[v]);
    RouteQuery q = make_route_query(QueryId{id}, g, s, t, 0U, gen);
    std::sort(tiles.begin(), tiles.end());
    q.selected_tiles = tiles;
    if (!validate_route_query(g, q).ok()) throw std::runtime_error{"invalid planner fixture query"};
    f.queries.push_back(std::move(q));
  };
  query(10, {0, 2}, {5}, 0, {tile_of(0), tile_of(2)});
  query(20, {0}, {4}, 1, {tile_of(0), tile_of(4)});
  query(30, {4}, {4}, 2, {tile_of(4)});
  query(40, {2}, {3}, 3, {tile_of(2), tile_of(3)});
  f.tile_runs = build_tile_run_layout(g);
  return f;
}

BatchPlannerPolicy planner_policy(std::uint32_t width) {
  BatchPlannerPolicy p;
  p.lane_width = width;
  p.maximum_union_inflation_numerator = 3U;
  p.maximum_union_inflation_denominator = 1U;
  return p;
}

ResourceClassId rc_of_test() { return ResourceClassId{1U}; }

// The greedy selection rule written out directly: every open batch scores
// every remaining query from scratch (Jaccard, shared terminal tiles,
// inflation, union edges / vertices / new tiles, gaps to the anchor, id) and
// admits the best until the width or the thresholds stop it.  The library
// computes the same plan incrementally; the two must agree exactly.
BatchPlan exhaustive_plan(const SelectedRegionIndex& index, const std::vector<BatchQueryFeatures>& qs,
                          const BatchPlannerPolicy& policy) {
  const auto frac_less = [](BatchFraction a, BatchFraction b) {
    return static_cast<unsigned __int128>(a.numerator) * b.denominator <
           static_cast<unsigned __int128>(b.numerator) * a.denominator;
  };
  const auto unite = [](const std::vector<TileId>& a, const std::vector<TileId>& b) {
    std::vector<TileId> out;
    std::set_union(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(out));
    return out;
  };
  const auto meet = [](const std::vector<TileId>& a, const std::vector<TileId>& b) {
    std::vector<TileId> out;
    std::set_intersection(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(out));
    return out.size();
  };
  const auto dist = [](std::uint32_t a, std::uint32_t b) { return a > b ? a - b : b - a; };
  BatchPlan plan;
  plan.policy = policy;
  plan.input_query_count = static_cast<std::uint32_t>(qs.size());
  std::vector<std::uint32_t> order(qs.size());
  for (std::uint32_t i = 0; i < order.size(); ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](std::uint32_t a, std::uint32_t b) {
    return std::make_tuple(qs[b].selected_edge_estimate, qs[b].selected_vertex_count, qs[b].selected_tiles.size(),
                           qs[a].expansion_generation, qs[a].query_id) <
           std::make_tuple(qs[a].selected_edge_estimate, qs[a].selected_vertex_count, qs[a].selected_tiles.size(),
                           qs[b].expansion_generation, qs[b].query_id);
  });
  const BatchFraction jmin{policy.minimum_jaccard_numerator, policy.minimum_jaccard_denominator};
  const BatchFraction imax{policy.maximum_union_inflation_numerator, policy.maximum_union_inflation_denominator};
  std::vector<bool> taken(qs.size(), false);
  for (std::uint32_t a : order) {
    if (taken[a]) continue;
    taken[a] = true;
    std::vector<std::uint32_t> lanes{a};
    std::vector<TileId> tiles = qs[a].selected_tiles, src = qs[a].source_tiles, tgt = qs[a].target_tiles;
    std::uint64_t lane_tiles = tiles.size();
    while (lanes.size() < policy.lane_width) {
      using Key = std::tuple<std::uint64_t, std::uint64_t, std::uint64_t, std::uint32_t, std::uint32_t,
                             std::uint32_t, std::uint32_t>;
      bool found = false;
      std::uint32_t best = 0;
      BatchFraction bj{}, bi{};
      std::uint64_t bs = 0;
      Key bk{};
      for (std::uint32_t q = 0; q < qs.size(); ++q) {
        if (taken[q]) continue;
        const auto u = unite(tiles, qs[q].selected_tiles);
        const BatchFraction j{meet(tiles, qs[q].selected_tiles), u.size()};
        const BatchFraction inf{(lanes.size() + 1U) * u.size(), lane_tiles + qs[q].selected_tiles.size()};
        if (frac_less(j, jmin) || frac_less(imax, inf)) continue;
        const std::uint64_t shared = meet(src, qs[q].source_tiles) + meet(tgt, qs[q].target_tiles);
        const Key k{index.selected_edge_count(u), index.selected_vertex_count(u), u.size() - tiles.size(),
                    dist(qs[q].expansion_generation, qs[a].expansion_generation),
                    dist(qs[q].source_count, qs[a].source_count), dist(qs[q].target_count, qs[a].target_count),
                    qs[q].query_id.value()};
        bool better = !found;
        if (!better) {
          if (frac_less(bj, j) || frac_less(j, bj)) better = frac_less(bj, j);
          else if (shared != bs) better = shared > bs;
          else if (frac_less(inf, bi) || frac_less(bi, inf)) better = frac_less(inf, bi);
          else better = k < bk;
        }
        if (better) found = true, best = q, bj = j, bi = inf, bs = shared, bk = k;
      }
      if (!found) break;
      taken[best] = true;
      lanes.push_back(best);
      lane_tiles += qs[best].selected_tiles.size();
      tiles = unite(tiles, qs[best].selected_tiles);
      src = unite(src, qs[best].source_tiles);
      tgt = unite(tgt, qs[best].target_tiles);
    }
    BatchPlanEntry e;
    e.lane_width = policy.lane_width;
    e.valid_lane_mask = lanes.size() >= 32 ? ~LaneMask{0} : (LaneMask{1} << lanes.size()) - 1U;
    e.query_indices_by_lane.assign(policy.lane_width, invalid_batch_query_index);
    e.query_ids_by_lane.assign(policy.lane_width, invalid_batch_query_id);
    e.expansion_generations_by_lane.assign(policy.lane_width, 0U);
    e.union_tiles = tiles;
    e.union_tile_lane_masks.assign(tiles.size(), 0U);
    for (std::uint32_t l = 0; l < lanes.size(); ++l) {
      const BatchQueryFeatures& f = qs[lanes[l]];
      e.query_indices_by_lane[l] = lanes[l];
      e.query_ids_by_lane[l] = f.query_id;
      e.expansion_generations_by_lane[l] = f.expansion_generation;
      e.selected_lane_vertex_count += f.selected_vertex_count;
      e.selected_lane_edge_estimate += f.selected_edge_estimate;
      for (std::size_t t = 0; t < tiles.size(); ++t)
        if (std::binary_search(f.selected_tiles.begin(), f.selected_tiles.end(), tiles[t]))
          e.union_tile_lane_masks[t] |= LaneMask{1} << l;
    }
    e.union_vertex_count = index.selected_vertex_count(tiles);
    e.union_edge_estimate = index.selected_edge_count(tiles);
    plan.batches.push_back(std::move(e));
  }
  return plan;
}

void planner(const char* golden_path) {
  const PlannerFixture fx = planner_fixture();
  const WeightedGraph& g = fx.partitioned.graph;
  const SelectedRegionIndex index{g, fx.tile_runs};
  const auto features = make_batch_query_features(g, index, fx.queries);
  CHECK(features.size() == 4U);
  CHECK(features[0].source_count == 2U && features[1].source_count == 1U);
  CHECK((features[0].source_tiles == std::vector<TileId>{TileId{0U}, TileId{1U}}));
  CHECK((features[0].target_tiles == std::vector<TileId>{TileId{1U}}));
  CHECK((features[1].target_tiles == std::vector<TileId>{TileId{3U}}));
  const std::uint64_t want_v[4] = {5, 4, 1, 3}, want_e[4] = {4, 3, 0, 0};
  for (int i = 0; i < 4; ++i) {
    CHECK(features[i].selected_vertex_count == want_v[i]);
    CHECK(features[i].selected_edge_estimate == want_e[i]);
    // brute force over the graph: vertices owned by, and edges with both ends
    // in, the selected tiles
    const auto& sel = features[i].selected_tiles;
    const auto in_sel = [&](TileId t) { return std::binary_search(sel.begin(), sel.end(), t); };
    std::uint64_t nv = 0, ne = 0;
    const auto out = g.outgoing();
    for (std::uint32_t v = 0; v < g.vertex_count(); ++v) {
      if (!in_sel(g.owner_tiles()[v])) continue;
      ++nv;
      for (auto e = out.row_offsets[v]; e < out.row_offsets[v + 1]; ++e)
        ne += in_sel(g.owner_tiles()[out.destinations[e].value()]) ? 1U : 0U;
    }
    CHECK(nv == features[i].selected_vertex_count && ne == features[i].selected_edge_estimate);
  }
  // An index of another (identical) graph object is refused.
  const PlannerFixture other = planner_fixture();
  const SelectedRegionIndex other_index{other.partitioned.graph, other.tile_runs};
  CHECK(throws<std::invalid_argument>([&] { (void)make_batch_query_features(g, other_index, fx.queries); }));

  // Policy rules and the exact Jaccard.
  for (std::uint32_t w : {1U, 8U, 16U, 32U})
    CHECK(validate_batch_planner_policy(planner_policy(w)) == BatchPlannerPolicyError::none);
  for (std::uint32_t w : {0U, 7U, 64U})
    CHECK(validate_batch_planner_policy(planner_policy(w)) == BatchPlannerPolicyError::unsupported_lane_width);
  BatchPlannerPolicy bad;
  bad.minimum_jaccard_denominator = 0U;
  CHECK(validate_batch_planner_policy(bad) == BatchPlannerPolicyError::invalid_jaccard_fraction);
  bad = BatchPlannerPolicy{};
  bad.maximum_union_inflation_denominator = 0U;
  CHECK(validate_batch_planner_policy(bad) == BatchPlannerPolicyError::invalid_union_inflation_fraction);
  const TileId l2[2] = {TileId{0U}, TileId{1U}}, r2[2] = {TileId{1U}, TileId{2U}};
  const BatchFraction j = tile_set_jaccard(l2, r2);
  CHECK(j.numerator * 3U == j.denominator);

  // Expanded selection of the same terminals: the larger region anchors.
  RouteQuery expanded = fx.queries.front();
  expanded.query_id = QueryId{11U};
  expanded.expansion_generation = 4U;
  expanded.selected_tiles.push_back(TileId{2U});
  const RouteQuery pair[2] = {fx.queries.front(), expanded};
  const auto pair_features = make_batch_query_features(g, index, pair);
  const BatchPlan grown = make_overlapping_batch_plan(index, pair_features, planner_policy(8U));
  CHECK(grown.batches.size() == 1U);
  CHECK(grown.batches[0].query_ids_by_lane[0] == QueryId{11U});
  CHECK(grown.batches[0].query_ids_by_lane[1] == QueryId{10U});
  CHECK((grown.batches[0].union_tile_lane_masks == std::vector<LaneMask>{3U, 3U, 1U}));

  // The golden width-8 plan.
  const BatchPlan plan = make_overlapping_batch_plan(index, features, planner_policy(8U));
  CHECK(validate_batch_plan(index, features, plan).ok());
  CHECK(plan.batches.size() == 1U);
  const BatchPlanEntry& b = plan.batches[0];
  CHECK(b.lane_width == 8U && b.valid_lane_mask == 0x0FU);
  CHECK((b.expansion_generations_by_lane == std::vector<std::uint32_t>{0, 1, 2, 3, 0, 0, 0, 0}));
  CHECK((b.union_tile_lane_masks == std::vector<LaneMask>{3U, 9U, 8U, 6U}));
  CHECK(b.union_vertex_count == 7U && b.union_edge_estimate == 8U);
  CHECK(b.selected_lane_vertex_count == 13U && b.selected_lane_edge_estimate == 7U);
  const std::uint32_t all_lanes[4] = {0, 1, 2, 3};
  const BatchFraction infl = batch_union_tile_inflation(features, all_lanes, b.union_tiles);
  CHECK(infl.numerator * 7U == infl.denominator * 16U);
  std::ifstream in(golden_path, std::ios::binary);
  std::stringstream golden;
  golden << in.rdbuf();
  CHECK(in.good() || in.eof());
  const std::string tsv = serialize_batch_plan_tsv(plan);
  CHECK(!golden.str().empty() && tsv == golden.str());
  CHECK(fingerprint_batch_plan(plan) == 16700197086403570866ULL);
  BatchPlan changed = plan;
  ++changed.batches[0].union_edge_estimate;
  CHECK(fingerprint_batch_plan(changed) != 16700197086403570866ULL);

  // The plan entry lays out as a lane-batched description in lane order.
  std::vector<RouteQuery> by_id = fx.queries;
  std::sort(by_id.begin(), by_id.end(),
            [](const RouteQuery& x, const RouteQuery& y) { return x.query_id < y.query_id; });
  BatchDeviceDescription desc;
  prepare_batch_device_description(g, by_id, features, b, desc);
  CHECK(desc.lane_width == 8U && desc.valid_lane_mask == 0x0FU);
  CHECK((desc.query_ids_by_lane == std::vector<std::uint32_t>{10, 20, 30, 40, 0, 0, 0, 0}));
  CHECK((desc.union_tiles == std::vector<std::uint32_t>{0, 1, 2, 3}));
  for (std::uint32_t t = 0; t < 4; ++t) CHECK(desc.tile_lane_masks[t] == b.union_tile_lane_masks[t]);
  CHECK(validate_batch_device_description(g, desc).ok());
  BatchPlanEntry wrong = b;
  wrong.query_ids_by_lane[2] = QueryId{40U};
  CHECK(throws<std::invalid_argument>([&] { prepare_batch_device_description(g, by_id, features, wrong, desc); }));

  // Damaged plans are refused with the matching code.
  using Code = BatchPlanValidationErrorCode;
  BatchPlan damaged = plan;
  damaged.batches[0].union_tile_lane_masks[0] ^= 1U;
  CHECK(validate_batch_plan(index, features, damaged).code == Code::tile_lane_mask_mismatch);
  damaged = plan;
  damaged.batches[0].query_indices_by_lane[5] = 0U;
  CHECK(validate_batch_plan(index, features, damaged).code == Code::invalid_padding);
  damaged = plan;
  damaged.batches[0].query_ids_by_lane[1] = QueryId{99U};
  CHECK(validate_batch_plan(index, features, damaged).code == Code::query_identity_mismatch);
  damaged = plan;
  damaged.batches[0].query_indices_by_lane[1] = 0U;
  damaged.batches[0].query_ids_by_lane[1] = QueryId{10U};
  damaged.batches[0].expansion_generations_by_lane[1] = 0U;
  CHECK(validate_batch_plan(index, features, damaged).code == Code::duplicate_assignment);
  damaged = plan;
  ++damaged.batches[0].union_vertex_count;
  CHECK(validate_batch_plan(index, features, damaged).code == Code::estimate_mismatch);
  damaged = plan;
  damaged.batches[0].valid_lane_mask = 0x0BU;
  CHECK(validate_batch_plan(index, features, damaged).code == Code::invalid_valid_lane_mask);

  // Narrow widths split the corpus; every query lands exactly once and the
  // family is the same policy at widths 32, 16 and 8.
  const BatchPlan singles = make_overlapping_batch_plan(index, features, planner_policy(1U));
  CHECK(singles.batches.size() == 4U && validate_batch_plan(index, features, singles).ok());
  const BatchPlanFamily family = make_standard_batch_plan_family(index, features, planner_policy(8U));
  CHECK(family.plans[0].policy.lane_width == 32U && family.plans[2] == plan);
  for (const BatchPlan& p : family.plans) CHECK(validate_batch_plan(index, features, p).ok());

  // A strict Jaccard floor keeps disjoint regions apart.
  BatchPlannerPolicy strict = planner_policy(8U);
  strict.minimum_jaccard_numerator = 1U;
  strict.minimum_jaccard_denominator = 1U;
  const BatchPlan apart = make_overlapping_batch_plan(index, features, strict);
  CHECK(apart.batches.size() == 4U && validate_batch_plan(index, features, apart).ok());

  // A random corpus on a 64x64 grid split into 8x8 tiles: every plan
  // validates, each admitted lane met the Jaccard floor and the inflation
  // cap against the union before it, and every query appears exactly once.
  std::vector<VertexMetadata> rv;
  std::vector<EdgeInputRecord> re;
  std::uint64_t state = 0x9E3779B97F4A7C15ULL;
  const auto next = [&] {
    state ^= state << 13, state ^= state >> 7, state ^= state << 17;
    return state;
  };
  for (int i = 0; i < 600; ++i)
    rv.push_back(VertexMetadata::located(static_cast<std::int32_t>(next() % 64),
                                         static_cast<std::int32_t>(next() % 64), rc_of_test()));
  for (std::uint32_t i = 0; i < 3000; ++i) {
    EdgeInputRecord e{};
    e.source = VertexId{static_cast<std::uint32_t>(next() % 600)};
    e.destination = VertexId{static_cast<std::uint32_t>(next() % 600)};
    e.weight = 1.0F + static_cast<float>(next() % 16);
    e.provenance = PhysicalProvenance{i};
    re.push_back(e);
  }
  const UniformGridPartitioner grid8{SpatialOrderConfig{0, 0, 8U, 8U}};
  const PartitionedGraph rg = grid8.partition(InputGraph{std::move(rv), std::move(re)});
  const TileRunLayout64 rruns = build_tile_run_layout(rg.graph);
  const SelectedRegionIndex rindex{rg.graph, rruns};
  std::vector<RouteQuery> rq;
  for (std::uint32_t i = 0; i < 70; ++i) {
    const VertexId s1[1] = {VertexId{static_cast<std::uint32_t>(next() % 600)}};
    const VertexId t1[1] = {VertexId{static_cast<std::uint32_t>(next() % 600)}};
    rq.push_back(make_route_query(QueryId{1000U + 7U * i}, rg.graph, s1, t1,
                                  static_cast<std::uint32_t>(next() % 3), static_cast<std::uint32_t>(i % 5)));
  }
  const auto rf = make_batch_query_features(rg.graph, rindex, rq);
  for (std::uint32_t w : {8U, 16U, 32U}) {
    const BatchPlannerPolicy pol = planner_policy(w);
    const BatchPlan rp = make_overlapping_batch_plan(rindex, rf, pol);
    CHECK(validate_batch_plan(rindex, rf, rp).ok());
    std::vector<int> seen(rf.size(), 0);
    for (const BatchPlanEntry& e : rp.batches) {
      std::vector<TileId> uni;
      std::uint64_t lane_tiles = 0;
      for (std::uint32_t lane = 0; lane < w; ++lane) {
        if (!(e.valid_lane_mask >> lane & 1U)) break;
        const auto& f = rf[e.query_indices_by_lane[lane]];
        ++seen[e.query_indices_by_lane[lane]];
        std::vector<TileId> next_uni;
        std::set_union(uni.begin(), uni.end(), f.selected_tiles.begin(), f.selected_tiles.end(),
                       std::back_inserter(next_uni));
        if (lane > 0) {
          std::vector<TileId> both;
          std::set_intersection(uni.begin(), uni.end(), f.selected_tiles.begin(),
                                f.selected_tiles.end(), std::back_inserter(both));
          // |both| / |next_uni| >= 1/8 and (lane+1)|next_uni| / tiles <= 3
          CHECK(8U * both.size() >= next_uni.size());
          CHECK((lane + 1U) * next_uni.size() <= 3U * (lane_tiles + f.selected_tiles.size()));
        }
        lane_tiles += f.selected_tiles.size();
        uni = std::move(next_uni);
      }
      CHECK(uni == e.union_tiles);
    }
    CHECK(std::all_of(seen.begin(), seen.end(), [](int n) { return n == 1; }));
    CHECK(rp == exhaustive_plan(rindex, rf, pol));
  }
  CHECK(make_overlapping_batch_plan(index, features, planner_policy(8U)) ==
        exhaustive_plan(index, features, planner_policy(8U)));

  // Overlapping boxes on a 24x24x16 synthetic grid (256 queries, generations
  // 0-2, boxes grown by 0-3 rings): the incremental planner equals the
  // written-out rule at every width and under loose and strict thresholds.
  SyntheticTileGridConfig sc;
  sc.tiles_x = 24;
  sc.tiles_y = 24;
  sc.nodes_per_tile = 16;
  sc.seed = 9;
  const SyntheticTileGraph sg = SyntheticTileGraph::generate(sc, 1);
  const WeightedGraph sgraph = sg.weighted_graph();
  const TileRunLayout64 sruns = build_tile_run_layout(sgraph);
  const SelectedRegionIndex sindex{sgraph, sruns};
  SyntheticQueryConfig sqc;
  sqc.count = 256;
  sqc.seed = 13;
  sqc.side_min = 2;
  sqc.side_max = 12;
  std::vector<RouteQuery> sq;
  std::uint32_t k = 0;
  for (const SyntheticQuery& q : generate_synthetic_queries(sc, sqc)) {
    const VertexId s1[1] = {VertexId{q.source}}, t1[1] = {VertexId{q.target}};
    sq.push_back(make_route_query(QueryId{q.query_id}, sgraph, s1, t1, k % 4U, k % 3U));
    ++k;
  }
  const auto sf = make_batch_query_features(sgraph, sindex, sq);
  for (std::uint32_t w : {8U, 16U, 32U}) {
    BatchPlannerPolicy loose;
    loose.lane_width = w;
    loose.minimum_jaccard_numerator = 0U;
    loose.maximum_union_inflation_numerator = 16U;
    BatchPlannerPolicy strict = loose;
    strict.minimum_jaccard_numerator = 1U;
    strict.minimum_jaccard_denominator = 4U;
    strict.maximum_union_inflation_numerator = 2U;
    for (const BatchPlannerPolicy& pol : {loose, strict, planner_policy(w)}) {
      const BatchPlan fast = make_overlapping_batch_plan(sindex, sf, pol);
      CHECK(fast == exhaustive_plan(sindex, sf, pol));
      CHECK(validate_batch_plan(sindex, sf, fast).ok());
    }
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: api_test cpu | planner GOLDEN.tsv | gpu OUT.bin\n");
    return 2;
  }
  int rc = 0;
  try {
    if (std::strcmp(argv[1], "cpu") == 0) {
      cpu_rules();
      cpu_layouts();
    } else if (std::strcmp(argv[1], "planner") == 0 && argc >= 3) {
      planner(argv[2]);
    } else if (std::strcmp(argv[1], "gpu") == 0 && argc >= 3) {
      rc = gpu(argv[2]);
    } else {
      std::fprintf(stderr, "unknown mode\n");
      return 2;
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "exception: %s\n", e.what());
    return 3;
  }
  std::printf("%s: %d failure(s)\n", argv[1], failures);
  return failures ? 1 : rc;
}
