// expansion_test.cpp -- run_batched_expansion, its options and schedule
// evidence, and the compact payload rules.
//
//   expansion_test host   the orchestrator over a host batch runner written
//                         here (restricted fp32 Dijkstra per lane), on the
//                         seven-route fixture of the reference's
//                         tests/batched_expansion_fixture_suite.hpp:17-131;
//                         no device call
//   expansion_test gpu    the same runs through hip::BatchedExpansionExecutor
//                         (every engine, schedule and transfer mode), checked
//                         against the host runner's outcomes and payloads
//
// The expected retry counts per schedule and route are the literals of the
// reference's tests/batched_expansion_test.cpp:111-147, :352-427 and
// :1177-1240; the rest states the orchestrator's own contract.
#include "bfnew/batched_expansion.hpp"
#include "bfnew/compact_paths.hpp"
#include "bfnew/device_layout.hpp"
#include "bfnew/graph.hpp"
#include "bfnew/hip/batched_expansion.hpp"
#include "bfnew/hip/engines.hpp"
#include "bfnew/hip/runtime.hpp"
#include "bfnew/no_congestion_pipeline.hpp"
#include "bfnew/workload_artifact.hpp"
#include "bfnew/query.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <functional>
#include <limits>
#include <map>
#include <queue>
#include <stdexcept>
#include <string>
#include <vector>

#include <unistd.h>

using namespace bfnew;

namespace {

int failures = 0;

void check(bool ok, const std::string& what, int line) {
  if (!ok) {
    std::fprintf(stderr, "FAIL line %d: %s\n", line, what.c_str());
    ++failures;
  }
}
#define CHECK(x) check((x), #x, __LINE__)
#define CHECK_MSG(x, msg) check((x), std::string(#x) + " [" + (msg) + "]", __LINE__)

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

bool same_bits(float a, float b) { return std::memcmp(&a, &b, 4) == 0; }

// ---------------------------------------------------------------------------
// The seven-route fixture: every route owns its own x range on a 10 x 10 grid.
constexpr std::uint32_t kImmediate = 1700, kTwoRing = 1701, kLongEdge = 1702, kSpill = 1703,
                        kUnreachable = 1704, kSecondMiss = 1705, kMultiSource = 1706;

struct Fixture {
  PartitionedGraph partitioned;
  TileDirectory directory;
  TileRunLayout64 tile_runs;
  std::vector<RouteQuery> queries;
  const WeightedGraph& graph() const { return partitioned.graph; }
  const RouteQuery& query(std::uint32_t id) const {
    for (const RouteQuery& q : queries)
      if (q.query_id.value() == id) return q;
    throw std::logic_error{"no such fixture query"};
  }
  std::vector<RouteQuery> subset(std::initializer_list<std::uint32_t> ids) const {
    std::vector<RouteQuery> out;
    for (std::uint32_t id : ids) out.push_back(query(id));
    return out;
  }
};

EdgeInputRecord edge(std::uint32_t s, std::uint32_t d, float w, std::uint32_t record) {
  EdgeInputRecord e{};
  e.source = VertexId{s};
  e.destination = VertexId{d};
  e.weight = w;
  e.provenance = PhysicalProvenance{0U, 0U, record};
  return e;
}

InputGraph fixture_input();

Fixture make_fixture() {
  const UniformGridPartitioner grid{SpatialOrderConfig{0, 0, 10U, 10U}};
  Fixture f{grid.partition(fixture_input()), {}, {}, {}};
  const WeightedGraph& g = f.graph();
  const auto to_new = g.old_to_new();
  const auto route = [&](std::uint32_t id, std::vector<std::uint32_t> src, std::uint32_t tgt) {
    std::vector<VertexId> s;
    for (std::uint32_t x : src) s.push_back(to_new[x]);
    const std::vector<VertexId> t{to_new[tgt]};
    f.queries.push_back(make_route_query(QueryId{id}, g, s, t));
  };
  route(kImmediate, {0}, 2);
  route(kTwoRing, {3}, 4);
  route(kLongEdge, {7}, 8);
  route(kSpill, {10}, 12);
  route(kUnreachable, {13}, 14);
  route(kSecondMiss, {15}, 16);
  route(kMultiSource, {3, 24}, 4);
  f.directory = build_tile_directory(g);
  f.tile_runs = build_tile_run_layout(g);
  return f;
}

InputGraph fixture_input() {
  const ResourceClassId rc{1U};
  const auto at = [rc](std::int32_t x, std::int32_t y) { return VertexMetadata::located(x, y, rc); };
  std::vector<VertexMetadata> v{
      at(40, 0),   at(50, 0),   at(60, 0),                 // 0-2 immediate chain
      at(0, 0),    at(20, 0),   at(0, 10),   at(10, 20),   // 3-6 two-ring northern detour
      at(80, 0),   at(100, 0),  at(80, 40),                // 7-9 jump to a fourth-ring vertex
      at(120, 0),  at(120, 20), VertexMetadata::unlocated(rc),  // 10-12 spill target
      at(160, 0),  at(180, 0),                             // 13-14 unreachable
      at(220, 0),  at(240, 0),  at(220, -10), at(230, -30),  // 15-18 southern detour
      at(80, 10),  at(80, 20),  at(80, 30),  at(120, 10), at(220, -20),  // 19-23 ring fillers
      at(1, 0)};                                           // 24 second northern source
  std::vector<EdgeInputRecord> e{
      edge(0, 1, 1.0F, 17000),  edge(1, 2, 1.0F, 17001),  edge(3, 5, 1.0F, 17002),
      edge(5, 6, 1.0F, 17003),  edge(6, 4, 1.0F, 17004),  edge(7, 9, 2.0F, 17005),
      edge(9, 8, 3.0F, 17006),  edge(10, 11, 4.0F, 17007), edge(11, 12, 5.0F, 17008),
      edge(15, 17, 1.0F, 17009), edge(17, 18, 1.0F, 17010), edge(18, 16, 1.0F, 17011),
      edge(24, 5, 0.5F, 17012)};
  return InputGraph{std::move(v), std::move(e)};
}

BatchedExpansionOptions options_of(EngineKind engine, ExpansionSchedulePolicy schedule,
                                   std::uint32_t maximum_expansions = 4U,
                                   ExpansionTerminalPolicy terminal = ExpansionTerminalPolicy::full_region_fallback) {
  BatchedExpansionOptions o;
  o.run_options.engine = engine;
  o.run_options.control_mode = ControlMode::per_round_host_poll;
  o.run_options.rounds_per_chunk = 1U;
  o.run_options.block_size = 256U;
  o.run_options.maximum_rounds = 64U;
  o.run_options.enable_per_lane_convergence = 1U;
  o.planner_policy.lane_width = 8U;
  o.planner_policy.minimum_jaccard_numerator = 0U;
  o.planner_policy.minimum_jaccard_denominator = 1U;
  o.planner_policy.maximum_union_inflation_numerator = 32U;
  o.planner_policy.maximum_union_inflation_denominator = 1U;
  o.execution_configuration_fingerprint = 0x1700000000000001ULL;
  o.schedule = schedule;
  o.maximum_expansions = maximum_expansions;
  o.terminal_policy = terminal;
  return o;
}

const std::array<ExpansionSchedulePolicy, 4> kSchedules{one_ring_expansion(), fixed_ring_expansion(2U),
                                                        doubling_margin_expansion(), hybrid_margin_expansion(2U)};

// Expansions each route needs (the reference test's table).
std::uint32_t expected_expansions(ExpansionScheduleKind kind, std::uint32_t id) {
  if (id == kImmediate) return 0U;
  if (id == kUnreachable) return 1U;  // a stalled first ring, then the full-region restart
  const bool short_route = id == kTwoRing || id == kSpill;
  switch (kind) {
    case ExpansionScheduleKind::one_geometric_ring: return short_route ? 2U : id == kLongEdge ? 4U : 3U;
    case ExpansionScheduleKind::fixed_larger_ring: return short_route ? 1U : 2U;
    default: return short_route ? 2U : 3U;
  }
}

// ---------------------------------------------------------------------------
// Host batch runner: fp32 Dijkstra over each lane's selected tiles.
struct LaneSolution {
  std::vector<float> dist;
  std::vector<std::uint64_t> via;  // CSR position of the improving edge, or ~0
  std::uint64_t region_edges{};
};

LaneSolution solve_lane(const WeightedGraph& g, const RouteQuery& q) {
  const std::uint32_t V = g.vertex_count();
  const auto owner = g.owner_tiles();
  const auto out = g.outgoing();
  std::vector<std::uint8_t> in(V, 0);
  for (std::uint32_t v = 0; v < V; ++v)
    in[v] = std::binary_search(q.selected_tiles.begin(), q.selected_tiles.end(), owner[v]) ? 1U : 0U;
  LaneSolution s;
  s.dist.assign(V, std::numeric_limits<float>::infinity());
  s.via.assign(V, ~std::uint64_t{0});
  using Item = std::pair<float, std::uint32_t>;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
  for (VertexId src : q.sources) {
    s.dist[src.value()] = 0.0F;
    heap.push({0.0F, src.value()});
  }
  std::vector<std::uint8_t> done(V, 0);
  while (!heap.empty()) {
    const auto [d, u] = heap.top();
    heap.pop();
    if (done[u]) continue;
    done[u] = 1U;
    for (std::uint64_t p = out.row_offsets[u]; p < out.row_offsets[u + 1]; ++p) {
      const std::uint32_t w = out.destinations[p].value();
      if (!in[w]) continue;
      const float nd = d + out.weights[p];
      if (nd < s.dist[w]) {
        s.dist[w] = nd;
        s.via[w] = p;
        heap.push({nd, w});
      }
    }
  }
  for (std::uint32_t u = 0; u < V; ++u)
    if (in[u])
      for (std::uint64_t p = out.row_offsets[u]; p < out.row_offsets[u + 1]; ++p)
        s.region_edges += in[out.destinations[p].value()];
  return s;
}

CompactPathPayload payload_of(const WeightedGraph& g, const RouteQuery& q, const LaneSolution& s) {
  const auto out = g.outgoing();
  // CSR position -> source vertex.
  const auto source_of = [&](std::uint64_t p) {
    return static_cast<std::uint32_t>(
        std::upper_bound(out.row_offsets.begin(), out.row_offsets.end(), p) - out.row_offsets.begin() - 1);
  };
  CompactPathPayload pl;
  pl.query_id = q.query_id;
  pl.expansion_generation = q.expansion_generation;
  pl.target_terminal_to_target = q.target_terminal_to_target;
  for (VertexId t : q.targets) {
    CompactTargetPath path;
    path.summary.target = t;
    path.summary.distance = s.dist[t.value()];
    if (!std::isfinite(path.summary.distance)) {
      path.summary.reached = CompactTargetReachStatus::not_reached;
      path.summary.reconstruction = CompactPathStatus::unreachable;
    } else {
      std::vector<std::uint32_t> rev{t.value()};
      std::vector<EdgeId> edges;
      for (std::uint32_t v = t.value(); s.via[v] != ~std::uint64_t{0};) {
        edges.push_back(out.edge_ids[s.via[v]]);
        v = source_of(s.via[v]);
        rev.push_back(v);
      }
      std::reverse(rev.begin(), rev.end());
      std::reverse(edges.begin(), edges.end());
      for (std::uint32_t v : rev) {
        path.vertices.push_back(VertexId{v});
        path.distance_labels.push_back(s.dist[v]);
      }
      path.edge_ids = edges;
      path.summary.selected_source = path.vertices.front();
      path.summary.has_selected_source = 1U;
      path.summary.path_length = static_cast<std::uint32_t>(edges.size());
      path.summary.reached = CompactTargetReachStatus::reached;
      path.summary.reconstruction = CompactPathStatus::complete;
    }
    pl.targets.push_back(std::move(path));
  }
  return pl;
}

struct HostRunner {
  const WeightedGraph& g;
  GpuRunOptions run;
  bool compact{};
  std::uint32_t calls{};

  ExpansionBatchExecution operator()(std::span<const RouteQuery> queries, std::span<const BatchQueryFeatures>,
                                     const BatchPlanEntry& batch, const ExpansionBatchContext&) {
    ++calls;
    ExpansionBatchExecution e;
    e.result.engine_kind = static_cast<std::uint32_t>(run.engine);
    e.result.control_mode = static_cast<std::uint32_t>(run.control_mode);
    DeviceRunStatus& st = e.result.status;
    st.valid_lane_mask = batch.valid_lane_mask;
    st.converged_lane_mask = batch.valid_lane_mask;
    st.converged = 1U;
    st.rounds_completed = 1U;
    st.stop_reason = static_cast<std::uint32_t>(DeviceStopReason::converged);
    const std::uint32_t V = g.vertex_count(), W = batch.lane_width;
    if (!compact) e.final_distances.assign(static_cast<std::size_t>(V) * W, std::numeric_limits<float>::infinity());
    e.work_evidence = ExpansionWorkEvidence::measured;
    for (std::uint32_t lane = 0; lane < W; ++lane) {
      if (!(batch.valid_lane_mask >> lane & 1U)) continue;
      const RouteQuery& q = queries[batch.query_indices_by_lane[lane]];
      const LaneSolution s = solve_lane(g, q);
      bool reached = true;
      for (VertexId t : q.targets) reached = reached && std::isfinite(s.dist[t.value()]);
      (reached ? st.reached_target_mask : st.bounding_box_miss_mask) |= LaneMask{1U} << lane;
      if (!compact)
        for (std::uint32_t v = 0; v < V; ++v) e.final_distances[static_cast<std::size_t>(v) * W + lane] = s.dist[v];
      e.logical_lane_edge_work += s.region_edges;
      e.shared_edge_work = std::max(e.shared_edge_work, s.region_edges);
      if (compact) e.compact_paths.push_back(payload_of(g, q, s));
    }
    e.result.work.edges_examined = e.logical_lane_edge_work;
    e.result.work.kernel_dispatches = 1U;
    if (compact) {
      std::sort(e.compact_paths.begin(), e.compact_paths.end(),
                [](const CompactPathPayload& a, const CompactPathPayload& b) { return a.query_id < b.query_id; });
      e.compact_transfer = measure_compact_transfer(e.compact_paths);
      e.compact_status_bytes = sizeof(DeviceRunStatus);
      e.compact_total_device_to_host_bytes = e.compact_transfer.total_bytes + e.compact_status_bytes;
      e.compact_overall_device_to_host_bytes = e.compact_total_device_to_host_bytes;
      e.compact_execution.host_timing = CompactStageTimingEvidence::measured;
    }
    return e;
  }
};

const ExpansionQueryOutcome& outcome_of(const BatchedExpansionRunResult& r, std::uint32_t id) {
  for (const ExpansionQueryOutcome& o : r.queries)
    if (o.final_query.query_id.value() == id) return o;
  throw std::logic_error{"no outcome for query"};
}

// Identities every run satisfies.
void check_metric_identities(const BatchedExpansionRunResult& r, const std::string& where) {
  const BatchedExpansionMetrics& m = r.metrics;
  CHECK_MSG(m.input_queries == r.queries.size(), where);
  CHECK_MSG(m.reached_queries + m.unreachable_full_region_queries + m.expansion_limit_queries +
                    m.stalled_region_queries + m.identity_or_count_overflow_queries + m.engine_failure_queries ==
                m.input_queries,
            where);
  CHECK_MSG(m.batches_executed == m.initial_batches_executed + m.retry_batches_executed && m.planning_passes != 0U,
            where);
  CHECK_MSG(m.retry_valid_lane_observations <= m.retry_lane_capacity &&
                m.failed_lane_observations <= m.failed_origin_valid_lane_observations,
            where);
  CHECK_MSG(m.device_work.expansion_count == m.scheduled_expansions + m.full_region_fallbacks, where);
  std::uint64_t hist = 0, scheduled = 0, fallbacks = 0, tiles = 0, vertices = 0, edges = 0;
  for (std::uint64_t c : m.expansion_count_histogram) hist += c;
  for (const ExpansionQueryOutcome& o : r.queries) {
    scheduled += o.scheduled_expansions;
    fallbacks += o.used_full_region_fallback;
    tiles += o.final_query.selected_tiles.size();
    vertices += o.selected_vertex_count;
    edges += o.selected_edge_count;
    if (o.disposition == ExpansionQueryDisposition::reached ||
        o.disposition == ExpansionQueryDisposition::unreachable_in_full_region)
      CHECK_MSG(o.attempts == o.total_expansions + 1U && o.final_query.expansion_generation == o.total_expansions,
                where);
  }
  CHECK_MSG(hist == m.input_queries && scheduled == m.scheduled_expansions && fallbacks == m.full_region_fallbacks,
            where);
  CHECK_MSG(tiles == m.final_selected_tile_count && vertices == m.final_selected_vertex_count &&
                edges == m.final_selected_edge_count,
            where);
  CHECK_MSG(m.schedule_comparison_fingerprint != 0U, where);
}

// Labels of a terminal outcome equal the host solve of its final region.
void check_final_labels(const WeightedGraph& g, const ExpansionQueryOutcome& o, const std::string& where) {
  CHECK_MSG(o.final_distances.size() == g.vertex_count(), where);
  if (o.final_distances.size() != g.vertex_count()) return;
  const LaneSolution s = solve_lane(g, o.final_query);
  bool equal = true, all_finite = true;
  for (std::uint32_t v = 0; v < g.vertex_count(); ++v) equal = equal && same_bits(o.final_distances[v], s.dist[v]);
  for (VertexId t : o.final_query.targets) all_finite = all_finite && std::isfinite(o.final_distances[t.value()]);
  CHECK_MSG(equal, where);
  CHECK_MSG(all_finite == o.reached(), where);
}

// The schedule matrix over the six single-source routes.
void check_schedule_matrix(const Fixture& f, EngineKind engine, const ExpansionBatchRunner& runner,
                           bool labels, const std::string& tag) {
  const WeightedGraph& g = f.graph();
  const auto queries = f.subset({kImmediate, kTwoRing, kLongEdge, kSpill, kUnreachable, kSecondMiss});
  std::array<ExpansionScheduleEvidence, 4> evidence{};
  for (std::size_t si = 0; si < kSchedules.size(); ++si) {
    const BatchedExpansionOptions o = options_of(engine, kSchedules[si]);
    const BatchedExpansionRunResult r =
        run_batched_expansion(g, f.directory, f.tile_runs, queries, o, runner);
    const std::string where = tag + " schedule " + std::to_string(si);
    CHECK_MSG(r.queries.size() == 6U && r.metrics.initial_reached_queries == 1U &&
                  r.metrics.reached_queries == 5U && r.metrics.unreachable_full_region_queries == 1U &&
                  r.metrics.engine_failure_queries == 0U,
              where);
    CHECK_MSG(!r.trace.empty() && r.trace.front().valid_lane_mask == 0x3FU &&
                  r.trace.front().reached_lane_mask == 0x1U && r.trace.front().miss_lane_mask == 0x3EU,
              where);
    for (const RouteQuery& q : queries) {
      const std::uint32_t id = q.query_id.value();
      const ExpansionQueryOutcome& oc = outcome_of(r, id);
      const std::uint32_t n = expected_expansions(kSchedules[si].kind, id);
      CHECK_MSG(oc.total_expansions == n && oc.attempts == n + 1U, where + " query " + std::to_string(id));
      if (id == kUnreachable)
        CHECK_MSG(oc.disposition == ExpansionQueryDisposition::unreachable_in_full_region &&
                      oc.used_full_region_fallback && oc.scheduled_expansions == 0U,
                  where);
      else
        CHECK_MSG(oc.reached() && !oc.used_full_region_fallback, where + " query " + std::to_string(id));
      if (id == kImmediate) CHECK_MSG(oc.final_query.expansion_generation == 0U, where);
      if (labels) check_final_labels(g, oc, where + " query " + std::to_string(id));
    }
    check_metric_identities(r, where);
    evidence[si] = make_expansion_schedule_evidence(kSchedules[si], r);
  }
  const auto chosen = select_expansion_schedule_from_evidence(evidence);
  CHECK_MSG(chosen.has_value(), tag);
  for (const ExpansionScheduleEvidence& ev : evidence)
    CHECK_MSG(ev.comparison_fingerprint == evidence[0].comparison_fingerprint, tag);
}

// ---------------------------------------------------------------------------
void host_mode() {
  const Fixture f = make_fixture();
  const WeightedGraph& g = f.graph();
  for (EngineKind engine : {EngineKind::jacobi_pull, EngineKind::dense_chaotic_push, EngineKind::frontier_push}) {
    HostRunner runner{g, options_of(engine, one_ring_expansion()).run_options};
    check_schedule_matrix(f, engine, std::ref(runner), true, "host engine " + std::to_string(int(engine)));
  }

  // Fallback after the one scheduled ring; explicit failure reports the limit.
  {
    const auto qs = f.subset({kImmediate, kLongEdge});
    BatchedExpansionOptions o =
        options_of(EngineKind::frontier_push, one_ring_expansion(), 1U, ExpansionTerminalPolicy::full_region_fallback);
    HostRunner runner{g, o.run_options};
    BatchedExpansionRunResult r = run_batched_expansion(g, f.directory, f.tile_runs, qs, o, std::ref(runner));
    const ExpansionQueryOutcome& fb = outcome_of(r, kLongEdge);
    CHECK(outcome_of(r, kImmediate).reached() && outcome_of(r, kImmediate).total_expansions == 0U);
    CHECK(fb.reached() && fb.scheduled_expansions == 1U && fb.total_expansions == 2U && fb.used_full_region_fallback);
    CHECK(r.metrics.initial_reached_queries == 1U && r.metrics.full_region_fallbacks == 1U);
    CHECK(fb.final_query.selected_tiles.size() == g.tile_coordinates().size());
    check_final_labels(g, fb, "fallback");
    check_metric_identities(r, "fallback");

    o.terminal_policy = ExpansionTerminalPolicy::explicit_failure;
    r = run_batched_expansion(g, f.directory, f.tile_runs, qs, o, std::ref(runner));
    const ExpansionQueryOutcome& lim = outcome_of(r, kLongEdge);
    CHECK(lim.disposition == ExpansionQueryDisposition::expansion_limit && lim.scheduled_expansions == 1U &&
          lim.total_expansions == 1U && !lim.used_full_region_fallback);
    CHECK(r.metrics.expansion_limit_queries == 1U && r.metrics.full_region_fallbacks == 0U);
    CHECK(lim.terminal_stop_reason == static_cast<std::uint32_t>(DeviceStopReason::converged));

    // A ring that adds no tile stalls instead of retrying forever.
    const auto lone = f.subset({kUnreachable});
    const BatchedExpansionOptions so =
        options_of(EngineKind::jacobi_pull, one_ring_expansion(), 4U, ExpansionTerminalPolicy::explicit_failure);
    HostRunner jacobi_runner{g, so.run_options};
    r = run_batched_expansion(g, f.directory, f.tile_runs, lone, so, std::ref(jacobi_runner));
    CHECK(r.queries.front().disposition == ExpansionQueryDisposition::region_stalled &&
          r.queries.front().total_expansions == 0U && r.metrics.stalled_region_queries == 1U);
  }

  // Input order does not matter; caller queries are not changed; identities
  // and terminal maps survive restarts.
  {
    const auto ordered = f.subset({kImmediate, kTwoRing, kSecondMiss});
    const std::vector<RouteQuery> shuffled{ordered[2], ordered[0], ordered[1]};
    const auto before = ordered;
    const BatchedExpansionOptions o = options_of(EngineKind::frontier_push, one_ring_expansion());
    HostRunner runner{g, o.run_options};
    BatchedExpansionOptions unbound = o;
    unbound.execution_configuration_fingerprint = 0U;
    CHECK(throws<std::invalid_argument>(
        [&] { (void)run_batched_expansion(g, f.directory, f.tile_runs, ordered, unbound, std::ref(runner)); }));
    const auto a = run_batched_expansion(g, f.directory, f.tile_runs, ordered, o, std::ref(runner));
    const auto b = run_batched_expansion(g, f.directory, f.tile_runs, shuffled, o, std::ref(runner));
    CHECK(ordered == before && a.trace == b.trace);
    CHECK(a.metrics.schedule_comparison_fingerprint == b.metrics.schedule_comparison_fingerprint);
    for (std::size_t i = 0; i < a.queries.size(); ++i) {
      const ExpansionQueryOutcome &x = a.queries[i], &y = b.queries[i];
      CHECK(x.final_query == y.final_query && x.disposition == y.disposition && x.attempts == y.attempts &&
            x.final_distances == y.final_distances);
      const RouteQuery& orig = f.query(x.final_query.query_id.value());
      CHECK(x.final_query.sources == orig.sources && x.final_query.targets == orig.targets &&
            x.final_query.source_terminals == orig.source_terminals &&
            x.final_query.target_terminal_to_target == orig.target_terminal_to_target);
    }
    CHECK(outcome_of(a, kTwoRing).final_query.selected_tiles != outcome_of(a, kSecondMiss).final_query.selected_tiles);
  }

  // Multi-source restart, identity limits, and measured retry work.
  {
    const BatchedExpansionOptions o = options_of(EngineKind::jacobi_pull, one_ring_expansion());
    HostRunner runner{g, o.run_options};
    const auto multi = f.subset({kMultiSource});
    auto r = run_batched_expansion(g, f.directory, f.tile_runs, multi, o, std::ref(runner));
    const ExpansionQueryOutcome& mo = r.queries.front();
    CHECK(mo.reached() && mo.attempts == 3U && mo.total_expansions == 2U && mo.final_query.sources.size() == 2U &&
          mo.final_query.sources == multi.front().sources);
    for (VertexId s : mo.final_query.sources) CHECK(same_bits(mo.final_distances[s.value()], 0.0F));
    check_final_labels(g, mo, "multi-source");

    RouteQuery high = f.query(kTwoRing);
    high.query_id = QueryId{0xFFFFFFFFu};
    high.expansion_generation = 7U;
    r = run_batched_expansion(g, f.directory, f.tile_runs, std::vector<RouteQuery>{high}, o, std::ref(runner));
    CHECK(r.queries.front().reached() && r.queries.front().final_query.expansion_generation == 9U &&
          r.queries.front().total_expansions == 2U);

    RouteQuery top = f.query(kTwoRing);
    top.expansion_generation = 0xFFFFFFFFu;
    r = run_batched_expansion(g, f.directory, f.tile_runs, std::vector<RouteQuery>{top}, o, std::ref(runner));
    CHECK(r.queries.front().disposition == ExpansionQueryDisposition::identity_or_count_overflow &&
          r.queries.front().attempts == 1U && r.queries.front().total_expansions == 0U &&
          r.metrics.identity_or_count_overflow_queries == 1U && r.metrics.scheduled_expansions == 0U &&
          r.metrics.full_region_fallbacks == 0U);

    // Fixed work per batch: 3 shared, 7 logical; two retries of one lane.
    std::uint32_t calls = 0;
    const ExpansionBatchRunner fixed = [&](std::span<const RouteQuery> cur, std::span<const BatchQueryFeatures> feat,
                                           const BatchPlanEntry& batch, const ExpansionBatchContext& ctx) {
      ++calls;
      CHECK(cur.size() == 1U && cur.front().sources == f.query(kTwoRing).sources &&
            ctx.planning_pass == calls - 1U && ctx.retry_pass == (calls > 1U));
      ExpansionBatchExecution e = runner(cur, feat, batch, ctx);
      e.final_distances.clear();
      e.shared_edge_work = 3U;
      e.logical_lane_edge_work = 7U;
      e.result.work = DeviceWorkStatistics{};
      e.result.work.edges_examined = 7U;
      e.result.work.kernel_dispatches = 1U;
      return e;
    };
    r = run_batched_expansion(g, f.directory, f.tile_runs, f.subset({kTwoRing}), o, fixed);
    const BatchedExpansionMetrics& m = r.metrics;
    CHECK(m.batches_executed == 3U && m.work_measured_batches == 3U && m.shared_edge_work == 9U &&
          m.logical_lane_edge_work == 21U && m.retry_work_measured_batches == 2U && m.retry_shared_edge_work == 6U &&
          m.retry_logical_lane_edge_work == 14U && m.failed_batch_work_measured_batches == 2U &&
          m.failed_batch_shared_edge_work == 6U && m.failed_batch_logical_lane_edge_work == 14U &&
          m.device_work.edges_examined == 21U && m.device_work.kernel_dispatches == 3U &&
          m.device_work.expansion_count == 2U);
    CHECK(r.trace.size() == 3U && r.trace[0].expansion_generations_by_lane[0] == 0U &&
          r.trace[1].expansion_generations_by_lane[0] == 1U && r.trace[2].expansion_generations_by_lane[0] == 2U);
    CHECK(r.trace[1].context.retry_pass && r.trace[2].context.execution_ordinal == 2U);
    CHECK((m.expansion_count_histogram == std::vector<std::uint64_t>{0U, 0U, 1U}));
    CHECK(r.queries.front().final_distances.empty());
  }

  // Engine stops that are not clean never expand; malformed runner results
  // are rejected.
  {
    const auto qs = f.subset({kImmediate, kTwoRing});
    const BatchedExpansionOptions o = options_of(EngineKind::frontier_push, one_ring_expansion());
    const auto status_run = [&](DeviceStopReason reason, LaneMask reached, LaneMask missed,
                                const std::function<void(ExpansionBatchExecution&)>& edit) {
      std::uint32_t calls = 0;
      const ExpansionBatchRunner r = [&](std::span<const RouteQuery>, std::span<const BatchQueryFeatures>,
                                         const BatchPlanEntry& batch, const ExpansionBatchContext&) {
        ++calls;
        ExpansionBatchExecution e;
        e.result.engine_kind = static_cast<std::uint32_t>(o.run_options.engine);
        e.result.control_mode = static_cast<std::uint32_t>(o.run_options.control_mode);
        DeviceRunStatus& st = e.result.status;
        st.valid_lane_mask = batch.valid_lane_mask;
        st.reached_target_mask = reached & batch.valid_lane_mask;
        st.bounding_box_miss_mask = missed & batch.valid_lane_mask;
        st.rounds_completed = 1U;
        st.stop_reason = static_cast<std::uint32_t>(reason);
        if (reason == DeviceStopReason::converged) {
          st.converged = 1U;
          st.converged_lane_mask = batch.valid_lane_mask;
        } else {
          st.active_lane_mask = batch.valid_lane_mask;
        }
        if (reason == DeviceStopReason::queue_overflow) st.error_bits = device_error::queue_overflow;
        e.work_evidence = ExpansionWorkEvidence::measured;
        e.shared_edge_work = 3U;
        e.logical_lane_edge_work = 7U;
        e.result.work.edges_examined = 7U;
        e.final_distances.assign(static_cast<std::size_t>(g.vertex_count()) * batch.lane_width, 1.0F);
        if (edit) edit(e);
        return e;
      };
      auto result = run_batched_expansion(g, f.directory, f.tile_runs, qs, o, r);
      return std::make_pair(std::move(result), calls);
    };
    for (DeviceStopReason reason : {DeviceStopReason::maximum_rounds, DeviceStopReason::queue_overflow}) {
      const auto [r, calls] = status_run(reason, 0U, 0U, {});
      CHECK(calls == 1U && r.metrics.engine_failure_queries == 2U && r.metrics.scheduled_expansions == 0U);
      for (const ExpansionQueryOutcome& oc : r.queries)
        CHECK(oc.disposition == ExpansionQueryDisposition::engine_failure && oc.final_distances.empty() &&
              oc.final_query.expansion_generation == 0U &&
              oc.terminal_stop_reason == static_cast<std::uint32_t>(reason));
    }
    CHECK(throws<std::logic_error>([&] { (void)status_run(DeviceStopReason::converged, 0x1U, 0x0U, {}); }));
    CHECK(throws<std::logic_error>([&] { (void)status_run(DeviceStopReason::maximum_rounds, 0x1U, 0x0U, {}); }));
    CHECK(throws<std::logic_error>([&] {
      (void)status_run(DeviceStopReason::converged, 0x1U, 0x2U, [](ExpansionBatchExecution& e) {
        e.final_distances.pop_back();
      });
    }));
    CHECK(throws<std::logic_error>([&] {
      (void)status_run(DeviceStopReason::converged, 0x1U, 0x2U, [](ExpansionBatchExecution& e) {
        e.work_evidence = static_cast<ExpansionWorkEvidence>(7);
      });
    }));
    CHECK(throws<std::logic_error>([&] {
      (void)status_run(DeviceStopReason::converged, 0x1U, 0x2U, [](ExpansionBatchExecution& e) {
        e.work_evidence = ExpansionWorkEvidence::unavailable;
        e.result.work = DeviceWorkStatistics{};
      });
    }));
    CHECK(throws<std::logic_error>([&] {
      (void)status_run(DeviceStopReason::converged, 0x1U, 0x2U, [](ExpansionBatchExecution& e) {
        e.work_evidence = ExpansionWorkEvidence::unavailable;
        e.shared_edge_work = e.logical_lane_edge_work = 0U;
      });
    }));
    CHECK(throws<std::logic_error>([&] {
      (void)status_run(DeviceStopReason::converged, 0x1U, 0x2U, [](ExpansionBatchExecution& e) {
        e.shared_edge_work = 8U;
      });
    }));
    CHECK(throws<std::logic_error>([&] {
      (void)status_run(DeviceStopReason::converged, 0x1U, 0x2U, [](ExpansionBatchExecution& e) {
        e.result.status.valid_lane_mask = 0x1U;
        e.result.status.converged_lane_mask = 0x1U;
        e.result.status.bounding_box_miss_mask = 0U;
      });
    }));
    CHECK(throws<std::logic_error>([&] {
      (void)status_run(DeviceStopReason::converged, 0x1U, 0x2U, [](ExpansionBatchExecution& e) {
        e.result.engine_kind = static_cast<std::uint32_t>(EngineKind::jacobi_pull);
      });
    }));
    CHECK(throws<std::logic_error>([&] {
      (void)status_run(DeviceStopReason::converged, 0x1U, 0x2U, [](ExpansionBatchExecution& e) {
        e.compact_status_bytes = sizeof(DeviceRunStatus);
        e.compact_total_device_to_host_bytes = sizeof(DeviceRunStatus);
        e.compact_overall_device_to_host_bytes = sizeof(DeviceRunStatus);
      });
    }));
    // Unavailable work evidence stays explicit.
    {
      const auto [r, calls] = status_run(DeviceStopReason::converged, 0x3U, 0x0U, [](ExpansionBatchExecution& e) {
        e.work_evidence = ExpansionWorkEvidence::unavailable;
        e.shared_edge_work = e.logical_lane_edge_work = 0U;
        e.result.work = DeviceWorkStatistics{};
      });
      CHECK(calls == 1U && r.metrics.work_evidence == ExpansionWorkEvidence::unavailable &&
            r.metrics.work_measured_batches == 0U && r.metrics.reached_queries == 2U);
    }
    // No queries: no batch, an empty histogram bin.
    std::uint32_t never = 0;
    const ExpansionBatchRunner none = [&](std::span<const RouteQuery>, std::span<const BatchQueryFeatures>,
                                          const BatchPlanEntry&, const ExpansionBatchContext&) {
      ++never;
      return ExpansionBatchExecution{};
    };
    const auto empty = run_batched_expansion(g, f.directory, f.tile_runs, std::vector<RouteQuery>{}, o, none);
    CHECK(never == 0U && empty.queries.empty() && empty.trace.empty() &&
          (empty.metrics.expansion_count_histogram == std::vector<std::uint64_t>{0U}));
    CHECK(throws<std::invalid_argument>(
        [&] { (void)run_batched_expansion(g, f.directory, f.tile_runs, qs, o, ExpansionBatchRunner{}); }));
    const std::vector<RouteQuery> twice{qs[0], qs[0]};
    CHECK(throws<std::invalid_argument>(
        [&] { (void)run_batched_expansion(g, f.directory, f.tile_runs, twice, o, none); }));
    TileRunLayout64 other_runs = f.tile_runs;
    other_runs.csr_run_destination_tiles.push_back(TileId{0U});
    CHECK(throws<std::invalid_argument>(
        [&] { (void)run_batched_expansion(g, f.directory, other_runs, qs, o, none); }));
  }

  // Options validation and evidence-only schedule selection.
  {
    using E = BatchedExpansionOptionsError;
    BatchedExpansionOptions o = options_of(EngineKind::jacobi_pull, one_ring_expansion());
    CHECK(validate_batched_expansion_options(o) == E::none);
    BatchedExpansionOptions x = o;
    x.schedule = {};
    CHECK(validate_batched_expansion_options(x) == E::unspecified_schedule);
    x.schedule.kind = static_cast<ExpansionScheduleKind>(9);
    CHECK(validate_batched_expansion_options(x) == E::invalid_schedule);
    x = o;
    x.execution_configuration_fingerprint = 0U;
    CHECK(validate_batched_expansion_options(x) == E::missing_execution_configuration_fingerprint);
    x = o;
    x.schedule = fixed_ring_expansion(1U);
    CHECK(validate_batched_expansion_options(x) == E::invalid_fixed_ring_size);
    x.schedule = ExpansionSchedulePolicy{ExpansionScheduleKind::one_geometric_ring, 2U, 0U};
    CHECK(validate_batched_expansion_options(x) == E::unexpected_fixed_ring_size);
    x.schedule = ExpansionSchedulePolicy{ExpansionScheduleKind::doubling_xy_margins, 0U, 1U};
    CHECK(validate_batched_expansion_options(x) == E::unexpected_hybrid_small_expansion_count);
    x.schedule = hybrid_margin_expansion(0U);
    CHECK(validate_batched_expansion_options(x) == E::invalid_hybrid_small_expansion_count);
    x.schedule = ExpansionSchedulePolicy{ExpansionScheduleKind::hybrid_small_then_doubling, 3U, 1U};
    CHECK(validate_batched_expansion_options(x) == E::unexpected_fixed_ring_size);
    x = o;
    x.run_options.block_size = 0U;
    CHECK(validate_batched_expansion_options(x) == E::invalid_run_options);
    x = o;
    x.planner_policy.lane_width = 7U;
    CHECK(validate_batched_expansion_options(x) == E::invalid_planner_policy);
    x = o;
    x.terminal_policy = static_cast<ExpansionTerminalPolicy>(5);
    CHECK(validate_batched_expansion_options(x) == E::invalid_terminal_policy);
    x = o;
    x.enable_compact_paths = 2U;
    CHECK(validate_batched_expansion_options(x) == E::invalid_compact_paths_flag);
    CHECK(o == options_of(EngineKind::jacobi_pull, one_ring_expansion()));
    x = o;
    x.maximum_expansions = 5U;
    CHECK(!(x == o));

    // Margin growth of each schedule, seen through the regions it selects:
    // the two-ring route under fixed ring 2 reaches in one expansion.
    ExpansionScheduleEvidence base{};
    base.input_queries = 6U;
    base.reached_queries = 5U;
    base.terminal_failures = 1U;
    base.batches_executed = 3U;
    base.total_expansions = 10U;
    base.attempted_selected_edge_estimate = 40U;
    base.work_evidence = ExpansionWorkEvidence::measured;
    base.shared_edge_work = 10U;
    base.logical_lane_edge_work = 20U;
    base.comparison_fingerprint = 0xABCDU;
    base.campaign_valid = true;
    std::array<ExpansionScheduleEvidence, 4> ev{};
    for (std::size_t i = 0; i < 4; ++i) {
      ev[i] = base;
      ev[i].schedule = kSchedules[i];
      ev[i].logical_lane_edge_work = 20U + i;
    }
    ev[2].logical_lane_edge_work = 19U;
    CHECK(select_expansion_schedule_from_evidence(ev) == doubling_margin_expansion());
    CHECK(!select_expansion_schedule_from_evidence(std::span(ev).first(3)).has_value());
    auto dup = ev;
    dup[3].schedule = kSchedules[0];
    CHECK(!select_expansion_schedule_from_evidence(dup).has_value());
    auto unlike = ev;
    unlike[1].comparison_fingerprint = 0x1U;
    CHECK(!select_expansion_schedule_from_evidence(unlike).has_value());
    auto failed = ev;
    failed[0].campaign_valid = false;
    CHECK(!select_expansion_schedule_from_evidence(failed).has_value());
    auto no_batch = ev;
    no_batch[0].batches_executed = 0U;
    CHECK(!select_expansion_schedule_from_evidence(no_batch).has_value());
    auto smuggled = ev;
    smuggled[0].work_evidence = ExpansionWorkEvidence::unavailable;
    CHECK(!select_expansion_schedule_from_evidence(smuggled).has_value());
    auto impossible = ev;
    impossible[0].logical_lane_edge_work = 1U;
    CHECK(!select_expansion_schedule_from_evidence(impossible).has_value());
    auto tie = ev;
    tie[2].logical_lane_edge_work = 20U;  // ties schedule 0
    CHECK(!select_expansion_schedule_from_evidence(tie).has_value());
    auto unmeasured = ev;  // unavailable work: selected-edge estimate decides
    for (auto& r : unmeasured) {
      r.work_evidence = ExpansionWorkEvidence::unavailable;
      r.shared_edge_work = r.logical_lane_edge_work = 0U;
    }
    unmeasured[1].attempted_selected_edge_estimate = 39U;
    CHECK(select_expansion_schedule_from_evidence(unmeasured) == fixed_ring_expansion(2U));
    BatchedExpansionRunResult saturated;
    saturated.metrics.input_queries = 1U;
    saturated.metrics.scheduled_expansions = ~std::uint64_t{0};
    saturated.metrics.full_region_fallbacks = 1U;
    const ExpansionScheduleEvidence sat = make_expansion_schedule_evidence(one_ring_expansion(), saturated);
    CHECK(sat.total_expansions == ~std::uint64_t{0} && !sat.campaign_valid && sat.terminal_failures == 1U);
    auto zero = ev;
    for (auto& r : zero) r.input_queries = r.reached_queries = r.terminal_failures = 0U;
    CHECK(!select_expansion_schedule_from_evidence(zero).has_value());
  }

  // Compact mode over the host runner: payloads retained for reached and
  // full-region outcomes, validated, and counted.
  {
    const auto qs = f.subset({kImmediate, kTwoRing, kLongEdge, kSpill, kUnreachable, kSecondMiss});
    BatchedExpansionOptions o = options_of(EngineKind::dense_chaotic_push, doubling_margin_expansion());
    o.enable_compact_paths = 1U;
    HostRunner runner{g, o.run_options, true};
    const auto r = run_batched_expansion(g, f.directory, f.tile_runs, qs, o, std::ref(runner));
    CHECK(r.metrics.reached_queries == 5U && r.metrics.unreachable_full_region_queries == 1U);
    CHECK(r.metrics.compact_host_timing == CompactStageTimingEvidence::measured);
    CHECK(r.metrics.compact_device_timing == CompactStageTimingEvidence::unavailable);
    std::uint64_t bytes = 0;
    for (const ExpansionQueryOutcome& oc : r.queries) {
      CHECK(oc.compact_paths.has_value() && oc.final_distances.empty());
      if (!oc.compact_paths) continue;
      const CompactQueryResult cr = make_compact_query_result(*oc.compact_paths, oc.disposition);
      CHECK(validate_compact_query_result(g, oc.final_query, cr).ok());
      CHECK(oc.compact_paths->expansion_generation == oc.final_query.expansion_generation);
      const LaneSolution s = solve_lane(g, oc.final_query);
      CHECK(validate_compact_query_result_against_distances(g, oc.final_query, cr, s.dist).ok());
      bytes += measure_compact_transfer(std::span(&*oc.compact_paths, 1)).total_bytes;
    }
    CHECK(r.metrics.compact_transfer.total_bytes >= bytes);
    CHECK(r.metrics.compact_total_device_to_host_bytes ==
          r.metrics.compact_transfer.total_bytes + r.metrics.batches_executed * sizeof(DeviceRunStatus));
    const ExpansionQueryOutcome& un = outcome_of(r, kUnreachable);
    CHECK(un.compact_paths && un.compact_paths->targets.front().summary.reconstruction ==
                                  CompactPathStatus::unreachable);

    // A payload that crosses lanes or fails validation is rejected.
    const ExpansionBatchRunner bad = [&](std::span<const RouteQuery> cur, std::span<const BatchQueryFeatures> feat,
                                         const BatchPlanEntry& batch, const ExpansionBatchContext& ctx) {
      ExpansionBatchExecution e = runner(cur, feat, batch, ctx);
      auto& p = e.compact_paths.front();
      if (!p.targets.front().vertices.empty()) p.targets.front().distance_labels.back() += 1.0F;
      return e;
    };
    CHECK(throws<std::logic_error>(
        [&] { (void)run_batched_expansion(g, f.directory, f.tile_runs, f.subset({kImmediate}), o, bad); }));
    const ExpansionBatchRunner untimed = [&](std::span<const RouteQuery> cur, std::span<const BatchQueryFeatures> feat,
                                             const BatchPlanEntry& batch, const ExpansionBatchContext& ctx) {
      ExpansionBatchExecution e = runner(cur, feat, batch, ctx);
      e.compact_execution.host_timing = CompactStageTimingEvidence::unavailable;
      return e;
    };
    CHECK(throws<std::logic_error>(
        [&] { (void)run_batched_expansion(g, f.directory, f.tile_runs, f.subset({kImmediate}), o, untimed); }));
  }

  // Payload checks on a reached route: every rule fires on its own mistake.
  {
    const RouteQuery q = f.query(kTwoRing);
    RouteQuery wide = q;
    wide.selected_tiles.clear();
    for (std::uint32_t t = 0; t < g.tile_coordinates().size(); ++t) wide.selected_tiles.push_back(TileId{t});
    const LaneSolution s = solve_lane(g, wide);
    const CompactPathPayload good = payload_of(g, wide, s);
    using C = CompactPathValidationErrorCode;
    CHECK(validate_compact_path_payload(g, wide, good).ok());
    CHECK(compact_path_payload_complete(good));
    CHECK(good.targets.front().vertices.size() == 4U && same_bits(good.targets.front().summary.distance, 3.0F));
    const auto code_of = [&](const std::function<void(CompactPathPayload&)>& edit, const RouteQuery& query) {
      CompactPathPayload p = good;
      edit(p);
      return validate_compact_path_payload(g, query, p).code;
    };
    CHECK(code_of([](auto& p) { p.query_id = QueryId{1U}; }, wide) == C::query_identity_mismatch);
    CHECK(code_of([](auto& p) { p.expansion_generation = 3U; }, wide) == C::expansion_generation_mismatch);
    CHECK(code_of([](auto& p) { p.target_terminal_to_target.push_back(0U); }, wide) == C::terminal_map_mismatch);
    CHECK(code_of([](auto& p) { p.targets.push_back(p.targets.front()); }, wide) == C::target_count_mismatch);
    CHECK(code_of([](auto& p) { p.targets[0].summary.target = VertexId{0U}; }, wide) == C::target_identity_mismatch);
    CHECK(code_of([](auto& p) { p.targets[0].summary.distance = std::nanf(""); }, wide) == C::invalid_distance);
    CHECK(code_of([](auto& p) { p.targets[0].edge_ids.pop_back(); }, wide) == C::path_shape_mismatch);
    CHECK(code_of([](auto& p) { p.targets[0].summary.path_length = 9U; }, wide) == C::path_length_mismatch);
    CHECK(code_of([](auto& p) { std::swap(p.targets[0].edge_ids[0], p.targets[0].edge_ids[1]); }, wide) ==
          C::edge_continuity_mismatch);
    CHECK(code_of([](auto& p) { p.targets[0].edge_ids[0] = EdgeId{999U}; }, wide) == C::edge_id_out_of_range);
    CHECK(code_of([](auto& p) { p.targets[0].distance_labels[1] = 0.75F; }, wide) == C::non_tight_edge);
    CHECK(code_of([](auto& p) { p.targets[0].summary.has_selected_source = 2U; }, wide) ==
          C::invalid_selected_source_flag);
    CHECK(code_of([](auto&) {}, q) == C::vertex_outside_selected_region);
    CHECK(code_of([](auto& p) {
            auto& t = p.targets[0];
            t.summary = CompactTargetSummary{t.summary.target};
            t.summary.reconstruction = CompactPathStatus::unreachable;
          }, wide) == C::inconsistent_unreachable_result);
    CHECK(code_of([](auto& p) {
            auto& t = p.targets[0];
            const VertexId target = t.summary.target;
            t = CompactTargetPath{};
            t.summary.target = target;
            t.summary.reconstruction = CompactPathStatus::query_terminal_failure;
          }, wide) == C::none);
    const CompactQueryResult as_failure = make_compact_query_result(good, ExpansionQueryDisposition::engine_failure);
    CHECK(validate_compact_query_result(g, wide, as_failure).code == C::failed_query_reported_complete);
    const CompactQueryResult as_limit = make_compact_query_result(good, ExpansionQueryDisposition::expansion_limit);
    CHECK(validate_compact_query_result(g, wide, as_limit).code == C::inconsistent_failure_result);
    std::vector<float> shifted = s.dist;
    shifted[good.targets[0].vertices[1].value()] += 1.0F;
    CHECK(validate_compact_query_result_against_distances(
              g, wide, make_compact_query_result(good, ExpansionQueryDisposition::reached), shifted)
              .code == C::target_distance_mismatch);
    const CompactTransferAccounting a = measure_compact_transfer(std::span(&good, 1));
    CHECK(a.summary_bytes == 28U && a.vertex_bytes == 16U && a.distance_label_bytes == 16U &&
          a.edge_id_bytes == 12U && a.total_bytes == 72U);
  }

  // The no-congestion pipeline over the host runner: results, accounting,
  // ledger and quality sample.
  {
    const auto qs = f.subset({kImmediate, kTwoRing, kLongEdge, kSpill, kUnreachable, kSecondMiss, kMultiSource});
    BatchedExpansionOptions o = options_of(EngineKind::frontier_push, one_ring_expansion());
    o.enable_compact_paths = 1U;
    HostRunner runner{g, o.run_options, true};
    NoCongestionPipelineOptions po;
    po.quality_sample_query_count = 5U;
    po.quality_selection_seed = 42U;
    po.cold_artifact_load = measured_host_stage(10U);
    po.graph_upload = measured_host_stage(20U);
    po.cold_execution = measured_host_stage(30U);
    const auto r = run_no_congestion_pipeline(g, f.directory, f.tile_runs, qs, o, std::ref(runner), po);
    const NoCongestionResultAccounting& a = r.result_accounting;
    CHECK(r.compact_results.size() == 7U && a.query_count == 7U && a.target_summary_count == 7U);
    CHECK(a.complete_path_count == 6U && a.unreachable_target_count == 1U && a.terminal_failure_target_count == 0U);
    CHECK(a.modeled_batch_status_count == r.expansion.metrics.batches_executed);
    CHECK(a.device_transfer_evidence == CompactStageTimingEvidence::measured &&
          a.actual_status_bytes == a.modeled_batch_status_bytes);
    std::uint64_t payload = 0;
    for (const CompactQueryResult& cr : r.compact_results) {
      CHECK(validate_compact_query_result(g, outcome_of(r.expansion, cr.query_id.value()).final_query, cr).ok());
      for (const CompactTargetPath& t : cr.targets) payload += 28U + 8U * t.vertices.size() + 4U * t.edge_ids.size();
    }
    CHECK(a.final_result_serialization.total_bytes == payload);
    CHECK(validate_no_congestion_stage_ledger(r.timing) == NoCongestionStageLedgerError::none);
    CHECK(r.timing.cold_pipeline.host_evidence == PipelineTimingEvidence::measured &&
          r.timing.cold_pipeline.host_nanoseconds == 60U);
    CHECK(r.timing.warm_all_query.host_nanoseconds == r.expansion.metrics.total_nanoseconds);
    // Quality: every sampled complete path is the unbounded shortest path here.
    CHECK(r.quality.sampled_query_count == 5U && r.quality.sampled_query_ids.size() == 5U);
    CHECK(r.quality.finite_target_pairs == r.quality.observations.size() && r.quality.finite_target_pairs >= 4U);
    for (const CompactPathQualityObservation& ob : r.quality.observations)
      CHECK(ob.cost_inflation_ratio == 1.0 && ob.path_length_inflation_ratio == 1.0 &&
            ob.absolute_cost_inflation == 0.0 && ob.absolute_path_length_inflation == 0);
    CHECK(r.quality.cost_ratio_p50 == 1.0 && r.quality.cost_ratio_max == 1.0);
    // The sample does not depend on input order.
    std::vector<RouteQuery> reversed(qs.rbegin(), qs.rend());
    const auto r2 = run_no_congestion_pipeline(g, f.directory, f.tile_runs, reversed, o, std::ref(runner), po);
    CHECK(r2.quality.sampled_query_ids == r.quality.sampled_query_ids &&
          r2.quality.observations == r.quality.observations);
    // A bounded detour shows as inflation: the two-ring route forced through
    // a longer region-limited path is compared with its unbounded one.
    std::vector<ExpansionQueryOutcome> outs{outcome_of(r.expansion, kTwoRing)};
    std::vector<CompactQueryResult> res{r.compact_results[1]};
    CHECK(res[0].query_id.value() == kTwoRing);
    res[0].targets[0].summary.distance = 3.0F;  // unchanged: ratio 1
    CHECK(sample_compact_path_quality(g, outs, res, 1U, 7U).cost_ratio_max == 1.0);

    // Rejections.
    CHECK(throws<std::invalid_argument>([&] {
      BatchedExpansionOptions no_paths = o;
      no_paths.enable_compact_paths = 0U;
      HostRunner plain{g, o.run_options};
      (void)run_no_congestion_pipeline(g, f.directory, f.tile_runs, qs, no_paths, std::ref(plain), po);
    }));
    std::vector<CompactQueryResult> swapped = r.compact_results;
    std::swap(swapped[0], swapped[1]);
    CHECK(throws<std::invalid_argument>([&] { (void)measure_no_congestion_result_transfer(swapped, 1U); }));
    std::vector<CompactQueryResult> broken = r.compact_results;
    broken[0].targets[0].edge_ids.push_back(EdgeId{0U});
    CHECK(throws<std::invalid_argument>([&] { (void)measure_no_congestion_result_transfer(broken, 1U); }));
    BatchedExpansionMetrics lying = r.expansion.metrics;
    lying.compact_status_bytes += 48U;
    CHECK(throws<std::invalid_argument>([&] {
      (void)measure_no_congestion_result_transfer(r.compact_results, r.expansion.metrics.batches_executed, &lying);
    }));
    CHECK(throws<std::invalid_argument>(
        [&] { (void)make_failed_compact_query_result(qs[0], ExpansionQueryDisposition::reached); }));
    const CompactQueryResult failed = make_failed_compact_query_result(qs[1], ExpansionQueryDisposition::expansion_limit);
    CHECK(validate_compact_query_result(g, qs[1], failed).ok());

    // Ledger rules, one per error code.
    using E = NoCongestionStageLedgerError;
    const auto ledger = [&](auto&& edit) {
      NoCongestionStageLedger l = r.timing;
      edit(l);
      return validate_no_congestion_stage_ledger(l);
    };
    CHECK(ledger([](auto& l) { l.sssp.host_evidence = static_cast<PipelineTimingEvidence>(9); }) == E::invalid_evidence);
    CHECK(ledger([](auto& l) { l.graph_upload = PipelineStageTiming{PipelineTimingEvidence::unavailable, 5U}; }) ==
          E::unavailable_host_has_value);
    CHECK(ledger([](auto& l) { l.sssp.device_milliseconds = 1.0; }) == E::unavailable_device_has_value);
    CHECK(ledger([](auto& l) {
            l.sssp.device_evidence = PipelineTimingEvidence::measured;
            l.sssp.device_milliseconds = -1.0;
          }) == E::invalid_device_duration);
    CHECK(ledger([](auto& l) {
            l.batch_planning.device_evidence = PipelineTimingEvidence::measured;
            l.batch_planning.device_milliseconds = 1.0;
          }) == E::cpu_only_device_evidence);
    CHECK(ledger([](auto& l) { l.warm_all_query = PipelineStageTiming{}; }) == E::warm_evidence_mismatch);
    CHECK(ledger([](auto& l) { l.warm_all_query.host_nanoseconds += 1U; }) == E::warm_sum_mismatch);
    CHECK(ledger([](auto& l) { l.cold_pipeline = PipelineStageTiming{}; }) == E::cold_evidence_mismatch);
    CHECK(ledger([](auto& l) { l.cold_pipeline.host_nanoseconds = 61U; }) == E::cold_sum_mismatch);
    CHECK(ledger([](auto& l) {
            l.cold_artifact_load.host_nanoseconds = ~std::uint64_t{0};
            l.cold_pipeline.host_nanoseconds = 0U;
          }) == E::cold_sum_overflow);
    CHECK(throws<std::invalid_argument>([&] {
      (void)make_no_congestion_stage_ledger({}, {}, measured_host_stage(5U), {}, {}, {}, {}, {},
                                            measured_host_stage(4U), {});
    }));
  }

  // Path quality sees region-bound detours: s -> b -> t (cost 10, b inside
  // the terminal rectangle) versus s -> a -> t (cost 2, a three tiles away).
  {
    const ResourceClassId rc{1U};
    std::vector<VertexMetadata> v{VertexMetadata::located(0, 0, rc), VertexMetadata::located(20, 0, rc),
                                  VertexMetadata::located(10, 30, rc), VertexMetadata::located(10, 0, rc)};
    std::vector<EdgeInputRecord> e{edge(0, 2, 1.0F, 1), edge(2, 1, 1.0F, 2), edge(0, 3, 5.0F, 3),
                                   edge(3, 1, 5.0F, 4)};
    const UniformGridPartitioner grid{SpatialOrderConfig{0, 0, 10U, 10U}};
    const PartitionedGraph pg = grid.partition(InputGraph{std::move(v), std::move(e)});
    const WeightedGraph& dg = pg.graph;
    const auto to_new = dg.old_to_new();
    const std::vector<VertexId> s1{to_new[0]}, t1{to_new[1]};
    const std::vector<RouteQuery> dq{make_route_query(QueryId{5U}, dg, s1, t1)};
    BatchedExpansionOptions o = options_of(EngineKind::jacobi_pull, one_ring_expansion());
    o.enable_compact_paths = 1U;
    HostRunner runner{dg, o.run_options, true};
    NoCongestionPipelineOptions po;
    po.quality_sample_query_count = 1U;
    const auto r = run_no_congestion_pipeline(dg, build_tile_directory(dg), build_tile_run_layout(dg), dq, o,
                                              std::ref(runner), po);
    CHECK(r.expansion.queries.front().reached() && r.expansion.queries.front().total_expansions == 0U);
    CHECK(r.quality.observations.size() == 1U);
    if (r.quality.observations.size() == 1U) {
      const CompactPathQualityObservation& ob = r.quality.observations.front();
      CHECK(ob.bounded_distance == 10.0F && ob.unbounded_distance == 2.0F && ob.cost_inflation_ratio == 5.0 &&
            ob.absolute_cost_inflation == 8.0 && ob.bounded_path_length == 2U && ob.unbounded_path_length == 2U &&
            ob.path_length_inflation_ratio == 1.0);
      CHECK(r.quality.cost_ratio_p50 == 5.0 && r.quality.absolute_cost_inflation_max == 8.0);
    }
  }

  // The spill tile joins a region only through a real directory adjacency.
  {
    const ResourceClassId rc{1U};
    std::vector<VertexMetadata> v{VertexMetadata::located(0, 0, rc), VertexMetadata::located(20, 0, rc),
                                  VertexMetadata::located(0, 20, rc), VertexMetadata::unlocated(rc),
                                  VertexMetadata::located(0, 10, rc)};
    std::vector<EdgeInputRecord> e{edge(0, 2, 1.0F, 17100), edge(2, 3, 1.0F, 17101), edge(3, 1, 1.0F, 17102)};
    const UniformGridPartitioner grid{SpatialOrderConfig{0, 0, 10U, 10U}};
    const PartitionedGraph pg = grid.partition(InputGraph{std::move(v), std::move(e)});
    const WeightedGraph& sg = pg.graph;
    const auto to_new = sg.old_to_new();
    const std::vector<VertexId> s{to_new[0]}, t{to_new[1]};
    const RouteQuery q = make_route_query(QueryId{1799U}, sg, s, t);
    const TileDirectory dir = build_tile_directory(sg);
    const TileRunLayout64 runs = build_tile_run_layout(sg);
    CHECK(!std::binary_search(q.selected_tiles.begin(), q.selected_tiles.end(), dir.spill_tile()));
    const BatchedExpansionOptions o =
        options_of(EngineKind::jacobi_pull, one_ring_expansion(), 2U, ExpansionTerminalPolicy::explicit_failure);
    HostRunner runner{sg, o.run_options};
    const auto r = run_batched_expansion(sg, dir, runs, std::vector<RouteQuery>{q}, o, std::ref(runner));
    const ExpansionQueryOutcome& oc = r.queries.front();
    CHECK(oc.reached() && oc.scheduled_expansions == 2U && !oc.used_full_region_fallback);
    CHECK(std::binary_search(oc.final_query.selected_tiles.begin(), oc.final_query.selected_tiles.end(),
                             dir.spill_tile()));
    CHECK(oc.final_distances.size() == sg.vertex_count() && same_bits(oc.final_distances[t[0].value()], 3.0F));
    // A directory built from another graph is rejected.
    CHECK(dir.built_from(sg) && !dir.built_from(g));
    // A copy holds the same data: the directory and the memoised region
    // index serve it, and the run on the copy equals the run on the
    // original.  A moved graph keeps them; the moved-from graph refuses.
    {
      const WeightedGraph copy = sg;
      CHECK(dir.built_from(copy) && graph_region_index(copy).matches_graph(copy) &&
            graph_region_index(copy).matches_graph(sg) && !graph_region_index(sg).matches_graph(g));
      HostRunner copy_runner{copy, o.run_options};
      const auto rc2 = run_batched_expansion(copy, dir, runs, std::vector<RouteQuery>{q}, o, std::ref(copy_runner));
      CHECK(rc2.trace == r.trace && rc2.queries.front().final_query == oc.final_query &&
            rc2.queries.front().scheduled_expansions == oc.scheduled_expansions);
      // The graph's memoised runs are taken without an element comparison.
      const auto rc3 = run_batched_expansion(copy, dir, graph_tile_runs(copy), std::vector<RouteQuery>{q}, o,
                                             std::ref(copy_runner));
      CHECK(rc3.trace == r.trace && graph_tile_runs(copy) == runs && &graph_tile_runs(sg) == &graph_tile_runs(copy));
      WeightedGraph moved_from = sg;
      const WeightedGraph moved = std::move(moved_from);
      CHECK(dir.built_from(moved) && graph_region_index(moved).matches_graph(moved));
      CHECK(!dir.built_from(moved_from));  // NOLINT(bugprone-use-after-move): the refusal is the check
      CHECK(throws<std::logic_error>([&] { (void)moved_from.memo(); }));  // NOLINT(bugprone-use-after-move)
    }
    const BatchedExpansionOptions fo = options_of(EngineKind::jacobi_pull, one_ring_expansion());
    HostRunner fixture_runner{g, fo.run_options};
    CHECK(throws<std::invalid_argument>([&] {
      (void)run_batched_expansion(g, dir, f.tile_runs, f.subset({kImmediate}), fo, std::ref(fixture_runner));
    }));
  }
}

// ---------------------------------------------------------------------------
// The orchestrator over the device executor: the same outcomes as the host
// runner, in every engine, schedule and transfer mode.
int gpu_mode() {
  const Fixture f = make_fixture();
  const WeightedGraph& g = f.graph();
  hip::HipStream stream;
  hip::ResidentDeviceGraph resident;
  resident.upload_once_async(hip::make_resident_graph_plan(build_device_graph_layout32(g, f.tile_runs)), stream);
  resident.wait_until_ready(stream);
  hip::ReusableBatchedWorkspace ws(stream);
  const auto queries = f.subset({kImmediate, kTwoRing, kLongEdge, kSpill, kUnreachable, kSecondMiss, kMultiSource});

  for (EngineKind engine : {EngineKind::jacobi_pull, EngineKind::dense_chaotic_push, EngineKind::frontier_push}) {
    const std::string tag = "engine " + std::to_string(static_cast<int>(engine));
    std::fprintf(stderr, "expansion gpu: %s\n", tag.c_str());
    const GpuRunOptions run = options_of(engine, one_ring_expansion()).run_options;
    hip::BatchedExpansionExecutor status_exec(g, f.tile_runs, resident, ws, stream, run);
    check_schedule_matrix(f, engine, std::ref(status_exec), false, tag + " compact_status");
    hip::BatchedExpansionExecutor work_exec(g, f.tile_runs, resident, ws, stream, run,
                                            hip::BatchedExpansionTransferMode::status_and_work_evidence);
    check_schedule_matrix(f, engine, std::ref(work_exec), false, tag + " work evidence");

    for (const ExpansionSchedulePolicy& schedule : kSchedules) {
      BatchedExpansionOptions o = options_of(engine, schedule);
      HostRunner host{g, run};
      const auto want = run_batched_expansion(g, f.directory, f.tile_runs, queries, o, std::ref(host));
      const auto got = run_batched_expansion(g, f.directory, f.tile_runs, queries, o, std::ref(work_exec));
      CHECK_MSG(got.metrics.work_evidence == ExpansionWorkEvidence::measured &&
                    got.metrics.logical_lane_edge_work >= got.metrics.shared_edge_work &&
                    got.metrics.shared_edge_work > 0U,
                tag);
      CHECK_MSG(got.trace.size() == want.trace.size(), tag);
      for (std::size_t i = 0; i < std::min(got.trace.size(), want.trace.size()); ++i)
        CHECK_MSG(got.trace[i].valid_lane_mask == want.trace[i].valid_lane_mask &&
                      got.trace[i].reached_lane_mask == want.trace[i].reached_lane_mask &&
                      got.trace[i].miss_lane_mask == want.trace[i].miss_lane_mask &&
                      got.trace[i].query_ids_by_lane == want.trace[i].query_ids_by_lane &&
                      got.trace[i].expansion_generations_by_lane == want.trace[i].expansion_generations_by_lane,
                  tag);
      for (std::size_t i = 0; i < want.queries.size(); ++i)
        CHECK_MSG(got.queries[i].final_query == want.queries[i].final_query &&
                      got.queries[i].disposition == want.queries[i].disposition &&
                      got.queries[i].attempts == want.queries[i].attempts,
                  tag);
      CHECK_MSG(got.metrics.schedule_comparison_fingerprint == want.metrics.schedule_comparison_fingerprint, tag);

      // One device run per planning pass: the same outcomes and traces.
      hip::PassBatchedExpansionExecutor pass_exec(g, f.tile_runs, resident, ws, stream, run);
      const auto by_pass = run_batched_expansion(g, f.directory, f.tile_runs, queries, o, std::ref(pass_exec));
      CHECK_MSG(pass_exec.device_runs() == by_pass.metrics.planning_passes, tag);
      CHECK_MSG(by_pass.trace.size() == want.trace.size(), tag);
      for (std::size_t i = 0; i < std::min(by_pass.trace.size(), want.trace.size()); ++i)
        CHECK_MSG(by_pass.trace[i].reached_lane_mask == want.trace[i].reached_lane_mask &&
                      by_pass.trace[i].miss_lane_mask == want.trace[i].miss_lane_mask &&
                      by_pass.trace[i].query_ids_by_lane == want.trace[i].query_ids_by_lane,
                  tag + " pass runner");
      for (std::size_t i = 0; i < want.queries.size(); ++i)
        CHECK_MSG(by_pass.queries[i].final_query == want.queries[i].final_query &&
                      by_pass.queries[i].disposition == want.queries[i].disposition &&
                      by_pass.queries[i].attempts == want.queries[i].attempts,
                  tag + " pass runner");

      // Compact paths: payloads equal the host runner's (the fixture has no
      // ties, so the tight path is unique).
      o.enable_compact_paths = 1U;
      hip::BatchedExpansionExecutor path_exec(g, f.tile_runs, resident, ws, stream, run,
                                              hip::BatchedExpansionTransferMode::compact_paths);
      HostRunner host_paths{g, run, true};
      const auto want_p = run_batched_expansion(g, f.directory, f.tile_runs, queries, o, std::ref(host_paths));
      const auto got_p = run_batched_expansion(g, f.directory, f.tile_runs, queries, o, std::ref(path_exec));
      CHECK_MSG(got_p.metrics.compact_host_timing == CompactStageTimingEvidence::measured, tag);
      CHECK_MSG(got_p.metrics.compact_transfer == want_p.metrics.compact_transfer, tag);
      CHECK_MSG(got_p.metrics.compact_total_device_to_host_bytes >
                    got_p.metrics.compact_transfer.total_bytes,
                tag);
      hip::PassBatchedExpansionExecutor pass_paths(g, f.tile_runs, resident, ws, stream, run,
                                                   hip::BatchedExpansionTransferMode::compact_paths);
      const auto pass_p = run_batched_expansion(g, f.directory, f.tile_runs, queries, o, std::ref(pass_paths));
      CHECK_MSG(pass_p.metrics.compact_transfer == want_p.metrics.compact_transfer, tag + " pass runner");
      for (std::size_t i = 0; i < want_p.queries.size(); ++i) {
        const ExpansionQueryOutcome &x = pass_p.queries[i], &y = want_p.queries[i];
        CHECK_MSG(x.final_query == y.final_query && x.disposition == y.disposition && x.attempts == y.attempts &&
                      x.compact_paths == y.compact_paths,
                  tag + " pass runner paths of query " + std::to_string(y.final_query.query_id.value()));
      }
      // The no-congestion pipeline on the device runner: the host runner's
      // results and quality, measured device bytes.
      {
        NoCongestionPipelineOptions po;
        po.quality_sample_query_count = 4U;
        po.quality_selection_seed = 9U;
        hip::PassBatchedExpansionExecutor pipe_exec(g, f.tile_runs, resident, ws, stream, run,
                                                    hip::BatchedExpansionTransferMode::compact_paths);
        HostRunner host_pipe{g, run, true};
        const auto dev = run_no_congestion_pipeline(g, f.directory, f.tile_runs, queries, o, std::ref(pipe_exec), po);
        const auto cpu = run_no_congestion_pipeline(g, f.directory, f.tile_runs, queries, o, std::ref(host_pipe), po);
        CHECK_MSG(dev.compact_results == cpu.compact_results, tag + " pipeline");
        CHECK_MSG(dev.quality.observations == cpu.quality.observations, tag + " pipeline");
        CHECK_MSG(dev.result_accounting.device_transfer_evidence == CompactStageTimingEvidence::measured &&
                      dev.result_accounting.final_result_serialization ==
                          cpu.result_accounting.final_result_serialization,
                  tag + " pipeline");
        CHECK_MSG(validate_no_congestion_stage_ledger(dev.timing) == NoCongestionStageLedgerError::none,
                  tag + " pipeline");
      }
      for (std::size_t i = 0; i < want_p.queries.size(); ++i) {
        const ExpansionQueryOutcome &x = got_p.queries[i], &y = want_p.queries[i];
        CHECK_MSG(x.final_query == y.final_query && x.disposition == y.disposition && x.attempts == y.attempts,
                  tag + " query " + std::to_string(y.final_query.query_id.value()));
        CHECK_MSG(x.compact_paths.has_value() == y.compact_paths.has_value(), tag);
        if (x.compact_paths && y.compact_paths)
          CHECK_MSG(*x.compact_paths == *y.compact_paths,
                    tag + " paths of query " + std::to_string(y.final_query.query_id.value()));
      }
    }
  }

  // From files: the fixture written as BFGRPH07 / BFQRY007, read back,
  // rebuilt, uploaded and run gives the in-memory run's outcomes.
  {
    const auto dir = std::filesystem::temp_directory_path() / ("bfnew_expansion_" + std::to_string(::getpid()));
    std::vector<MappedRouteQuery> mapped;
    for (const RouteQuery& q : queries) mapped.push_back(MappedRouteQuery{"net", q, {}, {}});
    for (MappedRouteQuery& m : mapped) {
      for (VertexId v : m.query.source_terminals) {
        MappedTerminal t{"site", "pin", g.new_to_old()[v.value()], v, g.vertices()[v.value()].x,
                         g.vertices()[v.value()].y, g.vertices()[v.value()].has_location, g.owner_tiles()[v.value()]};
        m.source_terminals.push_back(t);
      }
      for (VertexId v : m.query.target_terminals) {
        MappedTerminal t{"site", "pin", g.new_to_old()[v.value()], v, g.vertices()[v.value()].x,
                         g.vertices()[v.value()].y, g.vertices()[v.value()].has_location, g.owner_tiles()[v.value()]};
        m.target_terminals.push_back(t);
      }
    }
    write_graph_artifact(dir / "graph.bfgrph", fixture_input());
    write_query_artifact(dir / "queries.bfqry", mapped);
    const UniformGridPartitioner grid{SpatialOrderConfig{0, 0, 10U, 10U}};
    const PartitionedGraph loaded = grid.partition(read_graph_artifact(dir / "graph.bfgrph"));
    const WeightedGraph& lg = loaded.graph;
    std::vector<RouteQuery> lq;
    for (const MappedRouteQuery& m : read_query_artifact(dir / "queries.bfqry", lg)) lq.push_back(m.query);
    std::filesystem::remove_all(dir);
    const TileRunLayout64 lruns = build_tile_run_layout(lg);
    hip::ResidentDeviceGraph lresident;
    lresident.upload_once_async(hip::make_resident_graph_plan(build_device_graph_layout32(lg, lruns)), stream);
    lresident.wait_until_ready(stream);
    BatchedExpansionOptions o = options_of(EngineKind::frontier_push, hybrid_margin_expansion(2U));
    o.enable_compact_paths = 1U;
    hip::PassBatchedExpansionExecutor from_files(lg, lruns, lresident, ws, stream, o.run_options,
                                                 hip::BatchedExpansionTransferMode::compact_paths);
    hip::PassBatchedExpansionExecutor in_memory(g, f.tile_runs, resident, ws, stream, o.run_options,
                                                hip::BatchedExpansionTransferMode::compact_paths);
    const auto a = run_batched_expansion(lg, build_tile_directory(lg), lruns, lq, o, std::ref(from_files));
    const auto b = run_batched_expansion(g, f.directory, f.tile_runs, queries, o, std::ref(in_memory));
    CHECK(lq == queries && a.queries.size() == b.queries.size());
    for (std::size_t i = 0; i < std::min(a.queries.size(), b.queries.size()); ++i)
      CHECK(a.queries[i].final_query == b.queries[i].final_query &&
            a.queries[i].disposition == b.queries[i].disposition &&
            a.queries[i].compact_paths == b.queries[i].compact_paths);
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: expansion_test host | gpu\n");
    return 2;
  }
  int rc = 0;
  try {
    if (std::strcmp(argv[1], "host") == 0) {
      host_mode();
    } else if (std::strcmp(argv[1], "gpu") == 0) {
      rc = gpu_mode();
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
