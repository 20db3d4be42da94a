// expansion_bench.cpp -- times run_batched_expansion over the device
// executor on a synthetic tile grid, once per expansion schedule, and prints
// one JSON object (read by bench.py --expansion).
//
// Each query keeps its synthetic source and target but starts from its two
// terminal tiles only (--start terminals, the default) or from the terminal
// rectangle (--start rectangle, make_route_query with padding 0).  Routes
// whose region holds no path miss, expand around the terminals' box and are
// replanned: the workload of the reference's all-query expansion runs
// (tools/gpu_batched_expansion.hip.cpp).
//
//   bfnew_expansion_bench [--tiles X Y N] [--queries Q] [--seed S]
//                         [--engine E] [--control C] [--lane-width W]
//                         [--sweeps S] [--boundary B] [--graph G]
//                         [--graph-rounds R] [--claim K] [--min-waves M]
//                         [--mode compact_paths|compact_status] [--reps R]
//                         [--start terminals|rectangle] [--runner batch|pass]
//
// --runner batch: hip::BatchedExpansionExecutor, one device run per planned
// lane batch; --runner pass: hip::PassBatchedExpansionExecutor, one device run
// per planning pass.
#include "bfnew/batched_expansion.hpp"
#include "bfnew/device_layout.hpp"
#include "bfnew/hip/batched_expansion.hpp"
#include "bfnew/hip/runtime.hpp"
#include "bfnew/query.hpp"
#include "bfnew/spatial.hpp"
#include "bfnew/synthetic.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

using namespace bfnew;

namespace {

struct Args {
  std::uint32_t tx = 64, ty = 64, npt = 256, queries = 256, lane_width = 32, reps = 2;
  std::uint64_t seed = 2, query_seed = 71;
  std::uint32_t engine = 2, control = 1, sweeps = 0, boundary = 1, use_graph = 1, graph_rounds = 8,
                claim = 2, min_waves = 0;
  bool paths = true;
  bool terminal_tiles_only = true;
  bool per_pass = false;
};

Args parse(int argc, char** argv) {
  Args a;
  for (int i = 1; i < argc; ++i) {
    const std::string k = argv[i];
    const auto next = [&]() -> std::uint64_t {
      if (i + 1 >= argc) throw std::invalid_argument{"missing value for " + k};
      return std::strtoull(argv[++i], nullptr, 10);
    };
    if (k == "--tiles") {
      a.tx = static_cast<std::uint32_t>(next());
      a.ty = static_cast<std::uint32_t>(next());
      a.npt = static_cast<std::uint32_t>(next());
    } else if (k == "--queries") a.queries = static_cast<std::uint32_t>(next());
    else if (k == "--seed") a.seed = next();
    else if (k == "--query-seed") a.query_seed = next();
    else if (k == "--engine") a.engine = static_cast<std::uint32_t>(next());
    else if (k == "--control") a.control = static_cast<std::uint32_t>(next());
    else if (k == "--lane-width") a.lane_width = static_cast<std::uint32_t>(next());
    else if (k == "--sweeps") a.sweeps = static_cast<std::uint32_t>(next());
    else if (k == "--boundary") a.boundary = static_cast<std::uint32_t>(next());
    else if (k == "--graph") a.use_graph = static_cast<std::uint32_t>(next());
    else if (k == "--graph-rounds") a.graph_rounds = static_cast<std::uint32_t>(next());
    else if (k == "--claim") a.claim = static_cast<std::uint32_t>(next());
    else if (k == "--min-waves") a.min_waves = static_cast<std::uint32_t>(next());
    else if (k == "--reps") a.reps = static_cast<std::uint32_t>(next());
    else if (k == "--mode") {
      if (i + 1 >= argc) throw std::invalid_argument{"missing value for --mode"};
      a.paths = std::strcmp(argv[++i], "compact_status") != 0;
    } else if (k == "--runner") {
      if (i + 1 >= argc) throw std::invalid_argument{"missing value for --runner"};
      a.per_pass = std::strcmp(argv[++i], "pass") == 0;
    } else if (k == "--start") {
      if (i + 1 >= argc) throw std::invalid_argument{"missing value for --start"};
      a.terminal_tiles_only = std::strcmp(argv[++i], "rectangle") != 0;
    } else {
      throw std::invalid_argument{"unknown argument " + k};
    }
  }
  return a;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

const char* schedule_name(ExpansionScheduleKind k) {
  switch (k) {
    case ExpansionScheduleKind::one_geometric_ring: return "one_ring";
    case ExpansionScheduleKind::fixed_larger_ring: return "fixed_ring_2";
    case ExpansionScheduleKind::doubling_xy_margins: return "doubling";
    case ExpansionScheduleKind::hybrid_small_then_doubling: return "hybrid_2";
    default: return "unspecified";
  }
}

}  // namespace

int main(int argc, char** argv) {
  try {
    const Args a = parse(argc, argv);
    const auto t_build = std::chrono::steady_clock::now();
    SyntheticTileGridConfig cfg;
    cfg.tiles_x = a.tx;
    cfg.tiles_y = a.ty;
    cfg.nodes_per_tile = a.npt;
    cfg.seed = a.seed;
    const SyntheticTileGraph gen = SyntheticTileGraph::generate(cfg);
    const WeightedGraph g = gen.weighted_graph();
    // The graph's memoised runs: the orchestrator's per-call check that the
    // runs are the graph's is then an identity test.
    const TileRunLayout64& runs = graph_tile_runs(g);
    const TileDirectory dir = build_tile_directory(g);
    SyntheticQueryConfig qc;
    qc.count = a.queries;
    qc.seed = a.query_seed;
    qc.side_min = 4;
    qc.side_max = 32;
    std::vector<RouteQuery> queries;
    for (const SyntheticQuery& sq : generate_synthetic_queries(cfg, qc)) {
      const std::vector<VertexId> s{VertexId{sq.source}}, t{VertexId{sq.target}};
      RouteQuery q = make_route_query(QueryId{sq.query_id}, g, s, t, 0U);
      if (a.terminal_tiles_only) {
        q.selected_tiles = {g.owner_tiles()[sq.source], g.owner_tiles()[sq.target]};
        std::sort(q.selected_tiles.begin(), q.selected_tiles.end());
        q.selected_tiles.erase(std::unique(q.selected_tiles.begin(), q.selected_tiles.end()),
                               q.selected_tiles.end());
      }
      if (!validate_route_query(g, q).ok()) throw std::logic_error{"invalid synthetic route query"};
      queries.push_back(std::move(q));
    }
    const double build_ms = ms_since(t_build);

    hip::HipStream stream;
    hip::ResidentDeviceGraph resident;
    resident.upload_once_async(hip::make_resident_graph_plan(gen.layout()), stream);
    resident.wait_until_ready(stream);
    hip::ReusableBatchedWorkspace ws(stream);

    BatchedExpansionOptions o;
    o.run_options.engine = static_cast<EngineKind>(a.engine);
    o.run_options.control_mode = static_cast<ControlMode>(a.control);
    o.run_options.rounds_per_chunk = 8U;
    o.run_options.block_size = 256U;
    o.run_options.maximum_rounds = 1U << 20;
    o.planner_policy.lane_width = a.lane_width;
    o.execution_configuration_fingerprint =
        0xE0000000ULL | (a.paths ? 1U : 0U) | a.sweeps << 1 | a.boundary << 2 | a.use_graph << 3;
    o.maximum_expansions = 4U;
    o.terminal_policy = ExpansionTerminalPolicy::full_region_fallback;
    o.enable_compact_paths = a.paths ? 1U : 0U;
    const auto mode = a.paths ? hip::BatchedExpansionTransferMode::compact_paths
                              : hip::BatchedExpansionTransferMode::compact_status;
    hip::BatchedExpansionExecutor batch_exec(g, runs, resident, ws, stream, o.run_options, mode);
    hip::PassBatchedExpansionExecutor pass_exec(g, runs, resident, ws, stream, o.run_options, mode);
    bfn_tuning tune;
    bfn_tuning_default(&tune);
    tune.tile_sweeps = a.sweeps;
    tune.boundary_mode = a.boundary;
    tune.use_graph = a.use_graph;
    tune.graph_rounds = a.graph_rounds;
    tune.items_per_claim = a.claim;
    tune.min_waves = a.min_waves;
    batch_exec.set_tuning(tune);
    pass_exec.set_tuning(tune);
    const ExpansionBatchRunner runner =
        a.per_pass ? ExpansionBatchRunner{std::ref(pass_exec)} : ExpansionBatchRunner{std::ref(batch_exec)};

    const std::array<ExpansionSchedulePolicy, 4> schedules{one_ring_expansion(), fixed_ring_expansion(2U),
                                                           doubling_margin_expansion(),
                                                           hybrid_margin_expansion(2U)};
    std::array<ExpansionScheduleEvidence, 4> evidence{};
    std::string rows;
    // One untimed run first: device allocations and code objects warm up.
    o.schedule = schedules[0];
    (void)run_batched_expansion(g, dir, runs, queries, o, runner);
    for (std::size_t i = 0; i < schedules.size(); ++i) {
      o.schedule = schedules[i];
      BatchedExpansionRunResult best;
      double best_ms = 0.0;
      for (std::uint32_t r = 0; r < std::max(1U, a.reps); ++r) {
        const auto t0 = std::chrono::steady_clock::now();
        BatchedExpansionRunResult run = run_batched_expansion(g, dir, runs, queries, o, runner);
        const double ms = ms_since(t0);
        if (r == 0 || ms < best_ms) {
          best_ms = ms;
          best = std::move(run);
        }
      }
      const BatchedExpansionMetrics& m = best.metrics;
      evidence[i] = make_expansion_schedule_evidence(schedules[i], best);
      std::uint64_t path_vertices = 0;
      for (const ExpansionQueryOutcome& q : best.queries)
        if (q.compact_paths)
          for (const CompactTargetPath& t : q.compact_paths->targets) path_vertices += t.vertices.size();
      char row[2048];
      std::snprintf(
          row, sizeof row,
          "%s{\"schedule\": \"%s\", \"queries_per_s\": %.1f, \"wall_ms\": %.3f, \"execution_ms\": %.3f, "
          "\"planning_ms\": %.3f, \"expansion_host_ms\": %.3f, \"batches\": %llu, \"planning_passes\": %llu, "
          "\"reached\": %llu, \"unreachable_full_region\": %llu, \"other_failures\": %llu, "
          "\"initial_reached\": %llu, \"scheduled_expansions\": %llu, \"full_region_fallbacks\": %llu, "
          "\"attempted_selected_edge_estimate\": %llu, \"final_selected_tiles\": %llu, "
          "\"compact_d2h_bytes\": %llu, \"path_vertices\": %llu, \"histogram\": [",
          rows.empty() ? "" : ", ", schedule_name(schedules[i].kind),
          m.all_query_throughput_milliqueries_per_second / 1000.0, best_ms, m.execution_nanoseconds / 1e6,
          (m.initial_planning_nanoseconds + m.replanning_nanoseconds) / 1e6,
          m.geometric_expansion_host_nanoseconds / 1e6, static_cast<unsigned long long>(m.batches_executed),
          static_cast<unsigned long long>(m.planning_passes), static_cast<unsigned long long>(m.reached_queries),
          static_cast<unsigned long long>(m.unreachable_full_region_queries),
          static_cast<unsigned long long>(m.input_queries - m.reached_queries - m.unreachable_full_region_queries),
          static_cast<unsigned long long>(m.initial_reached_queries),
          static_cast<unsigned long long>(m.scheduled_expansions),
          static_cast<unsigned long long>(m.full_region_fallbacks),
          static_cast<unsigned long long>(m.attempted_selected_edge_estimate),
          static_cast<unsigned long long>(m.final_selected_tile_count),
          static_cast<unsigned long long>(m.compact_overall_device_to_host_bytes),
          static_cast<unsigned long long>(path_vertices));
      rows += row;
      for (std::size_t h = 0; h < m.expansion_count_histogram.size(); ++h)
        rows += (h ? ", " : "") + std::to_string(m.expansion_count_histogram[h]);
      rows += "]}";
      std::fprintf(stderr, "[expansion] %s done in %.1f ms\n", schedule_name(schedules[i].kind), best_ms);
    }
    const auto chosen = select_expansion_schedule_from_evidence(evidence);
    std::printf(
        "{\"workload\": \"expansion_%ux%ux%u_q%u\", \"runner\": \"%s\", \"start\": \"%s\", \"lane_width\": %u, \"engine\": %u, "
        "\"control\": %u, \"mode\": \"%s\", \"maximum_expansions\": 4, \"terminal_policy\": "
        "\"full_region_fallback\", \"host_build_ms\": %.1f, \"selected_schedule\": \"%s\", \"schedules\": [%s]}\n",
        a.tx, a.ty, a.npt, a.queries, a.per_pass ? "one device run per planning pass" : "one device run per lane batch",
        a.terminal_tiles_only ? "terminal_tiles" : "terminal_rectangle", a.lane_width, a.engine, a.control, a.paths ? "compact_paths" : "compact_status",
        build_ms, chosen ? schedule_name(chosen->kind) : "none (tie or incomplete evidence)", rows.c_str());
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "expansion bench: %s\n", e.what());
    return 3;
  }
}
