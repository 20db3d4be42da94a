// no_congestion_pipeline.cpp -- compact results, their accounting, the stage
// ledger and the path quality sample (reference behaviour:
// src/no_congestion_pipeline.cpp:62-760; the run is this library's, through
// any ExpansionBatchRunner).
#include "bfnew/no_congestion_pipeline.hpp"

#include "bfnew/graph.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <queue>
#include <stdexcept>
#include <string_view>

namespace bfnew {
namespace {

constexpr std::uint64_t kU64Max = std::numeric_limits<std::uint64_t>::max();
using Error = NoCongestionStageLedgerError;

void add(std::uint64_t& total, std::uint64_t v, const char* what) {
  if (v > kU64Max - total) throw std::overflow_error{what};
  total += v;
}

std::uint64_t times(std::uint64_t a, std::uint64_t b, const char* what) {
  if (a != 0U && b > kU64Max / a) throw std::overflow_error{what};
  return a * b;
}

bool known(PipelineTimingEvidence e) noexcept {
  return e == PipelineTimingEvidence::unavailable || e == PipelineTimingEvidence::measured;
}

Error check_stage(const PipelineStageTiming& s) noexcept {
  if (!known(s.host_evidence) || !known(s.device_evidence)) return Error::invalid_evidence;
  if (s.host_evidence == PipelineTimingEvidence::unavailable && s.host_nanoseconds != 0U)
    return Error::unavailable_host_has_value;
  if (s.device_evidence == PipelineTimingEvidence::unavailable && s.device_milliseconds != 0.0)
    return Error::unavailable_device_has_value;
  if (s.device_evidence == PipelineTimingEvidence::measured &&
      (!std::isfinite(s.device_milliseconds) || s.device_milliseconds < 0.0))
    return Error::invalid_device_duration;
  return Error::none;
}

bool all_measured(std::initializer_list<const PipelineStageTiming*> stages) noexcept {
  return std::all_of(stages.begin(), stages.end(),
                     [](const PipelineStageTiming* s) { return s->host_evidence == PipelineTimingEvidence::measured; });
}

// Sum of the stages' host nanoseconds (unmeasured stages hold zero).
bool host_sum(std::initializer_list<const PipelineStageTiming*> stages, std::uint64_t& total) noexcept {
  total = 0U;
  for (const PipelineStageTiming* s : stages) {
    if (s->host_nanoseconds > kU64Max - total) return false;
    total += s->host_nanoseconds;
  }
  return true;
}

std::uint64_t splitmix64(std::uint64_t v) noexcept {
  v += 0x9E3779B97F4A7C15ULL;
  v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ULL;
  v = (v ^ (v >> 27)) * 0x94D049BB133111EBULL;
  return v ^ (v >> 31);
}

double ratio(double bounded, double unbounded) {
  if (!std::isfinite(bounded) || !std::isfinite(unbounded) || bounded < 0.0 || unbounded < 0.0)
    throw std::invalid_argument{"path-quality ratios need finite nonnegative values"};
  if (unbounded == 0.0) return bounded == 0.0 ? 1.0 : std::numeric_limits<double>::infinity();
  return bounded / unbounded;
}

// Nearest rank: the ceil(p n / 100)-th smallest value; 0 for no values.
double nearest_rank(std::vector<double> v, std::uint32_t p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const std::uint64_t rank = (static_cast<std::uint64_t>(p) * v.size() + 99U) / 100U;
  return v[static_cast<std::size_t>(rank - 1U)];
}

// fp32 multi-source Dijkstra over the whole graph.
std::vector<float> unbounded_labels(const WeightedGraph& g, std::span<const VertexId> sources) {
  const auto out = g.outgoing();
  std::vector<float> d(g.vertex_count(), std::numeric_limits<float>::infinity());
  using Item = std::pair<float, std::uint32_t>;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
  for (VertexId s : sources) {
    d[s.value()] = 0.0F;
    heap.push({0.0F, s.value()});
  }
  while (!heap.empty()) {
    const auto [du, u] = heap.top();
    heap.pop();
    if (du > d[u]) continue;
    for (std::uint64_t p = out.row_offsets[u]; p < out.row_offsets[u + 1]; ++p) {
      const std::uint32_t v = out.destinations[p].value();
      const float nd = du + out.weights[p];
      if (nd < d[v]) {
        d[v] = nd;
        heap.push({nd, v});
      }
    }
  }
  return d;
}

// Edge count of the tight path to `target`: at each vertex the tight
// in-edges (label[src] + w == label[v]) in stable EdgeId order, skipping
// sources already on the path, backtracking when none is left; ends at a
// query source.  Returns false when no such path exists.
bool tight_path_length(const WeightedGraph& g, std::span<const float> d, std::span<const VertexId> sources,
                       VertexId target, std::uint32_t& length) {
  const auto in = g.incoming();
  std::vector<std::uint8_t> is_source(g.vertex_count(), 0U), on_path(g.vertex_count(), 0U);
  for (VertexId s : sources) is_source[s.value()] = 1U;
  struct Frame {
    std::uint32_t vertex;
    std::vector<std::pair<std::uint64_t, std::uint32_t>> options;  // (edge id, source)
    std::size_t next = 0;
  };
  const auto frame_of = [&](std::uint32_t v) {
    Frame f{v, {}};
    for (std::uint64_t p = in.column_offsets[v]; p < in.column_offsets[v + 1]; ++p) {
      const std::uint32_t u = in.sources[p].value();
      if (std::isfinite(d[u]) && d[u] + in.weights[p] == d[v]) f.options.push_back({in.edge_ids[p].value(), u});
    }
    std::sort(f.options.begin(), f.options.end());
    return f;
  };
  std::vector<Frame> stack;
  stack.push_back(frame_of(target.value()));
  on_path[target.value()] = 1U;
  while (!stack.empty()) {
    Frame& top = stack.back();
    if (is_source[top.vertex]) {
      length = static_cast<std::uint32_t>(stack.size() - 1U);
      return true;
    }
    if (top.next == top.options.size()) {
      on_path[top.vertex] = 0U;
      stack.pop_back();
      continue;
    }
    const std::uint32_t u = top.options[top.next++].second;
    if (on_path[u]) continue;
    on_path[u] = 1U;
    stack.push_back(frame_of(u));
  }
  return false;
}

}  // namespace

NoCongestionStageLedger make_no_congestion_stage_ledger(
    PipelineStageTiming cold_artifact_load, PipelineStageTiming graph_upload, PipelineStageTiming batch_planning,
    PipelineStageTiming sssp, PipelineStageTiming expansion, PipelineStageTiming controller_orchestration,
    PipelineStageTiming reconstruction, PipelineStageTiming result_transfer, PipelineStageTiming warm_all_query,
    PipelineStageTiming cold_execution) {
  NoCongestionStageLedger l{cold_artifact_load, graph_upload,   batch_planning, sssp,
                            expansion,          controller_orchestration, reconstruction, result_transfer,
                            warm_all_query,     cold_execution, {}};
  if (all_measured({&l.cold_artifact_load, &l.graph_upload, &l.cold_execution})) {
    l.cold_pipeline.host_evidence = PipelineTimingEvidence::measured;
    if (!host_sum({&l.cold_artifact_load, &l.graph_upload, &l.cold_execution}, l.cold_pipeline.host_nanoseconds))
      throw std::overflow_error{"cold pipeline timing sum overflow"};
  }
  if (validate_no_congestion_stage_ledger(l) != Error::none)
    throw std::invalid_argument{"invalid no-congestion stage timing input"};
  return l;
}

NoCongestionStageLedgerError validate_no_congestion_stage_ledger(const NoCongestionStageLedger& l) noexcept {
  for (const PipelineStageTiming* s :
       {&l.cold_artifact_load, &l.graph_upload, &l.batch_planning, &l.sssp, &l.expansion,
        &l.controller_orchestration, &l.reconstruction, &l.result_transfer, &l.warm_all_query, &l.cold_execution,
        &l.cold_pipeline})
    if (const Error e = check_stage(*s); e != Error::none) return e;
  for (const PipelineStageTiming* s : {&l.cold_artifact_load, &l.batch_planning, &l.expansion,
                                       &l.controller_orchestration, &l.warm_all_query, &l.cold_execution,
                                       &l.cold_pipeline})
    if (s->device_evidence != PipelineTimingEvidence::unavailable) return Error::cpu_only_device_evidence;
  std::uint64_t warm = 0;
  if (!host_sum({&l.batch_planning, &l.sssp, &l.expansion, &l.controller_orchestration, &l.reconstruction,
                 &l.result_transfer},
                warm))
    return Error::warm_sum_overflow;
  const bool warm_parts = all_measured({&l.batch_planning, &l.sssp, &l.expansion, &l.controller_orchestration,
                                        &l.reconstruction, &l.result_transfer});
  if (warm_parts) {
    if (l.warm_all_query.host_evidence != PipelineTimingEvidence::measured) return Error::warm_evidence_mismatch;
    if (l.warm_all_query.host_nanoseconds != warm) return Error::warm_sum_mismatch;
  } else if (l.warm_all_query.host_evidence == PipelineTimingEvidence::measured &&
             warm > l.warm_all_query.host_nanoseconds) {
    return Error::warm_sum_mismatch;
  }
  const bool cold_parts = all_measured({&l.cold_artifact_load, &l.graph_upload, &l.cold_execution});
  if ((l.cold_pipeline.host_evidence == PipelineTimingEvidence::measured) != cold_parts)
    return Error::cold_evidence_mismatch;
  if (cold_parts) {
    std::uint64_t cold = 0;
    if (!host_sum({&l.cold_artifact_load, &l.graph_upload, &l.cold_execution}, cold)) return Error::cold_sum_overflow;
    if (l.cold_pipeline.host_nanoseconds != cold) return Error::cold_sum_mismatch;
  }
  return Error::none;
}

NoCongestionResultAccounting measure_no_congestion_result_transfer(std::span<const CompactQueryResult> results,
                                                                   std::uint64_t modeled_batch_status_count,
                                                                   const BatchedExpansionMetrics* m) {
  NoCongestionResultAccounting a;
  a.query_count = results.size();
  a.modeled_batch_status_count = modeled_batch_status_count;
  a.modeled_batch_status_bytes = times(modeled_batch_status_count, sizeof(DeviceRunStatus), "status bytes overflow");
  std::uint64_t vertices = 0, edges = 0;
  for (std::size_t i = 0; i < results.size(); ++i) {
    const CompactQueryResult& r = results[i];
    if (i != 0U && !(results[i - 1].query_id < r.query_id))
      throw std::invalid_argument{"compact result accounting needs strictly increasing query ids"};
    if (r.targets.empty()) throw std::invalid_argument{"compact result accounting needs at least one target"};
    add(a.target_summary_count, r.targets.size(), "target count overflow");
    bool all_complete = true, any_unreachable = false, all_classified = true, any_failure = false,
         all_terminal = true;
    for (const CompactTargetPath& t : r.targets) {
      const CompactTargetSummary& s = t.summary;
      const bool no_arenas = t.vertices.empty() && t.distance_labels.empty() && t.edge_ids.empty();
      const bool no_path = s.has_selected_source == 0U && s.path_length == 0U && no_arenas;
      switch (s.reconstruction) {
        case CompactPathStatus::complete:
          all_terminal = false;
          if (s.reached != CompactTargetReachStatus::reached || s.has_selected_source != 1U ||
              !std::isfinite(s.distance) || s.distance < 0.0F || t.vertices.size() != t.distance_labels.size() ||
              t.vertices.size() != t.edge_ids.size() + 1U || t.edge_ids.size() != s.path_length)
            throw std::invalid_argument{"complete compact path has an inconsistent arena shape"};
          add(a.complete_path_count, 1U, "path count overflow");
          add(vertices, t.vertices.size(), "path vertex count overflow");
          add(edges, t.edge_ids.size(), "path edge count overflow");
          break;
        case CompactPathStatus::unreachable:
          all_complete = false, any_unreachable = true, all_terminal = false;
          if (s.reached != CompactTargetReachStatus::not_reached || !std::isinf(s.distance) || s.distance < 0.0F ||
              !no_path)
            throw std::invalid_argument{"unreachable compact summary has inconsistent fields"};
          add(a.unreachable_target_count, 1U, "target count overflow");
          break;
        case CompactPathStatus::query_terminal_failure:
          all_complete = false, all_classified = false, any_failure = true;
          if (s.reached != CompactTargetReachStatus::not_reached || !std::isinf(s.distance) || s.distance < 0.0F ||
              !no_path)
            throw std::invalid_argument{"terminal-failure compact summary has inconsistent fields"};
          add(a.terminal_failure_target_count, 1U, "target count overflow");
          break;
        case CompactPathStatus::no_tight_path:
        case CompactPathStatus::path_length_overflow:
          all_complete = false, all_classified = false, any_failure = true, all_terminal = false;
          if (s.reached != CompactTargetReachStatus::reached || !std::isfinite(s.distance) || s.distance < 0.0F ||
              !no_path)
            throw std::invalid_argument{"reconstruction-failure compact summary has inconsistent fields"};
          add(a.reconstruction_failure_target_count, 1U, "target count overflow");
          break;
        default:
          throw std::invalid_argument{"compact target has an unknown reconstruction status"};
      }
    }
    switch (r.disposition) {
      case ExpansionQueryDisposition::reached:
        if (!all_complete) throw std::invalid_argument{"reached compact result has an incomplete target"};
        break;
      case ExpansionQueryDisposition::unreachable_in_full_region:
        if (!any_unreachable || !all_classified)
          throw std::invalid_argument{"full-region-unreachable result has inconsistent targets"};
        break;
      case ExpansionQueryDisposition::engine_failure:
        if (!any_failure) throw std::invalid_argument{"engine-failure result has no failed target"};
        break;
      case ExpansionQueryDisposition::expansion_limit:
      case ExpansionQueryDisposition::region_stalled:
      case ExpansionQueryDisposition::identity_or_count_overflow:
        if (!all_terminal) throw std::invalid_argument{"terminal query failure has a non-failure target"};
        break;
      default:
        throw std::invalid_argument{"compact result has an unknown disposition"};
    }
  }
  CompactTransferAccounting& f = a.final_result_serialization;
  f.summary_bytes = times(a.target_summary_count, sizeof(CompactTargetSummary), "summary bytes overflow");
  f.vertex_bytes = times(vertices, 4U, "vertex bytes overflow");
  f.distance_label_bytes = times(vertices, 4U, "label bytes overflow");
  f.edge_id_bytes = times(edges, 4U, "edge id bytes overflow");
  for (std::uint64_t part : {f.summary_bytes, f.vertex_bytes, f.distance_label_bytes, f.edge_id_bytes})
    add(f.total_bytes, part, "result bytes overflow");
  a.modeled_final_transfer_bytes = a.modeled_batch_status_bytes;
  add(a.modeled_final_transfer_bytes, f.total_bytes, "result bytes overflow");

  if (m != nullptr && m->compact_total_device_to_host_bytes != 0U) {
    const CompactTransferAccounting& t = m->compact_transfer;
    std::uint64_t parts = 0;
    for (std::uint64_t part : {t.summary_bytes, t.vertex_bytes, t.distance_label_bytes, t.edge_id_bytes})
      add(parts, part, "device bytes overflow");
    if (parts != t.total_bytes) throw std::invalid_argument{"device compact payload byte parts are inconsistent"};
    if (m->compact_status_bytes != times(modeled_batch_status_count, sizeof(DeviceRunStatus), "status bytes"))
      throw std::invalid_argument{"device status bytes disagree with the batch count"};
    std::uint64_t total = t.total_bytes;
    add(total, m->compact_status_bytes, "device bytes overflow");
    add(total, m->compact_error_bytes, "device bytes overflow");
    if (total != m->compact_total_device_to_host_bytes)
      throw std::invalid_argument{"device compact transfer records are inconsistent"};
    if ((m->compact_controller_poll_count == 0U) != (m->compact_controller_poll_bytes == 0U))
      throw std::invalid_argument{"device controller polls and their bytes disagree"};
    add(total, m->compact_controller_poll_bytes, "device bytes overflow");
    if (total != m->compact_overall_device_to_host_bytes)
      throw std::invalid_argument{"device overall transfer record is inconsistent"};
    a.device_transfer_evidence = CompactStageTimingEvidence::measured;
    a.actual_compact_device_transfer = t;
    a.actual_status_bytes = m->compact_status_bytes;
    a.actual_error_bytes = m->compact_error_bytes;
    a.actual_compact_total_device_to_host_bytes = m->compact_total_device_to_host_bytes;
    a.actual_controller_poll_count = m->compact_controller_poll_count;
    a.actual_controller_poll_bytes = m->compact_controller_poll_bytes;
    a.actual_overall_device_to_host_bytes = m->compact_overall_device_to_host_bytes;
  } else if (m != nullptr && (m->compact_status_bytes || m->compact_error_bytes || m->compact_controller_poll_count ||
                              m->compact_controller_poll_bytes || m->compact_overall_device_to_host_bytes)) {
    throw std::invalid_argument{"device transfer records without a compact total"};
  }
  return a;
}

CompactPathQualitySample sample_compact_path_quality(const WeightedGraph& graph,
                                                     std::span<const ExpansionQueryOutcome> outcomes,
                                                     std::span<const CompactQueryResult> results,
                                                     std::size_t sample_query_count, std::uint64_t seed) {
  if (const char* why = validate_weighted_graph_message(graph); (why && !std::string_view{why}.empty()) ||
                                                                outcomes.size() != results.size())
    throw std::invalid_argument{"path quality needs a valid graph and one result per outcome"};
  struct Candidate {
    std::uint64_t key;
    QueryId id;
    std::size_t index;
  };
  std::vector<Candidate> candidates;
  for (std::size_t i = 0; i < outcomes.size(); ++i) {
    const ExpansionQueryOutcome& o = outcomes[i];
    const CompactQueryResult& r = results[i];
    if (o.final_query.query_id != r.query_id || o.final_query.expansion_generation != r.expansion_generation ||
        o.disposition != r.disposition || !validate_compact_query_result(graph, o.final_query, r).ok())
      throw std::invalid_argument{"path quality: an outcome and its result disagree"};
    candidates.push_back({splitmix64(seed ^ r.query_id.value()), r.query_id, i});
  }
  std::sort(candidates.begin(), candidates.end(), [](const Candidate& a, const Candidate& b) {
    return a.key != b.key ? a.key < b.key : a.id < b.id;
  });
  for (std::size_t i = 1; i < candidates.size(); ++i)
    if (candidates[i].id == candidates[i - 1].id) throw std::invalid_argument{"path quality: duplicate query id"};

  CompactPathQualitySample q;
  q.selection_seed = seed;
  q.population_query_count = outcomes.size();
  q.requested_sample_query_count = sample_query_count;
  const std::size_t n = std::min(sample_query_count, candidates.size());
  q.sampled_query_count = n;
  std::vector<double> cost_ratio, length_ratio, cost_gap, length_gap;
  for (std::size_t k = 0; k < n; ++k) {
    const ExpansionQueryOutcome& o = outcomes[candidates[k].index];
    const CompactQueryResult& r = results[candidates[k].index];
    q.sampled_query_ids.push_back(candidates[k].id);
    const std::vector<float> d = unbounded_labels(graph, o.final_query.sources);
    for (const CompactTargetPath& t : r.targets) {
      if (t.summary.reconstruction != CompactPathStatus::complete) continue;
      const float ud = d[t.summary.target.value()];
      std::uint32_t ulen = 0;
      if (!std::isfinite(ud) || !tight_path_length(graph, d, o.final_query.sources, t.summary.target, ulen))
        throw std::logic_error{"path quality: a bounded path has no unbounded counterpart"};
      CompactPathQualityObservation ob;
      ob.query_id = r.query_id;
      ob.target = t.summary.target;
      ob.bounded_distance = t.summary.distance;
      ob.unbounded_distance = ud;
      ob.bounded_path_length = t.summary.path_length;
      ob.unbounded_path_length = ulen;
      ob.absolute_cost_inflation = static_cast<double>(t.summary.distance) - static_cast<double>(ud);
      ob.absolute_path_length_inflation =
          static_cast<std::int64_t>(t.summary.path_length) - static_cast<std::int64_t>(ulen);
      ob.cost_inflation_ratio = ratio(t.summary.distance, ud);
      ob.path_length_inflation_ratio = ratio(t.summary.path_length, ulen);
      q.observations.push_back(ob);
      cost_ratio.push_back(ob.cost_inflation_ratio);
      length_ratio.push_back(ob.path_length_inflation_ratio);
      cost_gap.push_back(ob.absolute_cost_inflation);
      length_gap.push_back(static_cast<double>(ob.absolute_path_length_inflation));
    }
  }
  q.finite_target_pairs = q.observations.size();
  const auto fill = [](const std::vector<double>& v, double& p50, double& p95, double& p99, double& max) {
    p50 = nearest_rank(v, 50);
    p95 = nearest_rank(v, 95);
    p99 = nearest_rank(v, 99);
    max = nearest_rank(v, 100);
  };
  fill(cost_gap, q.absolute_cost_inflation_p50, q.absolute_cost_inflation_p95, q.absolute_cost_inflation_p99,
       q.absolute_cost_inflation_max);
  fill(cost_ratio, q.cost_ratio_p50, q.cost_ratio_p95, q.cost_ratio_p99, q.cost_ratio_max);
  fill(length_gap, q.absolute_path_length_inflation_p50, q.absolute_path_length_inflation_p95,
       q.absolute_path_length_inflation_p99, q.absolute_path_length_inflation_max);
  fill(length_ratio, q.path_length_ratio_p50, q.path_length_ratio_p95, q.path_length_ratio_p99,
       q.path_length_ratio_max);
  return q;
}

NoCongestionPipelineResult run_no_congestion_pipeline(const WeightedGraph& graph, const TileDirectory& directory,
                                                      const TileRunLayout64& tile_runs,
                                                      std::span<const RouteQuery> queries,
                                                      const BatchedExpansionOptions& options,
                                                      const ExpansionBatchRunner& runner,
                                                      const NoCongestionPipelineOptions& po) {
  if (options.enable_compact_paths != 1U)
    throw std::invalid_argument{"the no-congestion pipeline runs in compact-path mode"};
  NoCongestionPipelineResult out;
  out.expansion = run_batched_expansion(graph, directory, tile_runs, queries, options, runner);
  for (ExpansionQueryOutcome& o : out.expansion.queries) {
    if (o.compact_paths)
      out.compact_results.push_back(make_compact_query_result(*o.compact_paths, o.disposition));
    else
      out.compact_results.push_back(make_failed_compact_query_result(o.final_query, o.disposition));
  }
  const BatchedExpansionMetrics& m = out.expansion.metrics;
  out.result_accounting = measure_no_congestion_result_transfer(out.compact_results, m.batches_executed, &m);

  // Stages from the run's own records; orchestration is the remainder of the
  // run's wall time (validation, descriptions, polling, retries, outcomes).
  const bool host_stages = m.compact_host_timing == CompactStageTimingEvidence::measured;
  const PipelineStageTiming planning = measured_host_stage(m.initial_planning_nanoseconds + m.replanning_nanoseconds);
  const PipelineStageTiming expansion = measured_host_stage(m.geometric_expansion_host_nanoseconds);
  PipelineStageTiming sssp{}, reconstruction{}, transfer{}, orchestration{}, warm{};
  warm = measured_host_stage(m.total_nanoseconds);
  if (host_stages) {
    sssp = measured_host_stage(m.sssp_host_nanoseconds);
    reconstruction = measured_host_stage(m.reconstruction_host_nanoseconds);
    transfer = measured_host_stage(m.result_transfer_host_nanoseconds);
    if (m.compact_device_timing == CompactStageTimingEvidence::measured) {
      sssp.device_evidence = reconstruction.device_evidence = transfer.device_evidence =
          PipelineTimingEvidence::measured;
      sssp.device_milliseconds = m.sssp_device_nanoseconds / 1e6;
      reconstruction.device_milliseconds = m.reconstruction_device_nanoseconds / 1e6;
      transfer.device_milliseconds = m.result_transfer_device_nanoseconds / 1e6;
    }
    const unsigned __int128 parts = static_cast<unsigned __int128>(planning.host_nanoseconds) +
                                    expansion.host_nanoseconds + sssp.host_nanoseconds +
                                    reconstruction.host_nanoseconds + transfer.host_nanoseconds;
    if (parts <= m.total_nanoseconds)
      orchestration = measured_host_stage(m.total_nanoseconds - static_cast<std::uint64_t>(parts));
    else
      warm = PipelineStageTiming{};  // stages measured on overlapping clocks: no enclosing interval
  }
  out.timing = make_no_congestion_stage_ledger(po.cold_artifact_load, po.graph_upload, planning, sssp, expansion,
                                               orchestration, reconstruction, transfer, warm, po.cold_execution);
  out.quality = sample_compact_path_quality(graph, out.expansion.queries, out.compact_results,
                                            po.quality_sample_query_count, po.quality_selection_seed);
  return out;
}

}  // namespace bfnew
