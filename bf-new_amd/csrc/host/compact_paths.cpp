// compact_paths.cpp -- host rules for compact path payloads (reference
// behaviour: src/compact_paths.cpp:269-470 payload checks, :527-538
// completeness, :683-728 transfer bytes, :736-870 disposition and
// label-image checks).
//
// Edges of a path are found by scanning the CSR row of the path's previous
// vertex for the stable edge id, so a payload check costs O(path x degree)
// rather than an O(E) edge table per call.
#include "bfnew/compact_paths.hpp"

#include "bfnew/graph.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string_view>

namespace bfnew {
namespace {

using Code = CompactPathValidationErrorCode;

CompactPathValidationResult fail(Code code, std::size_t target = CompactPathValidationResult::no_position,
                                 std::size_t position = CompactPathValidationResult::no_position) {
  return CompactPathValidationResult{code, target, position};
}

bool known_reach(CompactTargetReachStatus s) noexcept {
  return s == CompactTargetReachStatus::not_reached || s == CompactTargetReachStatus::reached;
}

bool known_path_status(CompactPathStatus s) noexcept {
  return static_cast<std::uint32_t>(s) <= static_cast<std::uint32_t>(CompactPathStatus::path_length_overflow);
}

bool known_disposition(ExpansionQueryDisposition d) noexcept {
  return static_cast<std::uint32_t>(d) <= static_cast<std::uint32_t>(ExpansionQueryDisposition::engine_failure);
}

// Labels are +inf or finite non-negative.
bool label_value_ok(float v) noexcept { return !std::isnan(v) && v >= 0.0F; }

bool reconstruction_failed(CompactPathStatus s) noexcept {
  return s == CompactPathStatus::query_terminal_failure || s == CompactPathStatus::no_tight_path ||
         s == CompactPathStatus::path_length_overflow;
}

void add_checked(std::uint64_t& total, std::uint64_t value) {
  if (value > ~std::uint64_t{0} - total) throw std::overflow_error{"compact transfer bytes overflow"};
  total += value;
}

std::uint64_t bytes_of(std::size_t count, std::uint64_t width) {
  if (count > ~std::uint64_t{0} / width) throw std::overflow_error{"compact transfer bytes overflow"};
  return static_cast<std::uint64_t>(count) * width;
}

// A complete path: shape, termination, simplicity, region, continuity,
// tightness and cost.  `seen` is all-false on entry and on return.
CompactPathValidationResult check_complete_path(const WeightedGraph& graph, const RouteQuery& query,
                                                const CompactTargetPath& path, std::size_t t,
                                                std::vector<std::uint8_t>& seen) {
  const CompactTargetSummary& s = path.summary;
  const std::uint32_t V = graph.vertex_count();
  if (s.reached != CompactTargetReachStatus::reached || !std::isfinite(s.distance) ||
      s.has_selected_source != 1U)
    return fail(Code::path_shape_mismatch, t);
  if (s.selected_source.value() >= V) return fail(Code::selected_source_out_of_range, t);
  if (!std::binary_search(query.sources.begin(), query.sources.end(), s.selected_source))
    return fail(Code::selected_source_not_query_source, t);
  if (path.distance_labels.size() != path.vertices.size() ||
      path.edge_ids.size() + 1U != path.vertices.size())
    return fail(Code::path_shape_mismatch, t);
  if (path.edge_ids.size() != s.path_length) return fail(Code::path_length_mismatch, t);
  if (path.vertices.front() != s.selected_source) return fail(Code::source_termination_mismatch, t);
  if (path.vertices.back() != s.target) return fail(Code::target_termination_mismatch, t);

  const auto owner = graph.owner_tiles();
  CompactPathValidationResult verdict{};
  std::size_t marked = 0;
  for (; marked < path.vertices.size(); ++marked) {
    const std::uint32_t v = path.vertices[marked].value();
    if (v >= V) {
      verdict = fail(Code::vertex_out_of_range, t, marked);
      break;
    }
    if (!std::binary_search(query.selected_tiles.begin(), query.selected_tiles.end(), owner[v])) {
      verdict = fail(Code::vertex_outside_selected_region, t, marked);
      break;
    }
    if (seen[v]) {
      verdict = fail(Code::repeated_vertex, t, marked);
      break;
    }
    seen[v] = 1U;
    const float label = path.distance_labels[marked];
    if (!std::isfinite(label) || label < 0.0F) {
      verdict = fail(Code::invalid_distance, t, marked);
      ++marked;
      break;
    }
  }
  for (std::size_t i = 0; i < marked; ++i) {
    const std::uint32_t v = path.vertices[i].value();
    if (v < V) seen[v] = 0U;
  }
  if (!verdict.ok()) return verdict;
  if (path.distance_labels.front() != 0.0F || path.distance_labels.back() != s.distance)
    return fail(Code::target_distance_mismatch, t);

  const OutgoingCsrView out = graph.outgoing();
  float cost = 0.0F;
  for (std::size_t k = 0; k < path.edge_ids.size(); ++k) {
    const EdgeId e = path.edge_ids[k];
    if (e.value() >= graph.edge_count()) return fail(Code::edge_id_out_of_range, t, k);
    const std::uint32_t from = path.vertices[k].value();
    std::uint64_t at = out.row_offsets[from + 1U];
    for (std::uint64_t p = out.row_offsets[from]; p < out.row_offsets[from + 1U]; ++p)
      if (out.edge_ids[p] == e) {
        at = p;
        break;
      }
    if (at == out.row_offsets[from + 1U] || out.destinations[at] != path.vertices[k + 1U])
      return fail(Code::edge_continuity_mismatch, t, k);
    const float w = out.weights[at];
    if (path.distance_labels[k] + w != path.distance_labels[k + 1U]) return fail(Code::non_tight_edge, t, k);
    cost = cost + w;
  }
  if (cost != s.distance) return fail(Code::reported_cost_mismatch, t);
  return {};
}

CompactPathValidationResult check_payload(const WeightedGraph& graph, const RouteQuery& query,
                                          std::uint64_t query_id, std::uint32_t generation,
                                          const std::vector<std::uint32_t>& terminal_map,
                                          const std::vector<CompactTargetPath>& targets) {
  if (!validate_route_query(graph, query).ok()) return fail(Code::query_validation_failed);
  if (query_id != query.query_id.value()) return fail(Code::query_identity_mismatch);
  if (generation != query.expansion_generation) return fail(Code::expansion_generation_mismatch);
  if (terminal_map != query.target_terminal_to_target) return fail(Code::terminal_map_mismatch);
  if (targets.size() != query.targets.size()) return fail(Code::target_count_mismatch);

  std::vector<std::uint8_t> seen;
  for (std::size_t t = 0; t < targets.size(); ++t) {
    const CompactTargetPath& path = targets[t];
    const CompactTargetSummary& s = path.summary;
    if (s.target != query.targets[t]) return fail(Code::target_identity_mismatch, t);
    if (!known_reach(s.reached)) return fail(Code::invalid_reach_status, t);
    if (!known_path_status(s.reconstruction)) return fail(Code::invalid_reconstruction_status, t);
    if (s.has_selected_source > 1U) return fail(Code::invalid_selected_source_flag, t);
    if (!label_value_ok(s.distance)) return fail(Code::invalid_distance, t);
    const bool no_arenas = path.vertices.empty() && path.distance_labels.empty() && path.edge_ids.empty();
    const bool no_path = s.has_selected_source == 0U && s.path_length == 0U && no_arenas;
    switch (s.reconstruction) {
      case CompactPathStatus::complete: {
        if (seen.empty()) seen.assign(graph.vertex_count(), 0U);
        const CompactPathValidationResult r = check_complete_path(graph, query, path, t, seen);
        if (!r.ok()) return r;
        break;
      }
      case CompactPathStatus::unreachable:
        if (s.reached != CompactTargetReachStatus::not_reached || std::isfinite(s.distance) || !no_path)
          return fail(Code::inconsistent_unreachable_result, t);
        break;
      case CompactPathStatus::query_terminal_failure:
        if (s.reached != CompactTargetReachStatus::not_reached || std::isfinite(s.distance) || !no_path)
          return fail(Code::inconsistent_failure_result, t);
        break;
      case CompactPathStatus::no_tight_path:
      case CompactPathStatus::path_length_overflow:
        if (s.reached != CompactTargetReachStatus::reached || !std::isfinite(s.distance) || !no_path)
          return fail(Code::inconsistent_reconstruction_failure, t);
        break;
    }
  }
  return {};
}

bool graph_ok(const WeightedGraph& graph) {
  const char* message = validate_weighted_graph_message(graph);
  return message == nullptr || std::string_view{message}.empty();
}

}  // namespace

CompactQueryResult make_compact_query_result(CompactPathPayload payload,
                                             ExpansionQueryDisposition disposition) {
  CompactQueryResult r;
  r.query_id = payload.query_id;
  r.expansion_generation = payload.expansion_generation;
  r.disposition = disposition;
  r.target_terminal_to_target = std::move(payload.target_terminal_to_target);
  r.targets = std::move(payload.targets);
  return r;
}

bool compact_path_payload_complete(const CompactPathPayload& payload) noexcept {
  if (payload.targets.empty()) return false;
  for (const CompactTargetPath& t : payload.targets)
    if (t.summary.reached != CompactTargetReachStatus::reached ||
        t.summary.reconstruction != CompactPathStatus::complete)
      return false;
  return true;
}

CompactTransferAccounting measure_compact_transfer(std::span<const CompactPathPayload> payloads) {
  CompactTransferAccounting a;
  for (const CompactPathPayload& p : payloads) {
    add_checked(a.summary_bytes, bytes_of(p.targets.size(), sizeof(CompactTargetSummary)));
    for (const CompactTargetPath& t : p.targets) {
      add_checked(a.vertex_bytes, bytes_of(t.vertices.size(), 4U));
      add_checked(a.distance_label_bytes, bytes_of(t.distance_labels.size(), 4U));
      add_checked(a.edge_id_bytes, bytes_of(t.edge_ids.size(), 4U));
    }
  }
  for (std::uint64_t part : {a.summary_bytes, a.vertex_bytes, a.distance_label_bytes, a.edge_id_bytes})
    add_checked(a.total_bytes, part);
  return a;
}

namespace detail {
CompactPathValidationResult validate_compact_payload_trusted_graph(const WeightedGraph& graph,
                                                                   const RouteQuery& query,
                                                                   const CompactPathPayload& payload) {
  return check_payload(graph, query, payload.query_id.value(), payload.expansion_generation,
                       payload.target_terminal_to_target, payload.targets);
}
}  // namespace detail

CompactPathValidationResult validate_compact_path_payload(const WeightedGraph& graph,
                                                          const RouteQuery& query,
                                                          const CompactPathPayload& payload) {
  if (!graph_ok(graph)) return fail(Code::graph_validation_failed);
  return detail::validate_compact_payload_trusted_graph(graph, query, payload);
}

CompactPathValidationResult validate_compact_query_result(const WeightedGraph& graph,
                                                          const RouteQuery& query,
                                                          const CompactQueryResult& result) {
  if (!known_disposition(result.disposition)) return fail(Code::invalid_query_disposition);
  if (!graph_ok(graph)) return fail(Code::graph_validation_failed);
  const CompactPathValidationResult r =
      check_payload(graph, query, result.query_id.value(), result.expansion_generation,
                    result.target_terminal_to_target, result.targets);
  if (!r.ok()) return r;
  bool all_complete = true, any_unreachable = false, all_classified = true, any_failed = false,
       all_terminal_failure = true;
  for (const CompactTargetPath& t : result.targets) {
    const CompactPathStatus s = t.summary.reconstruction;
    all_complete = all_complete && s == CompactPathStatus::complete;
    any_unreachable = any_unreachable || s == CompactPathStatus::unreachable;
    all_classified = all_classified && (s == CompactPathStatus::complete || s == CompactPathStatus::unreachable);
    any_failed = any_failed || reconstruction_failed(s);
    all_terminal_failure = all_terminal_failure && s == CompactPathStatus::query_terminal_failure;
  }
  switch (result.disposition) {
    case ExpansionQueryDisposition::reached:
      if (!all_complete) return fail(Code::reached_query_incomplete);
      break;
    case ExpansionQueryDisposition::unreachable_in_full_region:
      if (!any_unreachable || !all_classified) return fail(Code::inconsistent_unreachable_result);
      break;
    case ExpansionQueryDisposition::engine_failure:
      if (!any_failed) return fail(Code::failed_query_reported_complete);
      break;
    case ExpansionQueryDisposition::expansion_limit:
    case ExpansionQueryDisposition::region_stalled:
    case ExpansionQueryDisposition::identity_or_count_overflow:
      if (!all_terminal_failure) return fail(Code::inconsistent_failure_result);
      break;
  }
  return {};
}

CompactPathValidationResult validate_compact_query_result_against_distances(
    const WeightedGraph& graph, const RouteQuery& query, const CompactQueryResult& result,
    std::span<const float> final_distances) {
  const CompactPathValidationResult structural = validate_compact_query_result(graph, query, result);
  if (!structural.ok()) return structural;
  if (final_distances.size() != graph.vertex_count()) return fail(Code::distance_image_size_mismatch);
  for (const float d : final_distances)
    if (!label_value_ok(d)) return fail(Code::invalid_distance_image);
  for (std::size_t t = 0; t < result.targets.size(); ++t) {
    const CompactTargetPath& path = result.targets[t];
    const float at_target = final_distances[path.summary.target.value()];
    switch (path.summary.reconstruction) {
      case CompactPathStatus::complete:
        for (std::size_t i = 0; i < path.vertices.size(); ++i)
          if (path.distance_labels[i] != final_distances[path.vertices[i].value()])
            return fail(Code::target_distance_mismatch, t, i);
        break;
      case CompactPathStatus::unreachable:
        if (std::isfinite(at_target)) return fail(Code::target_distance_mismatch, t);
        break;
      case CompactPathStatus::no_tight_path:
      case CompactPathStatus::path_length_overflow:
        if (path.summary.distance != at_target) return fail(Code::target_distance_mismatch, t);
        break;
      case CompactPathStatus::query_terminal_failure:
        break;
    }
  }
  return {};
}

}  // namespace bfnew
