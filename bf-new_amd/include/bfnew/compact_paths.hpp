// compact_paths.hpp -- checks and byte accounting of compact path payloads
// (reference: include/bfnew/compact_paths.hpp:14-162, src/compact_paths.cpp:269-870).
//
// Paths are produced on the device (hip::ReusableCompactPathWorkspace and the
// engines' run_compact_paths); this header holds the host-side rules a
// payload must satisfy before a caller retains it.
#pragma once

#include "bfnew/batched_expansion.hpp"
#include "bfnew/compact_path_types.hpp"

#include <cstddef>
#include <cstdint>
#include <limits>
#include <span>
#include <vector>

namespace bfnew {

// A payload with the disposition its query ended with.
struct CompactQueryResult {
  QueryId query_id{};
  std::uint32_t expansion_generation{};
  ExpansionQueryDisposition disposition{ExpansionQueryDisposition::engine_failure};
  std::vector<std::uint32_t> target_terminal_to_target;
  std::vector<CompactTargetPath> targets;  // canonical target order
  bool operator==(const CompactQueryResult&) const = default;
};

[[nodiscard]] CompactQueryResult make_compact_query_result(CompactPathPayload payload,
                                                           ExpansionQueryDisposition disposition);

// Every target reached with a complete path (and at least one target).
[[nodiscard]] bool compact_path_payload_complete(const CompactPathPayload& payload) noexcept;

// Host reconstruction of one query's compact paths from a graph-sized label
// image (reference: src/compact_paths.cpp:540-640): per canonical target,
// +inf -> unreachable; otherwise the first tight path in stable-EdgeId order
// whose vertices all lie in the query's selected tiles, found by a
// backtracking search that never revisits a vertex of the current path ->
// complete, or no_tight_path when there is none.  Throws
// std::invalid_argument for an invalid graph, query or label image (size,
// NaN, negative, a source not at 0), std::logic_error if the payload it
// built fails validate_compact_path_payload.
[[nodiscard]] CompactPathPayload reconstruct_compact_path_payload(
    const WeightedGraph& graph, const RouteQuery& query, std::span<const float> final_distances);

// The payload above with its disposition, checked against the disposition
// rules of validate_compact_query_result (std::logic_error otherwise).
[[nodiscard]] CompactQueryResult reconstruct_compact_query_paths(
    const WeightedGraph& graph, const RouteQuery& query, std::span<const float> final_distances,
    ExpansionQueryDisposition disposition);

// A label-free result for a query that ended in an expansion or engine
// failure: every target query_terminal_failure.  Throws
// std::invalid_argument for reached, unreachable_in_full_region or an
// unknown disposition.
[[nodiscard]] CompactQueryResult make_failed_compact_query_result(
    const RouteQuery& query, ExpansionQueryDisposition disposition);

// One result per expansion outcome, in order (outcomes sorted by query id):
// a retained compact payload is moved out and checked (and compared with the
// outcome's label image when it kept one); otherwise a label image is
// reconstructed (reached / unreachable_in_full_region) or a label-free
// failure is published.  Each outcome's payload and label image are released.
// Throws std::invalid_argument for an invalid graph or non-canonical
// outcomes, std::logic_error for a payload that fails its checks or a
// reached outcome without paths or labels.
[[nodiscard]] std::vector<CompactQueryResult> extract_host_compact_paths(
    const WeightedGraph& graph, std::span<ExpansionQueryOutcome> outcomes);

// Bytes the payloads occupy on the device-to-host wire: 28-byte summaries,
// 4-byte vertices, labels and edge ids (edge ids are 32-bit on the device).
// Throws std::overflow_error when a total does not fit 64 bits.
[[nodiscard]] CompactTransferAccounting measure_compact_transfer(
    std::span<const CompactPathPayload> payloads);

enum class CompactPathValidationErrorCode : std::uint8_t {
  none,
  graph_validation_failed,
  query_validation_failed,
  invalid_query_disposition,
  query_identity_mismatch,
  expansion_generation_mismatch,
  terminal_map_mismatch,
  target_count_mismatch,
  target_identity_mismatch,
  invalid_reach_status,
  invalid_reconstruction_status,
  invalid_distance,
  reached_query_incomplete,
  failed_query_reported_complete,
  inconsistent_unreachable_result,
  inconsistent_failure_result,
  inconsistent_reconstruction_failure,
  invalid_selected_source_flag,
  selected_source_out_of_range,
  selected_source_not_query_source,
  path_shape_mismatch,
  path_length_mismatch,
  vertex_out_of_range,
  vertex_outside_selected_region,
  repeated_vertex,
  source_termination_mismatch,
  target_termination_mismatch,
  edge_id_out_of_range,
  edge_continuity_mismatch,
  reported_cost_mismatch,
  distance_image_size_mismatch,
  invalid_distance_image,
  target_distance_mismatch,
  non_tight_edge,
};

struct CompactPathValidationResult {
  static constexpr std::size_t no_position = std::numeric_limits<std::size_t>::max();
  CompactPathValidationErrorCode code{CompactPathValidationErrorCode::none};
  std::size_t target_index{no_position};
  std::size_t path_index{no_position};
  [[nodiscard]] constexpr bool ok() const noexcept {
    return code == CompactPathValidationErrorCode::none;
  }
};

// Identity (query, generation, terminal map, canonical targets), status
// consistency per target, and for a complete path: it starts at a query
// source with label 0, ends at the target with the reported distance, is
// simple, stays in the selected tiles, follows stable edge ids, every edge is
// tight in fp32 (label[i] + w == label[i+1]) and the fp32 left-to-right sum of
// its weights equals the reported distance.
[[nodiscard]] CompactPathValidationResult validate_compact_path_payload(
    const WeightedGraph& graph, const RouteQuery& query, const CompactPathPayload& payload);

// The payload checks plus the disposition rule: reached => all complete;
// unreachable_in_full_region => some unreachable and all complete or
// unreachable; engine_failure => some reconstruction failure; the expansion
// failures => every target query_terminal_failure.
[[nodiscard]] CompactPathValidationResult validate_compact_query_result(
    const WeightedGraph& graph, const RouteQuery& query, const CompactQueryResult& result);

// The query-result checks plus agreement with a full label image of the
// query (one graph-sized lane): path labels equal the image, unreachable
// targets are +inf there, failed reconstructions report the image's label.
[[nodiscard]] CompactPathValidationResult validate_compact_query_result_against_distances(
    const WeightedGraph& graph, const RouteQuery& query, const CompactQueryResult& result,
    std::span<const float> final_distances);

namespace detail {
// validate_compact_path_payload without the O(V + E) graph check, for callers
// that have already validated the graph once.
[[nodiscard]] CompactPathValidationResult validate_compact_payload_trusted_graph(
    const WeightedGraph& graph, const RouteQuery& query, const CompactPathPayload& payload);
}  // namespace detail

}  // namespace bfnew
