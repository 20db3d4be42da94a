// compact_path_types.hpp -- per-target path results
// (reference: include/bfnew/compact_path_types.hpp:12-84).
#pragma once

#include "bfnew/types.hpp"
#include "bfnew_amd.h"

#include <cstdint>
#include <limits>
#include <type_traits>
#include <vector>

namespace bfnew {

enum class CompactTargetReachStatus : std::uint32_t {
  not_reached = 0U,
  reached = 1U,
};

enum class CompactPathStatus : std::uint32_t {
  complete = 0U,
  unreachable = 1U,
  query_terminal_failure = 2U,
  no_tight_path = 3U,
  path_length_overflow = 4U,
};

// path_length is an edge count, meaningful only for a complete path.
struct CompactTargetSummary {
  VertexId target{};
  VertexId selected_source{};
  float distance{std::numeric_limits<float>::infinity()};
  std::uint32_t path_length{};
  CompactTargetReachStatus reached{CompactTargetReachStatus::not_reached};
  CompactPathStatus reconstruction{CompactPathStatus::query_terminal_failure};
  std::uint32_t has_selected_source{};
  constexpr bool operator==(const CompactTargetSummary&) const noexcept = default;
};

static_assert(std::is_standard_layout_v<CompactTargetSummary>);
static_assert(std::is_trivially_copyable_v<CompactTargetSummary>);
static_assert(sizeof(CompactTargetSummary) == 28U && alignof(CompactTargetSummary) == 4U);
static_assert(sizeof(CompactTargetSummary) == sizeof(bfn_target_summary));

// A complete path: path_length + 1 vertices (global reordered ids, source
// first) with their final labels, and path_length stable edge ids.
struct CompactTargetPath {
  CompactTargetSummary summary{};
  std::vector<VertexId> vertices;
  std::vector<float> distance_labels;
  std::vector<EdgeId> edge_ids;
  bool operator==(const CompactTargetPath&) const = default;
};

struct CompactTransferAccounting {
  std::uint64_t summary_bytes{};
  std::uint64_t vertex_bytes{};
  std::uint64_t distance_label_bytes{};
  std::uint64_t edge_id_bytes{};
  std::uint64_t total_bytes{};
  constexpr bool operator==(const CompactTransferAccounting&) const noexcept = default;
};

// One query's compact paths: its targets in canonical (sorted) order and the
// map from the caller's target terminals to them
// (reference: include/bfnew/compact_path_types.hpp:65-72).
struct CompactPathPayload {
  QueryId query_id{};
  std::uint32_t expansion_generation{};
  std::vector<std::uint32_t> target_terminal_to_target;
  std::vector<CompactTargetPath> targets;
  bool operator==(const CompactPathPayload&) const = default;
};

enum class CompactStageTimingEvidence : std::uint32_t {
  unavailable = 0U,
  measured = 1U,
};

// Device-event and host-wall stage durations of one compact-path execution
// (reference: include/bfnew/compact_path_types.hpp:96-111).
struct CompactPathExecutionMetrics {
  CompactStageTimingEvidence device_timing{CompactStageTimingEvidence::unavailable};
  CompactStageTimingEvidence host_timing{CompactStageTimingEvidence::unavailable};
  std::uint64_t sssp_device_nanoseconds{};
  std::uint64_t reconstruction_device_nanoseconds{};
  std::uint64_t result_transfer_device_nanoseconds{};
  std::uint64_t sssp_host_nanoseconds{};
  std::uint64_t reconstruction_host_nanoseconds{};
  std::uint64_t result_transfer_host_nanoseconds{};
  std::uint64_t end_to_end_host_nanoseconds{};
  constexpr bool operator==(const CompactPathExecutionMetrics&) const noexcept = default;
};

}  // namespace bfnew
