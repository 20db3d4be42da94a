// batched_expansion.hpp -- the batch-runner boundary through which an
// expansion/replanning orchestrator drives the relaxation engines
// (reference: include/bfnew/batched_expansion.hpp:130-181).
//
// A runner receives one planned lane batch (BatchPlanEntry) with the queries
// and features it indexes, runs it, and returns one ExpansionBatchExecution.
#pragma once

#include "bfnew/batch_plan.hpp"
#include "bfnew/compact_path_types.hpp"
#include "bfnew/gpu_api.hpp"
#include "bfnew/query.hpp"

#include <cstdint>
#include <functional>
#include <span>
#include <vector>

namespace bfnew {

enum class ExpansionWorkEvidence : std::uint8_t {
  unavailable,
  measured,
};

struct ExpansionBatchContext {
  std::uint64_t execution_ordinal{};
  std::uint32_t planning_pass{};
  std::uint32_t batch_index{};
  bool retry_pass{};
  constexpr bool operator==(const ExpansionBatchContext&) const noexcept = default;
};

struct ExpansionBatchExecution {
  GpuSsspResult result{};
  ExpansionWorkEvidence work_evidence{ExpansionWorkEvidence::unavailable};
  // Edge records read by the device (shared by the lanes of a tile visit)
  // and per-lane admitted relaxations; meaningful when measured.
  std::uint64_t shared_edge_work{};
  std::uint64_t logical_lane_edge_work{};
  // Vertex-major, lane-contiguous labels; empty unless the runner downloads
  // them (graph.vertex_count() * batch.lane_width entries otherwise).
  std::vector<float> final_distances;
  // Compact mode: one payload per valid lane of a cleanly converged batch,
  // sorted by QueryId.
  std::vector<CompactPathPayload> compact_paths;
  CompactTransferAccounting compact_transfer{};
  std::uint64_t compact_status_bytes{};
  std::uint64_t compact_error_bytes{};
  std::uint64_t compact_total_device_to_host_bytes{};
  std::uint64_t compact_controller_poll_count{};
  std::uint64_t compact_controller_poll_bytes{};
  std::uint64_t compact_overall_device_to_host_bytes{};
  CompactPathExecutionMetrics compact_execution{};
};

using ExpansionBatchRunner = std::function<ExpansionBatchExecution(
    std::span<const RouteQuery> queries, std::span<const BatchQueryFeatures> features,
    const BatchPlanEntry& batch, const ExpansionBatchContext& context)>;

}  // namespace bfnew
