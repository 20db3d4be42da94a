// hip/batched_expansion.hpp -- ExpansionBatchRunner over the batched engines
// (reference: include/bfnew/hip/batched_expansion.hpp:14-104).
//
// One executor is bound to one engine (run_options.engine), its run options,
// one transfer mode and a reusable workspace; calls are sequential.  Every
// call lays the planned lanes out afresh and runs the engine from its reset
// labels: nothing of a previous call is reused.
#pragma once

#include "bfnew/batched_expansion.hpp"
#include "bfnew/hip/engines.hpp"

#include <cstdint>
#include <memory>
#include <span>

namespace bfnew::hip {

enum class BatchedExpansionTransferMode : std::uint8_t {
  compact_status = 0U,            // one DeviceRunStatus
  status_and_work_evidence = 1U,  // status plus measured work (light instrumentation)
  compact_paths = 2U,             // status plus compact paths of every valid lane
};

[[nodiscard]] constexpr bool valid_batched_expansion_transfer_mode(
    BatchedExpansionTransferMode mode) noexcept {
  return mode == BatchedExpansionTransferMode::compact_status ||
         mode == BatchedExpansionTransferMode::status_and_work_evidence ||
         mode == BatchedExpansionTransferMode::compact_paths;
}

class BatchedExpansionExecutor final {
 public:
  // Throws std::invalid_argument for invalid options, an invalid transfer
  // mode, or compact paths requested with instrumentation.
  BatchedExpansionExecutor(const WeightedGraph& host_graph, const TileRunLayout64& tile_runs,
                           const ResidentDeviceGraph& resident_graph,
                           ReusableBatchedWorkspace& workspace, const HipStream& stream,
                           GpuRunOptions run_options,
                           BatchedExpansionTransferMode transfer_mode =
                               BatchedExpansionTransferMode::compact_status);
  ~BatchedExpansionExecutor();
  BatchedExpansionExecutor(const BatchedExpansionExecutor&) = delete;
  BatchedExpansionExecutor& operator=(const BatchedExpansionExecutor&) = delete;
  BatchedExpansionExecutor(BatchedExpansionExecutor&&) = delete;
  BatchedExpansionExecutor& operator=(BatchedExpansionExecutor&&) = delete;

  [[nodiscard]] EngineKind kind() const noexcept;
  [[nodiscard]] BatchedExpansionTransferMode transfer_mode() const noexcept;
  // Engine knobs beyond GpuRunOptions (set_tuning of the engine).
  void set_tuning(const bfn_tuning& tuning) noexcept;

  // Signature-compatible with ExpansionBatchRunner (bind with std::ref).
  [[nodiscard]] ExpansionBatchExecution operator()(std::span<const RouteQuery> queries,
                                                   std::span<const BatchQueryFeatures> features,
                                                   const BatchPlanEntry& batch,
                                                   const ExpansionBatchContext& context);

 private:
  class Impl;
  std::unique_ptr<Impl> impl_;
};

}  // namespace bfnew::hip
