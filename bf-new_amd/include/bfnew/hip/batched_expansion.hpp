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

// The same runner contract with one device run per planning pass: the first
// call of a pass (batch_index 0, or a pass whose queries differ from the
// last run's) submits every query of the pass as one batch -- the engines
// keep labels per (query, selected tile), so a batch needs no lanes and a
// query's labels, rounds and classification do not depend on the others --
// and every planned lane batch of the pass is answered from that run.
// Outcomes, classifications and compact payloads equal
// BatchedExpansionExecutor's; the device sees a few large launches that fill
// the chip instead of one small launch per lane batch.  Transfer modes
// compact_status and compact_paths; the pass's host timing and its
// path-offset words are reported with the batch that ran it.
class PassBatchedExpansionExecutor final {
 public:
  // Throws std::invalid_argument for invalid options, a transfer mode other
  // than compact_status / compact_paths, or compact paths with
  // instrumentation.
  PassBatchedExpansionExecutor(const WeightedGraph& host_graph, const TileRunLayout64& tile_runs,
                               const ResidentDeviceGraph& resident_graph,
                               ReusableBatchedWorkspace& workspace, const HipStream& stream,
                               GpuRunOptions run_options,
                               BatchedExpansionTransferMode transfer_mode =
                                   BatchedExpansionTransferMode::compact_status);
  ~PassBatchedExpansionExecutor();
  PassBatchedExpansionExecutor(const PassBatchedExpansionExecutor&) = delete;
  PassBatchedExpansionExecutor& operator=(const PassBatchedExpansionExecutor&) = delete;
  PassBatchedExpansionExecutor(PassBatchedExpansionExecutor&&) = delete;
  PassBatchedExpansionExecutor& operator=(PassBatchedExpansionExecutor&&) = delete;

  [[nodiscard]] EngineKind kind() const noexcept;
  [[nodiscard]] BatchedExpansionTransferMode transfer_mode() const noexcept;
  void set_tuning(const bfn_tuning& tuning) noexcept;
  // Device runs made so far (one per planning pass served).
  [[nodiscard]] std::uint64_t device_runs() const noexcept;

  [[nodiscard]] ExpansionBatchExecution operator()(std::span<const RouteQuery> queries,
                                                   std::span<const BatchQueryFeatures> features,
                                                   const BatchPlanEntry& batch,
                                                   const ExpansionBatchContext& context);

 private:
  class Impl;
  std::unique_ptr<Impl> impl_;
};

}  // namespace bfnew::hip
