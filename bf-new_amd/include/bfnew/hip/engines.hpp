// engines.hpp -- the batched and single-query relaxation engines
// (reference: include/bfnew/hip/batched_jacobi.hpp:137-200,
// batched_dense_chaotic_push.hpp:105-165, batched_frontier_push.hpp:91-144,
// jacobi.hpp:64-125, dense_chaotic_push.hpp, frontier_push.hpp,
// compact_path_results.hpp:94-134).
//
// Same class names, constructor arguments, run methods and option checks as
// the reference.  Engines hold non-owning references to the graph, the
// workspace and the stream, and are neither copyable nor movable.
//
// Differences in what the engines compute: none -- every run reaches the
// unique fp32 fixed point, so distances, statuses and paths are those of the
// reference for the same inputs.  Jacobi runs point Jacobi by default, which
// also reproduces the reference's round counts; set_tuning() selects the
// faster block-Jacobi sweep or hipGraph replay.
#pragma once

#include "bfnew/batch_layout.hpp"
#include "bfnew/compact_path_types.hpp"
#include "bfnew/device_layout.hpp"
#include "bfnew/gpu_api.hpp"
#include "bfnew/graph.hpp"
#include "bfnew/hip/runtime.hpp"
#include "bfnew/query.hpp"
#include "bfnew_amd.h"

#include <cstdint>
#include <vector>

namespace bfnew::hip {

enum class BatchedJacobiLoadStrategy : std::uint32_t {
  compiler_uniform = 0U,
  explicit_wave_broadcast = 1U,
};

struct BatchedRunMetrics {
  float sssp_device_timeline_milliseconds{};
  float result_transfer_gpu_milliseconds{};
  double end_to_end_wall_milliseconds{};
  std::uint32_t lane_width{};
  std::uint32_t valid_lane_count{};
  std::uint64_t engine_round_dispatches{};
  std::uint64_t convergence_host_checks{};
  // Collected with light or debug instrumentation only.
  std::uint64_t tile_visits{};
  std::uint64_t edge_records_evaluated{};
  std::uint64_t algorithmic_bytes{};
  double batch_queries_per_second{};
};

struct BatchedRunOutput {
  GpuSsspResult result{};
  // Vertex-major final labels: distances[vertex * lane_width + lane]
  // (+inf outside the lane's selected tiles).  Empty unless downloaded.
  std::vector<float> distances;
  bool distances_downloaded{};
  // Per lane: rounds until the lane's labels stopped changing, plus the
  // final no-change round (0 for invalid lanes).
  std::vector<std::uint64_t> lane_convergence_rounds;
  std::vector<std::uint64_t> lane_executed_rounds;
  // Per lane: admitted edges whose source is reachable (Dijkstra's
  // relaxation count), filled with light or debug instrumentation.
  std::vector<std::uint64_t> lane_relaxations;
  BatchedRunMetrics metrics{};
};
using BatchedJacobiRunOutput = BatchedRunOutput;
using BatchedDenseChaoticPushRunOutput = BatchedRunOutput;
using BatchedFrontierPushRunOutput = BatchedRunOutput;

struct CompactPathBatchMetrics {
  float sssp_device_milliseconds{};            // HIP events around the batch's kernels
  double sssp_host_milliseconds{};             // host wall: submit to completion
  double reconstruction_host_milliseconds{};   // host wall: both path passes and their copies
  double end_to_end_wall_milliseconds{};
  // Device-to-host bytes beside the payload: the per-target path offsets and
  // the edge arena's one unused slot per path.
  std::uint64_t control_bytes{};
};

struct CompactPathBatchOutput {
  DeviceRunStatus status{};
  // Flattened in BatchDeviceDescription::targets order.
  std::vector<CompactTargetPath> targets;
  CompactTransferAccounting transfer{};
  CompactPathBatchMetrics metrics{};
};

// Shared machinery of the three batched engines.
class BatchedEngineBase {
 public:
  BatchedEngineBase(const BatchedEngineBase&) = delete;
  BatchedEngineBase& operator=(const BatchedEngineBase&) = delete;
  BatchedEngineBase(BatchedEngineBase&&) = delete;
  BatchedEngineBase& operator=(BatchedEngineBase&&) = delete;

  // Engine knobs beyond GpuRunOptions (tile sweeps, claim size, hipGraph
  // replay); bfn_tuning_default() plus point Jacobi when never called.
  void set_tuning(const bfn_tuning& tuning) noexcept { tuning_ = tuning; }
  [[nodiscard]] const bfn_tuning& tuning() const noexcept { return tuning_; }

 protected:
  BatchedEngineBase(EngineKind kind, const WeightedGraph& host_graph,
                    const TileRunLayout64& tile_runs, const ResidentDeviceGraph& resident_graph,
                    ReusableBatchedWorkspace& workspace, const HipStream& stream);
  ~BatchedEngineBase() = default;

  enum class Readback : std::uint8_t { status, compact_status, distances, compact_paths };
  BatchedRunOutput run_impl(const BatchDeviceDescription& batch, const GpuRunOptions& options,
                            Readback readback, CompactPathBatchOutput* paths);

 private:
  EngineKind kind_;
  const WeightedGraph& host_graph_;
  const TileRunLayout64& tile_runs_;
  const ResidentDeviceGraph& resident_graph_;
  ReusableBatchedWorkspace& workspace_;
  const HipStream& stream_;
  bfn_tuning tuning_{};
};

class BatchedJacobiPullEngine final : public BatchedEngineBase {
 public:
  BatchedJacobiPullEngine(const WeightedGraph& host_graph, const TileRunLayout64& tile_runs,
                          const ResidentDeviceGraph& resident_graph,
                          ReusableBatchedJacobiWorkspace& workspace, const HipStream& stream)
      : BatchedEngineBase(EngineKind::jacobi_pull, host_graph, tile_runs, resident_graph,
                          workspace, stream) {}
  [[nodiscard]] constexpr EngineKind kind() const noexcept { return EngineKind::jacobi_pull; }

  // The load strategy selects how the reference's kernel reads edge records;
  // any valid value runs the same kernel here.
  [[nodiscard]] BatchedJacobiRunOutput run_status_only(
      const BatchDeviceDescription& batch, const GpuRunOptions& options,
      BatchedJacobiLoadStrategy load_strategy = BatchedJacobiLoadStrategy::compiler_uniform);
  [[nodiscard]] DeviceRunStatus run_compact_status(
      const BatchDeviceDescription& batch, const GpuRunOptions& options,
      BatchedJacobiLoadStrategy load_strategy = BatchedJacobiLoadStrategy::compiler_uniform);
  [[nodiscard]] CompactPathBatchOutput run_compact_paths(
      const BatchDeviceDescription& batch, const GpuRunOptions& options,
      ReusableCompactPathWorkspace& compact_workspace,
      BatchedJacobiLoadStrategy load_strategy = BatchedJacobiLoadStrategy::compiler_uniform);
  [[nodiscard]] BatchedJacobiRunOutput run_with_distances(
      const BatchDeviceDescription& batch, const GpuRunOptions& options,
      BatchedJacobiLoadStrategy load_strategy = BatchedJacobiLoadStrategy::compiler_uniform);
};

class BatchedDenseChaoticPushEngine final : public BatchedEngineBase {
 public:
  BatchedDenseChaoticPushEngine(const WeightedGraph& host_graph, const TileRunLayout64& tile_runs,
                                const ResidentDeviceGraph& resident_graph,
                                ReusableBatchedDenseWorkspace& workspace, const HipStream& stream)
      : BatchedEngineBase(EngineKind::dense_chaotic_push, host_graph, tile_runs, resident_graph,
                          workspace, stream) {}
  [[nodiscard]] constexpr EngineKind kind() const noexcept {
    return EngineKind::dense_chaotic_push;
  }
  [[nodiscard]] BatchedDenseChaoticPushRunOutput run_status_only(
      const BatchDeviceDescription& batch, const GpuRunOptions& options);
  [[nodiscard]] DeviceRunStatus run_compact_status(const BatchDeviceDescription& batch,
                                                   const GpuRunOptions& options);
  [[nodiscard]] CompactPathBatchOutput run_compact_paths(
      const BatchDeviceDescription& batch, const GpuRunOptions& options,
      ReusableCompactPathWorkspace& compact_workspace);
  [[nodiscard]] BatchedDenseChaoticPushRunOutput run_with_distances(
      const BatchDeviceDescription& batch, const GpuRunOptions& options);
};

class BatchedFrontierPushEngine final : public BatchedEngineBase {
 public:
  BatchedFrontierPushEngine(const WeightedGraph& host_graph, const TileRunLayout64& tile_runs,
                            const ResidentDeviceGraph& resident_graph,
                            ReusableBatchedFrontierWorkspace& workspace, const HipStream& stream)
      : BatchedEngineBase(EngineKind::frontier_push, host_graph, tile_runs, resident_graph,
                          workspace, stream) {}
  [[nodiscard]] constexpr EngineKind kind() const noexcept { return EngineKind::frontier_push; }
  [[nodiscard]] BatchedFrontierPushRunOutput run_status_only(const BatchDeviceDescription& batch,
                                                             const GpuRunOptions& options);
  [[nodiscard]] DeviceRunStatus run_compact_status(const BatchDeviceDescription& batch,
                                                   const GpuRunOptions& options);
  [[nodiscard]] CompactPathBatchOutput run_compact_paths(
      const BatchDeviceDescription& batch, const GpuRunOptions& options,
      ReusableCompactPathWorkspace& compact_workspace);
  [[nodiscard]] BatchedFrontierPushRunOutput run_with_distances(
      const BatchDeviceDescription& batch, const GpuRunOptions& options);
};

// ---- single-query engines (plain mode) -----------------------------------
struct SingleQueryRunOutput {
  GpuSsspResult result{};
  // Full-graph labels (+inf outside the selected tiles), or the selected
  // tiles' labels concatenated in canonical tile order, or empty.
  std::vector<float> distances;
  bool distances_downloaded{};
  std::uint64_t relaxations{};
};
using JacobiRunOutput = SingleQueryRunOutput;
using DenseChaoticPushRunOutput = SingleQueryRunOutput;
using FrontierPushRunOutput = SingleQueryRunOutput;

class SingleQueryEngineBase : public GpuSsspEngine {
 public:
  SingleQueryEngineBase(const SingleQueryEngineBase&) = delete;
  SingleQueryEngineBase& operator=(const SingleQueryEngineBase&) = delete;

  [[nodiscard]] GpuSsspResult run(const RouteQuery& query, const GpuRunOptions& options) override;
  [[nodiscard]] SingleQueryRunOutput run_status_only(const RouteQuery& query,
                                                     const GpuRunOptions& options);
  [[nodiscard]] SingleQueryRunOutput run_with_distances(const RouteQuery& query,
                                                        const GpuRunOptions& options);
  [[nodiscard]] SingleQueryRunOutput run_with_selected_distances(const RouteQuery& query,
                                                                 const GpuRunOptions& options);
  void set_tuning(const bfn_tuning& tuning) noexcept { tuning_ = tuning; }

 protected:
  SingleQueryEngineBase(EngineKind kind, const WeightedGraph& host_graph,
                        const TileRunLayout64& tile_runs,
                        const ResidentDeviceGraph& resident_graph,
                        ReusableDeviceWorkspace& workspace, const HipStream& stream);

 private:
  enum class Readback : std::uint8_t { none, full_graph, selected_ranges };
  SingleQueryRunOutput run_impl(const RouteQuery& query, const GpuRunOptions& options,
                                Readback readback);

  EngineKind kind_;
  const WeightedGraph& host_graph_;
  const ResidentDeviceGraph& resident_graph_;
  ReusableDeviceWorkspace& workspace_;
  const HipStream& stream_;
  bfn_tuning tuning_{};
};

class JacobiPullEngine final : public SingleQueryEngineBase {
 public:
  JacobiPullEngine(const WeightedGraph& host_graph, const TileRunLayout64& tile_runs,
                   const ResidentDeviceGraph& resident_graph, ReusableDeviceWorkspace& workspace,
                   const HipStream& stream)
      : SingleQueryEngineBase(EngineKind::jacobi_pull, host_graph, tile_runs, resident_graph, workspace,
                              stream) {}
  [[nodiscard]] EngineKind kind() const noexcept override { return EngineKind::jacobi_pull; }
};

class DenseChaoticPushEngine final : public SingleQueryEngineBase {
 public:
  DenseChaoticPushEngine(const WeightedGraph& host_graph, const TileRunLayout64& tile_runs,
                         const ResidentDeviceGraph& resident_graph, ReusableDeviceWorkspace& workspace,
                         const HipStream& stream)
      : SingleQueryEngineBase(EngineKind::dense_chaotic_push, host_graph, tile_runs, resident_graph, workspace,
                              stream) {}
  [[nodiscard]] EngineKind kind() const noexcept override { return EngineKind::dense_chaotic_push; }
};

class FrontierPushEngine final : public SingleQueryEngineBase {
 public:
  FrontierPushEngine(const WeightedGraph& host_graph, const TileRunLayout64& tile_runs,
                     const ResidentDeviceGraph& resident_graph, ReusableDeviceWorkspace& workspace,
                     const HipStream& stream)
      : SingleQueryEngineBase(EngineKind::frontier_push, host_graph, tile_runs, resident_graph, workspace,
                              stream) {}
  [[nodiscard]] EngineKind kind() const noexcept override { return EngineKind::frontier_push; }
};

}  // namespace bfnew::hip
