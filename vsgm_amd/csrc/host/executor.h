// Pipeline executor: the per-device enumeration engine.
//
// Mirrors PipelineExecutor (src/executor/pipeline_executor.h:20-120) and
// Processor (src/executor/processor.h:20-280): for every plan the exec
// levels run as a depth-first recursion over levels with breadth-first,
// memory-budgeted batches inside each level. The level data (ImData,
// src/executor/im_data.h) lives in HBM as per-pattern-vertex columns; all
// device work goes through the C-ABI of include/vsgm_hip.h on one stream.
// Embedding counts accumulate in a device u64 per plan and are read once
// at the end of Match.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "common.h"
#include "graph.h"
#include "plan.h"
#include "view.h"
#include "vsgm_hip.h"

namespace vsgm {

struct LevelStat {
  uint64_t batches = 0;
  uint64_t paths = 0;    // partial embeddings entering the level
  uint64_t work = 0;     // elements visited by the level's kernels
  uint64_t outputs = 0;  // candidates / new paths produced
  uint64_t adj_entries = 0;  // backward adjacency volume of intersections
  double ms = 0;         // device time (profiling mode only)
};

struct ExecutorOptions {
  uint64_t mem_budget = 0;      // 0: 90% of free device memory
  uint64_t batch_bytes = 0;     // 0: budget / remaining levels (BatchManager::GetSafeBatchSize)
  uint64_t min_batch_bytes = 128ull << 20;
  bool profile = false;         // time every level op with HIP events
  // Dense membership index of the view's highest-degree lists (V/8 bytes
  // each): at most hub_bytes of it (~0: min(budget/8, 16 GiB); 0: off), over
  // vertices of view degree >= hub_min_degree.
  uint64_t hub_bytes = ~0ull;
  uint64_t hub_min_degree = 64;
};

class DeviceBuffer;

class PipelineExecutor {
 public:
  PipelineExecutor(int device, uint64_t vertex_count, std::vector<const Plan*> plans, uint64_t max_sources, ExecutorOptions opt = {});
  ~PipelineExecutor();
  PipelineExecutor(const PipelineExecutor&) = delete;
  PipelineExecutor& operator=(const PipelineExecutor&) = delete;

  // Upload a materialised view and its sources (pipeline_executor.h:64-81).
  void Transfer(const ViewBinHolder& vbh);
  // Upload a whole graph as one view whose sources are `sources` (all
  // vertices when empty); equivalent to a single-bin partition with
  // hop >= the graph diameter.
  void TransferGraph(const Graph& g, const std::vector<uintV>& sources = {});
  // Enumerate every plan over the current view (pipeline_executor.h:83-109).
  void Match(bool do_match = true);

  const std::vector<uintC>& GetTotalCounts() const { return total_counts_; }
  const std::vector<std::vector<LevelStat>>& GetLevelStats() const { return level_stats_; }
  const std::vector<double>& GetMatchMs() const { return match_ms_; }  // per plan, summed over Match calls
  void ResetCounts();
  void SetPlans(std::vector<const Plan*> plans);
  vsgm_ctx* Context() { return ctx_; }
  ExecutorOptions& Options() { return opt_; }
  uint64_t PeakDeviceBytes() const;
  uint32_t HubCount() const { return hub_.n_hubs; }

 private:
  struct Level;
  void BuildHubIndex(const uint64_t* h_row_ptrs);
  void ProcessLevel(size_t level, const Level& parent);
  void RunBatch(size_t level, const Level& slice);
  uint64_t LevelBatchBytes(size_t level) const;
  void FillArgs(size_t level, const Level& l, vsgm_level_args& a) const;
  void Check(int status, const char* what) const;

  int device_;
  uint64_t vertex_count_;
  std::vector<const Plan*> plans_;
  ExecutorOptions opt_;
  vsgm_ctx* ctx_ = nullptr;

  std::shared_ptr<DeviceBuffer> d_row_ptrs_, d_cols_, d_sources_;
  uint64_t cols_size_ = 0;
  uint64_t sources_num_ = 0;
  vsgm_hub_index hub_{nullptr, nullptr, 0, 0};

  const Plan* plan_ = nullptr;  // plan being matched
  size_t plan_index_ = 0;
  uint64_t* d_count_ = nullptr;
  std::vector<uintC> total_counts_;
  std::vector<std::vector<LevelStat>> level_stats_;
  std::vector<double> match_ms_;
};

}  // namespace vsgm
