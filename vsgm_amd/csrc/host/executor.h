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

#include <future>
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
  // Memory of the view-builder context used by PrefetchDeviceView (0: a
  // quarter of the executor's budget), taken out of the executor's budget
  // when the first prefetch starts.
  uint64_t prefetch_bytes = 0;
  // Match the plans of one call as a prefix tree: a level op that several
  // plans reach through identical levels runs once per batch for all of
  // them (plan-prefix sharing across queries, e.g. the -km graphlets).
  // Per-plan match times are then not measured (SharedMatchMs holds the
  // whole call's), and shared levels' statistics go to the first plan.
  bool share_prefixes = false;
};

// What determines a plan's level op and the state it leaves: the operator,
// its vertex, the materialised and pending vertex groups, and the backward
// set and conditions the op applies. Plans whose signatures agree on levels
// 0..l hold identical partial embeddings after level l.
std::string LevelSignature(const Plan& p, size_t level);
// Number of leading levels two plans share.
size_t SharedPrefix(const Plan& a, const Plan& b);

class DeviceBuffer;

// Launch-shape and budget choices for one plan (set by the autotuner).
struct PlanTuning {
  uint32_t isect_kernel = 4;
  uint32_t max_grid = 8192;
  uint64_t batch_bytes = 0;     // 0: the executor's option
  uint32_t cold_weight = 16;    // isect_kernel 3 / 4 pivot cost of a cold list test
  uint32_t ticket_chunks = 64;  // isect_kernel 3 / 4 chunks per ticket
};

class PipelineExecutor {
 public:
  PipelineExecutor(int device, uint64_t vertex_count, std::vector<const Plan*> plans, uint64_t max_sources, ExecutorOptions opt = {});
  ~PipelineExecutor();
  PipelineExecutor(const PipelineExecutor&) = delete;
  PipelineExecutor& operator=(const PipelineExecutor&) = delete;

  // Upload a materialised view and its sources (pipeline_executor.h:64-81);
  // replaces every resident view.
  void Transfer(const ViewBinHolder& vbh);
  // Upload a whole graph as one view whose sources are `sources` (all
  // vertices when empty); equivalent to a single-bin partition with
  // hop >= the graph diameter. Replaces every resident view.
  void TransferGraph(const Graph& g, const std::vector<uintV>& sources = {});
  // Keep one more view resident (all views stay in HBM together); returns
  // its index and makes it the current view.
  size_t AddView(const ViewBinHolder& vbh);
  size_t ViewCount() const { return views_.size(); }
  // Keep the whole graph resident in HBM; AddDeviceView then builds a bin's
  // view on the device from it (no host materialisation, only the sources
  // cross PCIe) and keeps it resident like AddView.
  void SetGraph(const Graph& g);
  size_t AddDeviceView(const std::vector<uintV>& sources, int hop, uint64_t root_degree);
  // Build one bin's view on the device like AddDeviceView, replacing every
  // resident view (the device-side counterpart of Transfer). A view
  // prefetched with the same arguments is taken instead of building it.
  void TransferDevice(const std::vector<uintV>& sources, int hop, uint64_t root_degree);
  // Start building a bin's view in the background: a builder thread drives
  // a second device context (its own stream, same priority, and its own
  // memory pool) while this thread goes on (typically with the current
  // bin's Match), so the next bin's view build overlaps the matching. The
  // next TransferDevice with the same arguments waits for it and installs
  // it (its membership index is built then, on the executor's context).
  // Replaces the reference's producer/consumer overlap of host view
  // materialisation with Transfer + Match (pattern_matching_pipeline.cu:
  // 103-172), here with both the build and the match on the device.
  void PrefetchDeviceView(const std::vector<uintV>& sources, int hop, uint64_t root_degree);
  bool PrefetchPending() const { return prefetch_pending_; }
  // Host copy of a resident view's CSR.
  void ViewCsr(size_t view, std::vector<uint64_t>& row_ptrs, std::vector<uint32_t>& cols) const;
  // Enumerate every plan over the current view (pipeline_executor.h:83-109).
  void Match(bool do_match = true);
  // Enumerate the plans in `which` (all when empty) over every resident view.
  void MatchAll(const std::vector<int>& which = {});

  // Batched mode: run MatchAll(which) once while recording the values the
  // host reads back, then capture the same work into one device graph.
  // Returns the graph's node count. ReplayGraph runs the graph once and adds
  // its counts to the totals.
  uint64_t CaptureAll(const std::vector<int>& which = {});
  void ReplayGraph();
  void ReleaseGraph();

  void SetPlanTuning(size_t plan, const PlanTuning& t);
  const PlanTuning& GetPlanTuning(size_t plan) const { return tuning_.at(plan); }

  const std::vector<uintC>& GetTotalCounts() const { return total_counts_; }
  const std::vector<std::vector<LevelStat>>& GetLevelStats() const { return level_stats_; }
  const std::vector<double>& GetMatchMs() const { return match_ms_; }  // per plan, summed over Match calls
  double SharedMatchMs() const { return shared_ms_; }                   // summed over prefix-shared Match calls
  uint64_t SharedLevelRuns() const { return shared_runs_; }             // level ops run once for several plans
  void ResetCounts();
  void SetPlans(std::vector<const Plan*> plans);
  vsgm_ctx* Context() { return ctx_; }
  ExecutorOptions& Options() { return opt_; }
  uint64_t PeakDeviceBytes() const;
  uint32_t HubCount() const { return views_.empty() ? 0 : views_[cur_].hub.n_hubs; }

 private:
  struct Level;
  struct View {
    std::shared_ptr<DeviceBuffer> row_ptrs, cols, sources;
    uint64_t sources_num = 0;
    vsgm_hub_index hub{nullptr, nullptr, 0, 0};
    bool hub_shared = false;  // hub is the whole graph's index (not freed with the view)
    bool loop_free = false;   // no v in adj(v): '!=' against backward vertices is implied
  };
  // A device view built on some context, before its membership index.
  struct BuiltView {
    View view;
  };
  BuiltView BuildViewOn(vsgm_ctx* ctx, const std::vector<uintV>& sources, int hop, uint64_t root_degree) const;
  void JoinPrefetch();
  void ClearViews();
  void IndexView(View& v, const uint64_t* h_row_ptrs, uint64_t hub_limit);
  void UseGraphIndex(View& v);
  View BuildDeviceView(const std::vector<uintV>& sources, int hop, uint64_t root_degree, uint64_t hub_limit);
  View UploadView(const uint64_t* row_ptrs, const uint32_t* cols, uint64_t n_cols, const uint32_t* sources, uint64_t n_sources,
                  uint64_t hub_limit);
  void MatchPlans(const std::vector<size_t>& views, const std::vector<int>& which);
  void ProcessLevel(size_t level, const Level& parent);
  void Descend(size_t level, const Level& parent);     // next level: of plan_, or of every plan in group_
  void ProcessShared(size_t level, const Level& parent);
  void RunBatch(size_t level, const Level& slice);
  uint64_t LevelBatchBytes(size_t level);
  void FillArgs(size_t level, const Level& l, vsgm_level_args& a) const;
  void CheckSources(const uint32_t* sources, uint64_t n_sources) const;
  void Check(int status, const char* what) const;
  void ApplyTuning(size_t plan);

  int device_;
  uint64_t vertex_count_;
  std::vector<const Plan*> plans_;
  ExecutorOptions opt_;
  vsgm_ctx* ctx_ = nullptr;

  // view prefetch: the builder context (created on the first prefetch), the
  // build in flight and its arguments; a joined build waits in prefetched_
  vsgm_ctx* builder_ = nullptr;
  std::future<BuiltView> prefetch_;
  bool prefetch_pending_ = false;
  std::unique_ptr<BuiltView> prefetched_;
  std::vector<uintV> prefetch_sources_;
  int prefetch_hop_ = 0;
  uint64_t prefetch_root_degree_ = 0;

  std::vector<View> views_;
  size_t cur_ = 0;
  std::shared_ptr<DeviceBuffer> graph_rp_, graph_cols_;  // whole graph (SetGraph)
  // Device views copy whole lists of the resident graph, so one membership
  // index of the graph's largest lists serves every one of them: built at
  // the first device view after SetGraph, freed with the graph.
  // the largest lists (degree, vertex), largest first: every index the
  // options can ask for at the default size and more
  std::vector<std::pair<uint64_t, uint32_t>> graph_hubs_;
  vsgm_hub_index graph_hub_{nullptr, nullptr, 0, 0};
  bool graph_hub_built_ = false;
  bool graph_loop_free_ = false;
  const View* view_ = nullptr;  // view being matched

  const Plan* plan_ = nullptr;  // plan being matched
  size_t plan_index_ = 0;
  uint64_t* d_counts_ = nullptr;  // u64 per plan
  uint64_t* d_count_ = nullptr;   // the matched plan's slot of d_counts_
  uint64_t* h_counts_ = nullptr;  // pinned landing slots of the graph's count copy
  std::vector<PlanTuning> tuning_;
  uint64_t batch_override_ = 0;   // the matched plan's batch budget (0: none)

  // batched mode: 0 eager, 1 recording, 2 capturing
  int tape_mode_ = 0;
  std::vector<uint64_t> batch_tape_;  // level budgets seen while recording
  size_t batch_pos_ = 0;
  bool graph_ready_ = false;
  std::vector<int> graph_plans_;
  double replay_ms_ = 0;

 public:
  double ReplayMs() const { return replay_ms_; }

 private:

  std::vector<uintC> total_counts_;
  std::vector<std::vector<LevelStat>> level_stats_;
  std::vector<double> match_ms_;
  // prefix sharing: the plans that reached the current level through the
  // same levels (plan_ is the first of them)
  bool sharing_ = false;
  std::vector<int> group_;
  double shared_ms_ = 0;
  uint64_t shared_runs_ = 0;
};

}  // namespace vsgm
