// Pipeline executor implementation. See executor.h.
#include "executor.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>

namespace vsgm {

// Pool allocation owned by a shared_ptr; level columns share it.
class DeviceBuffer {
 public:
  DeviceBuffer(vsgm_ctx* ctx, void* adopt) : ctx_(ctx), ptr_(adopt) {}
  DeviceBuffer(vsgm_ctx* ctx, uint64_t bytes) : ctx_(ctx) {
    int s = vsgm_alloc(ctx, bytes, &ptr_);
    if (s != VSGM_OK) throw Error(std::string("device allocation failed: ") + vsgm_last_error(ctx));
  }
  ~DeviceBuffer() {
    if (ptr_) vsgm_free(ctx_, ptr_);
  }
  void* get() const { return ptr_; }

 private:
  vsgm_ctx* ctx_;
  void* ptr_ = nullptr;
};

namespace {
template <typename T>
struct Column {
  std::shared_ptr<DeviceBuffer> owner;
  T* ptr = nullptr;
  uint64_t size = 0;
};

template <typename T>
Column<T> Adopt(vsgm_ctx* ctx, T* p, uint64_t size) {
  Column<T> c;
  c.owner = std::make_shared<DeviceBuffer>(ctx, (void*)p);
  c.ptr = p;
  c.size = size;
  return c;
}

bool Contains(const VTGroup& g, uintV u) { return std::find(g.begin(), g.end(), u) != g.end(); }

// Index bytes per view: the option, or min(budget/8, 16 GiB) shared by
// `views` resident views.
uint64_t HubLimit(vsgm_ctx* ctx, uint64_t option, size_t views) {
  if (option != ~0ull) return option;
  uint64_t budget = 0;
  vsgm_mem_info(ctx, nullptr, &budget, nullptr);
  return std::min<uint64_t>(budget / 8, 16ull << 30) / std::max<size_t>(views, 1);
}
}  // namespace

struct PipelineExecutor::Level {
  uint64_t path_num = 0;
  Column<uint32_t> inst[kMaxQueryVerticesNum];
  Column<uint32_t> cand[kMaxQueryVerticesNum];
  Column<uint64_t> cand_off[kMaxQueryVerticesNum];
  Column<uint64_t> cand_idx[kMaxQueryVerticesNum];  // ptr == nullptr: identity rows
  uint64_t cand_idx_base[kMaxQueryVerticesNum] = {0};

  // Non-owning batch slice [l, l+count) of the path-indexed columns
  // (ImData::CopyBatchData, im_data.h:22-45); candidate CSRs stay whole.
  Level Slice(uint64_t l, uint64_t count, const VTGroup& mat, const VTGroup& cu) const {
    Level s;
    s.path_num = count;
    for (auto u : mat) {
      s.inst[u] = inst[u];
      s.inst[u].ptr += l;
      s.inst[u].size = count;
    }
    for (auto v : cu) {
      s.cand[v] = cand[v];
      s.cand_off[v] = cand_off[v];
      s.cand_idx[v] = cand_idx[v];
      s.cand_idx_base[v] = cand_idx_base[v];
      if (cand_idx[v].ptr) {
        s.cand_idx[v].ptr += l;
        s.cand_idx[v].size = count;
      } else {
        s.cand_idx_base[v] += l;
      }
    }
    return s;
  }
};

PipelineExecutor::PipelineExecutor(int device, uint64_t vertex_count, std::vector<const Plan*> plans, uint64_t max_sources,
                                   ExecutorOptions opt)
    : device_(device), vertex_count_(vertex_count), opt_(opt) {
  (void)max_sources;
  int s = vsgm_ctx_create(device, opt_.mem_budget, &ctx_);
  if (s == VSGM_ERR_NO_DEVICE) throw Error("no HIP device is visible: the enumerator needs a GPU");
  if (s != VSGM_OK) throw Error("cannot create the device context on device " + std::to_string(device) + " (status " + std::to_string(s) + ")");
  SetPlans(std::move(plans));
}

PipelineExecutor::~PipelineExecutor() {
  vsgm_graph_release(ctx_);
  try {
    JoinPrefetch();
  } catch (const std::exception&) {
    // a failed background build leaves nothing behind
  }
  prefetched_.reset();
  ClearViews();
  vsgm_hub_index_free(ctx_, &graph_hub_);
  graph_rp_.reset();
  graph_cols_.reset();
  if (d_counts_) vsgm_free(ctx_, d_counts_);
  if (h_counts_) vsgm_host_free(h_counts_);
  vsgm_ctx_destroy(builder_);  // after every view it built is gone
  vsgm_ctx_destroy(ctx_);
}

void PipelineExecutor::SetPlans(std::vector<const Plan*> plans) {
  plans_ = std::move(plans);
  for (auto* p : plans_) VSGM_CHECK(p && p->GetVertexCount() <= (size_t)kMaxQueryVerticesNum, "bad plan");
  ReleaseGraph();
  if (d_counts_) vsgm_free(ctx_, d_counts_);
  if (h_counts_) vsgm_host_free(h_counts_);
  d_counts_ = nullptr;
  h_counts_ = nullptr;
  Check(vsgm_alloc(ctx_, std::max<size_t>(plans_.size(), 1) * sizeof(uint64_t), (void**)&d_counts_), "count allocation");
  Check(vsgm_host_alloc(std::max<size_t>(plans_.size(), 1) * sizeof(uint64_t), (void**)&h_counts_), "count staging allocation");
  tuning_.assign(plans_.size(), PlanTuning{});
  ResetCounts();
}

void PipelineExecutor::SetPlanTuning(size_t plan, const PlanTuning& t) {
  VSGM_CHECK(plan < plans_.size(), "plan index out of range");
  VSGM_CHECK(t.isect_kernel <= 4 && t.max_grid > 0 && t.max_grid <= (1u << 20), "bad plan tuning");
  VSGM_CHECK(t.cold_weight >= 1 && t.cold_weight <= 1024 && t.ticket_chunks >= 1 && t.ticket_chunks <= 64, "bad plan tuning");
  ReleaseGraph();
  tuning_[plan] = t;
}

void PipelineExecutor::ApplyTuning(size_t plan) {
  const PlanTuning& t = tuning_[plan];
  vsgm_tuning kt{t.isect_kernel, t.max_grid, t.cold_weight, t.ticket_chunks};
  Check(vsgm_set_tuning(ctx_, &kt), "tuning");
  batch_override_ = t.batch_bytes;
}

void PipelineExecutor::ResetCounts() {
  total_counts_.assign(plans_.size(), 0);
  match_ms_.assign(plans_.size(), 0.0);
  level_stats_.assign(plans_.size(), {});
  for (size_t i = 0; i < plans_.size(); ++i) level_stats_[i].assign(plans_[i]->GetExecuteOperations().size(), LevelStat{});
}

uint64_t PipelineExecutor::PeakDeviceBytes() const {
  uint64_t peak = 0;
  vsgm_mem_info(ctx_, nullptr, nullptr, &peak);
  return peak;
}

void PipelineExecutor::Check(int status, const char* what) const {
  if (status != VSGM_OK) throw Error(std::string(what) + " failed: " + vsgm_last_error(ctx_));
}

void PipelineExecutor::ClearViews() {
  // views built by the builder context are freed into its pool, which a
  // background build must not be using meanwhile
  JoinPrefetch();
  ReleaseGraph();
  for (auto& v : views_)
    if (!v.hub_shared) vsgm_hub_index_free(ctx_, &v.hub);
  views_.clear();
  cur_ = 0;
  view_ = nullptr;
}

// Upload one view (row_ptrs over the full vertex range, cols, sources) and
// index its highest-degree lists: vertices of degree >= hub_min_degree,
// largest first, as many as hub_limit bytes hold.
// Every source becomes a level-0 binding whose row_ptrs the kernels read, so
// ids outside the vertex range are refused before any upload or launch.
void PipelineExecutor::CheckSources(const uint32_t* sources, uint64_t n_sources) const {
  VSGM_CHECK(n_sources == 0 || sources, "null sources");
  for (uint64_t i = 0; i < n_sources; ++i)
    VSGM_CHECK(sources[i] < vertex_count_, "source " + std::to_string(sources[i]) + " (position " + std::to_string(i) +
                                               ") is not a vertex of the " + std::to_string(vertex_count_) + "-vertex graph");
}

PipelineExecutor::View PipelineExecutor::UploadView(const uint64_t* row_ptrs, const uint32_t* cols, uint64_t n_cols, const uint32_t* sources,
                                                    uint64_t n_sources, uint64_t hub_limit) {
  CheckSources(sources, n_sources);
  View v;
  v.row_ptrs = std::make_shared<DeviceBuffer>(ctx_, (vertex_count_ + 1) * sizeof(uintE));
  v.cols = std::make_shared<DeviceBuffer>(ctx_, std::max<uint64_t>(n_cols, 1) * sizeof(uintV));
  v.sources = std::make_shared<DeviceBuffer>(ctx_, std::max<uint64_t>(n_sources, 1) * sizeof(uintV));
  v.sources_num = n_sources;
  Check(vsgm_h2d(ctx_, v.row_ptrs->get(), row_ptrs, (vertex_count_ + 1) * sizeof(uintE)), "row_ptrs upload");
  Check(vsgm_h2d(ctx_, v.cols->get(), cols, n_cols * sizeof(uintV)), "cols upload");
  Check(vsgm_h2d(ctx_, v.sources->get(), sources, n_sources * sizeof(uintV)), "sources upload");
  Check(vsgm_ctx_sync(ctx_), "view upload");
  IndexView(v, row_ptrs, hub_limit);
  return v;
}

void PipelineExecutor::IndexView(View& v, const uint64_t* row_ptrs, uint64_t hub_limit) {
  uint64_t loops = 0;
  Check(vsgm_view_self_loops(ctx_, (const uint64_t*)v.row_ptrs->get(), (const uint32_t*)v.cols->get(), vertex_count_, &loops), "self-loop check");
  v.loop_free = loops == 0;
  const uint64_t row_bytes = (vertex_count_ + 63) / 64 * 8;
  const uint64_t cap = std::min<uint64_t>(hub_limit / row_bytes, 0xFFFFFFFEull);
  if (cap == 0 || vertex_count_ == 0) return;
  std::vector<std::pair<uint64_t, uintV>> cand;
  for (uint64_t u = 0; u < vertex_count_; ++u) {
    const uint64_t d = row_ptrs[u + 1] - row_ptrs[u];
    if (d >= std::max<uint64_t>(opt_.hub_min_degree, 1)) cand.emplace_back(d, (uintV)u);
  }
  if (cand.empty()) return;
  const size_t k = (size_t)std::min<uint64_t>(cap, cand.size());
  std::partial_sort(cand.begin(), cand.begin() + k, cand.end(),
                    [](const auto& x, const auto& y) { return x.first > y.first || (x.first == y.first && x.second < y.second); });
  std::vector<uint32_t> hubs(k);
  for (size_t i = 0; i < k; ++i) hubs[i] = cand[i].second;
  Check(vsgm_hub_index_build(ctx_, (const uint64_t*)v.row_ptrs->get(), (const uint32_t*)v.cols->get(), vertex_count_, hubs.data(),
                             (uint32_t)k, &v.hub),
        "hub index build");
}

void PipelineExecutor::SetGraph(const Graph& g) {
  VSGM_CHECK(g.vertex_count == vertex_count_, "graph vertex count does not match the executor");
  JoinPrefetch();  // a background build reads the resident graph
  prefetched_.reset();
  ReleaseGraph();
  bool shared = false;  // device views of the previous graph share its index
  for (const auto& v : views_) shared = shared || v.hub_shared;
  if (shared) ClearViews();
  vsgm_hub_index_free(ctx_, &graph_hub_);
  graph_hub_built_ = false;
  graph_rp_ = std::make_shared<DeviceBuffer>(ctx_, (vertex_count_ + 1) * sizeof(uintE));
  graph_cols_ = std::make_shared<DeviceBuffer>(ctx_, std::max<uint64_t>(g.cols.size(), 1) * sizeof(uintV));
  Check(vsgm_h2d(ctx_, graph_rp_->get(), g.row_ptrs.data(), (vertex_count_ + 1) * sizeof(uintE)), "row_ptrs upload");
  Check(vsgm_h2d(ctx_, graph_cols_->get(), g.cols.data(), g.cols.size() * sizeof(uintV)), "cols upload");
  Check(vsgm_ctx_sync(ctx_), "graph upload");
  uint64_t loops = 0;
  Check(vsgm_view_self_loops(ctx_, (const uint64_t*)graph_rp_->get(), (const uint32_t*)graph_cols_->get(), vertex_count_, &loops),
        "self-loop check");
  graph_loop_free_ = loops == 0;
  // index candidates in the order IndexView takes them (degree, then id),
  // kept up to 2^38 bytes of index rows (at least 65,536 rows)
  std::vector<std::pair<uint64_t, uintV>> cand;
  for (uint64_t u = 0; u < vertex_count_; ++u) {
    const uint64_t d = g.row_ptrs[u + 1] - g.row_ptrs[u];
    if (d > 0) cand.emplace_back(d, (uintV)u);
  }
  const uint64_t row_bytes = (vertex_count_ + 63) / 64 * 8;
  const size_t keep = (size_t)std::min<uint64_t>(cand.size(), std::max<uint64_t>(65536, (1ull << 38) / std::max<uint64_t>(row_bytes, 1)));
  auto by_degree = [](const auto& x, const auto& y) { return x.first > y.first || (x.first == y.first && x.second < y.second); };
  std::partial_sort(cand.begin(), cand.begin() + keep, cand.end(), by_degree);
  cand.resize(keep);
  graph_hubs_ = std::move(cand);
}

// A device view's membership index and loop flag are the resident graph's:
// its lists are copies of the graph's, and the kernels only probe the index
// rows of vertices whose lists the view holds.
void PipelineExecutor::UseGraphIndex(View& v) {
  if (!graph_hub_built_) {
    const uint64_t row_bytes = (vertex_count_ + 63) / 64 * 8;
    const uint64_t cap = std::min<uint64_t>(HubLimit(ctx_, opt_.hub_bytes, 1) / row_bytes, 0xFFFFFFFEull);
    std::vector<uint32_t> hubs;
    for (const auto& [d, u] : graph_hubs_) {
      if (hubs.size() >= cap || d < std::max<uint64_t>(opt_.hub_min_degree, 1)) break;
      hubs.push_back(u);
    }
    if (!hubs.empty())
      Check(vsgm_hub_index_build(ctx_, (const uint64_t*)graph_rp_->get(), (const uint32_t*)graph_cols_->get(), vertex_count_, hubs.data(),
                                 (uint32_t)hubs.size(), &graph_hub_),
            "hub index build");
    graph_hub_built_ = true;
  }
  v.hub = graph_hub_;
  v.hub_shared = true;
  v.loop_free = graph_loop_free_;
}

size_t PipelineExecutor::AddDeviceView(const std::vector<uintV>& sources, int hop, uint64_t root_degree) {
  ReleaseGraph();
  // the automatic index size assumes up to 8 resident views
  views_.push_back(BuildDeviceView(sources, hop, root_degree, HubLimit(ctx_, opt_.hub_bytes, 8)));
  cur_ = views_.size() - 1;
  return cur_;
}

void PipelineExecutor::TransferDevice(const std::vector<uintV>& sources, int hop, uint64_t root_degree) {
  JoinPrefetch();
  std::unique_ptr<BuiltView> b = std::move(prefetched_);
  const bool hit = b && hop == prefetch_hop_ && root_degree == prefetch_root_degree_ && sources == prefetch_sources_;
  if (!hit) b.reset();  // a prefetch of another bin is dropped
  ClearViews();
  if (hit) {
    UseGraphIndex(b->view);
    views_.push_back(std::move(b->view));
  } else {
    views_.push_back(BuildDeviceView(sources, hop, root_degree, HubLimit(ctx_, opt_.hub_bytes, 1)));
  }
  cur_ = 0;
}

void PipelineExecutor::PrefetchDeviceView(const std::vector<uintV>& sources, int hop, uint64_t root_degree) {
  VSGM_CHECK(graph_rp_ && graph_cols_, "device view before SetGraph");
  CheckSources(sources.data(), sources.size());
  JoinPrefetch();
  prefetched_.reset();
  if (!builder_) {
    uint64_t budget = 0, available = 0;
    vsgm_mem_info(ctx_, &available, &budget, nullptr);
    const uint64_t take = opt_.prefetch_bytes ? opt_.prefetch_bytes : budget / 4;
    VSGM_CHECK(take < available, "the view builder's memory does not fit in what the executor has left");
    Check(vsgm_ctx_set_budget(ctx_, budget - take), "budget split for the view builder");
    vsgm_ctx* b = nullptr;
    if (vsgm_ctx_create(device_, take, &b) != VSGM_OK) {
      vsgm_ctx_set_budget(ctx_, budget);
      throw Error("cannot create the view builder's device context");
    }
    builder_ = b;
  }
  prefetch_sources_ = sources;
  prefetch_hop_ = hop;
  prefetch_root_degree_ = root_degree;
  prefetch_ = std::async(std::launch::async, [this, sources, hop, root_degree] {
    if (vsgm_ctx_bind_thread(builder_) != VSGM_OK) throw Error("view builder thread: cannot select the device");
    return BuildViewOn(builder_, sources, hop, root_degree);
  });
  prefetch_pending_ = true;
}

void PipelineExecutor::JoinPrefetch() {
  if (!prefetch_pending_) return;
  prefetch_pending_ = false;
  prefetched_.reset(new BuiltView(prefetch_.get()));  // rethrows a failed build
}

// The view of one bin on context ctx (its stream and pool): BFS levels over
// the resident graph, the copy of the selected lists, and the row_ptrs read
// back for the membership index; returns once the device work is complete.
PipelineExecutor::BuiltView PipelineExecutor::BuildViewOn(vsgm_ctx* ctx, const std::vector<uintV>& sources, int hop,
                                                          uint64_t root_degree) const {
  auto check = [ctx](int status, const char* what) {
    if (status != VSGM_OK) throw Error(std::string(what) + " failed: " + vsgm_last_error(ctx));
  };
  BuiltView b;
  View& v = b.view;
  v.sources = std::make_shared<DeviceBuffer>(ctx, std::max<uint64_t>(sources.size(), 1) * sizeof(uintV));
  v.sources_num = sources.size();
  check(vsgm_h2d(ctx, v.sources->get(), sources.data(), sources.size() * sizeof(uintV)), "sources upload");
  vsgm_view_out o;
  check(vsgm_view_build(ctx, (const uint64_t*)graph_rp_->get(), (const uint32_t*)graph_cols_->get(), vertex_count_,
                        (const uint32_t*)v.sources->get(), sources.size(), hop, root_degree, &o),
        "view build");
  v.row_ptrs = std::make_shared<DeviceBuffer>(ctx, (void*)o.row_ptrs);
  v.cols = std::make_shared<DeviceBuffer>(ctx, (void*)o.cols);
  check(vsgm_ctx_sync(ctx), "view build");
  return b;
}

PipelineExecutor::View PipelineExecutor::BuildDeviceView(const std::vector<uintV>& sources, int hop, uint64_t root_degree,
                                                         uint64_t hub_limit) {
  VSGM_CHECK(graph_rp_ && graph_cols_, "device view before SetGraph");
  CheckSources(sources.data(), sources.size());
  (void)hub_limit;
  BuiltView b = BuildViewOn(ctx_, sources, hop, root_degree);
  UseGraphIndex(b.view);
  return std::move(b.view);
}

void PipelineExecutor::ViewCsr(size_t view, std::vector<uint64_t>& row_ptrs, std::vector<uint32_t>& cols) const {
  VSGM_CHECK(view < views_.size(), "view index out of range");
  const View& v = views_[view];
  row_ptrs.resize(vertex_count_ + 1);
  Check(vsgm_d2h(ctx_, row_ptrs.data(), v.row_ptrs->get(), row_ptrs.size() * sizeof(uint64_t)), "view readback");
  cols.resize(row_ptrs.back());
  Check(vsgm_d2h(ctx_, cols.data(), v.cols->get(), cols.size() * sizeof(uint32_t)), "view readback");
}

void PipelineExecutor::Transfer(const ViewBinHolder& vbh) {
  VSGM_CHECK(vbh.GetVertexCount() == vertex_count_, "view vertex count does not match the executor");
  ClearViews();
  views_.push_back(UploadView(vbh.GetRowPtrs(), vbh.GetCols(), vbh.GetTotalSize(), vbh.GetSources().data(), vbh.GetSourcesNum(),
                              HubLimit(ctx_, opt_.hub_bytes, 1)));
  cur_ = 0;
}

size_t PipelineExecutor::AddView(const ViewBinHolder& vbh) {
  VSGM_CHECK(vbh.GetVertexCount() == vertex_count_, "view vertex count does not match the executor");
  ReleaseGraph();
  // the automatic index size assumes up to 8 resident views
  views_.push_back(UploadView(vbh.GetRowPtrs(), vbh.GetCols(), vbh.GetTotalSize(), vbh.GetSources().data(), vbh.GetSourcesNum(),
                              HubLimit(ctx_, opt_.hub_bytes, 8)));
  cur_ = views_.size() - 1;
  return cur_;
}

void PipelineExecutor::TransferGraph(const Graph& g, const std::vector<uintV>& sources) {
  VSGM_CHECK(g.vertex_count == vertex_count_, "graph vertex count does not match the executor");
  ClearViews();
  std::vector<uintV> all;
  const std::vector<uintV>* src = &sources;
  if (sources.empty()) {
    all.resize(vertex_count_);
    for (uint64_t v = 0; v < vertex_count_; ++v) all[v] = (uintV)v;
    src = &all;
  }
  views_.push_back(UploadView(g.row_ptrs.data(), g.cols.data(), g.cols.size(), src->data(), src->size(), HubLimit(ctx_, opt_.hub_bytes, 1)));
  cur_ = 0;
}

void PipelineExecutor::Match(bool do_match) {
  if (!do_match) return;
  VSGM_CHECK(!views_.empty(), "Match before Transfer");
  MatchPlans({cur_}, {});
}

void PipelineExecutor::MatchAll(const std::vector<int>& which) {
  VSGM_CHECK(!views_.empty(), "Match before Transfer");
  std::vector<size_t> all(views_.size());
  for (size_t i = 0; i < all.size(); ++i) all[i] = i;
  MatchPlans(all, which);
}

// Every selected plan over every selected view, counts in d_counts_. Eager
// mode times each plan with HIP events and reads the counts back once;
// capture mode (tape_mode_ 2) leaves a copy of the counts into the pinned
// slots as the graph's last node.
std::string LevelSignature(const Plan& p, size_t level) {
  const auto& e = p.GetExecuteOperations().at(level);
  std::ostringstream s;
  auto group = [&s](const std::vector<uintV>& g) {
    s << "[";
    for (auto v : g) s << v << ",";
    s << "]";
  };
  auto conds = [&s](const std::vector<CondType>& c) {
    s << "{";
    for (const auto& x : c) s << (int)x.first << ":" << x.second << ",";
    s << "}";
  };
  const auto& mat = p.GetMaterializedVertices()[level];
  const auto& cu = p.GetComputedUnmaterializedVertices()[level];
  s << (int)e.first << " " << e.second << " ";
  group(mat);
  group(cu);
  switch (e.first) {
    case COMPUTE:
    case COMPUTE_COUNT:
    case COMPUTE_PATH_COUNT:
      group(p.GetBackwardConnectivity()[e.second]);
      conds(p.GetComputeConditions()[e.second]);
      if (e.first == COMPUTE_COUNT && cu.size() == 1) s << " o" << (int)p.ConditionBetween(e.second, cu[0]);
      break;
    case MATERIALIZE:
      conds(p.GetMaterializeConditions()[e.second]);
      break;
    case FILTER_COMPUTE:
      conds(p.GetFilterConditions()[e.second]);
      break;
    case COUNT:
      if (cu.size() == 2) s << " o" << (int)p.ConditionBetween(cu[0], cu[1]);
      break;
  }
  return s.str();
}

size_t SharedPrefix(const Plan& a, const Plan& b) {
  const size_t n = std::min(a.GetExecuteOperations().size(), b.GetExecuteOperations().size());
  size_t l = 0;
  while (l < n && LevelSignature(a, l) == LevelSignature(b, l)) ++l;
  return l;
}

void PipelineExecutor::MatchPlans(const std::vector<size_t>& views, const std::vector<int>& which) {
  std::vector<int> sel = which;
  if (sel.empty())
    for (size_t i = 0; i < plans_.size(); ++i) sel.push_back((int)i);
  for (int i : sel) VSGM_CHECK(i >= 0 && (size_t)i < plans_.size(), "plan index out of range");
  const bool capturing = tape_mode_ == 2;
  Check(vsgm_fill_u64(ctx_, d_counts_, plans_.size(), 0), "count reset");
  std::vector<std::pair<int, int>> copies;  // (plan, identical earlier plan): counted once, copied
  if (opt_.share_prefixes && sel.size() > 1) {
    std::vector<int> distinct;
    for (int i : sel) {
      int same = -1;
      const size_t n = plans_[i]->GetExecuteOperations().size();
      for (int j : distinct)
        if (plans_[j]->GetExecuteOperations().size() == n && SharedPrefix(*plans_[i], *plans_[j]) == n) same = j;
      if (same < 0) distinct.push_back(i);
      else copies.emplace_back(i, same);
    }
    if (!capturing) Check(vsgm_timer_start(ctx_, 0), "timer");
    sharing_ = true;
    try {
      for (size_t v : views) {
        view_ = &views_[v];
        group_ = distinct;
        ProcessShared(0, Level{});
      }
    } catch (...) {
      sharing_ = false;
      group_.clear();
      throw;
    }
    sharing_ = false;
    group_.clear();
    if (!capturing) {
      float ms = 0;
      Check(vsgm_timer_stop(ctx_, 0, &ms), "timer");
      shared_ms_ += ms;
    }
  } else {
    for (int i : sel) {
      plan_ = plans_[i];
      plan_index_ = i;
      d_count_ = d_counts_ + i;
      ApplyTuning(i);
      if (!capturing) Check(vsgm_timer_start(ctx_, 0), "timer");
      for (size_t v : views) {
        view_ = &views_[v];
        ProcessLevel(0, Level{});
      }
      if (!capturing) {
        float ms = 0;
        Check(vsgm_timer_stop(ctx_, 0, &ms), "timer");
        match_ms_[i] += ms;
      }
    }
  }
  plan_ = nullptr;
  view_ = nullptr;
  if (capturing) {
    Check(vsgm_d2h_async(ctx_, h_counts_, d_counts_, plans_.size() * sizeof(uint64_t)), "count copy");
    return;
  }
  std::vector<uint64_t> c(plans_.size(), 0);
  Check(vsgm_d2h(ctx_, c.data(), d_counts_, c.size() * sizeof(uint64_t)), "count readback");
  for (const auto& [i, j] : copies) c[i] = c[j];
  for (int i : sel) total_counts_[i] += c[i];
}

// Level `level` for every plan of group_ (all identical before it): plans
// whose op at this level is the same run it once, as a subgroup led by its
// first plan, and descend together.
void PipelineExecutor::ProcessShared(size_t level, const Level& parent) {
  const std::vector<int> group = group_;
  const Plan* saved_plan = plan_;
  const size_t saved_index = plan_index_;
  uint64_t* saved_count = d_count_;
  std::vector<std::string> keys;
  std::vector<std::vector<int>> subs;
  for (int p : group) {
    const std::string k = LevelSignature(*plans_[p], level);
    size_t i = 0;
    while (i < keys.size() && keys[i] != k) ++i;
    if (i == keys.size()) {
      keys.push_back(k);
      subs.emplace_back();
    }
    subs[i].push_back(p);
  }
  for (const auto& sub : subs) {
    group_ = sub;
    plan_ = plans_[sub[0]];
    plan_index_ = sub[0];
    d_count_ = d_counts_ + sub[0];
    ApplyTuning(sub[0]);
    if (sub.size() > 1) shared_runs_ += sub.size() - 1;
    ProcessLevel(level, parent);
  }
  // back to the caller's plan (its next batch runs its own level op)
  group_ = group;
  plan_ = saved_plan;
  plan_index_ = saved_index;
  d_count_ = saved_count;
  if (plan_) ApplyTuning(plan_index_);
}

void PipelineExecutor::Descend(size_t level, const Level& parent) {
  if (sharing_) ProcessShared(level, parent);
  else ProcessLevel(level, parent);
}

uint64_t PipelineExecutor::CaptureAll(const std::vector<int>& which) {
  VSGM_CHECK(!views_.empty(), "capture before Transfer");
  ReleaseGraph();
  batch_tape_.clear();
  tape_mode_ = 1;
  Check(vsgm_tape_record(ctx_), "tape record");
  try {
    MatchAll(which);
  } catch (...) {
    tape_mode_ = 0;
    vsgm_tape_stop(ctx_, nullptr);
    throw;
  }
  Check(vsgm_tape_stop(ctx_, nullptr), "tape stop");
  tape_mode_ = 2;
  batch_pos_ = 0;
  Check(vsgm_capture_begin(ctx_), "capture begin");
  uint64_t nodes = 0;
  try {
    MatchAll(which);
  } catch (...) {
    tape_mode_ = 0;
    vsgm_capture_end(ctx_, nullptr);
    vsgm_graph_release(ctx_);
    throw;
  }
  tape_mode_ = 0;
  Check(vsgm_capture_end(ctx_, &nodes), "capture end");
  VSGM_CHECK(batch_pos_ == batch_tape_.size(), "captured work diverged from the recording");
  graph_ready_ = true;
  graph_plans_ = which;
  return nodes;
}

void PipelineExecutor::ReplayGraph() {
  VSGM_CHECK(graph_ready_, "no captured graph");
  Check(vsgm_timer_start(ctx_, 0), "timer");
  Check(vsgm_graph_launch(ctx_), "graph launch");
  float ms = 0;
  Check(vsgm_timer_stop(ctx_, 0, &ms), "timer");
  replay_ms_ += ms;
  std::vector<int> sel = graph_plans_;
  if (sel.empty())
    for (size_t i = 0; i < plans_.size(); ++i) sel.push_back((int)i);
  for (int i : sel) total_counts_[i] += h_counts_[i];
}

void PipelineExecutor::ReleaseGraph() {
  if (graph_ready_) vsgm_graph_release(ctx_);
  graph_ready_ = false;
}

uint64_t PipelineExecutor::LevelBatchBytes(size_t level) {
  if (tape_mode_ == 2) {
    VSGM_CHECK(batch_pos_ < batch_tape_.size(), "captured work diverged from the recording");
    return batch_tape_[batch_pos_++];
  }
  uint64_t b;
  if (batch_override_) {
    b = batch_override_;
  } else if (opt_.batch_bytes) {
    b = opt_.batch_bytes;
  } else {
    uint64_t available = 0;
    vsgm_mem_info(ctx_, &available, nullptr, nullptr);
    size_t remaining = plan_->GetExecuteOperations().size() - level;
    b = std::max((uint64_t)(0.9 * (double)(available / std::max<size_t>(remaining, 1))), opt_.min_batch_bytes);
  }
  if (tape_mode_ == 1) batch_tape_.push_back(b);
  return b;
}

void PipelineExecutor::FillArgs(size_t level, const Level& l, vsgm_level_args& a) const {
  memset(&a, 0, sizeof(a));
  a.row_ptrs = (const uint64_t*)view_->row_ptrs->get();
  a.cols = (const uint32_t*)view_->cols->get();
  a.vertex_count = vertex_count_;
  a.path_num = l.path_num;
  a.hub_slot = view_->hub.slot;
  a.hub_bits = view_->hub.bits;
  a.hub_words = view_->hub.words;
  a.flags = view_->loop_free ? VSGM_LEVEL_LOOP_FREE : 0u;
  for (int u = 0; u < kMaxQueryVerticesNum; ++u) {
    a.inst[u] = l.inst[u].ptr;
    a.cand[u] = l.cand[u].ptr;
    a.cand_off[u] = l.cand_off[u].ptr;
    a.cand_idx[u] = l.cand_idx[u].ptr;
    a.cand_idx_base[u] = l.cand_idx_base[u];
  }
  const auto& e = plan_->GetExecuteOperations()[level];
  const auto& cu = plan_->GetComputedUnmaterializedVertices()[level];
  a.target = e.second;
  a.other = -1;
  auto set_conds = [&a](const std::vector<CondType>& c) {
    a.n_cond = (uint32_t)c.size();
    for (size_t i = 0; i < c.size(); ++i) a.cond[i] = vsgm_cond{(uint32_t)c[i].first, c[i].second};
  };
  switch (e.first) {
    case COMPUTE:
    case COMPUTE_COUNT:
    case COMPUTE_PATH_COUNT: {
      const auto& back = plan_->GetBackwardConnectivity()[e.second];
      a.n_back = (uint32_t)back.size();
      for (size_t i = 0; i < back.size(); ++i) a.back[i] = back[i];
      set_conds(plan_->GetComputeConditions()[e.second]);
      if (e.first == COMPUTE_COUNT && cu.size() == 1) {
        a.other = (int32_t)cu[0];
        a.other_op = (uint32_t)plan_->ConditionBetween(e.second, cu[0]);
      }
      VSGM_CHECK(e.first != COMPUTE_COUNT || cu.size() <= 1, "COMPUTE_COUNT with more than one pending candidate set");
    } break;
    case MATERIALIZE:
      set_conds(plan_->GetMaterializeConditions()[e.second]);
      break;
    case FILTER_COMPUTE:
      set_conds(plan_->GetFilterConditions()[e.second]);
      break;
    case COUNT:
      VSGM_CHECK(cu.size() == 1 || cu.size() == 2, "COUNT needs one or two candidate sets");
      a.target = cu[0];
      if (cu.size() == 2) {
        a.other = (int32_t)cu[1];
        a.other_op = (uint32_t)plan_->ConditionBetween(cu[0], cu[1]);
      }
      break;
  }
}

void PipelineExecutor::ProcessLevel(size_t level, const Level& parent) {
  const auto& ops = plan_->GetExecuteOperations();
  if (level == ops.size()) return;
  LevelStat unused;  // a capture only replays work already counted
  LevelStat& stat = tape_mode_ == 2 ? unused : level_stats_[plan_index_][level];
  const size_t remaining = ops.size() - level;

  if (level == 0) {
    // InitFirstLevel (materialize.h:11-18): root bound to the bin's sources,
    // split evenly by the level-0 cost (processor.h:207-213).
    const uint64_t P = view_->sources_num;
    if (P == 0) return;
    const uint64_t pf = sizeof(uintV) + (uint64_t)std::ceil(3.0 * sizeof(uint64_t) / remaining);
    const uint64_t bytes = LevelBatchBytes(0);
    uint64_t nb = std::max<uint64_t>(1, (P * pf + bytes - 1) / bytes);
    uint64_t per = (P + nb - 1) / nb;
    const uintV root = ops[0].second;
    for (uint64_t l = 0; l < P; l += per) {
      Level child;
      child.path_num = std::min(per, P - l);
      child.inst[root].owner = view_->sources;
      child.inst[root].ptr = (uint32_t*)view_->sources->get() + l;
      child.inst[root].size = child.path_num;
      stat.batches++;
      stat.paths += child.path_num;
      stat.outputs += child.path_num;
      Descend(1, child);
    }
    return;
  }

  const uint64_t P = parent.path_num;
  if (P == 0) return;
  const auto op = ops[level].first;
  const auto& mat = plan_->GetMaterializedVertices()[level];
  const auto& cu = plan_->GetComputedUnmaterializedVertices()[level];
  const uint64_t bytes = LevelBatchBytes(level);

  // cost model of this implementation's buffers (bytes per parent / child)
  int mode = 2;
  uint64_t pf = 16, cf = 0;
  switch (op) {
    case COMPUTE:
    case COMPUTE_PATH_COUNT:
      mode = 0;
      pf = 40;
      cf = 5;
      break;
    case MATERIALIZE:
      mode = 1;
      pf = 24;
      cf = 4 * (mat.size() + 1) + 8 * (cu.size() ? cu.size() - 1 : 0) + 17;
      break;
    case FILTER_COMPUTE:
      mode = 1;
      pf = 32;
      cf = 5;
      break;
    case COMPUTE_COUNT:
      pf = 16;
      break;
    case COUNT:
      pf = cu.size() == 2 ? 16 : 1;
      break;
  }

  std::vector<uint64_t> ends;
  if (mode == 2) {
    uint64_t nb = std::max<uint64_t>(1, (P * pf + bytes - 1) / bytes);
    uint64_t per = (P + nb - 1) / nb;
    for (uint64_t e = per; e < P; e += per) ends.push_back(e);
    ends.push_back(P);
  } else {
    vsgm_level_args a;
    FillArgs(level, parent, a);
    DeviceBuffer children(ctx_, P * sizeof(uint64_t));
    Check(vsgm_children_count(ctx_, &a, mode, (uint64_t*)children.get()), "children count");
    ends.resize(64);
    uint64_t nb = 0;
    int s = vsgm_split_batch(ctx_, (const uint64_t*)children.get(), P, pf, cf, bytes, ends.data(), ends.size(), &nb);
    if (s == VSGM_ERR_ARG && nb > ends.size()) {
      ends.resize(nb);
      s = vsgm_split_batch(ctx_, (const uint64_t*)children.get(), P, pf, cf, bytes, ends.data(), ends.size(), &nb);
    }
    Check(s, "batch split");
    ends.resize(nb);
  }

  uint64_t begin = 0;
  for (uint64_t end : ends) {
    Level slice = parent.Slice(begin, end - begin, mat, cu);
    RunBatch(level, slice);
    begin = end;
  }
}

void PipelineExecutor::RunBatch(size_t level, const Level& slice) {
  const auto& e = plan_->GetExecuteOperations()[level];
  const auto& mat = plan_->GetMaterializedVertices()[level];
  const auto& cu = plan_->GetComputedUnmaterializedVertices()[level];
  LevelStat unused;
  LevelStat& stat = tape_mode_ == 2 ? unused : level_stats_[plan_index_][level];
  const uintV u = e.second;
  vsgm_level_args a;
  FillArgs(level, slice, a);
  const bool timed = opt_.profile && tape_mode_ != 2;
  if (timed) Check(vsgm_timer_start(ctx_, 1), "timer");

  Level child;
  bool next = false;
  switch (e.first) {
    case COMPUTE: {
      vsgm_csr_out out;
      Check(vsgm_intersect_write(ctx_, &a, &out), "intersect");
      child = slice;
      child.cand[u] = Adopt(ctx_, out.values, out.count);
      child.cand_off[u] = Adopt(ctx_, out.offsets, slice.path_num + 1);
      child.cand_idx[u] = Column<uint64_t>{};
      child.cand_idx_base[u] = 0;
      next = out.count > 0;
    } break;
    case MATERIALIZE: {
      std::vector<uint32_t> carried;
      for (auto v : cu)
        if (v != u) carried.push_back(v);
      vsgm_expand_out out;
      Check(vsgm_expand(ctx_, &a, mat.data(), (uint32_t)mat.size(), carried.data(), (uint32_t)carried.size(), &out), "expand");
      child.path_num = out.path_num;
      child.inst[u] = Adopt(ctx_, out.inst[u], out.path_num);
      for (auto w : mat) child.inst[w] = Adopt(ctx_, out.inst[w], out.path_num);
      for (auto v : carried) {
        child.cand[v] = slice.cand[v];
        child.cand_off[v] = slice.cand_off[v];
        child.cand_idx[v] = Adopt(ctx_, out.cand_idx[v], out.path_num);
        child.cand_idx_base[v] = 0;
      }
      next = out.path_num > 0;
    } break;
    case FILTER_COMPUTE: {
      vsgm_csr_out out;
      Check(vsgm_filter(ctx_, &a, &out), "filter");
      child = slice;
      child.cand[u] = Adopt(ctx_, out.values, out.count);
      child.cand_off[u] = Adopt(ctx_, out.offsets, slice.path_num + 1);
      child.cand_idx[u] = Column<uint64_t>{};
      child.cand_idx_base[u] = 0;
      next = out.count > 0;
    } break;
    case COMPUTE_COUNT:
      Check(vsgm_intersect_count_total(ctx_, &a, d_count_), "intersect count");
      break;
    case COUNT:
      Check(vsgm_count_sets(ctx_, &a, d_count_), "count");
      break;
    case COMPUTE_PATH_COUNT:
      throw Error("COMPUTE_PATH_COUNT is not emitted by the planner");
  }
  vsgm_op_stats st;
  vsgm_last_op_stats(ctx_, &st);
  stat.batches++;
  stat.paths += slice.path_num;
  stat.work += st.work;
  stat.outputs += st.outputs;
  stat.adj_entries += st.adj_entries;
  if (timed) {
    float ms = 0;
    Check(vsgm_timer_stop(ctx_, 1, &ms), "timer");
    stat.ms += ms;
  }
  if (next) Descend(level + 1, child);
}

}  // namespace vsgm
