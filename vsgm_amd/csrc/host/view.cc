// View runtime implementation. See view.h.
#include "view.h"

#include <omp.h>

#include <algorithm>
#include <cstring>
#include <fstream>

#include "vsgm_hip.h"

namespace vsgm {

ViewBinHolder::ViewBinHolder(uint64_t vertex_count, uint64_t max_view_bin_bytes)
    : vertex_count_(vertex_count), capacity_entries_(max_view_bin_bytes / sizeof(uintV)) {
  void* rp = nullptr;
  void* cp = nullptr;
  // pinned when a device runtime is present, pageable otherwise (the copy
  // then goes through a staging buffer)
  if (vsgm_host_alloc((vertex_count_ + 1) * sizeof(uintE), &rp) == VSGM_OK &&
      vsgm_host_alloc(std::max<uint64_t>(capacity_entries_, 1) * sizeof(uintV), &cp) == VSGM_OK) {
    pinned_ = true;
  } else {
    if (rp) vsgm_host_free(rp);
    if (cp) vsgm_host_free(cp);
    rp = malloc((vertex_count_ + 1) * sizeof(uintE));
    cp = malloc(std::max<uint64_t>(capacity_entries_, 1) * sizeof(uintV));
    VSGM_CHECK(rp && cp, "cannot allocate view holder");
  }
  row_ptrs_ = (uintE*)rp;
  cols_ = (uintV*)cp;
}

ViewBinHolder::~ViewBinHolder() {
  if (pinned_) {
    vsgm_host_free(row_ptrs_);
    vsgm_host_free(cols_);
  } else {
    free(row_ptrs_);
    free(cols_);
  }
}

uint64_t ViewBin::GetComputationFactors() const {
  uint64_t s = 0;
  for (auto v : sources_) s += graph_->Degree(v);
  return s;
}

void ViewBin::Materialize(ViewBinHolder& h, uint64_t root_degree) const {
  VSGM_CHECK(h.vertex_count_ == graph_->vertex_count, "holder / graph vertex count mismatch");
  const uint64_t V = graph_->vertex_count;
  const int nt = threads_ > 0 ? threads_ : omp_get_max_threads();
  std::vector<char> added(V, 0);
  std::vector<uintV> frontier;
  for (auto s : sources_)
    if (root_degree == 0 || graph_->Degree(s) >= root_degree) {
      if (!added[s]) {
        added[s] = 1;
        frontier.push_back(s);
      }
    }
  // hop-1 further BFS levels
  for (int level = 1; level < hop_ && !frontier.empty(); ++level) {
    std::vector<std::vector<uintV>> next(nt);
#pragma omp parallel num_threads(nt)
    {
      auto& mine = next[omp_get_thread_num()];
#pragma omp for schedule(dynamic, 256)
      for (uint64_t i = 0; i < frontier.size(); ++i) {
        uintV v = frontier[i];
        for (uintE j = graph_->row_ptrs[v]; j < graph_->row_ptrs[v + 1]; ++j) {
          uintV w = graph_->cols[j];
          if (!__atomic_load_n(&added[w], __ATOMIC_RELAXED)) {  // other threads set marks concurrently
            char expected = 0;
            if (__atomic_compare_exchange_n(&added[w], &expected, (char)1, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) mine.push_back(w);
          }
        }
      }
    }
    frontier.clear();
    for (auto& part : next) frontier.insert(frontier.end(), part.begin(), part.end());
  }
  uint64_t total = 0, vertices = 0;
  for (uint64_t v = 0; v < V; ++v) {
    h.row_ptrs_[v] = total;
    if (added[v]) {
      total += graph_->Degree((uintV)v);
      ++vertices;
    }
  }
  h.row_ptrs_[V] = total;
  VSGM_CHECK(total <= h.capacity_entries_, "view bin " + std::to_string(id_) + " needs " + std::to_string(total * 4) +
                                               " bytes, more than the holder's buffer (-m)");
#pragma omp parallel for num_threads(nt) schedule(dynamic, 4096)
  for (uint64_t v = 0; v < V; ++v)
    if (h.row_ptrs_[v + 1] > h.row_ptrs_[v])
      memcpy(h.cols_ + h.row_ptrs_[v], graph_->cols.data() + graph_->row_ptrs[v], (h.row_ptrs_[v + 1] - h.row_ptrs_[v]) * sizeof(uintV));
  h.total_size_ = total;
  h.vertices_num_ = vertices;
  h.id_ = id_;
  h.sources_ = sources_;
}

void ViewBinManager::SetPartition(int n_bins, const std::vector<int>& assignment) {
  VSGM_CHECK(assignment.size() == graph_->vertex_count, "partition size does not match the graph");
  std::vector<std::vector<uintV>> parts(n_bins);
  for (uint64_t v = 0; v < assignment.size(); ++v) {
    VSGM_CHECK(assignment[v] >= 0 && assignment[v] < n_bins, "bin id out of range");
    parts[assignment[v]].push_back((uintV)v);
  }
  pool_.clear();
  for (int i = 0; i < n_bins; ++i) {
    pool_.emplace_back(new ViewBin(i, graph_, hop_, threads_));
    pool_.back()->SetSources(std::move(parts[i]));
  }
}

void ViewBinManager::LoadViewBinPartition(const std::string& filename, const std::vector<uintV>* new_id) {
  std::ifstream in(filename);
  VSGM_CHECK(in.good(), "cannot open view-bin partition " + filename);
  int n_bins = 0;
  in >> n_bins;
  VSGM_CHECK(n_bins > 0, "bad bin count in " + filename);
  std::vector<int> assignment(graph_->vertex_count);
  for (auto& a : assignment) {
    in >> a;
    VSGM_CHECK(!in.fail(), "short partition file " + filename);
  }
  if (new_id) {
    VSGM_CHECK(new_id->size() == assignment.size(), "id map does not match the graph");
    std::vector<int> mapped(assignment.size());
    for (size_t v = 0; v < assignment.size(); ++v) mapped[(*new_id)[v]] = assignment[v];
    assignment.swap(mapped);
  }
  SetPartition(n_bins, assignment);
}

uint64_t ViewBinManager::GetMaxPartitionedSourcesNum() const {
  uint64_t m = 0;
  for (auto& b : pool_) m = std::max<uint64_t>(m, b->GetSources().size());
  return m;
}

void ViewBinManager::Reorder() {
  // std::sort with the reference's comparator (view_bin_manager.h:136-142),
  // so bins of equal cost keep the reference's order too
  std::sort(pool_.begin(), pool_.end(), [](const auto& a, const auto& b) { return a->GetComputationFactors() > b->GetComputationFactors(); });
}

void ViewBinManager::Split(int device_num, int times) {
  int n = (int)pool_.size();
  int target = (n / device_num + times) * device_num;
  int splits = target - n;
  while (splits-- > 0) {
    uint64_t sum = 0;
    for (auto& b : pool_) sum += b->GetComputationFactors();
    uint64_t avg = sum / pool_.size();
    // the bin furthest above the average is split in two, alternating its
    // sources by descending degree
    uint64_t best_diff = 0;
    size_t best = 0;
    for (size_t i = 0; i < pool_.size(); ++i) {
      uint64_t f = pool_[i]->GetComputationFactors();
      if (f < avg) continue;
      if (f - avg > best_diff) {
        best_diff = f - avg;
        best = i;
      }
    }
    std::vector<std::pair<uintV, uint64_t>> by_degree;
    for (auto s : pool_[best]->GetSources()) by_degree.emplace_back(s, graph_->Degree(s));
    // std::sort as the reference sorts (view_bin_manager.h:100), so sources of
    // equal degree split the same way
    std::sort(by_degree.begin(), by_degree.end(), [](const auto& a, const auto& b) { return a.second > b.second; });
    std::vector<uintV> odd, even;
    for (size_t i = 0; i < by_degree.size(); ++i) (i % 2 ? odd : even).push_back(by_degree[i].first);
    pool_.erase(pool_.begin() + best);
    // ids above the removed position shift down (view_bin_manager.h:111-115
    // compares ids with the position, as here)
    for (auto& b : pool_)
      if (b->GetId() > (int)best) b->SetId(b->GetId() - 1);
    int id = (int)pool_.size();
    pool_.emplace_back(new ViewBin(id, graph_, hop_, threads_));
    pool_.back()->SetSources(std::move(odd));
    pool_.emplace_back(new ViewBin(id + 1, graph_, hop_, threads_));
    pool_.back()->SetSources(std::move(even));
  }
  Reorder();
}

void WriteViewBinPartition(const std::string& filename, int n_bins, const std::vector<int>& assignment) {
  std::ofstream out(filename);
  VSGM_CHECK(out.good(), "cannot create " + filename);
  out << n_bins << "\n";
  for (auto a : assignment) out << a << "\n";
}

std::string ViewBinPartitionName(const std::string& graph_file, int hop) { return graph_file + "." + std::to_string(hop) + ".hop.vbmap"; }

}  // namespace vsgm
