// View bins: a partition of the source vertices into bins, and for each bin
// a k-hop view CSR over the full vertex range.
//
// Follows the reference's src/view/view_bin_manager.h:30-142 (.vbmap
// parsing, Split, Reorder), src/view/view_bin.h:49-142 (materialisation)
// and src/view/view_bin_holder.h / view_bin_buffer.h (pinned holders).
// A view keeps the full adjacency list of every vertex within hop-1 of a
// source whose degree reaches root_degree, and an empty list for every
// other vertex. The reference's shared-"visited" DFS computes exactly that
// ball for hop <= 3 (view_bin_holder.h:55); here it is computed by a
// level-synchronous multi-source BFS, which is the same set for every hop.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "common.h"
#include "graph.h"

namespace vsgm {

// Pinned host buffers that hold one materialised view bin.
class ViewBinHolder {
 public:
  ViewBinHolder(uint64_t vertex_count, uint64_t max_view_bin_bytes);
  ~ViewBinHolder();
  ViewBinHolder(const ViewBinHolder&) = delete;
  ViewBinHolder& operator=(const ViewBinHolder&) = delete;

  int GetId() const { return id_; }
  const std::vector<uintV>& GetSources() const { return sources_; }
  uint64_t GetSourcesNum() const { return sources_.size(); }
  uint64_t GetTotalSize() const { return total_size_; }  // entries in cols
  uint64_t GetVerticesNum() const { return vertices_num_; }
  uint64_t GetVertexCount() const { return vertex_count_; }
  const uintE* GetRowPtrs() const { return row_ptrs_; }
  const uintV* GetCols() const { return cols_; }
  bool IsPinned() const { return pinned_; }

 private:
  friend class ViewBin;
  uint64_t vertex_count_;
  uint64_t capacity_entries_;
  int id_ = -1;
  std::vector<uintV> sources_;
  uint64_t total_size_ = 0;
  uint64_t vertices_num_ = 0;
  uintE* row_ptrs_ = nullptr;
  uintV* cols_ = nullptr;
  bool pinned_ = false;
};

class ViewBin {
 public:
  ViewBin(int id, const Graph* g, int hop, int threads) : id_(id), graph_(g), hop_(hop), threads_(threads) {}
  int GetId() const { return id_; }
  void SetId(int id) { id_ = id; }
  const std::vector<uintV>& GetSources() const { return sources_; }
  void SetSources(std::vector<uintV> s) { sources_ = std::move(s); }
  // Sum of the sources' degrees (view_bin.h:38-47).
  uint64_t GetComputationFactors() const;
  // Fill the holder with this bin's view (view_bin.h:66-142).
  void Materialize(ViewBinHolder& holder, uint64_t root_degree = 0) const;

 private:
  int id_;
  const Graph* graph_;
  int hop_;
  int threads_;
  std::vector<uintV> sources_;
};

class ViewBinManager {
 public:
  ViewBinManager(const Graph* g, int hop, int threads) : graph_(g), hop_(hop), threads_(threads) {}
  // Text file: bin count, then one bin id per vertex (view_bin_manager.h:30-67).
  // With new_id, line v of the file is the bin of vertex new_id[v] (a
  // partition written for the graph before RelabelByDegree).
  void LoadViewBinPartition(const std::string& filename, const std::vector<uintV>* new_id = nullptr);
  void SetPartition(int n_bins, const std::vector<int>& assignment);
  void Split(int device_num, int times);  // view_bin_manager.h:69-134
  void Reorder();                         // view_bin_manager.h:136-142
  int GetViewBinNumber() const { return (int)pool_.size(); }
  uint64_t GetMaxPartitionedSourcesNum() const;
  const std::vector<std::unique_ptr<ViewBin>>& GetViewBinPool() const { return pool_; }

 private:
  const Graph* graph_;
  int hop_;
  int threads_;
  std::vector<std::unique_ptr<ViewBin>> pool_;
};

void WriteViewBinPartition(const std::string& filename, int n_bins, const std::vector<int>& assignment);
// "<graph>.<hop>.hop.vbmap" (pattern_matching_pipeline.cu:94).
std::string ViewBinPartitionName(const std::string& graph_file, int hop);

}  // namespace vsgm
