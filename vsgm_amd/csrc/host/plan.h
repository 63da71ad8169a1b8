// Execution plan: matching order, per-level operator sequence and the
// conditions applied at each operator.
//
// Restates the reference's heuristic planner (src/query/plan.h:61-580) and
// the per-vertex condition tables it feeds to the device
// (src/query/dev_plan.h:130-146). Plan::FromOrder builds the same kind of
// plan from any connected matching order; the learned planner uses it.
#pragma once

#include <string>
#include <vector>

#include "common.h"
#include "query.h"

namespace vsgm {

enum TraversalOperation {
  COMPUTE = 0,
  COMPUTE_PATH_COUNT = 1,
  MATERIALIZE = 2,
  FILTER_COMPUTE = 3,
  COUNT = 4,
  COMPUTE_COUNT = 5,
};
const char* TraversalOperationName(TraversalOperation op);

typedef std::pair<TraversalOperation, uintV> TraversalEntry;
typedef std::vector<uintV> VTGroup;
typedef std::vector<VTGroup> MultiVTGroup;

class Plan {
 public:
  explicit Plan(const Query& query);

  // Heuristic optimisation (plan.h:548-561).
  void Optimize();
  // Same pipeline with a caller-chosen matching order (a permutation of the
  // pattern vertices whose every prefix is connected); seq[0] is the root.
  static Plan FromOrder(const Query& query, const std::vector<uintV>& seq);

  const Query& GetQuery() const { return query_; }
  size_t GetVertexCount() const { return vertex_count_; }
  size_t GetEdgeCount() const { return edge_count_; }
  int GetHop() const { return hop_; }
  size_t GetRootDegree() const { return root_degree_; }
  uintV GetRoot() const { return root_; }
  const std::vector<uintV>& GetSearchSequence() const { return seq_; }
  const std::vector<TraversalEntry>& GetExecuteOperations() const { return exec_seq_; }
  const AllConnType& GetBackwardConnectivity() const { return back_conn_; }
  const MultiVTGroup& GetMaterializedVertices() const { return materialized_vertices_; }
  const MultiVTGroup& GetComputedUnmaterializedVertices() const { return computed_unmaterialized_vertices_; }
  // Per pattern vertex condition lists (dev_plan.h:292-309).
  const AllCondType& GetComputeConditions() const { return computed_cond_; }
  const AllCondType& GetMaterializeConditions() const { return materialized_cond_; }
  const AllCondType& GetFilterConditions() const { return filter_cond_; }
  CondOperator ConditionBetween(uintV a, uintV b) const { return GetConditionType(a, b, order_); }

  // Textual dump used by the planner parity tests (one key per line).
  std::string Serialize() const;
  std::string Describe() const;

  // BFS-level depth from `root` that covers every pattern edge
  // (plan.h:156-183).
  int GetMinHopWithBFS(uintV root) const;
  int GetEdgeCoverHop(uintV root) const;  // plan.h:122-154

 private:
  void FindRoot();                             // plan.h:185-208
  void GenerateSearchSequence(uintV root);     // plan.h:210-265
  void GenerateBackwardConnectivity();         // plan.h:267-290
  void GenerateComputeExecuteSequence();       // plan.h:292-318
  void GenerateExecuteSequence();              // plan.h:383-430
  void GenerateCount1(const VTGroup& un, const std::vector<bool>& need_filter);  // plan.h:329-343
  void GenerateCount2(const VTGroup& un, const std::vector<bool>& need_filter);  // plan.h:345-381
  void GenerateConstraints();                  // plan.h:432-452
  void GenerateConditions();                   // plan.h:460-528
  size_t GetOperationIndex(TraversalOperation op, uintV u) const;
  void Finish();

  Query query_;
  size_t vertex_count_;
  size_t edge_count_;
  AllConnType conn_;
  AllConnType back_conn_;
  AllCondType order_;

  int hop_ = 0;
  uintV root_ = 0;
  size_t root_degree_ = 0;

  std::vector<uintV> seq_;
  std::vector<TraversalEntry> exec_seq_;
  MultiVTGroup materialized_vertices_;
  MultiVTGroup computed_unmaterialized_vertices_;
  AllCondType computed_cond_, materialized_cond_, filter_cond_;
};

// True if `seq` is a permutation of the pattern vertices with every prefix
// connected.
bool IsConnectedOrder(const Query& q, const std::vector<uintV>& seq);

}  // namespace vsgm
