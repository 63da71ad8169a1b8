// Query model: preset patterns, k-cliques, k-graphlets and arbitrary
// connected patterns, with the symmetry-breaking partial order.
//
// Behaviour follows the reference's src/query/query.h, pattern.h, clique.h,
// graphlet.h and partial_order.h. The automorphism group is enumerated by
// brute force over all vertex permutations (patterns have at most 7
// vertices, so at most 5040 permutations) instead of with bliss; the
// conditions derived from it depend only on the group as a set, so they are
// the same as the reference's (partial_order.h:84-144 sorts and dedups them).
#pragma once

#include <string>
#include <vector>

#include "common.h"

namespace vsgm {

// Operator between pattern vertices a and b, NON_EQUAL when the order
// lists no relation (query.h:43-53).
CondOperator GetConditionType(uintV a, uintV b, const AllCondType& order);
const char* CondOperatorName(CondOperator op);

// All automorphisms of an undirected pattern (adjacency lists), identity
// included, in lexicographic permutation order.
std::vector<std::vector<uintV>> PatternAutomorphisms(const AllConnType& conn);

// Symmetry-breaking pairs (first < second must hold), sorted and unique
// (partial_order.h:96-144).
std::vector<std::pair<uintV, uintV>> SymmetryConditions(const AllConnType& conn);

// True if the two patterns are isomorphic (replaces bliss canonical-form
// comparison of graphlet.h:280-295).
bool PatternsIsomorphic(const AllConnType& a, const AllConnType& b);

class Query {
 public:
  Query() : type_(PatternType), vertex_count_(0), edge_count_(0) {}

  // Preset pattern P0..P18 (pattern.h:9-383); pm_pipeline builds these
  // with hard_code_ordering=false and enable_ordering=true.
  static Query Preset(int preset, bool enable_ordering = true);
  // k-clique with the hard-coded chain 0<1<...<k-1 (clique.h:11-37).
  static Query Clique(int k);
  // Generic connected pattern from adjacency lists (query.h:62-78): the ids
  // are remapped and the symmetry conditions installed.
  static Query FromAdjacency(const std::vector<std::vector<int>>& adj, QueryType type = GraphletType);
  // All connected k-vertex graphlets, 3 <= k <= 5 (graphlet.h:15-136).
  static std::vector<Query> Graphlets(int k);

  QueryType GetQueryType() const { return type_; }
  const char* GetQueryTypeString() const;
  size_t GetVertexCount() const { return vertex_count_; }
  size_t GetEdgeCount() const { return edge_count_; }  // undirected edges
  const AllConnType& GetConnectivity() const { return conn_; }
  const AllCondType& GetOrder() const { return order_; }
  int GetPreset() const { return preset_; }
  std::string Describe() const;

 private:
  void ReMapVertexId();                                          // query.h:123-185
  void SetConditions(const std::vector<std::pair<uintV, uintV>>& c);  // query.h:195-203
  void SetAdjacencyFromEdges(int n, const std::vector<std::pair<int, int>>& edges);

  QueryType type_;
  int preset_ = -1;
  size_t vertex_count_;
  size_t edge_count_;
  AllConnType conn_;
  AllCondType order_;
};

}  // namespace vsgm
