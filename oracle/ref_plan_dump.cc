// Fixture generator over the REFERENCE planner (test infrastructure only).
//
// Built by oracle/Makefile against the reference's own headers
// (src/query/{query,pattern,clique,graphlet,plan}.h) and its vendored bliss
// sources, where they lie under /root/reference; output goes to
// oracle/_ref/. It reads one query spec per line from stdin:
//   preset <k>        Pattern((PresetPatternType)k, false)   (pattern_matching_pipeline.cu:69)
//   clique <k>        Clique(k)                              (pattern_matching_pipeline.cu:61)
//   graphlets <k>     every query of Graphlet(k)             (pattern_matching_pipeline.cu:65)
//   adj <n> <bits>    Query(n, m, adj) with an n*n 0/1 matrix in row-major order
// and writes, to the file named by argv[1], one record per query in the
// key/value format of vsgm_amd's Plan::Serialize, after "query: <spec>".
// <algorithm> first: the reference's partial_order.h relies on it being
// included before it (query.h does so).
#include <algorithm>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "query/clique.h"
#include "query/graphlet.h"
#include "query/pattern.h"
#include "query/plan.h"
#include "query/query.h"

static const char* OpName(CondOperator op) {
  switch (op) {
    case LESS_THAN: return "LESS_THAN";
    case LARGER_THAN: return "LARGER_THAN";
    case NON_EQUAL: return "NON_EQUAL";
    default: return "NONE";
  }
}

static void Group(std::ostream& os, const std::vector<uintV>& g) {
  os << "[";
  for (size_t i = 0; i < g.size(); ++i) os << (i ? "," : "") << g[i];
  os << "]";
}

static void Conds(std::ostream& os, const std::vector<CondType>& c) {
  os << "[";
  for (size_t i = 0; i < c.size(); ++i) os << (i ? "," : "") << OpName(c[i].first) << "(" << c[i].second << ")";
  os << "]";
}

static void Dump(std::ostream& os, const std::string& spec, Query* q) {
  Plan plan(q);
  plan.Optimize();
  os << "query: " << spec << "\n";
  os << "conn:";
  for (auto& l : q->GetConnectivity()) Group(os, l);
  os << "\norder:";
  AllCondType order = q->GetOrder();
  order.resize(q->GetVertexCount());
  for (auto& l : order) Conds(os, l);
  os << "\nhop:" << plan.GetHop() << "\nroot_degree:" << plan.GetRootDegree() << "\nexec:";
  for (auto& e : plan.GetExecuteOperations()) os << "(" << GetTraversalOperationString(e.first) << "," << e.second << ")";
  os << "\nmaterialized:";
  for (auto& g : plan.GetMaterializedVertices()) Group(os, g);
  os << "\ncomputed_unmaterialized:";
  for (auto& g : plan.GetComputedUnmaterializedVertices()) Group(os, g);
  os << "\nbackward:";
  for (auto& g : plan.GetBackwardConnectivity()) Group(os, g);
  AllCondType cc, mc, fc;
  plan.GetComputeCondition(cc);
  plan.GetMaterializeCondition(mc);
  plan.GetFilterCondition(fc);
  os << "\ncompute_cond:";
  for (auto& c : cc) Conds(os, c);
  os << "\nmaterialize_cond:";
  for (auto& c : mc) Conds(os, c);
  os << "\nfilter_cond:";
  for (auto& c : fc) Conds(os, c);
  os << "\n";
}

int main(int argc, char** argv) {
  if (argc != 2) {
    std::cerr << "usage: ref_plan_dump <output file> < specs" << std::endl;
    return 2;
  }
  std::ofstream out(argv[1]);
  std::string line;
  while (std::getline(std::cin, line)) {
    std::istringstream is(line);
    std::string kind;
    if (!(is >> kind)) continue;
    if (kind == "preset") {
      int k;
      is >> k;
      Pattern p((PresetPatternType)k, false);
      Dump(out, line, &p);
    } else if (kind == "clique") {
      int k;
      is >> k;
      Clique c(k);
      Dump(out, line, &c);
    } else if (kind == "graphlets") {
      int k;
      is >> k;
      Graphlet g(k);
      const auto& qs = g.GetQueries();
      for (size_t i = 0; i < qs.size(); ++i) Dump(out, line + " #" + std::to_string(i), qs[i]);
    } else if (kind == "adj") {
      int n;
      std::string bits;
      is >> n >> bits;
      std::vector<std::vector<int>> adj(n);
      size_t directed = 0;
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
          if (bits[i * n + j] == '1') {
            adj[i].push_back(j);
            ++directed;
          }
      Query q(n, directed / 2, adj);
      Dump(out, line, &q);
    }
  }
  return 0;
}
