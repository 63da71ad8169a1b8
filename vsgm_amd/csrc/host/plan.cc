// Planner implementation. See plan.h for the reference mapping.
#include "plan.h"

#include <algorithm>
#include <queue>
#include <set>
#include <sstream>
#include <tuple>

namespace vsgm {

const char* TraversalOperationName(TraversalOperation op) {
  switch (op) {
    case COMPUTE: return "COMPUTE";
    case COMPUTE_PATH_COUNT: return "COMPUTE_PATH_COUNT";
    case MATERIALIZE: return "MATERIALIZE";
    case FILTER_COMPUTE: return "FILTER_COMPUTE";
    case COUNT: return "COUNT";
    case COMPUTE_COUNT: return "COMPUTE_COUNT";
  }
  return "";
}

Plan::Plan(const Query& query) : query_(query) {
  vertex_count_ = query.GetVertexCount();
  edge_count_ = query.GetEdgeCount();
  conn_ = query.GetConnectivity();
  order_ = query.GetOrder();
  order_.resize(vertex_count_);
}

int Plan::GetEdgeCoverHop(uintV root) const {
  // BFS layer by layer; stop when every undirected edge has been touched.
  std::vector<char> seen(vertex_count_, 0);
  std::set<std::pair<uintV, uintV>> touched;
  std::vector<uintV> layer{root};
  seen[root] = 1;
  int hop = 0;
  while (!layer.empty()) {
    ++hop;
    std::vector<uintV> nxt;
    for (auto v : layer)
      for (auto w : conn_[v]) {
        if (!seen[w]) {
          seen[w] = 1;
          nxt.push_back(w);
        }
        touched.insert(std::make_pair(std::min(v, w), std::max(v, w)));
      }
    if (touched.size() == edge_count_) break;
    layer.swap(nxt);
  }
  return hop;
}

int Plan::GetMinHopWithBFS(uintV root) const {
  // Depth of the deepest BFS level, plus one if an edge joins two vertices
  // of that deepest level.
  std::vector<int> level(vertex_count_, -1);
  std::queue<uintV> q;
  q.push(root);
  level[root] = 0;
  int deepest_vertex = -1, deepest_same_level_edge = -1;
  while (!q.empty()) {
    uintV v = q.front();
    q.pop();
    deepest_vertex = level[v];
    for (auto w : conn_[v]) {
      if (level[w] < 0) {
        level[w] = level[v] + 1;
        q.push(w);
      } else if (level[w] == level[v]) {
        deepest_same_level_edge = level[w];
      }
    }
  }
  return deepest_vertex + (deepest_vertex == deepest_same_level_edge ? 1 : 0);
}

void Plan::FindRoot() {
  root_ = 0;
  hop_ = 10;
  root_degree_ = 0;
  for (uintV u = 0; u < vertex_count_; ++u) {
    int h = GetMinHopWithBFS(u);
    if (h < hop_ || (h == hop_ && root_degree_ < conn_[u].size())) {
      root_ = u;
      hop_ = h;
      root_degree_ = conn_[u].size();
    }
  }
}

void Plan::GenerateSearchSequence(uintV root) {
  // (BFS depth asc, degree desc, id asc)
  std::vector<std::tuple<int, int, uintV>> key(vertex_count_, std::make_tuple(0, 0, (uintV)0));
  std::vector<char> seen(vertex_count_, 0);
  seen[root] = 1;
  key[root] = std::make_tuple(0, (int)conn_[root].size(), root);
  std::vector<uintV> layer{root};
  int depth = 0;
  while (!layer.empty()) {
    ++depth;
    std::vector<uintV> nxt;
    for (auto v : layer)
      for (auto w : conn_[v])
        if (!seen[w]) {
          seen[w] = 1;
          nxt.push_back(w);
          key[w] = std::make_tuple(depth, (int)conn_[w].size(), w);
        }
    layer.swap(nxt);
  }
  std::sort(key.begin(), key.end(), [](const auto& a, const auto& b) {
    if (std::get<0>(a) != std::get<0>(b)) return std::get<0>(a) < std::get<0>(b);
    if (std::get<1>(a) != std::get<1>(b)) return std::get<1>(a) > std::get<1>(b);
    return std::get<2>(a) < std::get<2>(b);
  });
  seq_.clear();
  for (auto& k : key) seq_.push_back(std::get<2>(k));
}

void Plan::GenerateBackwardConnectivity() {
  back_conn_.assign(vertex_count_, {});
  std::vector<char> placed(vertex_count_, 0);
  for (auto u : seq_) {
    for (auto w : conn_[u])
      if (placed[w]) back_conn_[u].push_back(w);
    placed[u] = 1;
  }
}

void Plan::GenerateComputeExecuteSequence() {
  exec_seq_.clear();
  std::vector<char> mat(vertex_count_, 0);
  exec_seq_.emplace_back(MATERIALIZE, seq_[0]);
  mat[seq_[0]] = 1;
  for (size_t i = 1; i < seq_.size(); ++i) {
    uintV u = seq_[i];
    for (auto w : back_conn_[u])
      if (!mat[w]) {
        exec_seq_.emplace_back(MATERIALIZE, w);
        mat[w] = 1;
      }
    exec_seq_.emplace_back(COMPUTE, u);
  }
}

size_t Plan::GetOperationIndex(TraversalOperation op, uintV u) const {
  for (size_t i = 0; i < exec_seq_.size(); ++i)
    if (exec_seq_[i].first == op && exec_seq_[i].second == u) return i;
  return exec_seq_.size();
}

void Plan::GenerateCount1(const VTGroup& un, const std::vector<bool>& need_filter) {
  if (need_filter[0]) {
    exec_seq_.emplace_back(FILTER_COMPUTE, un[0]);
    exec_seq_.emplace_back(COUNT, (uintV)vertex_count_);
  } else {
    auto& last = exec_seq_.back();
    VSGM_CHECK(last.first == COMPUTE && last.second == un[0], "count-1 tail must be the COMPUTE of the last vertex");
    last.first = COMPUTE_COUNT;
  }
}

void Plan::GenerateCount2(const VTGroup& un, const std::vector<bool>& need_filter) {
  if (need_filter[0] && need_filter[1]) {
    exec_seq_.emplace_back(FILTER_COMPUTE, un[0]);
    exec_seq_.emplace_back(FILTER_COMPUTE, un[1]);
    exec_seq_.emplace_back(COUNT, (uintV)vertex_count_);
  } else if (!need_filter[0] && !need_filter[1]) {
    // the earliest operation naming each vertex; the later of the two must
    // be the tail and becomes COMPUTE_COUNT
    size_t i0 = exec_seq_.size(), i1 = exec_seq_.size();
    for (int j = (int)exec_seq_.size() - 1; j >= 0; --j) {
      if (exec_seq_[j].second == un[0]) i0 = j;
      if (exec_seq_[j].second == un[1]) i1 = j;
    }
    size_t tail = std::max(i0, i1);
    VSGM_CHECK(tail == exec_seq_.size() - 1, "count-2 tail must be the last operation");
    exec_seq_[tail].first = COMPUTE_COUNT;
  } else {
    uintV filtered = need_filter[0] ? un[0] : un[1];
    uintV direct = need_filter[0] ? un[1] : un[0];
    size_t idx = GetOperationIndex(COMPUTE, direct);
    exec_seq_.erase(exec_seq_.begin() + idx);
    exec_seq_.emplace_back(FILTER_COMPUTE, filtered);
    exec_seq_.emplace_back(COMPUTE_COUNT, direct);
  }
}

void Plan::GenerateExecuteSequence() {
  VTGroup un;
  std::vector<bool> need_filter;
  std::vector<char> mat(vertex_count_, 0);
  for (auto& e : exec_seq_)
    if (e.first == MATERIALIZE) mat[e.second] = 1;
  for (uintV u = 0; u < vertex_count_; ++u) {
    if (mat[u]) continue;
    un.push_back(u);
    size_t at = GetOperationIndex(COMPUTE, u);
    bool f = false;
    for (size_t i = at + 1; i < exec_seq_.size(); ++i)
      if (exec_seq_[i].first == MATERIALIZE) f = true;
    need_filter.push_back(f);
  }
  if (un.size() == 1) {
    GenerateCount1(un, need_filter);
  } else if (un.size() == 2) {
    GenerateCount2(un, need_filter);
  } else {
    for (size_t i = 2; i < un.size(); ++i) exec_seq_.emplace_back(MATERIALIZE, un[i]);
    VTGroup first_two(un.begin(), un.begin() + 2);
    GenerateCount2(first_two, std::vector<bool>(2, true));
  }
}

void Plan::GenerateConstraints() {
  std::vector<char> mat(vertex_count_, 0);
  VTGroup all_mat, all_comp;
  materialized_vertices_.assign(exec_seq_.size(), {});
  computed_unmaterialized_vertices_.assign(exec_seq_.size(), {});
  for (size_t i = 0; i < exec_seq_.size(); ++i) {
    materialized_vertices_[i] = all_mat;
    for (auto u : all_comp)
      if (!mat[u]) computed_unmaterialized_vertices_[i].push_back(u);
    auto op = exec_seq_[i].first;
    if (op == COMPUTE || op == COMPUTE_PATH_COUNT) {
      all_comp.push_back(exec_seq_[i].second);
    } else if (op == MATERIALIZE) {
      all_mat.push_back(exec_seq_[i].second);
      mat[exec_seq_[i].second] = 1;
    }
  }
}

void Plan::GenerateConditions() {
  computed_cond_.assign(vertex_count_, {});
  materialized_cond_.assign(vertex_count_, {});
  filter_cond_.assign(vertex_count_, {});
  // compute: against every vertex materialized before the op (plan.h:460-476)
  VTGroup mat_so_far;
  for (auto& e : exec_seq_) {
    if (e.first == COMPUTE || e.first == COMPUTE_COUNT || e.first == COMPUTE_PATH_COUNT) {
      auto& c = computed_cond_[e.second];
      c.clear();
      for (auto w : mat_so_far) c.emplace_back(GetConditionType(e.second, w, order_), w);
    } else if (e.first == MATERIALIZE) {
      mat_so_far.push_back(e.second);
    }
  }
  // materialize / filter: against the vertices materialized between the
  // COMPUTE of u and this op, nearest first (plan.h:478-528)
  for (size_t i = 0; i < exec_seq_.size(); ++i) {
    auto op = exec_seq_[i].first;
    if (op != MATERIALIZE && op != FILTER_COMPUTE) continue;
    uintV u = exec_seq_[i].second;
    auto& c = (op == MATERIALIZE) ? materialized_cond_[u] : filter_cond_[u];
    c.clear();
    for (int j = (int)i - 1; j >= 0; --j) {
      if (exec_seq_[j].first == COMPUTE && exec_seq_[j].second == u) break;
      if (exec_seq_[j].first == MATERIALIZE) c.emplace_back(GetConditionType(u, exec_seq_[j].second, order_), exec_seq_[j].second);
    }
  }
}

void Plan::Finish() {
  GenerateBackwardConnectivity();
  GenerateComputeExecuteSequence();
  GenerateExecuteSequence();
  GenerateConstraints();
  GenerateConditions();
}

void Plan::Optimize() {
  FindRoot();
  GenerateSearchSequence(root_);
  Finish();
}

bool IsConnectedOrder(const Query& q, const std::vector<uintV>& seq) {
  size_t n = q.GetVertexCount();
  if (seq.size() != n) return false;
  std::vector<char> placed(n, 0);
  for (size_t i = 0; i < n; ++i) {
    uintV u = seq[i];
    if (u >= n || placed[u]) return false;
    if (i > 0) {
      bool linked = false;
      for (auto w : q.GetConnectivity()[u]) linked |= placed[w] != 0;
      if (!linked) return false;
    }
    placed[u] = 1;
  }
  return true;
}

Plan Plan::FromOrder(const Query& query, const std::vector<uintV>& seq) {
  VSGM_CHECK(IsConnectedOrder(query, seq), "matching order must be a connected permutation");
  Plan p(query);
  p.root_ = seq[0];
  // The view must hold the adjacency list of every pattern vertex that an
  // intersection reads, i.e. every vertex with a neighbour later in the
  // order: hop >= 1 + its BFS depth from the root. A breadth-first order
  // (the heuristic's) needs no more than GetMinHopWithBFS; other orders
  // can read deeper lists.
  std::vector<int> depth(p.vertex_count_, -1);
  std::queue<uintV> bfs;
  bfs.push(p.root_);
  depth[p.root_] = 0;
  while (!bfs.empty()) {
    uintV v = bfs.front();
    bfs.pop();
    for (auto w : p.conn_[v])
      if (depth[w] < 0) {
        depth[w] = depth[v] + 1;
        bfs.push(w);
      }
  }
  int need = 0;
  for (size_t i = 0; i < seq.size(); ++i)
    for (size_t j = i + 1; j < seq.size(); ++j)
      if (std::find(p.conn_[seq[i]].begin(), p.conn_[seq[i]].end(), seq[j]) != p.conn_[seq[i]].end())
        need = std::max(need, depth[seq[i]] + 1);
  p.hop_ = std::max(p.GetMinHopWithBFS(p.root_), need);
  p.root_degree_ = p.conn_[p.root_].size();
  p.seq_ = seq;
  p.Finish();
  return p;
}

std::string Plan::Serialize() const {
  std::ostringstream os;
  auto group = [&](const VTGroup& g) {
    os << "[";
    for (size_t i = 0; i < g.size(); ++i) os << (i ? "," : "") << g[i];
    os << "]";
  };
  auto conds = [&](const std::vector<CondType>& c) {
    os << "[";
    for (size_t i = 0; i < c.size(); ++i) os << (i ? "," : "") << CondOperatorName(c[i].first) << "(" << c[i].second << ")";
    os << "]";
  };
  os << "conn:";
  for (auto& l : conn_) group(l);
  os << "\norder:";
  for (auto& l : order_) conds(l);
  os << "\nhop:" << hop_ << "\nroot_degree:" << root_degree_ << "\nexec:";
  for (auto& e : exec_seq_) os << "(" << TraversalOperationName(e.first) << "," << e.second << ")";
  os << "\nmaterialized:";
  for (auto& g : materialized_vertices_) group(g);
  os << "\ncomputed_unmaterialized:";
  for (auto& g : computed_unmaterialized_vertices_) group(g);
  os << "\nbackward:";
  for (auto& g : back_conn_) group(g);
  os << "\ncompute_cond:";
  for (auto& c : computed_cond_) conds(c);
  os << "\nmaterialize_cond:";
  for (auto& c : materialized_cond_) conds(c);
  os << "\nfilter_cond:";
  for (auto& c : filter_cond_) conds(c);
  os << "\n";
  return os.str();
}

std::string Plan::Describe() const { return query_.Describe() + Serialize(); }

}  // namespace vsgm
