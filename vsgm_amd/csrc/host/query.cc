// Query model implementation. See query.h for the reference mapping.
#include "query.h"

#include <algorithm>
#include <map>
#include <numeric>
#include <queue>
#include <set>
#include <sstream>
#include <tuple>

namespace vsgm {

CondOperator GetConditionType(uintV a, uintV b, const AllCondType& order) {
  // First entry of a's list that names b decides; default NON_EQUAL.
  if (a < order.size())
    for (const auto& c : order[a])
      if (c.second == b) return c.first;
  return NON_EQUAL;
}

const char* CondOperatorName(CondOperator op) {
  switch (op) {
    case LESS_THAN: return "LESS_THAN";
    case LARGER_THAN: return "LARGER_THAN";
    case NON_EQUAL: return "NON_EQUAL";
    default: return "NONE";
  }
}

namespace {
std::vector<std::vector<char>> AdjMatrix(const AllConnType& conn) {
  size_t n = conn.size();
  std::vector<std::vector<char>> m(n, std::vector<char>(n, 0));
  for (size_t i = 0; i < n; ++i)
    for (auto j : conn[i]) m[i][j] = m[j][i] = 1;
  return m;
}

// Orbit partition of the vertex set under the permutation set `perms`,
// keyed by the smallest member (partial_order.h:84-94).
std::map<uintV, std::set<uintV>> Orbits(const std::vector<std::vector<uintV>>& perms, size_t n) {
  std::map<uintV, std::set<uintV>> orbits;
  for (size_t i = 0; i < n; ++i) {
    std::set<uintV> images;
    for (const auto& p : perms) images.insert(p[i]);
    if (images.empty()) continue;
    orbits[*images.begin()].insert(images.begin(), images.end());
  }
  return orbits;
}
}  // namespace

std::vector<std::vector<uintV>> PatternAutomorphisms(const AllConnType& conn) {
  size_t n = conn.size();
  auto adj = AdjMatrix(conn);
  std::vector<uintV> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::vector<std::vector<uintV>> result;
  do {
    bool ok = true;
    for (size_t i = 0; i < n && ok; ++i)
      for (size_t j = i + 1; j < n && ok; ++j)
        if (adj[i][j] != adj[perm[i]][perm[j]]) ok = false;
    if (ok) result.push_back(perm);
  } while (std::next_permutation(perm.begin(), perm.end()));
  return result;
}

std::vector<std::pair<uintV, uintV>> SymmetryConditions(const AllConnType& conn) {
  size_t n = conn.size();
  auto group = PatternAutomorphisms(conn);
  std::vector<std::pair<uintV, uintV>> result;
  auto first_nontrivial = [](const std::map<uintV, std::set<uintV>>& orbits) {
    return std::find_if(orbits.begin(), orbits.end(), [](const auto& e) { return e.second.size() > 1; });
  };
  auto orbits = Orbits(group, n);
  auto it = first_nontrivial(orbits);
  while (it != orbits.end()) {
    const std::set<uintV> orbit = it->second;
    uintV n0 = *orbit.begin();
    for (const auto& p : group) {
      // the orbit member (other than n0) with the smallest image under p
      uintV best = 0;
      bool have = false;
      for (auto x : orbit) {
        if (x == n0) continue;
        if (!have || p[x] < p[best]) {
          best = x;
          have = true;
        }
      }
      result.emplace_back(n0, best);
    }
    // keep the stabiliser of n0
    group.erase(std::remove_if(group.begin(), group.end(), [n0](const std::vector<uintV>& p) { return p[n0] != n0; }), group.end());
    orbits = Orbits(group, n);
    it = first_nontrivial(orbits);
  }
  std::sort(result.begin(), result.end());
  result.erase(std::unique(result.begin(), result.end()), result.end());
  return result;
}

bool PatternsIsomorphic(const AllConnType& a, const AllConnType& b) {
  if (a.size() != b.size()) return false;
  size_t n = a.size();
  auto ma = AdjMatrix(a), mb = AdjMatrix(b);
  size_t ea = 0, eb = 0;
  std::vector<size_t> da(n, 0), db(n, 0);
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) {
      da[i] += ma[i][j];
      db[i] += mb[i][j];
    }
  for (size_t i = 0; i < n; ++i) {
    ea += da[i];
    eb += db[i];
  }
  if (ea != eb) return false;
  auto sa = da, sb = db;
  std::sort(sa.begin(), sa.end());
  std::sort(sb.begin(), sb.end());
  if (sa != sb) return false;
  std::vector<uintV> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  do {
    bool ok = true;
    for (size_t i = 0; i < n && ok; ++i)
      for (size_t j = i + 1; j < n && ok; ++j)
        if (ma[i][j] != mb[perm[i]][perm[j]]) ok = false;
    if (ok) return true;
  } while (std::next_permutation(perm.begin(), perm.end()));
  return false;
}

const char* Query::GetQueryTypeString() const {
  switch (type_) {
    case PatternType: return "Pattern";
    case CliqueType: return "Clique";
    case GraphletType: return "Graphlet";
  }
  return "";
}

void Query::SetAdjacencyFromEdges(int n, const std::vector<std::pair<int, int>>& edges) {
  vertex_count_ = n;
  edge_count_ = edges.size();
  conn_.assign(n, {});
  for (auto& e : edges) {
    conn_[e.first].push_back(e.second);
    conn_[e.second].push_back(e.first);
  }
  for (auto& l : conn_) std::sort(l.begin(), l.end());
}

void Query::ReMapVertexId() {
  const size_t n = vertex_count_;
  // root: first vertex of maximum degree
  uintV root = 0;
  size_t best = 0;
  for (uintV i = 0; i < n; ++i)
    if (conn_[i].size() > best) {
      best = conn_[i].size();
      root = i;
    }
  // level-by-level BFS; inside a level order by (degree desc, number of
  // already-discovered neighbours desc, id asc)
  std::vector<char> seen(n, 0);
  std::vector<uintV> frontier{root};
  seen[root] = 1;
  std::vector<uintV> new_id(n, 0);
  uintV next = 0;
  while (!frontier.empty()) {
    std::vector<uintV> upcoming;
    for (auto v : frontier)
      for (auto w : conn_[v])
        if (!seen[w]) {
          seen[w] = 1;
          upcoming.push_back(w);
        }
    std::vector<std::tuple<size_t, size_t, uintV>> keyed;
    for (auto v : frontier) {
      size_t linked = 0;
      for (auto w : conn_[v]) linked += seen[w] ? 1 : 0;
      keyed.emplace_back(conn_[v].size(), linked, v);
    }
    std::sort(keyed.begin(), keyed.end(), [](const auto& a, const auto& b) {
      if (std::get<0>(a) != std::get<0>(b)) return std::get<0>(a) > std::get<0>(b);
      if (std::get<1>(a) != std::get<1>(b)) return std::get<1>(a) > std::get<1>(b);
      return std::get<2>(a) < std::get<2>(b);
    });
    for (auto& k : keyed) new_id[std::get<2>(k)] = next++;
    frontier.swap(upcoming);
  }
  AllConnType remapped(n);
  for (uintV i = 0; i < n; ++i)
    for (auto j : conn_[i]) remapped[new_id[i]].push_back(new_id[j]);
  for (auto& l : remapped) std::sort(l.begin(), l.end());
  conn_ = std::move(remapped);
}

void Query::SetConditions(const std::vector<std::pair<uintV, uintV>>& c) {
  order_.resize(vertex_count_);
  for (auto& p : c) {
    order_[p.first].emplace_back(LESS_THAN, p.second);
    order_[p.second].emplace_back(LARGER_THAN, p.first);
  }
}

Query Query::Preset(int preset, bool enable_ordering) {
  // Edge lists of the preset shapes P0..P18 (pattern.h:36-361).
  static const std::vector<std::pair<int, std::vector<std::pair<int, int>>>> kPresets = {
      {3, {{0, 1}, {0, 2}, {1, 2}}},                                                    // P0 triangle
      {4, {{0, 1}, {0, 3}, {1, 2}, {2, 3}}},                                            // P1 square
      {4, {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {2, 3}}},                                    // P2 chordal square
      {5, {{0, 1}, {0, 2}, {1, 2}, {2, 3}, {3, 4}}},                                    // P3 2-tails triangle
      {5, {{0, 1}, {0, 2}, {1, 2}, {1, 3}, {2, 4}, {3, 4}}},                            // P4 house
      {5, {{0, 1}, {0, 2}, {0, 3}, {0, 4}, {1, 2}, {1, 3}, {2, 3}, {3, 4}}},            // P5 chordal house
      {5, {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 4}, {3, 4}}},                    // P6 chordal roof
      {5, {{0, 1}, {0, 2}, {0, 3}, {0, 4}, {1, 3}, {1, 2}, {2, 4}}},                    // P7 three triangles
      {5, {{0, 1}, {0, 2}, {0, 3}, {0, 4}, {1, 2}, {1, 4}, {2, 3}, {3, 4}}},            // P8 solar square
      {5, {{0, 1}, {0, 2}, {0, 3}, {0, 4}, {1, 2}, {1, 3}, {1, 4}, {2, 3}, {3, 4}}},    // P9 near 5-clique
      {6, {{0, 1}, {0, 2}, {0, 3}, {0, 4}, {0, 5}, {1, 2}, {2, 3}, {3, 4}, {4, 5}}},    // P10 four triangles
      {6, {{0, 1}, {0, 2}, {0, 3}, {0, 5}, {1, 2}, {1, 3}, {1, 4}, {2, 4}, {2, 5}}},    // P11 one in three triangles
      {6, {{0, 1}, {0, 2}, {1, 2}, {1, 3}, {1, 4}, {1, 5}, {2, 3}, {2, 4}, {2, 5}, {3, 4}, {4, 5}}},  // P12 near 6-clique
      {6, {{0, 1}, {0, 2}, {1, 3}, {2, 3}, {2, 4}, {2, 5}, {3, 4}, {3, 5}}},            // P13 square on top
      {7, {{0, 1}, {0, 2}, {0, 3}, {0, 4}, {0, 5}, {1, 2}, {1, 3}, {1, 5}, {2, 3}, {2, 5}, {3, 4}, {3, 5}, {3, 6}, {4, 5}, {5, 6}}},  // P14
      {7, {{0, 1}, {0, 2}, {0, 3}, {0, 4}, {1, 2}, {1, 3}, {1, 4}, {2, 3}, {2, 4}, {3, 4}, {3, 5}, {3, 6}, {4, 5}, {4, 6}}},  // P15
      {5, {{0, 1}, {0, 2}, {1, 3}, {2, 4}, {3, 4}}},                                    // P16 5-cycle
      {6, {{0, 1}, {0, 2}, {1, 3}, {2, 4}, {3, 5}, {4, 5}}},                            // P17 6-cycle
      {6, {{0, 1}, {0, 2}, {0, 4}, {1, 2}, {1, 5}, {2, 3}, {3, 4}, {3, 5}, {4, 5}}},    // P18 hourglass
  };
  VSGM_CHECK(preset >= 0 && preset < (int)kPresets.size(), "unknown preset pattern P" + std::to_string(preset));
  Query q;
  q.type_ = PatternType;
  q.preset_ = preset;
  q.SetAdjacencyFromEdges(kPresets[preset].first, kPresets[preset].second);
  if (!enable_ordering) {
    q.order_.assign(q.vertex_count_, {});
  } else {
    q.ReMapVertexId();
    q.SetConditions(SymmetryConditions(q.conn_));
  }
  return q;
}

Query Query::Clique(int k) {
  VSGM_CHECK(k >= 3 && k <= kMaxQueryVerticesNum, "clique size must be in [3, 7]");
  Query q;
  q.type_ = CliqueType;
  q.vertex_count_ = k;
  q.edge_count_ = (size_t)k * (k - 1) / 2;
  q.conn_.assign(k, {});
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j)
      if (i != j) q.conn_[i].push_back(j);
  q.order_.assign(k, {});
  for (int i = 0; i + 1 < k; ++i) {
    q.order_[i].emplace_back(LESS_THAN, i + 1);
    q.order_[i + 1].emplace_back(LARGER_THAN, i);
  }
  return q;
}

Query Query::FromAdjacency(const std::vector<std::vector<int>>& adj, QueryType type) {
  size_t n = adj.size();
  VSGM_CHECK(n >= 2 && n <= (size_t)kMaxQueryVerticesNum, "pattern must have 2..7 vertices");
  Query q;
  q.type_ = type;
  q.vertex_count_ = n;
  q.conn_.assign(n, {});
  size_t directed = 0;
  for (size_t i = 0; i < n; ++i)
    for (int j : adj[i]) {
      VSGM_CHECK(j >= 0 && (size_t)j < n && (size_t)j != i, "bad pattern adjacency");
      q.conn_[i].push_back(j);
      ++directed;
    }
  q.edge_count_ = directed / 2;
  q.ReMapVertexId();
  q.SetConditions(SymmetryConditions(q.conn_));
  return q;
}

std::vector<Query> Query::Graphlets(int k) {
  VSGM_CHECK(k >= 3 && k <= 5, "graphlet size must be in [3, 5]");
  // clique edge list in (v<u) lexicographic order
  std::vector<std::pair<int, int>> clique_edges;
  for (int v = 0; v < k; ++v)
    for (int u = v + 1; u < k; ++u) clique_edges.emplace_back(v, u);
  const int m_full = (int)clique_edges.size();
  const int m_min = k - 1;

  auto to_adj = [k](const std::vector<std::pair<int, int>>& edges) {
    std::vector<std::vector<int>> adj(k);
    for (auto& e : edges) {
      adj[e.first].push_back(e.second);
      adj[e.second].push_back(e.first);
    }
    return adj;
  };
  auto connected = [k](const std::vector<std::vector<int>>& adj) {
    std::vector<char> seen(k, 0);
    std::vector<int> stack{0};
    seen[0] = 1;
    while (!stack.empty()) {
      int v = stack.back();
      stack.pop_back();
      for (int w : adj[v])
        if (!seen[w]) {
          seen[w] = 1;
          stack.push_back(w);
        }
    }
    return std::all_of(seen.begin(), seen.end(), [](char c) { return c != 0; });
  };
  auto as_conn = [](const std::vector<std::vector<int>>& adj) {
    AllConnType c(adj.size());
    for (size_t i = 0; i < adj.size(); ++i)
      for (int j : adj[i]) c[i].push_back(j);
    return c;
  };

  std::vector<std::vector<std::vector<int>>> shapes;
  for (int removed = 0; removed <= m_full - m_min; ++removed) {
    if (removed == 0) {
      shapes.push_back(to_adj(clique_edges));
      continue;
    }
    // all index combinations of `removed` edges in lexicographic order; the
    // removal erases one index at a time from the shrinking list, as the
    // reference does (graphlet.h:76-77).
    std::vector<int> comb(removed);
    std::iota(comb.begin(), comb.end(), 0);
    std::vector<std::vector<std::vector<int>>> kept;  // connected shapes for this removal count
    while (true) {
      auto edges = clique_edges;
      for (int idx : comb) edges.erase(edges.begin() + idx);
      auto adj = to_adj(edges);
      if (connected(adj)) kept.push_back(adj);
      // next combination
      int i = removed - 1;
      while (i >= 0 && comb[i] == m_full - removed + i) --i;
      if (i < 0) break;
      ++comb[i];
      for (int j = i + 1; j < removed; ++j) comb[j] = comb[j - 1] + 1;
    }
    std::vector<size_t> uniques;
    for (size_t i = 0; i < kept.size(); ++i) {
      bool fresh = true;
      for (size_t u : uniques)
        if (PatternsIsomorphic(as_conn(kept[u]), as_conn(kept[i]))) {
          fresh = false;
          break;
        }
      if (fresh) {
        uniques.push_back(i);
        shapes.push_back(kept[i]);
      }
    }
  }
  std::vector<Query> queries;
  for (auto& s : shapes) queries.push_back(FromAdjacency(s, GraphletType));
  return queries;
}

std::string Query::Describe() const {
  std::ostringstream os;
  os << GetQueryTypeString() << " V=" << vertex_count_ << " E=" << edge_count_ << "\n";
  for (size_t i = 0; i < conn_.size(); ++i) {
    os << i << ":";
    for (auto j : conn_[i]) os << " " << j;
    os << "\n";
  }
  os << "conditions:\n";
  for (size_t i = 0; i < order_.size(); ++i) {
    os << i << ":";
    for (auto& c : order_[i]) os << " " << CondOperatorName(c.first) << "(" << c.second << ")";
    os << "\n";
  }
  return os.str();
}

}  // namespace vsgm
