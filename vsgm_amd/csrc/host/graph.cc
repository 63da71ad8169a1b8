// Graph IO and generators. See graph.h.
#include "graph.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <memory>
#include <sstream>

namespace vsgm {

namespace {
// VSGM_PROGRESS=<seconds> (> 0): the generators report each build phase on
// stderr, so that a minutes-long build of a billion-edge graph is visibly
// making progress.
void Phase(const char* what) {
  static const bool on = [] {
    const char* s = getenv("VSGM_PROGRESS");
    return s && atof(s) > 0;
  }();
  if (!on) return;
  static const double t0 = omp_get_wtime();
  fprintf(stderr, "vsgm: graph build: %s (%.1f s)\n", what, omp_get_wtime() - t0);
  fflush(stderr);
}
}  // namespace

uint64_t Graph::MaxDegree() const {
  uint64_t m = 0;
  for (uint64_t v = 0; v < vertex_count; ++v) m = std::max<uint64_t>(m, row_ptrs[v + 1] - row_ptrs[v]);
  return m;
}

std::string Graph::Validate(bool check_symmetry) const {
  std::ostringstream err;
  if (row_ptrs.size() != vertex_count + 1) return "row_ptrs size mismatch";
  if (row_ptrs[0] != 0 || row_ptrs[vertex_count] != cols.size()) return "row_ptrs bounds mismatch";
  for (uint64_t v = 0; v < vertex_count; ++v) {
    if (row_ptrs[v + 1] < row_ptrs[v]) return "row_ptrs not monotone";
    for (uint64_t j = row_ptrs[v]; j < row_ptrs[v + 1]; ++j) {
      if (cols[j] >= vertex_count) return "neighbour id out of range";
      if (cols[j] == v) return "self loop at " + std::to_string(v);
      if (j > row_ptrs[v] && cols[j - 1] >= cols[j]) return "unsorted or duplicate list at " + std::to_string(v);
    }
  }
  if (check_symmetry) {
    for (uint64_t v = 0; v < vertex_count; ++v)
      for (uint64_t j = row_ptrs[v]; j < row_ptrs[v + 1]; ++j) {
        uintV w = cols[j];
        if (!std::binary_search(cols.begin() + row_ptrs[w], cols.begin() + row_ptrs[w + 1], (uintV)v)) return "asymmetric edge";
      }
  }
  return "";
}

Graph ReadCSRBinFile(const std::string& filename) {
  FILE* f = fopen(filename.c_str(), "rb");
  VSGM_CHECK(f != nullptr, "cannot open " + filename);
  std::unique_ptr<FILE, int (*)(FILE*)> guard(f, fclose);
  uint64_t header[4];
  VSGM_CHECK(fread(header, sizeof(uint64_t), 4, f) == 4, "short header in " + filename);
  VSGM_CHECK(header[0] == sizeof(uintV) && header[1] == sizeof(uintE), "unexpected id sizes in " + filename);
  Graph g;
  g.vertex_count = header[2];
  g.row_ptrs.resize(g.vertex_count + 1);
  g.cols.resize(header[3]);
  VSGM_CHECK(fread(g.row_ptrs.data(), sizeof(uintE), g.vertex_count + 1, f) == g.vertex_count + 1, "short row_ptrs in " + filename);
  VSGM_CHECK(fread(g.cols.data(), sizeof(uintV), header[3], f) == header[3], "short cols in " + filename);
  VSGM_CHECK(fgetc(f) == EOF, "trailing bytes in " + filename);
  VSGM_CHECK(g.row_ptrs[g.vertex_count] == header[3], "row_ptrs do not match the entry count");
  return g;
}

void WriteCSRBinFile(const std::string& filename, const Graph& g) {
  FILE* f = fopen(filename.c_str(), "wb");
  VSGM_CHECK(f != nullptr, "cannot create " + filename);
  std::unique_ptr<FILE, int (*)(FILE*)> guard(f, fclose);
  uint64_t header[4] = {sizeof(uintV), sizeof(uintE), g.vertex_count, (uint64_t)g.cols.size()};
  VSGM_CHECK(fwrite(header, sizeof(uint64_t), 4, f) == 4, "write failed");
  VSGM_CHECK(fwrite(g.row_ptrs.data(), sizeof(uintE), g.row_ptrs.size(), f) == g.row_ptrs.size(), "write failed");
  VSGM_CHECK(fwrite(g.cols.data(), sizeof(uintV), g.cols.size(), f) == g.cols.size(), "write failed");
}

uint64_t Mix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

double UniformAt(uint64_t seed, uint64_t stream, uint64_t index) {
  uint64_t h = Mix64(Mix64(seed ^ Mix64(stream)) + index);
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

static int ThreadCount(int threads) { return threads > 0 ? threads : omp_get_max_threads(); }

Graph BuildFromEdges(uint64_t vertex_count, std::vector<std::pair<uintV, uintV>>& edges, int threads) {
  const int nt = ThreadCount(threads);
  Graph g;
  g.vertex_count = vertex_count;
  // count both directions, self loops dropped
  std::vector<std::atomic<uint64_t>> deg(vertex_count + 1);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (uint64_t v = 0; v <= vertex_count; ++v) deg[v].store(0, std::memory_order_relaxed);
  const uint64_t m = edges.size();
#pragma omp parallel for num_threads(nt) schedule(static)
  for (uint64_t i = 0; i < m; ++i) {
    auto e = edges[i];
    if (e.first == e.second) continue;
    deg[e.first].fetch_add(1, std::memory_order_relaxed);
    deg[e.second].fetch_add(1, std::memory_order_relaxed);
  }
  Phase("degrees counted");
  std::vector<uint64_t> start(vertex_count + 1, 0);
  for (uint64_t v = 0; v < vertex_count; ++v) start[v + 1] = start[v] + deg[v].load(std::memory_order_relaxed);
  std::vector<uintV> raw(start[vertex_count]);
  std::vector<std::atomic<uint64_t>> fillp(vertex_count);
#pragma omp parallel for num_threads(nt) schedule(static)
  for (uint64_t v = 0; v < vertex_count; ++v) fillp[v].store(start[v], std::memory_order_relaxed);
  for (uint64_t s0 = 0; s0 < m; s0 += (m + 7) / 8) {
    const uint64_t s1 = std::min(m, s0 + (m + 7) / 8);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (uint64_t i = s0; i < s1; ++i) {
      auto e = edges[i];
      if (e.first == e.second) continue;
      raw[fillp[e.first].fetch_add(1, std::memory_order_relaxed)] = e.second;
      raw[fillp[e.second].fetch_add(1, std::memory_order_relaxed)] = e.first;
    }
    Phase("lists filling (slice)");
  }
  std::vector<std::pair<uintV, uintV>>().swap(edges);
  Phase("lists filled");
  // sort + unique each list, then compact
  std::vector<uint64_t> kept(vertex_count + 1, 0);
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1024)
  for (uint64_t v = 0; v < vertex_count; ++v) {
    auto b = raw.begin() + start[v], e = raw.begin() + start[v + 1];
    std::sort(b, e);
    kept[v] = std::unique(b, e) - b;
  }
  Phase("lists sorted");
  g.row_ptrs.assign(vertex_count + 1, 0);
  for (uint64_t v = 0; v < vertex_count; ++v) g.row_ptrs[v + 1] = g.row_ptrs[v] + kept[v];
  g.cols.resize(g.row_ptrs[vertex_count]);
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1024)
  for (uint64_t v = 0; v < vertex_count; ++v) std::copy(raw.begin() + start[v], raw.begin() + start[v] + kept[v], g.cols.begin() + g.row_ptrs[v]);
  Phase("CSR compacted");
  return g;
}

Graph RelabelByDegree(const Graph& g, std::vector<uintV>* new_id_out, int threads) {
  const int nt = ThreadCount(threads);
  const uint64_t n = g.vertex_count;
  // counting sort by degree; ids ascending within one degree (stable)
  const uint64_t maxd = g.MaxDegree();
  std::vector<uint64_t> first(maxd + 2, 0);
  for (uint64_t v = 0; v < n; ++v) ++first[g.Degree((uintV)v) + 1];
  for (uint64_t d = 0; d <= maxd; ++d) first[d + 1] += first[d];
  std::vector<uintV> order(n), new_id(n);
  for (uint64_t v = 0; v < n; ++v) {
    const uint64_t r = first[g.Degree((uintV)v)]++;
    order[r] = (uintV)v;
    new_id[v] = (uintV)r;
  }
  Graph out;
  out.vertex_count = n;
  out.row_ptrs.assign(n + 1, 0);
  for (uint64_t r = 0; r < n; ++r) out.row_ptrs[r + 1] = out.row_ptrs[r] + g.Degree(order[r]);
  out.cols.resize(g.cols.size());
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1024)
  for (uint64_t r = 0; r < n; ++r) {
    const uintV v = order[r];
    uintV* dst = out.cols.data() + out.row_ptrs[r];
    const uint64_t b = g.row_ptrs[v], e = g.row_ptrs[v + 1];
    for (uint64_t i = b; i < e; ++i) dst[i - b] = new_id[g.cols[i]];
    std::sort(dst, dst + (e - b));
  }
  if (new_id_out) *new_id_out = std::move(new_id);
  return out;
}

Graph GenerateRMAT(int scale, int edge_factor, double a, double b, double c, uint64_t seed, int threads) {
  VSGM_CHECK(scale >= 1 && scale <= 31, "R-MAT scale must be in [1, 31]");
  VSGM_CHECK(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, "bad R-MAT probabilities");
  const uint64_t n = 1ull << scale;
  const uint64_t m = (uint64_t)edge_factor << scale;
  const int nt = ThreadCount(threads);
  std::vector<std::pair<uintV, uintV>> edges(m);
  const double ab = a + b, abc = a + b + c;
#pragma omp parallel for num_threads(nt) schedule(static)
  for (uint64_t i = 0; i < m; ++i) {
    uint64_t u = 0, v = 0;
    for (int bit = 0; bit < scale; ++bit) {
      double r = UniformAt(seed, (uint64_t)bit, i);
      uint64_t ub = 0, vb = 0;
      if (r < a) {
      } else if (r < ab) {
        vb = 1;
      } else if (r < abc) {
        ub = 1;
      } else {
        ub = 1;
        vb = 1;
      }
      u = (u << 1) | ub;
      v = (v << 1) | vb;
    }
    edges[i] = std::make_pair((uintV)u, (uintV)v);
  }
  return BuildFromEdges(n, edges, nt);
}

Graph GenerateChungLu(uint64_t vertex_count, uint64_t draws, double gamma, uint64_t seed, int threads) {
  VSGM_CHECK(vertex_count >= 2 && gamma > 1.0, "bad Chung-Lu parameters");
  const int nt = ThreadCount(threads);
  const double expo = -1.0 / (gamma - 1.0);
  const double i0 = 1.0;
  // cumulative weights for inverse-CDF sampling
  std::vector<double> cdf(vertex_count);
  double acc = 0;
  for (uint64_t i = 0; i < vertex_count; ++i) {
    acc += std::pow((double)i + i0, expo);
    cdf[i] = acc;
  }
  for (auto& x : cdf) x /= acc;
  cdf[vertex_count - 1] = 1.0;
  std::vector<std::pair<uintV, uintV>> edges(draws);
  Phase("Chung-Lu weights");
  // in eight slices, so that progress shows during a billion-draw build
  for (uint64_t s0 = 0; s0 < draws; s0 += (draws + 7) / 8) {
    const uint64_t s1 = std::min(draws, s0 + (draws + 7) / 8);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (uint64_t i = s0; i < s1; ++i) {
      double r0 = UniformAt(seed, 0, i), r1 = UniformAt(seed, 1, i);
      uint64_t u = std::upper_bound(cdf.begin(), cdf.end(), r0) - cdf.begin();
      uint64_t v = std::upper_bound(cdf.begin(), cdf.end(), r1) - cdf.begin();
      if (u >= vertex_count) u = vertex_count - 1;
      if (v >= vertex_count) v = vertex_count - 1;
      edges[i] = std::make_pair((uintV)u, (uintV)v);
    }
    Phase("Chung-Lu edges drawn (slice)");
  }
  return BuildFromEdges(vertex_count, edges, nt);
}

Graph GraphFromAdjacency(const std::vector<std::vector<uintV>>& adj) {
  std::vector<std::pair<uintV, uintV>> edges;
  for (uintV v = 0; v < adj.size(); ++v)
    for (auto w : adj[v]) edges.emplace_back(v, w);
  return BuildFromEdges(adj.size(), edges, 1);
}

}  // namespace vsgm
