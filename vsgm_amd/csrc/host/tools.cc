// The reference's offline partitioning tools over this library
// (src/tools/*.cc): kmeans on the GPU (one id per line to
// `<graph>.kmeans.<k>`, KMeans::Dump, kmeans.cc:559-572, one or more
// levels), features (`.1hop.bin` / `.2hop.bin`), fennel
// (`.fennel.part.<p>`) and query_test (plans).
#include "tools.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <numeric>
#include <random>
#include <string>

#include "ingest.h"
#include "plan.h"
#include "query.h"
#include "vsgm_hip.h"
#include "vsgm_partition.h"

namespace vsgm {

std::vector<int> GpuKMeans(const Graph& g, int k, int iters, uint64_t seed, int device) {
  vsgm_ctx* ctx = nullptr;
  if (vsgm_ctx_create(device, 0, &ctx) != VSGM_OK) throw Error("GPU k-means needs GPU " + std::to_string(device));
  void *rp = nullptr, *cols = nullptr;
  std::vector<int32_t> assign(g.vertex_count);
  std::vector<uint64_t> sizes(std::max(k, 1));
  auto check = [&](int s, const char* what) {
    if (s != VSGM_OK) {
      std::string msg = std::string(what) + " failed: " + vsgm_last_error(ctx);
      vsgm_ctx_destroy(ctx);
      throw Error(msg);
    }
  };
  check(vsgm_alloc(ctx, (g.vertex_count + 1) * sizeof(uint64_t), &rp), "k-means allocation");
  check(vsgm_alloc(ctx, std::max<size_t>(g.cols.size(), 1) * sizeof(uint32_t), &cols), "k-means allocation");
  check(vsgm_h2d(ctx, rp, g.row_ptrs.data(), (g.vertex_count + 1) * sizeof(uint64_t)), "graph upload");
  check(vsgm_h2d(ctx, cols, g.cols.data(), g.cols.size() * sizeof(uint32_t)), "graph upload");
  check(vsgm_kmeans(ctx, (const uint64_t*)rp, (const uint32_t*)cols, g.vertex_count, k, iters, seed, assign.data(), sizes.data(), nullptr),
        "k-means");
  vsgm_ctx_destroy(ctx);  // frees the pool
  return std::vector<int>(assign.begin(), assign.end());
}

namespace {

// old -> new cluster ids by decreasing count of distinct vertices in the
// closed neighbourhoods of each cluster's members (members[j] has local
// cluster a[j] < k); std::sort with the reference's comparator.
std::vector<int> NonZeroOrder(const Graph& g, const std::vector<uint32_t>* members, const std::vector<int>& a, int k) {
  const uint64_t V = g.vertex_count;
  std::vector<std::vector<uint64_t>> marks(k, std::vector<uint64_t>((V + 63) / 64, 0));
  for (uint64_t j = 0; j < a.size(); ++j) {
    const uint64_t v = members ? (*members)[j] : j;
    auto& m = marks[a[j]];
    m[v >> 6] |= 1ull << (v & 63);
    for (uint64_t i = g.row_ptrs[v]; i < g.row_ptrs[v + 1]; ++i) m[g.cols[i] >> 6] |= 1ull << (g.cols[i] & 63);
  }
  std::vector<unsigned> nonzero(k, 0);
  for (int c = 0; c < k; ++c)
    for (uint64_t w : marks[c]) nonzero[c] += __builtin_popcountll(w);
  std::vector<unsigned> new2old(k);
  std::iota(new2old.begin(), new2old.end(), 0u);
  std::sort(new2old.begin(), new2old.end(), [&](unsigned x, unsigned y) { return nonzero[x] > nonzero[y]; });
  std::vector<int> old2new(k);
  for (int i = 0; i < k; ++i) old2new[new2old[i]] = i;
  return old2new;
}

}  // namespace

void ReorderClusters(const Graph& g, int k, std::vector<int>& assignment) {
  VSGM_CHECK(assignment.size() == g.vertex_count, "assignment size does not match the graph");
  const std::vector<int> old2new = NonZeroOrder(g, nullptr, assignment, k);
  for (auto& a : assignment) a = old2new[a];
}

Graph ClosedNeighbourhoods(const Graph& g) {
  Graph out;
  out.vertex_count = g.vertex_count;
  out.row_ptrs.resize(g.vertex_count + 1);
  out.cols.resize(g.cols.size() + g.vertex_count);
  for (uint64_t v = 0; v <= g.vertex_count; ++v) out.row_ptrs[v] = g.row_ptrs[v] + v;
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint64_t v = 0; v < g.vertex_count; ++v) {
    auto b = out.cols.begin() + out.row_ptrs[v];
    std::copy(g.cols.begin() + g.row_ptrs[v], g.cols.begin() + g.row_ptrs[v + 1], b);
    b[g.row_ptrs[v + 1] - g.row_ptrs[v]] = (uint32_t)v;
    std::sort(b, out.cols.begin() + out.row_ptrs[v + 1]);
  }
  return out;
}

Graph TwoHopBalls(const Graph& g) {
  Graph out;
  out.vertex_count = g.vertex_count;
  std::vector<std::vector<uint32_t>> rows(g.vertex_count);
#pragma omp parallel
  {
    std::vector<uint8_t> seen(g.vertex_count, 0);
#pragma omp for schedule(dynamic, 256)
    for (uint64_t v = 0; v < g.vertex_count; ++v) {
      auto& r = rows[v];
      auto add = [&](uint32_t u) {
        if (!seen[u]) seen[u] = 1, r.push_back(u);
      };
      add((uint32_t)v);
      for (uint64_t i = g.row_ptrs[v]; i < g.row_ptrs[v + 1]; ++i) {
        const uint32_t w = g.cols[i];
        add(w);
        for (uint64_t j = g.row_ptrs[w]; j < g.row_ptrs[w + 1]; ++j) add(g.cols[j]);
      }
      for (uint32_t u : r) seen[u] = 0;
      std::sort(r.begin(), r.end());
    }
  }
  out.row_ptrs.assign(g.vertex_count + 1, 0);
  for (uint64_t v = 0; v < g.vertex_count; ++v) out.row_ptrs[v + 1] = out.row_ptrs[v] + rows[v].size();
  out.cols.reserve(out.row_ptrs.back());
  for (auto& r : rows) out.cols.insert(out.cols.end(), r.begin(), r.end());
  return out;
}

Graph Duplicate(const Graph& g, int copies) {
  VSGM_CHECK(copies >= 1 && (uint64_t)copies * g.vertex_count <= 0xFFFFFFFFull, "bad duplicate count");
  const uint64_t V = g.vertex_count, N = V * (uint64_t)copies;
  // copy d of vertex v is v + V * d, its list shifted alike (features.cc:89-125)
  auto old_row = [&](uint64_t u, std::vector<uint32_t>& out) {
    const uint64_t v = u % V, d = u / V;
    for (uint64_t i = g.row_ptrs[v]; i < g.row_ptrs[v + 1]; ++i) out.push_back((uint32_t)(g.cols[i] + V * d));
  };
  // ids shuffled with the standard library's default engine, default seed
  // (features.cc:131-137), so the file matches the reference tool's
  std::vector<uint32_t> new2old(N);
  std::iota(new2old.begin(), new2old.end(), 0u);
  std::shuffle(new2old.begin(), new2old.end(), std::default_random_engine{});
  std::vector<uint32_t> old2new(N);
  for (uint64_t u = 0; u < N; ++u) old2new[new2old[u]] = (uint32_t)u;
  Graph out;
  out.vertex_count = N;
  out.row_ptrs.assign(N + 1, 0);
  for (uint64_t u = 0; u < N; ++u) {
    const uint64_t v = new2old[u] % V;
    out.row_ptrs[u + 1] = out.row_ptrs[u] + (g.row_ptrs[v + 1] - g.row_ptrs[v]);
  }
  out.cols.resize(out.row_ptrs[N]);
#pragma omp parallel
  {
    std::vector<uint32_t> row;
#pragma omp for schedule(dynamic, 1024)
    for (uint64_t u = 0; u < N; ++u) {
      row.clear();
      old_row(new2old[u], row);
      for (auto& c : row) c = old2new[c];
      std::sort(row.begin(), row.end());
      std::copy(row.begin(), row.end(), out.cols.begin() + out.row_ptrs[u]);
    }
  }
  return out;
}

bool OpenIfClosed(Graph& g) {
  for (uint64_t v = 0; v < g.vertex_count; ++v)
    if (!std::binary_search(g.cols.begin() + g.row_ptrs[v], g.cols.begin() + g.row_ptrs[v + 1], (uint32_t)v)) return false;
  uint64_t w = 0;
  for (uint64_t v = 0; v < g.vertex_count; ++v) {
    const uint64_t b = g.row_ptrs[v], e = g.row_ptrs[v + 1];
    bool dropped = false;
    g.row_ptrs[v] = w;
    for (uint64_t i = b; i < e; ++i) {
      if (!dropped && g.cols[i] == v) {
        dropped = true;
        continue;
      }
      g.cols[w++] = g.cols[i];
    }
  }
  g.row_ptrs[g.vertex_count] = w;
  g.cols.resize(w);
  return true;
}

int ParseBisection(const std::string& ks) {
  if (ks.size() < 3 || ks.compare(0, 2, "2^") != 0) return 0;
  const std::string x = ks.substr(2);
  VSGM_CHECK(!x.empty() && x.size() < 3 && x.find_first_not_of("0123456789") == std::string::npos, "bad level count in \"" + ks + "\"");
  const int n = std::stoi(x);
  VSGM_CHECK(n >= 1 && n <= 16, "bad level count in \"" + ks + "\"");
  return n;
}

std::vector<int> ParseLevels(const std::string& ks) {
  auto number = [&](const std::string& t) {
    VSGM_CHECK(!t.empty() && t.size() < 6 && t.find_first_not_of("0123456789") == std::string::npos, "bad cluster count in \"" + ks + "\"");
    return std::stoi(t);
  };
  std::vector<int> levels;
  const size_t caret = ks.find('^');
  if (caret != std::string::npos) {
    const int a = number(ks.substr(0, caret)), x = number(ks.substr(caret + 1));
    // "2^x" is the recursive bisection of GpuKMeansBisect
    VSGM_CHECK(a != 2, "\"2^x\" is a recursive bisection, not a list of levels");
    VSGM_CHECK(x >= 1 && x <= 8, "bad level count in \"" + ks + "\"");
    levels.assign(x, a);
  } else {
    for (size_t start = 0;;) {
      const size_t pos = ks.find('x', start);
      levels.push_back(number(ks.substr(start, pos == std::string::npos ? std::string::npos : pos - start)));
      if (pos == std::string::npos) break;
      start = pos + 1;
    }
  }
  uint64_t total = 1;
  for (int k : levels) {
    VSGM_CHECK(k >= 1 && k <= VSGM_KMEANS_MAX_K, "cluster counts must be in [1, 16]");
    total *= (uint64_t)k;
  }
  VSGM_CHECK(total <= (1u << 30), "too many clusters");
  return levels;
}

namespace {

// Recursive bisection (KMeans::Run2x, kmeans.cc:284-383). Each node splits
// its vertices in two with the member-subset k-means, groups them cluster
// 0 first (the reference's in-place two-pointer swap), recurses on both
// halves, and then orients the halves so the two most similar child-level
// centroids meet at the boundary. A centroid is kept as its member list;
// the similarity of two is the dot product of their mean features
// (CosineSimilarity's dense-dense similarity, kmeans_utils.h:152-158).
class Bisector {
 public:
  Bisector(const Graph& g, vsgm_ctx* ctx, const void* rp, const void* cols, void* dm, int iters, uint64_t seed, int levels,
           std::vector<int>& global)
      : g_(g), ctx_(ctx), rp_(rp), cols_(cols), dm_(dm), iters_(iters), seed_(seed), levels_(levels), global_(global),
        acc_(g.vertex_count, 0.0) {}

  int Run(uint32_t* vids, uint64_t n, int level, std::vector<uint32_t> (&cent)[2]) {
    std::vector<int32_t> a(n, 0);
    uint64_t size0 = n;
    if (n >= 2) {
      std::vector<uint64_t> sz(2, 0);
      const uint64_t s = seed_ + 0x9E3779B97F4A7C15ull * (++nodes_);
      Check(vsgm_h2d(ctx_, dm_, vids, n * sizeof(uint32_t)), "member upload");
      Check(vsgm_kmeans_members(ctx_, (const uint64_t*)rp_, (const uint32_t*)cols_, g_.vertex_count, (const uint32_t*)dm_, n, 2, iters_, s,
                                a.data(), sz.data(), nullptr),
            "k-means");
      size0 = sz[0];
    }
    if (size0 == 0 || size0 == n) {  // cannot be split: one cluster, both centroids its own
      cent[0].assign(vids, vids + n);
      cent[1] = cent[0];
      return 1;
    }
    for (uint64_t i = 0, j = n - 1; i < j;) {
      while (i < j && a[i] == 0) ++i;
      while (i < j && a[j] == 1) --j;
      if (i < j) std::swap(vids[i++], vids[j--]);
    }
    const uint64_t size1 = n - size0;
    if (level == levels_) {
      for (uint64_t i = size0; i < n; ++i) global_[vids[i]] = 1;
      cent[0].assign(vids, vids + size0);
      cent[1].assign(vids + size0, vids + n);
      return 2;
    }
    std::vector<uint32_t> c0[2], c1[2];
    const int n0 = Run(vids, size0, level + 1, c0);
    const int n1 = Run(vids + size0, size1, level + 1, c1);
    const double d[4] = {Dot(c0[0], c1[0]), Dot(c0[0], c1[1]), Dot(c0[1], c1[0]), Dot(c0[1], c1[1])};
    const int min_idx = (int)(std::min_element(d, d + 4) - d);
    const int i0 = (min_idx & 2) != 0, i1 = (min_idx & 1) != 0;
    cent[0] = std::move(c0[1 - i0]);
    cent[1] = std::move(c1[1 - i1]);
    if (i0 != 1) std::reverse(vids, vids + size0);
    if (i1 != 0) std::reverse(vids + size0, vids + n);
    for (uint64_t i = size0; i < n; ++i) global_[vids[i]] += n0;
    return n0 + n1;
  }

 private:
  void Check(int st, const char* what) {
    if (st != VSGM_OK) throw Error(std::string(what) + " failed: " + vsgm_last_error(ctx_));
  }
  // mean feature of A . mean feature of B, features 1/sqrt(|N[v]|) on N[v] = row + v
  double Dot(const std::vector<uint32_t>& A, const std::vector<uint32_t>& B) {
    auto each = [&](uint32_t v, auto&& f) {
      const double w = std::sqrt(1.0 / (double)(g_.row_ptrs[v + 1] - g_.row_ptrs[v] + 1));
      for (uint64_t i = g_.row_ptrs[v]; i < g_.row_ptrs[v + 1]; ++i) f(g_.cols[i], w);
      f(v, w);
    };
    for (uint32_t v : A) each(v, [&](uint32_t u, double w) { acc_[u] += w; });
    double dot = 0.0;
    for (uint32_t v : B) each(v, [&](uint32_t u, double w) { dot += acc_[u] * w; });
    for (uint32_t v : A) each(v, [&](uint32_t u, double) { acc_[u] = 0.0; });
    return dot / ((double)A.size() * (double)B.size());
  }

  const Graph& g_;
  vsgm_ctx* ctx_;
  const void *rp_, *cols_;
  void* dm_;
  int iters_;
  uint64_t seed_;
  int levels_;
  std::vector<int>& global_;
  std::vector<double> acc_;
  uint64_t nodes_ = 0;
};

}  // namespace

std::vector<int> GpuKMeansBisect(const Graph& g, int levels, int iters, uint64_t seed, int device, int* clusters) {
  const uint64_t V = g.vertex_count;
  VSGM_CHECK(V > 0 && levels >= 1, "bad bisection arguments");
  vsgm_ctx* ctx = nullptr;
  if (vsgm_ctx_create(device, 0, &ctx) != VSGM_OK) throw Error("GPU k-means needs GPU " + std::to_string(device));
  std::vector<int> global(V, 0);
  int n = 0;
  try {
    void *rp = nullptr, *cols = nullptr, *dm = nullptr;
    auto check = [&](int st, const char* what) {
      if (st != VSGM_OK) throw Error(std::string(what) + " failed: " + vsgm_last_error(ctx));
    };
    check(vsgm_alloc(ctx, (V + 1) * sizeof(uint64_t), &rp), "k-means allocation");
    check(vsgm_alloc(ctx, std::max<size_t>(g.cols.size(), 1) * sizeof(uint32_t), &cols), "k-means allocation");
    check(vsgm_alloc(ctx, V * sizeof(uint32_t), &dm), "k-means allocation");
    check(vsgm_h2d(ctx, rp, g.row_ptrs.data(), (V + 1) * sizeof(uint64_t)), "graph upload");
    check(vsgm_h2d(ctx, cols, g.cols.data(), g.cols.size() * sizeof(uint32_t)), "graph upload");
    std::vector<uint32_t> all(V);
    std::iota(all.begin(), all.end(), 0u);
    std::vector<uint32_t> cent[2];
    Bisector b(g, ctx, rp, cols, dm, iters, seed, levels, global);
    n = b.Run(all.data(), V, 1, cent);
    // ids renumbered by runs along the final vertex order (kmeans.cc:364-376)
    int prev = global[all[0]];
    global[all[0]] = 0;
    for (uint64_t i = 1; i < V; ++i) {
      int& a = global[all[i]];
      if (prev != a) {
        prev = a;
        a = global[all[i - 1]] + 1;
      } else {
        a = global[all[i - 1]];
      }
    }
  } catch (...) {
    vsgm_ctx_destroy(ctx);
    throw;
  }
  vsgm_ctx_destroy(ctx);  // frees the pool
  if (clusters) *clusters = n;
  return global;
}

std::vector<int> GpuKMeansSpec(const Graph& g, const std::string& ks, int iters, uint64_t seed, int device, bool reorder, int* clusters) {
  if (const int x = ParseBisection(ks)) return GpuKMeansBisect(g, x, iters, seed, device, clusters);
  return GpuKMeansLevels(g, ParseLevels(ks), iters, seed, device, reorder, clusters);
}

std::vector<int> GpuKMeansLevels(const Graph& g, const std::vector<int>& levels, int iters, uint64_t seed, int device, bool reorder,
                                 int* clusters) {
  const uint64_t V = g.vertex_count;
  VSGM_CHECK(V > 0, "empty graph");
  vsgm_ctx* ctx = nullptr;
  if (vsgm_ctx_create(device, 0, &ctx) != VSGM_OK) throw Error("GPU k-means needs GPU " + std::to_string(device));
  void *rp = nullptr, *cols = nullptr, *dm = nullptr;
  auto check = [&](int s, const char* what) {
    if (s != VSGM_OK) {
      std::string msg = std::string(what) + " failed: " + vsgm_last_error(ctx);
      vsgm_ctx_destroy(ctx);
      throw Error(msg);
    }
  };
  check(vsgm_alloc(ctx, (V + 1) * sizeof(uint64_t), &rp), "k-means allocation");
  check(vsgm_alloc(ctx, std::max<size_t>(g.cols.size(), 1) * sizeof(uint32_t), &cols), "k-means allocation");
  check(vsgm_alloc(ctx, V * sizeof(uint32_t), &dm), "k-means allocation");
  check(vsgm_h2d(ctx, rp, g.row_ptrs.data(), (V + 1) * sizeof(uint64_t)), "graph upload");
  check(vsgm_h2d(ctx, cols, g.cols.data(), g.cols.size() * sizeof(uint32_t)), "graph upload");
  // KMeans::Run(std::vector<int>) (kmeans.cc:386-489): every level splits
  // each non-empty cluster of the previous one into K, ids a + c * K
  std::vector<int> global(V, 0);
  std::vector<uint64_t> sizes(1, V);
  for (size_t level = 0; level < levels.size(); ++level) {
    const int K = levels[level];
    if (K <= 1) continue;
    std::vector<std::vector<uint32_t>> parts(sizes.size());
    for (uint64_t v = 0; v < V; ++v) parts[global[v]].push_back((uint32_t)v);
    parts.erase(std::remove_if(parts.begin(), parts.end(), [](const std::vector<uint32_t>& p) { return p.empty(); }), parts.end());
    std::vector<uint64_t> next_sizes(parts.size() * K, 0);
    for (size_t c = 0; c < parts.size(); ++c) {
      const auto& part = parts[c];
      const uint64_t m = part.size();
      const int k = (int)std::min<uint64_t>(K, m);
      std::vector<int32_t> a(m);
      std::vector<uint64_t> sz(k, 0);
      if (m == V) {
        // the whole graph: the single-level run, same seed
        check(vsgm_kmeans(ctx, (const uint64_t*)rp, (const uint32_t*)cols, V, k, iters, seed, a.data(), sz.data(), nullptr), "k-means");
      } else {
        const uint64_t s = seed + 0x9E3779B97F4A7C15ull * ((uint64_t)level * 65536 + c + 1);
        check(vsgm_h2d(ctx, dm, part.data(), m * sizeof(uint32_t)), "member upload");
        check(vsgm_kmeans_members(ctx, (const uint64_t*)rp, (const uint32_t*)cols, V, (const uint32_t*)dm, m, k, iters, s, a.data(),
                                  sz.data(), nullptr),
              "k-means");
      }
      std::vector<int> local(a.begin(), a.end());
      std::vector<int> old2new(k);
      std::iota(old2new.begin(), old2new.end(), 0);
      if (reorder) old2new = NonZeroOrder(g, m == V ? nullptr : &part, local, k);
      for (uint64_t j = 0; j < m; ++j) global[part[j]] = old2new[local[j]] + (int)c * K;
      for (int x = 0; x < k; ++x) next_sizes[old2new[x] + c * K] = sz[x];
    }
    sizes.swap(next_sizes);
  }
  vsgm_ctx_destroy(ctx);  // frees the pool
  if (clusters) *clusters = (int)sizes.size();
  return global;
}

int RemoveEmptyClusters(int k, std::vector<int>& assignment) {
  std::vector<uint64_t> size(k, 0);
  for (int a : assignment) ++size[a];
  std::vector<int> old2new(k);
  int kept = 0;
  for (int c = 0; c < k; ++c)
    if (size[c]) old2new[c] = kept++;
  for (auto& a : assignment) a = old2new[a];
  return kept;
}

namespace {

const char* Flag(int argc, char** argv, const char* key) {
  for (int i = 1; i + 1 < argc; ++i)
    if (!strcmp(argv[i], key)) return argv[i + 1];
  return nullptr;
}

}  // namespace

int KMeansToolMain(int argc, char** argv) {
  if (argc == 1) {
    std::cout << "./kmeans -f GRAPH_FILENAME(bin format) -k num_clusters -i num_iterations -o result_output_file_path -d to_dump"
                 " -rc reorder_centroids -init 0 [-g GPU] [-s SEED]"
              << std::endl;
    return -1;
  }
  auto str = [&](const char* key, const std::string& d) { return Flag(argc, argv, key) ? std::string(Flag(argc, argv, key)) : d; };
  auto num = [&](const char* key, long d) { return Flag(argc, argv, key) ? atol(Flag(argc, argv, key)) : d; };
  const std::string filename = str("-f", "./data/com-dblp.ungraph.txt");
  const std::string ks = str("-k", "4");
  const int iters = (int)num("-i", 20);
  const std::string output = str("-o", filename + ".kmeans.");
  const int dump = (int)num("-d", 1);
  const int reorder = (int)num("-rc", 0);
  const int init = (int)num("-init", 0);
  const int device = (int)num("-g", 0);
  const uint64_t seed = (uint64_t)num("-s", 1);
  try {
    if (!ParseBisection(ks)) ParseLevels(ks);
  } catch (const std::exception& ex) {
    std::cerr << ex.what() << std::endl;
    return 2;
  }
  if (init != 0) {
    std::cerr << "only the farthest-first initialisation (-init 0) is supported" << std::endl;
    return 2;
  }
  Graph g = ReadCSRBinFile(filename);
  // The reference clusters the rows as given and is run on the closed rows
  // written by `features -h 1`; the GPU k-means closes every row itself,
  // so rows that already hold their own vertex are opened first.
  if (OpenIfClosed(g)) std::cout << "kmeans: rows hold their own vertex (a features -h 1 graph); clustering their closed rows" << std::endl;
  auto t0 = std::chrono::steady_clock::now();
  int k = 0;
  std::vector<int> assignment = GpuKMeansSpec(g, ks, iters, seed, device, reorder != 0, &k);
  const int kept = RemoveEmptyClusters(k, assignment);
  std::cout << "kmeans: " << kept << " non-empty clusters of " << k << " in "
            << std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() << " ms" << std::endl;
  if (dump) {
    std::ofstream os(output + ks);
    if (!os) throw Error("cannot write " + output + ks);
    for (int a : assignment) os << a << "\n";
  }
  return 0;
}

}  // namespace vsgm

extern "C" int vsgm_kmeans_tool(int argc, char** argv) {
  try {
    return vsgm::KMeansToolMain(argc, argv);
  } catch (const std::exception& ex) {
    std::cerr << "kmeans: " << ex.what() << std::endl;
    return 1;
  }
}

namespace vsgm {

int FeaturesToolMain(int argc, char** argv) {
  if (argc == 1) return -1;
  auto str = [&](const char* key, const std::string& d) { return Flag(argc, argv, key) ? std::string(Flag(argc, argv, key)) : d; };
  auto num = [&](const char* key, long d) { return Flag(argc, argv, key) ? atol(Flag(argc, argv, key)) : d; };
  const std::string filename = str("-f", "./data/com-dblp.ungraph.txt");
  const int hop = (int)num("-h", 2);
  const int duplicates = (int)num("-dn", 0);
  Graph g = LoadGraph(filename, -1, 0, false);
  if (duplicates > 0) {
    // "<file>.x<dn>.bin": dn shuffled copies of the graph, -h ignored (features.cc:89-173)
    const std::string out = filename + ".x" + std::to_string(duplicates) + ".bin";
    WriteCSRBinFile(out, Duplicate(g, duplicates));
    std::cout << "wrote " << out << std::endl;
    return 0;
  }
  if (hop != 1 && hop != 2) {
    std::cerr << "only -h 1 and -h 2 features are written (not the sampled -h -2)" << std::endl;
    return 2;
  }
  // "<file>.1hop.bin" (features.cc:176-207): each row plus its own vertex;
  // "<file>.2hop.bin" (:207-258): every vertex within two steps; sorted rows
  const std::string out = filename + (hop == 1 ? ".1hop.bin" : ".2hop.bin");
  WriteCSRBinFile(out, hop == 1 ? ClosedNeighbourhoods(g) : TwoHopBalls(g));
  std::cout << "wrote " << out << std::endl;
  return 0;
}

}  // namespace vsgm

extern "C" int vsgm_features_tool(int argc, char** argv) {
  try {
    return vsgm::FeaturesToolMain(argc, argv);
  } catch (const std::exception& ex) {
    std::cerr << "features: " << ex.what() << std::endl;
    return 1;
  }
}

namespace vsgm {

std::vector<int> FennelPartition(const Graph& g, int parts, int hop) {
  VSGM_CHECK(parts >= 1 && (hop == 1 || hop == 2), "fennel needs parts >= 1 and hop 1 or 2");
  // The reference's constants and arithmetic types (fennel.cc:40-44):
  // alpha and miu are floats, the balance term is double, scores are floats.
  const size_t vertex_num = g.vertex_count;
  const size_t edge_num = g.cols.size() / 2;
  const float gamma = 1.5;
  const float alpha = std::sqrt(parts) * edge_num / std::pow(vertex_num, gamma);
  const float slack = 1.1;
  const float miu = slack * vertex_num / parts;
  std::vector<int> part(vertex_num, -1);
  std::vector<long> size(parts, 0);
  // Streaming assignment in vertex order (fennel.cc:54-89). A neighbour
  // counts for part i when it was already placed there; the reference's
  // union-find holds exactly that relation (each placed vertex points at
  // its part's first vertex).
  auto in_part = [&](uintV u, int i) { return part[u] == i; };
  for (uint64_t v = 0; v < vertex_num; ++v) {
    float max_score = -FLT_MAX;
    int max_part = 0;
    for (int i = 0; i < parts; ++i) {
      if (!(size[i] <= miu)) continue;
      const double delta_c = alpha * (std::pow(size[i] + 1, gamma) - std::pow(size[i], gamma));
      float score = 0;
      for (uint64_t j = g.row_ptrs[v]; j < g.row_ptrs[v + 1]; ++j) {
        const uintV nid = g.cols[j];
        if (in_part(nid, i)) score += 1;
        if (hop == 2)
          for (uint64_t k = g.row_ptrs[nid]; k < g.row_ptrs[nid + 1]; ++k)
            if (in_part(g.cols[k], i)) score += 1;
      }
      score = score - delta_c;
      if (max_score < score) {
        max_score = score;
        max_part = i;
      }
    }
    size[max_part] += 1;
    part[v] = max_part;
  }
  return part;
}

int FennelToolMain(int argc, char** argv) {
  if (argc == 1) return -1;
  auto str = [&](const char* key, const std::string& d) { return Flag(argc, argv, key) ? std::string(Flag(argc, argv, key)) : d; };
  auto num = [&](const char* key, long d) { return Flag(argc, argv, key) ? atol(Flag(argc, argv, key)) : d; };
  const std::string filename = str("-f", "./data/com-dblp.ungraph.txt");
  const int parts = (int)num("-p", 4);
  const int hop = (int)num("-h", 1);
  Graph g = LoadGraph(filename, -1, 0, false);
  auto t0 = std::chrono::steady_clock::now();
  std::vector<int> part = FennelPartition(g, parts, hop);
  std::cout << "fennel: " << parts << " parts in "
            << std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() << " ms" << std::endl;
  // "<file>.fennel.part.<p>", one part per line (fennel.cc:94-99)
  const std::string out = filename + ".fennel.part." + std::to_string(parts);
  std::ofstream os(out);
  for (int p : part) os << p << "\n";
  if (!os) throw Error("write to " + out + " failed");
  return 0;
}

}  // namespace vsgm

extern "C" int vsgm_fennel_tool(int argc, char** argv) {
  try {
    return vsgm::FennelToolMain(argc, argv);
  } catch (const std::exception& ex) {
    std::cerr << "fennel: " << ex.what() << std::endl;
    return 1;
  }
}

namespace vsgm {

int QueryTestMain(int argc, char** argv) {
  if (argc == 1) return -1;
  auto num = [&](const char* key, long d) { return Flag(argc, argv, key) ? atol(Flag(argc, argv, key)) : d; };
  const int preset = (int)num("-p", 1);
  const int k_c = (int)num("-kc", 0);
  const int k_m = (int)num("-km", 0);
  std::vector<Query> queries;
  if (k_c > 0)
    queries.push_back(Query::Clique(k_c));
  else if (k_m > 0)
    queries = Query::Graphlets(k_m);
  else
    queries.push_back(Query::Preset(preset));
  int largest = 0;
  std::vector<int> hops;
  for (auto& q : queries) {
    std::cout << q.Describe();
    Plan p(q);
    p.Optimize();
    std::cout << p.Serialize();
    hops.push_back(p.GetHop());
    largest = std::max(largest, p.GetHop());
  }
  // the graphlet summary of query_test.cc:56-60
  if (k_c == 0 && k_m > 0) {
    std::cout << "-----------" << std::endl;
    for (size_t i = 0; i < hops.size(); ++i) std::cout << i << ": " << hops[i] << std::endl;
    std::cout << "largest hop: " << largest << std::endl;
  }
  return 0;
}

}  // namespace vsgm

extern "C" int vsgm_query_test(int argc, char** argv) {
  try {
    return vsgm::QueryTestMain(argc, argv);
  } catch (const std::exception& ex) {
    std::cerr << "query_test: " << ex.what() << std::endl;
    return 1;
  }
}
