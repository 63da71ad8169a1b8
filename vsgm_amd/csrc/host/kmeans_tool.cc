// The kmeans tool (src/tools/kmeans.cc:636-694) on the GPU: reads a .bin
// graph, clusters its vertices, writes one cluster id per line to
// `<graph>.kmeans.<k>` (KMeans::Dump, kmeans.cc:559-572).
#include "kmeans_tool.h"

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <numeric>
#include <string>

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

void ReorderClusters(const Graph& g, int k, std::vector<int>& assignment) {
  const uint64_t V = g.vertex_count;
  std::vector<std::vector<uint64_t>> marks(k, std::vector<uint64_t>((V + 63) / 64, 0));
  for (uint64_t v = 0; v < V; ++v) {
    auto& m = marks[assignment[v]];
    m[v >> 6] |= 1ull << (v & 63);
    for (uint64_t i = g.row_ptrs[v]; i < g.row_ptrs[v + 1]; ++i) m[g.cols[i] >> 6] |= 1ull << (g.cols[i] & 63);
  }
  std::vector<unsigned> nonzero(k, 0);
  for (int c = 0; c < k; ++c)
    for (uint64_t w : marks[c]) nonzero[c] += __builtin_popcountll(w);
  std::vector<unsigned> new2old(k);
  std::iota(new2old.begin(), new2old.end(), 0u);
  std::sort(new2old.begin(), new2old.end(), [&](unsigned a, unsigned b) { return nonzero[a] > nonzero[b]; });
  std::vector<int> old2new(k);
  for (int i = 0; i < k; ++i) old2new[new2old[i]] = i;
  for (auto& a : assignment) a = old2new[a];
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
  if (ks.empty() || ks.find_first_not_of("0123456789") != std::string::npos) {
    std::cerr << "multi-level clustering (\"axb\", \"a^x\") is not supported; give one cluster count" << std::endl;
    return 2;
  }
  if (init != 0) {
    std::cerr << "only the farthest-first initialisation (-init 0) is supported" << std::endl;
    return 2;
  }
  const int k = std::stoi(ks);
  Graph g = ReadCSRBinFile(filename);
  auto t0 = std::chrono::steady_clock::now();
  std::vector<int> assignment = GpuKMeans(g, k, iters, seed, device);
  if (reorder) ReorderClusters(g, k, assignment);
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
