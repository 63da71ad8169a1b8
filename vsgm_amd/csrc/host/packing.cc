// View-bin packing (src/tools/view_packing.cc:141-324, pool of
// src/tools/view_bin.h:65-171). The steps are kept as the reference takes
// them so that the same graph and partition give the same `.vbmap`:
//  * a bin's size is the sum of the list lengths of the vertices its
//    sources reach within hop-1 steps, counted by a depth-first walk that
//    marks a vertex `added` once counted and `visited` once expanded;
//  * a source is tried in every open bin and goes to the bin it grows
//    least among those it fits (first such bin on ties); the other trials
//    are undone by clearing the `added` marks of every vertex the trial
//    reached, expanded ones included (the reference's undo, which leaves
//    their `visited` marks set: later sizes depend on it);
//  * when it fits no open bin, a new bin opens (when pool_size bins are
//    open, the largest first-found one closes) and the source is tried again;
//  * the bins of all clusters, sorted by size, are merged next-fit.
#include "packing.h"

#include "view.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <utility>

namespace vsgm {
namespace {

class Bits {
 public:
  explicit Bits(uint64_t n) : w_((n + 63) / 64, 0) {}
  bool get(uint64_t i) const { return (w_[i >> 6] >> (i & 63)) & 1; }
  void set(uint64_t i) { w_[i >> 6] |= 1ull << (i & 63); }
  void clear(uint64_t i) { w_[i >> 6] &= ~(1ull << (i & 63)); }

 private:
  std::vector<uint64_t> w_;
};

struct Trial {
  uint64_t size = 0;
  std::vector<uint32_t> added, expanded;
};

struct Bin {
  Bin(uint64_t V) : added(V), visited(V) {}
  Bits added, visited;
};

class ClusterPacker {
 public:
  ClusterPacker(const Graph& g, int hop, uint64_t max_entries, int pool_size)
      : g_(g), hop_(hop), max_(max_entries), pool_size_(std::max(pool_size, 1)) {}

  void Pack(const std::vector<uint32_t>& order) {
    Open();
    for (uint32_t u : order) {
      if (!Insert(u)) {
        Reopen();
        if (!Insert(u)) throw Error("the view of source " + std::to_string(u) + " alone exceeds the bin size limit");
      }
    }
  }
  std::vector<std::pair<std::vector<uint32_t>, uint64_t>> Take() {
    std::vector<std::pair<std::vector<uint32_t>, uint64_t>> out;
    for (auto& b : all_) out.emplace_back(std::move(b.sources), b.size);
    return out;
  }

 private:
  struct Slot {
    std::vector<uint32_t> sources;
    uint64_t size = 0;
  };

  // Size the walk from v adds to bin b (view_bin.h:30-46 order).
  uint64_t Walk(Bin& b, uint32_t v, int hop, Trial& t) {
    uint64_t size = 0;
    if (!b.added.get(v)) {
      b.added.set(v);
      size = g_.row_ptrs[v + 1] - g_.row_ptrs[v];
      t.added.push_back(v);
    }
    if (hop > 1) {
      b.visited.set(v);
      t.expanded.push_back(v);
      for (uint64_t i = g_.row_ptrs[v]; i < g_.row_ptrs[v + 1]; ++i)
        if (!b.visited.get(g_.cols[i])) size += Walk(b, g_.cols[i], hop - 1, t);
    }
    return size;
  }

  static void Undo(Bin& b, const Trial& t) {
    for (uint32_t x : t.added) b.added.clear(x);
    for (uint32_t x : t.expanded) b.added.clear(x);
  }

  bool Insert(uint32_t s) {
    const size_t n = open_.size();
    std::vector<Trial> trials(n);
    int best = -1;
    for (size_t i = 0; i < n; ++i) {
      Bin& b = *bins_[open_[i]];
      trials[i].size = Walk(b, s, hop_, trials[i]);
      if (all_[open_[i]].size + trials[i].size <= max_ && (best < 0 || trials[i].size < trials[best].size)) best = (int)i;
    }
    for (size_t i = 0; i < n; ++i)
      if ((int)i != best) Undo(*bins_[open_[i]], trials[i]);
    if (best < 0) return false;
    all_[open_[best]].size += trials[best].size;
    all_[open_[best]].sources.push_back(s);
    return true;
  }

  void Open() {
    bins_.emplace_back(new Bin(g_.vertex_count));
    all_.emplace_back();
    open_.push_back(bins_.size() - 1);
  }

  void Reopen() {
    if ((int)open_.size() >= pool_size_) {
      size_t worst = 0;
      uint64_t worst_size = 0;
      for (size_t i = 0; i < open_.size(); ++i)
        if (all_[open_[i]].size > worst_size) worst_size = all_[open_[i]].size, worst = i;
      bins_[open_[worst]].reset();  // closed for good: its marks are no longer needed
      open_.erase(open_.begin() + worst);
    }
    Open();
  }

  const Graph& g_;
  int hop_;
  uint64_t max_;
  int pool_size_;
  std::vector<std::unique_ptr<Bin>> bins_;  // marks of open bins
  std::vector<Slot> all_;                   // sources and size of every bin
  std::vector<size_t> open_;
};

}  // namespace

PackedBins PackViewBins(const Graph& g, int hop, const std::vector<int>& partition, int first_layer, int second_layer,
                        uint64_t max_bin_entries, int pool_size, bool degree_order, int threads) {
  const uint64_t V = g.vertex_count;
  VSGM_CHECK(hop >= 1, "hop must be >= 1");
  VSGM_CHECK(first_layer >= 1 && second_layer >= 0, "bad partition layers");
  VSGM_CHECK(partition.size() == V, "partition size does not match the graph");
  const int parts = second_layer ? first_layer * second_layer : first_layer;
  for (int p : partition) VSGM_CHECK(p >= 0 && p < parts, "partition id out of range");
  auto by_degree = [&](std::vector<uint32_t>& vs) {
    std::sort(vs.begin(), vs.end(), [&](uint32_t a, uint32_t b) { return g.Degree(a) > g.Degree(b); });
  };
  // sources of each first-layer cluster, in vertex order (inner clusters in turn)
  std::vector<std::vector<uint32_t>> order(first_layer);
  if (!second_layer) {
    for (uint64_t u = 0; u < V; ++u) order[partition[u]].push_back((uint32_t)u);
    if (degree_order)
      for (auto& o : order) by_degree(o);
  } else {
    std::vector<std::vector<uint32_t>> inner(parts);
    for (uint64_t u = 0; u < V; ++u) inner[partition[u]].push_back((uint32_t)u);
    for (int i = 0; i < first_layer; ++i)
      for (int j = 0; j < second_layer; ++j) {
        auto& c = inner[i * second_layer + j];
        if (degree_order) by_degree(c);
        order[i].insert(order[i].end(), c.begin(), c.end());
      }
  }
  std::vector<std::vector<std::pair<std::vector<uint32_t>, uint64_t>>> packed(first_layer);
  std::vector<std::string> err(first_layer);
  const int nt = std::max(1, std::min(threads > 0 ? threads : omp_get_max_threads(), first_layer));
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
  for (int c = 0; c < first_layer; ++c) {
    try {
      ClusterPacker p(g, hop, max_bin_entries, pool_size);
      p.Pack(order[c]);
      packed[c] = p.Take();
    } catch (const std::exception& ex) {
      err[c] = ex.what();
    }
  }
  for (auto& e : err)
    if (!e.empty()) throw Error(e);

  // label in cluster order, then merge next-fit in increasing size
  std::vector<std::pair<std::vector<uint32_t>, uint64_t>*> bins;
  std::vector<std::pair<int, size_t>> id_size;
  for (auto& pc : packed)
    for (auto& b : pc) {
      id_size.emplace_back((int)bins.size(), (size_t)b.second);
      bins.push_back(&b);
    }
  std::sort(id_size.begin(), id_size.end(), [](const auto& a, const auto& b) { return a.second < b.second; });
  std::vector<std::vector<int>> merged(bins.size() + 1);
  int count = 0;
  uint64_t open_size = 0;
  for (auto& [id, size] : id_size) {
    if (open_size + size > max_bin_entries) {
      ++count;
      open_size = size;
    } else {
      open_size += size;
    }
    merged[count].push_back(id);
  }
  ++count;
  PackedBins out;
  out.bins = count;
  out.assignment.assign(V, 0);
  out.sources.assign(count, 0);
  out.sizes.assign(count, 0);
  for (int i = 0; i < count; ++i)
    for (int id : merged[i]) {
      for (uint32_t v : bins[id]->first) out.assignment[v] = i;
      out.sources[i] += bins[id]->first.size();
      out.sizes[i] += bins[id]->second;
    }
  return out;
}

std::vector<int> ReadPartitionFile(const std::string& filename, uint64_t vertex_count) {
  std::ifstream f(filename);
  VSGM_CHECK(f.is_open(), "cannot open partition file " + filename);
  std::vector<int> p;
  p.reserve(vertex_count);
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty()) continue;
    p.push_back(std::stoi(line));
  }
  VSGM_CHECK(p.size() == vertex_count, "partition file " + filename + " has " + std::to_string(p.size()) + " lines for " +
                                           std::to_string(vertex_count) + " vertices");
  return p;
}

void WriteVbstat(const std::string& filename, const PackedBins& p) {
  std::ofstream f(filename);
  VSGM_CHECK(f.is_open(), "cannot write " + filename);
  f << "view bin count: " << p.bins << "\n";
  for (int i = 0; i < p.bins; ++i) f << "view bin id: " << i << ", sources: " << p.sources[i] << ", total size: " << p.sizes[i] << "\n";
}

namespace {

double Ms(std::chrono::steady_clock::time_point a) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
}

const char* Flag(int argc, char** argv, const char* k) {
  for (int i = 1; i + 1 < argc; ++i)
    if (!strcmp(argv[i], k)) return argv[i + 1];
  return nullptr;
}

}  // namespace

int ViewPackingMain(int argc, char** argv) {
  if (argc == 1) return -1;
  auto str = [&](const char* k, const std::string& d) { return Flag(argc, argv, k) ? std::string(Flag(argc, argv, k)) : d; };
  auto num = [&](const char* k, double d) { return Flag(argc, argv, k) ? atof(Flag(argc, argv, k)) : d; };
  const std::string graph_file = str("-gf", "./data/com-dblp.ungraph.txt.bin");
  const std::string part_file = str("-pf", "./data/com-dblp.ungraph.txt.bin.kmeans.4");
  const int hop = (int)num("-h", 2);
  const double mem = num("-m", 10);
  const int threads = (int)num("-t", 16);
  const int pool_size = (int)num("-s", 1);
  const int degree_order = (int)num("-o", 0);
  const int dump = (int)num("-d", 1);
  // "<...>.K" or "<...>.KxM" names the partition's layers (view_packing.cc:147-160)
  const std::string suffix = part_file.substr(part_file.find_last_of('.') + 1);
  const size_t x = suffix.find('x');
  const int first = std::stoi(suffix.substr(0, x));
  const int second = x == std::string::npos ? 0 : std::stoi(suffix.substr(x + 1));
  if (first != threads) {
    std::cerr << "-t must equal the partition's first layer (" << first << ")" << std::endl;
    return 2;
  }
  const uint64_t max_entries = (uint64_t)((mem * 1000 * 1000 * 1000) / sizeof(uintV));
  std::cout << "mem: " << mem << "G, max view bin size: " << max_entries << std::endl;
  Graph g = ReadCSRBinFile(graph_file);
  // the degree cache beside the graph (view_packing.cc:40-114): written when
  // absent, checked against the graph when present
  const std::string degrees_file = graph_file + ".degrees.bin";
  if (FILE* f = fopen(degrees_file.c_str(), "rb")) {
    uint64_t n = 0;
    bool ok = fread(&n, sizeof(n), 1, f) == 1 && n == g.vertex_count;
    for (uint64_t v = 0; ok && v < n; ++v) {
      uint64_t d = 0;
      ok = fread(&d, sizeof(d), 1, f) == 1 && d == g.Degree((uintV)v);
    }
    fclose(f);
    if (!ok) throw Error(degrees_file + " does not match " + graph_file);
  } else {
    std::vector<uint64_t> out(g.vertex_count + 1);
    out[0] = g.vertex_count;
    for (uint64_t v = 0; v < g.vertex_count; ++v) out[v + 1] = g.Degree((uintV)v);
    FILE* w = fopen(degrees_file.c_str(), "wb");
    if (!w) throw Error("cannot write " + degrees_file);
    const bool ok = fwrite(out.data(), sizeof(uint64_t), out.size(), w) == out.size();
    if (fclose(w) != 0 || !ok) throw Error("write to " + degrees_file + " failed");
  }
  auto t0 = std::chrono::steady_clock::now();
  PackedBins p = PackViewBins(g, hop, ReadPartitionFile(part_file, g.vertex_count), first, second, max_entries, pool_size, degree_order != 0);
  std::cout << "total " << p.bins << " bins after merging (" << Ms(t0) << " ms)" << std::endl;
  if (dump) {
    const std::string base = graph_file + "." + std::to_string(hop) + ".hop";
    WriteVbstat(base + ".vbstat", p);
    WriteViewBinPartition(base + ".vbmap", p.bins, p.assignment);
  }
  return 0;
}

}  // namespace vsgm

extern "C" int vsgm_view_packing(int argc, char** argv) {
  try {
    return vsgm::ViewPackingMain(argc, argv);
  } catch (const std::exception& ex) {
    std::cerr << "view_packing: " << ex.what() << std::endl;
    return 1;
  }
}
