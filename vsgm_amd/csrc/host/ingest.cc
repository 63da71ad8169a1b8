// SNAP edge-list ingestion and the preprocess_graph tool
// (src/preprocess/preprocess_graph.cc:9-36): parse the text, build the
// CSR (on the host, or on a GPU through vsgm_csr_from_pairs), drop self
// loops, parallel edges and isolated vertices, and write `<file>.bin`.
#include "ingest.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <unordered_map>

#include <omp.h>

#include "vsgm_hip.h"

namespace vsgm {
namespace {

bool IsDigit(char c) { return c >= '0' && c <= '9'; }
bool IsBlank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// Parse the edge lines of text[b, e) (b at a line start) into `out`.
// A line that does not start with a digit is skipped
// (graph_io.h:207-208); otherwise it holds two unsigned ids separated by
// blanks, and whatever follows them on the line is ignored.
void ParseRange(const char* text, uint64_t b, uint64_t e, std::vector<std::pair<uint32_t, uint32_t>>& out) {
  uint64_t i = b;
  while (i < e) {
    uint64_t eol = i;
    while (eol < e && text[eol] != '\n') ++eol;
    if (eol > i && IsDigit(text[i])) {
      uint64_t ids[2];
      uint64_t p = i;
      for (int k = 0; k < 2; ++k) {
        while (p < eol && IsBlank(text[p])) ++p;
        if (p == eol || !IsDigit(text[p]))
          throw Error("malformed edge line at byte " + std::to_string(i) + ": expected two vertex ids");
        uint64_t x = 0;
        while (p < eol && IsDigit(text[p])) {
          x = x * 10 + (uint64_t)(text[p] - '0');
          if (x > 0xFFFFFFFFull) throw Error("vertex id above 2^32-1 at byte " + std::to_string(i));
          ++p;
        }
        ids[k] = x;
      }
      out.emplace_back((uint32_t)ids[0], (uint32_t)ids[1]);
    }
    i = eol + 1;
  }
}

class MappedFile {
 public:
  explicit MappedFile(const std::string& path) {
    fd_ = open(path.c_str(), O_RDONLY);
    if (fd_ < 0) throw Error("cannot open " + path);
    struct stat st;
    if (fstat(fd_, &st) != 0) {
      close(fd_);
      throw Error("cannot stat " + path);
    }
    size_ = (uint64_t)st.st_size;
    if (size_) {
      void* p = mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd_, 0);
      if (p == MAP_FAILED) {
        close(fd_);
        throw Error("cannot map " + path);
      }
      data_ = (const char*)p;
    }
  }
  ~MappedFile() {
    if (data_) munmap((void*)data_, size_);
    close(fd_);
  }
  MappedFile(const MappedFile&) = delete;
  MappedFile& operator=(const MappedFile&) = delete;
  const char* data() const { return data_; }
  uint64_t size() const { return size_; }

 private:
  int fd_ = -1;
  const char* data_ = nullptr;
  uint64_t size_ = 0;
};

}  // namespace

EdgeList ReadSNAPEdges(const std::string& filename, int threads) {
  MappedFile f(filename);
  const int nt = std::max(1, threads > 0 ? threads : omp_get_max_threads());
  // chunk boundaries moved forward to line starts
  std::vector<uint64_t> cut(nt + 1);
  for (int t = 0; t <= nt; ++t) {
    uint64_t c = f.size() * (uint64_t)t / (uint64_t)nt;
    if (t > 0 && t < nt)
      while (c > 0 && c < f.size() && f.data()[c - 1] != '\n') ++c;
    cut[t] = c;
  }
  for (int t = 1; t <= nt; ++t) cut[t] = std::max(cut[t], cut[t - 1]);
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> part(nt);
  std::vector<std::string> err(nt);
#pragma omp parallel for num_threads(nt) schedule(static, 1)
  for (int t = 0; t < nt; ++t) {
    try {
      ParseRange(f.data(), cut[t], cut[t + 1], part[t]);
    } catch (const std::exception& ex) {
      err[t] = ex.what();
    }
  }
  for (auto& e : err)
    if (!e.empty()) throw Error(filename + ": " + e);
  EdgeList out;
  uint64_t n = 0;
  uint32_t lo = 0xFFFFFFFFu, hi = 0;
  for (auto& p : part) {
    n += p.size();
    for (auto& e : p) {
      lo = std::min(lo, std::min(e.first, e.second));
      hi = std::max(hi, std::max(e.first, e.second));
    }
  }
  if (n == 0) throw Error(filename + ": no edge lines");
  // ids are shifted so that the smallest becomes 0 (graph_io.h:218, :233)
  out.vertex_count = (uint64_t)hi - lo + 1;
  out.edges.reserve(n);
  for (auto& p : part) {
    for (auto& e : p) out.edges.emplace_back(e.first - lo, e.second - lo);
    std::vector<std::pair<uint32_t, uint32_t>>().swap(p);
  }
  return out;
}

EdgeList ReadSNAPLargeVIDEdges(const std::string& filename) {
  MappedFile f(filename);
  const char* t = f.data();
  const uint64_t n = f.size();
  std::unordered_map<std::string, uint32_t> ids;
  std::vector<uint32_t> ends;
  uint64_t i = 0;
  while (i < n) {
    uint64_t eol = i;
    while (eol < n && t[eol] != '\n') ++eol;
    if (eol > i && IsDigit(t[i])) {
      // every maximal run of digits on the line is one id, numbered in order of first appearance
      for (uint64_t p = i; p < eol;) {
        if (!IsDigit(t[p])) {
          ++p;
          continue;
        }
        uint64_t q = p;
        while (q < eol && IsDigit(t[q])) ++q;
        auto it = ids.emplace(std::string(t + p, q - p), (uint32_t)ids.size()).first;
        if (ids.size() > 0xFFFFFFFFull) throw Error(filename + ": more than 2^32-1 distinct vertex ids");
        ends.push_back(it->second);
        p = q;
      }
    }
    i = eol + 1;
  }
  if (ends.size() % 2) throw Error(filename + ": odd number of vertex ids");
  if (ends.empty()) throw Error(filename + ": no edge lines");
  EdgeList out;
  out.vertex_count = ids.size();
  out.edges.resize(ends.size() / 2);
  for (size_t e = 0; e < out.edges.size(); ++e) out.edges[e] = {ends[2 * e], ends[2 * e + 1]};
  return out;
}

Graph DropIsolated(const Graph& g) {
  // every list entry must point at a vertex whose own list is not empty
  for (uint32_t c : g.cols) VSGM_CHECK(g.row_ptrs[c + 1] > g.row_ptrs[c], "isolated-vertex removal needs a symmetric graph");
  std::vector<uint32_t> id(g.vertex_count);
  uint64_t kept = 0;
  for (uint64_t u = 0; u < g.vertex_count; ++u)
    if (g.row_ptrs[u + 1] > g.row_ptrs[u]) id[u] = (uint32_t)kept++;
  Graph out;
  out.vertex_count = kept;
  out.row_ptrs.reserve(kept + 1);
  for (uint64_t u = 0; u < g.vertex_count; ++u)
    if (g.row_ptrs[u + 1] > g.row_ptrs[u]) out.row_ptrs.push_back(g.row_ptrs[u]);
  out.row_ptrs.push_back(g.row_ptrs[g.vertex_count]);
  out.cols.resize(g.cols.size());
  for (uint64_t i = 0; i < g.cols.size(); ++i) out.cols[i] = id[g.cols[i]];
  return out;
}

Graph BuildOnDevice(int device, uint64_t vertex_count, const std::vector<std::pair<uint32_t, uint32_t>>& edges, bool drop_isolated) {
  vsgm_ctx* ctx = nullptr;
  if (vsgm_ctx_create(device, 0, &ctx) != VSGM_OK) throw Error("cannot open GPU " + std::to_string(device));
  void* pairs = nullptr;
  auto check = [&](int s, const char* what) {
    if (s != VSGM_OK) {
      std::string msg = std::string(what) + " failed: " + vsgm_last_error(ctx);
      vsgm_ctx_destroy(ctx);
      throw Error(msg);
    }
  };
  static_assert(sizeof(std::pair<uint32_t, uint32_t>) == 8, "edge pairs must be two packed u32");
  check(vsgm_alloc(ctx, std::max<uint64_t>(edges.size(), 1) * 8, &pairs), "edge upload");
  check(vsgm_h2d(ctx, pairs, edges.data(), edges.size() * 8), "edge upload");
  vsgm_view_out o;
  check(vsgm_csr_from_pairs(ctx, (const uint32_t*)pairs, edges.size(), vertex_count, 1, drop_isolated ? 1 : 0, &o), "CSR build");
  Graph g;
  g.vertex_count = o.vertices;
  g.row_ptrs.resize(o.vertices + 1);
  g.cols.resize(o.entries);
  check(vsgm_d2h(ctx, g.row_ptrs.data(), o.row_ptrs, g.row_ptrs.size() * sizeof(uint64_t)), "CSR download");
  check(vsgm_d2h(ctx, g.cols.data(), o.cols, g.cols.size() * sizeof(uint32_t)), "CSR download");
  vsgm_ctx_destroy(ctx);  // frees the pool
  return g;
}

Graph ReadSNAPFile(const std::string& filename, bool directed, int device, int threads) {
  // The matcher reads undirected graphs only; the reference's renumbering
  // (preprocess_graph.h:67-82) would also leave the ids of directed
  // targets without out-edges out of range.
  if (directed) throw Error("directed edge lists are not supported; pattern matching reads undirected graphs");
  EdgeList el = ReadSNAPEdges(filename, threads);
  if (device >= 0) return BuildOnDevice(device, el.vertex_count, el.edges, true);
  return DropIsolated(BuildFromEdges(el.vertex_count, el.edges, threads));
}

Graph LoadGraph(const std::string& filename, int device, int threads, bool drop_isolated) {
  const std::string suffix = filename.substr(filename.rfind('.') + 1);
  if (suffix == "bin") {
    Graph g = ReadCSRBinFile(filename);
    if (!drop_isolated) return g;  // moved, not copied
    return DropIsolated(g);
  }
  EdgeList el;
  if (suffix == "txt")
    el = ReadSNAPEdges(filename, threads);
  else if (suffix == "lvid")
    el = ReadSNAPLargeVIDEdges(filename);
  else
    throw Error("unknown graph file suffix ." + suffix + " (expected .bin, .txt or .lvid)");
  if (device >= 0) return BuildOnDevice(device, el.vertex_count, el.edges, drop_isolated);
  Graph g = BuildFromEdges(el.vertex_count, el.edges, threads);
  if (!drop_isolated) return g;
  return DropIsolated(g);
}

void WriteSNAPText(const std::string& filename, const Graph& g) {
  std::ofstream f(filename);
  if (!f) throw Error("cannot write " + filename);
  for (uint64_t u = 0; u < g.vertex_count; ++u)
    for (uint64_t j = g.row_ptrs[u]; j < g.row_ptrs[u + 1]; ++j)
      if (u < g.cols[j]) f << u << " " << g.cols[j] << "\n";
  if (!f) throw Error("write to " + filename + " failed");
}

void WriteAdjText(const std::string& filename, const Graph& g) {
  std::ofstream f(filename);
  if (!f) throw Error("cannot write " + filename);
  for (uint64_t u = 0; u < g.vertex_count; ++u) {
    f << u << "\t";
    for (uint64_t j = g.row_ptrs[u]; j < g.row_ptrs[u + 1]; ++j) f << g.cols[j] << " ";
    f << "\n";
  }
  if (!f) throw Error("write to " + filename + " failed");
  std::ofstream info(filename + ".info");
  info << g.vertex_count << " " << g.cols.size();
  if (!info) throw Error("write to " + filename + ".info failed");
}

namespace {

struct Args {
  int argc;
  char** argv;
  const char* raw(const char* k) const {
    for (int i = 1; i + 1 < argc; ++i)
      if (!strcmp(argv[i], k)) return argv[i + 1];
    return nullptr;
  }
  std::string str(const char* k, const std::string& d) const { return raw(k) ? raw(k) : d; }
  int i(const char* k, int d) const { return raw(k) ? atoi(raw(k)) : d; }
};

double Ms(std::chrono::steady_clock::time_point a) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
}

}  // namespace

int PreprocessGraphMain(int argc, char** argv) {
  if (argc == 1) {
    std::cout << "./preprocess_graph -f GRAPH_FILENAME(SNAP format) -d DIRECTED -o OUTPUT_FORMAT [-g GPU] [-t THREADS] [-rl 1]" << std::endl;
    std::cout << "OUTPUT_FORMAT: 0 for .snap.txt, 1 (Default) for .bin, 3 for .adj.txt; -g >= 0 builds the CSR on that GPU;" << std::endl;
    std::cout << "-rl 1 numbers the vertices in ascending degree order (an isomorphic graph with the same pattern counts)" << std::endl;
    return -1;
  }
  Args a{argc, argv};
  const std::string filename = a.str("-f", "./data/com-dblp.ungraph.txt");
  const int directed = a.i("-d", 0);
  const int format = a.i("-o", 1);
  const int device = a.i("-g", -1);
  const int threads = a.i("-t", 0);
  const int relabel = a.i("-rl", 0);
  // GraphFormat numbering (graph_io.h:22): SNAP 0, BIN 1, SNAP_LVID 2 (an
  // input format only), ADJ 3
  if (format != 0 && format != 1 && format != 3) {
    std::cerr << "unknown output format " << format << " (0 .snap.txt, 1 .bin, 3 .adj.txt)" << std::endl;
    return 2;
  }
  if (directed) throw Error("directed edge lists are not supported; pattern matching reads undirected graphs");
  auto t0 = std::chrono::steady_clock::now();
  Graph g = LoadGraph(filename, device, threads, true);
  if (relabel) g = RelabelByDegree(g, nullptr, threads);
  std::cout << "finish preprocess, time=" << Ms(t0) << "ms, now vertex_count=" << g.vertex_count << ",edge_count=" << g.cols.size()
            << std::endl;
  const std::string bad = g.Validate();
  if (!bad.empty()) throw Error("preprocessed graph is invalid: " + bad);
  // output names of GraphIO::WriteDataFile (graph_io.h:297-384)
  const std::string out = filename + (format == 0 ? ".snap.txt" : format == 1 ? ".bin" : ".adj.txt");
  if (format == 0)
    WriteSNAPText(out, g);
  else if (format == 1)
    WriteCSRBinFile(out, g);
  else
    WriteAdjText(out, g);
  std::cout << "wrote " << out << std::endl;
  return 0;
}

}  // namespace vsgm

extern "C" int vsgm_preprocess_graph(int argc, char** argv) {
  try {
    return vsgm::PreprocessGraphMain(argc, argv);
  } catch (const std::exception& ex) {
    std::cerr << "preprocess_graph: " << ex.what() << std::endl;
    return 1;
  }
}
