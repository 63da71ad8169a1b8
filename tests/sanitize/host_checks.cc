// Host code of libvsgm under AddressSanitizer and UndefinedBehaviorSanitizer
// (built with g++ by tests/test_host_sanitizers.py; no GPU is touched): the
// planner over every preset, clique and graphlet query, graph IO and
// generators, SNAP ingestion (multi-threaded parse, large ids, text
// writers, error paths), view-bin materialisation with reorder and split,
// view-bin packing, and the C-ABI's host calls. Any report aborts the run
// with a non-zero status.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "graph.h"
#include "ingest.h"
#include "tools.h"
#include "packing.h"
#include "plan.h"
#include "query.h"
#include "view.h"
#include "vsgm.h"

using namespace vsgm;

namespace {

int failures = 0;

void Expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "check failed: %s\n", what);
    ++failures;
  }
}

template <class F>
void ExpectThrow(F&& f, const char* what) {
  try {
    f();
  } catch (const std::exception&) {
    return;
  }
  Expect(false, what);
}

void Planner() {
  std::vector<Query> qs;
  for (int p = 0; p <= 18; ++p) qs.push_back(Query::Preset(p));
  for (int k = 3; k <= 7; ++k) qs.push_back(Query::Clique(k));
  for (int k = 3; k <= 5; ++k)
    for (auto& q : Query::Graphlets(k)) qs.push_back(q);
  for (auto& q : qs) {
    Plan p(q);
    p.Optimize();
    Expect(!p.Serialize().empty(), "plan text");
    Plan o = Plan::FromOrder(q, p.GetSearchSequence());
    Expect(o.GetHop() >= 1, "FromOrder hop");
  }
}

bool SameGraph(const Graph& a, const Graph& b) {
  return a.vertex_count == b.vertex_count && a.row_ptrs == b.row_ptrs && a.cols == b.cols;
}

void GraphsAndIngest(const std::string& dir) {
  Graph g = GenerateRMAT(10, 8, 0.57, 0.19, 0.19, 1, 4);
  Expect(g.Validate().empty(), "R-MAT graph valid");
  WriteCSRBinFile(dir + "/g.bin", g);
  Expect(SameGraph(ReadCSRBinFile(dir + "/g.bin"), g), ".bin round trip");
  Expect(GenerateChungLu(3000, 9000, 2.3, 2, 4).Validate().empty(), "Chung-Lu graph valid");
  {
    std::vector<uintV> id;
    Graph r = RelabelByDegree(g, &id, 3);
    Expect(r.Validate().empty() && r.cols.size() == g.cols.size() && id.size() == g.vertex_count, "relabelled graph valid");
    bool ascending = true;
    for (uint64_t v = 0; v + 1 < r.vertex_count; ++v) ascending = ascending && r.Degree((uintV)v) <= r.Degree((uintV)(v + 1));
    Expect(ascending, "relabelled degrees ascend");
    Expect(SameGraph(RelabelByDegree(g, nullptr, 1), r), "relabel independent of threads");
    Graph empty;
    empty.row_ptrs.assign(1, 0);
    Expect(RelabelByDegree(empty).vertex_count == 0, "relabel of an empty graph");
  }

  // SNAP text: comments, CRLF, trailing fields, a self loop, duplicates
  std::mt19937_64 rng(5);
  std::string text = "# comment\r\n% other\n\n";
  for (int i = 0; i < 4000; ++i) text += std::to_string(17 + rng() % 700) + "\t" + std::to_string(17 + rng() % 700) + " 1.0\r\n";
  text += "20 20\n";
  std::ofstream(dir + "/g.txt") << text;
  EdgeList ref = ReadSNAPEdges(dir + "/g.txt", 1);
  for (int t : {2, 3, 16, 64}) Expect(ReadSNAPEdges(dir + "/g.txt", t).edges == ref.edges, "parse independent of threads");
  Graph full = LoadGraph(dir + "/g.txt", -1, 4, false);
  Graph kept = LoadGraph(dir + "/g.txt", -1, 4, true);
  Expect(full.Validate().empty() && kept.Validate().empty(), "text graphs valid");
  Expect(SameGraph(DropIsolated(full), kept), "drop isolated");
  WriteSNAPText(dir + "/g.snap.txt", kept);
  WriteAdjText(dir + "/g.adj.txt", kept);
  Expect(SameGraph(LoadGraph(dir + "/g.snap.txt", -1, 2, true), kept), "SNAP text round trip");

  std::ofstream(dir + "/g.lvid") << "100 2000000000000000000000\n2000000000000000000000 7\n7 100\n";
  Graph tri = LoadGraph(dir + "/g.lvid");
  Expect(tri.vertex_count == 3 && tri.cols.size() == 6, "large-id triangle");

  std::ofstream(dir + "/tiny.txt") << "1 2\n";
  Expect(ReadSNAPEdges(dir + "/tiny.txt", 64).edges.size() == 1, "tiny file, many threads");
  std::ofstream(dir + "/bad.txt") << "1 2\n3\n";
  ExpectThrow([&] { ReadSNAPEdges(dir + "/bad.txt", 2); }, "malformed line rejected");
  std::ofstream(dir + "/odd.lvid") << "1 2 3\n";
  ExpectThrow([&] { ReadSNAPLargeVIDEdges(dir + "/odd.lvid"); }, "odd id count rejected");
  ExpectThrow([&] { ReadCSRBinFile(dir + "/missing.bin"); }, "missing file rejected");
  ExpectThrow([&] { LoadGraph(dir + "/x.unknown"); }, "unknown suffix rejected");
}

void ViewsAndPacking() {
  Graph g = GenerateRMAT(11, 8, 0.57, 0.19, 0.19, 3, 4);
  std::mt19937 rng(7);
  std::vector<int> part(g.vertex_count);
  for (auto& p : part) p = (int)(rng() % 4);
  for (int hop : {1, 2, 3}) {
    ViewBinManager vbm(&g, hop, 4);
    vbm.SetPartition(4, part);
    vbm.Reorder();
    vbm.Split(2, 2);
    uint64_t sources = 0;
    ViewBinHolder holder(g.vertex_count, 64ull << 20);
    for (auto& bin : vbm.GetViewBinPool()) {
      bin->Materialize(holder, hop == 2 ? 3 : 0);
      sources += bin->GetSources().size();
      Expect(holder.GetRowPtrs()[g.vertex_count] == holder.GetTotalSize(), "view row_ptrs end at the entry count");
    }
    Expect(sources == g.vertex_count, "bins cover every vertex");
  }
  const uint64_t limit = g.cols.size();
  PackedBins one = PackViewBins(g, 2, part, 4, 0, limit, 2, true, 4);
  std::vector<int> two_layer(g.vertex_count);
  for (auto& p : two_layer) p = (int)(rng() % 6);
  PackedBins two = PackViewBins(g, 2, two_layer, 2, 3, limit, 1, false, 2);
  for (const PackedBins* p : {&one, &two}) {
    uint64_t s = 0;
    for (uint64_t x : p->sources) s += x;
    Expect(s == g.vertex_count, "packed bins cover every vertex");
  }
  ExpectThrow([&] { PackViewBins(g, 2, part, 4, 0, 10, 1, false, 1); }, "oversized view rejected");
  Expect(ParseLevels("4") == std::vector<int>{4} && ParseLevels("4x2") == std::vector<int>({4, 2}) &&
             ParseLevels("3^2") == std::vector<int>({3, 3}),
         "cluster level strings");
  for (const char* bad : {"", "2^3", "4x", "x4", "17", "4y2", "99999999"})
    ExpectThrow([&] { ParseLevels(bad); }, "bad cluster level string rejected");
  Expect(ParseBisection("2^3") == 3 && ParseBisection("4x2") == 0 && ParseBisection("3^2") == 0, "bisection strings");
  for (const char* bad : {"2^0", "2^17", "2^x", "2^999"}) ExpectThrow([&] { ParseBisection(bad); }, "bad bisection string rejected");
  Graph closed = ClosedNeighbourhoods(g);
  Expect(closed.cols.size() == g.cols.size() + g.vertex_count, "closed rows add one entry per vertex");
  Graph opened = closed;
  Expect(OpenIfClosed(opened) && SameGraph(opened, g), "opening closed rows restores the graph");
  Graph plain = g;
  Expect(!OpenIfClosed(plain) && SameGraph(plain, g), "open rows stay as they are");
  std::vector<int> a = part;
  ReorderClusters(g, 4, a);
  Expect(RemoveEmptyClusters(4, a) == 4, "no empty clusters");
}

// The C-ABI's host calls and their error paths; creating an executor
// without a device must fail cleanly.
void CApi(const std::string& dir) {
  vsgm_graph* g = nullptr;
  Expect(vsgm_graph_rmat(9, 8, 0.57, 0.19, 0.19, 4, 2, &g) == 0, "rmat");
  Expect(vsgm_graph_write_bin(g, (dir + "/c.bin").c_str()) == 0, "write");
  vsgm_graph* r = nullptr;
  Expect(vsgm_graph_read_bin((dir + "/missing.bin").c_str(), &r) != 0 && vsgm_error()[0], "missing file reported");
  const uint32_t bad_pairs[2] = {0, 1u << 20};
  Expect(vsgm_graph_from_edges(10, bad_pairs, 1, &r) != 0, "out-of-range edge reported");
  vsgm_query* q = nullptr;
  Expect(vsgm_query_preset(2, &q) == 0, "preset");
  vsgm_plan* p = nullptr;
  Expect(vsgm_plan_optimize(q, &p) == 0, "optimize");
  size_t need = 0;
  vsgm_plan_serialize(p, nullptr, 0, &need);
  std::vector<char> buf(need + 1);
  Expect(vsgm_plan_serialize(p, buf.data(), buf.size(), &need) == 0, "serialize");
  const vsgm_plan* plans[1] = {p};
  vsgm_executor* e = nullptr;
  int devices = 0;
  vsgm_device_count(&devices);
  if (devices == 0) Expect(vsgm_executor_create(0, 512, plans, 1, 512, 0, &e) != 0 && e == nullptr, "no-device executor rejected");
  std::vector<int32_t> assign(512), part(512, 0);
  int bins = 0;
  Expect(vsgm_view_pack(g, 2, part.data(), 1, 0, 1ull << 40, 1, 0, 1, assign.data(), &bins, nullptr, nullptr, 0) == 0 && bins == 1,
         "pack");
  part[3] = 5;
  Expect(vsgm_view_pack(g, 2, part.data(), 1, 0, 1ull << 40, 1, 0, 1, assign.data(), &bins, nullptr, nullptr, 0) != 0,
         "bad partition id reported");
  // null handles and outputs are errors, not crashes
  Expect(vsgm_graph_write_bin(nullptr, (dir + "/n.bin").c_str()) != 0, "null graph reported");
  Expect(vsgm_graph_read_bin(nullptr, &r) != 0 && vsgm_graph_read_bin((dir + "/c.bin").c_str(), nullptr) != 0, "null path / out reported");
  Expect(vsgm_plan_optimize(nullptr, &p) != 0 && vsgm_query_preset(1, nullptr) != 0, "null query / out reported");
  Expect(vsgm_executor_match(nullptr, 1) != 0 && vsgm_executor_counts(nullptr, nullptr, 0) != 0, "null executor reported");
  // executor arguments are checked before the device is looked for
  const vsgm_plan* null_plans[1] = {nullptr};
  Expect(vsgm_executor_create(0, 512, nullptr, 1, 512, 0, &e) == VSGM_ERR_ARG && e == nullptr, "null plan list reported");
  Expect(vsgm_executor_create(0, 512, null_plans, 1, 512, 0, &e) == VSGM_ERR_ARG && e == nullptr, "null plan reported");
  Expect(vsgm_executor_create(0, 512, plans, 1, 512, 0, nullptr) == VSGM_ERR_ARG, "null executor out reported");
  Expect(vsgm_executor_create(0, 512, plans, -1, 512, 0, &e) == VSGM_ERR_ARG && e == nullptr, "negative plan count reported");
  uint32_t u32 = 0;
  uint64_t u64v = 0;
  Expect(vsgm_executor_hub_count(nullptr, &u32) != 0 && vsgm_executor_peak_bytes(nullptr, &u64v) != 0, "null executor accessors reported");
  Expect(vsgm_executor_get_plan_tuning(nullptr, 0, &u32, &u32, &u64v, &u32, &u32) != 0, "null executor tuning reported");
  int n_graphlets = 0;
  Expect(vsgm_query_graphlets(4, nullptr, 0, &n_graphlets) == 0 && n_graphlets == 6, "graphlet count without storage");
  vsgm_plan_free(p);
  vsgm_query_free(q);
  vsgm_graph_free(g);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  try {
    Planner();
    GraphsAndIngest(argv[1]);
    ViewsAndPacking();
    CApi(argv[1]);
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "unexpected exception: %s\n", ex.what());
    return 1;
  }
  if (failures) return 1;
  std::printf("host checks passed\n");
  return 0;
}
