// C-ABI of the host library (include/vsgm.h) over the C++ classes.
#include "vsgm.h"

#include <cstring>
#include <memory>
#include <string>

#include "executor.h"
#include "graph.h"
#include "ingest.h"
#include "tools.h"
#include "packing.h"
#include "plan.h"
#include "query.h"
#include "view.h"

struct vsgm_graph {
  vsgm::Graph g;
};
struct vsgm_query {
  vsgm::Query q;
};
struct vsgm_plan {
  std::unique_ptr<vsgm::Plan> p;
};
struct vsgm_view_manager {
  std::unique_ptr<vsgm::ViewBinManager> m;
};
struct vsgm_view_holder {
  std::unique_ptr<vsgm::ViewBinHolder> h;
};
struct vsgm_executor {
  std::unique_ptr<vsgm::PipelineExecutor> e;
};

namespace {
thread_local std::string g_error;

template <typename F>
int Guard(F&& f) {
  try {
    g_error.clear();
    f();
    return VSGM_OK;
  } catch (const std::bad_alloc&) {
    g_error = "host allocation failed";
    return VSGM_ERR_OOM;
  } catch (const std::exception& ex) {
    g_error = ex.what();
    return VSGM_ERR_ARG;
  }
}
}  // namespace

extern "C" {

const char* vsgm_error(void) { return g_error.c_str(); }

int vsgm_graph_read_bin(const char* path, vsgm_graph** out) {
  return Guard([&] {
    VSGM_CHECK(path && out, "null argument");
    *out = new vsgm_graph{vsgm::ReadCSRBinFile(path)};
  });
}

int vsgm_graph_write_bin(const vsgm_graph* g, const char* path) {
  return Guard([&] {
    VSGM_CHECK(g && path, "null argument");
    vsgm::WriteCSRBinFile(path, g->g);
  });
}

int vsgm_graph_from_csr(uint64_t vertex_count, const uint64_t* row_ptrs, const uint32_t* cols, vsgm_graph** out) {
  return Guard([&] {
    VSGM_CHECK(out, "null argument");
    auto* h = new vsgm_graph;
    h->g.vertex_count = vertex_count;
    h->g.row_ptrs.assign(row_ptrs, row_ptrs + vertex_count + 1);
    h->g.cols.assign(cols, cols + row_ptrs[vertex_count]);
    *out = h;
  });
}

int vsgm_graph_from_edges(uint64_t vertex_count, const uint32_t* pairs, uint64_t n_pairs, vsgm_graph** out) {
  return Guard([&] {
    VSGM_CHECK(out, "null argument");
    std::vector<std::pair<vsgm::uintV, vsgm::uintV>> edges(n_pairs);
    for (uint64_t i = 0; i < n_pairs; ++i) {
      VSGM_CHECK(pairs[2 * i] < vertex_count && pairs[2 * i + 1] < vertex_count, "edge endpoint out of range");
      edges[i] = {pairs[2 * i], pairs[2 * i + 1]};
    }
    *out = new vsgm_graph{vsgm::BuildFromEdges(vertex_count, edges)};
  });
}

int vsgm_graph_read_snap(const char* path, int directed, int device, int threads, vsgm_graph** out) {
  return Guard([&] {
    VSGM_CHECK(path && out, "null argument");
    *out = new vsgm_graph{vsgm::ReadSNAPFile(path, directed != 0, device, threads)};
  });
}

int vsgm_graph_load(const char* path, int device, int threads, int drop_isolated, vsgm_graph** out) {
  return Guard([&] {
    VSGM_CHECK(path && out, "null argument");
    *out = new vsgm_graph{vsgm::LoadGraph(path, device, threads, drop_isolated != 0)};
  });
}

int vsgm_graph_drop_isolated(const vsgm_graph* g, vsgm_graph** out) {
  return Guard([&] {
    VSGM_CHECK(g && out, "null argument");
    *out = new vsgm_graph{vsgm::DropIsolated(g->g)};
  });
}

int vsgm_graph_relabel_degree(const vsgm_graph* g, int threads, uint32_t* new_id, vsgm_graph** out) {
  return Guard([&] {
    VSGM_CHECK(g && out, "null argument");
    std::vector<vsgm::uintV> id;
    *out = new vsgm_graph{vsgm::RelabelByDegree(g->g, new_id ? &id : nullptr, threads)};
    if (new_id) std::copy(id.begin(), id.end(), new_id);
  });
}

int vsgm_graph_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed, int threads, vsgm_graph** out) {
  return Guard([&] {
    VSGM_CHECK(out, "null argument");
    *out = new vsgm_graph{vsgm::GenerateRMAT(scale, edge_factor, a, b, c, seed, threads)};
  });
}

int vsgm_graph_chung_lu(uint64_t vertex_count, uint64_t draws, double gamma, uint64_t seed, int threads, vsgm_graph** out) {
  return Guard([&] {
    VSGM_CHECK(out, "null argument");
    *out = new vsgm_graph{vsgm::GenerateChungLu(vertex_count, draws, gamma, seed, threads)};
  });
}

int vsgm_graph_info(const vsgm_graph* g, uint64_t* vertex_count, uint64_t* entries, uint64_t* max_degree) {
  return Guard([&] {
    VSGM_CHECK(g, "null argument");
    if (vertex_count) *vertex_count = g->g.vertex_count;
    if (entries) *entries = g->g.cols.size();
    if (max_degree) *max_degree = g->g.MaxDegree();
  });
}

int vsgm_graph_csr(const vsgm_graph* g, const uint64_t** row_ptrs, const uint32_t** cols) {
  return Guard([&] {
    VSGM_CHECK(g && row_ptrs && cols, "null argument");
    *row_ptrs = g->g.row_ptrs.data();
    *cols = g->g.cols.data();
  });
}

int vsgm_graph_validate(const vsgm_graph* g, char* msg, size_t len) {
  return Guard([&] {
    VSGM_CHECK(g, "null argument");
    std::string m = g->g.Validate(true);
    if (len) {
      strncpy(msg, m.c_str(), len - 1);
      msg[len - 1] = 0;
    }
  });
}

void vsgm_graph_free(vsgm_graph* g) { delete g; }

int vsgm_query_preset(int preset, vsgm_query** out) {
  return Guard([&] {
    VSGM_CHECK(out, "null argument");
    *out = new vsgm_query{vsgm::Query::Preset(preset)};
  });
}

int vsgm_query_clique(int k, vsgm_query** out) {
  return Guard([&] {
    VSGM_CHECK(out, "null argument");
    *out = new vsgm_query{vsgm::Query::Clique(k)};
  });
}

int vsgm_query_from_matrix(int n, const uint8_t* adj, vsgm_query** out) {
  return Guard([&] {
    VSGM_CHECK(out, "null argument");
    std::vector<std::vector<int>> lists(n);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        if (i != j && adj[i * n + j]) {
          VSGM_CHECK(adj[j * n + i], "adjacency matrix must be symmetric");
          lists[i].push_back(j);
        }
    *out = new vsgm_query{vsgm::Query::FromAdjacency(lists)};
  });
}

int vsgm_query_graphlets(int k, vsgm_query** out, int cap, int* n_out) {
  return Guard([&] {
    VSGM_CHECK(n_out && (out || cap <= 0), "null argument");
    auto qs = vsgm::Query::Graphlets(k);
    *n_out = (int)qs.size();
    for (int i = 0; i < (int)qs.size() && i < cap; ++i) out[i] = new vsgm_query{qs[i]};
  });
}

int vsgm_query_info(const vsgm_query* q, int* n, int* edges, uint8_t* adj, uint8_t* cond) {
  return Guard([&] {
    VSGM_CHECK(q, "null argument");
    int k = (int)q->q.GetVertexCount();
    if (n) *n = k;
    if (edges) *edges = (int)q->q.GetEdgeCount();
    if (adj) {
      memset(adj, 0, k * k);
      for (int i = 0; i < k; ++i)
        for (auto j : q->q.GetConnectivity()[i]) adj[i * k + j] = 1;
    }
    if (cond)
      for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) cond[i * k + j] = (uint8_t)vsgm::GetConditionType(i, j, q->q.GetOrder());
  });
}

void vsgm_query_free(vsgm_query* q) { delete q; }

int vsgm_plan_optimize(const vsgm_query* q, vsgm_plan** out) {
  return Guard([&] {
    VSGM_CHECK(q && out, "null argument");
    auto p = std::make_unique<vsgm::Plan>(q->q);
    p->Optimize();
    *out = new vsgm_plan{std::move(p)};
  });
}

int vsgm_plan_from_order(const vsgm_query* q, const uint32_t* seq, int n, vsgm_plan** out) {
  return Guard([&] {
    VSGM_CHECK(q && out, "null argument");
    std::vector<vsgm::uintV> s(seq, seq + n);
    *out = new vsgm_plan{std::make_unique<vsgm::Plan>(vsgm::Plan::FromOrder(q->q, s))};
  });
}

int vsgm_plan_info(const vsgm_plan* p, int* hop, uint64_t* root_degree, int* n_ops) {
  return Guard([&] {
    VSGM_CHECK(p, "null argument");
    if (hop) *hop = p->p->GetHop();
    if (root_degree) *root_degree = p->p->GetRootDegree();
    if (n_ops) *n_ops = (int)p->p->GetExecuteOperations().size();
  });
}

int vsgm_plans_shared_prefix(const vsgm_plan* a, const vsgm_plan* b, int* levels) {
  return Guard([&] {
    VSGM_CHECK(a && b && levels, "null argument");
    *levels = (int)vsgm::SharedPrefix(*a->p, *b->p);
  });
}

int vsgm_plan_op(const vsgm_plan* p, int i, int* op, int* vertex) {
  return Guard([&] {
    VSGM_CHECK(p, "null argument");
    const auto& ops = p->p->GetExecuteOperations();
    VSGM_CHECK(i >= 0 && i < (int)ops.size(), "operation index out of range");
    *op = (int)ops[i].first;
    *vertex = (int)ops[i].second;
  });
}

int vsgm_plan_order(const vsgm_plan* p, uint32_t* seq, int cap) {
  return Guard([&] {
    VSGM_CHECK(p, "null argument");
    const auto& s = p->p->GetSearchSequence();
    VSGM_CHECK(cap >= (int)s.size(), "order buffer too small");
    for (size_t i = 0; i < s.size(); ++i) seq[i] = s[i];
  });
}

int vsgm_plan_serialize(const vsgm_plan* p, char* buf, size_t len, size_t* needed) {
  return Guard([&] {
    VSGM_CHECK(p, "null argument");
    std::string s = p->p->Serialize();
    if (needed) *needed = s.size() + 1;
    if (buf && len) {
      size_t n = std::min(len - 1, s.size());
      memcpy(buf, s.data(), n);
      buf[n] = 0;
    }
  });
}

void vsgm_plan_free(vsgm_plan* p) { delete p; }

int vsgm_view_pack(const vsgm_graph* g, int hop, const int32_t* partition, int first_layer, int second_layer, uint64_t max_bin_entries,
                   int pool_size, int degree_order, int threads, int32_t* out_assign, int* out_bins, uint64_t* out_sources,
                   uint64_t* out_sizes, int cap) {
  return Guard([&] {
    VSGM_CHECK(g, "null argument");
    VSGM_CHECK(g && partition && out_assign && out_bins, "null argument");
    const uint64_t V = g->g.vertex_count;
    vsgm::PackedBins p = vsgm::PackViewBins(g->g, hop, std::vector<int>(partition, partition + V), first_layer, second_layer,
                                            max_bin_entries, pool_size, degree_order != 0, threads);
    *out_bins = p.bins;
    VSGM_CHECK((!out_sources && !out_sizes) || cap >= p.bins, "bin statistics need room for " + std::to_string(p.bins) + " bins");
    std::copy(p.assignment.begin(), p.assignment.end(), out_assign);
    if (out_sources) std::copy(p.sources.begin(), p.sources.end(), out_sources);
    if (out_sizes) std::copy(p.sizes.begin(), p.sizes.end(), out_sizes);
  });
}

int vsgm_graph_kmeans(const vsgm_graph* g, const char* levels, int iters, uint64_t seed, int device, int reorder, int32_t* assignment,
                      int* clusters) {
  return Guard([&] {
    VSGM_CHECK(g && levels, "null argument");
    VSGM_CHECK(g && levels && assignment && clusters && iters >= 0, "bad arguments");
    std::vector<int> a = vsgm::GpuKMeansSpec(g->g, levels, iters, seed, device, reorder != 0, clusters);
    std::copy(a.begin(), a.end(), assignment);
  });
}

int vsgm_graph_fennel(const vsgm_graph* g, int parts, int hop, int32_t* assignment) {
  return Guard([&] {
    VSGM_CHECK(g, "null argument");
    VSGM_CHECK(g && assignment, "bad arguments");
    std::vector<int> a = vsgm::FennelPartition(g->g, parts, hop);
    std::copy(a.begin(), a.end(), assignment);
  });
}

int vsgm_clusters_reorder(const vsgm_graph* g, int k, int32_t* assignment) {
  return Guard([&] {
    VSGM_CHECK(g, "null argument");
    VSGM_CHECK(g && assignment && k >= 1, "bad arguments");
    const uint64_t V = g->g.vertex_count;
    std::vector<int> a(assignment, assignment + V);
    for (int x : a) VSGM_CHECK(x >= 0 && x < k, "cluster id out of range");
    vsgm::ReorderClusters(g->g, k, a);
    std::copy(a.begin(), a.end(), assignment);
  });
}

int vsgm_clusters_compact(int k, int32_t* assignment, uint64_t vertex_count, int* kept) {
  return Guard([&] {
    VSGM_CHECK(assignment && kept && k >= 1, "bad arguments");
    std::vector<int> a(assignment, assignment + vertex_count);
    for (int x : a) VSGM_CHECK(x >= 0 && x < k, "cluster id out of range");
    *kept = vsgm::RemoveEmptyClusters(k, a);
    std::copy(a.begin(), a.end(), assignment);
  });
}

int vsgm_vbmap_write(const char* path, int n_bins, const int* assignment, uint64_t vertex_count) {
  return Guard([&] {
    VSGM_CHECK(path, "null argument");
    vsgm::WriteViewBinPartition(path, n_bins, std::vector<int>(assignment, assignment + vertex_count));
  });
}

int vsgm_view_manager_load(const vsgm_graph* g, int hop, int threads, const char* vbmap, vsgm_view_manager** out) {
  return Guard([&] {
    VSGM_CHECK(g && vbmap && out, "null argument");
    auto m = std::make_unique<vsgm::ViewBinManager>(&g->g, hop, threads);
    m->LoadViewBinPartition(vbmap);
    *out = new vsgm_view_manager{std::move(m)};
  });
}

int vsgm_view_manager_from_assignment(const vsgm_graph* g, int hop, int threads, int n_bins, const int* assignment,
                                      vsgm_view_manager** out) {
  return Guard([&] {
    VSGM_CHECK(g && out, "null argument");
    auto m = std::make_unique<vsgm::ViewBinManager>(&g->g, hop, threads);
    m->SetPartition(n_bins, std::vector<int>(assignment, assignment + g->g.vertex_count));
    *out = new vsgm_view_manager{std::move(m)};
  });
}

int vsgm_view_manager_info(const vsgm_view_manager* m, int* n_bins, uint64_t* max_sources) {
  return Guard([&] {
    VSGM_CHECK(m, "null argument");
    if (n_bins) *n_bins = m->m->GetViewBinNumber();
    if (max_sources) *max_sources = m->m->GetMaxPartitionedSourcesNum();
  });
}

int vsgm_view_manager_reorder(vsgm_view_manager* m) {
  return Guard([&] {
    VSGM_CHECK(m, "null argument");
    m->m->Reorder();
  });
}

int vsgm_view_manager_split(vsgm_view_manager* m, int devices, int times) {
  return Guard([&] {
    VSGM_CHECK(m, "null argument");
    m->m->Split(devices, times);
  });
}

void vsgm_view_manager_free(vsgm_view_manager* m) { delete m; }

int vsgm_view_holder_create(uint64_t vertex_count, uint64_t max_bytes, vsgm_view_holder** out) {
  return Guard([&] {
    VSGM_CHECK(out, "null argument");
    *out = new vsgm_view_holder{std::make_unique<vsgm::ViewBinHolder>(vertex_count, max_bytes)};
  });
}

int vsgm_view_materialize(const vsgm_view_manager* m, int index, vsgm_view_holder* h, uint64_t root_degree) {
  return Guard([&] {
    VSGM_CHECK(m && h, "null argument");
    const auto& pool = m->m->GetViewBinPool();
    VSGM_CHECK(index >= 0 && index < (int)pool.size(), "view bin index out of range");
    pool[index]->Materialize(*h->h, root_degree);
  });
}

int vsgm_view_holder_info(const vsgm_view_holder* h, int* bin_id, uint64_t* sources, uint64_t* entries, uint64_t* vertices) {
  return Guard([&] {
    VSGM_CHECK(h, "null argument");
    if (bin_id) *bin_id = h->h->GetId();
    if (sources) *sources = h->h->GetSourcesNum();
    if (entries) *entries = h->h->GetTotalSize();
    if (vertices) *vertices = h->h->GetVerticesNum();
  });
}

int vsgm_view_holder_csr(const vsgm_view_holder* h, const uint64_t** row_ptrs, const uint32_t** cols, const uint32_t** sources) {
  return Guard([&] {
    VSGM_CHECK(h && row_ptrs && cols && sources, "null argument");
    if (row_ptrs) *row_ptrs = h->h->GetRowPtrs();
    if (cols) *cols = h->h->GetCols();
    if (sources) *sources = h->h->GetSources().data();
  });
}

void vsgm_view_holder_free(vsgm_view_holder* h) { delete h; }

int vsgm_executor_create(int device, uint64_t vertex_count, const vsgm_plan* const* plans, int n_plans, uint64_t max_sources,
                         uint64_t mem_budget, vsgm_executor** out) {
  int code = VSGM_OK;
  int s = Guard([&] {
    VSGM_CHECK(out, "null argument");
    VSGM_CHECK(n_plans >= 0 && (n_plans == 0 || plans), "null plans");
    std::vector<const vsgm::Plan*> ps;
    for (int i = 0; i < n_plans; ++i) {
      VSGM_CHECK(plans[i], "null plan " + std::to_string(i));
      ps.push_back(plans[i]->p.get());
    }
    vsgm::ExecutorOptions opt;
    opt.mem_budget = mem_budget;
    int n_dev = 0;
    vsgm_device_count(&n_dev);
    if (n_dev == 0) {
      code = VSGM_ERR_NO_DEVICE;
      throw vsgm::Error("no HIP device is visible: the enumerator needs a GPU");
    }
    *out = new vsgm_executor{std::make_unique<vsgm::PipelineExecutor>(device, vertex_count, ps, max_sources, opt)};
  });
  return code != VSGM_OK ? code : s;
}

int vsgm_executor_set_option(vsgm_executor* e, const char* key, double value) {
  return Guard([&] {
    VSGM_CHECK(e && key, "null argument");
    auto& o = e->e->Options();
    std::string k = key;
    if (k == "batch_bytes") o.batch_bytes = (uint64_t)value;
    else if (k == "min_batch_bytes") o.min_batch_bytes = (uint64_t)value;
    else if (k == "profile") o.profile = value != 0;
    else if (k == "hub_bytes") o.hub_bytes = value < 0 ? ~0ull : (uint64_t)value;
    else if (k == "hub_min_degree") o.hub_min_degree = (uint64_t)value;
    else if (k == "prefetch_bytes") o.prefetch_bytes = (uint64_t)value;
    else if (k == "share_prefixes") o.share_prefixes = value != 0;
    else if (k == "isect_kernel" || k == "max_grid" || k == "cold_weight" || k == "ticket_chunks") {
      // applies to every plan (the per-plan tuning of the autotuner refines it)
      vsgm_tuning t;
      vsgm_get_tuning(e->e->Context(), &t);
      VSGM_CHECK(value >= 0 && value <= 4294967295.0, "bad tuning value for " + k);
      if (k == "isect_kernel") t.isect_kernel = (uint32_t)value;
      else if (k == "max_grid") t.max_grid = (uint32_t)value;
      else if (k == "cold_weight") t.cold_weight = (uint32_t)value;
      else t.ticket_chunks = (uint32_t)value;
      if (vsgm_set_tuning(e->e->Context(), &t) != VSGM_OK) throw vsgm::Error("bad tuning value for " + k);
      for (size_t i = 0; i < e->e->GetTotalCounts().size(); ++i) {
        vsgm::PlanTuning pt = e->e->GetPlanTuning(i);
        pt.isect_kernel = t.isect_kernel;
        pt.max_grid = t.max_grid;
        pt.cold_weight = t.cold_weight;
        pt.ticket_chunks = t.ticket_chunks;
        e->e->SetPlanTuning(i, pt);
      }
    } else throw vsgm::Error("unknown executor option " + k);
  });
}

int vsgm_executor_set_plan_tuning(vsgm_executor* e, int plan, uint32_t isect_kernel, uint32_t max_grid, uint64_t batch_bytes,
                                  uint32_t cold_weight, uint32_t ticket_chunks) {
  return Guard([&] {
    VSGM_CHECK(e, "null argument");
    VSGM_CHECK(plan >= 0 && (size_t)plan < e->e->GetTotalCounts().size(), "plan index out of range");
    if (isect_kernel > 4 || max_grid == 0 || max_grid > (1u << 20)) throw vsgm::Error("bad plan tuning");
    if (cold_weight == 0 || cold_weight > 1024 || ticket_chunks == 0 || ticket_chunks > 64) throw vsgm::Error("bad plan tuning");
    e->e->SetPlanTuning((size_t)plan, vsgm::PlanTuning{isect_kernel, max_grid, batch_bytes, cold_weight, ticket_chunks});
  });
}

int vsgm_executor_get_plan_tuning(const vsgm_executor* e, int plan, uint32_t* isect_kernel, uint32_t* max_grid, uint64_t* batch_bytes,
                                  uint32_t* cold_weight, uint32_t* ticket_chunks) {
  return Guard([&] {
    VSGM_CHECK(e && isect_kernel && max_grid && batch_bytes && cold_weight && ticket_chunks, "null argument");
    VSGM_CHECK(plan >= 0 && (size_t)plan < e->e->GetTotalCounts().size(), "plan index out of range");
    const auto& t = e->e->GetPlanTuning((size_t)plan);
    *isect_kernel = t.isect_kernel;
    *max_grid = t.max_grid;
    *batch_bytes = t.batch_bytes;
    *cold_weight = t.cold_weight;
    *ticket_chunks = t.ticket_chunks;
  });
}

int vsgm_executor_add_view(vsgm_executor* e, const vsgm_view_holder* h, int* index) {
  return Guard([&] {
    VSGM_CHECK(e && h, "null argument");
    size_t i = e->e->AddView(*h->h);
    if (index) *index = (int)i;
  });
}

int vsgm_executor_set_graph(vsgm_executor* e, const vsgm_graph* g) {
  return Guard([&] {
    VSGM_CHECK(e && g, "null argument");
    e->e->SetGraph(g->g);
  });
}

int vsgm_executor_transfer_device(vsgm_executor* e, const uint32_t* sources, uint64_t n_sources, int hop, uint64_t root_degree) {
  return Guard([&] {
    VSGM_CHECK(e, "null argument");
    VSGM_CHECK(e && (sources || n_sources == 0), "bad arguments");
    e->e->TransferDevice(std::vector<vsgm::uintV>(sources, sources + n_sources), hop, root_degree);
  });
}

int vsgm_executor_prefetch_device(vsgm_executor* e, const uint32_t* sources, uint64_t n_sources, int hop, uint64_t root_degree) {
  return Guard([&] {
    VSGM_CHECK(e && (sources || n_sources == 0), "null argument");
    e->e->PrefetchDeviceView(std::vector<vsgm::uintV>(sources, sources + n_sources), hop, root_degree);
  });
}

int vsgm_executor_add_device_view(vsgm_executor* e, const uint32_t* sources, uint64_t n_sources, int hop, uint64_t root_degree,
                                  int* index) {
  return Guard([&] {
    VSGM_CHECK(e && (sources || n_sources == 0), "null argument");
    std::vector<vsgm::uintV> s(sources, sources + n_sources);
    size_t i = e->e->AddDeviceView(s, hop, root_degree);
    if (index) *index = (int)i;
  });
}

int vsgm_executor_view_csr(const vsgm_executor* e, int view, uint64_t* row_ptrs, uint32_t* cols, uint64_t cols_cap, uint64_t* entries) {
  return Guard([&] {
    VSGM_CHECK(e, "null argument");
    std::vector<uint64_t> rp;
    std::vector<uint32_t> cl;
    e->e->ViewCsr((size_t)view, rp, cl);
    if (entries) *entries = cl.size();
    if (row_ptrs) std::copy(rp.begin(), rp.end(), row_ptrs);
    if (cols) {
      if (cl.size() > cols_cap) throw vsgm::Error("view has more entries than the buffer holds");
      std::copy(cl.begin(), cl.end(), cols);
    }
  });
}

int vsgm_executor_match_all(vsgm_executor* e, const int* plans, int n_plans) {
  return Guard([&] {
    VSGM_CHECK(e && n_plans >= 0, "bad arguments");
    e->e->MatchAll(plans ? std::vector<int>(plans, plans + n_plans) : std::vector<int>{});
  });
}

int vsgm_executor_capture(vsgm_executor* e, const int* plans, int n_plans, uint64_t* n_nodes) {
  return Guard([&] {
    VSGM_CHECK(e && n_plans >= 0, "bad arguments");
    uint64_t n = e->e->CaptureAll(plans ? std::vector<int>(plans, plans + n_plans) : std::vector<int>{});
    if (n_nodes) *n_nodes = n;
  });
}

int vsgm_executor_replay(vsgm_executor* e, double* ms) {
  return Guard([&] {
    VSGM_CHECK(e, "null argument");
    const double before = e->e->ReplayMs();
    e->e->ReplayGraph();
    if (ms) *ms = e->e->ReplayMs() - before;
  });
}

int vsgm_executor_hub_count(const vsgm_executor* e, uint32_t* n) {
  return Guard([&] {
    VSGM_CHECK(e && n, "null argument");
    *n = e->e->HubCount();
  });
}

int vsgm_executor_transfer(vsgm_executor* e, const vsgm_view_holder* h) {
  return Guard([&] {
    VSGM_CHECK(e && h, "null argument");
    e->e->Transfer(*h->h);
  });
}

int vsgm_executor_transfer_graph(vsgm_executor* e, const vsgm_graph* g, const uint32_t* sources, uint64_t n_sources) {
  return Guard([&] {
    VSGM_CHECK(e && g && (sources || n_sources == 0), "null argument");
    std::vector<vsgm::uintV> s;
    if (sources) s.assign(sources, sources + n_sources);
    e->e->TransferGraph(g->g, s);
  });
}

int vsgm_executor_match(vsgm_executor* e, int do_match) {
  return Guard([&] {
    VSGM_CHECK(e, "null argument");
    e->e->Match(do_match != 0);
  });
}

int vsgm_executor_counts(const vsgm_executor* e, uint64_t* counts, int n) {
  return Guard([&] {
    VSGM_CHECK(e && counts, "null argument");
    const auto& c = e->e->GetTotalCounts();
    VSGM_CHECK(n >= (int)c.size(), "counts buffer too small");
    for (size_t i = 0; i < c.size(); ++i) counts[i] = c[i];
  });
}

int vsgm_executor_match_ms(const vsgm_executor* e, double* ms, int n) {
  return Guard([&] {
    VSGM_CHECK(e && ms, "null argument");
    const auto& c = e->e->GetMatchMs();
    VSGM_CHECK(n >= (int)c.size(), "buffer too small");
    for (size_t i = 0; i < c.size(); ++i) ms[i] = c[i];
  });
}

int vsgm_executor_shared_stats(const vsgm_executor* e, double* ms, uint64_t* saved_runs) {
  return Guard([&] {
    VSGM_CHECK(e && ms && saved_runs, "null argument");
    *ms = e->e->SharedMatchMs();
    *saved_runs = e->e->SharedLevelRuns();
  });
}

int vsgm_executor_reset(vsgm_executor* e) {
  return Guard([&] {
    VSGM_CHECK(e, "null argument");
    e->e->ResetCounts();
  });
}

int vsgm_executor_level_stats(const vsgm_executor* e, int plan, vsgm_level_stat* out, int cap, int* n_out) {
  return Guard([&] {
    VSGM_CHECK(e && n_out && (cap <= 0 || out), "null argument");
    const auto& all = e->e->GetLevelStats();
    VSGM_CHECK(plan >= 0 && plan < (int)all.size(), "plan index out of range");
    const auto& s = all[plan];
    *n_out = (int)s.size();
    for (int i = 0; i < (int)s.size() && i < cap; ++i) out[i] = vsgm_level_stat{s[i].batches, s[i].paths, s[i].work, s[i].outputs, s[i].adj_entries, s[i].ms};
  });
}

int vsgm_executor_context(vsgm_executor* e, vsgm_ctx** ctx) {
  return Guard([&] {
    VSGM_CHECK(e && ctx, "null argument");
    *ctx = e->e->Context();
  });
}

int vsgm_executor_peak_bytes(const vsgm_executor* e, uint64_t* bytes) {
  return Guard([&] {
    VSGM_CHECK(e && bytes, "null argument");
    *bytes = e->e->PeakDeviceBytes();
  });
}

void vsgm_executor_free(vsgm_executor* e) { delete e; }

}  // extern "C"
