/*
 * vsgm.h — C-ABI of the host library (libvsgm.so): graph IO and generators,
 * queries and the planner, view bins and the pipeline executor.
 *
 * It exposes, as plain functions over opaque handles, the C++ host API that
 * the reference's pm_pipeline drives in-process
 * (src/app/pattern_matching_pipeline.cu:58-198): Graph, Pattern / Clique /
 * Graphlet queries, Plan::Optimize, ViewBinManager / ViewBin, and
 * PipelineExecutor (ctor, Transfer, Match, GetTotalCounts). Functions return
 * 0 on success and a nonzero vsgm_status otherwise; the message is in
 * vsgm_error() (per calling thread). Handles are freed with their *_free.
 */
#ifndef VSGM_H_
#define VSGM_H_

#include <stddef.h>
#include <stdint.h>

#include "vsgm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vsgm_graph vsgm_graph;
typedef struct vsgm_query vsgm_query;
typedef struct vsgm_plan vsgm_plan;
typedef struct vsgm_view_manager vsgm_view_manager;
typedef struct vsgm_view_holder vsgm_view_holder;
typedef struct vsgm_executor vsgm_executor;

const char* vsgm_error(void);

/* ---- graph (src/graph/graph_io.h, src/graph/cpu_graph.h) ---------------- */
/* GraphIO::ReadCSRBinFile (graph_io.h:57-100). */
int vsgm_graph_read_bin(const char* path, vsgm_graph** out);
/* GraphIO::WriteCSRBinFile (graph_io.h:313-339); writes exactly `path`. */
int vsgm_graph_write_bin(const vsgm_graph* g, const char* path);
/* Graph from a CSR (copied); lists must be sorted and symmetric. */
int vsgm_graph_from_csr(uint64_t vertex_count, const uint64_t* row_ptrs, const uint32_t* cols, vsgm_graph** out);
/* Graph from undirected edge pairs (cleaned: symmetrised, deduplicated, no self loops). */
int vsgm_graph_from_edges(uint64_t vertex_count, const uint32_t* pairs, uint64_t n_pairs, vsgm_graph** out);
/* SNAP edge-list text -> preprocessed undirected graph (GraphIO::ReadSNAPFile,
 * src/graph/graph_io.h:193-274, then PreprocessGraph::Preprocess,
 * src/preprocess/preprocess_graph.h:12-92): lines that do not start with a
 * digit are skipped, ids are shifted so the smallest is 0, self loops,
 * parallel edges and isolated vertices are dropped and the rest renumbered
 * in order. device >= 0 builds the CSR on that GPU (vsgm_csr_from_pairs),
 * device < 0 on the host; threads <= 0 uses all cores. directed != 0 is
 * rejected (the matcher reads undirected graphs). */
int vsgm_graph_read_snap(const char* path, int directed, int device, int threads, vsgm_graph** out);
/* Any graph file by suffix (GraphIO::ReadDataFile, src/graph/graph_io.h:42-55):
 * .bin CSR, .txt SNAP (ids shifted by the smallest), .lvid large-id SNAP
 * (every digit run an id, numbered by first appearance, graph_io.h:106-191).
 * Text input is symmetrised and cleaned of self loops and parallel edges
 * (the reference keeps them when it matches a .txt file directly);
 * drop_isolated != 0 also removes isolated vertices. device >= 0 builds
 * text input on that GPU. */
int vsgm_graph_load(const char* path, int device, int threads, int drop_isolated, vsgm_graph** out);
/* Copy of g without its isolated vertices, ids renumbered in order
 * (preprocess_graph.h:67-82). */
int vsgm_graph_drop_isolated(const vsgm_graph* g, vsgm_graph** out);
/* Isomorphic copy of g with ids in ascending (degree, id) order; new_id
 * (optional, vertex_count entries) receives each vertex's id in the copy.
 * Pattern counts are unchanged; roots that must be the smallest id of an
 * embedding then start from low-degree vertices. No counterpart in the
 * reference (it matches ids as given): for callers whose plans root
 * symmetric patterns at the smallest id. */
int vsgm_graph_relabel_degree(const vsgm_graph* g, int threads, uint32_t* new_id, vsgm_graph** out);
int vsgm_graph_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed, int threads, vsgm_graph** out);
int vsgm_graph_chung_lu(uint64_t vertex_count, uint64_t draws, double gamma, uint64_t seed, int threads, vsgm_graph** out);
int vsgm_graph_info(const vsgm_graph* g, uint64_t* vertex_count, uint64_t* entries, uint64_t* max_degree);
/* Borrowed pointers into the graph's CSR (valid until vsgm_graph_free). */
int vsgm_graph_csr(const vsgm_graph* g, const uint64_t** row_ptrs, const uint32_t** cols);
/* Writes "" to msg when the CSR is sorted, loop-free, duplicate-free and symmetric. */
int vsgm_graph_validate(const vsgm_graph* g, char* msg, size_t len);
void vsgm_graph_free(vsgm_graph* g);

/* ---- queries (src/query/pattern.h, clique.h, graphlet.h, query.h) ------- */
int vsgm_query_preset(int preset, vsgm_query** out);  /* Pattern(Pk, false) */
int vsgm_query_clique(int k, vsgm_query** out);       /* Clique(k) */
/* Connected pattern from an n x n 0/1 adjacency matrix (Query(n, m, adj), query.h:62-78). */
int vsgm_query_from_matrix(int n, const uint8_t* adj, vsgm_query** out);
/* Graphlet(k).GetQueries() (graphlet.h:15-136): *n_out queries written to
 * out[0..cap); *n_out is set even when cap is too small. */
int vsgm_query_graphlets(int k, vsgm_query** out, int cap, int* n_out);
/* n, undirected edges, remapped adjacency matrix and condition matrix
 * (cond[a*n+b] = operator between a and b, VSGM_NON_EQUAL when unrelated). */
int vsgm_query_info(const vsgm_query* q, int* n, int* edges, uint8_t* adj, uint8_t* cond);
void vsgm_query_free(vsgm_query* q);

/* ---- plans (src/query/plan.h) -------------------------------------------- */
int vsgm_plan_optimize(const vsgm_query* q, vsgm_plan** out); /* Plan(q).Optimize() */
/* Plan over a caller-chosen connected matching order seq[0..n). */
int vsgm_plan_from_order(const vsgm_query* q, const uint32_t* seq, int n, vsgm_plan** out);
int vsgm_plan_info(const vsgm_plan* p, int* hop, uint64_t* root_degree, int* n_ops);
/* Operation i: op (TraversalOperation numbering of plan.h:14-21) and vertex. */
int vsgm_plan_op(const vsgm_plan* p, int i, int* op, int* vertex);
/* Matching order (search sequence) into seq[0..n). */
int vsgm_plan_order(const vsgm_plan* p, uint32_t* seq, int cap);
/* Key/value text of the whole plan (the planner parity fixtures' format);
 * returns the needed size in *needed. */
int vsgm_plan_serialize(const vsgm_plan* p, char* buf, size_t len, size_t* needed);
/* Number of leading exec levels two plans share (same operator, vertex,
 * vertex groups, backward set and conditions; identical partial embeddings
 * after them), the unit of prefix-shared matching. No counterpart in the
 * reference. */
int vsgm_plans_shared_prefix(const vsgm_plan* a, const vsgm_plan* b, int* levels);
void vsgm_plan_free(vsgm_plan* p);

/* ---- view bins (src/view) ------------------------------------------------ */
int vsgm_vbmap_write(const char* path, int n_bins, const int* assignment, uint64_t vertex_count);
int vsgm_view_manager_load(const vsgm_graph* g, int hop, int threads, const char* vbmap, vsgm_view_manager** out);
int vsgm_view_manager_from_assignment(const vsgm_graph* g, int hop, int threads, int n_bins, const int* assignment,
                                      vsgm_view_manager** out);
int vsgm_view_manager_info(const vsgm_view_manager* m, int* n_bins, uint64_t* max_sources);
int vsgm_view_manager_reorder(vsgm_view_manager* m);
int vsgm_view_manager_split(vsgm_view_manager* m, int devices, int times);
void vsgm_view_manager_free(vsgm_view_manager* m);
/* Pinned holder with room for max_bytes of view columns (ViewBinHolder). */
int vsgm_view_holder_create(uint64_t vertex_count, uint64_t max_bytes, vsgm_view_holder** out);
/* ViewBin::Materialize of the manager's bin at pool position `index`. */
int vsgm_view_materialize(const vsgm_view_manager* m, int index, vsgm_view_holder* h, uint64_t root_degree);
int vsgm_view_holder_info(const vsgm_view_holder* h, int* bin_id, uint64_t* sources, uint64_t* entries, uint64_t* vertices);
int vsgm_view_holder_csr(const vsgm_view_holder* h, const uint64_t** row_ptrs, const uint32_t** cols, const uint32_t** sources);
void vsgm_view_holder_free(vsgm_view_holder* h);

/* ---- executor (src/executor/pipeline_executor.h:20-120) ------------------ */
typedef struct vsgm_level_stat {
  uint64_t batches;
  uint64_t paths;
  uint64_t work;
  uint64_t outputs;
  uint64_t adj_entries;
  double ms;
} vsgm_level_stat;

int vsgm_executor_create(int device, uint64_t vertex_count, const vsgm_plan* const* plans, int n_plans, uint64_t max_sources,
                         uint64_t mem_budget, vsgm_executor** out);
/* Options: "batch_bytes" (0 = budget / remaining levels), "min_batch_bytes",
 * "profile" (1 = time every level op), "hub_bytes" (device bytes for the
 * dense membership index of the highest-degree view lists, built at the next
 * transfer; 0 = off, -1 = min(budget/8, 16 GiB), the default),
 * "hub_min_degree" (smallest indexed degree, default 64), the device
 * tuning knobs "isect_kernel", "max_grid", "cold_weight" and "ticket_chunks"
 * (vsgm_tuning in vsgm_hip.h), applied to every plan, "prefetch_bytes"
 * (memory of the background view builder, 0 = a quarter of the budget) and
 * "share_prefixes" (1 = match the plans of a call as a prefix tree: a level
 * op that several plans reach through identical levels runs once for all,
 * sharing the common plan prefixes of queries such as the -km graphlets of
 * src/query/graphlet.h:15-136; per-plan match times are then not measured,
 * vsgm_executor_shared_stats reports the whole call's). */
int vsgm_executor_set_option(vsgm_executor* e, const char* key, double value);
/* Prefix-shared matching: summed time of the shared calls and the number of
 * level-op runs saved (no counterpart in the reference, which runs every
 * plan alone: pipeline_executor.h:83-109). */
int vsgm_executor_shared_stats(const vsgm_executor* e, double* ms, uint64_t* saved_runs);
/* Number of view vertices in the current membership index. */
int vsgm_executor_hub_count(const vsgm_executor* e, uint32_t* n);
/* Launch shape and batch budget of one plan (0 = the executor option),
 * as chosen by the autotuner (vsgm_amd/autotune.py). */
int vsgm_executor_set_plan_tuning(vsgm_executor* e, int plan, uint32_t isect_kernel, uint32_t max_grid, uint64_t batch_bytes,
                                  uint32_t cold_weight, uint32_t ticket_chunks);
int vsgm_executor_get_plan_tuning(const vsgm_executor* e, int plan, uint32_t* isect_kernel, uint32_t* max_grid, uint64_t* batch_bytes,
                                  uint32_t* cold_weight, uint32_t* ticket_chunks);

/* ---- batched mode (no counterpart in the reference, which uploads one
 * view bin at a time and runs plans one after another with a host
 * synchronisation per level op, pipeline_executor.h:64-109) -------------
 * add_view keeps one more view bin resident next to the others (their
 * index rows share the automatic index budget); match_all runs the listed
 * plans (NULL = all) over every resident view; capture runs the same work
 * once, recording the host-visible values, then captures it into one device
 * graph; replay launches that graph (one launch for every plan and view) and
 * adds its counts to the totals. Counts are identical to match_all. */
int vsgm_executor_add_view(vsgm_executor* e, const vsgm_view_holder* h, int* index);
/* Device-built views (vsgm_view_build in vsgm_hip.h; replaces the host
 * ViewBin::Materialize + upload, view_bin.h:49-142, pipeline_executor.h:64-81):
 * set_graph keeps the whole graph resident, add_device_view builds and keeps
 * the view of one bin (its sources, hop, root degree) without host work;
 * view_csr copies a resident view back (row_ptrs u64[V+1]; cols up to
 * cols_cap entries, *entries = its size). */
int vsgm_executor_set_graph(vsgm_executor* e, const vsgm_graph* g);
int vsgm_executor_add_device_view(vsgm_executor* e, const uint32_t* sources, uint64_t n_sources, int hop, uint64_t root_degree,
                                  int* index);
/* The device-built counterpart of vsgm_executor_transfer: builds the bin's
 * view from the set_graph graph and replaces every resident view with it
 * (what pm_pipeline -dv 1 does per bin). */
int vsgm_executor_transfer_device(vsgm_executor* e, const uint32_t* sources, uint64_t n_sources, int hop, uint64_t root_degree);
/* Start building a bin's view in the background (a second device context
 * and a builder thread) so that it overlaps whatever the executor runs next,
 * typically the current bin's match; the next transfer_device with the same
 * arguments takes it. The overlap the reference gets from its producer
 * threads (view materialisation beside Transfer + Match,
 * src/app/pattern_matching_pipeline.cu:103-172), with the view built on the
 * device. */
int vsgm_executor_prefetch_device(vsgm_executor* e, const uint32_t* sources, uint64_t n_sources, int hop, uint64_t root_degree);
int vsgm_executor_view_csr(const vsgm_executor* e, int view, uint64_t* row_ptrs, uint32_t* cols, uint64_t cols_cap, uint64_t* entries);
int vsgm_executor_match_all(vsgm_executor* e, const int* plans, int n_plans);
int vsgm_executor_capture(vsgm_executor* e, const int* plans, int n_plans, uint64_t* n_nodes);
int vsgm_executor_replay(vsgm_executor* e, double* ms);
int vsgm_executor_transfer(vsgm_executor* e, const vsgm_view_holder* h);
/* Whole graph as one view; sources NULL = every vertex. */
int vsgm_executor_transfer_graph(vsgm_executor* e, const vsgm_graph* g, const uint32_t* sources, uint64_t n_sources);
int vsgm_executor_match(vsgm_executor* e, int do_match);
int vsgm_executor_counts(const vsgm_executor* e, uint64_t* counts, int n);
int vsgm_executor_match_ms(const vsgm_executor* e, double* ms, int n);
int vsgm_executor_reset(vsgm_executor* e);
int vsgm_executor_level_stats(const vsgm_executor* e, int plan, vsgm_level_stat* out, int cap, int* n_out);
int vsgm_executor_peak_bytes(const vsgm_executor* e, uint64_t* bytes);
/* The executor's device context, for the profiling calls of vsgm_hip.h. */
int vsgm_executor_context(vsgm_executor* e, vsgm_ctx** ctx);
void vsgm_executor_free(vsgm_executor* e);

/* ---- the pm_pipeline driver (src/app/pattern_matching_pipeline.cu:27-201) */
int vsgm_pm_pipeline(int argc, char** argv);

/* ---- view-bin packing (src/tools/view_packing.cc:141-324, src/tools/view_bin.h:65-171):
 * packs the sources of each first-layer cluster of `partition` (ids in
 * [0, first_layer), or [0, first_layer * second_layer) for a two-layer
 * partition) into bins of at most max_bin_entries list entries (their
 * hop-1 balls, as the reference's packing counts them), at most pool_size
 * bins open per cluster, sources in vertex order or by decreasing degree
 * (degree_order), then merges the bins next-fit in increasing size. Writes
 * the merged bin of every vertex to out_assign (int32[V]) and the bin count
 * to *out_bins; out_sources / out_sizes (u64[cap], may be NULL) receive
 * each bin's source count and size; a cap below the bin count is an error
 * after *out_bins is set. Fails if one source's view alone exceeds the limit. */
int vsgm_view_pack(const vsgm_graph* g, int hop, const int32_t* partition, int first_layer, int second_layer, uint64_t max_bin_entries,
                   int pool_size, int degree_order, int threads, int32_t* out_assign, int* out_bins, uint64_t* out_sources,
                   uint64_t* out_sizes, int cap);
/* The reference's view_packing tool: -gf graph.bin -pf partition.<K>[x<M>]
 * -h hop -m GB -t K -s pool -o degree_order -d dump; writes
 * <graph>.<hop>.hop.vbstat and .vbmap. */
int vsgm_view_packing(int argc, char** argv);

/* ---- the kmeans tool (src/tools/kmeans.cc:636-694) on the GPU: -f graph.bin
 * -k K -i iterations -o output prefix (default <graph>.kmeans.) -d dump
 * -rc reorder -init 0, plus -g GPU and -s seed (the reference seeds with
 * the clock); writes one cluster id per line to <prefix><K>, empty clusters
 * removed. Multi-level "axb" and "a^x" levels are clustered level by level;
 * "2^x" is the reference's recursive bisection (Run2x, kmeans.cc:284-383). */
int vsgm_kmeans_tool(int argc, char** argv);
/* The features tool (src/tools/features.cc:68-258): -f graph -h 1 writes
 * <graph>.1hop.bin (every row plus its own vertex), -h 2 <graph>.2hop.bin
 * (every vertex within two steps), rows sorted; -dn n writes
 * <graph>.x<n>.bin, n copies with shuffled ids. The sampled -h -2 is
 * rejected. The kmeans tool reads such a file as the reference's
 * does (rows that hold their own vertex are clustered as they are). */
int vsgm_features_tool(int argc, char** argv);
/* The query_test tool (src/tools/query_test.cc:12-65): -p preset, -kc k or
 * -km k; prints each query and its optimised plan (and, for graphlets, the
 * hop of every plan and the largest). */
int vsgm_query_test(int argc, char** argv);
/* The fennel tool (src/tools/fennel.cc:21-101): -f graph -p parts -h 1|2;
 * writes <graph>.fennel.part.<p>, one part per vertex, the same parts as
 * the reference's streaming partitioner. vsgm_graph_fennel is the same
 * partition as a library call (int32[V]). */
int vsgm_fennel_tool(int argc, char** argv);
int vsgm_graph_fennel(const vsgm_graph* g, int parts, int hop, int32_t* assignment);
/* Clustering of g's vertices on GPU `device` (the tool without its file
 * IO): levels "K", "AxBx..." or "A^X" (A != 2), KMeans::Run(std::string),
 * kmeans.cc:253-281 and Run(std::vector<int>), :386-489 — every level
 * splits each non-empty cluster c of the previous one into K with
 * vsgm_kmeans_members, ids a + c * K — or "2^X", the recursive bisection
 * of Run2x (:284-383). Writes int32[V] ids and the id range (*clusters;
 * empty clusters are not removed). */
int vsgm_graph_kmeans(const vsgm_graph* g, const char* levels, int iters, uint64_t seed, int device, int reorder, int32_t* assignment,
                      int* clusters);
/* The tool's two host steps on an assignment (int32[V], ids < k):
 * renumber clusters by decreasing count of distinct vertices in their
 * members' closed neighbourhoods (ReorderCentroids, kmeans.cc:435-452), and
 * drop empty clusters keeping the order (RemoveEmptyClusters,
 * kmeans.cc:491-510; *kept = clusters left). */
int vsgm_clusters_reorder(const vsgm_graph* g, int k, int32_t* assignment);
int vsgm_clusters_compact(int k, int32_t* assignment, uint64_t vertex_count, int* kept);

/* ---- the preprocess_graph tool (src/preprocess/preprocess_graph.cc:9-36):
 * -f graph file, -d directed (0 only), -o 0 (.snap.txt), 1 (.bin, the
 * default) or 3 (.adj.txt + .info), the reference's GraphFormat numbers,
 * plus -g GPU (>= 0 builds the CSR there) and -t threads; writes
 * <file> plus that suffix. */
int vsgm_preprocess_graph(int argc, char** argv);

#ifdef __cplusplus
}
#endif

#endif /* VSGM_H_ */
