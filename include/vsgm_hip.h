/*
 * vsgm_hip.h — C-ABI of the gfx950 device library (libvsgm_hip.so).
 *
 * This is the drop-in boundary that replaces the reference's device layer:
 * the moderngpu-based CudaContext / DeviceArray runtime (src/device) and
 * the GpuUtils primitives and executor ops (src/gpu, src/executor/ops).
 * Every entry point takes plain pointers and sizes, runs on the context's
 * HIP stream and returns a vsgm_status (0 = success). No entry point calls
 * exit(); the message of the last failure is in vsgm_last_error(ctx).
 *
 * Device pointers handed in are owned by the caller. Buffers that an op
 * allocates for its outputs come from the context's pool and are released
 * with vsgm_free(). One context per host thread per device; a context is
 * not thread-safe.
 */
#ifndef VSGM_HIP_H_
#define VSGM_HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VSGM_MAX_QV 7 /* max pattern vertices, src/common/meta.h:33 */

typedef enum vsgm_status {
  VSGM_OK = 0,
  VSGM_ERR_HIP = 1,     /* a HIP runtime call failed */
  VSGM_ERR_OOM = 2,     /* device allocation failed */
  VSGM_ERR_BUDGET = 3,  /* request exceeds the context's memory budget */
  VSGM_ERR_ARG = 4,     /* invalid argument */
  VSGM_ERR_NO_DEVICE = 5
} vsgm_status;

/* Condition operators, same numbering as src/common/meta.h:35. */
enum { VSGM_LESS_THAN = 0, VSGM_LARGER_THAN = 1, VSGM_NON_EQUAL = 2 };

typedef struct vsgm_ctx vsgm_ctx;

/* ---- context, memory, copies --------------------------------------------
 * Replaces CudaContext + DeviceMemoryInfo (src/device/cuda_context.h:139-229,
 * src/device/device_memory_info.h:3-42) and the per-executor stream of
 * PipelineExecutor (src/executor/pipeline_executor.h:29-35). mem_budget = 0
 * selects 90% of the device memory that is free at creation. */
int vsgm_ctx_create(int device, uint64_t mem_budget, vsgm_ctx** out);
int vsgm_ctx_destroy(vsgm_ctx* ctx);
int vsgm_ctx_sync(vsgm_ctx* ctx);
const char* vsgm_last_error(const vsgm_ctx* ctx);
/* Number of visible devices (0 on a host without a GPU). */
int vsgm_device_count(int* count);
/* Budget accounting: bytes still available under the budget, the budget,
 * and the high-water mark of live bytes (DeviceMemoryInfo::GetAvailableMemorySize). */
int vsgm_mem_info(const vsgm_ctx* ctx, uint64_t* available, uint64_t* budget, uint64_t* peak);
/* No counterpart in the reference (one context per consumer thread there).
 * vsgm_ctx_set_budget changes the budget of a live context (it must cover
 * the bytes already live); vsgm_ctx_bind_thread makes the context's device
 * current on the calling thread, for a context driven from a thread other
 * than the one that created it (the executor's view-builder thread). */
int vsgm_ctx_set_budget(vsgm_ctx* ctx, uint64_t mem_budget);
int vsgm_ctx_bind_thread(const vsgm_ctx* ctx);
/* Pool allocation (CudaContext::Malloc/Free, cuda_context.h:166-210). */
int vsgm_alloc(vsgm_ctx* ctx, uint64_t bytes, void** out);
int vsgm_free(vsgm_ctx* ctx, void* ptr);
/* Asynchronous host->device copy on the context stream (HToD,
 * src/device/cuda_context.h:243-253); the host buffer must stay valid
 * until the next vsgm_ctx_sync. */
int vsgm_h2d(vsgm_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* Synchronous device->host copy (DToH). */
int vsgm_d2h(vsgm_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* Pinned host memory (cudaMallocHost in view_bin_holder.h:18-19). */
int vsgm_host_alloc(uint64_t bytes, void** out);
int vsgm_host_free(void* ptr);
/* Stream timing with HIP events in one of VSGM_TIMER_SLOTS independent
 * slots: vsgm_timer_stop returns the ms elapsed on the stream since
 * vsgm_timer_start of the same slot (it synchronises). */
#define VSGM_TIMER_SLOTS 4
int vsgm_timer_start(vsgm_ctx* ctx, int slot);
int vsgm_timer_stop(vsgm_ctx* ctx, int slot, float* ms);

/* ---- primitives ----------------------------------------------------------
 * Replace the moderngpu / CUB / Thrust wrappers of src/gpu. */
/* out[0..n] = exclusive prefix sum of in[0..n-1], out[n] = total
 * (GpuUtils::Scan::ExclusiveSum, src/gpu/gpu_utils_scan.h:13-35).
 * in == out is allowed. */
int vsgm_scan_excl_u64(vsgm_ctx* ctx, const uint64_t* d_in, uint64_t* d_out, uint64_t n);
/* *d_total += sum(in[0..n-1]) (GpuUtils::Reduce::Sum, gpu_utils_reduce.h:13-21). */
int vsgm_reduce_u64(vsgm_ctx* ctx, const uint64_t* d_in, uint64_t n, uint64_t* d_total);
/* dst[i] = src[idx[i]] (GpuUtils::Copy::Gather, gpu_utils_copy.h:15-21). */
int vsgm_gather_u32(vsgm_ctx* ctx, const uint64_t* d_idx, uint64_t n, const uint32_t* d_src, uint32_t* d_dst);
int vsgm_gather_u64(vsgm_ctx* ctx, const uint64_t* d_idx, uint64_t n, const uint64_t* d_src, uint64_t* d_dst);
/* d[i] = base + i (GpuUtils::Transform::Sequence, gpu_utils_transform.h:140-144). */
int vsgm_iota_u64(vsgm_ctx* ctx, uint64_t* d, uint64_t n, uint64_t base);
/* d[i] = value, a kernel fill (no hipMemset, so it is safe under graph capture). */
int vsgm_fill_u64(vsgm_ctx* ctx, uint64_t* d, uint64_t n, uint64_t value);

/* ---- level operators -----------------------------------------------------
 * One exec level of the BFS join over P partial embeddings ("paths") of a
 * batch. The per-pattern-vertex state mirrors ImData
 * (src/executor/im_data.h:7-90):
 *   inst[u]      u32[P]  bound data vertex of materialized pattern vertex u;
 *   cand[u]      u32[]   candidate values of computed, unmaterialized u,
 *   cand_off[u]  u64[]   as a CSR whose rows are sorted ascending;
 *   cand_idx[u]  u64[P]  row of that CSR for each path; NULL means the
 *                        identity map row = cand_idx_base[u] + path.
 * The view is a CSR over the full vertex range (row_ptrs u64[V+1], cols
 * u32[...]) with sorted adjacency lists.
 * The op fields follow the DevPlan tables (src/query/dev_plan.h:128-313):
 *   target   pattern vertex the op works on;
 *   back[]   backward neighbours of target (intersections);
 *   cond[]   conditions "candidate OP inst[operand]" (computed / materialized
 *            / filter ordering of the target, depending on the op);
 *   other    second computed, unmaterialized vertex counted against the
 *            target's candidates (-1: none), with "target other_op other".
 *   hub_*    optional dense membership index of the same view (see
 *            vsgm_hub_index_build; hub_slot NULL: none). Membership tests
 *            against an indexed list read one bit instead of a binary search;
 *            results are identical either way. */
typedef struct vsgm_cond {
  uint32_t op;
  uint32_t operand;
} vsgm_cond;

typedef struct vsgm_level_args {
  const uint64_t* row_ptrs;
  const uint32_t* cols;
  uint64_t vertex_count;
  uint64_t path_num;
  const uint32_t* inst[VSGM_MAX_QV];
  const uint32_t* cand[VSGM_MAX_QV];
  const uint64_t* cand_off[VSGM_MAX_QV];
  const uint64_t* cand_idx[VSGM_MAX_QV];
  uint64_t cand_idx_base[VSGM_MAX_QV];
  uint32_t target;
  uint32_t n_back;
  uint32_t back[VSGM_MAX_QV];
  uint32_t n_cond;
  vsgm_cond cond[VSGM_MAX_QV];
  int32_t other;
  uint32_t other_op;
  const uint32_t* hub_slot;  /* u32[V]: bitmap row of an indexed vertex, ~0u otherwise */
  const uint32_t* hub_bits;  /* u32[n_hubs * hub_words]: bit c of row s = c in adj(hub s) */
  uint64_t hub_words;        /* u32 words per bitmap row */
  uint32_t flags;            /* VSGM_LEVEL_* properties of the view (0: none known) */
} vsgm_level_args;
/* The view has no self-loop, so a candidate drawn from or tested against
 * the list of a bound vertex v can never equal v: '!=' conditions against
 * backward vertices are then implied and the intersection maps skip them. */
#define VSGM_LEVEL_LOOP_FREE 1u

/* Dense membership index of the highest-degree lists of a view. No
 * counterpart in the reference, whose connectivity check always binary-
 * searches the list (ThreadCheckConnectivity / BinSearch,
 * src/gpu/gpu_check_constraints_common.h:11-63, gpu_utils_bin_search.h:17-29);
 * it trades HBM (V/8 bytes per indexed vertex) for one probe per test. */
typedef struct vsgm_hub_index {
  uint32_t* slot;      /* u32[vertex_count] */
  uint32_t* bits;      /* u32[n_hubs * words] */
  uint64_t words;
  uint32_t n_hubs;
} vsgm_hub_index;
/* Build the index of the n_hubs vertices h_hubs[] (host array, distinct ids
 * < vertex_count) over the device view (d_row_ptrs, d_cols); memory comes
 * from the context pool. */
int vsgm_hub_index_build(vsgm_ctx* ctx, const uint64_t* d_row_ptrs, const uint32_t* d_cols, uint64_t vertex_count, const uint32_t* h_hubs,
                         uint32_t n_hubs, vsgm_hub_index* out);
int vsgm_hub_index_free(vsgm_ctx* ctx, vsgm_hub_index* index);
/* Number of self-loops (v in adj(v)) of a device view, into *count (host).
 * No counterpart in the reference; it decides VSGM_LEVEL_LOOP_FREE. */
int vsgm_view_self_loops(vsgm_ctx* ctx, const uint64_t* d_row_ptrs, const uint32_t* d_cols, uint64_t vertex_count, uint64_t* count);

/* View of one bin built on the device from the whole graph resident in
 * HBM (replaces the host ViewBin::Materialize + ExtendFromSource,
 * src/view/view_bin.h:49-142, and the view upload of
 * PipelineExecutor::Transfer, pipeline_executor.h:64-81): the full adjacency
 * list of every vertex within hop-1 BFS levels of a source whose degree
 * reaches root_degree (0: every source), empty lists elsewhere, over the full
 * vertex range. row_ptrs u64[V+1] and cols u32[entries] are pool blocks
 * owned by the caller; vertices = vertices inside the view. */
typedef struct vsgm_view_out {
  uint64_t* row_ptrs;
  uint32_t* cols;
  uint64_t entries;
  uint64_t vertices;
} vsgm_view_out;
int vsgm_view_build(vsgm_ctx* ctx, const uint64_t* d_row_ptrs, const uint32_t* d_cols, uint64_t vertex_count, const uint32_t* d_sources,
                    uint64_t n_sources, int hop, uint64_t root_degree, vsgm_view_out* out);

/* ---- graph ingestion ------------------------------------------------------
 * Replaces GraphIO::ReadSNAPFile's CSR build + SortCSRArray
 * (src/graph/graph_io.h:193-274) and PreprocessGraph::Preprocess
 * (src/preprocess/preprocess_graph.h:12-92) on the device. d_pairs holds
 * n_pairs (u, v) pairs of u32 ids < vertex_count; symmetric != 0 adds
 * (v, u) for each. Self loops and duplicate entries are dropped and every
 * list is sorted. drop_isolated != 0 (symmetric only) removes vertices with
 * empty lists and renumbers the rest in order. out->row_ptrs
 * (u64[out->vertices + 1]) and out->cols (u32[out->entries]) are
 * pool-allocated; free them with vsgm_free. Needs two u64 keys per
 * emitted entry plus the sort's workspace while it runs. */
int vsgm_csr_from_pairs(vsgm_ctx* ctx, const uint32_t* d_pairs, uint64_t n_pairs, uint64_t vertex_count, int symmetric, int drop_isolated,
                        vsgm_view_out* out);

/* A candidate CSR produced by an op: offsets u64[path_num+1] and values
 * u32[count], both pool-allocated (free with vsgm_free). */
typedef struct vsgm_csr_out {
  uint64_t* offsets;
  uint32_t* values;
  uint64_t count;
} vsgm_csr_out;

/* Columns of the next level produced by vsgm_expand. */
typedef struct vsgm_expand_out {
  uint64_t path_num;
  uint32_t* inst[VSGM_MAX_QV];     /* new columns for the materialized set + target */
  uint64_t* cand_idx[VSGM_MAX_QV]; /* new row maps for the carried candidate sets */
} vsgm_expand_out;

/* Sizes seen by the last level op (work = candidate-vs-list elements
 * visited after range narrowing, outputs = surviving elements,
 * adj_entries = sum over paths of the degrees of all backward vertices of
 * an intersection, i.e. the list volume of the merge model). */
typedef struct vsgm_op_stats {
  uint64_t paths;
  uint64_t work;
  uint64_t outputs;
  uint64_t adj_entries;
} vsgm_op_stats;

/* Per-kernel profile: every level-op kernel launched while profiling is on
 * is bracketed with HIP events on the context stream. Kernels are named
 * "<stage>_<op>" with stage setup / map / scatter and op icount (COMPUTE_COUNT),
 * iwrite (COMPUTE), expand (MATERIALIZE), filter (FILTER_COMPUTE), count2
 * (two-set COUNT), ipath (per-path counts). algo_bytes sums the launches'
 * algorithmic HBM bytes (for intersections the streaming model: every
 * candidate element read once from its pivot list, plus each path's bound
 * vertices, row_ptrs and segment words). */
typedef struct vsgm_kernel_stat {
  char name[32];
  uint64_t launches;
  double total_ms;
  double max_ms;
  uint64_t algo_bytes;
  uint64_t work;
  uint64_t paths;
} vsgm_kernel_stat;
/* Launch-shape knobs of the level kernels, chosen by the autotuner:
 *   isect_kernel  0: every lane verifies its own candidate with its own
 *                    binary searches (block-strided elements);
 *                 1: a wavefront walks a contiguous element run; runs inside
 *                    one path stream the other lists through LDS tiles from
 *                    a per-wave cursor, other runs use branchless searches;
 *                 2: as 0, and each workgroup keeps in LDS one backward
 *                    adjacency list (<= 4096 entries) bound to the same vertex
 *                    across its chunk and searches it there;
 *                 3: COMPUTE_COUNT / per-path counts with two or more
 *                    backward lists: one 1024-thread workgroup per CU keeps
 *                    the longest list (<= 14336 entries) whose vertex stays
 *                    bound across a run of consecutive paths in an LDS hash
 *                    table, rebuilt when the run changes, and takes tickets
 *                    of ticket_chunks consecutive 8192-element chunks from a
 *                    device counter; the pivot choice then prefers streaming
 *                    the list that changes per path (other ops run as 0);
 *                 4: as 3, and COMPUTE_COUNT levels with two or three
 *                    backward lists, no second computed set and at most
 *                    three '!=' conditions run the flat map: each wavefront
 *                    walks its chunk in 1024-element blocks that may span
 *                    paths, with every path's probe plan resolved when a
 *                    window of 64 paths is loaded (lists up to 12800
 *                    entries held, power-of-two cuckoo tables);
 *   max_grid      cap on workgroups per map launch (grid-stride beyond);
 *   cold_weight   pivot cost of one test against a list that changes per
 *                 path, in units of one streamed pivot element (isect_kernel
 *                 3 and 4; 1..1024);
 *   ticket_chunks chunks per ticket of isect_kernel 3 and 4 (1..64).
 * Defaults: isect_kernel 4, max_grid 8192, cold_weight 16, ticket_chunks 64. */
typedef struct vsgm_tuning {
  uint32_t isect_kernel;
  uint32_t max_grid;
  uint32_t cold_weight;
  uint32_t ticket_chunks;
} vsgm_tuning;
int vsgm_set_tuning(vsgm_ctx* ctx, const vsgm_tuning* t);
int vsgm_get_tuning(const vsgm_ctx* ctx, vsgm_tuning* t);

/* ---- recorded replay and graph capture -----------------------------------
 * No counterpart in the reference, whose host side synchronises on every
 * level op (cudaMemcpy of totals in the src/executor/ops headers). The host decisions
 * of a Match depend only on a few values read back from the device (element
 * totals, output counts, batch ends). vsgm_tape_record .. vsgm_tape_stop
 * records them during an ordinary run; between vsgm_capture_begin and
 * vsgm_capture_end the same calls replay them without any device
 * synchronisation while the context stream is captured into one graph,
 * which vsgm_graph_launch then replays (same inputs -> same work, fresh
 * results). The captured work must issue the recorded sequence exactly;
 * capture_end fails otherwise. Buffers the graph reads but does not write
 * first (views, sources, index) must outlive it; vsgm_graph_release drops it.
 * vsgm_d2h_async enqueues a copy without waiting (for pinned destinations
 * inside a capture). */
int vsgm_d2h_async(vsgm_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int vsgm_tape_record(vsgm_ctx* ctx);
int vsgm_tape_stop(vsgm_ctx* ctx, uint64_t* n_values);
int vsgm_capture_begin(vsgm_ctx* ctx);
int vsgm_capture_end(vsgm_ctx* ctx, uint64_t* n_nodes);
int vsgm_graph_launch(vsgm_ctx* ctx);
int vsgm_graph_release(vsgm_ctx* ctx);

/* Turn profiling on/off; either way the records collected so far are dropped. */
int vsgm_profile_enable(vsgm_ctx* ctx, int enable);
/* Aggregate the records by kernel name (synchronises). */
int vsgm_profile_collect(vsgm_ctx* ctx, vsgm_kernel_stat* out, int cap, int* n_out);

/* COMPUTE (ComputeGeneral, src/executor/ops/compute.h:7-41 via
 * GpsmBinSearch, src/gpu/gpu_utils_gpsm_bin_search.h:134-171): for each
 * path, the sorted candidates of target = intersection of the adjacency
 * lists of inst[back[*]] filtered by cond[]. */
int vsgm_intersect_write(vsgm_ctx* ctx, const vsgm_level_args* a, vsgm_csr_out* out);
/* COMPUTE_COUNT (ComputeCount, src/executor/ops/compute_count.h:9-84):
 * *d_total += sum over paths and valid candidates of 1, or, with
 * other >= 0, of the number of other-candidates c2 with "cand other_op c2"
 * (ThreadCountWithComputedVertex, gpu_check_constraints_common.h:120-139). */
int vsgm_intersect_count_total(vsgm_ctx* ctx, const vsgm_level_args* a, uint64_t* d_total);
/* Allocation-free COMPUTE_COUNT. No counterpart in the reference; it is for
 * callers that cannot allocate device memory (code under HIP graph capture
 * by the caller, tight serving loops, harnesses that forbid hipMalloc).
 * Same result as vsgm_intersect_count_total (same kernels and tuning), but
 * every temporary (segment words, scans, chunk plans) comes from a 64 MiB
 * scratch array the library holds on each device for the life of the
 * process: no hipMalloc / hipFree, no pool block, no event. Limit:
 * path_num <= VSGM_NOALLOC_MAX_PATHS (VSGM_ERR_ARG beyond). Calls on one
 * device share that scratch, so calls that overlap on one device (several
 * contexts or threads at once) race; calls in sequence are safe (each
 * rewrites every scratch word it reads). Not available while the context
 * records a capture (vsgm_capture_begin). */
#define VSGM_NOALLOC_MAX_PATHS (1ull << 21)
int vsgm_intersect_count_total_noalloc(vsgm_ctx* ctx, const vsgm_level_args* a, uint64_t* d_total);
/* Per-path intersection counts (the count pass of GpsmBinSearch and
 * ComputePathCount, src/executor/ops/compute_path_count.h:7-49):
 * d_counts[p] = number of valid candidates of path p. */
int vsgm_intersect_count_paths(vsgm_ctx* ctx, const vsgm_level_args* a, uint64_t* d_counts);
/* MATERIALIZE (Materialize, src/executor/ops/materialize.h:20-126): every
 * candidate of target passing cond[] becomes a new path. New columns:
 * inst[target] = candidate, inst[u] = parent's inst[u] for the n_mat
 * vertices in mat[], cand_idx[v] = parent's row for the n_cu vertices in
 * cu[]. Paths stay in (parent, candidate) order. */
int vsgm_expand(vsgm_ctx* ctx, const vsgm_level_args* a, const uint32_t* mat, uint32_t n_mat, const uint32_t* cu, uint32_t n_cu,
                vsgm_expand_out* out);
/* FILTER_COMPUTE (FilterCompute, src/executor/ops/filter_compute.h:97-127):
 * a new candidate CSR for target, one row per path, keeping candidates
 * that pass cond[]. */
int vsgm_filter(vsgm_ctx* ctx, const vsgm_level_args* a, vsgm_csr_out* out);
/* COUNT (LIGHTCount, src/executor/ops/count.h:109-167): with other < 0,
 * *d_total += sum of the target's candidate-row sizes; otherwise
 * *d_total += number of pairs (x in cand(target), y in cand(other)) with
 * "x other_op y" (GeneralCountTwoComputedVertices, count.h:16-107). */
int vsgm_count_sets(vsgm_ctx* ctx, const vsgm_level_args* a, uint64_t* d_total);

/* Batch organisation (Processor::OrganizeBatch, src/executor/processor.h:198-263,
 * organize_batch.h:11-61). mode 0: pivot degree of the intersection
 * (BuildIntersectChildrenCount); mode 1: candidate-row size of target
 * (BuildGatherChildrenCount); mode 2: 1 per path. */
int vsgm_children_count(vsgm_ctx* ctx, const vsgm_level_args* a, int mode, uint64_t* d_out);
/* Split n parents into consecutive batches whose cost
 * sum(children*children_factor + parent_factor) stays within batch_bytes
 * (one parent may exceed it alone). Writes the exclusive end of each
 * batch to h_ends (host, capacity max_batches) and the count to
 * *n_batches (BatchManager / SplitBatchKernel, batch_manager.h:34-121,
 * split_batch.h:37-81). */
int vsgm_split_batch(vsgm_ctx* ctx, const uint64_t* d_children, uint64_t n, uint64_t parent_factor, uint64_t children_factor,
                     uint64_t batch_bytes, uint64_t* h_ends, uint64_t max_batches, uint64_t* n_batches);

int vsgm_last_op_stats(const vsgm_ctx* ctx, vsgm_op_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* VSGM_HIP_H_ */
