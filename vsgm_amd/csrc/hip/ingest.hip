// Graph ingestion on the device: an edge list becomes the sorted, symmetric,
// loop-free, duplicate-free CSR that the matcher reads (GraphIO::ReadSNAPFile
// + SortCSRArray, src/graph/graph_io.h:193-274, and PreprocessGraph::Preprocess,
// src/preprocess/preprocess_graph.h:12-92). Every entry is one u64 key
// (row << 32 | col); one radix sort over the bits that carry data orders the
// whole CSR at once, duplicates become neighbours, and self loops are mapped
// to a sentinel key that sorts past every row. Row boundaries are found in
// one sweep over the unique keys, each vertex's row_ptr written exactly once.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "vsgm_hip.h"

namespace vsgm_internal {
hipStream_t stream(vsgm_ctx* ctx);
int fail_msg(vsgm_ctx* ctx, int code, const char* msg);
}  // namespace vsgm_internal

namespace {

constexpr int kBlock = 256;
constexpr unsigned kMaxGrid = 16384;

unsigned grid_for(uint64_t n) { return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + kBlock - 1) / kBlock, kMaxGrid)); }

// keys[i] (and keys[n + i] for the reversed pair) from pairs[2i], pairs[2i+1];
// self loops become the sentinel V << 32; ids >= V are counted in *bad.
__global__ void k_emit(const uint32_t* __restrict__ pairs, uint64_t n, uint64_t V, int symmetric, uint64_t* __restrict__ keys,
                       unsigned long long* bad) {
  const uint64_t sentinel = V << 32;
  unsigned long long local_bad = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t u = pairs[2 * i], v = pairs[2 * i + 1];
    const bool ok = u < V && v < V;
    local_bad += ok ? 0 : 1;
    const bool drop = !ok || u == v;
    keys[i] = drop ? sentinel : (u << 32 | v);
    if (symmetric) keys[n + i] = drop ? sentinel : (v << 32 | u);
  }
  if (local_bad) atomicAdd(bad, local_bad);
}

// cols[i] = low half of key i; row_ptrs[x] = i for every x in
// (row(key[i-1]), row(key[i])], and M for every x past the last row.
__global__ void k_split(const uint64_t* __restrict__ keys, uint64_t M, uint64_t V, uint32_t* __restrict__ cols,
                        uint64_t* __restrict__ row_ptrs) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < M; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t key = keys[i];
    const uint64_t u = key >> 32;
    cols[i] = (uint32_t)key;
    const uint64_t first = i == 0 ? 0 : (keys[i - 1] >> 32) + 1;
    for (uint64_t x = first; x <= u; ++x) row_ptrs[x] = i;
    if (i + 1 == M)
      for (uint64_t x = u + 1; x <= V; ++x) row_ptrs[x] = M;
  }
}

__global__ void k_fill_u64(uint64_t* __restrict__ p, uint64_t n, uint64_t value) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = value;
}

__global__ void k_zero_ull(unsigned long long* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

// flag[u] = 1 when u's list is not empty (u < V); flag[V] = 0.
__global__ void k_nonempty(const uint64_t* __restrict__ rp, uint64_t V, uint32_t* __restrict__ flag) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u <= V; u += (uint64_t)gridDim.x * blockDim.x)
    flag[u] = u < V && rp[u + 1] > rp[u] ? 1u : 0u;
}

// Renumbered row_ptrs: the list of u (not empty) moves to row id[u]; the
// final entry (row id[V] = number of kept vertices) is M.
__global__ void k_compact_rows(const uint64_t* __restrict__ rp, const uint32_t* __restrict__ id, uint64_t V, uint64_t* __restrict__ out) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u <= V; u += (uint64_t)gridDim.x * blockDim.x)
    if (u == V || rp[u + 1] > rp[u]) out[id[u]] = rp[u];
}

__global__ void k_remap(uint32_t* __restrict__ cols, uint64_t M, const uint32_t* __restrict__ id) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < M; i += (uint64_t)gridDim.x * blockDim.x) cols[i] = id[cols[i]];
}

int check(vsgm_ctx* ctx, const char* what) {
  if (hipGetLastError() != hipSuccess) return vsgm_internal::fail_msg(ctx, VSGM_ERR_HIP, what);
  return VSGM_OK;
}

unsigned bits_for(uint64_t x) {
  unsigned b = 0;
  while (b < 64 && (x >> b) != 0) ++b;
  return b;
}

}  // namespace

extern "C" {

int vsgm_csr_from_pairs(vsgm_ctx* ctx, const uint32_t* d_pairs, uint64_t n_pairs, uint64_t vertex_count, int symmetric, int drop_isolated,
                        vsgm_view_out* out) {
  if (!ctx) return VSGM_ERR_ARG;
  if (!out || (n_pairs && !d_pairs) || vertex_count == 0 || vertex_count > 0xFFFFFFFFull)
    return vsgm_internal::fail_msg(ctx, VSGM_ERR_ARG, "bad edge-list arguments");
  if (drop_isolated && !symmetric)
    return vsgm_internal::fail_msg(ctx, VSGM_ERR_ARG, "isolated-vertex removal needs a symmetric graph");
  if (n_pairs > (1ull << 59)) return vsgm_internal::fail_msg(ctx, VSGM_ERR_ARG, "edge list too long");
  const uint64_t V = vertex_count;
  const uint64_t N = symmetric ? 2 * n_pairs : n_pairs;
  hipStream_t s = vsgm_internal::stream(ctx);
  std::vector<void*> owned;
  auto get = [&](uint64_t bytes, void** p) {
    int st = vsgm_alloc(ctx, std::max<uint64_t>(bytes, 8), p);
    if (st == VSGM_OK) owned.push_back(*p);
    return st;
  };
  auto release = [&](void* p) {
    owned.erase(std::find(owned.begin(), owned.end(), p));
    vsgm_free(ctx, p);
  };
  auto done = [&](int st) {
    for (void* p : owned) vsgm_free(ctx, p);
    return st;
  };
  uint64_t *keys_a = nullptr, *keys_b = nullptr, *counters = nullptr, *row_ptrs = nullptr;
  uint32_t* cols = nullptr;
  void* tmp = nullptr;
  int st = get(N * sizeof(uint64_t), (void**)&keys_a);
  if (st == VSGM_OK) st = get(N * sizeof(uint64_t), (void**)&keys_b);
  if (st == VSGM_OK) st = get(2 * sizeof(uint64_t), (void**)&counters);  // [0] bad ids, [1] unique count
  if (st != VSGM_OK) return done(st);
  unsigned long long* bad = (unsigned long long*)counters;

  // 1. keys
  hipLaunchKernelGGL(k_zero_ull, dim3(1), dim3(64), 0, s, bad);
  if (n_pairs) hipLaunchKernelGGL(k_emit, dim3(grid_for(n_pairs)), dim3(kBlock), 0, s, d_pairs, n_pairs, V, symmetric, keys_a, bad);
  if ((st = check(ctx, "edge key launch failed")) != VSGM_OK) return done(st);

  // 2. sort over the bits in use (rows < V plus the sentinel row V), 3. unique
  uint64_t M_all = 0;
  if (N) {
    const unsigned end_bit = 32 + bits_for(V);
    size_t sort_bytes = 0, uniq_bytes = 0;
    if (rocprim::radix_sort_keys(nullptr, sort_bytes, keys_a, keys_b, (size_t)N, 0, end_bit, s) != hipSuccess ||
        rocprim::unique(nullptr, uniq_bytes, keys_b, keys_a, counters + 1, (size_t)N, rocprim::equal_to<uint64_t>(), s) != hipSuccess)
      return done(vsgm_internal::fail_msg(ctx, VSGM_ERR_HIP, "sort workspace query failed"));
    if ((st = get(std::max(sort_bytes, uniq_bytes), &tmp)) != VSGM_OK) return done(st);
    size_t bytes = sort_bytes;
    if (rocprim::radix_sort_keys(tmp, bytes, keys_a, keys_b, (size_t)N, 0, end_bit, s) != hipSuccess)
      return done(vsgm_internal::fail_msg(ctx, VSGM_ERR_HIP, "edge key sort failed"));
    bytes = uniq_bytes;
    if (rocprim::unique(tmp, bytes, keys_b, keys_a, counters + 1, (size_t)N, rocprim::equal_to<uint64_t>(), s) != hipSuccess)
      return done(vsgm_internal::fail_msg(ctx, VSGM_ERR_HIP, "edge key deduplication failed"));
  }
  uint64_t h[2] = {0, 0};
  if ((st = vsgm_d2h(ctx, h, counters, sizeof(h))) != VSGM_OK) return done(st);
  if (h[0]) return done(vsgm_internal::fail_msg(ctx, VSGM_ERR_ARG, "edge list holds vertex ids >= vertex_count"));
  if (N) M_all = h[1];
  uint64_t M = M_all;
  if (M_all) {
    uint64_t last = 0;
    if ((st = vsgm_d2h(ctx, &last, keys_a + (M_all - 1), sizeof(last))) != VSGM_OK) return done(st);
    if ((last >> 32) == V) --M;  // the sentinel of the self loops
  }
  if (tmp) release(tmp);
  release(keys_b);

  // 4. rows and columns
  if ((st = get(M * sizeof(uint32_t), (void**)&cols)) != VSGM_OK) return done(st);
  if ((st = get((V + 1) * sizeof(uint64_t), (void**)&row_ptrs)) != VSGM_OK) return done(st);
  if (M)
    hipLaunchKernelGGL(k_split, dim3(grid_for(M)), dim3(kBlock), 0, s, keys_a, M, V, cols, row_ptrs);
  else
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(V + 1)), dim3(kBlock), 0, s, row_ptrs, V + 1, 0ull);
  if ((st = check(ctx, "row split launch failed")) != VSGM_OK) return done(st);
  release(keys_a);

  // 5. isolated vertices out, ids renumbered in order (preprocess_graph.h:67-82)
  uint64_t kept = V;
  if (drop_isolated) {
    uint32_t *flag = nullptr, *id = nullptr;
    uint64_t* rp2 = nullptr;
    if ((st = get((V + 1) * sizeof(uint32_t), (void**)&flag)) != VSGM_OK) return done(st);
    if ((st = get((V + 1) * sizeof(uint32_t), (void**)&id)) != VSGM_OK) return done(st);
    hipLaunchKernelGGL(k_nonempty, dim3(grid_for(V + 1)), dim3(kBlock), 0, s, row_ptrs, V, flag);
    if ((st = check(ctx, "isolated-vertex flag launch failed")) != VSGM_OK) return done(st);
    size_t scan_bytes = 0;
    if (rocprim::exclusive_scan(nullptr, scan_bytes, flag, id, 0u, (size_t)(V + 1), rocprim::plus<uint32_t>(), s) != hipSuccess)
      return done(vsgm_internal::fail_msg(ctx, VSGM_ERR_HIP, "scan workspace query failed"));
    if ((st = get(scan_bytes, &tmp)) != VSGM_OK) return done(st);
    size_t bytes = scan_bytes;
    if (rocprim::exclusive_scan(tmp, bytes, flag, id, 0u, (size_t)(V + 1), rocprim::plus<uint32_t>(), s) != hipSuccess)
      return done(vsgm_internal::fail_msg(ctx, VSGM_ERR_HIP, "vertex renumbering scan failed"));
    uint32_t k32 = 0;
    if ((st = vsgm_d2h(ctx, &k32, id + V, sizeof(k32))) != VSGM_OK) return done(st);
    kept = k32;
    release(tmp);
    release(flag);
    if ((st = get((kept + 1) * sizeof(uint64_t), (void**)&rp2)) != VSGM_OK) return done(st);
    hipLaunchKernelGGL(k_compact_rows, dim3(grid_for(V + 1)), dim3(kBlock), 0, s, row_ptrs, id, V, rp2);
    if (M) hipLaunchKernelGGL(k_remap, dim3(grid_for(M)), dim3(kBlock), 0, s, cols, M, id);
    if ((st = check(ctx, "vertex renumbering launch failed")) != VSGM_OK) return done(st);
    if ((st = vsgm_ctx_sync(ctx)) != VSGM_OK) return done(st);
    release(id);
    release(row_ptrs);
    row_ptrs = rp2;
  }
  if ((st = vsgm_ctx_sync(ctx)) != VSGM_OK) return done(st);
  release(counters);
  out->row_ptrs = row_ptrs;
  out->cols = cols;
  out->entries = M;
  out->vertices = kept;
  return VSGM_OK;  // row_ptrs and cols now belong to the caller
}

}  // extern "C"
