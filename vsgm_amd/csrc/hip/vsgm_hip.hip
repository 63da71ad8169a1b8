// libvsgm_hip.so — gfx950 kernels and the C-ABI declared in include/vsgm_hip.h.
//
// Work decomposition. Every level operator of the BFS join is a segmented
// map over P paths: path p owns a segment [begin_p, begin_p + len_p) of a
// sorted u32 array (the pivot adjacency list of an intersection, or the
// path's candidate row), and each element of every segment is one work
// item. Operators run as
//   setup   one thread per path: gather the path's bound vertices, pick the
//           segment and narrow it with the '<' / '>' conditions by binary
//           search (those conditions are monotone on a sorted list);
//   scan    exclusive u64 scan of the segment lengths -> element offsets;
//   map     256-thread workgroups over fixed chunks of 2048 elements. The
//           chunk's slice of the segment offsets is staged in LDS, so each
//           element finds its path with an LDS binary search
//           (load-balance search); the element is verified and either
//           counted (block reduction + one u64 atomic per chunk) or its
//           validity bit is written with a wave64 ballot (one u64 mask word
//           per wavefront) together with a per-chunk valid count;
//   scatter (compacting ops only) chunk counts are scanned, then each valid
//           element is written at chunk base + popcount of the valid bits
//           before it, so outputs keep (path, candidate) order — candidate
//           rows stay sorted, as the reference requires
//           (src/gpu/gpu_utils_gpsm_bin_search.h:110-126).
// All indices over elements and paths are 64-bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "vsgm_hip.h"

namespace {

constexpr int kBlock = 256;               // threads per workgroup (4 wavefronts)
constexpr int kWave = 64;                 // CDNA wavefront
constexpr int kItems = 8;                 // elements per thread per chunk
constexpr int kChunk = kBlock * kItems;   // elements per chunk (multiple of 64)
constexpr int kWordsPerChunk = kChunk / kWave;
constexpr int kPathCap = kChunk;          // paths whose offsets fit in LDS per chunk
constexpr int kPhaseItems = 4;            // elements a thread verifies in one phased group
constexpr uint64_t kBeginMask = (1ull << 56) - 1;
constexpr int kMaxGrid = 8192;

// LDS membership table of the isect_kernel 3 map (one workgroup per CU):
// two cuckoo tables of kCuckooSlots u32 = 112 KiB; lists of up to kHashCap
// entries keep the load factor <= 0.45.
constexpr uint32_t kCuckooSlots = 14336;
constexpr uint32_t kHashCap = 12800;
constexpr uint32_t kDefaultColdWeight = 16;
constexpr uint32_t kDefaultTicketChunks = 64;

enum OpKind { OP_ICOUNT = 0, OP_IWRITE = 1, OP_EXPAND = 2, OP_FILTER = 3, OP_COUNT2 = 4, OP_IPATH = 5 };

__host__ __device__ constexpr bool is_intersect(int op) { return op == OP_ICOUNT || op == OP_IWRITE || op == OP_IPATH; }

// ---------------------------------------------------------------------------
// device helpers

__device__ inline uint32_t lower_bound_u32(const uint32_t* a, uint32_t n, uint32_t key) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ inline uint32_t upper_bound_u32(const uint32_t* a, uint32_t n, uint32_t key) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] <= key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// index of the last entry <= key in the non-decreasing a[0..n) (a[0] <= key)
__device__ inline uint64_t last_le_u64(const uint64_t* a, uint64_t n, uint64_t key) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (a[mid] <= key) lo = mid + 1;
    else hi = mid;
  }
  return lo - 1;
}

__device__ inline uint64_t first_ge_u64(const uint64_t* a, uint64_t n, uint64_t key) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ inline uint32_t lds_last_le(const uint64_t* a, uint32_t n, uint64_t key) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] <= key) lo = mid + 1;
    else hi = mid;
  }
  return lo - 1;
}

__device__ inline bool cond_holds(uint32_t op, uint32_t c, uint32_t m) {
  return op == VSGM_LESS_THAN ? (c < m) : (op == VSGM_LARGER_THAN ? (c > m) : (c != m));
}

__device__ inline uint64_t cand_row(const vsgm_level_args& a, uint32_t u, uint64_t p) {
  return a.cand_idx[u] ? a.cand_idx[u][p] : a.cand_idx_base[u] + p;
}

// Number of y in sorted s[0..n) with "x op y" (ThreadCountWithComputedVertex).
__device__ inline uint64_t count_against(const uint32_t* s, uint32_t n, uint32_t x, uint32_t op) {
  if (n == 0) return 0;
  uint32_t pos = lower_bound_u32(s, n, x);
  uint32_t eq = (pos < n && s[pos] == x) ? 1u : 0u;
  if (op == VSGM_LESS_THAN) return n - (pos + eq);  // y > x
  if (op == VSGM_LARGER_THAN) return pos;           // y < x
  return n - eq;                                    // y != x
}

// Number of y in sorted s[0..n) with "y op x" (the mirrored side).
__device__ inline uint64_t count_against_mirror(const uint32_t* s, uint32_t n, uint32_t x, uint32_t op) {
  if (n == 0) return 0;
  uint32_t pos = lower_bound_u32(s, n, x);
  uint32_t eq = (pos < n && s[pos] == x) ? 1u : 0u;
  if (op == VSGM_LESS_THAN) return pos;               // y < x
  if (op == VSGM_LARGER_THAN) return n - (pos + eq);  // y > x
  return n - eq;
}

// Wavefront-scope barrier for LDS data exchanged between the lanes of one
// wavefront (the waves of a workgroup diverge here, so no __syncthreads).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr uint32_t kTile = 256;          // LDS tile of one wavefront (u32 entries)
constexpr int kCursorTiles = 4;          // tiles streamed from a cursor before searching HBM
constexpr uint32_t kStageCap = 4096;     // adjacency entries a workgroup can hold in LDS
constexpr int kCursorSlots = VSGM_MAX_QV + 1;

// Largest power of two <= n (0 for n == 0).
__device__ inline uint32_t top_step(uint32_t n) { return n ? (1u << (31 - __clz(n))) : 0u; }

__device__ inline uint32_t wave_max_u32(uint32_t v) {
  for (int d = kWave / 2; d > 0; d >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, d));
  return v;
}

__device__ inline uint64_t readfirstlane_u64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Binary lifting over a sorted a[0..n): a fixed sequence of steps top, top/2,
// ..., 1 with predicated probes, so lanes with different n and keys share one
// instruction stream (no divergent search loops). `top` is a wave-uniform
// power of two >= top_step of every lane's n; a[0] must be readable even
// when n == 0. Returns the number of entries < key (lower bound).
template <typename T>
__device__ inline uint32_t lift_lower_bound(const T* a, uint32_t n, T key, uint32_t top) {
  uint32_t pos = 0;
  for (uint32_t step = top; step; step >>= 1) {
    const uint32_t probe = pos + step;
    const bool in = probe <= n;
    const T v = a[in ? probe - 1 : 0];
    pos = (in && v < key) ? probe : pos;
  }
  return pos;
}

// Same, returning the number of entries <= key (upper bound).
template <typename T>
__device__ inline uint32_t lift_upper_bound(const T* a, uint32_t n, T key, uint32_t top) {
  uint32_t pos = 0;
  for (uint32_t step = top; step; step >>= 1) {
    const uint32_t probe = pos + step;
    const bool in = probe <= n;
    const T v = a[in ? probe - 1 : 0];
    pos = (in && v <= key) ? probe : pos;
  }
  return pos;
}

// Per-lane lower bound in per-lane lists (lanes may hold different lists):
// binary lifting with the wave's largest list length. `safe` is any readable
// u32 used as the base of empty lists.
__device__ inline uint32_t lane_lower_bound(const uint32_t* L, uint32_t n, uint32_t key, const uint32_t* safe, bool& found) {
  const uint32_t top = top_step(wave_max_u32(n));
  const uint32_t* base = n ? L : safe;
  const uint32_t pos = lift_lower_bound(base, n, key, top);
  const uint32_t v = base[pos < n ? pos : 0];
  found = pos < n && v == key;
  return pos;
}

// lower_bound of every lane's key in L[0..n) for a wavefront whose 64 keys
// ascend with the lane index, given a wave-uniform cursor <= lower_bound of
// the smallest key (the previous iteration of the same path left it at the
// lower_bound of its largest key). Tiles of L are streamed from the cursor
// through the wave's LDS tile and searched there until every lane has its
// answer; lanes still unplaced after kCursorTiles tiles (or when the first
// tile shows the cursor far behind) search the rest of L in HBM. Advances
// the cursor to the answer of lane 63.
__device__ inline uint32_t wave_lower_bound_cursor(const uint32_t* __restrict__ L, uint32_t n, uint32_t key, uint32_t& cursor, uint32_t* tile,
                                                   bool& found) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint32_t kmin = __shfl(key, 0);
  uint32_t t = cursor;
  uint32_t pos = n;
  bool placed = false;
  found = false;
  for (int k = 0; k < kCursorTiles && t < n; ++k) {
    const uint32_t m = min(kTile, n - t);
#pragma unroll
    for (uint32_t j = 0; j < kTile / kWave; ++j) {
      const uint32_t i = lane + j * kWave;
      if (i < m) tile[i] = L[t + i];
    }
    wave_sync();
    const uint32_t last = tile[m - 1];
    const bool far = (k == 0) && kmin > last && (n - t > m);
    if (!far) {
      const uint32_t j = lift_lower_bound(tile, m, key, kTile);
      const bool here = !placed && key <= last;  // then j < m
      pos = here ? t + j : pos;
      found = here ? (tile[j < m ? j : 0] == key) : found;
      placed = placed || here;
    }
    wave_sync();
    t += m;
    if (far || __shfl((int)placed, kWave - 1)) break;
  }
  if (t < n && __ballot(!placed) != 0) {
    // lanes past the streamed tiles: binary lifting over the rest in HBM
    const uint32_t r = t + lift_lower_bound(L + t, n - t, key, top_step(n - t));
    const uint32_t rv = L[r < n ? r : n - 1];
    found = placed ? found : (r < n && rv == key);
    pos = placed ? pos : r;
  }
  cursor = __shfl(pos, kWave - 1);
  return pos;
}

// Narrow [lo, hi) of the sorted array arr with the '<' / '>' conditions.
__device__ inline void narrow_range(const vsgm_level_args& a, const uint32_t* arr, uint64_t p, uint64_t& lo, uint64_t& hi) {
  for (uint32_t i = 0; i < a.n_cond; ++i) {
    uint32_t op = a.cond[i].op;
    if (op != VSGM_LESS_THAN && op != VSGM_LARGER_THAN) continue;
    if (hi <= lo) break;
    uint32_t m = a.inst[a.cond[i].operand][p];
    uint32_t n = (uint32_t)(hi - lo);
    if (op == VSGM_LESS_THAN) hi = lo + lower_bound_u32(arr + lo, n, m);
    else lo = lo + upper_bound_u32(arr + lo, n, m);
  }
  if (hi < lo) hi = lo;
}

__device__ inline bool non_equal_ok(const vsgm_level_args& a, uint64_t p, uint32_t c) {
  for (uint32_t i = 0; i < a.n_cond; ++i)
    if (a.cond[i].op == VSGM_NON_EQUAL && a.inst[a.cond[i].operand][p] == c) return false;
  return true;
}

template <typename T>
__device__ inline T wave_sum(T v) {
  for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

// exclusive scan across the workgroup; s_tmp needs kBlock/kWave entries
__device__ inline uint64_t block_excl_scan(uint64_t v, uint64_t* s_tmp, uint64_t& total) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  uint64_t x = v;
  for (int d = 1; d < kWave; d <<= 1) {
    uint64_t y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  if (lane == kWave - 1) s_tmp[wave] = x;
  __syncthreads();
  uint64_t before = 0, all = 0;
  for (int w = 0; w < kBlock / kWave; ++w) {
    uint64_t t = s_tmp[w];
    if (w < wave) before += t;
    all += t;
  }
  __syncthreads();
  total = all;
  return before + x - v;
}

__device__ inline uint64_t block_sum(uint64_t v, uint64_t* s_tmp) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  if (lane == 0) s_tmp[wave] = v;
  __syncthreads();
  uint64_t all = 0;
  for (int w = 0; w < kBlock / kWave; ++w) all += s_tmp[w];
  __syncthreads();
  return all;
}

inline unsigned grid_for(uint64_t units) {
  uint64_t g = units == 0 ? 1 : units;
  return (unsigned)std::min<uint64_t>(g, kMaxGrid);
}

// ---------------------------------------------------------------------------
// primitives

__global__ __launch_bounds__(kBlock) void k_tile_sum(const uint64_t* __restrict__ in, uint64_t n, uint64_t* __restrict__ partial,
                                                     uint64_t n_tiles) {
  __shared__ uint64_t s_tmp[kBlock / kWave];
  for (uint64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    uint64_t base = t * kChunk, s = 0;
    for (int i = 0; i < kItems; ++i) {
      uint64_t e = base + (uint64_t)i * kBlock + threadIdx.x;
      if (e < n) s += in[e];
    }
    s = block_sum(s, s_tmp);
    if (threadIdx.x == 0) partial[t] = s;
  }
}

// Exclusive scan of one tile per workgroup, offset by tile_prefix[t];
// the workgroup owning the last tile writes out[n] = total.
__global__ __launch_bounds__(kBlock) void k_tile_scan(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t n_tiles,
                                                      const uint64_t* __restrict__ tile_prefix) {
  __shared__ uint64_t s_data[kChunk];
  __shared__ uint64_t s_tmp[kBlock / kWave];
  for (uint64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    uint64_t base = t * kChunk;
    for (int i = 0; i < kItems; ++i) {
      uint64_t e = base + (uint64_t)i * kBlock + threadIdx.x;
      s_data[i * kBlock + threadIdx.x] = e < n ? in[e] : 0;
    }
    __syncthreads();
    uint64_t local[kItems];
    uint64_t sum = 0;
    for (int i = 0; i < kItems; ++i) {
      local[i] = s_data[threadIdx.x * kItems + i];
      sum += local[i];
    }
    uint64_t total;
    uint64_t run = block_excl_scan(sum, s_tmp, total);
    uint64_t pre = tile_prefix ? tile_prefix[t] : 0;
    run += pre;
    for (int i = 0; i < kItems; ++i) {
      s_data[threadIdx.x * kItems + i] = run;
      run += local[i];
    }
    __syncthreads();
    for (int i = 0; i < kItems; ++i) {
      uint64_t e = base + (uint64_t)i * kBlock + threadIdx.x;
      if (e < n) out[e] = s_data[i * kBlock + threadIdx.x];
    }
    if (t == n_tiles - 1 && threadIdx.x == 0) out[n] = pre + total;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void k_reduce(const uint64_t* __restrict__ in, uint64_t n, unsigned long long* total) {
  __shared__ uint64_t s_tmp[kBlock / kWave];
  uint64_t s = 0;
  for (uint64_t e = (uint64_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += (uint64_t)gridDim.x * kBlock) s += in[e];
  s = block_sum(s, s_tmp);
  if (threadIdx.x == 0 && s) atomicAdd(total, (unsigned long long)s);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather(const uint64_t* __restrict__ idx, uint64_t n, const T* __restrict__ src,
                                                   T* __restrict__ dst) {
  for (uint64_t e = (uint64_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += (uint64_t)gridDim.x * kBlock) dst[e] = src[idx[e]];
}

__global__ __launch_bounds__(kBlock) void k_iota(uint64_t* d, uint64_t n, uint64_t base) {
  for (uint64_t e = (uint64_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += (uint64_t)gridDim.x * kBlock) d[e] = base + e;
}

__global__ __launch_bounds__(kBlock) void k_fill(uint64_t* d, uint64_t n, uint64_t v) {
  for (uint64_t e = (uint64_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += (uint64_t)gridDim.x * kBlock) d[e] = v;
}

// Hub index: slot[hubs[i]] = i (slot pre-filled with ~0u), then one
// workgroup per hub sets the bits of its adjacency list. Lists are sorted,
// so neighbouring lanes often hit the same word: atomicOr merges them.
__global__ __launch_bounds__(kBlock) void k_self_loops(const uint64_t* __restrict__ row_ptrs, const uint32_t* __restrict__ cols, uint64_t V,
                                                      unsigned long long* __restrict__ count) {
  uint64_t n = 0;
  for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < V; v += (uint64_t)gridDim.x * kBlock) {
    const uint64_t rb = row_ptrs[v];
    const uint32_t len = (uint32_t)(row_ptrs[v + 1] - rb);
    const uint32_t pos = lower_bound_u32(cols + rb, len, (uint32_t)v);
    n += (pos < len && cols[rb + pos] == (uint32_t)v) ? 1 : 0;
  }
  n = wave_sum(n);
  if ((threadIdx.x & (kWave - 1)) == 0 && n) atomicAdd(count, (unsigned long long)n);
}

__global__ __launch_bounds__(kBlock) void k_hub_slots(const uint32_t* __restrict__ hubs, uint32_t n_hubs, uint32_t* __restrict__ slot) {
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n_hubs; i += gridDim.x * kBlock) slot[hubs[i]] = i;
}

__global__ __launch_bounds__(kBlock) void k_hub_bits(const uint64_t* __restrict__ row_ptrs, const uint32_t* __restrict__ cols,
                                                     uint64_t vertex_count, const uint32_t* __restrict__ hubs, uint32_t n_hubs,
                                                     uint32_t* __restrict__ bits, uint64_t words) {
  for (uint32_t s = blockIdx.x; s < n_hubs; s += gridDim.x) {
    const uint32_t v = hubs[s];
    const uint64_t b = row_ptrs[v], e = row_ptrs[v + 1];
    uint32_t* row = bits + (uint64_t)s * words;
    for (uint64_t j = b + threadIdx.x; j < e; j += kBlock) {
      const uint32_t c = cols[j];
      if (c < vertex_count) atomicOr(row + (c >> 5), 1u << (c & 31));
    }
  }
}

// ---------------------------------------------------------------------------
// level operators

// Relative cost of testing one candidate against a list bound to a vertex
// that changes from path to path (its index row or search path is cold:
// one fetched line per test) versus one bound to the same vertex as the
// neighbouring paths (hot in cache); streaming one pivot element costs 1.
constexpr uint64_t kColdTest = 4;

// Is the vertex bound to backward position i the same in a neighbouring path?
__device__ inline bool stable_binding(const vsgm_level_args& a, uint32_t i, uint64_t p, uint32_t v) {
  const uint32_t* col = a.inst[a.back[i]];
  return (p > 0 && col[p - 1] == v) || (p + 1 < a.path_num && col[p + 1] == v);
}

// Per-path segment: begin (low 56 bits, with the pivot position or the
// count side in the top byte) and length. For intersections, *adj_sum
// accumulates sum over paths of the degrees of all backward vertices (the
// merge-model list volume used for the algorithmic byte count).
//
// Pivot choice. The reference streams the shortest backward list
// (ThreadChoosePivotLevel, gpu_check_constraints_common.h:11-23). Any
// backward list gives the same candidates in the same (ascending) order, so
// here the pivot minimises deg(pivot) * (1 + sum of the other lists' test
// costs): a list whose vertex is shared with the neighbouring paths is
// cheap to test against (weight 1), one that changes per path costs
// kColdTest. This streams the fresh list and probes the shared one when the
// fresh list is at most a few times longer.
//
// With hash_cap > 0 (the map kernel keeps one stable list of up to hash_cap
// entries in an LDS table) testing against that list is nearly free: the
// stable list with the most entries that fit gets weight 0, and a list that
// changes per path costs `cold` (one fetched line per test, against 4 bytes
// per streamed pivot element).
template <int OP>
__global__ __launch_bounds__(kBlock) void k_setup(vsgm_level_args a, uint64_t* __restrict__ seg_begin, uint64_t* __restrict__ seg_len,
                                                  unsigned long long* __restrict__ adj_sum, uint32_t hash_cap, uint32_t cold) {
  __shared__ uint64_t s_tmp[kBlock / kWave];
  const uint64_t P = a.path_num;
  const uint64_t cold_w = hash_cap ? cold : kColdTest;
  uint64_t deg_sum = 0;
  for (uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x; p < P; p += (uint64_t)gridDim.x * kBlock) {
    uint64_t lo = 0, hi = 0, tag = 0;
    if (is_intersect(OP)) {
      uint64_t w[VSGM_MAX_QV];
      uint64_t w_total = 0;
      if (a.n_back > 1) {
        uint32_t hashed = ~0u;
        uint64_t hashed_n = 0;
        for (uint32_t i = 0; i < a.n_back; ++i) {
          const uint32_t v = a.inst[a.back[i]][p];
          const bool stable = stable_binding(a, i, p, v);
          w[i] = stable ? 1 : cold_w;
          const uint64_t n = a.row_ptrs[v + 1] - a.row_ptrs[v];
          if (hash_cap && stable && n <= hash_cap && (hashed == ~0u || n > hashed_n)) {
            hashed = i;
            hashed_n = n;
          }
        }
        if (hashed != ~0u) w[hashed] = 0;
        for (uint32_t i = 0; i < a.n_back; ++i) w_total += w[i];
      } else {
        w[0] = 0;
      }
      uint64_t best_cost = 0;
      for (uint32_t i = 0; i < a.n_back; ++i) {
        const uint32_t v = a.inst[a.back[i]][p];
        const uint64_t b = a.row_ptrs[v], e = a.row_ptrs[v + 1];
        deg_sum += e - b;
        const uint64_t cost = (e - b) * (1 + w_total - w[i]);
        if (i == 0 || cost < best_cost) {  // first minimum wins
          best_cost = cost;
          lo = b;
          hi = e;
          tag = i;
        }
      }
      narrow_range(a, a.cols, p, lo, hi);
    } else if (OP == OP_EXPAND || OP == OP_FILTER) {
      uint64_t r = cand_row(a, a.target, p);
      lo = a.cand_off[a.target][r];
      hi = a.cand_off[a.target][r + 1];
      narrow_range(a, a.cand[a.target], p, lo, hi);
    } else {  // OP_COUNT2
      uint64_t r0 = cand_row(a, a.target, p), r1 = cand_row(a, (uint32_t)a.other, p);
      uint64_t b0 = a.cand_off[a.target][r0], n0 = a.cand_off[a.target][r0 + 1] - b0;
      uint64_t b1 = a.cand_off[a.other][r1], n1 = a.cand_off[a.other][r1 + 1] - b1;
      if (n0 <= n1) {
        lo = b0;
        hi = b0 + n0;
        tag = 0;
      } else {
        lo = b1;
        hi = b1 + n1;
        tag = 1;
      }
    }
    seg_begin[p] = lo | (tag << 56);
    seg_len[p] = hi - lo;
  }
  if (is_intersect(OP)) {
    uint64_t s = block_sum(deg_sum, s_tmp);
    if (threadIdx.x == 0 && s) atomicAdd(adj_sum, (unsigned long long)s);
  }
}

// out[0] = *total (the scan total), out[1] stays: packs the two values read
// back after a setup + scan into one 16-byte copy.
__global__ void k_copy1(const uint64_t* __restrict__ src, uint64_t* __restrict__ dst) { dst[0] = src[0]; }

// ---------------------------------------------------------------------------
// device view builder (the view of ViewBin::Materialize, built from the
// graph resident in HBM): dist[v] = BFS level from the qualifying sources
// (0xFF: outside), then the lists of every vertex within hop-1 levels.

constexpr uint8_t kFar = 0xFF;

__global__ __launch_bounds__(kBlock) void k_view_seed(const uint64_t* __restrict__ row_ptrs, const uint32_t* __restrict__ sources,
                                                      uint64_t n_sources, uint64_t V, uint64_t root_degree, uint8_t* __restrict__ dist) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n_sources; i += (uint64_t)gridDim.x * kBlock) {
    const uint32_t s = sources[i];
    if (s >= V) continue;  // not a vertex of this graph
    if (root_degree == 0 || row_ptrs[s + 1] - row_ptrs[s] >= root_degree) dist[s] = 0;
  }
}

// Lists of at least kViewBig entries are not walked by one wavefront (a hub
// of millions of entries would leave one wave running long after the rest):
// the per-vertex pass queues them (up to kViewQueue per pass; more are
// walked in place) and k_view_big spreads each queued list over the grid.
constexpr uint64_t kViewBig = 8192;
constexpr uint32_t kViewQueue = 16384;
constexpr unsigned kViewBigGrid = 1024;

// Queues v (wave-uniform) when its list is long and the queue has room.
__device__ inline bool view_queue_big(uint32_t v, uint64_t n, uint32_t* queue, unsigned int* qn, int lane) {
  if (n < kViewBig) return false;
  uint32_t slot = 0;
  if (lane == 0) slot = atomicAdd(qn, 1u);
  slot = __shfl(slot, 0);
  if (slot >= kViewQueue) return false;
  if (lane == 0) queue[slot] = v;
  return true;
}

// One wavefront per vertex of the previous level; lanes stride its list.
// Several lanes may mark the same vertex: they all store the same level.
__global__ __launch_bounds__(kBlock) void k_view_expand(const uint64_t* __restrict__ row_ptrs, const uint32_t* __restrict__ cols,
                                                        uint64_t V, uint8_t level, uint8_t* __restrict__ dist, uint32_t* __restrict__ queue,
                                                        unsigned int* __restrict__ qn) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t waves = (uint64_t)gridDim.x * (kBlock / kWave);
  for (uint64_t v = (uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; v < V; v += waves) {
    if (dist[v] != (uint8_t)(level - 1)) continue;
    const uint64_t b = row_ptrs[v], e = row_ptrs[v + 1];
    if (view_queue_big((uint32_t)v, e - b, queue, qn, lane)) continue;
    for (uint64_t j = b + lane; j < e; j += kWave) {
      const uint32_t w = cols[j];
      if (dist[w] == kFar) dist[w] = level;
    }
  }
}

// The queued lists, each spread over every wavefront of the grid (COPY:
// copied into the view at vrp[v]; else: their far vertices marked level).
// The starting wave is rotated per list so short queued lists land on
// different waves.
template <bool COPY>
__global__ __launch_bounds__(kBlock) void k_view_big(const uint64_t* __restrict__ row_ptrs, const uint32_t* __restrict__ cols,
                                                     const uint32_t* __restrict__ queue, const unsigned int* __restrict__ qn, uint8_t level,
                                                     uint8_t* __restrict__ dist, const uint64_t* __restrict__ vrp, uint32_t* __restrict__ vcols) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t waves = (uint64_t)gridDim.x * (kBlock / kWave);
  const uint64_t wave = (uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  const uint32_t n = min(*qn, kViewQueue);
  for (uint32_t q = 0; q < n; ++q) {
    const uint32_t v = queue[q];
    const uint64_t b = row_ptrs[v], len = row_ptrs[v + 1] - b;
    const uint64_t w0 = (wave + (uint64_t)q * 37) % waves;
    for (uint64_t j = w0 * kWave + lane; j < len; j += waves * kWave) {
      if (COPY) {
        vcols[vrp[v] + j] = cols[b + j];
      } else {
        const uint32_t w = cols[b + j];
        if (dist[w] == kFar) dist[w] = level;
      }
    }
  }
}

// len[v] = deg(v) inside the view, 0 outside; *n_in += vertices inside.
__global__ __launch_bounds__(kBlock) void k_view_len(const uint64_t* __restrict__ row_ptrs, const uint8_t* __restrict__ dist, uint64_t V,
                                                     uint64_t* __restrict__ len, unsigned long long* __restrict__ n_in) {
  __shared__ uint64_t s_tmp[kBlock / kWave];
  uint64_t cnt = 0;
  for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < V; v += (uint64_t)gridDim.x * kBlock) {
    const bool in = dist[v] != kFar;
    len[v] = in ? row_ptrs[v + 1] - row_ptrs[v] : 0;
    cnt += in ? 1 : 0;
  }
  const uint64_t s = block_sum(cnt, s_tmp);
  if (threadIdx.x == 0 && s) atomicAdd(n_in, (unsigned long long)s);
}

// One wavefront per vertex inside the view copies its list (coalesced);
// long lists are queued for k_view_big.
__global__ __launch_bounds__(kBlock) void k_view_copy(const uint64_t* __restrict__ row_ptrs, const uint32_t* __restrict__ cols,
                                                      const uint8_t* __restrict__ dist, uint64_t V, const uint64_t* __restrict__ vrp,
                                                      uint32_t* __restrict__ vcols, uint32_t* __restrict__ queue, unsigned int* __restrict__ qn) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t waves = (uint64_t)gridDim.x * (kBlock / kWave);
  for (uint64_t v = (uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; v < V; v += waves) {
    if (dist[v] == kFar) continue;
    const uint64_t b = row_ptrs[v], n = row_ptrs[v + 1] - b, o = vrp[v];
    if (view_queue_big((uint32_t)v, n, queue, qn, lane)) continue;
    for (uint64_t j = lane; j < n; j += kWave) vcols[o + j] = cols[b + j];
  }
}

struct ExpandCols {
  uint32_t n_mat;
  uint32_t mat[VSGM_MAX_QV];
  uint32_t n_cu;
  uint32_t cu[VSGM_MAX_QV];
  uint32_t* out_inst[VSGM_MAX_QV];
  uint64_t* out_idx[VSGM_MAX_QV];
};

template <int OP>
__device__ inline const uint32_t* seg_source(const vsgm_level_args& a, uint64_t tag) {
  if (is_intersect(OP)) return a.cols;
  if (OP == OP_COUNT2) return tag ? a.cand[a.other] : a.cand[a.target];
  return a.cand[a.target];
}

// Verify element c of path p. Returns its contribution (count ops) and sets
// valid.
template <int OP>
__device__ inline uint64_t eval_element(const vsgm_level_args& a, uint64_t p, uint64_t tag, uint32_t c, bool& valid,
                                        uint32_t stage_v = ~0u, const uint32_t* s_list = nullptr, uint32_t stage_n = 0) {
  if (is_intersect(OP)) {
    bool ok = true;
    for (uint32_t i = 0; i < a.n_back && ok; ++i) {
      if (i == tag) continue;
      uint32_t v = a.inst[a.back[i]][p];
      if (a.hub_slot) {
        // indexed list: one bit probe instead of a binary search
        const uint32_t s = a.hub_slot[v];
        if (s != ~0u) {
          ok = (a.hub_bits[(uint64_t)s * a.hub_words + (c >> 5)] >> (c & 31)) & 1u;
          continue;
        }
      }
      if (v == stage_v) {
        // the workgroup holds this adjacency list in LDS
        uint32_t pos = lower_bound_u32(s_list, stage_n, c);
        ok = pos < stage_n && s_list[pos] == c;
        continue;
      }
      uint64_t b = a.row_ptrs[v];
      uint32_t n = (uint32_t)(a.row_ptrs[v + 1] - b);
      uint32_t pos = lower_bound_u32(a.cols + b, n, c);
      ok = pos < n && a.cols[b + pos] == c;
    }
    ok = ok && non_equal_ok(a, p, c);
    valid = ok;
    if (!ok) return 0;
    if (OP == OP_ICOUNT && a.other >= 0) {
      uint64_t r = cand_row(a, (uint32_t)a.other, p);
      uint64_t b = a.cand_off[a.other][r];
      uint32_t n = (uint32_t)(a.cand_off[a.other][r + 1] - b);
      return count_against(a.cand[a.other] + b, n, c, a.other_op);
    }
    return 1;
  } else if (OP == OP_EXPAND || OP == OP_FILTER) {
    valid = non_equal_ok(a, p, c);
    return valid ? 1 : 0;
  } else {  // OP_COUNT2: c comes from the smaller side
    valid = true;
    if (tag == 0) {
      uint64_t r = cand_row(a, (uint32_t)a.other, p);
      uint64_t b = a.cand_off[a.other][r];
      uint32_t n = (uint32_t)(a.cand_off[a.other][r + 1] - b);
      return count_against(a.cand[a.other] + b, n, c, a.other_op);
    } else {
      uint64_t r = cand_row(a, a.target, p);
      uint64_t b = a.cand_off[a.target][r];
      uint32_t n = (uint32_t)(a.cand_off[a.target][r + 1] - b);
      return count_against_mirror(a.cand[a.target] + b, n, c, a.other_op);
    }
  }
}

// Same verification as eval_element for a wavefront whose 64 lanes all hold
// consecutive candidates of one path: every list lookup goes through
// wave_lower_bound (LDS-staged windows of the other lists).
template <int OP>
__device__ inline uint64_t eval_element_wave(const vsgm_level_args& a, uint64_t p, uint64_t tag, uint32_t c, bool& valid, uint32_t* tile,
                                             uint32_t* cursor) {
  if (is_intersect(OP)) {
    bool ok = true;
    for (uint32_t i = 0; i < a.n_back; ++i) {
      if (i == tag) continue;
      if (__ballot(ok) == 0) break;
      const uint32_t v = a.inst[a.back[i]][p];
      const uint64_t b = a.row_ptrs[v];
      const uint32_t n = (uint32_t)(a.row_ptrs[v + 1] - b);
      bool f;
      wave_lower_bound_cursor(a.cols + b, n, c, cursor[i], tile, f);
      ok = ok && f;
    }
    ok = ok && non_equal_ok(a, p, c);
    valid = ok;
    if (OP == OP_ICOUNT && a.other >= 0) {
      if (__ballot(ok) == 0) return 0;
      const uint64_t r = cand_row(a, (uint32_t)a.other, p);
      const uint64_t b = a.cand_off[a.other][r];
      const uint32_t n = (uint32_t)(a.cand_off[a.other][r + 1] - b);
      bool f;
      const uint32_t pos = wave_lower_bound_cursor(a.cand[a.other] + b, n, c, cursor[VSGM_MAX_QV], tile, f);
      if (!ok) return 0;
      const uint32_t eq = f ? 1u : 0u;
      if (a.other_op == VSGM_LESS_THAN) return n - (pos + eq);
      if (a.other_op == VSGM_LARGER_THAN) return pos;
      return n - eq;
    }
    return ok ? 1 : 0;
  } else {  // OP_COUNT2
    valid = true;
    const uint32_t u = tag == 0 ? (uint32_t)a.other : a.target;
    const uint64_t r = cand_row(a, u, p);
    const uint64_t b = a.cand_off[u][r];
    const uint32_t n = (uint32_t)(a.cand_off[u][r + 1] - b);
    bool f;
    const uint32_t pos = wave_lower_bound_cursor(a.cand[u] + b, n, c, cursor[0], tile, f);
    const uint32_t eq = f ? 1u : 0u;
    if (tag == 0) {  // c from the target, counted against the other: "c op y"
      if (a.other_op == VSGM_LESS_THAN) return n - (pos + eq);
      if (a.other_op == VSGM_LARGER_THAN) return pos;
    } else {  // c from the other, counted against the target: "y op c"
      if (a.other_op == VSGM_LESS_THAN) return pos;
      if (a.other_op == VSGM_LARGER_THAN) return n - (pos + eq);
    }
    return n - eq;
  }
}

// Verification for a wavefront whose lanes hold different paths. Every lane
// executes it (inactive lanes pass active = false and a readable element);
// all list searches are binary lifting with the wave's longest list, so the
// lanes never diverge in a search loop.
template <int OP>
__device__ inline uint64_t eval_lanes(const vsgm_level_args& a, uint64_t p, uint32_t tag, uint32_t c, bool active, bool& valid) {
  if (is_intersect(OP)) {
    bool ok = active;
    for (uint32_t i = 0; i < a.n_back; ++i) {
      if (__ballot(ok) == 0) break;
      const bool skip = (i == tag);
      const uint32_t v = a.inst[a.back[i]][p];
      const uint64_t b = a.row_ptrs[v];
      const uint32_t n = (ok && !skip) ? (uint32_t)(a.row_ptrs[v + 1] - b) : 0u;
      bool f;
      lane_lower_bound(a.cols + b, n, c, a.cols, f);
      ok = skip ? ok : (ok && f);
    }
    ok = ok && non_equal_ok(a, p, c);
    valid = ok;
    if (OP == OP_ICOUNT && a.other >= 0) {
      if (__ballot(ok) == 0) return 0;
      const uint64_t r = cand_row(a, (uint32_t)a.other, p);
      const uint64_t b = a.cand_off[a.other][r];
      const uint32_t n_all = (uint32_t)(a.cand_off[a.other][r + 1] - b);
      bool f;
      const uint32_t pos = lane_lower_bound(a.cand[a.other] + b, ok ? n_all : 0u, c, a.cand[a.other], f);
      if (!ok) return 0;
      const uint32_t eq = f ? 1u : 0u;
      if (a.other_op == VSGM_LESS_THAN) return n_all - (pos + eq);
      if (a.other_op == VSGM_LARGER_THAN) return pos;
      return n_all - eq;
    }
    return ok ? 1 : 0;
  } else {  // OP_COUNT2
    valid = active;
    const uint32_t u = tag == 0 ? (uint32_t)a.other : a.target;
    const uint64_t r = cand_row(a, u, p);
    const uint64_t b = a.cand_off[u][r];
    const uint32_t n_all = (uint32_t)(a.cand_off[u][r + 1] - b);
    bool f;
    const uint32_t pos = lane_lower_bound(a.cand[u] + b, active ? n_all : 0u, c, a.cand[u], f);
    if (!active) return 0;
    const uint32_t eq = f ? 1u : 0u;
    if (tag == 0) {
      if (a.other_op == VSGM_LESS_THAN) return n_all - (pos + eq);
      if (a.other_op == VSGM_LARGER_THAN) return pos;
    } else {
      if (a.other_op == VSGM_LESS_THAN) return pos;
      if (a.other_op == VSGM_LARGER_THAN) return n_all - (pos + eq);
    }
    return n_all - eq;
  }
}

// Chunk bookkeeping shared by the map and scatter kernels: the paths whose
// segments overlap elements [e0, e1) and (if they fit) their offsets in LDS.
struct ChunkPaths {
  uint64_t first;  // path of element e0
  uint64_t last;   // path of element e1-1
  uint32_t staged; // number of offsets staged in LDS (0: search global memory)
};

__device__ inline void stage_chunk(const uint64_t* seg_off, uint64_t P, uint64_t e0, uint64_t e1, ChunkPaths& cp, uint64_t* s_off,
                                   ChunkPaths* s_cp) {
  if (threadIdx.x == 0) {
    ChunkPaths c;
    c.first = last_le_u64(seg_off, P + 1, e0);
    c.last = last_le_u64(seg_off, P + 1, e1 - 1);
    uint64_t span = c.last - c.first + 2;  // offsets first..last+1
    c.staged = span <= (uint64_t)kPathCap + 1 ? (uint32_t)span : 0;
    *s_cp = c;
  }
  __syncthreads();
  cp = *s_cp;
  for (uint32_t i = threadIdx.x; i < cp.staged; i += kBlock) s_off[i] = seg_off[cp.first + i];
  __syncthreads();
}

__device__ inline uint64_t path_of(uint64_t e, const ChunkPaths& cp, const uint64_t* s_off, const uint64_t* seg_off, uint64_t& start) {
  if (cp.staged) {
    uint32_t i = lds_last_le(s_off, cp.staged, e);
    start = s_off[i];
    return cp.first + i;
  }
  uint64_t i = last_le_u64(seg_off + cp.first, cp.last - cp.first + 2, e);
  start = seg_off[cp.first + i];
  return cp.first + i;
}

// Element -> path map of one chunk: the segment starts of the chunk's paths,
// relative to the chunk start, are staged in LDS once per chunk
// (s_rel[i] = start of path first+i, negative for the path that began
// earlier), and each element finds its path with a binary search there
// (load-balanced search over an LDS-resident offset window). A chunk that
// touches more than kPathCap paths searches the global offsets instead.
struct ChunkMap {
  uint64_t first;    // path of the chunk's first element
  uint64_t last;     // path of the chunk's last element
  uint32_t mapped;   // 1: s_rel valid; 0: more paths than kPathCap
};

// The chunks one workgroup processes, in increasing order: groups of kGroup
// consecutive chunks (counted from c_lo) dealt round-robin to the
// workgroups. (Placing neighbouring chunks on one XCD, so that they share an
// L2, measured no faster: the index rows and lists they share are served
// from the device-wide cache level either way.)
template <uint64_t kGroup>
__device__ inline uint64_t block_chunk(uint64_t k, uint64_t c_lo = 0) {
  return c_lo + ((uint64_t)blockIdx.x + (k / kGroup) * gridDim.x) * kGroup + k % kGroup;
}

// First and last path (and, for the CSR-writing scatter, the range of paths
// whose segment starts in the chunk) of the workgroup's next kBlock chunks,
// one chunk per thread: the binary searches over the path offsets run in
// parallel instead of one thread's serial searches per chunk.
struct ChunkTable {
  uint64_t first[kBlock];
  uint64_t last[kBlock];
  uint64_t p_lo[kBlock];
  uint64_t p_hi[kBlock];
};

template <uint64_t kGroup, bool kStarts>
__device__ inline void fill_chunk_table(const uint64_t* __restrict__ seg_off, uint64_t P, uint64_t W, uint64_t n_chunks, uint64_t k0,
                                        ChunkTable* tab, uint64_t c_lo = 0, uint64_t c_hi = ~0ull) {
  const uint64_t chunk = block_chunk<kGroup>(k0 + threadIdx.x, c_lo);
  if (chunk < n_chunks && chunk < c_hi) {
    const uint64_t e0 = chunk * kChunk;
    const uint64_t e1 = min(e0 + (uint64_t)kChunk, W);
    tab->first[threadIdx.x] = last_le_u64(seg_off, P + 1, e0);
    tab->last[threadIdx.x] = last_le_u64(seg_off, P + 1, e1 - 1);
    if (kStarts) {
      tab->p_lo[threadIdx.x] = first_ge_u64(seg_off, P + 1, e0);
      tab->p_hi[threadIdx.x] = chunk == n_chunks - 1 ? P + 1 : first_ge_u64(seg_off, P + 1, e1);
    }
  }
  __syncthreads();
}

__device__ inline void build_chunk_map(const uint64_t* __restrict__ seg_off, uint64_t first, uint64_t last, uint64_t e0, ChunkMap& cm,
                                       int32_t* s_rel) {
  cm.first = first;
  cm.last = last;
  cm.mapped = (last - first + 1) <= (uint64_t)kPathCap ? 1u : 0u;
  if (!cm.mapped) return;
  const uint32_t n_paths = (uint32_t)(cm.last - cm.first + 1);
  for (uint32_t i = threadIdx.x; i <= n_paths; i += kBlock) s_rel[i] = (int32_t)((int64_t)seg_off[cm.first + i] - (int64_t)e0);
  __syncthreads();
}

// Path of element e in [e0, e1) and the element's rank inside its segment.
__device__ inline uint64_t chunk_path(uint64_t e, uint64_t e0, const ChunkMap& cm, const int32_t* s_rel, const uint64_t* __restrict__ seg_off,
                                      uint64_t& rank) {
  if (cm.mapped) {
    // last staged path whose segment starts at or before the element
    const int32_t j = (int32_t)(e - e0);
    uint32_t lo = 0, hi = (uint32_t)(cm.last - cm.first + 2);
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_rel[mid] <= j) lo = mid + 1;
      else hi = mid;
    }
    const uint32_t i = lo - 1;
    rank = (uint64_t)((int64_t)j - (int64_t)s_rel[i]);
    return cm.first + i;
  }
  const uint64_t i = last_le_u64(seg_off + cm.first, cm.last - cm.first + 2, e);
  rank = e - seg_off[cm.first + i];
  return cm.first + i;
}

// LDS membership table of one adjacency list (the map kernel of
// isect_kernel 3 keeps the list of a vertex bound across a run of paths
// here): two-choice cuckoo hashing over two tables of m u32 slots, a key
// sitting in t0[h0(key)] or t1[h1(key)], so a lookup is two independent LDS
// reads with no probe loop.
constexpr uint32_t kEmptySlot = ~0u;
constexpr int kCuckooRounds = 64;  // evictions one insertion may cause before the build gives up

struct HashView {
  const uint32_t* t0;
  const uint32_t* t1;
  uint32_t m;  // slots in use per table
  uint32_t v;  // vertex whose list the tables hold (~0u: none)
};

__device__ inline uint32_t cuckoo_h0(uint32_t c, uint32_t m) { return (uint32_t)(((uint64_t)(c * 0x9E3779B1u) * m) >> 32); }

__device__ inline uint32_t cuckoo_h1(uint32_t c, uint32_t m) {
  uint32_t x = c ^ (c >> 16);
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  return (uint32_t)(((uint64_t)(x * 0xC2B2AE35u) * m) >> 32);
}

__device__ inline bool hash_contains(const HashView& h, uint32_t c) {
  const uint32_t a = h.t0[cuckoo_h0(c, h.m)];
  const uint32_t b = h.t1[cuckoo_h1(c, h.m)];
  return a == c || b == c;
}

// Intersection verification of a lane's elements base + it * 64 + lane of
// the wavefront's run (it in [it0, it0 + kN)), phase by phase: every phase
// issues its kN independent loads back to back before any result is used,
// so one wavefront keeps up to kN dependent-load chains in flight (path ->
// segment -> candidate -> bound vertex -> index row -> bit) instead of one.
// Sets valid[it]; returns the summed count contribution (OP_ICOUNT) and
// bumps path_counts (OP_IPATH).
template <int OP, int kN>
__device__ inline uint64_t verify_items(const vsgm_level_args& a, const uint64_t* __restrict__ seg_begin, const uint64_t* __restrict__ seg_off,
                                        uint64_t e0, uint64_t base, uint64_t e1, const ChunkMap& cm, const int32_t* s_rel, int it0, bool* valid,
                                        unsigned long long* __restrict__ path_counts) {
  uint64_t p[kN];
  uint32_t c[kN], pivot[kN];
  bool ok[kN];
  {
    uint32_t rank[kN];
#pragma unroll
    for (int it = 0; it < kN; ++it) {
      const uint64_t e = base + (uint64_t)(it0 + it) * kWave + (threadIdx.x & (kWave - 1));
      ok[it] = e < e1;
      uint64_t r = 0;
      p[it] = ok[it] ? chunk_path(e, e0, cm, s_rel, seg_off, r) : cm.first;
      rank[it] = (uint32_t)r;
    }
    uint64_t packed[kN];
#pragma unroll
    for (int it = 0; it < kN; ++it) packed[it] = ok[it] ? seg_begin[p[it]] : 0;
#pragma unroll
    for (int it = 0; it < kN; ++it) {
      pivot[it] = (uint32_t)(packed[it] >> 56);
      c[it] = ok[it] ? a.cols[(packed[it] & kBeginMask) + rank[it]] : 0u;
    }
  }
  for (uint32_t i = 0; i < a.n_back; ++i) {
    const uint32_t* col = a.inst[a.back[i]];
    uint32_t v[kN], s[kN], w[kN];
#pragma unroll
    for (int it = 0; it < kN; ++it) v[it] = (ok[it] && pivot[it] != i) ? col[p[it]] : ~0u;
    if (a.hub_slot) {
#pragma unroll
      for (int it = 0; it < kN; ++it) s[it] = v[it] != ~0u ? a.hub_slot[v[it]] : ~0u;
#pragma unroll
      for (int it = 0; it < kN; ++it) w[it] = s[it] != ~0u ? a.hub_bits[(uint64_t)s[it] * a.hub_words + (c[it] >> 5)] : 0u;
    } else {
#pragma unroll
      for (int it = 0; it < kN; ++it) s[it] = ~0u, w[it] = 0u;
    }
#pragma unroll
    for (int it = 0; it < kN; ++it) {
      if (v[it] == ~0u) continue;
      if (s[it] != ~0u) {
        ok[it] = (w[it] >> (c[it] & 31)) & 1u;
      } else {
        const uint64_t b = a.row_ptrs[v[it]];
        const uint32_t n = (uint32_t)(a.row_ptrs[v[it] + 1] - b);
        const uint32_t pos = lower_bound_u32(a.cols + b, n, c[it]);
        ok[it] = pos < n && a.cols[b + pos] == c[it];
      }
    }
  }
  for (uint32_t i = 0; i < a.n_cond; ++i) {
    if (a.cond[i].op != VSGM_NON_EQUAL) continue;
    const uint32_t* col = a.inst[a.cond[i].operand];
#pragma unroll
    for (int it = 0; it < kN; ++it)
      if (ok[it] && col[p[it]] == c[it]) ok[it] = false;
  }
  uint64_t acc = 0;
#pragma unroll
  for (int it = 0; it < kN; ++it) {
    valid[it0 + it] = ok[it];
    if (!ok[it]) continue;
    if (OP == OP_IPATH) {
      atomicAdd(path_counts + p[it], 1ull);
    } else if (OP == OP_ICOUNT) {
      if (a.other >= 0) {
        const uint64_t r = cand_row(a, (uint32_t)a.other, p[it]);
        const uint64_t b = a.cand_off[a.other][r];
        const uint32_t n = (uint32_t)(a.cand_off[a.other][r + 1] - b);
        acc += count_against(a.cand[a.other] + b, n, c[it], a.other_op);
      } else {
        acc += 1;
      }
    }
  }
  return acc;
}

// Intersection verification of one wavefront's run of kItems * 64
// consecutive elements [base, min(base + 512, e1)), element base + it * 64 +
// lane per lane. When the whole run lies in one path (the common case on
// hub-heavy levels, whose segments are long) the path's segment, bound
// vertices, index rows and condition operands are wave-uniform and read once
// with scalar loads; each lane then only streams its candidates (coalesced)
// and probes the index bits or searches the lists. Other runs take the
// per-lane phased path. Sets valid[it] for every item.
template <int OP>
__device__ inline uint64_t verify_run(const vsgm_level_args& a, const uint64_t* __restrict__ seg_begin, const uint64_t* __restrict__ seg_off,
                                      uint64_t e0, uint64_t base, uint64_t e1, const ChunkMap& cm, const int32_t* s_rel, bool (&valid)[kItems],
                                      unsigned long long* __restrict__ path_counts) {
  const int lane = threadIdx.x & (kWave - 1);
  if (base >= e1) {
#pragma unroll
    for (int it = 0; it < kItems; ++it) valid[it] = false;
    return 0;
  }
  const uint64_t last = min(base + (uint64_t)(kItems * kWave), e1) - 1;
  uint64_t r_first = 0, r_last = 0;
  const uint64_t p = readfirstlane_u64(chunk_path(base, e0, cm, s_rel, seg_off, r_first));
  const uint64_t p_last = readfirstlane_u64(chunk_path(last, e0, cm, s_rel, seg_off, r_last));
  if (p != p_last) {
    uint64_t acc = 0;
#pragma unroll
    for (int it0 = 0; it0 < kItems; it0 += kPhaseItems)
      acc += verify_items<OP, kPhaseItems>(a, seg_begin, seg_off, e0, base, e1, cm, s_rel, it0, valid, path_counts);
    return acc;
  }
  // one path: candidate of element e is cols[begin + (e - start)]
  const uint64_t packed = seg_begin[p];
  const uint32_t pivot = (uint32_t)(packed >> 56);
  const uint64_t start = base - readfirstlane_u64(r_first);
  const uint64_t begin = packed & kBeginMask;
  uint32_t c[kItems];
  bool ok[kItems];
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint64_t e = base + (uint64_t)it * kWave + lane;
    ok[it] = e <= last;
    c[it] = ok[it] ? a.cols[begin + (e - start)] : 0u;
  }
  for (uint32_t i = 0; i < a.n_back; ++i) {
    if (i == pivot) continue;
    const uint32_t v = a.inst[a.back[i]][p];
    const uint32_t s = a.hub_slot ? a.hub_slot[v] : ~0u;
    if (s != ~0u) {
      const uint32_t* row = a.hub_bits + (uint64_t)s * a.hub_words;
      uint32_t w[kItems];
#pragma unroll
      for (int it = 0; it < kItems; ++it) w[it] = ok[it] ? row[c[it] >> 5] : 0u;
#pragma unroll
      for (int it = 0; it < kItems; ++it) ok[it] = ok[it] && ((w[it] >> (c[it] & 31)) & 1u);
    } else {
      const uint64_t b = a.row_ptrs[v];
      const uint32_t n = (uint32_t)(a.row_ptrs[v + 1] - b);
#pragma unroll
      for (int it = 0; it < kItems; ++it) {
        if (!ok[it]) continue;
        const uint32_t pos = lower_bound_u32(a.cols + b, n, c[it]);
        ok[it] = pos < n && a.cols[b + pos] == c[it];
      }
    }
  }
  for (uint32_t i = 0; i < a.n_cond; ++i) {
    if (a.cond[i].op != VSGM_NON_EQUAL) continue;
    const uint32_t m = a.inst[a.cond[i].operand][p];
#pragma unroll
    for (int it = 0; it < kItems; ++it) ok[it] = ok[it] && c[it] != m;
  }
  uint64_t acc = 0;
  if (OP == OP_ICOUNT && a.other >= 0) {
    const uint64_t r = cand_row(a, (uint32_t)a.other, p);
    const uint64_t b = a.cand_off[a.other][r];
    const uint32_t n = (uint32_t)(a.cand_off[a.other][r + 1] - b);
#pragma unroll
    for (int it = 0; it < kItems; ++it)
      if (ok[it]) acc += count_against(a.cand[a.other] + b, n, c[it], a.other_op);
  } else {
#pragma unroll
    for (int it = 0; it < kItems; ++it) acc += ok[it] ? 1 : 0;
  }
#pragma unroll
  for (int it = 0; it < kItems; ++it) valid[it] = ok[it];
  if (OP == OP_IPATH) {
    // one atomic per wavefront: every valid element belongs to path p
    const uint64_t n_valid = wave_sum(acc);
    if (lane == 0 && n_valid) atomicAdd(path_counts + p, (unsigned long long)n_valid);
    return 0;
  }
  return acc;
}

// Map: count (OP_ICOUNT, OP_COUNT2, OP_IPATH) or mark valid elements
// (OP_IWRITE, OP_EXPAND, OP_FILTER). Element e of a chunk goes to thread
// (e mod 256) in iteration e / 256, and every lane verifies its own element
// with its own binary searches. With STAGE, an intersection workgroup also
// keeps one backward adjacency list that is bound to the same vertex across
// its chunk in LDS and searches it there. One launch covers the chunk
// window [c_lo, c_hi) of the n_chunks chunks.
template <int OP, bool STAGE = false>
__global__ __launch_bounds__(kBlock) void k_map(vsgm_level_args a, const uint64_t* __restrict__ seg_begin, const uint64_t* __restrict__ seg_off,
                                                uint64_t W, uint64_t n_chunks, uint64_t c_lo, uint64_t c_hi, uint64_t* __restrict__ mask,
                                                uint64_t* __restrict__ chunk_cnt, unsigned long long* __restrict__ total,
                                                unsigned long long* __restrict__ path_counts) {
  __shared__ int32_t s_rel[kPathCap + 1];
  __shared__ uint64_t s_tmp[kBlock / kWave];
  __shared__ ChunkTable s_tab;
  __shared__ uint32_t s_list[(STAGE && is_intersect(OP)) ? kStageCap : 1];
  __shared__ uint32_t s_want[2];
  constexpr bool kCompact = (OP == OP_IWRITE || OP == OP_EXPAND || OP == OP_FILTER);
  const uint64_t P = a.path_num;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  uint64_t acc = 0;
  // Groups of kGroup consecutive chunks are dealt to the workgroups
  // round-robin: per-element cost varies strongly along the element range
  // (hub-heavy paths come first), so round-robin keeps the workgroups
  // balanced, while a group's consecutive chunks usually continue one run of
  // paths and reuse the staged list.
  constexpr uint64_t kGroup = STAGE ? 16 : 1;
  uint32_t stage_v = ~0u, stage_n = 0;  // vertex whose adjacency list s_list holds
  for (uint64_t k0 = 0; block_chunk<kGroup>(k0, c_lo) < c_hi; k0 += kBlock) {
  fill_chunk_table<kGroup, false>(seg_off, P, W, n_chunks, k0, &s_tab, c_lo, c_hi);
  for (uint32_t t = 0; t < (uint32_t)kBlock; ++t) {
    const uint64_t chunk = block_chunk<kGroup>(k0 + t, c_lo);
    if (chunk >= c_hi) break;
    const uint64_t e0 = chunk * kChunk;
    const uint64_t e1 = min(e0 + (uint64_t)kChunk, W);
    ChunkMap cm;
    build_chunk_map(seg_off, s_tab.first[t], s_tab.last[t], e0, cm, s_rel);
    if (STAGE && is_intersect(OP)) {
      // stage the largest backward list (<= kStageCap entries) that is bound
      // to the same vertex at both ends of the chunk
      if (threadIdx.x == 0) {
        uint32_t best_v = ~0u, best_n = 0;
        for (uint32_t i = 0; i < a.n_back; ++i) {
          const uint32_t v = a.inst[a.back[i]][cm.first];
          if (v != a.inst[a.back[i]][cm.last]) continue;
          const uint32_t n = (uint32_t)(a.row_ptrs[v + 1] - a.row_ptrs[v]);
          if (n <= kStageCap && n > best_n) {
            best_v = v;
            best_n = n;
          }
        }
        s_want[0] = best_v;
        s_want[1] = best_n;
      }
      __syncthreads();
      const uint32_t want_v = s_want[0], want_n = s_want[1];
      if (want_v != ~0u && want_v != stage_v) {
        const uint64_t b = a.row_ptrs[want_v];
        for (uint32_t j = threadIdx.x; j < want_n; j += kBlock) s_list[j] = a.cols[b + j];
        __syncthreads();
        stage_v = want_v;
        stage_n = want_n;
      }
    }
    uint64_t valid_count = 0;
    if constexpr (is_intersect(OP) && !STAGE) {
      // wavefront w owns the contiguous run [e0 + 512 w, e0 + 512 (w + 1))
      bool valid[kItems];
      const uint64_t base = e0 + (uint64_t)wave * (kItems * kWave);
      acc += verify_run<OP>(a, seg_begin, seg_off, e0, base, e1, cm, s_rel, valid, path_counts);
      if (kCompact) {
#pragma unroll
        for (int it = 0; it < kItems; ++it) {
          const uint64_t bits = __ballot(valid[it]);
          const uint64_t word_first = base + (uint64_t)it * kWave;
          if (lane == 0 && word_first < W) mask[word_first / kWave] = bits;
          valid_count += valid[it] ? 1 : 0;
        }
      }
    } else
    for (int it = 0; it < kItems; ++it) {
      const uint64_t e = e0 + (uint64_t)it * kBlock + threadIdx.x;
      bool valid = false;
      if (e < e1) {
        uint64_t rank;
        const uint64_t p = chunk_path(e, e0, cm, s_rel, seg_off, rank);
        const uint64_t packed = seg_begin[p];
        const uint64_t tag = packed >> 56;
        const uint32_t c = seg_source<OP>(a, tag)[(packed & kBeginMask) + rank];
        const uint64_t contrib = eval_element<OP>(a, p, tag, c, valid, stage_v, s_list, stage_n);
        if (OP == OP_IPATH) {
          if (valid) atomicAdd(path_counts + p, 1ull);
        } else if (!kCompact) {
          acc += contrib;
        }
      }
      if (kCompact) {
        uint64_t bits = __ballot(valid);
        uint64_t word_first = e0 + (uint64_t)it * kBlock + (uint64_t)wave * kWave;
        if (lane == 0 && word_first < W) mask[word_first / kWave] = bits;
        valid_count += valid ? 1 : 0;
      }
    }
    if (kCompact) {
      uint64_t s = block_sum(valid_count, s_tmp);
      if (threadIdx.x == 0) chunk_cnt[chunk] = s;
    }
    __syncthreads();
  }
  }
  if (!kCompact && OP != OP_IPATH) {
    uint64_t s = block_sum(acc, s_tmp);
    if (threadIdx.x == 0 && s) atomicAdd(total, (unsigned long long)s);
  }
}

// Map kernel of the intersection and two-set count operators. Each wavefront
// walks a contiguous run of kItems * 64 elements. An iteration whose 64
// elements all lie in one path ("uniform", the common case on hub-heavy
// levels) makes the path wave-uniform, so its bound vertices, row_ptrs and
// segment come from scalar loads, and verifies the 64 ascending candidates
// against each other list with the LDS cursor tiles; other iterations verify
// lane by lane with branchless binary lifting.
template <int OP>
__global__ __launch_bounds__(kBlock) void k_map_isect(vsgm_level_args a, const uint64_t* __restrict__ seg_begin,
                                                      const uint64_t* __restrict__ seg_off, uint64_t W, uint64_t n_chunks, uint64_t c_lo,
                                                      uint64_t c_hi, uint64_t* __restrict__ mask, uint64_t* __restrict__ chunk_cnt,
                                                      unsigned long long* __restrict__ total, unsigned long long* __restrict__ path_counts) {
  __shared__ uint64_t s_off[kPathCap + 1];
  __shared__ uint64_t s_tmp[kBlock / kWave];
  __shared__ uint32_t s_tile[kBlock / kWave][kTile];
  __shared__ uint32_t s_cursor[kBlock / kWave][kCursorSlots];
  __shared__ ChunkPaths s_cp;
  constexpr bool kCompact = (OP == OP_IWRITE);
  const uint64_t P = a.path_num;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  uint64_t acc = 0;
  uint64_t cursor_path = ~0ull;  // path whose list positions s_cursor[wave] holds
  (void)n_chunks;
  for (uint64_t chunk = c_lo + blockIdx.x; chunk < c_hi; chunk += gridDim.x) {
    const uint64_t e0 = chunk * kChunk;
    const uint64_t e1 = min(e0 + (uint64_t)kChunk, W);
    ChunkPaths cp;
    stage_chunk(seg_off, P, e0, e1, cp, s_off, &s_cp);
    uint64_t valid_count = 0;
    const uint64_t wave_base = e0 + (uint64_t)wave * (kItems * kWave);
    for (int it = 0; it < kItems; ++it) {
      const uint64_t ef = wave_base + (uint64_t)it * kWave;  // first element of this iteration
      if (ef >= e1) break;
      const uint64_t e = ef + lane;
      const bool active = e < e1;
      const uint64_t ee = active ? e : ef;  // inactive lanes mirror the first element
      uint64_t p, start;
      bool uniform;
      if (cp.staged) {
        const uint32_t ns = cp.staged - 1;  // s_off[1..ns] are the path ends
        const uint32_t top = top_step(ns);
        const uint32_t i0 = lift_upper_bound(s_off + 1, ns, ef, top);
        uniform = (ef + kWave <= e1) && (ef + kWave <= s_off[i0 + 1]);
        const uint32_t i = uniform ? i0 : lift_upper_bound(s_off + 1, ns, ee, top);
        p = cp.first + i;
        start = s_off[i];
      } else {
        // more paths than fit in LDS (mostly empty segments): global search
        p = cp.first + last_le_u64(seg_off + cp.first, cp.last - cp.first + 2, ee);
        start = seg_off[p];
        uniform = __ballot(!(active && p == __shfl(p, 0))) == 0;
      }
      bool valid = false;
      uint64_t contrib = 0;
      if (uniform) {
        p = readfirstlane_u64(p);
        start = readfirstlane_u64(start);
        const uint64_t packed = seg_begin[p];
        const uint32_t tag = (uint32_t)(packed >> 56);
        const uint32_t c = seg_source<OP>(a, tag)[(packed & kBeginMask) + (e - start)];
        if (p != cursor_path) {
          cursor_path = p;
          if (lane < kCursorSlots) s_cursor[wave][lane] = 0;
          wave_sync();
        }
        contrib = eval_element_wave<OP>(a, p, tag, c, valid, s_tile[wave], s_cursor[wave]);
      } else {
        cursor_path = ~0ull;
        const uint64_t packed = seg_begin[p];
        const uint32_t tag = (uint32_t)(packed >> 56);
        const uint32_t c = seg_source<OP>(a, tag)[(packed & kBeginMask) + (ee - start)];
        contrib = eval_lanes<OP>(a, p, tag, c, active, valid);
      }
      valid = valid && active;
      if (OP == OP_IPATH) {
        if (valid) atomicAdd(path_counts + p, 1ull);
      } else if (!kCompact) {
        acc += active ? contrib : 0;
      }
      if (kCompact) {
        const uint64_t bits = __ballot(valid);
        if (lane == 0) mask[ef / kWave] = bits;
        valid_count += valid ? 1 : 0;
      }
    }
    if (kCompact) {
      uint64_t s = block_sum(valid_count, s_tmp);
      if (threadIdx.x == 0) chunk_cnt[chunk] = s;
    }
    __syncthreads();
  }
  if (!kCompact && OP != OP_IPATH) {
    uint64_t s = block_sum(acc, s_tmp);
    if (threadIdx.x == 0 && s) atomicAdd(total, (unsigned long long)s);
  }
}

// ---------------------------------------------------------------------------
// Intersection map with the stable list in LDS (isect_kernel 3).
//
// Paths arrive in materialization order, so a backward vertex bound before
// the last one stays the same over long runs of consecutive paths (the
// 4-cycle's last level: v1 is fixed while v2 walks v0's list). Probing that
// list through the membership index costs one fetched line per test. Here
// one 1024-thread workgroup per CU keeps that list in an LDS hash table and
// rebuilds it only when the run's vertex changes; candidates are still
// streamed coalesced from the pivot lists. Workgroups take tickets of
// consecutive chunks from a device counter (dynamic balance, and a ticket's
// chunks mostly continue one run). A chunk is 8192 elements, 512 per
// wavefront as in k_map.

constexpr int kHThreads = 1024;
constexpr int kHWaves = kHThreads / kWave;
constexpr uint64_t kHChunk = (uint64_t)kHThreads * kItems;
constexpr int kTicketCap = 64;  // chunks per ticket (upper bound of the tuning knob)

// Per chunk of the launch window: its first and last path, and the vertex
// whose list the map kernel should hold in LDS for it (~0u: none) with that
// list's length. The hashed list is the longest backward list of at most
// kHashCap entries that is bound to the same vertex at both ends of the chunk
// and is not streamed as the pivot at both ends.
struct ChunkPlan {
  uint64_t first;
  uint64_t last;
  uint32_t want_v;
  uint32_t want_n;
};

__global__ __launch_bounds__(kBlock) void k_chunk_plan(vsgm_level_args a, const uint64_t* __restrict__ seg_begin,
                                                       const uint64_t* __restrict__ seg_off, uint64_t W, uint64_t c_lo, uint64_t c_hi,
                                                       ChunkPlan* __restrict__ plan) {
  const uint64_t P = a.path_num;
  for (uint64_t k = c_lo + (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < c_hi; k += (uint64_t)gridDim.x * kBlock) {
    const uint64_t e0 = k * kHChunk, e1 = min(e0 + kHChunk, W);
    ChunkPlan cp;
    cp.first = last_le_u64(seg_off, P + 1, e0);
    cp.last = last_le_u64(seg_off, P + 1, e1 - 1);
    cp.want_v = ~0u;
    cp.want_n = 0;
    if (a.n_back > 1) {
      const uint32_t t0 = (uint32_t)(seg_begin[cp.first] >> 56), t1 = (uint32_t)(seg_begin[cp.last] >> 56);
      for (uint32_t i = 0; i < a.n_back; ++i) {
        if (i == t0 && i == t1) continue;
        const uint32_t v = a.inst[a.back[i]][cp.first];
        if (v != a.inst[a.back[i]][cp.last]) continue;
        const uint64_t n = a.row_ptrs[v + 1] - a.row_ptrs[v];
        if (n <= kHashCap && (cp.want_v == ~0u || n > cp.want_n)) {
          cp.want_v = v;
          cp.want_n = (uint32_t)n;
        }
      }
    }
    plan[k - c_lo] = cp;
  }
}

// Insert the n entries of list into the cuckoo tables t0 / t1 of m slots
// each (every thread of the workgroup takes part; barriers on both sides).
// An insertion that is still evicting after kCuckooRounds steps leaves its
// key out and sets *fail; the tables are then not used.
__device__ inline bool cuckoo_build(uint32_t* t0, uint32_t* t1, uint32_t m, const uint32_t* __restrict__ list, uint32_t n, uint32_t* fail) {
  for (uint32_t i = threadIdx.x; i < m; i += kHThreads) {
    t0[i] = kEmptySlot;
    t1[i] = kEmptySlot;
  }
  if (threadIdx.x == 0) *fail = 0;
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < n; j += kHThreads) {
    uint32_t x = list[j];
    uint32_t t = 0;
    bool placed = false;
    for (int r = 0; r < kCuckooRounds && !placed; ++r) {
      uint32_t* slot = t == 0 ? &t0[cuckoo_h0(x, m)] : &t1[cuckoo_h1(x, m)];
      x = atomicExch(slot, x);  // the key evicted from this table moves to its slot in the other one
      placed = x == kEmptySlot;
      t ^= 1u;
    }
    if (!placed) *fail = 1;
  }
  __syncthreads();
  return *fail == 0;
}

// Columns a level's verification reads per path: every backward vertex and
// every '!=' operand, each staged once (stage = 0: more distinct columns than
// kStageCols, never staged).
constexpr uint32_t kStageCols = VSGM_MAX_QV - 1;

struct StageCols {
  uint32_t stage;
  uint32_t n;
  uint32_t col[kStageCols];       // pattern vertex of staged column k
  uint32_t back_col[VSGM_MAX_QV]; // staged column of backward position i
  uint32_t n_ne;
  uint32_t ne_col[VSGM_MAX_QV];   // staged column of the k-th '!=' condition
};

// Per-wavefront window of the paths it is walking: end offset of each path
// relative to the chunk start, packed segment begin, and the staged columns.
constexpr uint32_t kWin = kWave;
constexpr int kWinItems = 4;  // elements per lane per step (4: no spills at 128 VGPRs)
struct PathWindow {
  int32_t end[kWin];
  int32_t start0;  // start of the window's first path, relative to the chunk
  uint32_t n;      // paths in the window
  uint64_t beg[kWin];
  uint32_t col[kStageCols][kWin];
};

// Verification of up to 512 elements [e, wend) of a chunk starting at e0,
// all in the wave's current window (paths first + i, i < win.n): element e
// + it * 64 + lane per lane. One global load (the candidate) per element;
// the path, its segment and bound vertices come from the window, a list
// held in the LDS table is probed there, other lists through the membership
// index or a binary search.
template <int OP, int NI>
__device__ inline uint64_t verify_window(const vsgm_level_args& a, const StageCols& sc, uint64_t first, uint64_t e0, uint64_t e,
                                         uint64_t wend, const PathWindow& win, const HashView& hs, unsigned long long* __restrict__ path_counts) {
  const int lane = threadIdx.x & (kWave - 1);
  uint32_t pi[NI], c[NI], piv[NI];
  bool ok[NI];
  uint32_t lo = 0;
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const uint64_t x = e + (uint64_t)it * kWave + lane;
    ok[it] = x < wend;
    const int32_t j = (int32_t)(x - e0);
    // first window path whose end lies beyond j; items ascend, so start at the previous item's path
    uint32_t l = lo, h = win.n - 1;
    if (ok[it])
      while (l < h) {
        const uint32_t mid = (l + h) >> 1;
        if (win.end[mid] > j) h = mid;
        else l = mid + 1;
      }
    lo = l;
    pi[it] = l;
    const uint64_t packed = win.beg[l];
    piv[it] = (uint32_t)(packed >> 56);
    const int32_t st = l ? win.end[l - 1] : win.start0;
    c[it] = ok[it] ? a.cols[(packed & kBeginMask) + (uint64_t)(j - st)] : 0u;
  }
  for (uint32_t b = 0; b < a.n_back; ++b) {
    const uint32_t* colv = win.col[sc.back_col[b]];
    uint32_t v[NI], sl[NI], w[NI];
#pragma unroll
    for (int it = 0; it < NI; ++it) {
      v[it] = (ok[it] && piv[it] != b) ? colv[pi[it]] : ~0u;
      if (v[it] != ~0u && v[it] == hs.v) {
        ok[it] = hash_contains(hs, c[it]);
        v[it] = ~0u;
      }
    }
    if (a.hub_slot) {
#pragma unroll
      for (int it = 0; it < NI; ++it) sl[it] = v[it] != ~0u ? a.hub_slot[v[it]] : ~0u;
#pragma unroll
      for (int it = 0; it < NI; ++it) w[it] = sl[it] != ~0u ? a.hub_bits[(uint64_t)sl[it] * a.hub_words + (c[it] >> 5)] : 0u;
    } else {
#pragma unroll
      for (int it = 0; it < NI; ++it) sl[it] = ~0u, w[it] = 0u;
    }
#pragma unroll
    for (int it = 0; it < NI; ++it) {
      if (v[it] == ~0u) continue;
      if (sl[it] != ~0u) {
        ok[it] = (w[it] >> (c[it] & 31)) & 1u;
      } else {
        const uint64_t rb = a.row_ptrs[v[it]];
        const uint32_t n = (uint32_t)(a.row_ptrs[v[it] + 1] - rb);
        const uint32_t pos = lower_bound_u32(a.cols + rb, n, c[it]);
        ok[it] = pos < n && a.cols[rb + pos] == c[it];
      }
    }
  }
  for (uint32_t k = 0; k < sc.n_ne; ++k) {
    const uint32_t* colm = win.col[sc.ne_col[k]];
#pragma unroll
    for (int it = 0; it < NI; ++it) ok[it] = ok[it] && c[it] != colm[pi[it]];
  }
  uint64_t acc = 0;
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    if (!ok[it]) continue;
    if (OP == OP_IPATH) {
      atomicAdd(path_counts + first + pi[it], 1ull);
    } else if (a.other >= 0) {
      const uint64_t r = cand_row(a, (uint32_t)a.other, first + pi[it]);
      const uint64_t cb = a.cand_off[a.other][r];
      const uint32_t n = (uint32_t)(a.cand_off[a.other][r + 1] - cb);
      acc += count_against(a.cand[a.other] + cb, n, c[it], a.other_op);
    } else {
      acc += 1;
    }
  }
  return acc;
}

// One wavefront walks chunk [e0, e1) (paths first..last) alone: it loads a
// window of up to 64 consecutive paths (one lane per path, coalesced), then
// verifies the window's elements 512 at a time, then slides on.
template <int OP>
__device__ inline uint64_t walk_chunk(const vsgm_level_args& a, const StageCols& sc, const uint64_t* __restrict__ seg_begin,
                                      const uint64_t* __restrict__ seg_off, uint64_t e0, uint64_t e1, uint64_t first, uint64_t last,
                                      PathWindow& win, const HashView& hs, unsigned long long* __restrict__ path_counts) {
  const int lane = threadIdx.x & (kWave - 1);
  uint64_t acc = 0;
  uint64_t p = first;  // first path of the window
  uint64_t e = e0;     // next element to verify
  while (e < e1 && p <= last) {
    const uint32_t n = (uint32_t)min<uint64_t>(kWin, last - p + 1);
    if ((uint32_t)lane < n) {
      const uint64_t q = p + lane;
      win.end[lane] = (int32_t)((int64_t)seg_off[q + 1] - (int64_t)e0);
      win.beg[lane] = seg_begin[q];
      for (uint32_t k = 0; k < sc.n; ++k) win.col[k][lane] = a.inst[sc.col[k]][q];
      if (lane == 0) {
        win.start0 = (int32_t)((int64_t)seg_off[q] - (int64_t)e0);
        win.n = n;
      }
    }
    wave_sync();
    const uint64_t wend = min(e1, e0 + (uint64_t)(int64_t)win.end[n - 1]);
    for (; e < wend; e += (uint64_t)kWinItems * kWave) acc += verify_window<OP, kWinItems>(a, sc, p, e0, e, wend, win, hs, path_counts);
    e = wend;
    p += n;
    wave_sync();  // the window is rewritten next
  }
  return acc;
}

// Register window: lane l of the wavefront holds path p + l's segment start
// and end, packed segment begin and up to kRegCols staged columns. A path
// with at least kUniTail elements left is verified in wave-uniform steps
// (its data in scalar registers, up to kUniLarge candidate loads in flight
// per lane, a masked step for its tail). Shorter stretches go 64-element
// block by block: each block resolves its lanes' paths with a uniform loop
// over the path boundaries inside it (readlane + select, no LDS). Either
// way an element costs its candidate load, ALU work and its probes: one
// pair of LDS reads when the tested list is the held one.
constexpr uint32_t kRegCols = 3;
constexpr int kRegItems = 4;  // elements per lane per step (4: no spills)
constexpr int kUniLarge = 16;  // loads per lane of a wave-uniform step over a long path
constexpr int kUniSmall = 8;
constexpr uint64_t kUniTail = 64;  // path elements from which a masked uniform step beats the per-lane path walk

__device__ inline uint64_t readlane_u64(uint64_t v, uint32_t l) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

__device__ inline uint32_t pick3(uint32_t i, uint32_t c0, uint32_t c1, uint32_t c2) { return i == 0 ? c0 : (i == 1 ? c1 : c2); }

// Wave-uniform steps over one path: while at least UI * 64 of its elements
// remain, the path's segment, bound vertices and probe targets are uniform
// (scalar registers) and each lane keeps UI candidate loads in flight.
// Verifies elements [x, x + k * UI * 64) for the largest such k; with TAIL,
// one more step over the rest [x, pend) (at most UI * 64 elements, lanes
// past pend idle). Returns the count contribution and advances x.
template <int OP, int UI, int NCOL, bool TAIL = false>
__device__ inline uint64_t uniform_steps(const vsgm_level_args& a, const StageCols& sc, uint64_t q, uint64_t st, uint64_t packed, uint32_t k0,
                                         uint32_t k1, uint32_t k2, uint64_t pend, uint64_t& x, const HashView& hs,
                                         unsigned long long* __restrict__ path_counts) {
  constexpr uint64_t kStep = (uint64_t)UI * kWave;
  const uint32_t lane = threadIdx.x & (kWave - 1);
  if (TAIL ? (pend <= x) : (pend - x < kStep)) return 0;
  const uint32_t piv = (uint32_t)(packed >> 56);
  const uint32_t* seg = a.cols + (packed & kBeginMask) - st;  // element y's candidate is seg[y]
  // probe plan of the tested lists (wave-uniform): 0 held in LDS, 1 index row, 2 binary search
  // (a level of this kernel has at most kRegCols backward lists: they are
  // distinct staged columns; loops run to the constant bound so the plan
  // stays in registers)
  uint32_t mode[kRegCols];
  const uint32_t* tgt[kRegCols];
  uint32_t tlen[kRegCols];
#pragma unroll
  for (uint32_t b = 0; b < kRegCols; ++b) {
    mode[b] = 3;  // the pivot, or no list: not tested
    tgt[b] = nullptr;
    tlen[b] = 0;
    if (b >= a.n_back || b == piv) continue;
    const uint32_t v = pick3(sc.back_col[b], k0, k1, k2);
    if (v == hs.v) {
      mode[b] = 0;
      continue;
    }
    const uint32_t sl = a.hub_slot ? a.hub_slot[v] : ~0u;
    if (sl != ~0u) {
      mode[b] = 1;
      tgt[b] = a.hub_bits + (uint64_t)sl * a.hub_words;
    } else {
      mode[b] = 2;
      const uint64_t rb = a.row_ptrs[v];
      tgt[b] = a.cols + rb;
      tlen[b] = (uint32_t)(a.row_ptrs[v + 1] - rb);
    }
  }
  uint64_t acc = 0;
  for (; TAIL ? (x < pend) : (pend - x >= kStep); x += kStep) {
    uint32_t c[UI];
    bool ok[UI];
#pragma unroll
    for (int it = 0; it < UI; ++it) {
      const uint64_t y = x + (uint64_t)it * kWave + lane;
      ok[it] = !TAIL || y < pend;
      c[it] = ok[it] ? seg[y] : 0u;
    }
#pragma unroll
    for (uint32_t b = 0; b < kRegCols; ++b) {
      if (mode[b] == 0) {
#pragma unroll
        for (int it = 0; it < UI; ++it) ok[it] = ok[it] && hash_contains(hs, c[it]);
      } else if (mode[b] == 1) {
        uint32_t w[UI];
#pragma unroll
        for (int it = 0; it < UI; ++it) w[it] = tgt[b][c[it] >> 5];
#pragma unroll
        for (int it = 0; it < UI; ++it) ok[it] = ok[it] && ((w[it] >> (c[it] & 31)) & 1u);
      } else if (mode[b] == 2) {
#pragma unroll
        for (int it = 0; it < UI; ++it) {
          if (!ok[it]) continue;
          const uint32_t pos = lower_bound_u32(tgt[b], tlen[b], c[it]);
          ok[it] = pos < tlen[b] && tgt[b][pos] == c[it];
        }
      }
    }
#pragma unroll
    for (uint32_t k = 0; k < kRegCols; ++k) {
      if (k >= sc.n_ne) break;
      const uint32_t m = pick3(sc.ne_col[k], k0, k1, k2);
#pragma unroll
      for (int it = 0; it < UI; ++it) ok[it] = ok[it] && c[it] != m;
    }
    uint64_t valid = 0;
#pragma unroll
    for (int it = 0; it < UI; ++it) {
      if (!ok[it]) continue;
      if (OP != OP_IPATH && a.other >= 0) {
        const uint64_t r = cand_row(a, (uint32_t)a.other, q);
        const uint64_t cb = a.cand_off[a.other][r];
        const uint32_t nn = (uint32_t)(a.cand_off[a.other][r + 1] - cb);
        acc += count_against(a.cand[a.other] + cb, nn, c[it], a.other_op);
      } else {
        valid += 1;
      }
    }
    if (OP == OP_IPATH) {
      const uint64_t n_valid = wave_sum(valid);
      if (lane == 0 && n_valid) atomicAdd(path_counts + q, (unsigned long long)n_valid);
    } else {
      acc += valid;
    }
    if (TAIL) {
      x = pend;
      break;
    }
  }
  return acc;
}

template <int OP, int NI, int NCOL>
__device__ inline uint64_t walk_chunk_reg(const vsgm_level_args& a, const StageCols& sc, const uint64_t* __restrict__ seg_begin,
                                          const uint64_t* __restrict__ seg_off, uint64_t e0, uint64_t e1, uint64_t first, uint64_t last,
                                          const HashView& hs, unsigned long long* __restrict__ path_counts) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  uint64_t acc = 0;
  uint64_t p = first;  // first path of the window
  uint64_t x = e0;     // next element to verify
  while (x < e1 && p <= last) {
    const uint32_t n = (uint32_t)min<uint64_t>(kWin, last - p + 1);
    const bool in = lane < n;
    const uint64_t q = p + lane;
    const uint64_t w_start = in ? seg_off[q] : ~0ull;
    const uint64_t w_end = in ? seg_off[q + 1] : ~0ull;
    const uint64_t w_beg = in ? seg_begin[q] : 0;
    const uint32_t w_c0 = (in && NCOL > 0) ? a.inst[sc.col[0]][q] : 0u;
    const uint32_t w_c1 = (in && NCOL > 1) ? a.inst[sc.col[1]][q] : 0u;
    const uint32_t w_c2 = (in && NCOL > 2) ? a.inst[sc.col[2]][q] : 0u;
    const uint64_t wend = min(e1, readlane_u64(w_end, n - 1));
    uint32_t ci = 0;  // window path containing the current block's first element
    while (x < wend) {
      // long stretch of one path: wave-uniform steps
      while (readlane_u64(w_end, ci) <= x) ++ci;
      {
        const uint64_t pend = min(readlane_u64(w_end, ci), wend);
        if (pend - x >= kUniTail) {
          const uint64_t st = readlane_u64(w_start, ci), packed = readlane_u64(w_beg, ci);
          const uint32_t k0 = NCOL > 0 ? __builtin_amdgcn_readlane(w_c0, ci) : 0u;
          const uint32_t k1 = NCOL > 1 ? __builtin_amdgcn_readlane(w_c1, ci) : 0u;
          const uint32_t k2 = NCOL > 2 ? __builtin_amdgcn_readlane(w_c2, ci) : 0u;
          acc += uniform_steps<OP, kUniLarge, NCOL>(a, sc, p + ci, st, packed, k0, k1, k2, pend, x, hs, path_counts);
          acc += uniform_steps<OP, kUniSmall, NCOL>(a, sc, p + ci, st, packed, k0, k1, k2, pend, x, hs, path_counts);
          if (pend - x >= kUniTail) acc += uniform_steps<OP, kUniSmall, NCOL, true>(a, sc, p + ci, st, packed, k0, k1, k2, pend, x, hs, path_counts);
          continue;
        }
      }
      uint32_t c[NI], piv[NI], c0[NI], c1[NI], c2[NI], pk[NI];
      bool ok[NI];
#pragma unroll
      for (int it = 0; it < NI; ++it) {
        const uint64_t B = x + (uint64_t)it * kWave;
        const uint64_t xe = B + lane;
        ok[it] = xe < wend;
        if (B >= wend) {
          c[it] = piv[it] = c0[it] = c1[it] = c2[it] = pk[it] = 0;
          continue;
        }
        while (readlane_u64(w_end, ci) <= B) ++ci;
        uint64_t beg = readlane_u64(w_beg, ci), st = readlane_u64(w_start, ci);
        uint32_t k0 = NCOL > 0 ? __builtin_amdgcn_readlane(w_c0, ci) : 0u;
        uint32_t k1 = NCOL > 1 ? __builtin_amdgcn_readlane(w_c1, ci) : 0u;
        uint32_t k2 = NCOL > 2 ? __builtin_amdgcn_readlane(w_c2, ci) : 0u;
        uint32_t pi = ci;
        for (uint32_t l = ci + 1; l < n; ++l) {
          const uint64_t sl = readlane_u64(w_start, l);
          if (sl >= B + kWave || sl >= wend) break;
          if (xe >= sl) {
            st = sl;
            beg = readlane_u64(w_beg, l);
            if (NCOL > 0) k0 = __builtin_amdgcn_readlane(w_c0, l);
            if (NCOL > 1) k1 = __builtin_amdgcn_readlane(w_c1, l);
            if (NCOL > 2) k2 = __builtin_amdgcn_readlane(w_c2, l);
            pi = l;
          }
        }
        piv[it] = (uint32_t)(beg >> 56);
        c0[it] = k0;
        c1[it] = NCOL > 1 ? k1 : 0u;
        c2[it] = NCOL > 2 ? k2 : 0u;
        pk[it] = pi;
        c[it] = ok[it] ? a.cols[(beg & kBeginMask) + (xe - st)] : 0u;
      }
      for (uint32_t b = 0; b < a.n_back; ++b) {
        const uint32_t bc = sc.back_col[b];
        uint32_t v[NI], sl[NI], w[NI];
        bool held[NI];
        // index slots depend only on the path: issued before the candidates are used
#pragma unroll
        for (int it = 0; it < NI; ++it) {
          v[it] = (ok[it] && piv[it] != b) ? pick3(bc, c0[it], c1[it], c2[it]) : ~0u;
          held[it] = v[it] != ~0u && v[it] == hs.v;
          sl[it] = (a.hub_slot && v[it] != ~0u && !held[it]) ? a.hub_slot[v[it]] : ~0u;
        }
#pragma unroll
        for (int it = 0; it < NI; ++it) {
          if (held[it]) {
            ok[it] = hash_contains(hs, c[it]);
            v[it] = ~0u;
          }
        }
#pragma unroll
        for (int it = 0; it < NI; ++it) w[it] = sl[it] != ~0u ? a.hub_bits[(uint64_t)sl[it] * a.hub_words + (c[it] >> 5)] : 0u;
#pragma unroll
        for (int it = 0; it < NI; ++it) {
          if (v[it] == ~0u) continue;
          if (sl[it] != ~0u) {
            ok[it] = (w[it] >> (c[it] & 31)) & 1u;
          } else {
            const uint64_t rb = a.row_ptrs[v[it]];
            const uint32_t nn = (uint32_t)(a.row_ptrs[v[it] + 1] - rb);
            const uint32_t pos = lower_bound_u32(a.cols + rb, nn, c[it]);
            ok[it] = pos < nn && a.cols[rb + pos] == c[it];
          }
        }
      }
      for (uint32_t k = 0; k < sc.n_ne; ++k) {
        const uint32_t nc = sc.ne_col[k];
#pragma unroll
        for (int it = 0; it < NI; ++it) ok[it] = ok[it] && c[it] != pick3(nc, c0[it], c1[it], c2[it]);
      }
#pragma unroll
      for (int it = 0; it < NI; ++it) {
        if (!ok[it]) continue;
        if (OP == OP_IPATH) {
          atomicAdd(path_counts + p + pk[it], 1ull);
        } else if (a.other >= 0) {
          const uint64_t r = cand_row(a, (uint32_t)a.other, p + pk[it]);
          const uint64_t cb = a.cand_off[a.other][r];
          const uint32_t nn = (uint32_t)(a.cand_off[a.other][r + 1] - cb);
          acc += count_against(a.cand[a.other] + cb, nn, c[it], a.other_op);
        } else {
          acc += 1;
        }
      }
      x += (uint64_t)NI * kWave;
    }
    x = wend;
    p += n;
  }
  return acc;
}

// NCOL: the level's distinct staged columns (1..kRegCols: register window;
// 0: more columns, LDS window).
template <int OP, int NCOL>
__global__ __launch_bounds__(kHThreads) void k_map_hash(vsgm_level_args a, StageCols sc, const uint64_t* __restrict__ seg_begin,
                                                        const uint64_t* __restrict__ seg_off, uint64_t W, uint64_t c_lo, uint64_t c_hi,
                                                        const ChunkPlan* __restrict__ plan, uint32_t ticket_chunks,
                                                        unsigned long long* __restrict__ tickets, unsigned long long* __restrict__ total,
                                                        unsigned long long* __restrict__ path_counts) {
  __shared__ uint32_t s_cuckoo[2 * kCuckooSlots];
  __shared__ uint32_t s_fail;
  __shared__ PathWindow s_win[NCOL ? 1 : kHWaves];
  __shared__ ChunkPlan s_plan[kTicketCap];
  __shared__ uint32_t s_votes[kTicketCap];
  __shared__ uint64_t s_tmp[kHWaves];
  __shared__ uint64_t s_ticket;
  __shared__ uint32_t s_next;
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t n_tickets = (c_hi - c_lo + ticket_chunks - 1) / ticket_chunks;
  uint32_t hv = ~0u, hm = 1;  // vertex whose list the cuckoo tables hold, slots in use per table
  uint64_t acc = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      s_ticket = atomicAdd(tickets, 1ull);
      s_next = 0;
    }
    __syncthreads();
    const uint64_t t = s_ticket;
    if (t >= n_tickets) break;
    const uint64_t k_lo = c_lo + t * ticket_chunks, k_hi = min(k_lo + ticket_chunks, c_hi);
    const uint32_t nk = (uint32_t)(k_hi - k_lo);
    if (threadIdx.x < nk) s_plan[threadIdx.x] = plan[k_lo - c_lo + threadIdx.x];
    __syncthreads();
    // the list to hold for this ticket: the wanted vertex of the most chunks
    if (threadIdx.x < nk) {
      const uint32_t v = s_plan[threadIdx.x].want_v;
      uint32_t votes = 0;
      if (v != ~0u)
        for (uint32_t i = 0; i < nk; ++i) votes += s_plan[i].want_v == v ? 1u : 0u;
      s_votes[threadIdx.x] = votes;
    }
    __syncthreads();
    uint32_t best = 0;
    for (uint32_t i = 1; i < nk; ++i)
      if (s_votes[i] > s_votes[best]) best = i;
    const uint32_t want = s_votes[best] ? s_plan[best].want_v : ~0u;
    if (want != ~0u && want != hv) {
      const uint32_t n = s_plan[best].want_n;
      hm = min(kCuckooSlots, max(64u, n + n / 8 + 64));  // load factor n / 2m <= 0.45 for n <= kHashCap
      hv = cuckoo_build(s_cuckoo, s_cuckoo + kCuckooSlots, hm, a.cols + a.row_ptrs[want], n, &s_fail) ? want : ~0u;
    }
    const HashView hs{s_cuckoo, s_cuckoo + kCuckooSlots, hm, hv};
    // waves take the ticket's chunks one at a time and walk them alone
    for (;;) {
      uint32_t j = 0;
      if (lane == 0) j = atomicAdd(&s_next, 1u);
      j = __shfl(j, 0);
      if (j >= nk) break;
      const ChunkPlan cp = s_plan[j];
      const uint64_t k = k_lo + j;
      const uint64_t e0 = k * kHChunk, e1 = min(e0 + kHChunk, W);
      if constexpr (NCOL > 0)
        acc += walk_chunk_reg<OP, kRegItems, NCOL>(a, sc, seg_begin, seg_off, e0, e1, cp.first, cp.last, hs, path_counts);
      else
        acc += walk_chunk<OP>(a, sc, seg_begin, seg_off, e0, e1, cp.first, cp.last, s_win[wave], hs, path_counts);
    }
    __syncthreads();  // every wave is done with s_plan and the table
  }
  if (OP != OP_IPATH) {
    acc = wave_sum(acc);
    if (lane == 0) s_tmp[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint64_t sum = 0;
      for (int w = 0; w < kHWaves; ++w) sum += s_tmp[w];
      if (sum) atomicAdd(total, (unsigned long long)sum);
    }
  }
}

// ---------------------------------------------------------------------------
// Flat hashed map (isect_kernel 4): COMPUTE_COUNT with no second computed
// set and two or three backward lists.
//
// Same tickets, chunk plan and held list as k_map_hash, but a wavefront
// walks its chunk in blocks of kFlatRows x 64 consecutive elements that may
// span several paths:
// * Each path's probe plan (per backward position: held in the LDS table,
//   membership-index row, binary search, or the pivot itself) and its
//   segment base are resolved when a window of 64 paths is loaded, one lane
//   per path, so a path costs no dependent load of its own.
// * A block inside one path reads everything uniform (scalar registers) and
//   loads its candidates with immediate offsets from one address; a block
//   over several paths finds each element's path by a uniform loop over the
//   block's path boundaries and gathers the path data across lanes.
// * The table hashes are 24-bit multiplicative (v_mul_u32_u24, full rate)
//   into power-of-two tables; hits are counted by ballot and popcount.

constexpr uint32_t kFlatSlots = 16384;  // slots per cuckoo table: two tables = 128 KiB of LDS
constexpr int kFlatRows = 12;           // rows of 64 elements per block (two buffers in flight)
constexpr uint64_t kFlatBlock = (uint64_t)kFlatRows * kWave;
constexpr uint32_t kFlatK0 = 0x9E3779u;  // odd 24-bit multipliers
constexpr uint32_t kFlatK1 = 0xC2B2AFu;
constexpr uint32_t kModeHeld = 0, kModeIndex = 1, kModeSearch = 2, kModePivot = 3;

// Slot of c in table 0 / 1 of 2^(32 - sh) slots. Table 0 hashes the low 24
// bits of c, table 1 the low 24 bits of c rotated by 16; two keys that agree
// in both are equal. Both operands of each product fit in 24 bits, so it
// compiles to the full-rate v_mul_u32_u24 (the inputs are masked in plain
// C: the 24-bit multiply builtins leave wider inputs unspecified, and the
// build and the lookups must hash alike).
__device__ inline uint32_t flat_h0(uint32_t c, uint32_t sh) { return ((c & 0xFFFFFFu) * kFlatK0) >> sh; }
__device__ inline uint32_t flat_h1(uint32_t c, uint32_t sh) { return ((__funnelshift_r(c, c, 16) & 0xFFFFFFu) * kFlatK1) >> sh; }

struct FlatTable {
  const uint32_t* t0;
  const uint32_t* t1;
  uint32_t sh;
  uint32_t v;  // vertex whose list the tables hold (~0u: none)
};

__device__ inline bool flat_contains(const FlatTable& h, uint32_t c) {
  return h.t0[flat_h0(c, h.sh)] == c || h.t1[flat_h1(c, h.sh)] == c;
}

// Cuckoo build of the n entries of list into t0 / t1 of 2^(32 - sh) slots
// each, by every thread of the workgroup (barriers on both sides); false
// when an insertion is still evicting after kCuckooRounds steps.
__device__ inline bool flat_build(uint32_t* t0, uint32_t* t1, uint32_t sh, const uint32_t* __restrict__ list, uint32_t n, uint32_t* fail) {
  const uint32_t m = 1u << (32 - sh);
  for (uint32_t i = threadIdx.x; i < m; i += kHThreads) {
    t0[i] = kEmptySlot;
    t1[i] = kEmptySlot;
  }
  if (threadIdx.x == 0) *fail = 0;
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < n; j += kHThreads) {
    uint32_t x = list[j];
    uint32_t t = 0;
    bool placed = false;
    for (int r = 0; r < kCuckooRounds && !placed; ++r) {
      uint32_t* slot = t == 0 ? &t0[flat_h0(x, sh)] : &t1[flat_h1(x, sh)];
      x = atomicExch(slot, x);
      placed = x == kEmptySlot;
      t ^= 1u;
    }
    if (!placed) *fail = 1;
  }
  __syncthreads();
  return *fail == 0;
}

// NE operands of a level that the flat map checks (pattern vertices), and
// the backward vertices.
struct FlatCols {
  uint32_t n_ne;
  uint32_t ne[VSGM_MAX_QV];
};

__device__ inline uint64_t shfl_u64(uint64_t v, uint32_t l) {
  const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, (int)l);
  const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), (int)l);
  return ((uint64_t)hi << 32) | lo;
}

// Membership of c in a tested list under one probe mode (kModeHeld: the
// table; kModeIndex: bit c of the index row at word offset t; kModeSearch:
// binary search of the n entries at cols + t).
__device__ inline bool flat_test(const vsgm_level_args& a, const FlatTable& h, uint32_t mode, uint64_t t, uint32_t n, uint32_t c) {
  if (mode == kModeHeld) return flat_contains(h, c);
  if (mode == kModeIndex) return (a.hub_bits[t + (c >> 5)] >> (c & 31)) & 1u;
  const uint32_t pos = lower_bound_u32(a.cols + t, n, c);
  return pos < n && a.cols[t + pos] == c;
}

// Rows of candidates: c[it] is the candidate of element it * 64 + lane of
// the block, or kNoCand where there is none or a test has failed (no vertex
// id equals kNoCand: views have fewer than kNoCand vertices).
constexpr uint32_t kNoCand = 0xFFFFFFFEu;

// Tests of the rows c[] against one backward position under a probe plan
// that is the same for every element of the block; failing rows become
// kNoCand.
template <int R>
__device__ inline void flat_rows_test(const vsgm_level_args& a, const FlatTable& h, uint32_t mode, uint64_t t, uint32_t n, uint32_t* c) {
  if (mode == kModeHeld) {
    // both tables' reads of every row are issued before any is compared, so
    // the rows' LDS latencies overlap (a short-circuit per row would wait
    // for each row's first read before issuing its second and the next row)
    uint32_t s0[R], s1[R];
#pragma unroll
    for (int it = 0; it < R; ++it) s0[it] = h.t0[flat_h0(c[it], h.sh)];
#pragma unroll
    for (int it = 0; it < R; ++it) s1[it] = h.t1[flat_h1(c[it], h.sh)];
#pragma unroll
    for (int it = 0; it < R; ++it) c[it] = ((s0[it] == c[it]) | (s1[it] == c[it])) ? c[it] : kNoCand;
  } else if (mode == kModeIndex) {
    // every row's word is loaded (a row without a candidate reads word 0 and
    // stays kNoCand), so the loads issue back to back without branches
    const uint32_t* row = a.hub_bits + t;
    uint32_t w[R];
#pragma unroll
    for (int it = 0; it < R; ++it) w[it] = row[(c[it] != kNoCand ? c[it] : 0u) >> 5];
#pragma unroll
    for (int it = 0; it < R; ++it) c[it] = ((w[it] >> (c[it] & 31)) & 1u) ? c[it] : kNoCand;
  } else if (mode == kModeSearch) {
    // one list for all rows: binary lifting, every row's probe of a step in
    // flight together. Probe indices are clamped into the list, so each
    // step's loads are unconditional and issue back to back; a probe past
    // the end is loaded but not taken. n is the same for every lane.
    const uint32_t* L = a.cols + t;
    if (n == 0) {
#pragma unroll
      for (int it = 0; it < R; ++it) c[it] = kNoCand;
      return;
    }
    uint32_t pos[R];
#pragma unroll
    for (int it = 0; it < R; ++it) pos[it] = 0;
    for (uint32_t step = top_step(n); step; step >>= 1) {
      uint32_t v[R];
#pragma unroll
      for (int it = 0; it < R; ++it) v[it] = L[min(pos[it] + step, n) - 1];
#pragma unroll
      for (int it = 0; it < R; ++it) {
        const uint32_t probe = pos[it] + step;
        pos[it] = (probe <= n && v[it] < c[it]) ? probe : pos[it];
      }
    }
    uint32_t f[R];
#pragma unroll
    for (int it = 0; it < R; ++it) f[it] = L[min(pos[it], n - 1)];
#pragma unroll
    for (int it = 0; it < R; ++it) c[it] = (pos[it] < n && f[it] == c[it]) ? c[it] : kNoCand;
  }
}

// A block of one probe plan whose candidate loads are in flight (every
// field wave-uniform).
template <int NB, int NNE>
struct FlatBlock {
  uint32_t md;
  uint64_t t[NB];
  uint32_t nn[NB];
  uint32_t ne[NNE > 0 ? NNE : 1];
};

// Tests and count of a block whose candidates c[] have been loaded.
template <int NB, int NNE>
__device__ inline uint64_t flat_finish(const vsgm_level_args& a, const FlatTable& h, const FlatBlock<NB, NNE>& bk, uint32_t* c) {
#pragma unroll
  for (int k = 0; k < NNE; ++k)
#pragma unroll
    for (int it = 0; it < kFlatRows; ++it) c[it] = c[it] != bk.ne[k] ? c[it] : kNoCand;
#pragma unroll
  for (int b = 0; b < NB; ++b) flat_rows_test<kFlatRows>(a, h, (bk.md >> (2 * b)) & 3u, bk.t[b], bk.nn[b], c);
  uint64_t n = 0;
#pragma unroll
  for (int it = 0; it < kFlatRows; ++it) n += (uint64_t)__popcll(__ballot(c[it] != kNoCand));
  return n;
}

// Path window of a wavefront: lane l describes path p + l (l < n).
template <int NB, int NNE>
struct FlatWindow {
  uint32_t n;
  uint64_t st, en, base;  // segment start / end (elements), cols index of element e: base + e
  uint32_t md;            // probe mode of each backward position (2 bits each)
  uint64_t t[NB];
  uint32_t nn[NB];
  uint32_t ne[NNE > 0 ? NNE : 1];
};

// One block [x, bend) of the window: with one probe plan and uniform '!='
// operands its loads are issued into c and its plan written to bk (returns
// true); otherwise it is walked row by row with each element's own plan and
// its count added to acc (returns false).
template <int NB, int NNE>
__device__ inline bool flat_block(const vsgm_level_args& a, const FlatTable& h, const FlatWindow<NB, NNE>& w, uint64_t x, uint64_t bend,
                                  uint32_t* c, FlatBlock<NB, NNE>& bk, uint64_t& acc) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t lim = (uint32_t)(bend - x);
  const uint32_t ci = (uint32_t)__popcll(__ballot(w.en <= x));
  const uint32_t cj = (uint32_t)__popcll(__ballot(w.en < bend));
  const bool span = lane >= ci && lane <= cj;
  bk.md = __builtin_amdgcn_readlane(w.md, ci);
  bool same = w.md == bk.md;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    bk.t[b] = readlane_u64(w.t[b], ci);
    bk.nn[b] = __builtin_amdgcn_readlane(w.nn[b], ci);
    same = same && w.t[b] == bk.t[b] && w.nn[b] == bk.nn[b];
  }
#pragma unroll
  for (int k = 0; k < NNE; ++k) {
    bk.ne[k] = __builtin_amdgcn_readlane(w.ne[k], ci);
    same = same && w.ne[k] == bk.ne[k];
  }
  if (__ballot(span && !same) == 0) {
    if (ci == cj) {
      const uint32_t* src = a.cols + readlane_u64(w.base, ci) + x;
      if (lim == (uint32_t)kFlatBlock) {  // a full block (uniform): no per-row bound
#pragma unroll
        for (int it = 0; it < kFlatRows; ++it) c[it] = src[(uint32_t)it * kWave + lane];
      } else {
        // rows past the end load the block's last element and are dropped
#pragma unroll
        for (int it = 0; it < kFlatRows; ++it) {
          const uint32_t r = (uint32_t)it * kWave + lane;
          const uint32_t v = src[min(r, lim - 1)];
          c[it] = r < lim ? v : kNoCand;
        }
      }
    } else {
      // element r of the block lies in path ci + #{path starts in (x, x + r]}
      uint32_t pi[kFlatRows];
#pragma unroll
      for (int it = 0; it < kFlatRows; ++it) pi[it] = ci;
      for (uint32_t l = ci + 1; l <= cj; ++l) {
        const uint32_t s = (uint32_t)(readlane_u64(w.st, l) - x);
#pragma unroll
        for (int it = 0; it < kFlatRows; ++it) pi[it] += ((uint32_t)it * kWave + lane >= s) ? 1u : 0u;
      }
      // every row's segment base is gathered first (with every lane active),
      // then every row's candidate is loaded; a row past the end loads the
      // block's last element (it lies in path cj, which such a row's pi is)
      // and is dropped
      uint64_t bse[kFlatRows];
#pragma unroll
      for (int it = 0; it < kFlatRows; ++it) bse[it] = shfl_u64(w.base, pi[it]);
      const uint32_t* src = a.cols + x;
#pragma unroll
      for (int it = 0; it < kFlatRows; ++it) {
        const uint32_t r = (uint32_t)it * kWave + lane;
        const uint32_t v = src[bse[it] + min(r, lim - 1)];
        c[it] = r < lim ? v : kNoCand;
      }
    }
    return true;
  }
  // paths with different probe plans: row by row, each element with its own path's plan
#pragma unroll 1
  for (uint32_t r0 = 0; r0 < lim; r0 += kWave) {
    const uint32_t r = r0 + lane;
    uint32_t pi = ci;
    for (uint32_t l = ci + 1; l <= cj; ++l) pi += r >= (uint32_t)(readlane_u64(w.st, l) - x) ? 1u : 0u;
    bool ok = r < lim;
    const uint64_t bse = shfl_u64(w.base, pi);  // every lane active
    const uint32_t cv = ok ? a.cols[bse + x + r] : 0u;
    const uint32_t mdi = (uint32_t)__shfl((int)w.md, (int)pi);
#pragma unroll
    for (int k = 0; k < NNE; ++k) {
      const uint32_t nv = (uint32_t)__shfl((int)w.ne[k], (int)pi);  // every lane active
      ok = ok && cv != nv;
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const uint32_t mode = (mdi >> (2 * b)) & 3u;
      const uint64_t tb = shfl_u64(w.t[b], pi);
      const uint32_t nb = (uint32_t)__shfl((int)w.nn[b], (int)pi);
      if (ok && mode != kModePivot) ok = flat_test(a, h, mode, tb, nb, cv);
    }
    acc += (uint64_t)__popcll(__ballot(ok));
  }
  return false;
}

// One wavefront's walk over chunk [e0, e1), paths first..last. Returns the
// wave's count (uniform across the wave). Blocks are software-pipelined
// over two register buffers: a block's candidate loads are issued before
// the previous block's tests run.
template <int NB, int NNE>
__device__ inline uint64_t flat_walk(const vsgm_level_args& a, const FlatCols& fc, const uint64_t* __restrict__ seg_begin,
                                     const uint64_t* __restrict__ seg_off, uint64_t e0, uint64_t e1, uint64_t first, uint64_t last,
                                     const FlatTable& h) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  uint64_t acc = 0;
  uint64_t p = first;
  uint64_t x = e0;
  uint32_t cA[kFlatRows], cB[kFlatRows];
  FlatBlock<NB, NNE> bA, bB;
  bool pA = false, pB = false;  // block whose loads are in flight in cA / cB
  while (x < e1 && p <= last) {
    FlatWindow<NB, NNE> w;
    w.n = (uint32_t)min<uint64_t>(kWave, last - p + 1);
    const bool in = lane < w.n;
    const uint64_t q = p + lane;
    w.st = in ? seg_off[q] : ~0ull;
    w.en = in ? seg_off[q + 1] : ~0ull;
    const uint64_t sb = in ? seg_begin[q] : 0;
    w.base = (sb & kBeginMask) - w.st;
    const uint32_t piv = (uint32_t)(sb >> 56);
    uint32_t v[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) v[b] = in ? a.inst[a.back[b]][q] : 0u;
#pragma unroll
    for (int k = 0; k < NNE; ++k) w.ne[k] = in ? a.inst[fc.ne[k]][q] : 0u;
    uint32_t sl[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const bool tested = in && (uint32_t)b != piv && v[b] != h.v;
      sl[b] = (tested && a.hub_slot) ? a.hub_slot[v[b]] : ~0u;
    }
    w.md = 0;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      uint32_t mode;
      w.t[b] = 0;
      w.nn[b] = 0;
      if ((uint32_t)b == piv || !in) {
        mode = kModePivot;
      } else if (v[b] == h.v) {
        mode = kModeHeld;
      } else if (sl[b] != ~0u) {
        mode = kModeIndex;
        w.t[b] = (uint64_t)sl[b] * a.hub_words;
      } else {
        mode = kModeSearch;
        const uint64_t rb = a.row_ptrs[v[b]];
        w.t[b] = rb;
        w.nn[b] = (uint32_t)(a.row_ptrs[v[b] + 1] - rb);
      }
      w.md |= mode << (2 * b);
    }
    const uint64_t wend = min(e1, readlane_u64(w.en, w.n - 1));
    while (x < wend) {
      {
        const uint64_t bend = min(x + kFlatBlock, wend);
        const bool fa = flat_block<NB, NNE>(a, h, w, x, bend, cA, bA, acc);
        if (pB) acc += flat_finish<NB, NNE>(a, h, bB, cB);
        pB = false;
        pA = fa;
        x = bend;
      }
      if (x >= wend) break;
      {
        const uint64_t bend = min(x + kFlatBlock, wend);
        const bool fb = flat_block<NB, NNE>(a, h, w, x, bend, cB, bB, acc);
        if (pA) acc += flat_finish<NB, NNE>(a, h, bA, cA);
        pA = false;
        pB = fb;
        x = bend;
      }
    }
    // the next window starts with buffer A again: nothing may stay in flight
    if (pA) acc += flat_finish<NB, NNE>(a, h, bA, cA);
    if (pB) acc += flat_finish<NB, NNE>(a, h, bB, cB);
    pA = pB = false;
    x = wend;
    p += w.n;
  }
  return acc;
}

template <int NB, int NNE>
__global__ __launch_bounds__(kHThreads) void k_map_flat(vsgm_level_args a, FlatCols fc, const uint64_t* __restrict__ seg_begin,
                                                        const uint64_t* __restrict__ seg_off, uint64_t W, uint64_t c_lo, uint64_t c_hi,
                                                        const ChunkPlan* __restrict__ plan, uint32_t ticket_chunks,
                                                        unsigned long long* __restrict__ tickets, unsigned long long* __restrict__ total) {
  __shared__ uint32_t s_cuckoo[2 * kFlatSlots];
  __shared__ uint32_t s_fail;
  __shared__ ChunkPlan s_plan[kTicketCap];
  __shared__ uint32_t s_votes[kTicketCap];
  __shared__ uint64_t s_tmp[kHWaves];
  __shared__ uint64_t s_ticket;
  __shared__ uint32_t s_next;
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t n_tickets = (c_hi - c_lo + ticket_chunks - 1) / ticket_chunks;
  uint32_t hv = ~0u, hsh = 26;  // held vertex, hash shift (64 slots)
  uint64_t acc = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      s_ticket = atomicAdd(tickets, 1ull);
      s_next = 0;
    }
    __syncthreads();
    const uint64_t t = readfirstlane_u64(s_ticket);
    if (t >= n_tickets) break;
    const uint64_t k_lo = c_lo + t * ticket_chunks, k_hi = min(k_lo + ticket_chunks, c_hi);
    const uint32_t nk = (uint32_t)(k_hi - k_lo);
    if (threadIdx.x < nk) s_plan[threadIdx.x] = plan[k_lo - c_lo + threadIdx.x];
    __syncthreads();
    if (threadIdx.x < nk) {
      const uint32_t v = s_plan[threadIdx.x].want_v;
      uint32_t votes = 0;
      if (v != ~0u)
        for (uint32_t i = 0; i < nk; ++i) votes += s_plan[i].want_v == v ? 1u : 0u;
      s_votes[threadIdx.x] = votes;
    }
    __syncthreads();
    uint32_t best = 0;
    for (uint32_t i = 1; i < nk; ++i)
      if (s_votes[i] > s_votes[best]) best = i;
    const uint32_t want = __builtin_amdgcn_readfirstlane(s_votes[best] ? s_plan[best].want_v : ~0u);
    if (want != ~0u && want != hv) {
      const uint32_t n = s_plan[best].want_n;
      // smallest power of two >= 1.125 n + 64 slots per table (load <= 0.45)
      const uint32_t need = n + n / 8 + 64;
      uint32_t bits = 6;
      while ((1u << bits) < need) ++bits;
      hsh = 32 - bits;
      hv = flat_build(s_cuckoo, s_cuckoo + kFlatSlots, hsh, a.cols + a.row_ptrs[want], n, &s_fail) ? want : ~0u;
    }
    const FlatTable h{s_cuckoo, s_cuckoo + kFlatSlots, hsh, hv};
    for (;;) {
      uint32_t j = 0;
      if (lane == 0) j = atomicAdd(&s_next, 1u);
      j = __builtin_amdgcn_readfirstlane(j);  // uniform (scalar) from here on
      if (j >= nk) break;
      const uint64_t first = readfirstlane_u64(s_plan[j].first), last = readfirstlane_u64(s_plan[j].last);
      const uint64_t k = k_lo + j;
      const uint64_t e0 = k * kHChunk, e1 = min(e0 + kHChunk, W);
      acc += flat_walk<NB, NNE>(a, fc, seg_begin, seg_off, e0, e1, first, last, h);
    }
    __syncthreads();
  }
  if (lane == 0) s_tmp[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t sum = 0;
    for (int w = 0; w < kHWaves; ++w) sum += s_tmp[w];
    if (sum) atomicAdd(total, (unsigned long long)sum);
  }
}

// Scatter the valid elements of the compacting ops; chunk_off is the
// exclusive scan of the chunk counts (n_chunks + 1 entries).
template <int OP>
__global__ __launch_bounds__(kBlock) void k_scatter(vsgm_level_args a, const uint64_t* __restrict__ seg_begin,
                                                    const uint64_t* __restrict__ seg_off, uint64_t W, uint64_t n_chunks,
                                                    const uint64_t* __restrict__ mask, const uint64_t* __restrict__ chunk_off,
                                                    uint32_t* __restrict__ out_vals, uint64_t* __restrict__ out_off, ExpandCols cols) {
  __shared__ int32_t s_rel[kPathCap + 1];
  __shared__ uint64_t s_word[kWordsPerChunk + 1];
  __shared__ uint64_t s_wpre[kWordsPerChunk + 1];
  __shared__ ChunkTable s_tab;
  constexpr bool kStarts = (OP != OP_EXPAND);
  const uint64_t P = a.path_num;
  const int lane = threadIdx.x & (kWave - 1);
  for (uint64_t k0 = 0; block_chunk<1>(k0) < n_chunks; k0 += kBlock) {
  fill_chunk_table<1, kStarts>(seg_off, P, W, n_chunks, k0, &s_tab);
  for (uint32_t t = 0; t < (uint32_t)kBlock; ++t) {
    const uint64_t chunk = block_chunk<1>(k0 + t);
    if (chunk >= n_chunks) break;
    const uint64_t e0 = chunk * kChunk;
    const uint64_t e1 = min(e0 + (uint64_t)kChunk, W);
    const uint32_t n_words = (uint32_t)((e1 - e0 + kWave - 1) / kWave);
    const uint64_t base = chunk_off[chunk];
    ChunkMap cm;
    if (threadIdx.x <= kWordsPerChunk) s_word[threadIdx.x] = threadIdx.x < n_words ? mask[e0 / kWave + threadIdx.x] : 0;
    build_chunk_map(seg_off, s_tab.first[t], s_tab.last[t], e0, cm, s_rel);
    __syncthreads();
    if (threadIdx.x == 0) {
      uint64_t run = 0;
      for (uint32_t w = 0; w <= kWordsPerChunk; ++w) {
        s_wpre[w] = run;
        run += __popcll(s_word[w]);
      }
    }
    __syncthreads();
    for (int it = 0; it < kItems; ++it) {
      const uint64_t e = e0 + (uint64_t)it * kBlock + threadIdx.x;
      if (e >= e1) continue;
      const uint32_t w = (uint32_t)((e - e0) / kWave);
      const uint64_t word = s_word[w];
      if (!((word >> lane) & 1ull)) continue;
      const uint64_t pos = base + s_wpre[w] + __popcll(word & ((1ull << lane) - 1ull));
      uint64_t rank;
      const uint64_t p = chunk_path(e, e0, cm, s_rel, seg_off, rank);
      const uint64_t packed = seg_begin[p];
      const uint32_t c = seg_source<OP>(a, packed >> 56)[(packed & kBeginMask) + rank];
      if (OP == OP_EXPAND) {
        cols.out_inst[a.target][pos] = c;
        for (uint32_t i = 0; i < cols.n_mat; ++i) cols.out_inst[cols.mat[i]][pos] = a.inst[cols.mat[i]][p];
        for (uint32_t i = 0; i < cols.n_cu; ++i) cols.out_idx[cols.cu[i]][pos] = cand_row(a, cols.cu[i], p);
      } else {
        out_vals[pos] = c;
      }
    }
    if (kStarts) {
      // new row offsets of the paths whose segment starts in this chunk
      const uint64_t p_lo = s_tab.p_lo[t];
      const uint64_t p_hi = s_tab.p_hi[t];
      for (uint64_t p = p_lo + threadIdx.x; p < p_hi; p += kBlock) {
        uint64_t rel = seg_off[p] - e0;
        uint32_t w = (uint32_t)(rel / kWave);
        uint32_t bit = (uint32_t)(rel % kWave);
        uint64_t word = w < n_words ? s_word[w] : 0;
        out_off[p] = base + s_wpre[w] + __popcll(word & ((1ull << bit) - 1ull));
      }
    }
    __syncthreads();
  }
  }
}

__global__ __launch_bounds__(kBlock) void k_count1(vsgm_level_args a, unsigned long long* total) {
  __shared__ uint64_t s_tmp[kBlock / kWave];
  uint64_t s = 0;
  for (uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x; p < a.path_num; p += (uint64_t)gridDim.x * kBlock) {
    uint64_t r = cand_row(a, a.target, p);
    s += a.cand_off[a.target][r + 1] - a.cand_off[a.target][r];
  }
  s = block_sum(s, s_tmp);
  if (threadIdx.x == 0 && s) atomicAdd(total, (unsigned long long)s);
}

__global__ __launch_bounds__(kBlock) void k_children(vsgm_level_args a, int mode, uint64_t* __restrict__ out) {
  for (uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x; p < a.path_num; p += (uint64_t)gridDim.x * kBlock) {
    uint64_t v = 1;
    if (mode == 0) {
      uint64_t best = 0;
      for (uint32_t i = 0; i < a.n_back; ++i) {
        uint32_t x = a.inst[a.back[i]][p];
        uint64_t d = a.row_ptrs[x + 1] - a.row_ptrs[x];
        if (i == 0 || d < best) best = d;
      }
      v = best;
    } else if (mode == 1) {
      uint64_t r = cand_row(a, a.target, p);
      v = a.cand_off[a.target][r + 1] - a.cand_off[a.target][r];
    }
    out[p] = v;
  }
}

__global__ __launch_bounds__(kBlock) void k_cost(const uint64_t* __restrict__ children, uint64_t n, uint64_t pf, uint64_t cf,
                                                 uint64_t* __restrict__ cost) {
  for (uint64_t e = (uint64_t)blockIdx.x * kBlock + threadIdx.x; e < n; e += (uint64_t)gridDim.x * kBlock) cost[e] = children[e] * cf + pf;
}

// ends[k] = number of parents whose cost prefix stays <= (k+1) * batch_bytes
__global__ __launch_bounds__(kBlock) void k_split(const uint64_t* __restrict__ prefix, uint64_t n, uint64_t batch_bytes, uint64_t nb,
                                                  uint64_t* __restrict__ ends) {
  for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < nb; k += (uint64_t)gridDim.x * kBlock) {
    uint64_t target = (k + 1) * batch_bytes;
    ends[k] = last_le_u64(prefix, n + 1, target);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// context and pool

// Device scratch of the allocation-free entry point
// (vsgm_intersect_count_total_noalloc): every device this library runs on
// holds one copy for the life of the process. Host code never passes this
// name to a kernel (on the host it is only a stand-in for the symbol);
// vsgm_ctx_create looks up the device address once per context, with the
// context's device current, and everything uses that address.
constexpr uint64_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

struct vsgm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start[VSGM_TIMER_SLOTS] = {};
  hipEvent_t ev_stop[VSGM_TIMER_SLOTS] = {};
  uint64_t budget = 0;
  uint64_t live = 0;
  uint64_t peak = 0;
  std::multimap<uint64_t, void*> cached;  // size class -> free block
  std::map<void*, uint64_t> blocks;       // live block -> size class
  uint64_t cached_bytes = 0;
  uint64_t* h_pinned = nullptr;           // small pinned staging for D2H of totals
  uint64_t* d_stats = nullptr;            // device u64[4]: scan total, adjacency degree sum, ticket counter, spare
  // device address of `scratch` on this context's device, and the bump arena
  // that pool temporaries come from while an allocation-free call runs
  uint8_t* scratch_ptr = nullptr;
  bool arena_on = false;
  uint64_t arena_off = 0;
  int n_cu = 0;                           // compute units of the device
  std::string err;
  vsgm_op_stats stats{0, 0, 0, 0};
  // kernel profiling: HIP events around each launch on the context stream
  struct KernelRecord {
    const char* name;
    hipEvent_t start, stop;
    uint64_t bytes, work, paths;
  };
  bool profiling = false;
  std::vector<KernelRecord> records;
  std::vector<hipEvent_t> spare_events;
  vsgm_tuning tuning{4, 8192, kDefaultColdWeight, kDefaultTicketChunks};
  // VSGM_PROGRESS=<seconds>: long map launches report their progress on
  // stderr at most this often (0: silent)
  double progress_s = 0;
  // Host-visible values (element totals, output counts, batch ends) are the
  // only data that steer the host side of a level op. Tape mode 1 records
  // them as they are read back; mode 2 replays them without touching the
  // device, which is what lets a whole Match be captured into one graph.
  int tape_mode = 0;
  std::vector<uint64_t> tape;
  size_t tape_pos = 0;
  bool capturing = false;
  bool saved_profiling = false;
  hipGraphExec_t graph_exec = nullptr;
  uint64_t graph_nodes = 0;
};

namespace {

int fail(vsgm_ctx* ctx, int code, const char* fmt, ...) {
  if (ctx) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    ctx->err = buf;
  }
  return code;
}

#define HIP_TRY(ctx, call)                                                                              \
  do {                                                                                                  \
    hipError_t _e = (call);                                                                             \
    if (_e != hipSuccess) return fail(ctx, VSGM_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), __FILE__, __LINE__); \
  } while (0)

#define VSGM_TRY(call)          \
  do {                          \
    int _s = (call);            \
    if (_s != VSGM_OK) return _s; \
  } while (0)

uint64_t size_class(uint64_t bytes) {
  if (bytes == 0) bytes = 1;
  if (bytes <= (1ull << 20)) return (bytes + 511) & ~511ull;
  // quarter steps between powers of two: at most 25% slack
  int lg = 63 - __builtin_clzll(bytes);
  uint64_t step = (1ull << lg) / 4;
  return (bytes + step - 1) / step * step;
}

int pool_alloc(vsgm_ctx* ctx, uint64_t bytes, void** out) {
  uint64_t cls = size_class(bytes);
  if (ctx->live + cls > ctx->budget)
    return fail(ctx, VSGM_ERR_BUDGET, "allocation of %llu bytes exceeds the budget (%llu live of %llu)", (unsigned long long)cls,
                (unsigned long long)ctx->live, (unsigned long long)ctx->budget);
  auto it = ctx->cached.find(cls);
  void* p = nullptr;
  if (it != ctx->cached.end()) {
    p = it->second;
    ctx->cached.erase(it);
    ctx->cached_bytes -= cls;
  } else {
    hipError_t e = hipMalloc(&p, cls);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      // cached blocks may be referenced by a captured graph: keep them
      if (ctx->capturing || ctx->graph_exec)
        return fail(ctx, VSGM_ERR_OOM, "hipMalloc(%llu) failed while a captured graph holds the pool", (unsigned long long)cls);
      // give cached blocks back and retry once
      HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
      for (auto& kv : ctx->cached) (void)hipFree(kv.second);
      ctx->cached.clear();
      ctx->cached_bytes = 0;
      e = hipMalloc(&p, cls);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, VSGM_ERR_OOM, "hipMalloc(%llu) failed: %s", (unsigned long long)cls, hipGetErrorString(e));
      }
    }
  }
  ctx->blocks[p] = cls;
  ctx->live += cls;
  ctx->peak = std::max(ctx->peak, ctx->live);
  *out = p;
  return VSGM_OK;
}

int pool_free(vsgm_ctx* ctx, void* p) {
  if (!p) return VSGM_OK;
  auto it = ctx->blocks.find(p);
  if (it == ctx->blocks.end()) return fail(ctx, VSGM_ERR_ARG, "vsgm_free of a pointer the pool does not own");
  uint64_t cls = it->second;
  ctx->blocks.erase(it);
  ctx->live -= cls;
  // Blocks are reused only by later work on the same stream, so no sync.
  ctx->cached.emplace(cls, p);
  ctx->cached_bytes += cls;
  return VSGM_OK;
}

// RAII holder of pool temporaries inside one entry point.
// While ctx->arena_on, temporaries are carved from the context's scratch
// (256-byte aligned, bump order, never freed individually) and the pool is
// not touched.
struct Temps {
  vsgm_ctx* ctx;
  std::vector<void*> ptrs;
  explicit Temps(vsgm_ctx* c) : ctx(c) {}
  ~Temps() {
    for (void* p : ptrs) pool_free(ctx, p);
  }
  template <typename T>
  int get(uint64_t n, T** out) {
    const uint64_t bytes = std::max<uint64_t>(n, 1) * sizeof(T);
    if (ctx->arena_on) {
      const uint64_t off = (ctx->arena_off + 255) & ~255ull;
      if (off + bytes > kScratchBytes)
        return fail(ctx, VSGM_ERR_BUDGET, "temporaries of %llu bytes exceed the %llu-byte scratch", (unsigned long long)(off + bytes),
                    (unsigned long long)kScratchBytes);
      ctx->arena_off = off + bytes;
      *out = (T*)(ctx->scratch_ptr + off);
      return VSGM_OK;
    }
    void* p = nullptr;
    int s = pool_alloc(ctx, bytes, &p);
    if (s != VSGM_OK) return s;
    ptrs.push_back(p);
    *out = (T*)p;
    return VSGM_OK;
  }
};

hipEvent_t take_event(vsgm_ctx* ctx) {
  hipEvent_t e = nullptr;
  if (!ctx->spare_events.empty()) {
    e = ctx->spare_events.back();
    ctx->spare_events.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    (void)hipGetLastError();
    e = nullptr;
  }
  return e;
}

// Brackets one kernel launch with events when profiling is on; `bytes` is
// the launch's algorithmic byte count.
struct ProfScope {
  vsgm_ctx* ctx;
  long idx = -1;
  ProfScope(vsgm_ctx* c, const char* name, uint64_t bytes, uint64_t work, uint64_t paths) : ctx(c) {
    if (!ctx->profiling) return;
    hipEvent_t a = take_event(ctx), b = take_event(ctx);
    if (!a || !b) return;
    (void)hipEventRecord(a, ctx->stream);
    ctx->records.push_back({name, a, b, bytes, work, paths});
    idx = (long)ctx->records.size() - 1;
  }
  ~ProfScope() {
    if (idx >= 0) (void)hipEventRecord(ctx->records[idx].stop, ctx->stream);
  }
};

// Read n u64 values the host side needs from device memory: a synchronous
// copy (recorded on the tape in mode 1), or the next n tape values in
// replay mode 2, with no device work at all.
int host_read(vsgm_ctx* ctx, const uint64_t* d, uint64_t* h, uint64_t n) {
  if (n == 0) return VSGM_OK;
  if (ctx->tape_mode == 2) {
    if (ctx->tape_pos + n > ctx->tape.size()) return fail(ctx, VSGM_ERR_ARG, "replay ran past the recorded values");
    std::copy(ctx->tape.begin() + ctx->tape_pos, ctx->tape.begin() + ctx->tape_pos + n, h);
    ctx->tape_pos += n;
    return VSGM_OK;
  }
  if (n <= 8) {
    HIP_TRY(ctx, hipMemcpyAsync(ctx->h_pinned, d, n * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    std::copy(ctx->h_pinned, ctx->h_pinned + n, h);
  } else {
    HIP_TRY(ctx, hipMemcpyAsync(h, d, n * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  }
  if (ctx->tape_mode == 1) ctx->tape.insert(ctx->tape.end(), h, h + n);
  return VSGM_OK;
}

int read_u64(vsgm_ctx* ctx, const uint64_t* d, uint64_t* h) { return host_read(ctx, d, h, 1); }

int launch_check(vsgm_ctx* ctx, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ctx, VSGM_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
  return VSGM_OK;
}

int scan_impl(vsgm_ctx* ctx, const uint64_t* in, uint64_t* out, uint64_t n) {
  uint64_t n_tiles = (n + kChunk - 1) / kChunk;
  if (n == 0) {
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(kBlock), 0, ctx->stream, out, 1, 0);
    return launch_check(ctx, "k_fill");
  }
  if (n_tiles == 1) {
    hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(kBlock), 0, ctx->stream, in, out, n, n_tiles, (const uint64_t*)nullptr);
    return launch_check(ctx, "k_tile_scan");
  }
  Temps t(ctx);
  uint64_t* partial;
  VSGM_TRY(t.get(n_tiles + 1, &partial));
  hipLaunchKernelGGL(k_tile_sum, dim3(grid_for(n_tiles)), dim3(kBlock), 0, ctx->stream, in, n, partial, n_tiles);
  VSGM_TRY(launch_check(ctx, "k_tile_sum"));
  VSGM_TRY(scan_impl(ctx, partial, partial, n_tiles));
  hipLaunchKernelGGL(k_tile_scan, dim3(grid_for(n_tiles)), dim3(kBlock), 0, ctx->stream, in, out, n, n_tiles, (const uint64_t*)partial);
  return launch_check(ctx, "k_tile_scan");
}

int check_args(vsgm_ctx* ctx, const vsgm_level_args* a) {
  if (!ctx) return VSGM_ERR_ARG;
  if (!a) return fail(ctx, VSGM_ERR_ARG, "null level args");
  if (a->n_back > VSGM_MAX_QV || a->n_cond > VSGM_MAX_QV || a->target >= VSGM_MAX_QV) return fail(ctx, VSGM_ERR_ARG, "level args out of range");
  if (a->other >= VSGM_MAX_QV) return fail(ctx, VSGM_ERR_ARG, "other vertex out of range");
  for (uint32_t i = 0; i < a->n_cond; ++i)
    if (a->cond[i].operand >= VSGM_MAX_QV || !a->inst[a->cond[i].operand] || a->cond[i].op > VSGM_NON_EQUAL)
      return fail(ctx, VSGM_ERR_ARG, "condition %u names an unmaterialized vertex", i);
  return VSGM_OK;
}

int check_intersect(vsgm_ctx* ctx, const vsgm_level_args* a) {
  if (a->n_back == 0) return fail(ctx, VSGM_ERR_ARG, "intersection without backward neighbours");
  if (!a->row_ptrs || !a->cols) return fail(ctx, VSGM_ERR_ARG, "intersection without a view");
  for (uint32_t i = 0; i < a->n_back; ++i)
    if (a->back[i] >= VSGM_MAX_QV || !a->inst[a->back[i]]) return fail(ctx, VSGM_ERR_ARG, "backward vertex %u is not materialized", i);
  return VSGM_OK;
}

int check_cand(vsgm_ctx* ctx, const vsgm_level_args* a, uint32_t u) {
  if (u >= VSGM_MAX_QV || !a->cand_off[u] || !a->cand[u]) return fail(ctx, VSGM_ERR_ARG, "vertex %u has no candidate set", u);
  return VSGM_OK;
}

// Kernel names used by the profiler, per operator.
template <int OP>
constexpr const char* kname(int stage) {
  // stage 0: setup, 1: map, 2: scatter
  return OP == OP_ICOUNT   ? (stage == 0 ? "setup_icount" : "map_icount")
         : OP == OP_IWRITE ? (stage == 0 ? "setup_iwrite" : stage == 1 ? "map_iwrite" : "scatter_iwrite")
         : OP == OP_EXPAND ? (stage == 0 ? "setup_expand" : stage == 1 ? "map_expand" : "scatter_expand")
         : OP == OP_FILTER ? (stage == 0 ? "setup_filter" : stage == 1 ? "map_filter" : "scatter_filter")
         : OP == OP_COUNT2 ? (stage == 0 ? "setup_count2" : "map_count2")
                           : (stage == 0 ? "setup_ipath" : "map_ipath");
}

// Algorithmic bytes one path costs the setup kernel: its bound vertices,
// their row_ptrs (or candidate rows) and the two segment words written.
template <int OP>
uint64_t setup_bytes_per_path(const vsgm_level_args& a) {
  if (is_intersect(OP)) return (4 + 16) * (uint64_t)a.n_back + 16;
  if (OP == OP_COUNT2) return 2 * (8 + 16) + 16;
  return 8 + 16 + 16;
}

// Shared driver: setup + scan of one level operator; returns the element
// count W and the merge-model adjacency volume S (intersections).
template <int OP>
int run_setup_scan(vsgm_ctx* ctx, const vsgm_level_args& a, Temps& t, uint64_t** seg_begin, uint64_t** seg_off, uint64_t* W) {
  const uint64_t P = a.path_num;
  VSGM_TRY(t.get(P, seg_begin));
  VSGM_TRY(t.get(P + 1, seg_off));
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->d_stats, 2, 0);
  {
    ProfScope ps(ctx, kname<OP>(0), P * setup_bytes_per_path<OP>(a), P, P);
    const bool hashed = (OP == OP_ICOUNT || OP == OP_IPATH) && ctx->tuning.isect_kernel >= 3 && a.n_back > 1;
    hipLaunchKernelGGL(k_setup<OP>, dim3(grid_for((P + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, a, *seg_begin, *seg_off,
                       (unsigned long long*)(ctx->d_stats + 1), hashed ? kHashCap : 0u, ctx->tuning.cold_weight);
  }
  VSGM_TRY(launch_check(ctx, "k_setup"));
  VSGM_TRY(scan_impl(ctx, *seg_off, *seg_off, P));
  hipLaunchKernelGGL(k_copy1, dim3(1), dim3(1), 0, ctx->stream, *seg_off + P, ctx->d_stats);
  VSGM_TRY(launch_check(ctx, "k_copy1"));
  uint64_t vals[2];
  VSGM_TRY(host_read(ctx, ctx->d_stats, vals, 2));
  *W = vals[0];
  ctx->stats.adj_entries = vals[1];
  return VSGM_OK;
}

// Algorithmic bytes of the map kernel (the streaming model): every one of
// the W candidate elements is read once from its pivot list, and each path's
// bound vertices, row_ptrs and segment words once. Membership tests against
// the other lists are not counted: they are answered from the LDS table or
// the membership index, whose traffic rocprof measures. (The merge model,
// every backward list read whole per path, 4 * adj_entries, over-counts
// what a probing intersection must move, so it is not used as the bound.)
template <int OP>
uint64_t map_bytes(const vsgm_ctx* ctx, const vsgm_level_args& a, uint64_t W, uint64_t n_chunks) {
  (void)ctx;
  const uint64_t P = a.path_num;
  uint64_t b;
  if (is_intersect(OP)) {
    b = 4 * W + P * ((4 + 16) * (uint64_t)a.n_back + 16);
    if (OP == OP_ICOUNT && a.other >= 0) b += P * (8 + 16);
  } else if (OP == OP_COUNT2) {
    b = 4 * W + P * (2 * (8 + 16) + 16);
  } else {
    b = 4 * W + P * 16;
  }
  if (OP == OP_IWRITE || OP == OP_EXPAND || OP == OP_FILTER) b += (W + 7) / 8 + 8 * n_chunks;
  return b;
}

// Chunks per map launch: a level op with more elements runs as several
// launches over consecutive chunk windows (about 8.6e9 elements each), so
// no single dispatch runs for minutes and progress can be reported between
// windows.
constexpr uint64_t kWindowChunks = 1ull << 22;

// One map launch over chunks [c_lo, c_hi): k_map_isect or k_map, as the
// tuning selects.
template <int OP>
void launch_map_window(vsgm_ctx* ctx, const vsgm_level_args& a, const uint64_t* seg_begin, const uint64_t* seg_off, uint64_t W,
                       uint64_t n_chunks, uint64_t c_lo, uint64_t c_hi, uint64_t* mask, uint64_t* chunk_cnt, unsigned long long* total,
                       unsigned long long* path_counts) {
  const unsigned grid = (unsigned)std::min<uint64_t>(std::max<uint64_t>(c_hi - c_lo, 1), ctx->tuning.max_grid);
  if constexpr (is_intersect(OP) || OP == OP_COUNT2) {
    if (ctx->tuning.isect_kernel == 1) {
      hipLaunchKernelGGL(k_map_isect<OP>, dim3(grid), dim3(kBlock), 0, ctx->stream, a, seg_begin, seg_off, W, n_chunks, c_lo, c_hi, mask,
                         chunk_cnt, total, path_counts);
      return;
    }
  }
  if constexpr (is_intersect(OP)) {
    if (ctx->tuning.isect_kernel == 2) {
      hipLaunchKernelGGL((k_map<OP, true>), dim3(grid), dim3(kBlock), 0, ctx->stream, a, seg_begin, seg_off, W, n_chunks, c_lo, c_hi, mask,
                         chunk_cnt, total, path_counts);
      return;
    }
  }
  hipLaunchKernelGGL(k_map<OP>, dim3(grid), dim3(kBlock), 0, ctx->stream, a, seg_begin, seg_off, W, n_chunks, c_lo, c_hi, mask, chunk_cnt,
                     total, path_counts);
}

// The whole map of one level op, window by window. Each window is one
// profiled launch carrying its share (by chunks) of the op's algorithmic
// bytes `bytes`.
// The hashed map over chunks [c_lo, c_hi) (2048-element units, as the
// window loop counts them): the chunk plan pass, then k_map_hash on one
// workgroup per CU pulling tickets from the context's counter.
template <int OP>
int launch_map_hash(vsgm_ctx* ctx, const vsgm_level_args& a, const uint64_t* seg_begin, const uint64_t* seg_off, uint64_t W, uint64_t c_lo,
                    uint64_t c_hi, ChunkPlan* plan, uint64_t bytes, uint64_t paths, unsigned long long* total, unsigned long long* path_counts) {
  constexpr uint64_t kRatio = kHChunk / kChunk;
  const uint64_t h_lo = c_lo / kRatio, h_hi = (c_hi + kRatio - 1) / kRatio;
  const uint64_t e_lo = c_lo * kChunk, e_hi = std::min(W, c_hi * kChunk);
  hipLaunchKernelGGL(k_chunk_plan, dim3(grid_for((h_hi - h_lo + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, a, seg_begin, seg_off, W,
                     h_lo, h_hi, plan);
  VSGM_TRY(launch_check(ctx, "k_chunk_plan"));
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->d_stats + 2, 1, 0);
  VSGM_TRY(launch_check(ctx, "k_fill"));
  StageCols sc{};
  sc.stage = 1;
  auto col_of = [&sc](uint32_t u) {
    for (uint32_t q = 0; q < sc.n; ++q)
      if (sc.col[q] == u) return q;
    if (sc.n == kStageCols) {
      sc.stage = 0;
      return 0u;
    }
    sc.col[sc.n] = u;
    return sc.n++;
  };
  for (uint32_t i = 0; i < a.n_back; ++i) sc.back_col[i] = col_of(a.back[i]);
  for (uint32_t i = 0; i < a.n_cond; ++i)
    if (a.cond[i].op == VSGM_NON_EQUAL) sc.ne_col[sc.n_ne++] = col_of(a.cond[i].operand);
  const uint32_t tc = ctx->tuning.ticket_chunks;
  const uint64_t n_tickets = (h_hi - h_lo + tc - 1) / tc;
  const unsigned grid = (unsigned)std::min<uint64_t>(std::min<uint64_t>(n_tickets, (uint64_t)ctx->n_cu), ctx->tuning.max_grid);
  {
    ProfScope ps(ctx, kname<OP>(1), bytes, e_hi - e_lo, paths);
    const int ncol = (sc.stage && sc.n >= 1 && sc.n <= kRegCols) ? (int)sc.n : 0;
#define VSGM_LAUNCH_HASH(N)                                                                                                           \
  hipLaunchKernelGGL((k_map_hash<OP, N>), dim3(grid), dim3(kHThreads), 0, ctx->stream, a, sc, seg_begin, seg_off, W, h_lo, h_hi, plan, tc, \
                     (unsigned long long*)(ctx->d_stats + 2), total, path_counts)
    if (ncol == 1) VSGM_LAUNCH_HASH(1);
    else if (ncol == 2) VSGM_LAUNCH_HASH(2);
    else if (ncol == 3) VSGM_LAUNCH_HASH(3);
    else VSGM_LAUNCH_HASH(0);
#undef VSGM_LAUNCH_HASH
  }
  return launch_check(ctx, kname<OP>(1));
}

// Is 'candidate != vertex bound to u' implied? On a view without self-loops
// it is whenever u is a backward vertex: the candidate lies in u's list.
bool ne_implied(const vsgm_level_args& a, uint32_t u) {
  if (!(a.flags & VSGM_LEVEL_LOOP_FREE)) return false;
  for (uint32_t i = 0; i < a.n_back; ++i)
    if (a.back[i] == u) return true;
  return false;
}

// Levels the flat map runs: COMPUTE_COUNT with two or three backward lists,
// no second computed set, at most three '!=' conditions.
bool flat_eligible(const vsgm_level_args& a) {
  if (a.other >= 0 || a.n_back < 2 || a.n_back > 3 || a.vertex_count >= kNoCand) return false;
  uint32_t n_ne = 0;
  for (uint32_t i = 0; i < a.n_cond; ++i) n_ne += (a.cond[i].op == VSGM_NON_EQUAL && !ne_implied(a, a.cond[i].operand)) ? 1u : 0u;
  return n_ne <= 3;
}

// The flat map over chunks [c_lo, c_hi) (2048-element units): the chunk plan
// pass, then k_map_flat on one workgroup per CU pulling tickets.
int launch_map_flat(vsgm_ctx* ctx, const vsgm_level_args& a, const uint64_t* seg_begin, const uint64_t* seg_off, uint64_t W, uint64_t c_lo,
                    uint64_t c_hi, ChunkPlan* plan, uint64_t bytes, uint64_t paths, unsigned long long* total) {
  constexpr uint64_t kRatio = kHChunk / kChunk;
  const uint64_t h_lo = c_lo / kRatio, h_hi = (c_hi + kRatio - 1) / kRatio;
  const uint64_t e_lo = c_lo * kChunk, e_hi = std::min(W, c_hi * kChunk);
  hipLaunchKernelGGL(k_chunk_plan, dim3(grid_for((h_hi - h_lo + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, a, seg_begin, seg_off, W,
                     h_lo, h_hi, plan);
  VSGM_TRY(launch_check(ctx, "k_chunk_plan"));
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->d_stats + 2, 1, 0);
  VSGM_TRY(launch_check(ctx, "k_fill"));
  FlatCols fc{};
  for (uint32_t i = 0; i < a.n_cond; ++i)
    if (a.cond[i].op == VSGM_NON_EQUAL && !ne_implied(a, a.cond[i].operand)) fc.ne[fc.n_ne++] = a.cond[i].operand;
  const uint32_t tc = ctx->tuning.ticket_chunks;
  const uint64_t n_tickets = (h_hi - h_lo + tc - 1) / tc;
  const unsigned grid = (unsigned)std::min<uint64_t>(std::min<uint64_t>(n_tickets, (uint64_t)ctx->n_cu), ctx->tuning.max_grid);
  {
    ProfScope ps(ctx, kname<OP_ICOUNT>(1), bytes, e_hi - e_lo, paths);
#define VSGM_LAUNCH_FLAT(NB, NNE)                                                                                                        \
  hipLaunchKernelGGL((k_map_flat<NB, NNE>), dim3(grid), dim3(kHThreads), 0, ctx->stream, a, fc, seg_begin, seg_off, W, h_lo, h_hi, plan, tc, \
                     (unsigned long long*)(ctx->d_stats + 2), total)
    const uint32_t key = a.n_back * 4 + fc.n_ne;
    switch (key) {
      case 8: VSGM_LAUNCH_FLAT(2, 0); break;
      case 9: VSGM_LAUNCH_FLAT(2, 1); break;
      case 10: VSGM_LAUNCH_FLAT(2, 2); break;
      case 11: VSGM_LAUNCH_FLAT(2, 3); break;
      case 12: VSGM_LAUNCH_FLAT(3, 0); break;
      case 13: VSGM_LAUNCH_FLAT(3, 1); break;
      case 14: VSGM_LAUNCH_FLAT(3, 2); break;
      default: VSGM_LAUNCH_FLAT(3, 3); break;
    }
#undef VSGM_LAUNCH_FLAT
  }
  return launch_check(ctx, kname<OP_ICOUNT>(1));
}

template <int OP>
int launch_map(vsgm_ctx* ctx, const vsgm_level_args& a, const uint64_t* seg_begin, const uint64_t* seg_off, uint64_t W, uint64_t n_chunks,
               uint64_t bytes, uint64_t* mask, uint64_t* chunk_cnt, unsigned long long* total, unsigned long long* path_counts) {
  const uint64_t n_windows = (n_chunks + kWindowChunks - 1) / kWindowChunks;
  constexpr bool kHashable = OP == OP_ICOUNT || OP == OP_IPATH;
  const bool hashed = kHashable && ctx->tuning.isect_kernel >= 3 && a.n_back > 1;
  const bool flat = OP == OP_ICOUNT && ctx->tuning.isect_kernel == 4 && flat_eligible(a);
  Temps tmp(ctx);
  ChunkPlan* plan = nullptr;
  if (hashed) VSGM_TRY(tmp.get((std::min(n_chunks, kWindowChunks) * kChunk + kHChunk - 1) / kHChunk, &plan));
  const bool report = ctx->progress_s > 0 && n_windows > 1 && !ctx->capturing;
  const auto t0 = std::chrono::steady_clock::now();
  auto last = t0;
  for (uint64_t w = 0; w < n_windows; ++w) {
    const uint64_t c_lo = w * kWindowChunks, c_hi = std::min(n_chunks, c_lo + kWindowChunks);
    const uint64_t e_lo = c_lo * kChunk, e_hi = std::min(W, c_hi * kChunk);
    const uint64_t share = (uint64_t)((unsigned __int128)bytes * (c_hi - c_lo) / n_chunks);
    bool done = false;
    if constexpr (kHashable) {
      if (flat) {
        VSGM_TRY(launch_map_flat(ctx, a, seg_begin, seg_off, W, c_lo, c_hi, plan, share, w == 0 ? a.path_num : 0, total));
        done = true;
      } else if (hashed) {
        VSGM_TRY(launch_map_hash<OP>(ctx, a, seg_begin, seg_off, W, c_lo, c_hi, plan, share, w == 0 ? a.path_num : 0, total, path_counts));
        done = true;
      }
    }
    if (!done) {
      {
        ProfScope ps(ctx, kname<OP>(1), share, e_hi - e_lo, w == 0 ? a.path_num : 0);
        launch_map_window<OP>(ctx, a, seg_begin, seg_off, W, n_chunks, c_lo, c_hi, mask, chunk_cnt, total, path_counts);
      }
      VSGM_TRY(launch_check(ctx, kname<OP>(1)));
    }
    if (report) {
      HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
      const auto now = std::chrono::steady_clock::now();
      if (std::chrono::duration<double>(now - last).count() >= ctx->progress_s || w + 1 == n_windows) {
        last = now;
        fprintf(stderr, "vsgm: %s window %llu/%llu, %.3g of %.3g elements, %.1f s\n", kname<OP>(1), (unsigned long long)(w + 1),
                (unsigned long long)n_windows, (double)e_hi, (double)W, std::chrono::duration<double>(now - t0).count());
        fflush(stderr);
      }
    }
  }
  return VSGM_OK;
}

template <int OP>
int run_count(vsgm_ctx* ctx, const vsgm_level_args& a, uint64_t* d_total, uint64_t* d_path_counts) {
  ctx->stats = {a.path_num, 0, 0, 0};
  if (a.path_num == 0) return VSGM_OK;
  Temps t(ctx);
  uint64_t *seg_begin, *seg_off, W;
  VSGM_TRY(run_setup_scan<OP>(ctx, a, t, &seg_begin, &seg_off, &W));
  ctx->stats.work = W;
  if (W == 0) return VSGM_OK;
  uint64_t n_chunks = (W + kChunk - 1) / kChunk;
  return launch_map<OP>(ctx, a, seg_begin, seg_off, W, n_chunks, map_bytes<OP>(ctx, a, W, n_chunks), nullptr, nullptr,
                        (unsigned long long*)d_total, (unsigned long long*)d_path_counts);
}

// Compacting ops: write the valid elements (count of them: C).
template <int OP>
int run_compact(vsgm_ctx* ctx, const vsgm_level_args& a, uint32_t* vals_or_null, uint64_t* out_off_or_null, ExpandCols cols,
                Temps& t, uint64_t W, uint64_t* seg_begin, uint64_t* seg_off, uint64_t* mask, uint64_t* chunk_off) {
  uint64_t n_chunks = (W + kChunk - 1) / kChunk;
  const uint64_t P = a.path_num, C = ctx->stats.outputs;
  uint64_t bytes = (W + 7) / 8 + 8 * n_chunks + 16 * P + 4 * C;
  if (OP == OP_EXPAND) bytes += C * (4 * (cols.n_mat + 1) + 8 * cols.n_cu) + C * (4 * cols.n_mat + 8 * cols.n_cu);
  else bytes += 4 * C + 8 * (P + 1);
  {
    ProfScope ps(ctx, kname<OP>(2), bytes, W, P);
    hipLaunchKernelGGL(k_scatter<OP>, dim3(grid_for(n_chunks)), dim3(kBlock), 0, ctx->stream, a, seg_begin, seg_off, W, n_chunks, mask,
                       chunk_off, vals_or_null, out_off_or_null, cols);
  }
  return launch_check(ctx, "k_scatter");
}

template <int OP>
int run_mark(vsgm_ctx* ctx, const vsgm_level_args& a, Temps& t, uint64_t W, uint64_t* seg_begin, uint64_t* seg_off, uint64_t** mask,
             uint64_t** chunk_off, uint64_t* count) {
  uint64_t n_chunks = (W + kChunk - 1) / kChunk;
  VSGM_TRY(t.get((W + kWave - 1) / kWave, mask));
  VSGM_TRY(t.get(n_chunks + 1, chunk_off));
  VSGM_TRY(launch_map<OP>(ctx, a, seg_begin, seg_off, W, n_chunks, map_bytes<OP>(ctx, a, W, n_chunks), *mask, *chunk_off, nullptr, nullptr));
  VSGM_TRY(scan_impl(ctx, *chunk_off, *chunk_off, n_chunks));
  VSGM_TRY(read_u64(ctx, *chunk_off + n_chunks, count));
  ctx->stats.outputs = *count;
  return VSGM_OK;
}

template <int OP>
int run_csr(vsgm_ctx* ctx, const vsgm_level_args& a, vsgm_csr_out* out) {
  const uint64_t P = a.path_num;
  ctx->stats = {P, 0, 0, 0};
  out->offsets = nullptr;
  out->values = nullptr;
  out->count = 0;
  uint64_t* offsets = nullptr;
  VSGM_TRY(pool_alloc(ctx, (P + 1) * sizeof(uint64_t), (void**)&offsets));
  Temps t(ctx);
  uint64_t *seg_begin = nullptr, *seg_off = nullptr, W = 0;
  int s = P ? run_setup_scan<OP>(ctx, a, t, &seg_begin, &seg_off, &W) : VSGM_OK;
  if (s != VSGM_OK) {
    pool_free(ctx, offsets);
    return s;
  }
  ctx->stats.work = W;
  uint32_t* values = nullptr;
  uint64_t count = 0;
  if (W == 0) {
    hipLaunchKernelGGL(k_fill, dim3(grid_for((P + 1 + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, offsets, P + 1, 0);
    s = launch_check(ctx, "k_fill");
    if (s == VSGM_OK) s = pool_alloc(ctx, sizeof(uint32_t), (void**)&values);
  } else {
    uint64_t *mask, *chunk_off;
    s = run_mark<OP>(ctx, a, t, W, seg_begin, seg_off, &mask, &chunk_off, &count);
    if (s == VSGM_OK) s = pool_alloc(ctx, std::max<uint64_t>(count, 1) * sizeof(uint32_t), (void**)&values);
    if (s == VSGM_OK) s = run_compact<OP>(ctx, a, values, offsets, ExpandCols{}, t, W, seg_begin, seg_off, mask, chunk_off);
  }
  if (s != VSGM_OK) {
    pool_free(ctx, offsets);
    pool_free(ctx, values);
    return s;
  }
  ctx->stats.outputs = count;
  out->offsets = offsets;
  out->values = values;
  out->count = count;
  return VSGM_OK;
}

}  // namespace

// Context services for the other device sources of this library (mlp.hip).
namespace vsgm_internal {
hipStream_t stream(vsgm_ctx* ctx) { return ctx->stream; }
int fail_msg(vsgm_ctx* ctx, int code, const char* msg) { return fail(ctx, code, "%s", msg); }
}  // namespace vsgm_internal

// ---------------------------------------------------------------------------
// C-ABI

extern "C" {

int vsgm_device_count(int* count) {
  if (!count) return VSGM_ERR_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    n = 0;
  }
  *count = n;
  return VSGM_OK;
}

int vsgm_ctx_create(int device, uint64_t mem_budget, vsgm_ctx** out) {
  if (!out) return VSGM_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return VSGM_ERR_NO_DEVICE;
  }
  if (device < 0 || device >= n) return VSGM_ERR_ARG;
  vsgm_ctx* ctx = new vsgm_ctx();
  ctx->device = device;
  auto bail = [&](int code) {
    vsgm_ctx_destroy(ctx);  // releases whatever was created so far
    return code;
  };
  if (hipSetDevice(device) != hipSuccess) return bail(VSGM_ERR_HIP);
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return bail(VSGM_ERR_HIP);
  for (int i = 0; i < VSGM_TIMER_SLOTS; ++i)
    if (hipEventCreate(&ctx->ev_start[i]) != hipSuccess || hipEventCreate(&ctx->ev_stop[i]) != hipSuccess) return bail(VSGM_ERR_HIP);
  if (hipHostMalloc((void**)&ctx->h_pinned, 64, hipHostMallocDefault) != hipSuccess) return bail(VSGM_ERR_HIP);
  if (hipMalloc((void**)&ctx->d_stats, 4 * sizeof(uint64_t)) != hipSuccess) return bail(VSGM_ERR_HIP);
  if (hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || ctx->n_cu <= 0) return bail(VSGM_ERR_HIP);
  if (hipGetSymbolAddress((void**)&ctx->scratch_ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ctx->scratch_ptr) {
    (void)hipGetLastError();
    return bail(VSGM_ERR_HIP);
  }
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return bail(VSGM_ERR_HIP);
  ctx->budget = mem_budget ? mem_budget : (uint64_t)(0.9 * (double)free_b);
  if (const char* s = getenv("VSGM_PROGRESS")) ctx->progress_s = std::max(0.0, atof(s));
  *out = ctx;
  return VSGM_OK;
}

int vsgm_ctx_destroy(vsgm_ctx* ctx) {
  if (!ctx) return VSGM_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
  for (auto& kv : ctx->cached) (void)hipFree(kv.second);
  for (auto& kv : ctx->blocks) (void)hipFree(kv.first);
  if (ctx->h_pinned) (void)hipHostFree(ctx->h_pinned);
  if (ctx->d_stats) (void)hipFree(ctx->d_stats);
  for (auto& r : ctx->records) {
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  for (auto e : ctx->spare_events) (void)hipEventDestroy(e);
  for (int i = 0; i < VSGM_TIMER_SLOTS; ++i) {
    if (ctx->ev_start[i]) (void)hipEventDestroy(ctx->ev_start[i]);
    if (ctx->ev_stop[i]) (void)hipEventDestroy(ctx->ev_stop[i]);
  }
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return VSGM_OK;
}

int vsgm_ctx_sync(vsgm_ctx* ctx) {
  if (!ctx) return VSGM_ERR_ARG;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return VSGM_OK;
}

const char* vsgm_last_error(const vsgm_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int vsgm_mem_info(const vsgm_ctx* ctx, uint64_t* available, uint64_t* budget, uint64_t* peak) {
  if (!ctx) return VSGM_ERR_ARG;
  if (available) *available = ctx->budget > ctx->live ? ctx->budget - ctx->live : 0;
  if (budget) *budget = ctx->budget;
  if (peak) *peak = ctx->peak;
  return VSGM_OK;
}

int vsgm_ctx_set_budget(vsgm_ctx* ctx, uint64_t mem_budget) {
  if (!ctx) return VSGM_ERR_ARG;
  if (mem_budget < ctx->live)
    return fail(ctx, VSGM_ERR_ARG, "budget %llu is below the %llu bytes live", (unsigned long long)mem_budget, (unsigned long long)ctx->live);
  ctx->budget = mem_budget;
  return VSGM_OK;
}

int vsgm_ctx_bind_thread(const vsgm_ctx* ctx) {
  if (!ctx) return VSGM_ERR_ARG;
  return hipSetDevice(ctx->device) == hipSuccess ? VSGM_OK : VSGM_ERR_HIP;
}

int vsgm_alloc(vsgm_ctx* ctx, uint64_t bytes, void** out) {
  if (!ctx || !out) return VSGM_ERR_ARG;
  return pool_alloc(ctx, bytes, out);
}

int vsgm_free(vsgm_ctx* ctx, void* ptr) {
  if (!ctx) return VSGM_ERR_ARG;
  return pool_free(ctx, ptr);
}

int vsgm_h2d(vsgm_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (!ctx) return VSGM_ERR_ARG;
  if (bytes == 0) return VSGM_OK;
  HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  return VSGM_OK;
}

int vsgm_d2h(vsgm_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (!ctx) return VSGM_ERR_ARG;
  if (bytes == 0) return VSGM_OK;
  HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return VSGM_OK;
}

int vsgm_d2h_async(vsgm_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (!ctx) return VSGM_ERR_ARG;
  if (bytes == 0) return VSGM_OK;
  HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  return VSGM_OK;
}

int vsgm_tape_record(vsgm_ctx* ctx) {
  if (!ctx) return VSGM_ERR_ARG;
  if (ctx->capturing) return fail(ctx, VSGM_ERR_ARG, "tape record during a capture");
  ctx->tape.clear();
  ctx->tape_pos = 0;
  ctx->tape_mode = 1;
  return VSGM_OK;
}

int vsgm_tape_stop(vsgm_ctx* ctx, uint64_t* n_values) {
  if (!ctx) return VSGM_ERR_ARG;
  if (ctx->capturing) return fail(ctx, VSGM_ERR_ARG, "tape stop during a capture");
  ctx->tape_mode = 0;
  if (n_values) *n_values = ctx->tape.size();
  return VSGM_OK;
}

int vsgm_capture_begin(vsgm_ctx* ctx) {
  if (!ctx) return VSGM_ERR_ARG;
  if (ctx->capturing || ctx->tape_mode == 1) return fail(ctx, VSGM_ERR_ARG, "capture already running or tape recording");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->graph_exec) {
    HIP_TRY(ctx, hipGraphExecDestroy(ctx->graph_exec));
    ctx->graph_exec = nullptr;
  }
  // Blocks cached before the capture are referenced by no captured node:
  // return them now, since during the capture (and while its graph lives)
  // the pool cannot hand cached memory back to recover from a failed
  // allocation.
  for (auto& kv : ctx->cached) (void)hipFree(kv.second);
  ctx->cached.clear();
  ctx->cached_bytes = 0;
  ctx->saved_profiling = ctx->profiling;
  ctx->profiling = false;
  ctx->tape_mode = 2;
  ctx->tape_pos = 0;
  hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) {
    ctx->tape_mode = 0;
    ctx->profiling = ctx->saved_profiling;
    return fail(ctx, VSGM_ERR_HIP, "hipStreamBeginCapture failed: %s", hipGetErrorString(e));
  }
  ctx->capturing = true;
  return VSGM_OK;
}

int vsgm_capture_end(vsgm_ctx* ctx, uint64_t* n_nodes) {
  if (!ctx) return VSGM_ERR_ARG;
  if (!ctx->capturing) return fail(ctx, VSGM_ERR_ARG, "no capture running");
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  ctx->capturing = false;
  ctx->profiling = ctx->saved_profiling;
  const bool consumed = ctx->tape_pos == ctx->tape.size();
  ctx->tape_mode = 0;
  if (e != hipSuccess || !g) return fail(ctx, VSGM_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
  size_t n = 0;
  (void)hipGraphGetNodes(g, nullptr, &n);
  if (!consumed) {
    (void)hipGraphDestroy(g);
    return fail(ctx, VSGM_ERR_ARG, "the captured work read %llu of %llu recorded values", (unsigned long long)ctx->tape_pos,
                (unsigned long long)ctx->tape.size());
  }
  e = hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    ctx->graph_exec = nullptr;
    return fail(ctx, VSGM_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
  }
  ctx->graph_nodes = n;
  if (n_nodes) *n_nodes = n;
  return VSGM_OK;
}

int vsgm_graph_launch(vsgm_ctx* ctx) {
  if (!ctx) return VSGM_ERR_ARG;
  if (!ctx->graph_exec) return fail(ctx, VSGM_ERR_ARG, "no captured graph");
  HIP_TRY(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
  return VSGM_OK;
}

int vsgm_graph_release(vsgm_ctx* ctx) {
  if (!ctx) return VSGM_ERR_ARG;
  if (ctx->graph_exec) {
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipGraphExecDestroy(ctx->graph_exec));
    ctx->graph_exec = nullptr;
    ctx->graph_nodes = 0;
  }
  return VSGM_OK;
}

int vsgm_host_alloc(uint64_t bytes, void** out) {
  if (!out) return VSGM_ERR_ARG;
  if (hipHostMalloc(out, std::max<uint64_t>(bytes, 1), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    return VSGM_ERR_OOM;
  }
  return VSGM_OK;
}

int vsgm_host_free(void* ptr) {
  if (ptr) (void)hipHostFree(ptr);
  return VSGM_OK;
}

int vsgm_timer_start(vsgm_ctx* ctx, int slot) {
  if (!ctx || slot < 0 || slot >= VSGM_TIMER_SLOTS) return VSGM_ERR_ARG;
  HIP_TRY(ctx, hipEventRecord(ctx->ev_start[slot], ctx->stream));
  return VSGM_OK;
}

int vsgm_timer_stop(vsgm_ctx* ctx, int slot, float* ms) {
  if (!ctx || !ms || slot < 0 || slot >= VSGM_TIMER_SLOTS) return VSGM_ERR_ARG;
  HIP_TRY(ctx, hipEventRecord(ctx->ev_stop[slot], ctx->stream));
  HIP_TRY(ctx, hipEventSynchronize(ctx->ev_stop[slot]));
  HIP_TRY(ctx, hipEventElapsedTime(ms, ctx->ev_start[slot], ctx->ev_stop[slot]));
  return VSGM_OK;
}

int vsgm_scan_excl_u64(vsgm_ctx* ctx, const uint64_t* d_in, uint64_t* d_out, uint64_t n) {
  if (!ctx || !d_out || (n && !d_in)) return VSGM_ERR_ARG;
  return scan_impl(ctx, d_in, d_out, n);
}

int vsgm_reduce_u64(vsgm_ctx* ctx, const uint64_t* d_in, uint64_t n, uint64_t* d_total) {
  if (!ctx || !d_total || (n && !d_in)) return VSGM_ERR_ARG;
  if (n == 0) return VSGM_OK;
  hipLaunchKernelGGL(k_reduce, dim3(grid_for((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, d_in, n,
                     (unsigned long long*)d_total);
  return launch_check(ctx, "k_reduce");
}

int vsgm_gather_u32(vsgm_ctx* ctx, const uint64_t* d_idx, uint64_t n, const uint32_t* d_src, uint32_t* d_dst) {
  if (!ctx) return VSGM_ERR_ARG;
  if (n == 0) return VSGM_OK;
  hipLaunchKernelGGL(k_gather<uint32_t>, dim3(grid_for((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, d_idx, n, d_src, d_dst);
  return launch_check(ctx, "k_gather");
}

int vsgm_gather_u64(vsgm_ctx* ctx, const uint64_t* d_idx, uint64_t n, const uint64_t* d_src, uint64_t* d_dst) {
  if (!ctx) return VSGM_ERR_ARG;
  if (n == 0) return VSGM_OK;
  hipLaunchKernelGGL(k_gather<uint64_t>, dim3(grid_for((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, d_idx, n, d_src, d_dst);
  return launch_check(ctx, "k_gather");
}

int vsgm_iota_u64(vsgm_ctx* ctx, uint64_t* d, uint64_t n, uint64_t base) {
  if (!ctx) return VSGM_ERR_ARG;
  if (n == 0) return VSGM_OK;
  hipLaunchKernelGGL(k_iota, dim3(grid_for((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, d, n, base);
  return launch_check(ctx, "k_iota");
}

int vsgm_fill_u64(vsgm_ctx* ctx, uint64_t* d, uint64_t n, uint64_t value) {
  if (!ctx) return VSGM_ERR_ARG;
  if (n == 0) return VSGM_OK;
  hipLaunchKernelGGL(k_fill, dim3(grid_for((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, d, n, value);
  return launch_check(ctx, "k_fill");
}

int vsgm_hub_index_build(vsgm_ctx* ctx, const uint64_t* d_row_ptrs, const uint32_t* d_cols, uint64_t vertex_count, const uint32_t* h_hubs,
                         uint32_t n_hubs, vsgm_hub_index* out) {
  if (!ctx) return VSGM_ERR_ARG;
  if (!out || !d_row_ptrs || !d_cols || (n_hubs && !h_hubs)) return fail(ctx, VSGM_ERR_ARG, "null hub index argument");
  if (vertex_count == 0 || vertex_count > 0xFFFFFFFFull) return fail(ctx, VSGM_ERR_ARG, "vertex count out of range");
  *out = vsgm_hub_index{nullptr, nullptr, 0, 0};
  {
    std::vector<bool> seen(vertex_count, false);
    for (uint32_t i = 0; i < n_hubs; ++i) {
      if (h_hubs[i] >= vertex_count || seen[h_hubs[i]]) return fail(ctx, VSGM_ERR_ARG, "hub %u is out of range or repeated", h_hubs[i]);
      seen[h_hubs[i]] = true;
    }
  }
  const uint64_t slot_pairs = (vertex_count + 1) / 2;          // u64 words holding the u32 slots
  const uint64_t words = ((vertex_count + 63) / 64) * 2;       // u32 words per row, even
  uint32_t *slot = nullptr, *bits = nullptr, *d_hubs = nullptr;
  VSGM_TRY(pool_alloc(ctx, slot_pairs * sizeof(uint64_t), (void**)&slot));
  auto undo = [&](int code) {
    pool_free(ctx, slot);
    pool_free(ctx, bits);
    pool_free(ctx, d_hubs);
    return code;
  };
  int s = pool_alloc(ctx, std::max<uint64_t>((uint64_t)n_hubs * words, 2) * sizeof(uint32_t), (void**)&bits);
  if (s == VSGM_OK) s = pool_alloc(ctx, std::max<uint32_t>(n_hubs, 1) * sizeof(uint32_t), (void**)&d_hubs);
  if (s != VSGM_OK) return undo(s);
  hipLaunchKernelGGL(k_fill, dim3(grid_for((slot_pairs + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, (uint64_t*)slot, slot_pairs,
                     ~0ull);
  if ((s = launch_check(ctx, "k_fill")) != VSGM_OK) return undo(s);
  if (n_hubs) {
    const uint64_t bit_pairs = (uint64_t)n_hubs * words / 2;
    hipLaunchKernelGGL(k_fill, dim3(grid_for((bit_pairs + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, (uint64_t*)bits, bit_pairs,
                       0ull);
    if ((s = launch_check(ctx, "k_fill")) != VSGM_OK) return undo(s);
    if (hipMemcpyAsync(d_hubs, h_hubs, (uint64_t)n_hubs * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
      return undo(fail(ctx, VSGM_ERR_HIP, "hub list upload failed"));
    hipLaunchKernelGGL(k_hub_slots, dim3(grid_for((n_hubs + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, d_hubs, n_hubs, slot);
    if ((s = launch_check(ctx, "k_hub_slots")) != VSGM_OK) return undo(s);
    hipLaunchKernelGGL(k_hub_bits, dim3(std::min<uint32_t>(n_hubs, (uint32_t)kMaxGrid)), dim3(kBlock), 0, ctx->stream, d_row_ptrs, d_cols,
                       vertex_count, d_hubs, n_hubs, bits, words);
    if ((s = launch_check(ctx, "k_hub_bits")) != VSGM_OK) return undo(s);
  }
  // the host hub list may be released once this returns
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return undo(fail(ctx, VSGM_ERR_HIP, "hub index build failed"));
  pool_free(ctx, d_hubs);
  out->slot = slot;
  out->bits = bits;
  out->words = words;
  out->n_hubs = n_hubs;
  return VSGM_OK;
}

int vsgm_view_build(vsgm_ctx* ctx, const uint64_t* d_row_ptrs, const uint32_t* d_cols, uint64_t vertex_count, const uint32_t* d_sources,
                    uint64_t n_sources, int hop, uint64_t root_degree, vsgm_view_out* out) {
  if (!ctx) return VSGM_ERR_ARG;
  if (!out || !d_row_ptrs || !d_cols || (n_sources && !d_sources)) return fail(ctx, VSGM_ERR_ARG, "null view argument");
  if (hop < 1 || hop > 250) return fail(ctx, VSGM_ERR_ARG, "hop out of range");
  if (vertex_count == 0 || vertex_count > 0xFFFFFFFFull) return fail(ctx, VSGM_ERR_ARG, "vertex count out of range");
  *out = vsgm_view_out{nullptr, nullptr, 0, 0};
  const uint64_t V = vertex_count;
  Temps t(ctx);
  uint8_t* dist;
  uint64_t* vrp = nullptr;
  uint32_t* queue;
  uint64_t* qword;  // the queue's counter, a u64 word cleared by k_fill
  VSGM_TRY(t.get((V + 7) / 8 * 8, &dist));
  VSGM_TRY(t.get(kViewQueue, &queue));
  VSGM_TRY(t.get(1, &qword));
  unsigned int* qn = (unsigned int*)qword;
  VSGM_TRY(pool_alloc(ctx, (V + 1) * sizeof(uint64_t), (void**)&vrp));
  auto undo = [&](int code) {
    pool_free(ctx, vrp);
    return code;
  };
  int s;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(((V + 7) / 8 + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, (uint64_t*)dist, (V + 7) / 8,
                     ~0ull);
  if ((s = launch_check(ctx, "k_fill")) != VSGM_OK) return undo(s);
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->d_stats, 2, 0);
  if ((s = launch_check(ctx, "k_fill")) != VSGM_OK) return undo(s);
  if (n_sources) {
    hipLaunchKernelGGL(k_view_seed, dim3(grid_for((n_sources + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, d_row_ptrs, d_sources,
                       n_sources, V, root_degree, dist);
    if ((s = launch_check(ctx, "k_view_seed")) != VSGM_OK) return undo(s);
  }
  const unsigned wave_grid = grid_for((V + kBlock / kWave - 1) / (kBlock / kWave));
  auto clear_queue = [&]() {
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(kBlock), 0, ctx->stream, qword, 1, 0);
    return launch_check(ctx, "k_fill");
  };
  for (int level = 1; level < hop; ++level) {
    if ((s = clear_queue()) != VSGM_OK) return undo(s);
    hipLaunchKernelGGL(k_view_expand, dim3(wave_grid), dim3(kBlock), 0, ctx->stream, d_row_ptrs, d_cols, V, (uint8_t)level, dist, queue, qn);
    if ((s = launch_check(ctx, "k_view_expand")) != VSGM_OK) return undo(s);
    hipLaunchKernelGGL(k_view_big<false>, dim3(kViewBigGrid), dim3(kBlock), 0, ctx->stream, d_row_ptrs, d_cols, queue, qn, (uint8_t)level,
                       dist, (const uint64_t*)nullptr, (uint32_t*)nullptr);
    if ((s = launch_check(ctx, "k_view_big")) != VSGM_OK) return undo(s);
  }
  hipLaunchKernelGGL(k_view_len, dim3(grid_for((V + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, d_row_ptrs, dist, V, vrp,
                     (unsigned long long*)(ctx->d_stats + 1));
  if ((s = launch_check(ctx, "k_view_len")) != VSGM_OK) return undo(s);
  if ((s = scan_impl(ctx, vrp, vrp, V)) != VSGM_OK) return undo(s);
  hipLaunchKernelGGL(k_copy1, dim3(1), dim3(1), 0, ctx->stream, vrp + V, ctx->d_stats);
  if ((s = launch_check(ctx, "k_copy1")) != VSGM_OK) return undo(s);
  uint64_t vals[2];
  if ((s = host_read(ctx, ctx->d_stats, vals, 2)) != VSGM_OK) return undo(s);
  uint32_t* vcols = nullptr;
  if ((s = pool_alloc(ctx, std::max<uint64_t>(vals[0], 1) * sizeof(uint32_t), (void**)&vcols)) != VSGM_OK) return undo(s);
  if ((s = clear_queue()) != VSGM_OK) {
    pool_free(ctx, vcols);
    return undo(s);
  }
  hipLaunchKernelGGL(k_view_copy, dim3(wave_grid), dim3(kBlock), 0, ctx->stream, d_row_ptrs, d_cols, dist, V, vrp, vcols, queue, qn);
  if ((s = launch_check(ctx, "k_view_copy")) == VSGM_OK) {
    hipLaunchKernelGGL(k_view_big<true>, dim3(kViewBigGrid), dim3(kBlock), 0, ctx->stream, d_row_ptrs, d_cols, queue, qn, (uint8_t)0, dist,
                       vrp, vcols);
    s = launch_check(ctx, "k_view_big");
  }
  if (s != VSGM_OK) {
    pool_free(ctx, vcols);
    return undo(s);
  }
  out->row_ptrs = vrp;
  out->cols = vcols;
  out->entries = vals[0];
  out->vertices = vals[1];
  return VSGM_OK;
}

int vsgm_view_self_loops(vsgm_ctx* ctx, const uint64_t* d_row_ptrs, const uint32_t* d_cols, uint64_t vertex_count, uint64_t* count) {
  if (!ctx) return VSGM_ERR_ARG;
  if (!count || (vertex_count && (!d_row_ptrs || !d_cols))) return fail(ctx, VSGM_ERR_ARG, "null view or output");
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->d_stats, 1, 0);
  VSGM_TRY(launch_check(ctx, "k_fill"));
  if (vertex_count) {
    hipLaunchKernelGGL(k_self_loops, dim3(grid_for((vertex_count + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, d_row_ptrs, d_cols,
                       vertex_count, (unsigned long long*)ctx->d_stats);
    VSGM_TRY(launch_check(ctx, "k_self_loops"));
  }
  return read_u64(ctx, ctx->d_stats, count);
}

int vsgm_hub_index_free(vsgm_ctx* ctx, vsgm_hub_index* index) {
  if (!ctx || !index) return VSGM_ERR_ARG;
  VSGM_TRY(pool_free(ctx, index->slot));
  VSGM_TRY(pool_free(ctx, index->bits));
  *index = vsgm_hub_index{nullptr, nullptr, 0, 0};
  return VSGM_OK;
}

int vsgm_intersect_write(vsgm_ctx* ctx, const vsgm_level_args* a, vsgm_csr_out* out) {
  VSGM_TRY(check_args(ctx, a));
  if (!out) return fail(ctx, VSGM_ERR_ARG, "null output");
  VSGM_TRY(check_intersect(ctx, a));
  return run_csr<OP_IWRITE>(ctx, *a, out);
}

int vsgm_intersect_count_total(vsgm_ctx* ctx, const vsgm_level_args* a, uint64_t* d_total) {
  VSGM_TRY(check_args(ctx, a));
  if (!d_total) return fail(ctx, VSGM_ERR_ARG, "null total");
  VSGM_TRY(check_intersect(ctx, a));
  if (a->other >= 0) VSGM_TRY(check_cand(ctx, a, (uint32_t)a->other));
  return run_count<OP_ICOUNT>(ctx, *a, d_total, nullptr);
}

int vsgm_intersect_count_total_noalloc(vsgm_ctx* ctx, const vsgm_level_args* a, uint64_t* d_total) {
  VSGM_TRY(check_args(ctx, a));
  if (!d_total) return fail(ctx, VSGM_ERR_ARG, "null total");
  VSGM_TRY(check_intersect(ctx, a));
  if (a->other >= 0) VSGM_TRY(check_cand(ctx, a, (uint32_t)a->other));
  if (a->path_num > VSGM_NOALLOC_MAX_PATHS)
    return fail(ctx, VSGM_ERR_ARG, "path_num %llu exceeds the allocation-free limit of %llu paths", (unsigned long long)a->path_num,
                (unsigned long long)VSGM_NOALLOC_MAX_PATHS);
  if (ctx->capturing) return fail(ctx, VSGM_ERR_ARG, "the allocation-free count is not available during graph capture");
  // no profiling events either: the call creates nothing
  const bool prof = ctx->profiling;
  ctx->profiling = false;
  ctx->arena_on = true;
  ctx->arena_off = 0;
  const int s = run_count<OP_ICOUNT>(ctx, *a, d_total, nullptr);
  ctx->arena_on = false;
  ctx->profiling = prof;
  return s;
}

int vsgm_intersect_count_paths(vsgm_ctx* ctx, const vsgm_level_args* a, uint64_t* d_counts) {
  VSGM_TRY(check_args(ctx, a));
  if (!d_counts && a->path_num) return fail(ctx, VSGM_ERR_ARG, "null counts");
  VSGM_TRY(check_intersect(ctx, a));
  VSGM_TRY(vsgm_fill_u64(ctx, d_counts, a->path_num, 0));
  return run_count<OP_IPATH>(ctx, *a, nullptr, d_counts);
}

int vsgm_expand(vsgm_ctx* ctx, const vsgm_level_args* a, const uint32_t* mat, uint32_t n_mat, const uint32_t* cu, uint32_t n_cu,
                vsgm_expand_out* out) {
  VSGM_TRY(check_args(ctx, a));
  if (!out || n_mat > VSGM_MAX_QV || n_cu > VSGM_MAX_QV) return fail(ctx, VSGM_ERR_ARG, "bad expand outputs");
  VSGM_TRY(check_cand(ctx, a, a->target));
  ExpandCols cols{};
  cols.n_mat = n_mat;
  cols.n_cu = n_cu;
  for (uint32_t i = 0; i < n_mat; ++i) {
    if (mat[i] >= VSGM_MAX_QV || !a->inst[mat[i]] || mat[i] == a->target) return fail(ctx, VSGM_ERR_ARG, "bad materialized column");
    cols.mat[i] = mat[i];
  }
  for (uint32_t i = 0; i < n_cu; ++i) {
    if (cu[i] >= VSGM_MAX_QV || cu[i] == a->target) return fail(ctx, VSGM_ERR_ARG, "bad candidate column");
    cols.cu[i] = cu[i];
  }
  memset(out, 0, sizeof(*out));
  const uint64_t P = a->path_num;
  ctx->stats = {P, 0, 0, 0};
  Temps t(ctx);
  uint64_t *seg_begin = nullptr, *seg_off = nullptr, W = 0;
  if (P) VSGM_TRY(run_setup_scan<OP_EXPAND>(ctx, *a, t, &seg_begin, &seg_off, &W));
  ctx->stats.work = W;
  uint64_t count = 0, *mask = nullptr, *chunk_off = nullptr;
  if (W) VSGM_TRY(run_mark<OP_EXPAND>(ctx, *a, t, W, seg_begin, seg_off, &mask, &chunk_off, &count));
  ctx->stats.outputs = count;
  // output columns
  std::vector<void*> made;
  auto undo = [&](int code) {
    for (void* p : made) pool_free(ctx, p);
    memset(out, 0, sizeof(*out));
    return code;
  };
  const uint64_t n_alloc = std::max<uint64_t>(count, 1);
  int s = pool_alloc(ctx, n_alloc * sizeof(uint32_t), (void**)&cols.out_inst[a->target]);
  if (s != VSGM_OK) return undo(s);
  made.push_back(cols.out_inst[a->target]);
  for (uint32_t i = 0; i < n_mat; ++i) {
    s = pool_alloc(ctx, n_alloc * sizeof(uint32_t), (void**)&cols.out_inst[mat[i]]);
    if (s != VSGM_OK) return undo(s);
    made.push_back(cols.out_inst[mat[i]]);
  }
  for (uint32_t i = 0; i < n_cu; ++i) {
    s = pool_alloc(ctx, n_alloc * sizeof(uint64_t), (void**)&cols.out_idx[cu[i]]);
    if (s != VSGM_OK) return undo(s);
    made.push_back(cols.out_idx[cu[i]]);
  }
  if (count) {
    s = run_compact<OP_EXPAND>(ctx, *a, nullptr, nullptr, cols, t, W, seg_begin, seg_off, mask, chunk_off);
    if (s != VSGM_OK) return undo(s);
  }
  out->path_num = count;
  for (int u = 0; u < VSGM_MAX_QV; ++u) {
    out->inst[u] = cols.out_inst[u];
    out->cand_idx[u] = cols.out_idx[u];
  }
  return VSGM_OK;
}

int vsgm_filter(vsgm_ctx* ctx, const vsgm_level_args* a, vsgm_csr_out* out) {
  VSGM_TRY(check_args(ctx, a));
  if (!out) return fail(ctx, VSGM_ERR_ARG, "null output");
  VSGM_TRY(check_cand(ctx, a, a->target));
  return run_csr<OP_FILTER>(ctx, *a, out);
}

int vsgm_count_sets(vsgm_ctx* ctx, const vsgm_level_args* a, uint64_t* d_total) {
  VSGM_TRY(check_args(ctx, a));
  if (!d_total) return fail(ctx, VSGM_ERR_ARG, "null total");
  VSGM_TRY(check_cand(ctx, a, a->target));
  if (a->other < 0) {
    ctx->stats = {a->path_num, a->path_num, 0, 0};
    if (a->path_num == 0) return VSGM_OK;
    hipLaunchKernelGGL(k_count1, dim3(grid_for((a->path_num + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, *a,
                       (unsigned long long*)d_total);
    return launch_check(ctx, "k_count1");
  }
  VSGM_TRY(check_cand(ctx, a, (uint32_t)a->other));
  return run_count<OP_COUNT2>(ctx, *a, d_total, nullptr);
}

int vsgm_children_count(vsgm_ctx* ctx, const vsgm_level_args* a, int mode, uint64_t* d_out) {
  VSGM_TRY(check_args(ctx, a));
  if (mode < 0 || mode > 2 || (!d_out && a->path_num)) return fail(ctx, VSGM_ERR_ARG, "bad children-count mode");
  if (mode == 0) VSGM_TRY(check_intersect(ctx, a));
  if (mode == 1) VSGM_TRY(check_cand(ctx, a, a->target));
  if (a->path_num == 0) return VSGM_OK;
  hipLaunchKernelGGL(k_children, dim3(grid_for((a->path_num + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, *a, mode, d_out);
  return launch_check(ctx, "k_children");
}

int vsgm_split_batch(vsgm_ctx* ctx, const uint64_t* d_children, uint64_t n, uint64_t parent_factor, uint64_t children_factor,
                     uint64_t batch_bytes, uint64_t* h_ends, uint64_t max_batches, uint64_t* n_batches) {
  if (!ctx || !n_batches || batch_bytes == 0) return VSGM_ERR_ARG;
  *n_batches = 0;
  if (n == 0) return VSGM_OK;
  Temps t(ctx);
  uint64_t* prefix;
  VSGM_TRY(t.get(n + 1, &prefix));
  hipLaunchKernelGGL(k_cost, dim3(grid_for((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, d_children, n, parent_factor,
                     children_factor, prefix);
  VSGM_TRY(launch_check(ctx, "k_cost"));
  VSGM_TRY(scan_impl(ctx, prefix, prefix, n));
  uint64_t total;
  VSGM_TRY(read_u64(ctx, prefix + n, &total));
  uint64_t nb = (total + batch_bytes - 1) / batch_bytes;
  std::vector<uint64_t> ends;
  if (nb > 1) {
    uint64_t* d_ends;
    VSGM_TRY(t.get(nb - 1, &d_ends));
    hipLaunchKernelGGL(k_split, dim3(grid_for((nb - 1 + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream, prefix, n, batch_bytes, nb - 1,
                       d_ends);
    VSGM_TRY(launch_check(ctx, "k_split"));
    std::vector<uint64_t> raw(nb - 1);
    VSGM_TRY(host_read(ctx, d_ends, raw.data(), nb - 1));
    uint64_t prev = 0;
    for (uint64_t e : raw) {
      if (e <= prev) e = prev + 1;  // every batch keeps at least one parent
      if (e >= n) break;
      ends.push_back(e);
      prev = e;
    }
  }
  ends.push_back(n);
  *n_batches = ends.size();
  if (!h_ends || ends.size() > max_batches) return fail(ctx, VSGM_ERR_ARG, "split needs %llu batch slots", (unsigned long long)ends.size());
  std::copy(ends.begin(), ends.end(), h_ends);
  return VSGM_OK;
}

int vsgm_set_tuning(vsgm_ctx* ctx, const vsgm_tuning* t) {
  if (!ctx || !t) return VSGM_ERR_ARG;
  if (t->isect_kernel > 4 || t->max_grid == 0 || t->max_grid > (1u << 20)) return fail(ctx, VSGM_ERR_ARG, "bad tuning");
  if (t->cold_weight == 0 || t->cold_weight > 1024 || t->ticket_chunks == 0 || t->ticket_chunks > (uint32_t)kTicketCap)
    return fail(ctx, VSGM_ERR_ARG, "bad tuning: cold_weight in [1, 1024], ticket_chunks in [1, %d]", kTicketCap);
  ctx->tuning = *t;
  return VSGM_OK;
}

int vsgm_get_tuning(const vsgm_ctx* ctx, vsgm_tuning* t) {
  if (!ctx || !t) return VSGM_ERR_ARG;
  *t = ctx->tuning;
  return VSGM_OK;
}

int vsgm_profile_enable(vsgm_ctx* ctx, int enable) {
  if (!ctx) return VSGM_ERR_ARG;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  for (auto& r : ctx->records) {
    ctx->spare_events.push_back(r.start);
    ctx->spare_events.push_back(r.stop);
  }
  ctx->records.clear();
  ctx->profiling = enable != 0;
  return VSGM_OK;
}

int vsgm_profile_collect(vsgm_ctx* ctx, vsgm_kernel_stat* out, int cap, int* n_out) {
  if (!ctx || !n_out) return VSGM_ERR_ARG;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<vsgm_kernel_stat> agg;
  for (auto& r : ctx->records) {
    float ms = 0;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, r.start, r.stop));
    auto it = std::find_if(agg.begin(), agg.end(), [&](const vsgm_kernel_stat& s) { return strcmp(s.name, r.name) == 0; });
    if (it == agg.end()) {
      vsgm_kernel_stat s;
      memset(&s, 0, sizeof(s));
      strncpy(s.name, r.name, sizeof(s.name) - 1);
      agg.push_back(s);
      it = agg.end() - 1;
    }
    it->launches += 1;
    it->total_ms += ms;
    it->max_ms = std::max(it->max_ms, (double)ms);
    it->algo_bytes += r.bytes;
    it->work += r.work;
    it->paths += r.paths;
  }
  *n_out = (int)agg.size();
  for (int i = 0; i < (int)agg.size() && i < cap; ++i) out[i] = agg[i];
  return VSGM_OK;
}

int vsgm_last_op_stats(const vsgm_ctx* ctx, vsgm_op_stats* out) {
  if (!ctx || !out) return VSGM_ERR_ARG;
  *out = ctx->stats;
  return VSGM_OK;
}

}  // extern "C"
