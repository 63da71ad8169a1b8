// k-means partitioning of the vertices (include/vsgm_partition.h) with
// gfx950 kernels. Each pass is a sweep over the CSR: one wavefront per
// vertex, lanes striding its list. Centroids are double [V][K]: the K
// values a list entry touches share one cache line, so an assignment pass
// reads each list once and one line per entry. On symmetric graphs the
// centroid update gathers (dimension w sums its closed neighbours' shares,
// one wavefront per w, fixed order) and the column norms are folded from
// per-wave partials in wave order, so a seed fixes the partition bit for
// bit; other graphs scatter 1/sqrt(|N[v]|) into the member's column with
// hardware float64 atomics. Both then normalise by cluster size.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "vsgm_hip.h"
#include "vsgm_partition.h"

namespace vsgm_internal {
hipStream_t stream(vsgm_ctx* ctx);
int fail_msg(vsgm_ctx* ctx, int code, const char* msg);
}  // namespace vsgm_internal

namespace {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr unsigned kMaxGrid = 8192;

unsigned grid_for(uint64_t units) { return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(units, kMaxGrid)); }
unsigned wave_grid(uint64_t V) { return grid_for((V + kWavesPerBlock - 1) / kWavesPerBlock); }

__device__ inline double wave_sum_d(double v) {
  for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

// Sum over the closed neighbourhood of v of cent[w][0..K) (lane 0 returns
// the totals).
template <int K>
__device__ inline void neighbourhood_sums(const uint64_t* __restrict__ rp, const uint32_t* __restrict__ cols, const double* __restrict__ cent,
                                          uint64_t v, int lane, double (&acc)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (uint64_t j = rp[v] + lane; j < rp[v + 1]; j += kWave) {
    const uint64_t w = cols[j];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] += cent[w * K + k];
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] += cent[v * K + k];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = wave_sum_d(acc[k]);
}

// assign[v] = argmax_k sim, first maximum wins; sim[v] = its value.
template <int K>
__global__ __launch_bounds__(kBlock) void k_km_assign(const uint64_t* __restrict__ rp, const uint32_t* __restrict__ cols, uint64_t V,
                                                      const double* __restrict__ cent, const double* __restrict__ norm,
                                                      int32_t* __restrict__ assign, double* __restrict__ sim) {
  const int lane = threadIdx.x & (kWave - 1);
  for (uint64_t v = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; v < V; v += (uint64_t)gridDim.x * kWavesPerBlock) {
    double acc[K];
    neighbourhood_sums<K>(rp, cols, cent, v, lane, acc);
    if (lane == 0) {
      const double rs = sqrt(1.0 / (double)(rp[v + 1] - rp[v] + 1));
      int best = 0;
      double bs = acc[0] * rs / norm[0];
#pragma unroll
      for (int k = 1; k < K; ++k) {
        const double s = acc[k] * rs / norm[k];
        if (s > bs) {
          bs = s;
          best = k;
        }
      }
      assign[v] = best;
      sim[v] = bs;
    }
  }
}

// Initialisation step: maxsim[v] = max(maxsim[v], sim(v, column col)), the
// vertex chosen last counts as infinitely similar (never chosen again).
template <int K>
__global__ __launch_bounds__(kBlock) void k_km_pp(const uint64_t* __restrict__ rp, const uint32_t* __restrict__ cols, uint64_t V,
                                                  const double* __restrict__ cent, const double* __restrict__ norm, int col, uint64_t last,
                                                  double* __restrict__ maxsim) {
  const int lane = threadIdx.x & (kWave - 1);
  for (uint64_t v = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; v < V; v += (uint64_t)gridDim.x * kWavesPerBlock) {
    double a = 0.0;
    for (uint64_t j = rp[v] + lane; j < rp[v + 1]; j += kWave) a += cent[(uint64_t)cols[j] * K + col];
    if (lane == 0) a += cent[v * K + col];
    a = wave_sum_d(a);
    if (lane == 0) {
      const double s = v == last ? DBL_MAX : a * sqrt(1.0 / (double)(rp[v + 1] - rp[v] + 1)) / norm[col];
      if (s > maxsim[v]) maxsim[v] = s;
    }
  }
}

// Column col := feature of vertex u (1/sqrt(|N[u]|) on N[u]).
template <int K>
__global__ void k_km_set(const uint64_t* __restrict__ rp, const uint32_t* __restrict__ cols, uint64_t u, int col, double* __restrict__ cent) {
  const double r = sqrt(1.0 / (double)(rp[u + 1] - rp[u] + 1));
  for (uint64_t j = rp[u] + threadIdx.x; j < rp[u + 1]; j += blockDim.x) cent[(uint64_t)cols[j] * K + col] = r;
  if (threadIdx.x == 0) cent[u * K + col] = r;
}

// Centroid update, pass 1: every member adds its feature to its column.
template <int K>
__global__ __launch_bounds__(kBlock) void k_km_scatter(const uint64_t* __restrict__ rp, const uint32_t* __restrict__ cols, uint64_t V,
                                                       const int32_t* __restrict__ assign, double* __restrict__ cent) {
  const int lane = threadIdx.x & (kWave - 1);
  for (uint64_t v = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; v < V; v += (uint64_t)gridDim.x * kWavesPerBlock) {
    const int k = assign[v];
    if (k < 0 || k >= K) continue;  // a caller's assignment out of range joins no cluster
    const double r = sqrt(1.0 / (double)(rp[v + 1] - rp[v] + 1));
    for (uint64_t j = rp[v] + lane; j < rp[v + 1]; j += kWave) unsafeAtomicAdd(&cent[(uint64_t)cols[j] * K + k], r);
    if (lane == 0) unsafeAtomicAdd(&cent[v * K + k], r);
  }
}

__global__ __launch_bounds__(kBlock) void k_km_sizes(const int32_t* __restrict__ assign, uint64_t V, int K, unsigned long long* __restrict__ sizes) {
  __shared__ unsigned long long s_hist[VSGM_KMEANS_MAX_K];
  if (threadIdx.x < VSGM_KMEANS_MAX_K) s_hist[threadIdx.x] = 0;
  __syncthreads();
  for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < V; v += (uint64_t)gridDim.x * kBlock) {
    const int k = assign[v];
    if (k >= 0 && k < K) atomicAdd(&s_hist[k], 1ull);
  }
  __syncthreads();
  if (threadIdx.x < (unsigned)K && s_hist[threadIdx.x]) atomicAdd(&sizes[threadIdx.x], s_hist[threadIdx.x]);
}

// Pass 2: divide by the cluster size; each wave writes its share of the
// squared column norms to partials[wave][K] (folded by k_km_fold, so the
// norms do not depend on the order waves finish in).
template <int K>
__global__ __launch_bounds__(kBlock) void k_km_finish(double* __restrict__ cent, uint64_t V, const unsigned long long* __restrict__ sizes,
                                                      double* __restrict__ partials) {
  double part[K];
#pragma unroll
  for (int k = 0; k < K; ++k) part[k] = 0.0;
  for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < V; v += (uint64_t)gridDim.x * kBlock) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double c = cent[v * K + k];
      if (sizes[k]) c /= (double)sizes[k];
      cent[v * K + k] = c;
      part[k] += c * c;
    }
  }
  const uint64_t wave = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double s = wave_sum_d(part[k]);
    if ((threadIdx.x & (kWave - 1)) == 0) partials[wave * K + k] = s;
  }
}

template <int K>
__global__ __launch_bounds__(kBlock) void k_km_colsq(const double* __restrict__ cent, uint64_t V, double* __restrict__ partials) {
  double part[K];
#pragma unroll
  for (int k = 0; k < K; ++k) part[k] = 0.0;
  for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < V; v += (uint64_t)gridDim.x * kBlock)
#pragma unroll
    for (int k = 0; k < K; ++k) part[k] += cent[v * K + k] * cent[v * K + k];
  const uint64_t wave = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double s = wave_sum_d(part[k]);
    if ((threadIdx.x & (kWave - 1)) == 0) partials[wave * K + k] = s;
  }
}

// out[k] = sqrt(sum over waves of partials[wave][k]), summed in wave order.
__global__ void k_km_fold(const double* __restrict__ partials, uint64_t waves, int K, double* __restrict__ out) {
  const int k = threadIdx.x;
  if (k >= K) return;
  double s = 0.0;
  for (uint64_t w = 0; w < waves; ++w) s += partials[w * K + k];
  out[k] = sqrt(s);
}

// Centroid update for symmetric graphs as a gather: dimension w of
// centroid k sums 1/sqrt(|N[v]|) over the members v of k in N[w] (the
// vertices whose closed neighbourhood holds w), one wavefront per w, in a
// fixed order; writes every entry, so no atomics and no zeroing.
template <int K>
__global__ __launch_bounds__(kBlock) void k_km_gather_cent(const uint64_t* __restrict__ rp, const uint32_t* __restrict__ cols, uint64_t V,
                                                           const int32_t* __restrict__ assign, double* __restrict__ cent) {
  const int lane = threadIdx.x & (kWave - 1);
  for (uint64_t w = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; w < V; w += (uint64_t)gridDim.x * kWavesPerBlock) {
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    auto add = [&](uint64_t v) {
      const int k = assign[v];
      const double r = sqrt(1.0 / (double)(rp[v + 1] - rp[v] + 1));
#pragma unroll
      for (int kk = 0; kk < K; ++kk) acc[kk] += kk == k ? r : 0.0;
    };
    for (uint64_t j = rp[w] + lane; j < rp[w + 1]; j += kWave) add(cols[j]);
    if (lane == 0) add(w);
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = wave_sum_d(acc[k]);
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < K; ++k) cent[w * K + k] = acc[k];
    }
  }
}

// Counts list entries (v, w) whose reverse entry (w, v) is missing; rows
// must be sorted (an unsorted row only makes the graph look asymmetric).
__global__ __launch_bounds__(kBlock) void k_km_asym(const uint64_t* __restrict__ rp, const uint32_t* __restrict__ cols, uint64_t V,
                                                    unsigned long long* __restrict__ missing) {
  const int lane = threadIdx.x & (kWave - 1);
  unsigned long long local = 0;
  for (uint64_t v = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; v < V; v += (uint64_t)gridDim.x * kWavesPerBlock)
    for (uint64_t j = rp[v] + lane; j < rp[v + 1]; j += kWave) {
      const uint64_t w = cols[j];
      uint64_t lo = rp[w], hi = rp[w + 1];
      while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if (cols[mid] < v)
          lo = mid + 1;
        else
          hi = mid;
      }
      if (lo == rp[w + 1] || cols[lo] != v) ++local;
    }
  for (int o = kWave / 2; o > 0; o >>= 1) local += __shfl_xor(local, o);
  if (lane == 0 && local) atomicAdd(missing, local);
}

__global__ __launch_bounds__(kBlock) void k_zero_d(double* d, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) d[i] = 0.0;
}

__global__ __launch_bounds__(kBlock) void k_fill_d(double* d, uint64_t n, double x) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) d[i] = x;
}

__global__ __launch_bounds__(kBlock) void k_sum_d(const double* __restrict__ d, uint64_t n, double* __restrict__ out) {
  double s = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) s += d[i];
  s = wave_sum_d(s);
  if ((threadIdx.x & (kWave - 1)) == 0 && s != 0.0) unsafeAtomicAdd(out, s);
}

// Order-preserving u64 key of a double (no NaNs).
__device__ inline uint64_t dkey(double x) {
  const uint64_t u = (uint64_t)__double_as_longlong(x);
  return (u >> 63) ? ~u : (u | (1ull << 63));
}

__global__ __launch_bounds__(kBlock) void k_min_key(const double* __restrict__ d, uint64_t n, unsigned long long* __restrict__ out) {
  uint64_t m = ~0ull;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) m = min(m, dkey(d[i]));
  for (int o = kWave / 2; o > 0; o >>= 1) m = min(m, (uint64_t)__shfl_xor((unsigned long long)m, o));
  if ((threadIdx.x & (kWave - 1)) == 0) atomicMin(out, (unsigned long long)m);
}

// Among vertices whose key equals *kmin: the largest degree, then the smallest id.
__global__ __launch_bounds__(kBlock) void k_pick(const double* __restrict__ d, uint64_t n, const uint64_t* __restrict__ rp,
                                                 const unsigned long long* __restrict__ kmin, unsigned long long* __restrict__ out) {
  uint64_t best = 0;
  const uint64_t target = *kmin;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
    if (dkey(d[i]) == target) best = max(best, ((rp[i + 1] - rp[i]) << 32) | (0xFFFFFFFFull - i));
  for (int o = kWave / 2; o > 0; o >>= 1) best = max(best, (uint64_t)__shfl_xor((unsigned long long)best, o));
  if ((threadIdx.x & (kWave - 1)) == 0 && best) atomicMax(out, (unsigned long long)best);
}

// Member subsets: mask[v] = 1 for members (v < V guaranteed by the caller's
// id check), assignments of non-members set to -1 (they join no cluster),
// a double array set to x outside the members, and the members' entries
// gathered in member order.
__global__ __launch_bounds__(kBlock) void k_km_mark(const uint32_t* __restrict__ members, uint64_t m, uint64_t V, uint8_t* __restrict__ mask,
                                                    unsigned long long* __restrict__ bad) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < m; i += (uint64_t)gridDim.x * kBlock) {
    const uint64_t v = members[i];
    if (v < V)
      mask[v] = 1;
    else
      atomicAdd(bad, 1ull);
  }
}

__global__ __launch_bounds__(kBlock) void k_km_mask_i(int32_t* __restrict__ a, const uint8_t* __restrict__ mask, uint64_t V) {
  for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < V; v += (uint64_t)gridDim.x * kBlock)
    if (!mask[v]) a[v] = -1;
}

__global__ __launch_bounds__(kBlock) void k_km_mask_d(double* __restrict__ d, const uint8_t* __restrict__ mask, uint64_t V, double x) {
  for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < V; v += (uint64_t)gridDim.x * kBlock)
    if (!mask[v]) d[v] = x;
}

__global__ __launch_bounds__(kBlock) void k_km_gather(const int32_t* __restrict__ a, const uint32_t* __restrict__ members, uint64_t m,
                                                      int32_t* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < m; i += (uint64_t)gridDim.x * kBlock) out[i] = a[members[i]];
}

__global__ void k_zero_u8(uint8_t* __restrict__ d, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) d[i] = 0;
}

__global__ void k_set_u64(unsigned long long* d, unsigned long long a, unsigned long long b) {
  d[0] = a;
  d[1] = b;
}

uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int check(vsgm_ctx* ctx, const char* what) {
  if (hipGetLastError() != hipSuccess) return vsgm_internal::fail_msg(ctx, VSGM_ERR_HIP, what);
  return VSGM_OK;
}

#define VSGM_KM_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)

int assign_pass(vsgm_ctx* ctx, const uint64_t* rp, const uint32_t* cols, uint64_t V, int K, const double* cent, const double* norm,
                int32_t* assign, double* sim) {
  hipStream_t s = vsgm_internal::stream(ctx);
  switch (K) {
#define X(N) \
  case N:    \
    hipLaunchKernelGGL(k_km_assign<N>, dim3(wave_grid(V)), dim3(kBlock), 0, s, rp, cols, V, cent, norm, assign, sim); \
    break;
    VSGM_KM_CASES(X)
#undef X
    default:
      return vsgm_internal::fail_msg(ctx, VSGM_ERR_ARG, "k out of range");
  }
  return check(ctx, "k_km_assign launch failed");
}

// Waves of the norm kernels' grid: the rows of their partials.
uint64_t fold_waves(uint64_t V) { return (uint64_t)grid_for((V + kBlock - 1) / kBlock) * kWavesPerBlock; }

// *symmetric = every list entry (v, w) has its reverse (w, v).
int symmetric_graph(vsgm_ctx* ctx, const uint64_t* rp, const uint32_t* cols, uint64_t V, bool* symmetric) {
  hipStream_t s = vsgm_internal::stream(ctx);
  unsigned long long* missing = nullptr;
  int st = vsgm_alloc(ctx, 2 * sizeof(unsigned long long), (void**)&missing);  // k_set_u64 writes two words
  if (st != VSGM_OK) return st;
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s, missing, 0ull, 0ull);
  hipLaunchKernelGGL(k_km_asym, dim3(wave_grid(V)), dim3(kBlock), 0, s, rp, cols, V, missing);
  unsigned long long h = 1;
  if ((st = check(ctx, "symmetry check launch failed")) == VSGM_OK) st = vsgm_d2h(ctx, &h, missing, sizeof(h));
  vsgm_free(ctx, missing);
  *symmetric = st == VSGM_OK && h == 0;
  return st;
}

// Centroids and norms from an assignment. Symmetric graphs use the gather
// form (no atomics: the same assignment gives the same centroids bit for
// bit); others scatter with float64 atomics. partials holds fold_waves(V)
// x K doubles.
int update_pass(vsgm_ctx* ctx, const uint64_t* rp, const uint32_t* cols, uint64_t V, int K, const int32_t* assign, double* cent, double* norm,
                uint64_t* sizes, bool symmetric, double* partials) {
  hipStream_t s = vsgm_internal::stream(ctx);
  if (!symmetric) hipLaunchKernelGGL(k_zero_d, dim3(grid_for((V * K + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, cent, V * K);
  hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(kBlock), 0, s, (double*)sizes, (uint64_t)K);
  hipLaunchKernelGGL(k_km_sizes, dim3(grid_for((V + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, assign, V, K, (unsigned long long*)sizes);
  switch (K) {
#define X(N)                                                                                                                       \
  case N:                                                                                                                          \
    if (symmetric)                                                                                                                 \
      hipLaunchKernelGGL(k_km_gather_cent<N>, dim3(wave_grid(V)), dim3(kBlock), 0, s, rp, cols, V, assign, cent);                  \
    else                                                                                                                           \
      hipLaunchKernelGGL(k_km_scatter<N>, dim3(wave_grid(V)), dim3(kBlock), 0, s, rp, cols, V, assign, cent);                      \
    hipLaunchKernelGGL(k_km_finish<N>, dim3(grid_for((V + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, cent, V,                     \
                       (const unsigned long long*)sizes, partials);                                                                \
    break;
    VSGM_KM_CASES(X)
#undef X
    default:
      return vsgm_internal::fail_msg(ctx, VSGM_ERR_ARG, "k out of range");
  }
  hipLaunchKernelGGL(k_km_fold, dim3(1), dim3(64), 0, s, partials, fold_waves(V), K, norm);
  return check(ctx, "k-means update launch failed");
}

}  // namespace

extern "C" {

int vsgm_kmeans_assign(vsgm_ctx* ctx, const uint64_t* d_row_ptrs, const uint32_t* d_cols, uint64_t vertex_count, int k,
                       const double* d_cent, const double* d_norm, int32_t* d_assign, double* d_sim) {
  if (!ctx) return VSGM_ERR_ARG;
  if (k < 1 || k > VSGM_KMEANS_MAX_K) return vsgm_internal::fail_msg(ctx, VSGM_ERR_ARG, "k out of range");
  if (vertex_count && (!d_row_ptrs || !d_cols || !d_cent || !d_norm || !d_assign || !d_sim))
    return vsgm_internal::fail_msg(ctx, VSGM_ERR_ARG, "null k-means argument");
  if (vertex_count == 0) return VSGM_OK;
  return assign_pass(ctx, d_row_ptrs, d_cols, vertex_count, k, d_cent, d_norm, d_assign, d_sim);
}

int vsgm_kmeans_update(vsgm_ctx* ctx, const uint64_t* d_row_ptrs, const uint32_t* d_cols, uint64_t vertex_count, int k,
                       const int32_t* d_assign, double* d_cent, double* d_norm, uint64_t* d_sizes) {
  if (!ctx) return VSGM_ERR_ARG;
  if (k < 1 || k > VSGM_KMEANS_MAX_K) return vsgm_internal::fail_msg(ctx, VSGM_ERR_ARG, "k out of range");
  if (vertex_count && (!d_row_ptrs || !d_cols || !d_cent || !d_norm || !d_assign || !d_sizes))
    return vsgm_internal::fail_msg(ctx, VSGM_ERR_ARG, "null k-means argument");
  if (vertex_count == 0) return VSGM_OK;
  bool symmetric = false;
  double* partials = nullptr;
  int st = symmetric_graph(ctx, d_row_ptrs, d_cols, vertex_count, &symmetric);
  if (st == VSGM_OK) st = vsgm_alloc(ctx, fold_waves(vertex_count) * k * sizeof(double), (void**)&partials);
  if (st == VSGM_OK) st = update_pass(ctx, d_row_ptrs, d_cols, vertex_count, k, d_assign, d_cent, d_norm, d_sizes, symmetric, partials);
  if (partials) vsgm_free(ctx, partials);
  return st;
}

}  // extern "C"

namespace {

// The whole run over every vertex (mask == nullptr) or over the vertices
// with mask[v] != 0 only: non-members are never chosen as initial centroids
// (their maxsim is pinned at the maximum), never join a cluster and add
// nothing to the objective. `first` is the seeded first centroid. The
// assignment stays on the device in d_assign.
int run_kmeans(vsgm_ctx* ctx, const uint64_t* d_row_ptrs, const uint32_t* d_cols, uint64_t V, int K, int iters, uint64_t first,
               const uint8_t* mask, int32_t* assign, uint64_t* h_sizes, double* h_objective) {
  hipStream_t s = vsgm_internal::stream(ctx);
  double *cent = nullptr, *norm = nullptr, *maxsim = nullptr, *sim = nullptr, *obj = nullptr;
  uint64_t *sizes = nullptr, *keys = nullptr;
  std::vector<void*> owned;
  auto get = [&](uint64_t bytes, void** p) {
    int st = vsgm_alloc(ctx, bytes, p);
    if (st == VSGM_OK) owned.push_back(*p);
    return st;
  };
  auto done = [&](int st) {
    for (void* p : owned) vsgm_free(ctx, p);
    return st;
  };
  int st = get(V * K * sizeof(double), (void**)&cent);
  if (st == VSGM_OK) st = get(K * sizeof(double), (void**)&norm);
  if (st == VSGM_OK) st = get(V * sizeof(double), (void**)&maxsim);
  if (st == VSGM_OK) st = get(V * sizeof(double), (void**)&sim);
  if (st == VSGM_OK) st = get(sizeof(double), (void**)&obj);
  if (st == VSGM_OK) st = get(K * sizeof(uint64_t), (void**)&sizes);
  if (st == VSGM_OK) st = get(2 * sizeof(uint64_t), (void**)&keys);
  double* partials = nullptr;
  if (st == VSGM_OK) st = get(fold_waves(V) * K * sizeof(double), (void**)&partials);
  if (st != VSGM_OK) return done(st);
  bool symmetric = false;
  if ((st = symmetric_graph(ctx, d_row_ptrs, d_cols, V, &symmetric)) != VSGM_OK) return done(st);

  // initialisation: farthest-first from the seeded vertex
  hipLaunchKernelGGL(k_zero_d, dim3(grid_for((V * K + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, cent, V * K);
  hipLaunchKernelGGL(k_fill_d, dim3(grid_for((V + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, maxsim, V, -DBL_MAX);
  if (mask) hipLaunchKernelGGL(k_km_mask_d, dim3(grid_for((V + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, maxsim, mask, V, DBL_MAX);
  uint64_t chosen = first;
  for (int c = 0; c < K; ++c) {
    switch (K) {
#define X(N)                                                                                                    \
  case N:                                                                                                       \
    hipLaunchKernelGGL(k_km_set<N>, dim3(1), dim3(kBlock), 0, s, d_row_ptrs, d_cols, chosen, c, cent);          \
    hipLaunchKernelGGL(k_km_colsq<N>, dim3(grid_for((V + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, cent, V, partials); \
    hipLaunchKernelGGL(k_km_fold, dim3(1), dim3(64), 0, s, partials, fold_waves(V), K, norm);                  \
    if (c + 1 < K)                                                                                              \
      hipLaunchKernelGGL(k_km_pp<N>, dim3(wave_grid(V)), dim3(kBlock), 0, s, d_row_ptrs, d_cols, V, cent, norm, c, chosen, maxsim); \
    break;
      VSGM_KM_CASES(X)
#undef X
    }
    if ((st = check(ctx, "k-means initialisation launch failed")) != VSGM_OK) return done(st);
    if (c + 1 == K) break;
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s, (unsigned long long*)keys, ~0ull, 0ull);
    hipLaunchKernelGGL(k_min_key, dim3(grid_for((V + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, maxsim, V, (unsigned long long*)keys);
    hipLaunchKernelGGL(k_pick, dim3(grid_for((V + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, maxsim, V, d_row_ptrs,
                       (const unsigned long long*)keys, (unsigned long long*)keys + 1);
    if ((st = check(ctx, "k-means initialisation launch failed")) != VSGM_OK) return done(st);
    uint64_t h[2];
    if ((st = vsgm_d2h(ctx, h, keys, sizeof(h))) != VSGM_OK) return done(st);
    chosen = 0xFFFFFFFFull - (h[1] & 0xFFFFFFFFull);
  }

  for (int it = 0; it <= iters; ++it) {
    if ((st = assign_pass(ctx, d_row_ptrs, d_cols, V, K, cent, norm, assign, sim)) != VSGM_OK) return done(st);
    if (mask) {
      hipLaunchKernelGGL(k_km_mask_i, dim3(grid_for((V + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, assign, mask, V);
      hipLaunchKernelGGL(k_km_mask_d, dim3(grid_for((V + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, sim, mask, V, 0.0);
    }
    if ((st = update_pass(ctx, d_row_ptrs, d_cols, V, K, assign, cent, norm, sizes, symmetric, partials)) != VSGM_OK) return done(st);
    if (h_objective) {
      hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(kBlock), 0, s, obj, (uint64_t)1);
      hipLaunchKernelGGL(k_sum_d, dim3(grid_for((V + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, sim, V, obj);
      if ((st = check(ctx, "k_sum_d launch failed")) != VSGM_OK) return done(st);
      if ((st = vsgm_d2h(ctx, &h_objective[it], obj, sizeof(double))) != VSGM_OK) return done(st);
    }
  }
  if ((st = vsgm_d2h(ctx, h_sizes, sizes, K * sizeof(uint64_t))) != VSGM_OK) return done(st);
  return done(VSGM_OK);
}

}  // namespace

extern "C" {

int vsgm_kmeans(vsgm_ctx* ctx, const uint64_t* d_row_ptrs, const uint32_t* d_cols, uint64_t vertex_count, int k, int iters,
                uint64_t seed, int32_t* h_assign, uint64_t* h_sizes, double* h_objective) {
  if (!ctx) return VSGM_ERR_ARG;
  if (k < 1 || k > VSGM_KMEANS_MAX_K || iters < 0) return vsgm_internal::fail_msg(ctx, VSGM_ERR_ARG, "k or iterations out of range");
  if (!d_row_ptrs || !d_cols || !h_assign || !h_sizes || vertex_count == 0 || vertex_count > 0xFFFFFFFFull)
    return vsgm_internal::fail_msg(ctx, VSGM_ERR_ARG, "bad k-means arguments");
  const uint64_t V = vertex_count;
  int32_t* assign = nullptr;
  int st = vsgm_alloc(ctx, V * sizeof(int32_t), (void**)&assign);
  if (st != VSGM_OK) return st;
  st = run_kmeans(ctx, d_row_ptrs, d_cols, V, k, iters, mix64(seed) % V, nullptr, assign, h_sizes, h_objective);
  if (st == VSGM_OK) st = vsgm_d2h(ctx, h_assign, assign, V * sizeof(int32_t));
  vsgm_free(ctx, assign);
  return st;
}

int vsgm_kmeans_members(vsgm_ctx* ctx, const uint64_t* d_row_ptrs, const uint32_t* d_cols, uint64_t vertex_count, const uint32_t* d_members,
                        uint64_t n_members, int k, int iters, uint64_t seed, int32_t* h_assign, uint64_t* h_sizes, double* h_objective) {
  if (!ctx) return VSGM_ERR_ARG;
  if (k < 1 || k > VSGM_KMEANS_MAX_K || iters < 0) return vsgm_internal::fail_msg(ctx, VSGM_ERR_ARG, "k or iterations out of range");
  if (!d_row_ptrs || !d_cols || !d_members || !h_assign || !h_sizes || vertex_count == 0 || vertex_count > 0xFFFFFFFFull ||
      n_members == 0 || n_members > vertex_count)
    return vsgm_internal::fail_msg(ctx, VSGM_ERR_ARG, "bad k-means arguments");
  if ((uint64_t)k > n_members) return vsgm_internal::fail_msg(ctx, VSGM_ERR_ARG, "more clusters than members");
  const uint64_t V = vertex_count, m = n_members;
  hipStream_t s = vsgm_internal::stream(ctx);
  uint8_t* mask = nullptr;
  int32_t *assign = nullptr, *out = nullptr;
  unsigned long long* bad = nullptr;
  auto done = [&](int st) {
    for (void* p : {(void*)mask, (void*)assign, (void*)out, (void*)bad})
      if (p) vsgm_free(ctx, p);
    return st;
  };
  int st = vsgm_alloc(ctx, V, (void**)&mask);
  if (st == VSGM_OK) st = vsgm_alloc(ctx, V * sizeof(int32_t), (void**)&assign);
  if (st == VSGM_OK) st = vsgm_alloc(ctx, m * sizeof(int32_t), (void**)&out);
  if (st == VSGM_OK) st = vsgm_alloc(ctx, 2 * sizeof(uint64_t), (void**)&bad);  // k_set_u64 writes two words
  if (st != VSGM_OK) return done(st);
  hipLaunchKernelGGL(k_zero_u8, dim3(grid_for((V + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, mask, V);
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s, bad, 0ull, 0ull);
  hipLaunchKernelGGL(k_km_mark, dim3(grid_for((m + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, d_members, m, V, mask, bad);
  if ((st = check(ctx, "member mask launch failed")) != VSGM_OK) return done(st);
  // the seeded first centroid is a member; ids are checked before any use
  uint64_t h[2] = {0, 0};
  uint32_t first = 0;
  if ((st = vsgm_d2h(ctx, h, bad, sizeof(h))) != VSGM_OK) return done(st);
  if (h[0]) return done(vsgm_internal::fail_msg(ctx, VSGM_ERR_ARG, "member id out of range"));
  if ((st = vsgm_d2h(ctx, &first, d_members + mix64(seed) % m, sizeof(first))) != VSGM_OK) return done(st);
  st = run_kmeans(ctx, d_row_ptrs, d_cols, V, k, iters, first, mask, assign, h_sizes, h_objective);
  if (st != VSGM_OK) return done(st);
  hipLaunchKernelGGL(k_km_gather, dim3(grid_for((m + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, assign, d_members, m, out);
  if ((st = check(ctx, "member gather launch failed")) != VSGM_OK) return done(st);
  if ((st = vsgm_d2h(ctx, h_assign, out, m * sizeof(int32_t))) != VSGM_OK) return done(st);
  return done(VSGM_OK);
}
}  // extern "C"
