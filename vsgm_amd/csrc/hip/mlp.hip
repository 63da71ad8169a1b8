// The learned planner's cost model (include/vsgm_mlp.h): a 3-layer MLP of
// hidden width 64 trained and evaluated with hand-written gfx950 kernels.
//
// Mapping: one wavefront per sample and one lane per hidden unit (the
// hidden width equals the wave width), four samples in flight per
// 256-thread workgroup. The BF16 weights are staged in LDS once per
// workgroup in the two layouts the passes read without bank conflicts
// (W1 and W2 transposed for the forward pass, W2 as stored for the
// backward pass); inputs and the previous layer's activations are
// broadcast from per-wave LDS rows; every sum is FP32. The backward pass
// accumulates the workgroup's W1 / W2 gradients in LDS with LDS float
// atomics (lane-rotated so a wave's 64 updates hit distinct banks) and adds
// them to the global gradient once per workgroup.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "vsgm_hip.h"
#include "vsgm_mlp.h"

namespace vsgm_internal {
hipStream_t stream(vsgm_ctx* ctx);
int fail_msg(vsgm_ctx* ctx, int code, const char* msg);
}  // namespace vsgm_internal

namespace {

constexpr int H = VSGM_MLP_HIDDEN;
constexpr int kMaxIn = VSGM_MLP_MAX_IN;
constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kWaves = kBlock / kWave;
static_assert(H == kWave, "one lane per hidden unit");

// Parameter offsets in the FP32 vector.
struct Layout {
  uint64_t w1, b1, w2, b2, w3, b3, n;
  __host__ __device__ explicit Layout(int d) {
    w1 = 0;
    b1 = (uint64_t)H * d;
    w2 = b1 + H;
    b2 = w2 + (uint64_t)H * H;
    w3 = b2 + H;
    b3 = w3 + H;
    n = b3 + 1;
  }
};

// FP32 -> BF16, round to nearest even (inputs are finite).
__host__ __device__ inline uint16_t to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ inline float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline float wave_sum(float v) {
  for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

// Forward (y != nullptr) and/or backward (grad != nullptr) over a batch.
// wq: BF16 copies W1 [H][D] | W2 [H][H] | w3 [H]; params: FP32 vector (for
// the biases). scale = 1 / batch (mean loss).
__global__ __launch_bounds__(kBlock) void k_mlp(const float* __restrict__ x, const float* __restrict__ t, uint64_t batch, int D,
                                                const uint16_t* __restrict__ wq, const float* __restrict__ params, float* __restrict__ y,
                                                float* __restrict__ grad, float* __restrict__ loss, float scale) {
  __shared__ uint16_t s_w1t[kMaxIn * H];  // [d][j] = W1[j][d]
  __shared__ uint16_t s_w2t[H * H];       // [k][j] = W2[j][k]
  __shared__ uint16_t s_w2[H * H];        // [j][k]
  __shared__ float s_x[kWaves][kMaxIn];
  __shared__ float s_h1[kWaves][H];
  __shared__ float s_dh2[kWaves][H];
  __shared__ float s_g1[H * kMaxIn];      // [k][d]
  __shared__ float s_g2[H * H];           // [j][k]
  const Layout L(D);
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const bool backward = grad != nullptr;
  for (int i = threadIdx.x; i < H * D; i += kBlock) s_w1t[(i % D) * H + i / D] = wq[i];
  for (int i = threadIdx.x; i < H * H; i += kBlock) {
    const uint16_t w = wq[H * D + i];
    s_w2[i] = w;
    s_w2t[(i % H) * H + i / H] = w;
  }
  if (backward) {
    for (int i = threadIdx.x; i < H * kMaxIn; i += kBlock) s_g1[i] = 0.f;
    for (int i = threadIdx.x; i < H * H; i += kBlock) s_g2[i] = 0.f;
  }
  __syncthreads();
  const float b1 = params[L.b1 + lane], b2 = params[L.b2 + lane], b3 = params[L.b3];
  const float w3 = from_bf16(wq[H * D + H * H + lane]);
  float g3 = 0.f, gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, lsum = 0.f;
  for (uint64_t s = (uint64_t)blockIdx.x * kWaves + wave; s < batch; s += (uint64_t)gridDim.x * kWaves) {
    for (int d = lane; d < D; d += kWave) s_x[wave][d] = x[s * D + d];
    wave_sync();
    float z1 = b1;
    for (int d = 0; d < D; ++d) z1 += from_bf16(s_w1t[d * H + lane]) * s_x[wave][d];
    const float h1 = z1 > 0.f ? z1 : 0.f;
    s_h1[wave][lane] = h1;
    wave_sync();
    float z2 = b2;
    for (int k = 0; k < H; ++k) z2 += from_bf16(s_w2t[k * H + lane]) * s_h1[wave][k];
    const float h2 = z2 > 0.f ? z2 : 0.f;
    const float yv = wave_sum(w3 * h2) + b3;
    if (y && lane == 0) y[s] = yv;
    if (backward) {
      const float err = yv - t[s];
      const float dy = err * scale;
      if (lane == 0) {
        lsum += 0.5f * err * err * scale;
        gb3 += dy;
      }
      g3 += dy * h2;
      const float dh2 = z2 > 0.f ? dy * w3 : 0.f;
      gb2 += dh2;
      s_dh2[wave][lane] = dh2;
      wave_sync();
      // dW2[j][k] += dh2[j] * h1[k]; lane j, k rotated by j
      for (int i = 0; i < H; ++i) {
        const int k = (i + lane) & (H - 1);
        atomicAdd(&s_g2[lane * H + k], dh2 * s_h1[wave][k]);
      }
      // dh1[k] = sum_j W2[j][k] dh2[j]; lane k
      float a = 0.f;
      for (int j = 0; j < H; ++j) a += from_bf16(s_w2[j * H + lane]) * s_dh2[wave][j];
      const float dh1 = z1 > 0.f ? a : 0.f;
      gb1 += dh1;
      // dW1[k][d] += dh1[k] * x[d]; lane k, d rotated by k
      for (int i = 0; i < D; ++i) {
        int d = i + lane;
        while (d >= D) d -= D;
        atomicAdd(&s_g1[lane * kMaxIn + d], dh1 * s_x[wave][d]);
      }
    }
    wave_sync();
  }
  if (!backward) return;
  __syncthreads();
  for (int i = threadIdx.x; i < H * D; i += kBlock) {
    const float v = s_g1[(i / D) * kMaxIn + i % D];
    if (v != 0.f) atomicAdd(&grad[L.w1 + i], v);
  }
  for (int i = threadIdx.x; i < H * H; i += kBlock) {
    const float v = s_g2[i];
    if (v != 0.f) atomicAdd(&grad[L.w2 + i], v);
  }
  atomicAdd(&grad[L.b1 + lane], gb1);
  atomicAdd(&grad[L.b2 + lane], gb2);
  atomicAdd(&grad[L.w3 + lane], g3);
  if (lane == 0) {
    atomicAdd(&grad[L.b3], gb3);
    atomicAdd(loss, lsum);
  }
}

// ---------------------------------------------------------------------------
// Matrix-core path: every product of the forward and backward passes as
// bf16 MFMA tiles with FP32 accumulation (v_mfma_f32_32x32x8_bf16). A
// workgroup takes tiles of 64 samples; each of its 4 wavefronts computes one
// 32 x 32 block of every 64 x 64 product (and two blocks of the 64 x D
// weight gradient), keeping the weight-gradient blocks in accumulator
// registers across tiles. Activations and backward deltas enter the
// products rounded to BF16 (the sums and the bias / w3 terms stay FP32):
// x, h1 = bf16(relu(z1)), dz2 = bf16(dy w3 [z2 > 0]), dz1 = bf16((dz2 W2)
// [h1 > 0]).
//
// Operand layout of v_mfma_f32_32x32x8_bf16 (A 32 x 8, B 8 x 32): lane l
// supplies A[l % 32][4 (l / 32) + i] and B[4 (l / 32) + i][l % 32], i < 4,
// and holds C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32] in accumulator r < 16.
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kT = 64;               // samples per tile
constexpr int kXS = kMaxIn + 8;      // LDS row stride (elements) of x / W1 tiles
constexpr int kHS = H + 8;           // LDS row stride of 64-wide BF16 tiles
constexpr int kH2S = H + 1;          // LDS row stride of the FP32 tile

__device__ inline bf16x4 row4(const uint16_t* p) {  // p[0..3], 8-byte aligned
  const uint2 v = *reinterpret_cast<const uint2*>(p);
  bf16x4 r;
  r[0] = (short)(v.x & 0xFFFFu);
  r[1] = (short)(v.x >> 16);
  r[2] = (short)(v.y & 0xFFFFu);
  r[3] = (short)(v.y >> 16);
  return r;
}

__device__ inline bf16x4 col4(const uint16_t* p, int stride) {  // p[0], p[stride], p[2 stride], p[3 stride]
  bf16x4 r;
#pragma unroll
  for (int i = 0; i < 4; ++i) r[i] = (short)p[i * stride];
  return r;
}

// acc (32 x 32) += A (32 x K) B (K x 32). Operand access: A by rows (a[i][k]
// at a + i * lda + k) or by columns (a[i][k] at a + k * lda + i); likewise B
// by rows of its transpose (b[k][j] at b + j * ldb + k) or directly
// (b[k][j] at b + k * ldb + j).
template <bool A_ROWS, bool B_T>
__device__ inline void mma(f32x16& acc, const uint16_t* a, int lda, const uint16_t* b, int ldb, int K, int lane) {
  const int i = lane & 31, kq = (lane >> 5) * 4;
  for (int k = 0; k < K; k += 8) {
    const bf16x4 av = A_ROWS ? row4(a + i * lda + k + kq) : col4(a + (k + kq) * lda + i, lda);
    const bf16x4 bv = B_T ? row4(b + i * ldb + k + kq) : col4(b + (k + kq) * ldb + i, ldb);
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
  }
}

__device__ inline int acc_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

__global__ __launch_bounds__(kBlock) void k_mlp_mfma(const float* __restrict__ x, const float* __restrict__ t, uint64_t batch, int D,
                                                     const uint16_t* __restrict__ wq, const float* __restrict__ params,
                                                     float* __restrict__ y, float* __restrict__ grad, float* __restrict__ loss,
                                                     float scale) {
  __shared__ __align__(16) uint16_t s_w1[H * kXS];   // W1 [j][d], zero beyond D
  __shared__ __align__(16) uint16_t s_w2[H * kHS];   // W2 [j][k]
  __shared__ __align__(16) uint16_t s_x[kT * kXS];   // x tile [s][d]
  __shared__ __align__(16) uint16_t s_h1[kT * kHS];  // h1 tile [s][k]
  __shared__ __align__(16) uint16_t s_dz2[kT * kHS];
  __shared__ __align__(16) uint16_t s_dz1[kT * kHS];
  __shared__ float s_f[kT * kH2S];                   // h2 tile, then dz1 (FP32)
  __shared__ float s_dy[kT];
  __shared__ float s_b1[H], s_b2[H], s_w3[H];
  const Layout L(D);
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const bool backward = grad != nullptr;
  const int Dp = (D + 7) & ~7;
  for (int i = tid; i < H * kXS; i += kBlock) {
    const int j = i / kXS, d = i % kXS;
    s_w1[i] = d < D ? wq[j * D + d] : (uint16_t)0;
  }
  for (int i = tid; i < H * kHS; i += kBlock) {
    const int j = i / kHS, k = i % kHS;
    s_w2[i] = k < H ? wq[H * D + j * H + k] : (uint16_t)0;
  }
  if (tid < H) {
    s_b1[tid] = params[L.b1 + tid];
    s_b2[tid] = params[L.b2 + tid];
    s_w3[tid] = from_bf16(wq[H * D + H * H + tid]);
  }
  const float b3 = params[L.b3];
  const int rb = (wave >> 1) * 32, cb = (wave & 1) * 32;  // this wave's block of a 64 x 64 product
  f32x16 g2 = {0}, g1a = {0}, g1b = {0};                   // dW2 block (rb, cb); dW1 blocks (rb, cb) and (rb, cb + 64)
  float gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f, lsum = 0.f;
  __syncthreads();
  for (uint64_t base = (uint64_t)blockIdx.x * kT; base < batch; base += (uint64_t)gridDim.x * kT) {
    for (int i = tid; i < kT * kXS; i += kBlock) {
      const int s = i / kXS, d = i % kXS;
      const uint64_t g = base + s;
      s_x[i] = (g < batch && d < D) ? to_bf16(x[g * D + d]) : (uint16_t)0;
    }
    __syncthreads();
    f32x16 acc = {0};
    mma<true, true>(acc, s_x + rb * kXS, kXS, s_w1 + cb * kXS, kXS, Dp, lane);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int s = rb + acc_row(r, lane), k = cb + (lane & 31);
      s_h1[s * kHS + k] = to_bf16(fmaxf(acc[r] + s_b1[k], 0.f));
    }
    __syncthreads();
    acc = f32x16{0};
    mma<true, true>(acc, s_h1 + rb * kHS, kHS, s_w2 + cb * kHS, kHS, H, lane);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int s = rb + acc_row(r, lane), j = cb + (lane & 31);
      s_f[s * kH2S + j] = fmaxf(acc[r] + s_b2[j], 0.f);
    }
    __syncthreads();
    if (tid < kT) {
      float yv = b3;
      for (int j = 0; j < H; ++j) yv += s_f[tid * kH2S + j] * s_w3[j];
      const uint64_t g = base + tid;
      float dy = 0.f;
      if (g < batch) {
        if (y) y[g] = yv;
        if (backward) {
          const float err = yv - t[g];
          dy = err * scale;
          lsum += 0.5f * err * err * scale;
        }
      }
      s_dy[tid] = dy;
    }
    if (!backward) {
      __syncthreads();
      continue;
    }
    __syncthreads();
    for (int i = tid; i < kT * H; i += kBlock) {
      const int s = i / H, j = i % H;
      s_dz2[s * kHS + j] = to_bf16(s_f[s * kH2S + j] > 0.f ? s_dy[s] * s_w3[j] : 0.f);
    }
    if (tid < H) {
      float a = 0.f, b = 0.f;
      for (int s = 0; s < kT; ++s) {
        const float h2 = s_f[s * kH2S + tid];
        a += s_dy[s] * h2;
        b += h2 > 0.f ? s_dy[s] * s_w3[tid] : 0.f;
      }
      gw3 += a;
      gb2 += b;
    }
    if (tid == 0)
      for (int s = 0; s < kT; ++s) gb3 += s_dy[s];
    __syncthreads();
    // dW2 += dz2^T h1 ; dh1 = dz2 W2
    mma<false, false>(g2, s_dz2 + rb, kHS, s_h1 + cb, kHS, kT, lane);
    acc = f32x16{0};
    mma<true, false>(acc, s_dz2 + rb * kHS, kHS, s_w2 + cb, kHS, H, lane);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int s = rb + acc_row(r, lane), k = cb + (lane & 31);
      const float v = from_bf16(s_h1[s * kHS + k]) > 0.f ? acc[r] : 0.f;
      s_f[s * kH2S + k] = v;
      s_dz1[s * kHS + k] = to_bf16(v);
    }
    __syncthreads();
    if (tid < H)
      for (int s = 0; s < kT; ++s) gb1 += s_f[s * kH2S + tid];
    // dW1 += dz1^T x
    mma<false, false>(g1a, s_dz1 + rb, kHS, s_x + cb, kXS, kT, lane);
    if (Dp > 64) mma<false, false>(g1b, s_dz1 + rb, kHS, s_x + cb + 64, kXS, kT, lane);
    __syncthreads();  // the next tile overwrites every tile buffer
  }
  if (!backward) return;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int j = rb + acc_row(r, lane), k = cb + (lane & 31);
    if (g2[r] != 0.f) atomicAdd(&grad[L.w2 + j * H + k], g2[r]);
    const int d0 = cb + (lane & 31), d1 = d0 + 64;
    if (d0 < D && g1a[r] != 0.f) atomicAdd(&grad[L.w1 + j * D + d0], g1a[r]);
    if (d1 < D && g1b[r] != 0.f) atomicAdd(&grad[L.w1 + j * D + d1], g1b[r]);
  }
  if (tid < H) {
    atomicAdd(&grad[L.b1 + tid], gb1);
    atomicAdd(&grad[L.b2 + tid], gb2);
    atomicAdd(&grad[L.w3 + tid], gw3);
  }
  if (tid == 0) {
    atomicAdd(&grad[L.b3], gb3);
  }
  if (tid < kT) atomicAdd(loss, lsum);
}

__global__ __launch_bounds__(kBlock) void k_zero_f32(float* d, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) d[i] = 0.f;
}

// BF16 copies of the weight blocks of the FP32 vector.
__device__ inline void refresh_bf16(const Layout& L, int D, uint64_t i, float p, uint16_t* wq) {
  if (i < L.b1) wq[i] = to_bf16(p);
  else if (i >= L.w2 && i < L.b2) wq[(uint64_t)H * D + (i - L.w2)] = to_bf16(p);
  else if (i >= L.w3 && i < L.b3) wq[(uint64_t)H * D + H * H + (i - L.w3)] = to_bf16(p);
}

__global__ __launch_bounds__(kBlock) void k_refresh(const float* __restrict__ params, int D, uint16_t* __restrict__ wq) {
  const Layout L(D);
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < L.n; i += (uint64_t)gridDim.x * kBlock) refresh_bf16(L, D, i, params[i], wq);
}

// Adam with bias corrections c1 = 1 / (1 - beta1^t), c2 = 1 / (1 - beta2^t).
__global__ __launch_bounds__(kBlock) void k_adam(float* __restrict__ params, const float* __restrict__ grad, float* __restrict__ m,
                                                 float* __restrict__ v, int D, float lr, float beta1, float beta2, float eps, float c1,
                                                 float c2, uint16_t* __restrict__ wq) {
  const Layout L(D);
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < L.n; i += (uint64_t)gridDim.x * kBlock) {
    const float g = grad[i];
    const float mi = beta1 * m[i] + (1.f - beta1) * g;
    const float vi = beta2 * v[i] + (1.f - beta2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float p = params[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    params[i] = p;
    refresh_bf16(L, D, i, p, wq);
  }
}

uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

unsigned grid_for(uint64_t units, unsigned cap = 1024) { return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(units, cap)); }

}  // namespace

struct vsgm_mlp {
  vsgm_ctx* ctx = nullptr;
  bool mfma = true;  // matrix-core kernels (else one wavefront per sample, FP32 activations)
  int D = 0;
  uint64_t n = 0;
  float* params = nullptr;
  float* grad = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  float* loss = nullptr;
  uint16_t* wq = nullptr;
  int steps = 0;
};

namespace {

int launch_ok(vsgm_ctx* ctx, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return vsgm_internal::fail_msg(ctx, VSGM_ERR_HIP, what);
  return VSGM_OK;
}

int zero(vsgm_mlp* m, float* d, uint64_t n) {
  hipLaunchKernelGGL(k_zero_f32, dim3(grid_for((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, vsgm_internal::stream(m->ctx), d, n);
  return launch_ok(m->ctx, "k_zero_f32 launch failed");
}

}  // namespace

extern "C" {

int vsgm_mlp_create(vsgm_ctx* ctx, int in_dim, uint64_t seed, vsgm_mlp** out) {
  if (!ctx || !out) return VSGM_ERR_ARG;
  *out = nullptr;
  if (in_dim < 1 || in_dim > kMaxIn) return vsgm_internal::fail_msg(ctx, VSGM_ERR_ARG, "mlp input width out of range");
  vsgm_mlp* m = new vsgm_mlp();
  m->ctx = ctx;
  m->D = in_dim;
  m->n = Layout(in_dim).n;
  const uint64_t nq = (uint64_t)H * in_dim + H * H + H;
  int s = vsgm_alloc(ctx, m->n * 4, (void**)&m->params);
  if (s == VSGM_OK) s = vsgm_alloc(ctx, m->n * 4, (void**)&m->grad);
  if (s == VSGM_OK) s = vsgm_alloc(ctx, m->n * 4, (void**)&m->m);
  if (s == VSGM_OK) s = vsgm_alloc(ctx, m->n * 4, (void**)&m->v);
  if (s == VSGM_OK) s = vsgm_alloc(ctx, 4, (void**)&m->loss);
  if (s == VSGM_OK) s = vsgm_alloc(ctx, nq * 2, (void**)&m->wq);
  if (s != VSGM_OK) {
    vsgm_mlp_free(m);
    return s;
  }
  // He-uniform weights from SplitMix64 keyed by (seed, index); zero biases
  const Layout L(in_dim);
  std::vector<float> p(m->n, 0.f);
  auto fill = [&](uint64_t off, uint64_t count, int fan_in) {
    const double a = std::sqrt(6.0 / fan_in);
    for (uint64_t i = 0; i < count; ++i) {
      const double u = (double)(mix64(seed * 0x100000001B3ull + off + i) >> 11) * (1.0 / 9007199254740992.0);
      p[off + i] = (float)((2.0 * u - 1.0) * a);
    }
  };
  fill(L.w1, (uint64_t)H * in_dim, in_dim);
  fill(L.w2, (uint64_t)H * H, H);
  fill(L.w3, H, H);
  s = vsgm_mlp_set_params(m, p.data(), m->n);
  if (s != VSGM_OK) {
    vsgm_mlp_free(m);
    return s;
  }
  *out = m;
  return VSGM_OK;
}

int vsgm_mlp_free(vsgm_mlp* m) {
  if (!m) return VSGM_OK;
  for (void* p : {(void*)m->params, (void*)m->grad, (void*)m->m, (void*)m->v, (void*)m->loss, (void*)m->wq})
    if (p) vsgm_free(m->ctx, p);
  delete m;
  return VSGM_OK;
}

int vsgm_mlp_param_count(const vsgm_mlp* m, uint64_t* n) {
  if (!m || !n) return VSGM_ERR_ARG;
  *n = m->n;
  return VSGM_OK;
}

int vsgm_mlp_get_params(vsgm_mlp* m, float* h_params, uint64_t n) {
  if (!m || !h_params) return VSGM_ERR_ARG;
  if (n != m->n) return vsgm_internal::fail_msg(m->ctx, VSGM_ERR_ARG, "parameter count mismatch");
  return vsgm_d2h(m->ctx, h_params, m->params, n * 4);
}

int vsgm_mlp_set_params(vsgm_mlp* m, const float* h_params, uint64_t n) {
  if (!m || !h_params) return VSGM_ERR_ARG;
  if (n != m->n) return vsgm_internal::fail_msg(m->ctx, VSGM_ERR_ARG, "parameter count mismatch");
  int s = vsgm_h2d(m->ctx, m->params, h_params, n * 4);
  if (s == VSGM_OK) s = zero(m, m->m, n);
  if (s == VSGM_OK) s = zero(m, m->v, n);
  if (s != VSGM_OK) return s;
  hipLaunchKernelGGL(k_refresh, dim3(grid_for((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, vsgm_internal::stream(m->ctx), m->params, m->D, m->wq);
  if ((s = launch_ok(m->ctx, "k_refresh launch failed")) != VSGM_OK) return s;
  m->steps = 0;
  return vsgm_ctx_sync(m->ctx);
}

int vsgm_mlp_forward(vsgm_mlp* m, const float* d_x, uint64_t batch, float* d_y) {
  if (!m || (batch && (!d_x || !d_y))) return VSGM_ERR_ARG;
  if (batch == 0) return VSGM_OK;
  if (m->mfma)
    hipLaunchKernelGGL(k_mlp_mfma, dim3(grid_for((batch + kT - 1) / kT)), dim3(kBlock), 0, vsgm_internal::stream(m->ctx), d_x,
                       (const float*)nullptr, batch, m->D, m->wq, m->params, d_y, (float*)nullptr, (float*)nullptr, 0.f);
  else
    hipLaunchKernelGGL(k_mlp, dim3(grid_for((batch + kWaves - 1) / kWaves)), dim3(kBlock), 0, vsgm_internal::stream(m->ctx), d_x,
                       (const float*)nullptr, batch, m->D, m->wq, m->params, d_y, (float*)nullptr, (float*)nullptr, 0.f);
  return launch_ok(m->ctx, "k_mlp launch failed");
}

int vsgm_mlp_backward(vsgm_mlp* m, const float* d_x, const float* d_t, uint64_t batch, float* h_loss, float* h_grad, uint64_t n) {
  if (!m || !d_x || !d_t || batch == 0) return VSGM_ERR_ARG;
  if (h_grad && n != m->n) return vsgm_internal::fail_msg(m->ctx, VSGM_ERR_ARG, "parameter count mismatch");
  int s = zero(m, m->grad, m->n);
  if (s == VSGM_OK) s = zero(m, m->loss, 1);
  if (s != VSGM_OK) return s;
  // at most 256 workgroups: each adds its gradient once
  if (m->mfma)
    hipLaunchKernelGGL(k_mlp_mfma, dim3(grid_for((batch + kT - 1) / kT, 256)), dim3(kBlock), 0, vsgm_internal::stream(m->ctx), d_x, d_t,
                       batch, m->D, m->wq, m->params, (float*)nullptr, m->grad, m->loss, (float)(1.0 / (double)batch));
  else
    hipLaunchKernelGGL(k_mlp, dim3(grid_for((batch + kWaves - 1) / kWaves, 256)), dim3(kBlock), 0, vsgm_internal::stream(m->ctx), d_x,
                       d_t, batch, m->D, m->wq, m->params, (float*)nullptr, m->grad, m->loss, (float)(1.0 / (double)batch));
  if ((s = launch_ok(m->ctx, "k_mlp launch failed")) != VSGM_OK) return s;
  if (h_loss && (s = vsgm_d2h(m->ctx, h_loss, m->loss, 4)) != VSGM_OK) return s;
  if (h_grad && (s = vsgm_d2h(m->ctx, h_grad, m->grad, n * 4)) != VSGM_OK) return s;
  return VSGM_OK;
}

int vsgm_mlp_set_mfma(vsgm_mlp* m, int on) {
  if (!m) return VSGM_ERR_ARG;
  m->mfma = on != 0;
  return VSGM_OK;
}

int vsgm_mlp_adam_step(vsgm_mlp* m, float lr, float beta1, float beta2, float eps) {
  if (!m) return VSGM_ERR_ARG;
  m->steps += 1;
  const float c1 = (float)(1.0 / (1.0 - std::pow((double)beta1, m->steps)));
  const float c2 = (float)(1.0 / (1.0 - std::pow((double)beta2, m->steps)));
  hipLaunchKernelGGL(k_adam, dim3(grid_for((m->n + kBlock - 1) / kBlock)), dim3(kBlock), 0, vsgm_internal::stream(m->ctx), m->params, m->grad,
                     m->m, m->v, m->D, lr, beta1, beta2, eps, c1, c2, m->wq);
  return launch_ok(m->ctx, "k_adam launch failed");
}

}  // extern "C"
