// Small MLP cost model: n_in -> 64 -> 64 -> n_out, ReLU hidden layers, linear outputs.
// Weights are used in BF16 (round-to-nearest-even from FP32 master copies), every product
// is accumulated in FP32. Training minimises the mean squared error with Adam on the
// device. Gradients are deterministic: each block writes the partial gradient of its
// sample range, and a second kernel sums the partials in block order (no float atomics).
//
// Parameter layout (FP32 master and BF16 copy): W1[64][n_in], b1[64], W2[64][64], b2[64],
// W3[n_out][64], b3[n_out].
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "mg_device.hpp"
#include "mg_internal.hpp"

namespace mgamd {
namespace {

constexpr int kH = 64;        // hidden width
constexpr int kMlpBlock = 64;  // one wave per block

// IEEE binary32 -> bfloat16, round to nearest even (NaN kept quiet).
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return static_cast<uint16_t>((u >> 16) | 0x40);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h) {
  const uint32_t u = static_cast<uint32_t>(h) << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

struct Dims {
  uint32_t n_in, n_out;
  uint32_t o_w1, o_b1, o_w2, o_b2, o_w3, o_b3, n_params;
};

Dims make_dims(uint32_t n_in, uint32_t n_out) {
  Dims d{};
  d.n_in = n_in;
  d.n_out = n_out;
  d.o_w1 = 0;
  d.o_b1 = d.o_w1 + kH * n_in;
  d.o_w2 = d.o_b1 + kH;
  d.o_b2 = d.o_w2 + kH * kH;
  d.o_w3 = d.o_b2 + kH;
  d.o_b3 = d.o_w3 + n_out * kH;
  d.n_params = d.o_b3 + n_out;
  return d;
}

// Lane j of a wave is hidden unit j of both hidden layers; the wave streams samples.
// Lane j keeps row j of W1 and W2, column j of W2 (for the backward pass) and column j of
// W3, all as FP32 values of the BF16 weights. Activations are broadcast with readlane;
// outputs are wave sums.
template <int NI, int NO>
struct LaneWeights {
  float w1[NI];
  float w2r[kH];  // W2[j][i], i = 0..63
  float w2c[kH];  // W2[i][j], i = 0..63
  float w3[NO];
  float b3[NO];
  float b1, b2;
};

__device__ inline float lane_f(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ inline float wave_sum_f(float v) {
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) v += __shfl_xor(v, m, kWave);
  return v;
}

template <int NI, int NO>
__device__ inline void load_lane_weights(const Dims& d, const uint16_t* __restrict__ w16, uint32_t j,
                                         LaneWeights<NI, NO>& L) {
#pragma unroll
  for (int i = 0; i < NI; ++i) L.w1[i] = bf16_to_f32(w16[d.o_w1 + j * NI + i]);
#pragma unroll
  for (int i = 0; i < kH; ++i) {
    L.w2r[i] = bf16_to_f32(w16[d.o_w2 + j * kH + i]);
    L.w2c[i] = bf16_to_f32(w16[d.o_w2 + i * kH + j]);
  }
#pragma unroll
  for (int o = 0; o < NO; ++o) {
    L.w3[o] = bf16_to_f32(w16[d.o_w3 + o * kH + j]);
    L.b3[o] = bf16_to_f32(w16[d.o_b3 + o]);
  }
  L.b1 = bf16_to_f32(w16[d.o_b1 + j]);
  L.b2 = bf16_to_f32(w16[d.o_b2 + j]);
}

// Forward of sample x (wave-uniform pointer): h1, h2 of this lane's unit, outputs y (uniform).
template <int NI, int NO>
__device__ inline void forward_wave(const LaneWeights<NI, NO>& L, const float* __restrict__ x, float& h1, float& h2,
                                    float* y) {
  float a = L.b1;
#pragma unroll
  for (int i = 0; i < NI; ++i) a += L.w1[i] * x[i];
  h1 = a > 0.f ? a : 0.f;
  a = L.b2;
#pragma unroll
  for (int i = 0; i < kH; ++i) a += L.w2r[i] * lane_f(h1, i);
  h2 = a > 0.f ? a : 0.f;
#pragma unroll
  for (int o = 0; o < NO; ++o) y[o] = L.b3[o] + wave_sum_f(L.w3[o] * h2);
}

template <int NI, int NO>
__global__ void __launch_bounds__(kMlpBlock) mlp_forward_kernel(Dims d, const uint16_t* __restrict__ w16,
                                                                const float* __restrict__ x, uint32_t n,
                                                                float* __restrict__ y) {
  const uint32_t j = lane_id();
  LaneWeights<NI, NO> L;
  load_lane_weights<NI, NO>(d, w16, j, L);
  const uint32_t per_wave = 64;
  const uint32_t s0 = blockIdx.x * per_wave;
  const uint32_t s1 = min(n, s0 + per_wave);
  for (uint32_t s = s0; s < s1; ++s) {
    float h1, h2, yo[NO];
    forward_wave<NI, NO>(L, x + static_cast<uint64_t>(s) * NI, h1, h2, yo);
#pragma unroll
    for (int o = 0; o < NO; ++o)
      if (j == static_cast<uint32_t>(o)) y[static_cast<uint64_t>(s) * NO + o] = yo[o];
  }
}

// Partial gradient of L = scale * sum over samples and outputs of (y - t)^2 for samples
// [block*per_block, ...), accumulated in sample order by one wave; written to
// partial[block][param].
template <int NI, int NO>
__global__ void __launch_bounds__(kMlpBlock) mlp_grad_kernel(Dims d, const uint16_t* __restrict__ w16,
                                                             const float* __restrict__ x, const float* __restrict__ t,
                                                             uint32_t n, uint32_t per_block, float scale,
                                                             float* __restrict__ partial, float* __restrict__ loss_part) {
  const uint32_t j = lane_id();
  LaneWeights<NI, NO> L;
  load_lane_weights<NI, NO>(d, w16, j, L);
  float g1[NI], g2[kH], g3[NO], gb3[NO];
#pragma unroll
  for (int i = 0; i < NI; ++i) g1[i] = 0.f;
#pragma unroll
  for (int i = 0; i < kH; ++i) g2[i] = 0.f;
#pragma unroll
  for (int o = 0; o < NO; ++o) { g3[o] = 0.f; gb3[o] = 0.f; }
  float gb1 = 0.f, gb2 = 0.f, loss = 0.f;
  const uint32_t s0 = blockIdx.x * per_block;
  const uint32_t s1 = min(n, s0 + per_block);
  for (uint32_t s = s0; s < s1; ++s) {
    const float* xs = x + static_cast<uint64_t>(s) * NI;
    float h1, h2, yo[NO], dy[NO];
    forward_wave<NI, NO>(L, xs, h1, h2, yo);
    float dh2 = 0.f;
#pragma unroll
    for (int o = 0; o < NO; ++o) {
      const float e = yo[o] - t[static_cast<uint64_t>(s) * NO + o];
      loss += e * e;
      dy[o] = 2.f * e * scale;
      dh2 += L.w3[o] * dy[o];
      g3[o] += dy[o] * h2;
      gb3[o] += dy[o];
    }
    dh2 = h2 > 0.f ? dh2 : 0.f;
    // dh1_j = relu'(h1_j) * sum_i W2[i][j] dh2_i
    float a = 0.f;
#pragma unroll
    for (int i = 0; i < kH; ++i) a += L.w2c[i] * lane_f(dh2, i);
    const float dh1 = h1 > 0.f ? a : 0.f;
#pragma unroll
    for (int i = 0; i < kH; ++i) g2[i] += dh2 * lane_f(h1, i);
    gb2 += dh2;
#pragma unroll
    for (int i = 0; i < NI; ++i) g1[i] += dh1 * xs[i];
    gb1 += dh1;
  }
  float* out = partial + static_cast<uint64_t>(blockIdx.x) * d.n_params;
#pragma unroll
  for (int i = 0; i < NI; ++i) out[d.o_w1 + j * NI + i] = g1[i];
  out[d.o_b1 + j] = gb1;
#pragma unroll
  for (int i = 0; i < kH; ++i) out[d.o_w2 + j * kH + i] = g2[i];
  out[d.o_b2 + j] = gb2;
#pragma unroll
  for (int o = 0; o < NO; ++o) {
    out[d.o_w3 + o * kH + j] = g3[o];
    if (j == static_cast<uint32_t>(o)) out[d.o_b3 + o] = gb3[o];
  }
  if (j == 0) loss_part[blockIdx.x] = loss;
}

// Supported (n_in, n_out) shapes; F(ni, no) receives the compile-time pair.
template <class F>
void dispatch_shape(uint32_t ni, uint32_t no, F&& f) {
#define MG_MLP_SHAPE(A, B) \
  if (ni == A && no == B) return f(std::integral_constant<int, A>{}, std::integral_constant<int, B>{});
  MG_MLP_SHAPE(4, 1) MG_MLP_SHAPE(4, 2) MG_MLP_SHAPE(8, 1) MG_MLP_SHAPE(8, 2) MG_MLP_SHAPE(16, 1) MG_MLP_SHAPE(16, 2)
#undef MG_MLP_SHAPE
  fail(MG_ERR_UNSUPPORTED, "mlp: (n_in, n_out) must be in {4, 8, 16} x {1, 2}");
}

// grad[i] = sum over blocks (in block order) of partial[b][i]; loss likewise.
__global__ void mlp_reduce_kernel(const float* __restrict__ partial, const float* __restrict__ loss_part,
                                  uint32_t n_blocks, uint32_t n_params, float* __restrict__ grad,
                                  float* __restrict__ loss) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_params) {
    float a = 0.f;
    for (uint32_t b = 0; b < n_blocks; ++b) a += partial[static_cast<uint64_t>(b) * n_params + i];
    grad[i] = a;
  }
  if (i == 0) {
    float l = 0.f;
    for (uint32_t b = 0; b < n_blocks; ++b) l += loss_part[b];
    *loss = l;
  }
}

// Adam step on the FP32 masters, then the BF16 copy.
__global__ void mlp_adam_kernel(uint32_t n_params, const float* __restrict__ grad, float* __restrict__ w,
                                float* __restrict__ m, float* __restrict__ v, uint16_t* __restrict__ w16, float lr,
                                float b1, float b2, float eps, float c1, float c2) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_params) return;
  const float gi = grad[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] -= upd;
  w16[i] = f32_to_bf16(w[i]);
}

__global__ void to_bf16_kernel(const float* __restrict__ w, uint32_t n, uint16_t* __restrict__ w16) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w16[i] = f32_to_bf16(w[i]);
}

}  // namespace
}  // namespace mgamd

using namespace mgamd;

struct mg_mlp {
  int device = 0;
  Dims d{};
  DevBuf<float> w, m, v, grad, loss;
  DevBuf<uint16_t> w16;
  DevBuf<float> partial, loss_part;
  uint32_t partial_blocks = 0;
  uint64_t steps = 0;
};

namespace {

void ensure_partials(mg_mlp* p, uint32_t blocks) {
  if (blocks <= p->partial_blocks) return;
  p->partial.alloc(static_cast<size_t>(blocks) * p->d.n_params);
  p->loss_part.alloc(blocks);
  p->partial_blocks = blocks;
}

// Gradient of the mean loss over n samples; returns through p->grad and p->loss.
void compute_grad(mg_context* ctx, mg_mlp* p, const float* x, const float* t, uint32_t n) {
  const uint32_t per_block = 256;  // samples per wave
  const uint32_t blocks = (n + per_block - 1) / per_block;
  ensure_partials(p, blocks);
  const float scale = 1.f / (static_cast<float>(n) * p->d.n_out);
  dispatch_shape(p->d.n_in, p->d.n_out, [&](auto ni, auto no) {
    mlp_grad_kernel<decltype(ni)::value, decltype(no)::value><<<blocks, kMlpBlock, 0, ctx->stream>>>(
        p->d, p->w16.get(), x, t, n, per_block, scale, p->partial.get(), p->loss_part.get());
  });
  MG_LAUNCH_CHECK();
  mlp_reduce_kernel<<<grid_for(p->d.n_params, 256), 256, 0, ctx->stream>>>(p->partial.get(), p->loss_part.get(),
                                                                            blocks, p->d.n_params, p->grad.get(),
                                                                            p->loss.get());
  MG_LAUNCH_CHECK();
}

}  // namespace

namespace mgamd {
void mlp_dims(const mg_mlp* p, uint32_t* n_in, uint32_t* n_out) {
  *n_in = p->d.n_in;
  *n_out = p->d.n_out;
}
}  // namespace mgamd

extern "C" int mg_mlp_create(mg_context* ctx, uint32_t n_in, uint32_t n_out, uint64_t seed, mg_mlp** out) {
  return guarded([&] {
    MG_REQUIRE(ctx && out, "mg_mlp_create: NULL argument");
    *out = nullptr;
    dispatch_shape(n_in, n_out, [](auto, auto) {});  // rejects unsupported shapes
    DeviceGuard g(ctx->device);
    auto p = std::make_unique<mg_mlp>();
    p->device = ctx->device;
    p->d = make_dims(n_in, n_out);
    const uint32_t np = p->d.n_params;
    // He-uniform initial weights from the SplitMix64 stream of `seed`; zero biases.
    std::vector<float> w(np, 0.f);
    const uint64_t key = stream_key(seed, 5);
    auto fill = [&](uint32_t off, uint32_t count, uint32_t fan_in) {
      const float lim = std::sqrt(6.f / static_cast<float>(fan_in));
      for (uint32_t i = 0; i < count; ++i) {
        const double u = static_cast<double>(sm64_at(key, off + i) >> 11) * (1.0 / 9007199254740992.0);
        w[off + i] = static_cast<float>((2.0 * u - 1.0) * lim);
      }
    };
    fill(p->d.o_w1, kH * n_in, n_in);
    fill(p->d.o_w2, kH * kH, kH);
    fill(p->d.o_w3, n_out * kH, kH);
    p->w.alloc(np);
    p->m.alloc(np);
    p->v.alloc(np);
    p->grad.alloc(np);
    p->loss.alloc(1);
    p->w16.alloc(np);
    MG_HIP(hipMemcpyAsync(p->w.get(), w.data(), sizeof(float) * np, hipMemcpyHostToDevice, ctx->stream));
    MG_HIP(hipMemsetAsync(p->m.get(), 0, sizeof(float) * np, ctx->stream));
    MG_HIP(hipMemsetAsync(p->v.get(), 0, sizeof(float) * np, ctx->stream));
    to_bf16_kernel<<<grid_for(np, 256), 256, 0, ctx->stream>>>(p->w.get(), np, p->w16.get());
    MG_LAUNCH_CHECK();
    MG_HIP(hipStreamSynchronize(ctx->stream));
    *out = p.release();
  });
}

extern "C" int mg_mlp_num_params(const mg_mlp* p, uint32_t* n_params) {
  return guarded([&] {
    MG_REQUIRE(p && n_params, "mg_mlp_num_params: NULL argument");
    *n_params = p->d.n_params;
  });
}

extern "C" int mg_mlp_get_params(mg_context* ctx, const mg_mlp* p, float* master, uint16_t* bf16) {
  return guarded([&] {
    MG_REQUIRE(ctx && p, "mg_mlp_get_params: NULL argument");
    DeviceGuard g(ctx->device);
    if (master)
      MG_HIP(hipMemcpyAsync(master, p->w.get(), sizeof(float) * p->d.n_params, hipMemcpyDeviceToHost, ctx->stream));
    if (bf16)
      MG_HIP(hipMemcpyAsync(bf16, p->w16.get(), sizeof(uint16_t) * p->d.n_params, hipMemcpyDeviceToHost, ctx->stream));
    MG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

extern "C" int mg_mlp_set_params(mg_context* ctx, mg_mlp* p, const float* master) {
  return guarded([&] {
    MG_REQUIRE(ctx && p && master, "mg_mlp_set_params: NULL argument");
    DeviceGuard g(ctx->device);
    MG_HIP(hipMemcpyAsync(p->w.get(), master, sizeof(float) * p->d.n_params, hipMemcpyHostToDevice, ctx->stream));
    to_bf16_kernel<<<grid_for(p->d.n_params, 256), 256, 0, ctx->stream>>>(p->w.get(), p->d.n_params, p->w16.get());
    MG_LAUNCH_CHECK();
    MG_HIP(hipStreamSynchronize(ctx->stream));
  });
}

extern "C" int mg_mlp_forward(mg_context* ctx, const mg_mlp* p, const float* x, uint32_t n, float* y) {
  return guarded([&] {
    MG_REQUIRE(ctx && p && (n == 0 || (x && y)), "mg_mlp_forward: NULL argument");
    if (n == 0) return;
    DeviceGuard g(ctx->device);
    dispatch_shape(p->d.n_in, p->d.n_out, [&](auto ni, auto no) {
      mlp_forward_kernel<decltype(ni)::value, decltype(no)::value>
          <<<grid_for(n, 64, 1u << 30), kMlpBlock, 0, ctx->stream>>>(p->d, p->w16.get(), x, n, y);
    });
    MG_LAUNCH_CHECK();
  });
}

extern "C" int mg_mlp_loss_grad(mg_context* ctx, mg_mlp* p, const float* x, const float* t, uint32_t n, float* loss,
                                float* grad) {
  return guarded([&] {
    MG_REQUIRE(ctx && p && x && t && n > 0, "mg_mlp_loss_grad: bad argument");
    DeviceGuard g(ctx->device);
    compute_grad(ctx, p, x, t, n);
    float l = 0.f;
    MG_HIP(hipMemcpyAsync(&l, p->loss.get(), sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    if (grad) MG_HIP(hipMemcpyAsync(grad, p->grad.get(), sizeof(float) * p->d.n_params, hipMemcpyDeviceToHost, ctx->stream));
    MG_HIP(hipStreamSynchronize(ctx->stream));
    if (loss) *loss = l / (static_cast<float>(n) * p->d.n_out);
  });
}

extern "C" int mg_mlp_train(mg_context* ctx, mg_mlp* p, const float* x, const float* t, uint32_t n, uint32_t epochs,
                            uint32_t batch, float lr, float* final_loss) {
  return guarded([&] {
    MG_REQUIRE(ctx && p && x && t && n > 0 && batch > 0, "mg_mlp_train: bad argument");
    DeviceGuard g(ctx->device);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (uint32_t ep = 0; ep < epochs; ++ep) {
      for (uint32_t s0 = 0; s0 < n; s0 += batch) {
        const uint32_t nb = std::min(batch, n - s0);
        compute_grad(ctx, p, x + static_cast<uint64_t>(s0) * p->d.n_in, t + static_cast<uint64_t>(s0) * p->d.n_out, nb);
        ++p->steps;
        const float c1 = 1.f - std::pow(b1, static_cast<float>(p->steps));
        const float c2 = 1.f - std::pow(b2, static_cast<float>(p->steps));
        mlp_adam_kernel<<<grid_for(p->d.n_params, 256), 256, 0, ctx->stream>>>(
            p->d.n_params, p->grad.get(), p->w.get(), p->m.get(), p->v.get(), p->w16.get(), lr, b1, b2, eps, c1, c2);
        MG_LAUNCH_CHECK();
      }
    }
    if (final_loss) {
      compute_grad(ctx, p, x, t, n);
      float l = 0.f;
      MG_HIP(hipMemcpyAsync(&l, p->loss.get(), sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
      MG_HIP(hipStreamSynchronize(ctx->stream));
      *final_loss = l / (static_cast<float>(n) * p->d.n_out);
    }
  });
}

extern "C" int mg_mlp_destroy(mg_mlp* p) {
  return guarded([&] {
    if (!p) return;
    DeviceGuard g(p->device);
    delete p;
  });
}
