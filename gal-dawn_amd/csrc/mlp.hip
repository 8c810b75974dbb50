// FP32 MLP 8 -> 32 -> 32 -> 3 (ReLU) for the step scheduler: forward,
// mean-squared-error gradient and Adam training, as hand-written HIP
// kernels.  No counterpart in the reference (its push/pull choice is fixed).
//
// Gradient in two deterministic phases (no floating-point atomics):
//   phase 1  one thread per sample: forward pass, stores activations and the
//            back-propagated deltas of the three layers;
//   phase 2  one thread per parameter: sums delta x activation over samples.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "host_util.hpp"

namespace dawn {
namespace {

constexpr int IN = 8, H = 32, OUT = 3;
constexpr int OFF_W1 = 0, OFF_B1 = H * IN, OFF_W2 = OFF_B1 + H, OFF_B2 = OFF_W2 + H * H,
              OFF_W3 = OFF_B2 + H, OFF_B3 = OFF_W3 + OUT * H, NP = OFF_B3 + OUT;
static_assert(NP == DAWN_SCHED_PARAMS, "parameter layout");
// Per-sample record of phase 1: h1[32] h2[32] d1[32] d2[32] d3[3] loss[1].
constexpr int REC = H + H + H + H + OUT + 1;

__device__ void forward_row(const float* p, const float* x, float* h1, float* h2, float* o) {
  for (int i = 0; i < H; i++) {
    float a = p[OFF_B1 + i];
    for (int j = 0; j < IN; j++) a = fmaf(p[OFF_W1 + i * IN + j], x[j], a);
    h1[i] = a > 0.f ? a : 0.f;
  }
  for (int i = 0; i < H; i++) {
    float a = p[OFF_B2 + i];
    for (int j = 0; j < H; j++) a = fmaf(p[OFF_W2 + i * H + j], h1[j], a);
    h2[i] = a > 0.f ? a : 0.f;
  }
  for (int i = 0; i < OUT; i++) {
    float a = p[OFF_B3 + i];
    for (int j = 0; j < H; j++) a = fmaf(p[OFF_W3 + i * H + j], h2[j], a);
    o[i] = a;
  }
}

__global__ void mlp_forward_kernel(const float* p, const float* x, int N, float* out) {
  __shared__ float sp[NP];
  for (int i = threadIdx.x; i < NP; i += blockDim.x) sp[i] = p[i];
  __syncthreads();
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= N) return;
  float h1[H], h2[H], o[OUT];
  forward_row(sp, x + (int64_t)r * IN, h1, h2, o);
  for (int i = 0; i < OUT; i++) out[(int64_t)r * OUT + i] = o[i];
}

// Phase 1 over samples idx[0..B) (idx null: 0..B).
__global__ void mlp_phase1_kernel(const float* p, const float* x, const float* y,
                                  const int32_t* idx, int B, float scale, float* rec) {
  __shared__ float sp[NP];
  for (int i = threadIdx.x; i < NP; i += blockDim.x) sp[i] = p[i];
  __syncthreads();
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= B) return;
  const int64_t s = idx ? idx[r] : r;
  const float* xr = x + s * IN;
  float h1[H], h2[H], o[OUT], d1[H], d2[H], d3[OUT];
  forward_row(sp, xr, h1, h2, o);
  float loss = 0.f;
  for (int i = 0; i < OUT; i++) {
    const float e = o[i] - y[s * OUT + i];
    loss += e * e * scale;
    d3[i] = 2.f * e * scale;
  }
  for (int j = 0; j < H; j++) {
    float a = 0.f;
    for (int i = 0; i < OUT; i++) a = fmaf(sp[OFF_W3 + i * H + j], d3[i], a);
    d2[j] = h2[j] > 0.f ? a : 0.f;
  }
  for (int j = 0; j < H; j++) {
    float a = 0.f;
    for (int i = 0; i < H; i++) a = fmaf(sp[OFF_W2 + i * H + j], d2[i], a);
    d1[j] = h1[j] > 0.f ? a : 0.f;
  }
  float* rr = rec + (int64_t)r * REC;
  for (int i = 0; i < H; i++) {
    rr[i] = h1[i];
    rr[H + i] = h2[i];
    rr[2 * H + i] = d1[i];
    rr[3 * H + i] = d2[i];
  }
  for (int i = 0; i < OUT; i++) rr[4 * H + i] = d3[i];
  rr[4 * H + OUT] = loss;
}

// Phase 2: thread t < NP sums its parameter's gradient; thread NP sums the
// loss.  grad[t] = sum_r delta_r[i] * act_r[j].
__global__ void mlp_phase2_kernel(const float* x, const int32_t* idx, int B, const float* rec,
                                  float* grad, float* loss) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > NP) return;
  float acc = 0.f;
  if (t == NP) {
    for (int r = 0; r < B; r++) acc += rec[(int64_t)r * REC + 4 * H + OUT];
    *loss = acc;
    return;
  }
  for (int r = 0; r < B; r++) {
    const float* rr = rec + (int64_t)r * REC;
    float g;
    if (t < OFF_B1) {
      const int i = t / IN, j = t % IN;
      const int64_t s = idx ? idx[r] : r;
      g = rr[2 * H + i] * x[s * IN + j];
    } else if (t < OFF_W2) {
      g = rr[2 * H + (t - OFF_B1)];
    } else if (t < OFF_B2) {
      const int u = t - OFF_W2, i = u / H, j = u % H;
      g = rr[3 * H + i] * rr[j];
    } else if (t < OFF_W3) {
      g = rr[3 * H + (t - OFF_B2)];
    } else if (t < OFF_B3) {
      const int u = t - OFF_W3, i = u / H, j = u % H;
      g = rr[4 * H + i] * rr[H + j];
    } else {
      g = rr[4 * H + (t - OFF_B3)];
    }
    acc += g;
  }
  grad[t] = acc;
}

__global__ void adam_kernel(float* p, const float* g, float* m, float* v, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= NP) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.f - b1) * gt;
  const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

struct MlpDev {
  DevBuf<float> p, x, y, out, rec, grad, m, v, loss;
  DevBuf<int32_t> idx;
};

int upload(DevBuf<float>& b, const float* h, size_t n) {
  DAWN_TRY(b.alloc(n));
  DAWN_HIP(hipMemcpy(b.p, h, sizeof(float) * n, hipMemcpyHostToDevice));
  return DAWN_OK;
}

int check_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    return DAWN_ERR_NO_DEVICE;
  }
  return DAWN_OK;
}

}  // namespace
}  // namespace dawn

using namespace dawn;

extern "C" {

int dawn_mlp_forward(const float* params, const float* x, int32_t N, float* out) {
  if (!params || N < 0 || (N > 0 && (!x || !out))) return DAWN_ERR_INVALID_ARG;
  DAWN_TRY(check_device());
  if (N == 0) return DAWN_OK;
  MlpDev d;
  DAWN_TRY(upload(d.p, params, NP));
  DAWN_TRY(upload(d.x, x, (size_t)N * IN));
  DAWN_TRY(d.out.alloc((size_t)N * OUT));
  mlp_forward_kernel<<<(N + 127) / 128, 128>>>(d.p.p, d.x.p, N, d.out.p);
  DAWN_HIP(hipGetLastError());
  DAWN_HIP(hipMemcpy(out, d.out.p, sizeof(float) * N * OUT, hipMemcpyDeviceToHost));
  return DAWN_OK;
}

int dawn_mlp_loss_grad(const float* params, const float* x, const float* y, int32_t N,
                       float* loss, float* grad) {
  if (!params || !x || !y || N <= 0 || !loss || !grad) return DAWN_ERR_INVALID_ARG;
  DAWN_TRY(check_device());
  MlpDev d;
  DAWN_TRY(upload(d.p, params, NP));
  DAWN_TRY(upload(d.x, x, (size_t)N * IN));
  DAWN_TRY(upload(d.y, y, (size_t)N * OUT));
  DAWN_TRY(d.rec.alloc((size_t)N * REC));
  DAWN_TRY(d.grad.alloc(NP));
  DAWN_TRY(d.loss.alloc(1));
  const float scale = 1.f / ((float)N * OUT);
  mlp_phase1_kernel<<<(N + 127) / 128, 128>>>(d.p.p, d.x.p, d.y.p, nullptr, N, scale, d.rec.p);
  mlp_phase2_kernel<<<(NP + 1 + 127) / 128, 128>>>(d.x.p, nullptr, N, d.rec.p, d.grad.p, d.loss.p);
  DAWN_HIP(hipGetLastError());
  DAWN_HIP(hipMemcpy(grad, d.grad.p, sizeof(float) * NP, hipMemcpyDeviceToHost));
  DAWN_HIP(hipMemcpy(loss, d.loss.p, sizeof(float), hipMemcpyDeviceToHost));
  return DAWN_OK;
}

int dawn_mlp_train(float* params, const float* x, const float* y, int32_t N, int32_t epochs,
                   int32_t batch, float lr, uint64_t seed, float* final_loss) {
  if (!params || !x || !y || N <= 0 || epochs < 0 || batch <= 0) return DAWN_ERR_INVALID_ARG;
  DAWN_TRY(check_device());
  MlpDev d;
  DAWN_TRY(upload(d.p, params, NP));
  DAWN_TRY(upload(d.x, x, (size_t)N * IN));
  DAWN_TRY(upload(d.y, y, (size_t)N * OUT));
  const int B = std::min(batch, N);
  DAWN_TRY(d.rec.alloc((size_t)std::max(B, N) * REC));
  DAWN_TRY(d.grad.alloc(NP));
  DAWN_TRY(d.m.alloc(NP));
  DAWN_TRY(d.v.alloc(NP));
  DAWN_TRY(d.loss.alloc(1));
  DAWN_TRY(d.idx.alloc((size_t)N));
  DAWN_HIP(hipMemset(d.m.p, 0, sizeof(float) * NP));
  DAWN_HIP(hipMemset(d.v.p, 0, sizeof(float) * NP));
  std::vector<int32_t> perm(N);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  long long t = 0;
  for (int e = 0; e < epochs; e++) {
    for (int32_t i = 0; i < N; i++) perm[i] = i;
    for (int32_t i = N - 1; i > 0; i--) {  // seeded Fisher-Yates
      uint64_t r = rng_u64(seed, (uint64_t)e, (uint64_t)i);
      int32_t j = (int32_t)(r % (uint64_t)(i + 1));
      std::swap(perm[i], perm[j]);
    }
    DAWN_HIP(hipMemcpy(d.idx.p, perm.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice));
    for (int32_t b0 = 0; b0 < N; b0 += B) {
      const int nb = std::min(B, N - b0);
      const float scale = 1.f / ((float)nb * OUT);
      mlp_phase1_kernel<<<(nb + 127) / 128, 128>>>(d.p.p, d.x.p, d.y.p, d.idx.p + b0, nb, scale,
                                                    d.rec.p);
      mlp_phase2_kernel<<<(NP + 1 + 127) / 128, 128>>>(d.x.p, d.idx.p + b0, nb, d.rec.p,
                                                       d.grad.p, d.loss.p);
      t++;
      const float c1 = 1.f - std::pow(b1, (float)t), c2 = 1.f - std::pow(b2, (float)t);
      adam_kernel<<<(NP + 127) / 128, 128>>>(d.p.p, d.grad.p, d.m.p, d.v.p, lr, b1, b2, eps, c1,
                                             c2);
    }
  }
  DAWN_HIP(hipGetLastError());
  if (final_loss) {
    const float scale = 1.f / ((float)N * OUT);
    mlp_phase1_kernel<<<(N + 127) / 128, 128>>>(d.p.p, d.x.p, d.y.p, nullptr, N, scale, d.rec.p);
    mlp_phase2_kernel<<<(NP + 1 + 127) / 128, 128>>>(d.x.p, nullptr, N, d.rec.p, d.grad.p,
                                                     d.loss.p);
    DAWN_HIP(hipMemcpy(final_loss, d.loss.p, sizeof(float), hipMemcpyDeviceToHost));
  }
  DAWN_HIP(hipMemcpy(params, d.p.p, sizeof(float) * NP, hipMemcpyDeviceToHost));
  return DAWN_OK;
}

}  // extern "C"
