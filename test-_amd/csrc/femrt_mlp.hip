// femrt_mlp.hip -- training and inference of the library's networks (the
// Newton-guess network and the isosurface-bracket network).
//
//   k_mlp_pack     FP32 master parameters -> BF16 operand layouts
//   k_mlp_infer    batched inference, 64 queries per wave (the same device
//                  function the element walk calls for its guesses)
//   k_mlp_train    forward + backward of 64 samples per wave on BF16 MFMA
//                  with FP32 accumulation; gradient sums by atomics
//                  (squared error for the guess network, pinball loss at the
//                  5% and 95% quantiles for the bracket network)
//   k_mlp_adam     Adam step on the master parameters from the per-block partial
//                  gradients, summed in block order (reproducible training)
//   k_mlp_samples  training pairs from the tracer's own point inversions:
//                  features of x = X_e(rst) -> Newton's preimage of x
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "femrt_mlp.hpp"
#include "femrt_runtime.hpp"

using namespace femrt;

// ================================================================== kernels

template <int KIND>
__global__ void k_mlp_pack(const float* __restrict__ p, void* __restrict__ pack)
{
  using N = Net<KIND>;
  constexpr int R1 = N::W1_ROWS, OUT = N::OUT;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  __bf16* pb = (__bf16*)pack;
  float* pf = (float*)(pb + PK_NBF16);
  const float* W1 = p + N::OFF_W1;
  const float* W2 = p + N::OFF_W2;
  const float* W3 = p + N::OFF_W3;
  if (t < PK_W2OP) {
    const int x = t & 7, lane = (t >> 3) & 63, q = (t >> 9) & 1, a = t >> 10;
    const int k = 16 * q + 8 * (lane >> 5) + x, u = 32 * a + (lane & 31);
    pb[t] = (__bf16)(k < R1 ? W1[k * 64 + u] : 0.f);
  } else if (t < PK_W3OP) {
    const int idx = t - PK_W2OP;
    const int x = idx & 7, lane = (idx >> 3) & 63, q = (idx >> 9) & 3, a = idx >> 11;
    pb[t] = (__bf16)W2[mlp_unit(q, lane >> 5, x) * 64 + 32 * a + (lane & 31)];
  } else if (t < PK_W1B) {
    const int idx = t - PK_W3OP;
    const int x = idx & 7, lane = (idx >> 3) & 63, q = idx >> 9, o = lane & 31;
    pb[t] = (__bf16)(o < OUT ? W3[mlp_unit(q, lane >> 5, x) * OUT + o] : 0.f);
  } else if (t < PK_W2B) {
    const int idx = t - PK_W1B, k = idx >> 6, u = idx & 63;
    pb[t] = (__bf16)(k < R1 ? W1[k * 64 + u] : 0.f);
  } else if (t < PK_W3B) {
    pb[t] = (__bf16)W2[t - PK_W2B];
  } else if (t < PK_NBF16) {
    const int idx = t - PK_W3B, u = idx >> 5, o = idx & 31;
    pb[t] = (__bf16)(o < OUT ? W3[u * OUT + o] : 0.f);
  } else if (t < PK_NBF16 + PK_NF32_ALL) {
    const int idx = t - PK_NBF16;
    float v = 0.f;
    if (idx < 64) v = p[N::OFF_B2 + idx];
    else if (idx < 64 + OUT) v = p[N::OFF_B3 + idx - 64];
    else if (idx >= PK_B1 && N::OFF_B1 >= 0) v = p[N::OFF_B1 + idx - PK_B1];
    pf[idx] = v;
  }
}

// 256 threads = 4 waves; every lane reaches wave_mlp_forward.
template <int KIND>
__global__ __launch_bounds__(256) void k_mlp_infer(int n, const float* __restrict__ feats,
                                                   const void* __restrict__ pack,
                                                   float* __restrict__ out)
{
  constexpr int IN = Net<KIND>::IN, OUT = Net<KIND>::OUT;
  __shared__ bf16x8 ops[PK_INFER_BF16 / 8];
  __shared__ bf16x8 xs[4][64 * 4];
  __shared__ float bias[PK_NF32_ALL];
  const bf16x8* g = (const bf16x8*)pack;
  for (int t = threadIdx.x; t < PK_INFER_BF16 / 8; t += blockDim.x) ops[t] = g[t];
  const float* gf = (const float*)((const __bf16*)pack + PK_NBF16);
  for (int t = threadIdx.x; t < PK_NF32_ALL; t += blockDim.x) bias[t] = gf[t];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  float f[IN];
  for (int k = 0; k < IN; ++k) f[k] = s < n ? feats[(size_t)s * IN + k] : 0.f;
  mlp_stage_inputs<KIND>(xs[wave], lane, f);
  __syncthreads();
  float y0[OUT], y1[OUT];
  wave_mlp_forward<KIND>(xs[wave], ops, bias, bias + 64, y0, y1, bias + PK_B1);
  if (lane < 32) {
    const int base = blockIdx.x * blockDim.x + wave * 64;
    const int s0 = base + lane, s1 = base + 32 + lane;
    if (s0 < n)
      for (int c = 0; c < OUT; ++c) out[(size_t)s0 * OUT + c] = y0[c];
    if (s1 < n)
      for (int c = 0; c < OUT; ++c) out[(size_t)s1 * OUT + c] = y1[c];
  }
}

// One wave (64 samples) per block.  Samples are the positions
// [base + 64 blockIdx, ...) < end of the permutation j -> (pa j + pb) mod n.
// Loss: inv_count * sum over samples and outputs of (y - target)^2 (guess
// network) or of the pinball loss rho_tau(target - y), rho_tau(r) =
// max(tau r, (tau - 1) r), at tau = 0.05 and 0.95 (bracket network).
template <int KIND>
__global__ __launch_bounds__(64) void k_mlp_train(const float* __restrict__ feats,
                                                  const float* __restrict__ targets, int n,
                                                  int base, int end, unsigned pa, unsigned pbias,
                                                  float inv_count, const void* __restrict__ pack,
                                                  float* __restrict__ part)
{
  // this block's partial gradient (and loss in the last slot): every entry is
  // written once by the wave, so the sum over blocks (k_mlp_adam,
  // k_mlp_reduce) runs in block order and training is reproducible
  float* __restrict__ grad = part + (size_t)blockIdx.x * (Net<KIND>::NPARAM + 1);
  __shared__ __bf16 Xb[64 * 32], H1b[64 * 64], H2b[64 * 64], GZb[64 * 64], GYb[64 * 32];
  __shared__ float H1f[64 * 64], H2f[64 * 64], T[64 * 4], GY[64 * 4];
  using N = Net<KIND>;
  constexpr int IN = N::IN, OUT = N::OUT;
  const int lane = threadIdx.x, i = lane & 31;
  const __bf16* pk = (const __bf16*)pack;
  const __bf16* W1b = pk + PK_W1B;
  const __bf16* W2b = pk + PK_W2B;
  const __bf16* W3b = pk + PK_W3B;
  const float* b2 = (const float*)(pk + PK_NBF16);
  const float* b3 = b2 + 64;
  const float* b1 = b2 + PK_B1;
  const int row0 = base + blockIdx.x * 64;
  {
    const int j = row0 + lane;
    float f[IN], t[OUT];
    for (int c = 0; c < OUT; ++c) t[c] = 0.f;
    if (j < end) {
      const size_t src = (size_t)(((unsigned long long)pa * (unsigned)j + pbias) % (unsigned)n);
      for (int k = 0; k < IN; ++k) f[k] = feats[src * IN + k];
      for (int c = 0; c < OUT; ++c) t[c] = targets[src * OUT + c];
    } else {
      for (int k = 0; k < IN; ++k) f[k] = 0.f;
    }
    for (int k = 0; k < 32; ++k)
      Xb[lane * 32 + k] = (__bf16)(k < IN ? f[k] : (N::BIAS_INPUT && k == IN && j < end ? 1.f : 0.f));
    for (int c = 0; c < OUT; ++c) T[lane * 4 + c] = t[c];
  }
  __syncthreads();
  // layer 1: Z1 = X W1 (+ b1; the guess network's bias through its constant input)
  for (int st = 0; st < 2; ++st)
    for (int ut = 0; ut < 2; ++ut) {
      f32x16 acc = {};
      if constexpr (!N::BIAS_INPUT)
        for (int v = 0; v < 16; ++v) acc[v] = b1[ut * 32 + i];
      acc = tile_gemm(Xb + st * 32 * 32, 32, 1, W1b + ut * 32, 64, 1, 32, acc);
      for (int v = 0; v < 16; ++v) {
        const int s = st * 32 + tile_row(v, lane), u = ut * 32 + i;
        const float hv = tanh_fast(acc[v]);
        H1f[s * 64 + u] = hv;
        H1b[s * 64 + u] = (__bf16)hv;
      }
    }
  __syncthreads();
  // layer 2
  for (int st = 0; st < 2; ++st)
    for (int ut = 0; ut < 2; ++ut) {
      f32x16 acc;
      for (int v = 0; v < 16; ++v) acc[v] = b2[ut * 32 + i];
      acc = tile_gemm(H1b + st * 32 * 64, 64, 1, W2b + ut * 32, 64, 1, 64, acc);
      for (int v = 0; v < 16; ++v) {
        const int s = st * 32 + tile_row(v, lane), u = ut * 32 + i;
        const float hv = tanh_fast(acc[v]);
        H2f[s * 64 + u] = hv;
        H2b[s * 64 + u] = (__bf16)hv;
      }
    }
  __syncthreads();
  // layer 3 and the loss gradient: dL/dy = 2 (y - t) inv_count (squared
  // error), -tau inv_count where t > y else (1 - tau) inv_count (pinball)
  float lsum = 0.f;
  for (int st = 0; st < 2; ++st) {
    f32x16 acc;
    for (int v = 0; v < 16; ++v) acc[v] = i < OUT ? b3[i] : 0.f;
    acc = tile_gemm(H2b + st * 32 * 64, 64, 1, W3b, 32, 1, 64, acc);
    for (int v = 0; v < 16; ++v) {
      const int s = st * 32 + tile_row(v, lane);
      float g = 0.f;
      if (i < OUT && row0 + s < end) {
        if constexpr (N::PINBALL) {
          const float tau = i == 0 ? BRACKET_TAU0 : BRACKET_TAU1;
          const float r = T[s * 4 + i] - acc[v];
          lsum += fmaxf(tau * r, (tau - 1.f) * r);
          g = (r > 0.f ? -tau : 1.f - tau) * inv_count;
        } else {
          const float diff = acc[v] - T[s * 4 + i];
          lsum += diff * diff;
          g = 2.f * diff * inv_count;
        }
      }
      if (i < OUT) GY[s * 4 + i] = g;
      GYb[s * 32 + i] = (__bf16)g;
    }
  }
  __syncthreads();
  // dW3 = H2^T GY, db3 = sum GY
  for (int ut = 0; ut < 2; ++ut) {
    f32x16 acc = {};
    acc = tile_gemm(H2b + ut * 32, 1, 64, GYb, 32, 1, 64, acc);
    if (i < OUT)
      for (int v = 0; v < 16; ++v)
        grad[N::OFF_W3 + (ut * 32 + tile_row(v, lane)) * OUT + i] = acc[v];
  }
  if (lane < OUT) {
    float s = 0.f;
    for (int r = 0; r < 64; ++r) s += GY[r * 4 + lane];
    grad[N::OFF_B3 + lane] = s;
  }
  // GZ2 = (GY W3^T) * (1 - H2^2)
  for (int st = 0; st < 2; ++st)
    for (int ut = 0; ut < 2; ++ut) {
      f32x16 acc = {};
      acc = tile_gemm(GYb + st * 32 * 32, 32, 1, W3b + ut * 32 * 32, 1, 32, 32, acc);
      for (int v = 0; v < 16; ++v) {
        const int s = st * 32 + tile_row(v, lane), u = ut * 32 + i;
        const float hv = H2f[s * 64 + u];
        const float gz = acc[v] * (1.f - hv * hv);
        H2f[s * 64 + u] = gz;
        GZb[s * 64 + u] = (__bf16)gz;
      }
    }
  __syncthreads();
  {
    float s = 0.f;
    for (int r = 0; r < 64; ++r) s += H2f[r * 64 + lane];
    grad[N::OFF_B2 + lane] = s;
  }
  // dW2 = H1^T GZ2
  for (int it = 0; it < 2; ++it)
    for (int ot = 0; ot < 2; ++ot) {
      f32x16 acc = {};
      acc = tile_gemm(H1b + it * 32, 1, 64, GZb + ot * 32, 64, 1, 64, acc);
      for (int v = 0; v < 16; ++v)
        grad[N::OFF_W2 + (it * 32 + tile_row(v, lane)) * 64 + ot * 32 + i] = acc[v];
    }
  // GZ1 = (GZ2 W2^T) * (1 - H1^2)
  f32x16 gh[2][2];
  for (int st = 0; st < 2; ++st)
    for (int it = 0; it < 2; ++it) {
      f32x16 acc = {};
      gh[st][it] = tile_gemm(GZb + st * 32 * 64, 64, 1, W2b + it * 32 * 64, 1, 64, 64, acc);
    }
  __syncthreads();
  for (int st = 0; st < 2; ++st)
    for (int it = 0; it < 2; ++it)
      for (int v = 0; v < 16; ++v) {
        const int s = st * 32 + tile_row(v, lane), u = it * 32 + i;
        const float hv = H1f[s * 64 + u];
        const float gz = gh[st][it][v] * (1.f - hv * hv);
        if constexpr (!N::BIAS_INPUT) H1f[s * 64 + u] = gz;
        GZb[s * 64 + u] = (__bf16)gz;
      }
  __syncthreads();
  if constexpr (!N::BIAS_INPUT) {
    // db1 = sum over samples of GZ1
    float s = 0.f;
    for (int r = 0; r < 64; ++r) s += H1f[r * 64 + lane];
    grad[N::OFF_B1 + lane] = s;
  }
  // dW1 = X^T GZ1 (the guess network's row 27 is its first-layer bias)
  for (int ut = 0; ut < 2; ++ut) {
    f32x16 acc = {};
    acc = tile_gemm(Xb, 1, 32, GZb + ut * 32, 64, 1, 64, acc);
    for (int v = 0; v < 16; ++v) {
      const int k = tile_row(v, lane);
      if (k < N::W1_ROWS) grad[N::OFF_W1 + k * 64 + ut * 32 + i] = acc[v];
    }
  }
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
  if (lane == 0) grad[N::NPARAM] = lsum * inv_count;
}

// Sum of the per-block partial gradients (stride n + 1, the loss last), in
// block order.
__device__ __forceinline__ float part_sum(const float* __restrict__ part, int blocks, int n, int t)
{
  float s = 0.f;
  for (int b = 0; b < blocks; ++b) s += part[(size_t)b * (n + 1) + t];
  return s;
}

__global__ void k_mlp_reduce(const float* __restrict__ part, int blocks, int n, float* __restrict__ out)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t <= n) out[t] = part_sum(part, blocks, n, t);
}

// Adam on the master parameters from the step's partial gradients.  Thread n
// adds the step's loss to loss_acc.
__global__ void k_mlp_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ part, int blocks, int n, float lr, float c1, float c2,
                           float* __restrict__ loss_acc)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  if (t < n) {
    const float gi = part_sum(part, blocks, n, t);
    const float mi = b1 * m[t] + (1.f - b1) * gi;
    const float vi = b2 * v[t] + (1.f - b2) * gi * gi;
    m[t] = mi;
    v[t] = vi;
    p[t] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  } else if (t == n) {
    if (loss_acc) *loss_acc += part_sum(part, blocks, n, n);
  }
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ float unit_float(unsigned long long z)
{
  return (float)(z >> 40) * (1.f / 16777216.f);  // [0, 1)
}

template <int NG>
__global__ void k_mlp_samples(TraceMesh M, int n_elem, int n, unsigned long long seed,
                              float* __restrict__ feats, float* __restrict__ targets,
                              long long* __restrict__ work)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const unsigned long long h0 = mix64(seed ^ mix64((unsigned long long)s));
  const int e = (int)(mix64(h0 ^ 1) % (unsigned long long)n_elem);
  float rt[3];
  for (int c = 0; c < 3; ++c) rt[c] = 2.f * unit_float(mix64(h0 + 2 + c)) - 1.f;
  const GeoMode* gc = M.geom + (size_t)e * (NG * NG * NG);
  float x[3];
  geom_point<float, NG>(gc, rt[0], rt[1], rt[2], x);
  float rst[3] = {0.f, 0.f, 0.f}, J[9];
  unsigned it = 0;
  const bool ok = invert_in_cube_f32<NG>(gc, x, rst, J, it);
  float f[MLP_IN];
  mlp_features(M.boxes + (size_t)e * FEMRT_BOX_STRIDE, M.verts + (size_t)e * 24, x, f);
  for (int k = 0; k < MLP_IN; ++k) feats[(size_t)s * MLP_IN + k] = f[k];
  for (int c = 0; c < 3; ++c) targets[(size_t)s * 3 + c] = ok ? rst[c] : rt[c];
  if (work) atomicAdd((unsigned long long*)&work[3], (unsigned long long)it);
}

// ================================================================== host

// Shapes of the network kinds (FEMRT_MLP_GUESS, FEMRT_MLP_BRACKET).
static int net_nparam(int kind) { return kind == NET_BRACKET ? Net<NET_BRACKET>::NPARAM : Net<NET_GUESS>::NPARAM; }
static int net_in(int kind) { return kind == NET_BRACKET ? Net<NET_BRACKET>::IN : Net<NET_GUESS>::IN; }
static int net_out(int kind) { return kind == NET_BRACKET ? Net<NET_BRACKET>::OUT : Net<NET_GUESS>::OUT; }
static void mlp_free(femrt_mlp* m)
{
  if (m->d_param) (void)hipFree(m->d_param);
  if (m->d_m) (void)hipFree(m->d_m);
  if (m->d_v) (void)hipFree(m->d_v);
  if (m->d_grad) (void)hipFree(m->d_grad);
  if (m->d_part) (void)hipFree(m->d_part);
  if (m->d_pack) (void)hipFree(m->d_pack);
}

// Per-block partial gradients for `blocks` training blocks (grown on demand).
static float* mlp_part(femrt_mlp* m, int blocks)
{
  if (blocks > m->part_blocks) {
    if (m->d_part) (void)hipFree(m->d_part);
    m->d_part = nullptr;
    m->part_blocks = 0;
    if (hipMalloc(&m->d_part, (size_t)blocks * (net_nparam(m->kind) + 1) * sizeof(float)) != hipSuccess)
      return nullptr;
    m->part_blocks = blocks;
  }
  return m->d_part;
}

#define FEMRT_NET_DISPATCH(KINDV, BODY)                                     \
  if ((KINDV) == NET_BRACKET) { constexpr int K_ = NET_BRACKET; BODY; }   \
  else { constexpr int K_ = NET_GUESS; BODY; }

static int mlp_pack(femrt_mlp* m)
{
  femrt_ctx* c = m->ctx;
  const int nt = PK_NBF16 + PK_NF32_ALL;
  return timed_launch(c, "k_mlp_pack", [&] {
    FEMRT_NET_DISPATCH(m->kind, (k_mlp_pack<K_><<<(nt + 255) / 256, 256, 0, c->stream>>>(m->d_param, m->d_pack)));
  });
}

static int mlp_reset_state(femrt_mlp* m)
{
  femrt_ctx* c = m->ctx;
  const int np = net_nparam(m->kind);
  HIPCHK(c, hipMemsetAsync(m->d_m, 0, np * sizeof(float), c->stream));
  HIPCHK(c, hipMemsetAsync(m->d_v, 0, np * sizeof(float), c->stream));
  m->adam_t = 0;
  return mlp_pack(m);
}

// Glorot-uniform weights from a keyed generator, zero biases.
static void mlp_init_params(int kind, unsigned long long seed, std::vector<float>& p)
{
  p.assign(net_nparam(kind), 0.f);
  auto mix = [](unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  auto fill = [&](int off, int rows, int cols, int fan_in, int fan_out) {
    const double a = std::sqrt(6.0 / (fan_in + fan_out));
    for (int k = 0; k < rows * cols; ++k) {
      const double u = (double)(mix(seed ^ mix((unsigned long long)(off + k))) >> 11) * 0x1.0p-53;
      p[off + k] = (float)((2.0 * u - 1.0) * a);
    }
  };
  if (kind == NET_BRACKET) {
    using N = Net<NET_BRACKET>;
    fill(N::OFF_W1, 32, 64, 32, 64);
    fill(N::OFF_W2, 64, 64, 64, 64);
    fill(N::OFF_W3, 64, 2, 64, 2);
    return;
  }
  fill(MLP_OFF_W1, 27, 64, 27, 64);  // row 27 (bias) stays zero
  fill(MLP_OFF_W2, 64, 64, 64, 64);
  fill(MLP_OFF_W3, 64, 3, 64, 3);
}

template <int NG>
static void launch_samples(femrt_mesh* mesh, int n, unsigned long long seed, float* feats,
                           float* targets)
{
  TraceMesh TM{mesh->d_geom, mesh->d_field, mesh->d_boxes, mesh->d_verts, mesh->d_adj,
               mesh->d_nodes, mesh->d_prims};
  k_mlp_samples<NG><<<(n + 255) / 256, 256, 0, mesh->ctx->stream>>>(TM, mesh->n_elem, n, seed,
                                                                     feats, targets, nullptr);
}

static int make_samples_dev(femrt_mesh* mesh, int n, unsigned long long seed, float* feats,
                            float* targets)
{
  femrt_ctx* c = mesh->ctx;
  return timed_launch(c, "k_mlp_samples", [&] {
    switch (mesh->NG) {
      case 4: launch_samples<4>(mesh, n, seed, feats, targets); break;
      case 5: launch_samples<5>(mesh, n, seed, feats, targets); break;
      default: launch_samples<7>(mesh, n, seed, feats, targets); break;
    }
  });
}

static unsigned gcd_u(unsigned a, unsigned b)
{
  while (b) {
    const unsigned t = a % b;
    a = b;
    b = t;
  }
  return a;
}

extern "C" {

FEMRT_API int femrt_mlp_num_params(int kind)
{
  return kind == FEMRT_MLP_GUESS || kind == FEMRT_MLP_BRACKET ? net_nparam(kind) : FEMRT_EUNSUP;
}

FEMRT_API int femrt_mlp_create(femrt_ctx* c, int kind, unsigned long long seed, femrt_mlp** out)
{
  if (!c || !out) return FEMRT_EINVAL;
  if (kind != FEMRT_MLP_GUESS && kind != FEMRT_MLP_BRACKET)
    return fail(c, FEMRT_EUNSUP, "femrt_mlp_create: unknown network kind");
  *out = nullptr;
  HIPCHK(c, hipSetDevice(c->device));
  femrt_mlp* m = new femrt_mlp;
  m->ctx = c;
  m->kind = kind;
  const int np = net_nparam(kind);
  if (hipMalloc(&m->d_param, np * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->d_m, np * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->d_v, np * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->d_grad, (np + 1) * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->d_pack, PK_BYTES) != hipSuccess) {
    mlp_free(m);
    delete m;
    return fail(c, FEMRT_ENOMEM, "femrt_mlp_create: device allocation failed");
  }
  std::vector<float> p;
  mlp_init_params(kind, seed, p);
  int rc = FEMRT_OK;
  if (hipMemcpyAsync(m->d_param, p.data(), np * sizeof(float), hipMemcpyHostToDevice,
                     c->stream) != hipSuccess)
    rc = fail(c, FEMRT_EHIP, "femrt_mlp_create: upload failed");
  if (rc == FEMRT_OK) rc = mlp_reset_state(m);
  if (rc == FEMRT_OK && hipStreamSynchronize(c->stream) != hipSuccess)
    rc = fail(c, FEMRT_EHIP, "femrt_mlp_create: stream synchronisation failed");
  if (rc != FEMRT_OK) {
    mlp_free(m);
    delete m;
    return rc;
  }
  *out = m;
  return FEMRT_OK;
}

FEMRT_API int femrt_mlp_destroy(femrt_mlp* m)
{
  if (!m) return FEMRT_OK;
  (void)hipStreamSynchronize(m->ctx->stream);
  mlp_free(m);
  delete m;
  return FEMRT_OK;
}

FEMRT_API int femrt_mlp_get_params(femrt_mlp* m, float* params)
{
  if (!m || !params) return FEMRT_EINVAL;
  femrt_ctx* c = m->ctx;
  HIPCHK(c, hipMemcpyAsync(params, m->d_param, net_nparam(m->kind) * sizeof(float), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FEMRT_OK;
}

FEMRT_API int femrt_mlp_set_params(femrt_mlp* m, const float* params)
{
  if (!m || !params) return FEMRT_EINVAL;
  femrt_ctx* c = m->ctx;
  HIPCHK(c, hipMemcpyAsync(m->d_param, params, net_nparam(m->kind) * sizeof(float), hipMemcpyHostToDevice,
                           c->stream));
  const int rc = mlp_reset_state(m);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FEMRT_OK;
}

FEMRT_API int femrt_mlp_forward(femrt_mlp* m, int n, const float* feats, float* out)
{
  if (!m || n < 0 || (n && (!feats || !out))) return FEMRT_EINVAL;
  if (!n) return FEMRT_OK;
  femrt_ctx* c = m->ctx;
  const int IN = net_in(m->kind), OUT = net_out(m->kind);
  float *d_in = nullptr, *d_out = nullptr;
  if (hipMalloc(&d_in, (size_t)n * IN * sizeof(float)) != hipSuccess ||
      hipMalloc(&d_out, (size_t)n * OUT * sizeof(float)) != hipSuccess) {
    if (d_in) (void)hipFree(d_in);
    return fail(c, FEMRT_ENOMEM, "femrt_mlp_forward: device allocation failed");
  }
  int rc = FEMRT_OK;
  if (hipMemcpyAsync(d_in, feats, (size_t)n * IN * sizeof(float), hipMemcpyHostToDevice,
                     c->stream) != hipSuccess)
    rc = fail(c, FEMRT_EHIP, "femrt_mlp_forward: upload failed");
  if (rc == FEMRT_OK)
    rc = timed_launch(c, "k_mlp_infer", [&] {
      FEMRT_NET_DISPATCH(m->kind, (k_mlp_infer<K_><<<(n + 255) / 256, 256, 0, c->stream>>>(n, d_in, m->d_pack, d_out)));
    });
  if (rc == FEMRT_OK && (hipMemcpyAsync(out, d_out, (size_t)n * OUT * sizeof(float),
                                        hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                         hipStreamSynchronize(c->stream) != hipSuccess))
    rc = fail(c, FEMRT_EHIP, "femrt_mlp_forward: download failed");
  (void)hipFree(d_in);
  (void)hipFree(d_out);
  return rc;
}

FEMRT_API int femrt_mlp_loss_grad(femrt_mlp* m, int n, const float* feats, const float* targets,
                                  float* loss, float* grad)
{
  if (!m || n <= 0 || !feats || !targets || !loss || !grad) return FEMRT_EINVAL;
  femrt_ctx* c = m->ctx;
  const int IN = net_in(m->kind), OUT = net_out(m->kind), NP = net_nparam(m->kind);
  float *d_in = nullptr, *d_t = nullptr;
  if (hipMalloc(&d_in, (size_t)n * IN * sizeof(float)) != hipSuccess ||
      hipMalloc(&d_t, (size_t)n * OUT * sizeof(float)) != hipSuccess) {
    if (d_in) (void)hipFree(d_in);
    return fail(c, FEMRT_ENOMEM, "femrt_mlp_loss_grad: device allocation failed");
  }
  int rc = FEMRT_OK;
  if (hipMemcpyAsync(d_in, feats, (size_t)n * IN * sizeof(float), hipMemcpyHostToDevice,
                     c->stream) != hipSuccess ||
      hipMemcpyAsync(d_t, targets, (size_t)n * OUT * sizeof(float), hipMemcpyHostToDevice,
                     c->stream) != hipSuccess)
    rc = fail(c, FEMRT_EHIP, "femrt_mlp_loss_grad: upload failed");
  const int blocks = (n + 63) / 64;
  float* part = rc == FEMRT_OK ? mlp_part(m, blocks) : nullptr;
  if (rc == FEMRT_OK && !part) rc = fail(c, FEMRT_ENOMEM, "femrt_mlp_loss_grad: device allocation failed");
  if (rc == FEMRT_OK)
    rc = timed_launch(c, "k_mlp_train", [&] {
      FEMRT_NET_DISPATCH(m->kind, (k_mlp_train<K_><<<blocks, 64, 0, c->stream>>>(
                                       d_in, d_t, n, 0, n, 1u, 0u, 1.f / n, m->d_pack, part)));
    });
  if (rc == FEMRT_OK)
    rc = timed_launch(c, "k_mlp_reduce", [&] {
      k_mlp_reduce<<<(NP + 1 + 255) / 256, 256, 0, c->stream>>>(part, blocks, NP, m->d_grad);
    });
  std::vector<float> g(NP + 1);
  if (rc == FEMRT_OK && (hipMemcpyAsync(g.data(), m->d_grad, (NP + 1) * sizeof(float),
                                        hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                         hipStreamSynchronize(c->stream) != hipSuccess))
    rc = fail(c, FEMRT_EHIP, "femrt_mlp_loss_grad: download failed");
  (void)hipFree(d_in);
  (void)hipFree(d_t);
  if (rc) return rc;
  std::memcpy(grad, g.data(), NP * sizeof(float));
  *loss = g[NP];
  return FEMRT_OK;
}

FEMRT_API int femrt_mlp_make_samples(femrt_mesh* mesh, int n, unsigned long long seed, float* feats,
                                     float* targets)
{
  if (!mesh || n <= 0 || !feats || !targets) return FEMRT_EINVAL;
  femrt_ctx* c = mesh->ctx;
  float *d_in = nullptr, *d_t = nullptr;
  if (hipMalloc(&d_in, (size_t)n * MLP_IN * sizeof(float)) != hipSuccess ||
      hipMalloc(&d_t, (size_t)n * 3 * sizeof(float)) != hipSuccess) {
    if (d_in) (void)hipFree(d_in);
    return fail(c, FEMRT_ENOMEM, "femrt_mlp_make_samples: device allocation failed");
  }
  int rc = make_samples_dev(mesh, n, seed, d_in, d_t);
  if (rc == FEMRT_OK &&
      (hipMemcpyAsync(feats, d_in, (size_t)n * MLP_IN * sizeof(float), hipMemcpyDeviceToHost,
                      c->stream) != hipSuccess ||
       hipMemcpyAsync(targets, d_t, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost,
                      c->stream) != hipSuccess ||
       hipStreamSynchronize(c->stream) != hipSuccess))
    rc = fail(c, FEMRT_EHIP, "femrt_mlp_make_samples: download failed");
  (void)hipFree(d_in);
  (void)hipFree(d_t);
  return rc;
}

FEMRT_API void femrt_mlp_train_params_default(femrt_mlp_train_params* p)
{
  if (!p) return;
  p->n_samples = 1 << 21;
  p->batch = 4096;
  p->epochs = 4;
  p->lr = 2e-3f;
  p->lr_final = 1e-4f;
  p->seed = 1;
}

}  // extern "C"

// Adam epochs over n training pairs already on the device (d_in n x IN,
// d_t n x OUT), in batches, each epoch in a fresh affine permutation of the
// pair order; the mean batch loss of each epoch goes to epoch_loss.
static int train_pairs_dev(femrt_mlp* m, int n, const float* d_in, const float* d_t,
                           const femrt_mlp_train_params* p, float* epoch_loss)
{
  femrt_ctx* c = m->ctx;
  const int NP = net_nparam(m->kind);
  float* d_loss = nullptr;
  if (hipMalloc(&d_loss, p->epochs * sizeof(float)) != hipSuccess)
    return fail(c, FEMRT_ENOMEM, "femrt_mlp_train: device allocation failed");
  int rc = FEMRT_OK;
  if (hipMemsetAsync(d_loss, 0, p->epochs * sizeof(float), c->stream) != hipSuccess)
    rc = fail(c, FEMRT_EHIP, "femrt_mlp_train: memset failed");
  const int steps = (n + p->batch - 1) / p->batch;
  const long long total = (long long)steps * p->epochs;
  unsigned long long st = p->seed * 0x9E3779B97F4A7C15ull + 12345u;
  long long step = 0;
  for (int ep = 0; ep < p->epochs && rc == FEMRT_OK; ++ep) {
    // a fresh affine permutation of the sample order each epoch
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    unsigned pa = (unsigned)(st >> 33) % (unsigned)n;
    if (pa == 0) pa = 1;
    while (gcd_u(pa, (unsigned)n) != 1) pa = pa + 1 < (unsigned)n ? pa + 1 : 1;
    const unsigned pb = (unsigned)(st >> 7) % (unsigned)n;
    for (int s = 0; s < steps && rc == FEMRT_OK; ++s, ++step) {
      const int base = s * p->batch;
      const int end = std::min(n, base + p->batch);
      const int cnt = end - base;
      const int blocks = (cnt + 63) / 64;
      float* part = mlp_part(m, blocks);
      if (!part) {
        rc = fail(c, FEMRT_ENOMEM, "femrt_mlp_train: device allocation failed");
        break;
      }
      rc = timed_launch(c, "k_mlp_train", [&] {
        FEMRT_NET_DISPATCH(m->kind, (k_mlp_train<K_><<<blocks, 64, 0, c->stream>>>(
                                         d_in, d_t, n, base, end, pa, pb, 1.f / cnt, m->d_pack, part)));
      });
      if (rc) break;
      // cosine decay from lr to lr_final
      const double frac = total > 1 ? (double)step / (double)(total - 1) : 1.0;
      const float lr = (float)(p->lr_final + 0.5 * (p->lr - p->lr_final) * (1.0 + std::cos(3.14159265358979323846 * frac)));
      ++m->adam_t;
      const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)m->adam_t)));
      const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)m->adam_t)));
      rc = timed_launch(c, "k_mlp_adam", [&] {
        k_mlp_adam<<<(NP + 1 + 255) / 256, 256, 0, c->stream>>>(
            m->d_param, m->d_m, m->d_v, part, blocks, NP, lr, c1, c2, d_loss + ep);
      });
      if (rc == FEMRT_OK) rc = mlp_pack(m);
    }
  }
  std::vector<float> hl(p->epochs);
  if (rc == FEMRT_OK && (hipMemcpyAsync(hl.data(), d_loss, p->epochs * sizeof(float),
                                        hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                         hipStreamSynchronize(c->stream) != hipSuccess))
    rc = fail(c, FEMRT_EHIP, "femrt_mlp_train: download failed");
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(d_loss);
  if (rc) return rc;
  if (epoch_loss)
    for (int e = 0; e < p->epochs; ++e) epoch_loss[e] = hl[e] / steps;
  return FEMRT_OK;
}

static bool train_params_ok(const femrt_mlp_train_params* p, int n)
{
  return p && n >= 64 && p->batch >= 64 && p->batch % 64 == 0 && p->epochs >= 1 && p->lr > 0.f &&
         p->lr_final > 0.f;
}

extern "C" {

FEMRT_API int femrt_mlp_train(femrt_mlp* m, femrt_mesh* mesh, const femrt_mlp_train_params* p,
                              float* epoch_loss)
{
  if (!m || !mesh || !p || !train_params_ok(p, p->n_samples))
    return m ? fail(m->ctx, FEMRT_EINVAL, "femrt_mlp_train: bad parameters") : FEMRT_EINVAL;
  if (m->kind != FEMRT_MLP_GUESS)
    return fail(m->ctx, FEMRT_EINVAL, "femrt_mlp_train: the mesh makes training pairs for the guess network only "
                                      "(train the bracket network with femrt_mlp_train_pairs)");
  if (mesh->ctx != m->ctx) return fail(m->ctx, FEMRT_EINVAL, "femrt_mlp_train: mesh and network on different contexts");
  femrt_ctx* c = m->ctx;
  const int n = p->n_samples;
  float *d_in = nullptr, *d_t = nullptr;
  if (hipMalloc(&d_in, (size_t)n * MLP_IN * sizeof(float)) != hipSuccess ||
      hipMalloc(&d_t, (size_t)n * 3 * sizeof(float)) != hipSuccess) {
    if (d_in) (void)hipFree(d_in);
    return fail(c, FEMRT_ENOMEM, "femrt_mlp_train: device allocation failed");
  }
  int rc = make_samples_dev(mesh, n, p->seed, d_in, d_t);
  if (rc == FEMRT_OK) rc = train_pairs_dev(m, n, d_in, d_t, p, epoch_loss);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(d_in);
  (void)hipFree(d_t);
  return rc;
}

FEMRT_API int femrt_mlp_train_pairs(femrt_mlp* m, int n, const float* feats, const float* targets,
                                    const femrt_mlp_train_params* p, float* epoch_loss)
{
  if (!m || !feats || !targets || !train_params_ok(p, n))
    return m ? fail(m->ctx, FEMRT_EINVAL, "femrt_mlp_train_pairs: bad parameters") : FEMRT_EINVAL;
  femrt_ctx* c = m->ctx;
  const int IN = net_in(m->kind), OUT = net_out(m->kind);
  float *d_in = nullptr, *d_t = nullptr;
  if (hipMalloc(&d_in, (size_t)n * IN * sizeof(float)) != hipSuccess ||
      hipMalloc(&d_t, (size_t)n * OUT * sizeof(float)) != hipSuccess) {
    if (d_in) (void)hipFree(d_in);
    return fail(c, FEMRT_ENOMEM, "femrt_mlp_train_pairs: device allocation failed");
  }
  int rc = FEMRT_OK;
  if (hipMemcpyAsync(d_in, feats, (size_t)n * IN * sizeof(float), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(d_t, targets, (size_t)n * OUT * sizeof(float), hipMemcpyHostToDevice, c->stream) != hipSuccess)
    rc = fail(c, FEMRT_EHIP, "femrt_mlp_train_pairs: upload failed");
  if (rc == FEMRT_OK) rc = train_pairs_dev(m, n, d_in, d_t, p, epoch_loss);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(d_in);
  (void)hipFree(d_t);
  return rc;
}

FEMRT_API int femrt_mesh_set_guess_mlp(femrt_mesh* mesh, const femrt_mlp* m)
{
  if (!mesh) return FEMRT_EINVAL;
  if (m && m->ctx != mesh->ctx) return fail(mesh->ctx, FEMRT_EINVAL, "femrt_mesh_set_guess_mlp: different contexts");
  if (m && m->kind != FEMRT_MLP_GUESS) return fail(mesh->ctx, FEMRT_EINVAL, "femrt_mesh_set_guess_mlp: not a guess network");
  mesh->guess_mlp = m;
  return FEMRT_OK;
}

FEMRT_API int femrt_mesh_set_bracket_mlp(femrt_mesh* mesh, const femrt_mlp* m)
{
  if (!mesh) return FEMRT_EINVAL;
  if (m && m->ctx != mesh->ctx) return fail(mesh->ctx, FEMRT_EINVAL, "femrt_mesh_set_bracket_mlp: different contexts");
  if (m && m->kind != FEMRT_MLP_BRACKET)
    return fail(mesh->ctx, FEMRT_EINVAL, "femrt_mesh_set_bracket_mlp: not a bracket network");
  mesh->bracket_mlp = m;
  return FEMRT_OK;
}

}  // extern "C"
