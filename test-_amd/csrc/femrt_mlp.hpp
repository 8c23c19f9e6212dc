// femrt_mlp.hpp -- the library's two networks on gfx950 matrix cores.
//
// Guess network (FEMRT_MLP_GUESS): 27 inputs (hit point and the 8 corner
// vertices, each normalised to the element's bounding box as
// 2 (x - lo) / (hi - lo) - 1) plus a constant 1 for the first-layer bias
// -> 64 tanh -> 64 tanh -> 3 reference coordinates (squared-error loss).
// Bracket network (FEMRT_MLP_BRACKET): 32 inputs (the reference coordinates
// where the ray enters and leaves the element segment, and the 26 low-order
// modal coefficients of u - iso on the element, scaled to unit maximum)
// -> 64 tanh (explicit bias b1) -> 64 tanh -> 2: the 5% and 95% quantiles of
// the crossing's position s in [0, 1] along the segment (pinball loss).
// Operands are BF16, products accumulate in FP32 on v_mfma_f32_32x32x16_bf16.
//
// Inference works on one wave of 64 queries at a time in the transposed form
// H^T = W^T X^T: sample j of a 32-wide tile is the matrix column held by lane
// j % 32, so each lane keeps its own hidden activations in the accumulator
// registers and those registers are directly the B operand of the next layer
// (the hidden units of a k step are permuted to the accumulator order, and the
// weights are stored in that order).  Training uses row-major LDS tiles and
// generic strided operand loads.
#pragma once

#include <hip/hip_runtime.h>

namespace femrt {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int MLP_IN = 27;
constexpr int MLP_HID = 64;
constexpr int MLP_OUT = 3;
// flat FP32 parameter vector of the guess network
constexpr int MLP_OFF_W1 = 0;                        // [28][64] (input 27 = bias)
constexpr int MLP_OFF_W2 = 28 * 64;                  // [64][64] (in, out)
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + 64 * 64;     // [64]
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + 64;          // [64][3] (in, out)
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + 64 * 3;      // [3]
constexpr int MLP_NPARAM = MLP_OFF_B3 + 3;           // 6147

// Network kinds (the values of FEMRT_MLP_GUESS and FEMRT_MLP_BRACKET) and
// their shapes and flat parameter layouts.
constexpr int NET_GUESS = 0, NET_BRACKET = 1;
template <int KIND>
struct Net;
template <>
struct Net<NET_GUESS> {
  static constexpr int IN = MLP_IN, OUT = 3, W1_ROWS = 28;
  static constexpr bool BIAS_INPUT = true, PINBALL = false;
  static constexpr int OFF_W1 = MLP_OFF_W1, OFF_B1 = -1, OFF_W2 = MLP_OFF_W2, OFF_B2 = MLP_OFF_B2,
                       OFF_W3 = MLP_OFF_W3, OFF_B3 = MLP_OFF_B3, NPARAM = MLP_NPARAM;
};
template <>
struct Net<NET_BRACKET> {
  static constexpr int IN = 32, OUT = 2, W1_ROWS = 32;
  static constexpr bool BIAS_INPUT = false, PINBALL = true;
  static constexpr int OFF_W1 = 0, OFF_B1 = 32 * 64, OFF_W2 = OFF_B1 + 64, OFF_B2 = OFF_W2 + 64 * 64,
                       OFF_W3 = OFF_B2 + 64, OFF_B3 = OFF_W3 + 64 * 2, NPARAM = OFF_B3 + 2;  // 6402
};
// quantiles of the bracket network's two outputs
constexpr float BRACKET_TAU0 = 0.05f, BRACKET_TAU1 = 0.95f;

// Packed operand buffer: bf16 section, then FP32 biases.
constexpr int PK_W1OP = 0;       // [2 a][2 q][64 lane] x 8  inference A operands, layer 1
constexpr int PK_W2OP = 2048;    // [2 a][4 q][64 lane] x 8  layer 2
constexpr int PK_W3OP = 6144;    // [4 q][64 lane] x 8       layer 3 (outputs 0..2 of 32 rows)
constexpr int PK_W1B = 8192;     // [32][64] row-major (training), rows 28..31 zero
constexpr int PK_W2B = 10240;    // [64][64]
constexpr int PK_W3B = 14336;    // [64][32], columns 3..31 zero
constexpr int PK_NBF16 = 16384;
constexpr int PK_INFER_BF16 = 8192;  // inference operands only
constexpr int PK_NF32 = 68;          // b2[64], b3[<= 3], pad (what the guess network reads)
constexpr int PK_B1 = 68;            // b1[64] of the bracket network
constexpr int PK_NF32_ALL = PK_B1 + 64;
constexpr int PK_BYTES = PK_NBF16 * 2 + PK_NF32_ALL * 4;

// Hidden unit fed by k slot (q, h, x) of a layer-2/3 step: lane half h,
// element x of k step q (16 units per step) -- the accumulator order of
// v_mfma_f32_32x32x16: register v of lane (h, i) holds row
// 8 (v / 4) + 4 h + v % 4 of its 32-row tile.
__host__ __device__ inline int mlp_unit(int q, int h, int x)
{
  return 16 * q + 8 * (x >> 2) + 4 * h + (x & 3);
}

__device__ __forceinline__ float tanh_fast(float x)
{
  // tanh x = 1 - 2 / (1 + e^{2x}); e^{2x} = inf gives 1, e^{2x} = 0 gives -1
  const float e = __expf(2.f * x);
  return 1.f - 2.f * __builtin_amdgcn_rcpf(1.f + e);
}

__device__ __forceinline__ f32x16 mfma_bf16(bf16x8 a, bf16x8 b, f32x16 c)
{
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Writes this lane's 32 BF16 network inputs to row `lane` of the wave's LDS
// input block xs (64 rows x 4 bf16x8): the guess network's 27 features, 1,
// zeros; the bracket network's 32 features.
template <int KIND = NET_GUESS>
__device__ __forceinline__ void mlp_stage_inputs(bf16x8* xs, int lane, const float* feat)
{
  bf16x8 c0, c1, c2, c3;
  for (int x = 0; x < 8; ++x) {
    c0[x] = (__bf16)feat[x];
    c1[x] = (__bf16)feat[8 + x];
    c2[x] = (__bf16)feat[16 + x];
    if constexpr (KIND == NET_GUESS)
      c3[x] = (__bf16)(x < 3 ? feat[24 + x] : (x == 3 ? 1.f : 0.f));
    else
      c3[x] = (__bf16)feat[24 + x];
  }
  xs[lane * 4 + 0] = c0;
  xs[lane * 4 + 1] = c1;
  xs[lane * 4 + 2] = c2;
  xs[lane * 4 + 3] = c3;
}

// Forward pass of 64 queries (one per lane, inputs staged in xs).  Every lane
// of the wave must call it together: the matrix instructions take operands
// from all 64 lanes.  Lanes 0..31 receive the outputs of query `lane` in y0
// and of query `lane + 32` in y1 (lanes 32..63 receive unused rows).
// b1 (the bracket network's first-layer bias) is read when KIND has one.
template <int KIND = NET_GUESS>
__device__ __forceinline__ void wave_mlp_forward(const bf16x8* xs, const bf16x8* ops,
                                                 const float* b2, const float* b3, float* y0,
                                                 float* y1, const float* b1 = nullptr)
{
  constexpr int OUT = Net<KIND>::OUT;
  const int lane = __lane_id();
  const int i = lane & 31, h = lane >> 5;
  const bf16x8* w1 = ops + PK_W1OP / 8;
  const bf16x8* w2 = ops + PK_W2OP / 8;
  const bf16x8* w3 = ops + PK_W3OP / 8;
  for (int c = 0; c < 2; ++c) {
    bf16x8 hb[2][2];  // layer-1 activations as layer-2 B operands [a][q % 2]
    {
      const bf16x8 x0 = xs[(32 * c + i) * 4 + h];
      const bf16x8 x1 = xs[(32 * c + i) * 4 + 2 + h];
      for (int a = 0; a < 2; ++a) {
        f32x16 acc = {};
        if constexpr (!Net<KIND>::BIAS_INPUT)
          for (int v = 0; v < 16; ++v) acc[v] = b1[32 * a + 8 * (v >> 2) + 4 * h + (v & 3)];
        acc = mfma_bf16(w1[(a * 2 + 0) * 64 + lane], x0, acc);
        acc = mfma_bf16(w1[(a * 2 + 1) * 64 + lane], x1, acc);
        for (int x = 0; x < 8; ++x) {
          hb[a][0][x] = (__bf16)tanh_fast(acc[x]);
          hb[a][1][x] = (__bf16)tanh_fast(acc[8 + x]);
        }
      }
    }
    bf16x8 gb[2][2];
    for (int a = 0; a < 2; ++a) {
      f32x16 acc;
      for (int v = 0; v < 16; ++v) acc[v] = b2[32 * a + 8 * (v >> 2) + 4 * h + (v & 3)];
      for (int q = 0; q < 4; ++q) acc = mfma_bf16(w2[(a * 4 + q) * 64 + lane], hb[q >> 1][q & 1], acc);
      for (int x = 0; x < 8; ++x) {
        gb[a][0][x] = (__bf16)tanh_fast(acc[x]);
        gb[a][1][x] = (__bf16)tanh_fast(acc[8 + x]);
      }
    }
    f32x16 acc;
    for (int v = 0; v < 16; ++v) {
      const int u = 8 * (v >> 2) + 4 * h + (v & 3);
      acc[v] = u < OUT ? b3[u] : 0.f;
    }
    for (int q = 0; q < 4; ++q) acc = mfma_bf16(w3[q * 64 + lane], gb[q >> 1][q & 1], acc);
    float* y = c ? y1 : y0;
    for (int o = 0; o < OUT; ++o) y[o] = acc[o];
  }
}

// Network inputs of a point x in an element with bounding box bx (lo.xyz,
// hi.xyz) and corner vertices vt (8 x 3).
__device__ __forceinline__ void mlp_features(const float* bx, const float* vt, const float* x,
                                             float* feat)
{
  float s[3], o[3];
  for (int c = 0; c < 3; ++c) {
    s[c] = 2.f / fmaxf(bx[3 + c] - bx[c], 1e-30f);
    o[c] = bx[c];
  }
  for (int c = 0; c < 3; ++c) feat[c] = (x[c] - o[c]) * s[c] - 1.f;
  for (int k = 0; k < 8; ++k)
    for (int c = 0; c < 3; ++c) feat[3 + 3 * k + c] = (vt[3 * k + c] - o[c]) * s[c] - 1.f;
}

// One 32 x 32 output tile of a product of row-major LDS (or global) BF16
// matrices with general strides: D(i, j) = sum_k A(i, k) B(k, j), where
// A(i, k) = A[i * a_si + k * a_sk], B(k, j) = B[k * b_sk + j * b_sj], and
// K a multiple of 16.  Register v of lane (h, i) holds D(8 (v / 4) + 4 h +
// v % 4, i).
__device__ __forceinline__ f32x16 tile_gemm(const __bf16* A, int a_si, int a_sk, const __bf16* B,
                                            int b_sk, int b_sj, int K, f32x16 acc)
{
  const int lane = __lane_id();
  const int i = lane & 31, h = lane >> 5;
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x8 a, b;
    for (int x = 0; x < 8; ++x) {
      const int k = k0 + 8 * h + x;
      a[x] = A[i * a_si + k * a_sk];
      b[x] = B[k * b_sk + i * b_sj];
    }
    acc = mfma_bf16(a, b, acc);
  }
  return acc;
}

__device__ __forceinline__ int tile_row(int v, int lane) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }

}  // namespace femrt
