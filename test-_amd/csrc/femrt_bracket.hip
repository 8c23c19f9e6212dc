// femrt_bracket.hip -- the isosurface-bracket network between the walk and
// the FP64 endgame.
//
//   k_bracket_infer<NF>  one query per (hit ray, isovalue) bracket the walk
//                        recorded: the network's 5% and 95% quantiles of the
//                        crossing's position in the scan's bracket [ta, tb],
//                        as a predicted [t_lo, t_hi] for the learned
//                        bisection endgame (polish_iso_bisect);
//   k_bracket_pairs<NF>  training pairs from a finished trace: the same
//                        inputs, and the position of the FP64 root the
//                        endgame found, s = (t - ta) / (tb - ta).
// Inputs (32): rst where the ray enters and leaves the bracket (at ta and
// tb: 6), the secant weight g(ta) / (g(ta) - g(tb)) of the FP32 samples
// there (1), and 25 modal coefficients of u - iso -- mode indices (i, j, k)
// in {0, 1, 2}^3 other than (2, 2, 2) and (2, 2, 1) -- scaled by their
// largest magnitude.  The isovalue enters through the 8 vertex modes:
// 1 = (phi_0 + phi_1)(r) (phi_0 + phi_1)(s) (phi_0 + phi_1)(t) in the
// Modified_A basis (fem/ModalBasis.cuh:16-70), so subtracting iso subtracts
// it from each vertex-mode coefficient.
// The reference has no counterpart: its bracket is the first sign change of
// 16 uniform samples, refined by Newton / bisection
// (fem/RayIsosurface.cu:67-171).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "femrt_internal.hpp"
#include "femrt_mlp.hpp"
#include "femrt_runtime.hpp"

namespace {

constexpr int BR_IN = Net<NET_BRACKET>::IN;

// s0 = (rst(ta), rst(tb).x), s1 = (rst(tb).yz, g(ta), g(tb)) as the walk
// recorded them.
template <int NF>
__device__ void bracket_features(const float* __restrict__ fc, float iso, float4 s0, float4 s1, float* f)
{
  f[0] = s0.x;
  f[1] = s0.y;
  f[2] = s0.z;
  f[3] = s0.w;
  f[4] = s1.x;
  f[5] = s1.y;
  const float den = s1.z - s1.w;
  f[6] = den != 0.f ? fminf(fmaxf(s1.z / den, 0.f), 1.f) : 0.5f;
  float mx = 0.f;
  int q = 7;
  for (int k = 0; k < 3; ++k)
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) {
        if (i == 2 && j == 2 && k >= 1) continue;
        const float c = fc[(k * NF + j) * NF + i] - ((i < 2 && j < 2 && k < 2) ? iso : 0.f);
        f[q++] = c;
        mx = fmaxf(mx, fabsf(c));
      }
  const float sc = mx > 0.f ? 1.f / mx : 1.f;
  for (int p = 7; p < BR_IN; ++p) f[p] *= sc;
}

// The (hit ray, isovalue) slots are scanned 64 per wave and step (XCD-sliced
// as the polish is); the slots holding a bracket are compacted into a
// per-wave LDS list by ballot and prefix count, and the network runs once per
// 64 listed brackets rather than once per 64 slots.  Control flow is
// wave-uniform: the network takes operands from all 64 lanes, and lanes
// without a listed bracket feed zeros and write nothing.
template <int NF>
__global__ __launch_bounds__(256) void k_bracket_infer(int n, const float* __restrict__ field, IsoParams P,
                                                       const int* __restrict__ hit_queue,
                                                       const int* __restrict__ counts,
                                                       const int* __restrict__ iso_elem,
                                                       const float4* __restrict__ iso_f32,
                                                       const float4* __restrict__ iso_seg,
                                                       const void* __restrict__ pack, float2* __restrict__ pred)
{
  __shared__ bf16x8 ops[PK_INFER_BF16 / 8];
  __shared__ bf16x8 xs[4][64 * 4];
  __shared__ float bias[PK_NF32_ALL];
  __shared__ long long list[4][64];
  const bf16x8* g = (const bf16x8*)pack;
  for (int t = threadIdx.x; t < PK_INFER_BF16 / 8; t += blockDim.x) ops[t] = g[t];
  const float* gf = (const float*)((const __bf16*)pack + PK_NBF16);
  for (int t = threadIdx.x; t < PK_NF32_ALL; t += blockDim.x) bias[t] = gf[t];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int count = counts[0] * P.n_iso;
  XcdStride it(count);
  int nlist = 0;  // wave-uniform
  auto wave_sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  };
  for (;;) {
    const bool more = __any(it.q < it.end);
    long long sl = -1;
    if (it.q < it.end) {
      const int r = hit_queue[it.q / P.n_iso], k = it.q % P.n_iso;
      const long long s = (long long)k * n + r;
      if (iso_elem[s] >= 0) sl = s;
    }
    const unsigned long long mask = __ballot(sl >= 0);
    const int c = __popcll(mask);
    if (nlist > 0 && (nlist + c > 64 || !more)) {
      // the network on the listed brackets
      const long long s = lane < nlist ? list[wave][lane] : -1;
      float f[BR_IN];
      float ta = 0.f, tb = 0.f;
      if (s >= 0) {
        const int k = (int)(s / n), e = iso_elem[s];
        const float4 b = iso_f32[s];
        ta = b.x;
        tb = b.y;
        bracket_features<NF>(field + (size_t)e * (NF * NF * NF), P.iso[k], iso_seg[2 * (size_t)s],
                             iso_seg[2 * (size_t)s + 1], f);
      } else {
        for (int p = 0; p < BR_IN; ++p) f[p] = 0.f;
      }
      mlp_stage_inputs<NET_BRACKET>(xs[wave], lane, f);
      wave_sync();
      float y0[2], y1[2];
      wave_mlp_forward<NET_BRACKET>(xs[wave], ops, bias, bias + 64, y0, y1, bias + PK_B1);
      // lane j < 32 holds the outputs of queries j and j + 32
      const float a0 = __shfl(y0[0], lane & 31), a1 = __shfl(y0[1], lane & 31);
      const float b0 = __shfl(y1[0], lane & 31), b1 = __shfl(y1[1], lane & 31);
      const float s_lo = lane < 32 ? a0 : b0, s_hi = lane < 32 ? a1 : b1;
      if (s >= 0) {
        const float h = tb - ta;
        const float lo = fminf(fmaxf(fminf(s_lo, s_hi), 0.f), 1.f), hi = fminf(fmaxf(fmaxf(s_lo, s_hi), 0.f), 1.f);
        pred[s] = make_float2(ta + lo * h, ta + hi * h);
      }
      nlist = 0;
      wave_sync();  // the list and the input block are rewritten next
    }
    if (!more) break;
    if (sl >= 0) list[wave][nlist + __popcll(mask & ((1ull << lane) - 1ull))] = sl;
    nlist += c;
    wave_sync();
    it.q += it.step;
  }
}

template <int NF>
__global__ __launch_bounds__(256) void k_bracket_pairs(int n, int n_iso, const float* __restrict__ field,
                                                       const float* __restrict__ iso32,
                                                       const int* __restrict__ hit_queue,
                                                       const int* __restrict__ counts,
                                                       const int* __restrict__ iso_elem,
                                                       const double* __restrict__ iso_t,
                                                       const float4* __restrict__ iso_f32,
                                                       const float4* __restrict__ iso_seg, int max_out,
                                                       int* __restrict__ n_out, float* __restrict__ feats,
                                                       float* __restrict__ targets)
{
  const int count = counts[0] * n_iso;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < count; q += gridDim.x * blockDim.x) {
    const int r = hit_queue[q / n_iso], k = q % n_iso;
    const size_t sl = (size_t)k * n + r;
    const int e = iso_elem[sl];
    if (e < 0) continue;
    const float4 s0 = iso_seg[2 * sl], s1 = iso_seg[2 * sl + 1];
    const float4 b = iso_f32[sl];
    const double h = (double)b.y - (double)b.x;
    if (!(h > 0.0)) continue;
    const int o = atomicAdd(n_out, 1);
    if (o >= max_out) continue;
    float f[BR_IN];
    bracket_features<NF>(field + (size_t)e * (NF * NF * NF), iso32[k], s0, s1, f);
    for (int p = 0; p < BR_IN; ++p) feats[(size_t)o * BR_IN + p] = f[p];
    const float s = (float)((iso_t[sl] - (double)b.x) / h);
    targets[2 * (size_t)o] = s;
    targets[2 * (size_t)o + 1] = s;
  }
}

}  // namespace

int bracket_infer_launch(femrt_ctx* c, femrt_mesh* m, femrt_frame* f, const TraceMesh& TM, const IsoParams& ip,
                         int blocks)
{
  (void)TM;
  int rc = FEMRT_EUNSUP;
  FEMRT_DISPATCH_NF(m->NF, {
    rc = timed_launch(c, "k_bracket_infer", [&] {
      k_bracket_infer<NF_><<<blocks, 256, 0, c->stream>>>(f->n, m->d_field, ip, f->d_queue_hit, f->d_counts,
                                                           f->d_iso_elem, f->d_iso_f32, f->d_iso_seg,
                                                           m->bracket_mlp->d_pack,
                                                           f->d_iso_pred);
    });
  });
  return rc;
}

extern "C" {

FEMRT_API int femrt_frame_bracket_pairs(femrt_mesh* m, femrt_frame* f, int n_iso, const double* iso, int max_pairs,
                                        float* feats, float* targets, int* n_pairs)
{
  if (!m || !f || !iso || n_iso < 1 || n_iso > FEMRT_MAX_ISO || max_pairs < 0 || !n_pairs ||
      (max_pairs > 0 && (!feats || !targets)))
    return FEMRT_EINVAL;
  femrt_ctx* c = m->ctx;
  if (f->ctx != c) return fail(c, FEMRT_EINVAL, "bracket_pairs: mesh and frame belong to different contexts");
  if (!f->seg_valid)
    return fail(c, FEMRT_ESTATE, "bracket_pairs: the last trace recorded no segments (FEMRT_TRACE_SEG_INFO)");
  HIPCHK(c, hipSetDevice(c->device));
  float *d_f = nullptr, *d_t = nullptr, *d_iso = nullptr;
  int* d_n = nullptr;
  const size_t cap = (size_t)std::max(max_pairs, 1);
  auto cleanup = [&]() {
    (void)hipFree(d_f);
    (void)hipFree(d_t);
    (void)hipFree(d_iso);
    (void)hipFree(d_n);
  };
  float iso32[FEMRT_MAX_ISO];
  for (int k = 0; k < n_iso; ++k) iso32[k] = (float)iso[k];
  if (hipMalloc(&d_f, cap * BR_IN * 4) != hipSuccess || hipMalloc(&d_t, cap * 2 * 4) != hipSuccess ||
      hipMalloc(&d_iso, FEMRT_MAX_ISO * 4) != hipSuccess || hipMalloc(&d_n, 4) != hipSuccess) {
    cleanup();
    return fail(c, FEMRT_ENOMEM, "bracket_pairs: device allocation failed");
  }
  int rc = FEMRT_OK;
  if (hipMemcpyAsync(d_iso, iso32, n_iso * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemsetAsync(d_n, 0, 4, c->stream) != hipSuccess)
    rc = fail(c, FEMRT_EHIP, "bracket_pairs: upload failed");
  if (rc == FEMRT_OK) {
    rc = FEMRT_EUNSUP;
    FEMRT_DISPATCH_NF(m->NF, {
      rc = timed_launch(c, "k_bracket_pairs", [&] {
        k_bracket_pairs<NF_><<<4 * c->cus, 256, 0, c->stream>>>(f->n, n_iso, m->d_field, d_iso, f->d_queue_hit,
                                                                 f->d_counts, f->d_iso_elem, f->d_iso_t, f->d_iso_f32,
                                                                 f->d_iso_seg,
                                                                 max_pairs, d_n, d_f, d_t);
      });
    });
  }
  int total = 0;
  if (rc == FEMRT_OK && (hipMemcpyAsync(&total, d_n, 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                         hipStreamSynchronize(c->stream) != hipSuccess))
    rc = fail(c, FEMRT_EHIP, "bracket_pairs: download failed");
  const int got = std::min(total, max_pairs);
  if (rc == FEMRT_OK && got > 0 &&
      (hipMemcpy(feats, d_f, (size_t)got * BR_IN * 4, hipMemcpyDeviceToHost) != hipSuccess ||
       hipMemcpy(targets, d_t, (size_t)got * 2 * 4, hipMemcpyDeviceToHost) != hipSuccess))
    rc = fail(c, FEMRT_EHIP, "bracket_pairs: download failed");
  cleanup();
  if (rc == FEMRT_OK) *n_pairs = got;
  return rc;
}

}  // extern "C"
