// femrt_march.hpp -- the persistent element walk (k_march) and its helpers.
//
// The kernel is instantiated per geometry mode count in femrt_march_ng*.hip
// (one translation unit each, so they compile side by side); the trace entry
// point launches it through march_launch<NG>.
// Reference: fem/RayIsosurface.cu:67-188 (iso search along the ray),
// raytracer.cu:102-254 (traversal + hex branch).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "femrt_mlp.hpp"
#include "femrt_runtime.hpp"
#include "femrt_trace_core.hpp"

// Bracket output of the walk: FP32 bracket and estimate per (isovalue, ray),
// and, when seg is set, the bracket's ends for the bracket network (two
// float4 per slot: rst at ta, rst at tb, g at ta, g at tb).
struct GlobalSink {
  int* elem;
  float4* f32;
  float4* rst32;
  size_t n;
  int r;
  float4* seg;
  __device__ void put(int k, const IsoBracket& b)
  {
    const size_t slot = (size_t)k * n + r;
    elem[slot] = b.elem;
    f32[slot] = make_float4(b.ta, b.tb, b.tm, b.ga);
    rst32[slot] = make_float4(b.rst[0], b.rst[1], b.rst[2], 0.f);
    if (seg) {
      seg[2 * slot] = make_float4(b.seg[0], b.seg[1], b.seg[2], b.seg[3]);
      seg[2 * slot + 1] = make_float4(b.seg[4], b.seg[5], b.seg[6], b.seg[7]);
    }
  }
};

// Network guesses of the interior samples of the lane's current segment.
struct LdsGuess {
  const _Float16* g;  // (samples - 1) x 3
  int lead;           // samples 1..lead carry a guess
  __device__ bool get(int j, float* rst) const
  {
    if (j > lead) return false;
    rst[0] = (float)g[(j - 1) * 3 + 0];
    rst[1] = (float)g[(j - 1) * 3 + 1];
    rst[2] = (float)g[(j - 1) * 3 + 2];
    return true;
  }
};

// Orders this wave's LDS writes before its following LDS reads (lanes read
// what other lanes of the same wave wrote).
__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// LDS of one wave for the cooperative scan: the SegInfo of each lane's
// segment (word-major), the sample values of each lane's segment, and the
// lanes with a segment to scan in this round.
#define FEMRT_SEG_WORDS 23
struct CoopWave {
  float seg[FEMRT_SEG_WORDS][64];
  float u[FEMRT_COOP_MAX_SAMPLES + 1][64];
  int list[64];
};

__device__ __forceinline__ void seg_store(CoopWave& L, int lane, const SegInfo& g)
{
  float w[FEMRT_SEG_WORDS] = {__int_as_float(g.e), __uint_as_float(g.open), g.lip, g.t_in, g.t_out,
                              g.r_in[0], g.r_in[1], g.r_in[2], g.r_out[0], g.r_out[1], g.r_out[2],
                              g.m_in[0], g.m_in[1], g.m_in[2], g.m_out[0], g.m_out[1], g.m_out[2],
                              g.o[0], g.o[1], g.o[2], g.d[0], g.d[1], g.d[2]};
#pragma unroll
  for (int k = 0; k < FEMRT_SEG_WORDS; ++k) L.seg[k][lane] = w[k];
}

__device__ __forceinline__ void seg_load(const CoopWave& L, int lane, SegInfo& g)
{
  g.e = __float_as_int(L.seg[0][lane]);
  g.open = __float_as_uint(L.seg[1][lane]);
  g.lip = L.seg[2][lane];
  g.t_in = L.seg[3][lane];
  g.t_out = L.seg[4][lane];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    g.r_in[c] = L.seg[5 + c][lane];
    g.r_out[c] = L.seg[8 + c][lane];
    g.m_in[c] = L.seg[11 + c][lane];
    g.m_out[c] = L.seg[14 + c][lane];
    g.o[c] = L.seg[17 + c][lane];
    g.d[c] = L.seg[20 + c][lane];
  }
}

// The scan step of one walk round with the samples spread over the wave:
// the lanes whose segment is open (an exit found, an isovalue still within
// the element's bounds of u) are taken four at a time, 16 lanes per segment
// -- lanes 0..14 of a group invert and evaluate interior samples 1..15, lane
// 15 evaluates the two ends -- with one Newton iteration per trip of the
// loop, so a lane that finishes early waits at most for its wave's slowest
// sample of the pass.
// The sample values go to LDS; each lane then scans its own segment.
// Returns (on the owner lanes) whether the segment must be redone.  Every
// lane of the wave must call it.
template <int NG, int NF, typename W, typename Sink>
__device__ bool coop_scan_round(const TraceMesh& M, const IsoParams& P, const float* o, const float* d,
                                MarchState& S, bool active, CoopWave& L, Sink& sink, W& w)
{
  const int lane = threadIdx.x & 63;
  const int Ms = P.samples;
  bool scan = false;
  if (active && S.ok) {
    scan = seg_may_cross(M, P, S);
    if (!scan) ++w.skipped;
  }
  if (scan) {
    SegInfo g;
    seg_info<NG>(M, P, o, d, S, g, w);
    seg_store(L, lane, g);
  }
  const unsigned long long mask = __ballot(scan);
  const int nscan = __popcll(mask);
  if (scan) L.list[__popcll(mask & ((1ull << lane) - 1ull))] = lane;
  wave_lds_sync();
  const int grp = lane >> 4, jj = lane & 15;
  for (int base = 0; base < nscan; base += 4) {
    const int slot = base + grp;
    const int owner = slot < nscan ? L.list[slot] : -1;
    SegInfo G;
    if (owner >= 0) seg_load(L, owner, G);
    else G.e = 0;
    const GeoMode* gc = M.geom + (size_t)G.e * (NG * NG * NG);
    const float* fc = M.field + (size_t)G.e * (NF * NF * NF);
    const int j = jj + 1;
    bool live = owner >= 0 && jj < 15 && j < Ms;
    int ep = (owner >= 0 && jj == 15) ? 0 : 2;  // the end samples still to evaluate
    SampleJob J;
    if (live) sample_job_begin(J, G, j, Ms, P.guess);
    while (__any(live || ep < 2)) {
      SampleTrip T;
      bool need = false;
      float rr[3];
      if (live) {
        sample_job_newton<NG>(J, gc, G.lip, T, w.inv_it);
        need = sample_trip_needs_u(T);
        rr[0] = J.rst[0];
        rr[1] = J.rst[1];
        rr[2] = J.rst[2];
      } else if (ep < 2) {
        need = true;
        for (int c = 0; c < 3; ++c) rr[c] = ep == 0 ? G.r_in[c] : G.r_out[c];
      }
      float u = 0.f;
      if (need) {
        u = field_value<float, NF>(fc, rr[0], rr[1], rr[2]);
        ++w.field;
      }
      if (live) {
        const int rc = sample_job_after(J, T, u, P, G.open, G.lip);
        if (rc != SJ_CONTINUE) {
          L.u[j][owner] = rc == SJ_DONE ? u : NAN;
          live = false;
        }
      } else if (ep < 2) {
        L.u[ep == 0 ? 0 : Ms][owner] = u;
        ++ep;
      }
    }
  }
  wave_lds_sync();
  bool redo = false;
  if (scan) {
    SegInfo g;
    seg_load(L, lane, g);
    redo = seg_scan(P, g, S, &L.u[0][lane], 64, sink, w);
  }
  wave_lds_sync();  // the next round rewrites the LDS
  return redo;
}

// Persistent element walk.  Each round every lane of a wave advances its ray
// by one element segment (exit, samples, next element); a lane whose ray
// ended takes the next hit ray from the queue (one atomic per wave), so the
// lanes stay busy until the queue drains.  With LEARNED, the point-inversion
// guesses of the first P.mlp_lead interior samples of the round's segments
// come from the network, evaluated on the matrix cores for the whole wave
// between the exit and sample steps (later samples extrapolate).
// COUNT selects the work counters (validation runs) or none (timed runs).
template <int NG, int NF, bool LEARNED, bool COUNT, bool COOP>
__global__ __launch_bounds__(256, 2) void k_march(
    const float* __restrict__ rays, int n, TraceMesh M, IsoParams P, const int* __restrict__ queue,
    int* __restrict__ counts, const int* __restrict__ hit_elem, const int* __restrict__ hit_face,
    const float4* __restrict__ hit_f32, int* __restrict__ iso_elem, float4* __restrict__ iso_f32,
    float4* __restrict__ iso_rst32, long long* __restrict__ work,
    const void* __restrict__ mlp_pack, _Float16* __restrict__ mlp_g, float4* __restrict__ iso_seg)
{
  const int count = counts[0];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  typename std::conditional<COUNT, Work, NoWork>::type w;
  GlobalSink sink{iso_elem, iso_f32, iso_rst32, (size_t)n, 0, iso_seg};
  bf16x8* ops = nullptr;
  float* bias = nullptr;
  bf16x8* xs = nullptr;
  _Float16* G = nullptr;  // this lane's guesses, gs halfs per lane
  int gs = 3;
  CoopWave* coop = nullptr;
  if constexpr (COOP) {
    __shared__ CoopWave s_coop[4];
    coop = &s_coop[wave];
  }
  static_assert(!(LEARNED && COOP), "the cooperative scan takes no network guesses");
  if constexpr (LEARNED) {
    __shared__ bf16x8 s_ops[PK_INFER_BF16 / 8];
    __shared__ float s_bias[PK_NF32];
    __shared__ bf16x8 s_xs[4][64 * 4];
    __shared__ _Float16 s_g[4][64 * 3];
    const bf16x8* gp = (const bf16x8*)mlp_pack;
    for (int t = threadIdx.x; t < PK_INFER_BF16 / 8; t += blockDim.x) s_ops[t] = gp[t];
    const float* gf = (const float*)((const __bf16*)mlp_pack + PK_NBF16);
    for (int t = threadIdx.x; t < PK_NF32; t += blockDim.x) s_bias[t] = gf[t];
    __syncthreads();
    ops = s_ops;
    bias = s_bias;
    xs = s_xs[wave];
    if (P.mlp_lead > 1) {
      // every interior sample guessed: per-lane rows of the global buffer
      // (the LDS holds one guess per lane, which keeps three blocks per CU)
      gs = (FEMRT_MLP_MAX_SAMPLES - 1) * 3;
      G = mlp_g + ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * gs;
    } else {
      G = &s_g[wave][lane * 3];
    }
  }
  bool active = false;
  float o[3] = {0.f, 0.f, 0.f}, d[3] = {1.f, 1.f, 1.f}, id[3] = {1.f, 1.f, 1.f};
  MarchState S;
  march_begin(S, HitF32());
  // The hit queue is split into FEMRT_XCDS x q contiguous segments (tile
  // order, so each covers one band of the image) with a counter each: XCD x
  // owns segments x q .. x q + q - 1.  A block starts on a segment of its XCD
  // (workgroups are dispatched round-robin over the XCDs, so block b runs on
  // XCD b % FEMRT_XCDS), its XCD's blocks spread over the q sub-bands, so
  // that each XCD's L2 holds the elements of one band; a wave that finds its
  // segment drained moves on to the next one.
  const int q_x = P.qseg_per_xcd, nseg = FEMRT_XCDS * q_x;
  int seg = (blockIdx.x % FEMRT_XCDS) * q_x + (blockIdx.x / FEMRT_XCDS) % q_x, drained = 0;  // wave-uniform
  // timeline with the work counters: start (min), queue dry (min), end (max)
  const bool timeline = work != nullptr && lane == 0;
  if (timeline) atomicMin((unsigned long long*)&work[16], (unsigned long long)wall_clock64());
  bool saw_dry = false;
  for (;;) {
    bool need = !active;
    while (drained < nseg && __any(need)) {
      const int lo = (int)((long long)count * seg / nseg);
      const int hi = (int)((long long)count * (seg + 1) / nseg);
      const int q = lo + wave_append(need, &counts[FEMRT_QSEG_CTR + seg]);
      const bool got = need && q < hi;
      bool took = false;
      // (a hit that k_retrace found false and could not replace is skipped)
      if (got && hit_elem[queue[q]] >= 0) {
        const int r = queue[q];
        took = true;
        sink.r = r;
        load_ray(rays, n, r, o, d);
        for (int c = 0; c < 3; ++c) id[c] = 1.f / d[c];
        HitF32 h;
        h.elem = hit_elem[r];
        h.face = hit_face[r];
        const float4 hf = hit_f32[r];
        h.u = hf.x;
        h.v = hf.y;
        h.t = hf.z;
        march_begin(S, h);
        active = true;
      }
      if (__any(need && !got)) {
        seg = seg + 1 == nseg ? 0 : seg + 1;
        ++drained;
      }
      need = need && !took;
    }
    if (timeline && drained == nseg && !saw_dry) {
      saw_dry = true;
      atomicMin((unsigned long long*)&work[17], (unsigned long long)wall_clock64());
    }
    if (!__any(active)) break;
    if (active) march_exit<NG>(M, P, o, d, id, S, w);
    bool coop_redo = false;
    if constexpr (COOP) coop_redo = coop_scan_round<NG, NF>(M, P, o, d, S, active, *coop, sink, w);
    if constexpr (LEARNED && !COOP) {
      // every lane of the wave is here: the network runs on the whole wave
      // (a segment whose first sample starts from the Hermite interpolant
      // needs no guess there; guesses of later samples still do)
      const bool seg = active && S.ok && (P.mlp_lead > 1 || !(S.have_m && P.hermite));
      const float* bx = M.boxes + (size_t)(seg ? S.e : 0) * FEMRT_BOX_STRIDE;
      const float* vt = M.verts + (size_t)(seg ? S.e : 0) * 24;
      // (wave-uniform: the forward pass is a wave-wide matrix product)
      const int jmax = __any(seg) ? min(P.mlp_lead, P.samples - 1) : 0;
      for (int j = 1; j <= jmax; ++j) {
        float f[MLP_IN];
        if (seg) {
          const float t = march_sample_t(S, j, P.samples);
          const float x[3] = {fmaf(t, d[0], o[0]), fmaf(t, d[1], o[1]), fmaf(t, d[2], o[2])};
          mlp_features(bx, vt, x, f);
          ++w.mlp;
        } else {
          for (int k = 0; k < MLP_IN; ++k) f[k] = 0.f;
        }
        mlp_stage_inputs(xs, lane, f);
        wave_lds_sync();
        float y0[3], y1[3];
        wave_mlp_forward(xs, ops, bias, bias + 64, y0, y1);
        if (lane < 32) {
          _Float16* g0 = &G[(j - 1) * 3];
          _Float16* g1 = g0 + 32 * gs;
          for (int c = 0; c < 3; ++c) {
            g0[c] = (_Float16)y0[c];
            g1[c] = (_Float16)y1[c];
          }
        }
        wave_lds_sync();
      }
    }
    if (active) {
      bool redo;
      if constexpr (COOP)
        redo = coop_redo;
      else if constexpr (LEARNED)
        redo = S.ok && march_samples<NG, NF>(M, P, o, d, S, LdsGuess{G, P.mlp_lead}, sink, w);
      else
        redo = S.ok && march_samples<NG, NF>(M, P, o, d, S, NoGuess(), sink, w);
      if (redo) {
        S.redo = true;
      } else if (!march_advance<NG>(M, P, o, d, S, w)) {
        active = false;
        if (work) {
          const int len = max(1, S.step & 0xFFFF);
          add_work(work, 19 + min(7, 31 - __clz(len)), 1);
          if (march_skips(S) > 0 || march_stalled(S)) {
            // a walk that stalled: keep the ray id for inspection
            const unsigned long long k = atomicAdd((unsigned long long*)&work[FEMRT_NCOUNTERS], 1ull);
            if (k < FEMRT_MAX_STALLED_IDS) work[FEMRT_NCOUNTERS + 1 + k] = sink.r;
          }
        }
        // a bracket written by a pass that was redone is not a result
        for (int k = 0; k < P.n_iso; ++k)
          if (!((S.found >> k) & 1u)) iso_elem[(size_t)k * n + sink.r] = -1;
      }
    }
  }
  if (timeline) atomicMax((unsigned long long*)&work[18], (unsigned long long)wall_clock64());
  if constexpr (COUNT) flush_work(work, w);
}


// Arguments of one launch of the walk.
struct MarchArgs {
  int blocks = 0;
  hipStream_t stream = nullptr;
  int nf = 0;
  bool learned = false, count = false, coop = false;
  const float* rays = nullptr;
  int n = 0;
  TraceMesh M;
  IsoParams P;
  const int* queue = nullptr;
  int* counts = nullptr;
  const int* hit_elem = nullptr;
  const int* hit_face = nullptr;
  const float4* hit_f32 = nullptr;
  int* iso_elem = nullptr;
  float4* iso_f32 = nullptr;
  float4* iso_rst32 = nullptr;
  long long* work = nullptr;
  const void* mlp_pack = nullptr;
  _Float16* mlp_g = nullptr;
  float4* iso_seg = nullptr;
};

template <int NG, int NF>
void march_launch_nf(const MarchArgs& a)
{
  auto go = [&](auto L, auto CNT, auto CO) {
    k_march<NG, NF, decltype(L)::value, decltype(CNT)::value, decltype(CO)::value><<<a.blocks, 256, 0, a.stream>>>(
        a.rays, a.n, a.M, a.P, a.queue, a.counts, a.hit_elem, a.hit_face, a.hit_f32, a.iso_elem, a.iso_f32,
        a.iso_rst32, a.work, a.mlp_pack, a.mlp_g, a.iso_seg);
  };
  using T1 = std::true_type;
  using F0 = std::false_type;
  if (a.coop && a.count) go(F0{}, T1{}, T1{});
  else if (a.coop) go(F0{}, F0{}, T1{});
  else if (a.learned && a.count) go(T1{}, T1{}, F0{});
  else if (a.learned) go(T1{}, F0{}, F0{});
  else if (a.count) go(F0{}, T1{}, F0{});
  else go(F0{}, F0{}, F0{});
}

// Launch of k_march<NG, a.nf, ...> (a.nf one of the instantiated field mode
// counts).  Defined in femrt_march_ng<NG>.hip.
template <int NG>
void march_launch(const MarchArgs& a);
template <>
void march_launch<4>(const MarchArgs& a);
template <>
void march_launch<5>(const MarchArgs& a);
template <>
void march_launch<7>(const MarchArgs& a);

#define FEMRT_MARCH_LAUNCH_DEF(NGV)                    \
  template <>                                          \
  void march_launch<NGV>(const MarchArgs& a)           \
  {                                                    \
    switch (a.nf) {                                    \
      case 3: march_launch_nf<NGV, 3>(a); break;       \
      case 5: march_launch_nf<NGV, 5>(a); break;       \
      case 7: march_launch_nf<NGV, 7>(a); break;       \
      case 9: march_launch_nf<NGV, 9>(a); break;       \
      default: break;                                  \
    }                                                  \
  }
