// femrt_walk.hip -- the element walk of the isosurface search in rounds,
// with the active rays binned by their current element.
//
// Every round advances each active ray by one element segment (exit,
// samples, next element).  Before the round the rays are sorted by element
// (histogram, prefix sum, scatter), so the lanes of a wave mostly work in the
// same element: their coefficient loads hit the same cache lines and their
// Newton solves take similar iteration counts.  Rays that leave the mesh find
// their next element through the BVH in a separate kernel, so the traversal
// does not stall the waves of the walk.  With learned guesses the guess
// network runs on the matrix cores for the whole wave between the exit and
// sample steps.
//
//   k_walk_init     walk state of every hit ray, round list = hit queue
//   k_walk_hist     rays per element
//   k_walk_scan     exclusive prefix sum of the histogram
//   k_walk_sort     ray ids in element order
//   k_walk_segment  one segment per ray; continuing rays to the next list,
//                   rays that left the mesh to the re-entry list
//   k_walk_reenter  BVH re-entry; rays that meet another element continue
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "femrt_mlp.hpp"
#include "femrt_runtime.hpp"

using namespace femrt;

namespace {

struct GlobalSink {
  int* elem;
  float4* f32;
  float4* rst32;
  size_t n;
  int r;
  __device__ void put(int k, const IsoBracket& b)
  {
    const size_t slot = (size_t)k * n + r;
    elem[slot] = b.elem;
    f32[slot] = make_float4(b.ta, b.tb, b.tm, b.ga);
    rst32[slot] = make_float4(b.rst[0], b.rst[1], b.rst[2], 0.f);
  }
};

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

__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ void load_state(const int4* si, const float4* sf, int r, MarchState& S)
{
  const int4 a = si[r];
  const float4 b = sf[r];
  S.e = a.x;
  S.f_in = a.y;
  S.step = a.z;
  S.found = (unsigned)a.w & 0xFFu;
  S.redo = ((unsigned)a.w >> 8) & 1u;
  S.t_in = b.x;
  S.u_in = b.y;
  S.v_in = b.z;
  S.ok = false;
  S.f_out = -1;
  S.t_out = S.u_out = S.v_out = 0.f;
}

__device__ __forceinline__ void store_state(int4* si, float4* sf, int r, const MarchState& S)
{
  si[r] = make_int4(S.e, S.f_in, S.step, (int)(S.found | (S.redo ? 256u : 0u)));
  sf[r] = make_float4(S.t_in, S.u_in, S.v_in, 0.f);
}

// Finished ray: brackets of isovalues not found cleared (a pass that was
// redone may have written one); the endgame takes the others.
__device__ __forceinline__ void finish_ray(bool fin, int r, unsigned found, int n_iso, int n, int* iso_elem)
{
  for (int k = 0; k < n_iso; ++k)
    if (fin && !((found >> k) & 1u)) iso_elem[(size_t)k * n + r] = -1;
}

}  // namespace

__global__ void k_walk_init(int* __restrict__ counts, const int* __restrict__ queue_hit,
                            const int* __restrict__ hit_elem, const int* __restrict__ hit_face,
                            const float4* __restrict__ hit_f32, int4* __restrict__ si,
                            float4* __restrict__ sf, int* __restrict__ list)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= counts[0]) return;
  const int r = queue_hit[q];
  if (hit_elem[r] < 0) return;  // a false hit that the re-trace could not replace
  const float4 h = hit_f32[r];
  si[r] = make_int4(hit_elem[r], hit_face[r], 0, 0);
  sf[r] = make_float4(h.z, h.x, h.y, 0.f);
  list[atomicAdd(&counts[5], 1)] = r;
}

__global__ void k_walk_zero(int* __restrict__ a, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

__global__ void k_walk_hist(const int* __restrict__ list, int n, const int4* __restrict__ si,
                            int* __restrict__ hist)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&hist[si[list[i]].x], 1);
}

// In-place exclusive prefix sum of n counts, one block of 1024 threads:
// chunk sums, a block scan of the chunk sums, then the chunks.
__global__ __launch_bounds__(1024) void k_walk_scan(int* __restrict__ a, int n)
{
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int chunk = (n + 1023) / 1024;
  const int lo = min(n, t * chunk), hi = min(n, lo + chunk);
  int s = 0;
  for (int i = lo; i < hi; ++i) s += a[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - s;  // exclusive offset of this chunk
  for (int i = lo; i < hi; ++i) {
    const int c = a[i];
    a[i] = run;
    run += c;
  }
}

__global__ void k_walk_sort(const int* __restrict__ list, int n, const int4* __restrict__ si,
                            int* __restrict__ cursor, int* __restrict__ sorted)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = list[i];
  sorted[atomicAdd(&cursor[si[r].x], 1)] = r;
}

// counts[2]: rays continuing to the next round (list next); counts[3]: rays
// to re-enter (list reent).  Every lane reaches the wave-level calls.
template <int NG, int NF, bool LEARNED>
__global__ __launch_bounds__(256, 2) void k_walk_segment(
    const float* __restrict__ rays, int n, TraceMesh M, IsoParams P, const int* __restrict__ sorted,
    int n_active, int4* __restrict__ si, float4* __restrict__ sf, int* __restrict__ next,
    int* __restrict__ reent, int* __restrict__ counts, int* __restrict__ iso_elem,
    float4* __restrict__ iso_f32, float4* __restrict__ iso_rst32, long long* __restrict__ work,
    const void* __restrict__ mlp_pack)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < n_active;
  Work w;
  int r = 0;
  float o[3] = {0.f, 0.f, 0.f}, d[3] = {1.f, 1.f, 1.f}, id[3] = {1.f, 1.f, 1.f};
  MarchState S;
  march_begin(S, HitF32());
  if (active) {
    r = sorted[i];
    load_ray(rays, n, r, o, d);
    for (int c = 0; c < 3; ++c) id[c] = 1.f / d[c];
    load_state(si, sf, r, S);
    march_exit<NG>(M, P, o, d, id, S, w);
  }
  GlobalSink sink{iso_elem, iso_f32, iso_rst32, (size_t)n, r};
  bool redo = false;
  if constexpr (LEARNED) {
    __shared__ bf16x8 s_ops[PK_INFER_BF16 / 8];
    __shared__ float s_bias[PK_NF32];
    __shared__ bf16x8 s_xs[4][64 * 4];
    __shared__ _Float16 s_g[4][64][(FEMRT_MLP_MAX_SAMPLES - 1) * 3];
    const bf16x8* gp = (const bf16x8*)mlp_pack;
    for (int t = threadIdx.x; t < PK_INFER_BF16 / 8; t += blockDim.x) s_ops[t] = gp[t];
    const float* gf = (const float*)((const __bf16*)mlp_pack + PK_NBF16);
    for (int t = threadIdx.x; t < PK_NF32; t += blockDim.x) s_bias[t] = gf[t];
    __syncthreads();
    _Float16* G = &s_g[wave][lane][0];
    const bool seg = active && S.ok;
    const float* bx = M.boxes + (size_t)(seg ? S.e : 0) * FEMRT_BOX_STRIDE;
    const float* vt = M.verts + (size_t)(seg ? S.e : 0) * 24;
    const int jmax = min(P.mlp_lead, P.samples - 1);
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
      mlp_stage_inputs(s_xs[wave], lane, f);
      wave_lds_sync();
      float y0[3], y1[3];
      wave_mlp_forward(s_xs[wave], s_ops, s_bias, s_bias + 64, y0, y1);
      if (lane < 32) {
        _Float16* g0 = &G[(j - 1) * 3];
        _Float16* g1 = g0 + 32 * (FEMRT_MLP_MAX_SAMPLES - 1) * 3;
        for (int c = 0; c < 3; ++c) {
          g0[c] = (_Float16)y0[c];
          g1[c] = (_Float16)y1[c];
        }
      }
      wave_lds_sync();
    }
    if (seg) redo = march_samples<NG, NF>(M, P, o, d, S, LdsGuess{G, P.mlp_lead}, sink, w);
  } else {
    if (active && S.ok) redo = march_samples<NG, NF>(M, P, o, d, S, NoGuess(), sink, w);
  }
  int k = MARCH_DONE;
  if (active) {
    if (redo) {
      S.redo = true;
      k = MARCH_NEXT;  // same element again, exit searched with the entry face
    } else {
      const float t_from = S.ok ? S.t_out : S.t_in;
      k = march_next(M, P, S);
      if (k == MARCH_REENTER) {
        // the re-entry starts beyond the last resolved point
        S.t_in = t_from;
      }
    }
    if (k != MARCH_DONE) store_state(si, sf, r, S);
  }
  const int sn = wave_append(active && k == MARCH_NEXT, &counts[2]);
  if (active && k == MARCH_NEXT) next[sn] = r;
  const int sr = wave_append(active && k == MARCH_REENTER, &counts[3]);
  if (active && k == MARCH_REENTER) reent[sr] = r;
  finish_ray(active && k == MARCH_DONE, r, S.found, P.n_iso, n, iso_elem);
  flush_work(work, w);
}

template <int NG>
__global__ __launch_bounds__(256) void k_walk_reenter(
    const float* __restrict__ rays, int n, TraceMesh M, IsoParams P, const int* __restrict__ reent,
    int4* __restrict__ si, float4* __restrict__ sf, int* __restrict__ next, int* __restrict__ counts,
    int* __restrict__ iso_elem, long long* __restrict__ work)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < counts[3];
  Work w;
  int r = 0;
  bool cont = false;
  MarchState S;
  march_begin(S, HitF32());
  if (active) {
    r = reent[i];
    float o[3], d[3];
    load_ray(rays, n, r, o, d);
    load_state(si, sf, r, S);
    cont = march_reenter<NG>(M, P, o, d, S, w);
    if (cont) store_state(si, sf, r, S);
  }
  const int sn = wave_append(active && cont, &counts[2]);
  if (active && cont) next[sn] = r;
  finish_ray(active && !cont, r, S.found, P.n_iso, n, iso_elem);
  flush_work(work, w);
}

__global__ void k_walk_reset(int* __restrict__ counts)
{
  if (threadIdx.x == 0) {
    counts[2] = 0;
    counts[3] = 0;
  }
}

// The rounds of the walk for a trace whose first-hit queue is in
// f->d_queue_hit / counts[0].  One host synchronisation per round (the next
// round's ray count).
int walk_run(femrt_ctx* c, femrt_mesh* m, femrt_frame* f, const TraceMesh& TM, const IsoParams& ip,
             bool learned, long long* work)
{
  const int n = f->n;
  if (f->walk_cap < n || f->walk_hist_cap < m->n_elem + 1) {
    void* ptrs[] = {f->d_walk_si, f->d_walk_sf, f->d_walk_a, f->d_walk_b, f->d_walk_sorted,
                    f->d_walk_reent, f->d_walk_hist};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    f->d_walk_si = nullptr;
    f->d_walk_sf = nullptr;
    f->d_walk_a = f->d_walk_b = f->d_walk_sorted = f->d_walk_reent = f->d_walk_hist = nullptr;
    if (hipMalloc(&f->d_walk_si, (size_t)n * 16) != hipSuccess ||
        hipMalloc(&f->d_walk_sf, (size_t)n * 16) != hipSuccess ||
        hipMalloc(&f->d_walk_a, (size_t)n * 4) != hipSuccess ||
        hipMalloc(&f->d_walk_b, (size_t)n * 4) != hipSuccess ||
        hipMalloc(&f->d_walk_sorted, (size_t)n * 4) != hipSuccess ||
        hipMalloc(&f->d_walk_reent, (size_t)n * 4) != hipSuccess ||
        hipMalloc(&f->d_walk_hist, (size_t)(m->n_elem + 1) * 4) != hipSuccess)
      return fail(c, FEMRT_ENOMEM, "trace: walk buffers");
    f->walk_cap = n;
    f->walk_hist_cap = m->n_elem + 1;
  }
  const int nb = (n + 255) / 256;
  int rc = timed_launch(c, "k_walk_init", [&] {
    k_walk_init<<<nb, 256, 0, c->stream>>>(f->d_counts, f->d_queue_hit, f->d_hit_elem, f->d_hit_face,
                                           f->d_hit_f32, f->d_walk_si, f->d_walk_sf, f->d_walk_a);
  });
  if (rc) return rc;
  int n_act = 0;
  HIPCHK(c, hipMemcpyAsync(&n_act, f->d_counts + 5, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int* cur = f->d_walk_a;
  int* nxt = f->d_walk_b;
  const int ne = m->n_elem;
  const int max_rounds = 2 * std::max(1, ip.max_march) + 8;
  for (int round = 0; n_act > 0 && round < max_rounds; ++round) {
    const int ab = (n_act + 255) / 256;
    rc = timed_launch(c, "k_walk_zero", [&] {
      k_walk_zero<<<(ne + 1 + 255) / 256, 256, 0, c->stream>>>(f->d_walk_hist, ne + 1);
    });
    if (!rc) rc = timed_launch(c, "k_walk_hist", [&] {
      k_walk_hist<<<ab, 256, 0, c->stream>>>(cur, n_act, f->d_walk_si, f->d_walk_hist);
    });
    if (!rc) rc = timed_launch(c, "k_walk_scan", [&] {
      k_walk_scan<<<1, 1024, 0, c->stream>>>(f->d_walk_hist, ne + 1);
    });
    if (!rc) rc = timed_launch(c, "k_walk_sort", [&] {
      k_walk_sort<<<ab, 256, 0, c->stream>>>(cur, n_act, f->d_walk_si, f->d_walk_hist, f->d_walk_sorted);
    });
    if (!rc) rc = timed_launch(c, "k_walk_reset", [&] { k_walk_reset<<<1, 64, 0, c->stream>>>(f->d_counts); });
    if (rc) return rc;
    rc = FEMRT_EUNSUP;
    FEMRT_DISPATCH_NG(m->NG, {
      FEMRT_DISPATCH_NF(m->NF, {
        rc = timed_launch(c, "k_walk_segment", [&] {
          if (learned)
            k_walk_segment<NG_, NF_, true><<<ab, 256, 0, c->stream>>>(
                f->d_rays, n, TM, ip, f->d_walk_sorted, n_act, f->d_walk_si, f->d_walk_sf, nxt,
                f->d_walk_reent, f->d_counts, f->d_iso_elem, f->d_iso_f32, f->d_iso_rst32, work,
                m->guess_mlp->d_pack);
          else
            k_walk_segment<NG_, NF_, false><<<ab, 256, 0, c->stream>>>(
                f->d_rays, n, TM, ip, f->d_walk_sorted, n_act, f->d_walk_si, f->d_walk_sf, nxt,
                f->d_walk_reent, f->d_counts, f->d_iso_elem, f->d_iso_f32, f->d_iso_rst32, work,
                nullptr);
        });
      });
      if (rc) break;
      rc = timed_launch(c, "k_walk_reenter", [&] {
        k_walk_reenter<NG_><<<ab, 256, 0, c->stream>>>(f->d_rays, n, TM, ip, f->d_walk_reent, f->d_walk_si,
                                                        f->d_walk_sf, nxt, f->d_counts, f->d_iso_elem,
                                                        work);
      });
    });
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(&n_act, f->d_counts + 2, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::swap(cur, nxt);
  }
  return FEMRT_OK;
}
