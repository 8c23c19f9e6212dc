// femrt_kernels.hip -- gfx950 kernels of the curved-hex ray tracer and the
// C ABI that drives them.
//
// Pipeline of one trace (all on the context stream):
//   k_first_hit<NG,FLAT> one lane per ray: BVH traversal over element boxes,
//                        per-face box cull, face-constrained Newton in FP32;
//                        hit rays are compacted into a queue with a wave
//                        ballot + prefix and one atomic per wave.
//   k_polish_hit<NG,NF>  FP64 Newton polish of each hit on its face, FP64
//                        field value at the hit; chord bin of each hit.
//   k_walk_order         per queue segment, hits sorted by descending chord
//                        bin (stable): the walk starts the longest rays first.
//   k_march<NG,NF>       per hit ray: element walk through face adjacency;
//                        in each element the exit face root (FP32), uniform
//                        samples of g = u - iso over [t_in, t_out] with point
//                        inversion + field evaluation (FP32), first sign
//                        change per isovalue kept as a bracket in the frame.
//   k_polish_iso<NG,NF>  FP64 safeguarded Newton/bisection endgame per
//                        (hit ray, isovalue) bracket, in hit-queue order.
//   k_shade              RGBA from the AOVs.
// Reference: raytracer.cu:102-254 (traversal + hex branch), high_order.h:
// 189-304, 409-465 (face Newton, nearest face), fem/HexElement.cu:227-272
// (point inversion), fem/RayIsosurface.cu:67-188 (iso search).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

#include "femrt_internal.hpp"
#include "femrt_math.hpp"
#include "femrt_trace_core.hpp"

using namespace femrt;

#include "femrt_march.hpp"
#include "femrt_mlp.hpp"
#include "femrt_point.hpp"
#include "femrt_runtime.hpp"


// ================================================================== kernels

__global__ void k_raygen(float* __restrict__ rays, int W, int H, float ox, float oy, float oz,
                         float lx, float ly, float lz, float hx, float hy, float hz, float vx,
                         float vy, float vz)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = W * H;
  if (p >= n) return;
  const int y = p / W, x = p - y * W;
  const float s = (x + 0.5f) / W, t = (y + 0.5f) / H;
  float d0 = lx + s * hx + t * vx - ox;
  float d1 = ly + s * hy + t * vy - oy;
  float d2 = lz + s * hz + t * vz - oz;
  const float il = 1.f / sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
  rays[p] = ox;
  rays[n + p] = oy;
  rays[2 * n + p] = oz;
  rays[3 * n + p] = d0 * il;
  rays[4 * n + p] = d1 * il;
  rays[5 * n + p] = d2 * il;
}

// Thread -> ray order of the first-hit pass.  With an image layout (W > 0)
// consecutive threads cover square tiles of 2^sh x 2^sh pixels (sh = 2, 3
// or 4, a launch shape; the spp samples of a pixel adjacent), tiles in row
// order, so the lanes of a wave trace neighbouring pixels and the hit queue,
// which the walk consumes in order, stays tile-coherent; otherwise ray order.
// The order only decides which lane traces which ray: every ray's result is
// the same under any tile size.
struct TileOrder {
  int W, H, spp, sh, tiles_x;
  __host__ __device__ int threads(int n) const
  {
    if (W <= 0) return n;
    const int te = 1 << sh;
    return tiles_x * ((H + te - 1) >> sh) * (te * te) * spp;
  }
  __device__ bool ray(int t, int n, int& r) const
  {
    if (W <= 0) {
      r = t;
      return t < n;
    }
    const int te = 1 << sh;
    const int s = t % spp, pix = t / spp;
    const int tile = pix >> (2 * sh), in = pix & (te * te - 1);
    const int x = (tile % tiles_x) * te + (in & (te - 1)), y = (tile / tiles_x) * te + (in >> sh);
    r = (y * W + x) * spp + s;
    return x < W && y < H;
  }
};

// FLAT: the face-Newton work of the sub-patch seeds flattened across the
// wave (first_hit_flat); else each lane's search to completion
// (first_hit_ray, the reference's seed grid with PLAIN guesses).  Same hits.
template <int NG, bool FLAT>
__global__ __launch_bounds__(256) void k_first_hit(const float* __restrict__ rays, int n, TraceMesh M,
                                                   int guess, float tmin, float tmax, TileOrder order,
                                                   int* __restrict__ hit_elem, int* __restrict__ hit_face,
                                                   float4* __restrict__ hit_f32, int* __restrict__ queue,
                                                   int* __restrict__ counts, long long* __restrict__ work)
{
  int r = 0;
  const bool active = order.ray(blockIdx.x * blockDim.x + threadIdx.x, n, r);
  Work w;
  HitF32 h;
  if (active) {
    float o[3], d[3];
    load_ray(rays, n, r, o, d);
    if constexpr (FLAT)
      h = first_hit_flat<NG>(M, o, d, tmin, tmax, w);
    else
      h = first_hit_ray<NG>(M, o, d, guess, tmin, tmax, w);
    hit_elem[r] = h.elem;
    hit_face[r] = h.face;
    hit_f32[r] = make_float4(h.u, h.v, h.t, 0.f);
  }
  const bool hit = active && h.elem >= 0;
  const int slot = wave_append(hit, &counts[0]);
  if (hit) queue[slot] = r;
  if (work && active) {
    add_work(work, 0, 1);
    add_work(work, 1, hit ? 1 : 0);
    flush_work(work, w, true);
  }
}

// FP64 face Newton polish of each queued hit; FP64 field value at the hit.
// Chord bins of the walk order: the ray's length from its hit to where it
// leaves the mesh's bounding box, in FEMRT_QBINS steps of the box diagonal.
// A ray can walk no more elements than its chord crosses, so long chords
// are the candidates for long walks.
#define FEMRT_QBINS 8

__device__ __forceinline__ int chord_bin(const TraceMesh& M, const float* o, const float* d, float t_hit)
{
  if (!M.nodes) return 0;
  const float4 lo = M.nodes[0], hi = M.nodes[1];
  const float l[3] = {lo.x, lo.y, lo.z}, h[3] = {hi.x, hi.y, hi.z};
  float t_exit = 3.4e38f, dn = 0.f, diag = 0.f;
  for (int c = 0; c < 3; ++c) {
    if (d[c] != 0.f) t_exit = fminf(t_exit, fmaxf((l[c] - o[c]) / d[c], (h[c] - o[c]) / d[c]));
    dn += d[c] * d[c];
    diag += (h[c] - l[c]) * (h[c] - l[c]);
  }
  const float len = fmaxf(t_exit - t_hit, 0.f) * sqrtf(dn);
  const float x = diag > 0.f ? len * rsqrtf(diag) * FEMRT_QBINS : 0.f;
  return x >= FEMRT_QBINS - 1 ? FEMRT_QBINS - 1 : (int)x;
}

// FP64 results of a polished hit.
template <int NF>
__device__ __forceinline__ void store_hit64(const TraceMesh& M, int n, int r, int e, int f, double u, double v,
                                            double tau, double* __restrict__ hit_t, double* __restrict__ hit_rst,
                                            double* __restrict__ hit_u)
{
  double rst[3];
  face_to_rst_t<double>(f, u, v, rst);
  hit_t[r] = tau;
  hit_rst[r] = rst[0];
  hit_rst[(size_t)n + r] = rst[1];
  hit_rst[2 * (size_t)n + r] = rst[2];
  hit_u[r] = field_value<double, NF>(M.field + (size_t)e * (NF * NF * NF), rst[0], rst[1], rst[2]);
}

template <int NG, int NF>
__global__ __launch_bounds__(256) void k_polish_hit(
    const float* __restrict__ rays, int n, TraceMesh M, float tmin, const int* __restrict__ queue,
    int* __restrict__ counts, const int* __restrict__ hit_elem,
    const int* __restrict__ hit_face, float4* __restrict__ hit_f32,
    double* __restrict__ hit_t, double* __restrict__ hit_rst, double* __restrict__ hit_u,
    unsigned char* __restrict__ qbin, int* __restrict__ retrace, long long* __restrict__ work)
{
  const int count = counts[0];
  Work w;
  for (XcdStride it(count); it.q < it.end; it.q += it.step) {
    const int q = it.q;
    const int r = queue[q];
    float of[3], df[3];
    load_ray(rays, n, r, of, df);
    qbin[q] = (unsigned char)chord_bin(M, of, df, hit_f32[r].z);
    const double o[3] = {of[0], of[1], of[2]}, d[3] = {df[0], df[1], df[2]};
    const int e = hit_elem[r], f = hit_face[r];
    const float4 h = hit_f32[r];
    double u = h.x, v = h.y, tau = h.z;
    if (polish_hit<NG>(M.geom + (size_t)e * (NG * NG * NG), o, d, f, u, v, tau, (double)tmin, w)) {
      store_hit64<NF>(M, n, r, e, f, u, v, tau, hit_t, hit_rst, hit_u);
      // the walk starts at the polished root: where the FP32 root was a near
      // miss just outside the face and the polish moved on to the face's real
      // crossing, the FP32 point lies outside the element
      hit_f32[r] = make_float4((float)u, (float)v, (float)tau, 0.f);
      ++w.field64;
    } else {
      // no FP64 root on the face: a near miss, traced again by k_retrace
      retrace[atomicAdd(&counts[FEMRT_RETRACE_CTR], 1)] = r;
    }
  }
  flush_work(work, w);
}

// Rays whose FP32 hit had no FP64 root on its face (a grazing near miss that
// the FP32 face Newton accepted within its residual tolerance): traced again
// with that (element, face) excluded, and the new hit polished; up to four
// exclusions, then the ray counts as a miss.  The walk starts from the new
// FP32 hit; a ray left without a hit is skipped by the walk.
template <int NG, int NF>
__global__ __launch_bounds__(64) void k_retrace(const float* __restrict__ rays, int n, TraceMesh M, int guess,
                                                float tmin, float tmax, const int* __restrict__ retrace,
                                                const int* __restrict__ counts, int* __restrict__ hit_elem,
                                                int* __restrict__ hit_face, float4* __restrict__ hit_f32,
                                                double* __restrict__ hit_t, double* __restrict__ hit_rst,
                                                double* __restrict__ hit_u, long long* __restrict__ work)
{
  const int count = counts[FEMRT_RETRACE_CTR];
  Work w;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < count; q += gridDim.x * blockDim.x) {
    const int r = retrace[q];
    float of[3], df[3];
    load_ray(rays, n, r, of, df);
    const double o[3] = {of[0], of[1], of[2]}, d[3] = {df[0], df[1], df[2]};
    HitExcl ex;
    ex.add(hit_elem[r], hit_face[r]);
    int e = -1;
    for (int attempt = 0; attempt < 4; ++attempt) {
      const HitF32 h = first_hit_ray<NG>(M, of, df, guess, tmin, tmax, w, -1, &ex);
      if (h.elem < 0) break;
      double u = h.u, v = h.v, tau = h.t;
      if (polish_hit<NG>(M.geom + (size_t)h.elem * (NG * NG * NG), o, d, h.face, u, v, tau, (double)tmin, w)) {
        e = h.elem;
        hit_face[r] = h.face;
        hit_f32[r] = make_float4((float)u, (float)v, (float)tau, 0.f);   // the polished root, as k_polish_hit
        store_hit64<NF>(M, n, r, h.elem, h.face, u, v, tau, hit_t, hit_rst, hit_u);
        break;
      }
      ex.add(h.elem, h.face);
    }
    hit_elem[r] = e;
  }
  flush_work(work, w);
}

// Walk order of the hit queue.  The queue is split into nseg (one block
// each) contiguous segments (tile order: each covers one band of the image),
// as the walk reads it; within each segment the hits are put in descending
// chord bin, tile order kept within a bin (a stable counting sort, one block
// per segment, each thread a contiguous chunk).  The walk then starts the
// rays that may walk longest first, and the rays still walking when the
// queue drains are short ones.
#define FEMRT_ORDER_THREADS 1024

__global__ __launch_bounds__(FEMRT_ORDER_THREADS) void k_walk_order(
    const int* __restrict__ queue, const unsigned char* __restrict__ qbin,
    const int* __restrict__ counts, int* __restrict__ out, int nseg)
{
  __shared__ int s_cnt[FEMRT_QBINS][FEMRT_ORDER_THREADS];
  __shared__ int s_base[FEMRT_QBINS];
  const int count = counts[0], seg = blockIdx.x, t = threadIdx.x;
  const int lo = (int)((long long)count * seg / nseg);
  const int hi = (int)((long long)count * (seg + 1) / nseg);
  const int c0 = lo + (int)((long long)(hi - lo) * t / FEMRT_ORDER_THREADS);
  const int c1 = lo + (int)((long long)(hi - lo) * (t + 1) / FEMRT_ORDER_THREADS);
  int cnt[FEMRT_QBINS];
  for (int b = 0; b < FEMRT_QBINS; ++b) cnt[b] = 0;
  for (int q = c0; q < c1; ++q) {
    const int b = qbin[q];
    for (int k = 0; k < FEMRT_QBINS; ++k) cnt[k] += (k == b) ? 1 : 0;
  }
  for (int b = 0; b < FEMRT_QBINS; ++b) s_cnt[b][t] = cnt[b];
  __syncthreads();
  if (t < FEMRT_QBINS) {
    // exclusive prefix over the threads' chunks for bin t; s_base[t] = total
    int acc = 0;
    for (int i = 0; i < FEMRT_ORDER_THREADS; ++i) {
      const int v = s_cnt[t][i];
      s_cnt[t][i] = acc;
      acc += v;
    }
    s_base[t] = acc;
  }
  __syncthreads();
  if (t == 0) {
    // bins in descending order: the longest chords first
    int acc = 0;
    for (int b = FEMRT_QBINS - 1; b >= 0; --b) {
      const int v = s_base[b];
      s_base[b] = acc;
      acc += v;
    }
  }
  __syncthreads();
  int off[FEMRT_QBINS];
  for (int b = 0; b < FEMRT_QBINS; ++b) off[b] = lo + s_base[b] + s_cnt[b][t];
  for (int q = c0; q < c1; ++q) {
    const int b = qbin[q];
    int dst = 0;
    for (int k = 0; k < FEMRT_QBINS; ++k)
      if (k == b) dst = off[k]++;
    out[dst] = queue[q];
  }
}

template <int NG, int NF>
__global__ __launch_bounds__(256, 2) void k_polish_iso(
    const float* __restrict__ rays, int n, TraceMesh M, IsoParams P, const int* __restrict__ hit_queue,
    const int* __restrict__ counts, const float4* __restrict__ iso_f32, const float4* __restrict__ iso_rst32,
    int* __restrict__ iso_elem, double* __restrict__ iso_t, double* __restrict__ iso_rst,
    double* __restrict__ iso_u, long long* __restrict__ work, const float2* __restrict__ iso_pred)
{
  // (hit ray, isovalue) pairs in hit-queue order, which is tile order, so
  // neighbouring lanes polish roots in the same elements
  const int count = counts[0] * P.n_iso;
  Work w;
  for (XcdStride it(count); it.q < it.end; it.q += it.step) {
    const int q = it.q;
    const int r = hit_queue[q / P.n_iso], k = q % P.n_iso;
    const size_t slot = (size_t)k * n + r;
    if (iso_elem[slot] < 0) continue;
    float of[3], df[3];
    load_ray(rays, n, r, of, df);
    const double o[3] = {of[0], of[1], of[2]}, d[3] = {df[0], df[1], df[2]};
    IsoBracket br;
    br.elem = iso_elem[slot];
    const float4 b4 = iso_f32[slot];
    const float4 r4 = iso_rst32[slot];
    br.ta = b4.x;
    br.tb = b4.y;
    br.tm = b4.z;
    br.ga = b4.w;
    br.rst[0] = r4.x;
    br.rst[1] = r4.y;
    br.rst[2] = r4.z;
    double t, rst[3], u;
    bool ok;
    if (P.endgame == ISO_NEWTON) {
      ok = polish_iso<NG, NF>(M, o, d, P.iso64[k], br, t, rst, u, w);
    } else {
      const bool pred = P.endgame == ISO_LEARNED_BISECT;
      const float2 pr = pred ? iso_pred[slot] : make_float2(0.f, 0.f);
      ok = polish_iso_bisect<NG, NF>(M, o, d, P.iso64[k], br, pred, pr.x, pr.y, t, rst, u, w);
    }
    if (ok) {
      iso_t[slot] = t;
      iso_rst[(size_t)k * 3 * n + r] = rst[0];
      iso_rst[(size_t)k * 3 * n + n + r] = rst[1];
      iso_rst[(size_t)k * 3 * n + 2 * n + r] = rst[2];
      iso_u[slot] = u;
    } else {
      iso_elem[slot] = -1;
    }
  }
  flush_work(work, w);
}

__global__ void k_clear_frame(int n, int n_iso, int* hit_elem, int* iso_elem, int* counts)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) {
    hit_elem[r] = -1;
    for (int k = 0; k < n_iso; ++k) iso_elem[(size_t)k * n + r] = -1;
  }
  if (r < FEMRT_NCOUNTS) counts[r] = 0;
}

__global__ void k_shade(int n, int n_iso, const float* __restrict__ rays,
                        const int* __restrict__ hit_elem, const double* __restrict__ hit_u,
                        const int* __restrict__ iso_elem, uint32_t* __restrict__ rgba)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  float c[3];
  int first = -1;
  for (int k = 0; k < n_iso; ++k)
    if (iso_elem[(size_t)k * n + r] >= 0) { first = k; break; }
  if (first >= 0) {
    const float a = n_iso > 1 ? (float)first / (n_iso - 1) : 0.5f;
    c[0] = 0.2f + 0.7f * a;
    c[1] = 0.3f;
    c[2] = 0.9f - 0.7f * a;
  } else if (hit_elem[r] >= 0) {
    const float u = (float)hit_u[r];
    const float a = fminf(fmaxf(0.5f + 0.5f * u, 0.f), 1.f);
    c[0] = c[1] = c[2] = 0.3f + 0.5f * a;
  } else {
    // sky gradient (raytracer.cu:398-402)
    const float dy = rays[4 * (size_t)n + r];
    const float t = 0.5f * (dy + 1.f);
    c[0] = (1.f - t) + t * 0.5f;
    c[1] = (1.f - t) + t * 0.7f;
    c[2] = 1.f;
  }
  uint32_t px = 0xFF000000u;
  for (int i = 0; i < 3; ++i) {
    const float g = sqrtf(fminf(fmaxf(c[i], 0.f), 1.f));
    px |= (uint32_t)(255.99f * g) << (8 * i);
  }
  rgba[r] = px;
}

// Point kernels: field value + world gradient (femrt_point.hpp); inversion.
template <int NG>
__global__ void k_invert(int n, const int* __restrict__ elem, const float* __restrict__ x,
                         const GeoMode* __restrict__ geom, double* __restrict__ rst_out,
                         int* __restrict__ iters_out)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int e = elem[p];
  const GeoMode* gc = geom + (size_t)e * (NG * NG * NG);
  const float xf[3] = {x[p], x[n + p], x[2 * n + p]};
  float rst[3] = {0.f, 0.f, 0.f}, J[9];
  unsigned it = 0, it64 = 0;
  const bool ok = invert_in_cube_f32<NG>(gc, xf, rst, J, it);
  double rd[3] = {rst[0], rst[1], rst[2]}, Jd[9];
  const double xd[3] = {xf[0], xf[1], xf[2]};
  const bool ok64 = ok && invert_f64<NG>(gc, xd, rd, Jd, it64);
  rst_out[p] = rd[0];
  rst_out[n + p] = rd[1];
  rst_out[2 * n + p] = rd[2];
  if (iters_out) iters_out[p] = ok64 ? (int)it : -1;
}

// ============================================================ mode dispatch

static int pad_ng(int ng)
{
  if (ng <= 4) return 4;
  if (ng == 5) return 5;
  if (ng <= 7) return 7;
  return 0;
}
static int pad_nf(int nf)
{
  if (nf <= 3) return 3;
  if (nf <= 5) return 5;
  if (nf <= 7) return 7;
  if (nf <= 9) return 9;
  return 0;
}

// Pad mode counts: coefficient (i,j,k) of an n-mode expansion moves to the
// same (i,j,k) of an N-mode one; higher modes are zero, so the map is
// unchanged.
static void pad_coeffs(const float* src, int n, int N, int comps, size_t n_elem, std::vector<float>& dst,
                       int stride_out)
{
  const size_t n3 = (size_t)n * n * n, N3 = (size_t)N * N * N;
  for (size_t e = 0; e < n_elem; ++e)
    for (int c = 0; c < comps; ++c)
      for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < n; ++i)
            dst[(e * N3 + ((size_t)k * N + j) * N + i) * stride_out + c] =
                src[(e * comps + c) * n3 + ((size_t)k * n + j) * n + i];
}

// ================================================================== C ABI

extern "C" {

FEMRT_API int femrt_version(void) { return 100; }

FEMRT_API int femrt_ctx_create(int device, femrt_ctx** out)
{
  if (!out) return FEMRT_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return FEMRT_EHIP;
  if (device < 0 || device >= ndev) return FEMRT_EINVAL;
  femrt_ctx* c = new femrt_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return FEMRT_EHIP;
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
    c->cus = cus;
  *out = c;
  return FEMRT_OK;
}

FEMRT_API int femrt_ctx_destroy(femrt_ctx* c)
{
  if (!c) return FEMRT_EINVAL;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  collect_timing(c);
  for (auto e : c->pool) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return FEMRT_OK;
}

FEMRT_API const char* femrt_last_error(const femrt_ctx* c) { return c ? c->err.c_str() : "null context"; }

FEMRT_API int femrt_sync(femrt_ctx* c)
{
  if (!c) return FEMRT_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  return FEMRT_OK;
}

FEMRT_API int femrt_timing_enable(femrt_ctx* c, int on)
{
  if (!c) return FEMRT_EINVAL;
  c->timing = on != 0;
  return FEMRT_OK;
}

FEMRT_API int femrt_timing_reset(femrt_ctx* c)
{
  if (!c) return FEMRT_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  c->stats.clear();
  return FEMRT_OK;
}

FEMRT_API int femrt_timing_read(femrt_ctx* c, int cap, char* names48, double* total_ms,
                                long long* launches)
{
  if (!c) return FEMRT_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  int i = 0;
  for (auto& kv : c->stats) {
    if (i < cap) {
      if (names48) {
        std::memset(names48 + 48 * i, 0, 48);
        std::strncpy(names48 + 48 * i, kv.first.c_str(), 47);
      }
      if (total_ms) total_ms[i] = kv.second.first;
      if (launches) launches[i] = kv.second.second;
    }
    ++i;
  }
  return i;
}

FEMRT_API int femrt_mesh_upload_hex(femrt_ctx* c, int n_elem, int ng, int nf, const float* vertices,
                                    const float* geom, const float* field, const int* adjacency,
                                    femrt_mesh** out)
{
  if (!c || !out || !vertices || !geom || !field || n_elem <= 0) return fail(c, FEMRT_EINVAL, "mesh_upload_hex: bad argument");
  *out = nullptr;
  const int NG = pad_ng(ng), NF = pad_nf(nf);
  if (ng < 2 || nf < 1 || !NG || !NF)
    return fail(c, FEMRT_EUNSUP, "mesh_upload_hex: geometry modes must be 2..7 and field modes 1..9");
  HIPCHK(c, hipSetDevice(c->device));
  femrt_mesh* m = new femrt_mesh();
  m->ctx = c;
  m->n_elem = n_elem;
  m->ng = ng;
  m->nf = nf;
  m->NG = NG;
  m->NF = NF;
  const size_t N3 = (size_t)NG * NG * NG, F3 = (size_t)NF * NF * NF;
  std::vector<float> g4((size_t)n_elem * N3 * 3, 0.f), fp((size_t)n_elem * F3, 0.f);
  pad_coeffs(geom, ng, NG, 3, n_elem, g4, 3);
  pad_coeffs(field, nf, NF, 1, n_elem, fp, 1);
  m->h_boxes.resize((size_t)n_elem * FEMRT_BOX_STRIDE);
  bernstein_boxes(n_elem, ng, geom, m->h_boxes.data());
  std::vector<float> lip(n_elem), frange(2 * (size_t)n_elem), jinv(n_elem);
  field_lipschitz(n_elem, nf, field, lip.data());
  geom_jinv_bound(n_elem, ng, geom, jinv.data());
  field_range(n_elem, nf, field, frange.data());
  auto cleanup = [&](const std::string& msg) {
    (void)hipFree(m->d_geom); (void)hipFree(m->d_field); (void)hipFree(m->d_verts); (void)hipFree(m->d_boxes); (void)hipFree(m->d_adj);
    (void)hipFree(m->d_lip);
    (void)hipFree(m->d_jinv);
    (void)hipFree(m->d_frange);
    delete m;
    return fail(c, FEMRT_ENOMEM, msg);
  };
  if (hipMalloc(&m->d_geom, g4.size() * 4) != hipSuccess) return cleanup("hipMalloc geom");
  if (hipMalloc(&m->d_field, fp.size() * 4) != hipSuccess) return cleanup("hipMalloc field");
  if (hipMalloc(&m->d_verts, (size_t)n_elem * 24 * 4) != hipSuccess) return cleanup("hipMalloc verts");
  if (hipMalloc(&m->d_boxes, m->h_boxes.size() * 4) != hipSuccess) return cleanup("hipMalloc boxes");
  if (adjacency && hipMalloc(&m->d_adj, (size_t)n_elem * 12 * 4) != hipSuccess) return cleanup("hipMalloc adj");
  if (hipMalloc(&m->d_lip, (size_t)n_elem * 4) != hipSuccess) return cleanup("hipMalloc lip");
  if (hipMalloc(&m->d_jinv, (size_t)n_elem * 4) != hipSuccess) return cleanup("hipMalloc jinv");
  if (hipMalloc(&m->d_frange, (size_t)n_elem * 8) != hipSuccess) return cleanup("hipMalloc field range");
  bool ok = hipMemcpy(m->d_geom, g4.data(), g4.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(m->d_field, fp.data(), fp.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(m->d_verts, vertices, (size_t)n_elem * 24 * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(m->d_boxes, m->h_boxes.data(), m->h_boxes.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(m->d_lip, lip.data(), lip.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(m->d_jinv, jinv.data(), jinv.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(m->d_frange, frange.data(), frange.size() * 4, hipMemcpyHostToDevice) == hipSuccess;
  if (ok && adjacency) {
    std::vector<int> adj((size_t)n_elem * 12);
    encode_adjacency(n_elem, vertices, adjacency, adj.data());
    ok = hipMemcpy(m->d_adj, adj.data(), adj.size() * 4, hipMemcpyHostToDevice) == hipSuccess;
  }
  if (!ok) return cleanup("hipMemcpy mesh");
  *out = m;
  return FEMRT_OK;
}

FEMRT_API int femrt_mesh_upload_quads(femrt_ctx* c, int n_quad, int ng, int nf, const float* vertices,
                                      const float* geom, const float* field, femrt_mesh** out)
{
  if (!c || !out || !vertices || n_quad <= 0) return fail(c, FEMRT_EINVAL, "mesh_upload_quads: bad argument");
  if (!geom) ng = 2;
  if (!field) nf = 1;
  if (ng < 2 || ng > 7 || nf < 1 || nf > 9)
    return fail(c, FEMRT_EUNSUP, "mesh_upload_quads: geometry modes must be 2..7 and field modes 1..9");
  // Each quad becomes the face t = -1 of a map constant in t: the mode
  // (i, j) coefficient goes to modes (i, j, 0) and (i, j, 1), since
  // phi_0(t) + phi_1(t) = 1.  Corners: quad vertex q = i + 2 j sits at
  // (r, s) = (2 i - 1, 2 j - 1) (QuadElement.cu ShapeWeight); hex corners are
  // in XML order (r, s) = (-1,-1), (1,-1), (1,1), (-1,1), twice.
  const size_t G2 = (size_t)ng * ng, G3 = G2 * ng, F2 = (size_t)nf * nf, F3 = F2 * nf;
  std::vector<float> v8((size_t)n_quad * 24), g3((size_t)n_quad * 3 * G3, 0.f), f3((size_t)n_quad * F3, 0.f);
  const int qof[4] = {0, 1, 3, 2};
  for (int e = 0; e < n_quad; ++e) {
    const float* q = vertices + (size_t)e * 12;
    for (int c8 = 0; c8 < 8; ++c8)
      for (int d = 0; d < 3; ++d) v8[(size_t)e * 24 + 3 * c8 + d] = q[3 * qof[c8 & 3] + d];
    for (int d = 0; d < 3; ++d) {
      float* dst = &g3[((size_t)e * 3 + d) * G3];
      for (int j = 0; j < ng; ++j)
        for (int i = 0; i < ng; ++i) {
          const float v = geom ? geom[((size_t)e * 3 + d) * G2 + (size_t)j * ng + i] : q[3 * (i + 2 * j) + d];
          dst[(size_t)j * ng + i] = v;                 // k = 0
          dst[G2 + (size_t)j * ng + i] = v;            // k = 1
        }
    }
    if (field)
      for (int j = 0; j < nf; ++j)
        for (int i = 0; i < nf; ++i) {
          const float v = field[(size_t)e * F2 + (size_t)j * nf + i];
          f3[(size_t)e * F3 + (size_t)j * nf + i] = v;
          if (nf > 1) f3[(size_t)e * F3 + F2 + (size_t)j * nf + i] = v;
        }
  }
  const int rc = femrt_mesh_upload_hex(c, n_quad, ng, nf, v8.data(), g3.data(), f3.data(), nullptr, out);
  if (rc == FEMRT_OK) (*out)->surface = true;
  return rc;
}

FEMRT_API int femrt_host_prism_to_hex(int n, const int* geom_modes, const int* field_modes, const float* vertices,
                                      const float* geom, const float* field, int* ng, int* nf,
                                      float* hex_vertices, double* hex_geom, double* hex_field)
{
  if (n <= 0 || !vertices || !ng || !nf) return FEMRT_EINVAL;
  if ((geom && !geom_modes) || (field && !field_modes)) return FEMRT_EINVAL;
  const int G = geom ? std::max(geom_modes[0], std::max(geom_modes[1], geom_modes[2])) : 2;
  const int F = field ? std::max(field_modes[0], std::max(field_modes[1], field_modes[2])) : 1;
  if (geom && (std::min(geom_modes[0], std::min(geom_modes[1], geom_modes[2])) < 1 || G < 2 || G > 7))
    return FEMRT_EUNSUP;
  if (field && (std::min(field_modes[0], std::min(field_modes[1], field_modes[2])) < 1 || F > 9))
    return FEMRT_EUNSUP;
  *ng = G;
  *nf = F;
  // corners of the hex embedding in XML order: the reference's linear prism
  // (ReferenceToWorldLinearPrism, fem/PrismElement.cu:194-226) is the
  // trilinear map of v0 v1 v2 v3 v4 v4 v5 v5 (the t = +1 face collapsed onto
  // the edge v4 v5)
  static const int kCorner[8] = {0, 1, 2, 3, 4, 4, 5, 5};
  if (hex_vertices)
    for (int e = 0; e < n; ++e)
      for (int c = 0; c < 8; ++c)
        for (int d = 0; d < 3; ++d) hex_vertices[((size_t)e * 8 + c) * 3 + d] = vertices[((size_t)e * 6 + kCorner[c]) * 3 + d];
  if (hex_geom) {
    if (geom) {
      femrt::prism_to_hex(n, geom_modes, G, geom, 3, hex_geom);
    } else {
      // trilinear coefficients: mode (i, j, k) is the corner (r, s, t) =
      // (2i-1, 2j-1, 2k-1), XML corner 0 1 3 2 4 5 7 6 of i + 2 j + 4 k
      static const int kXml[8] = {0, 1, 3, 2, 4, 5, 7, 6};
      for (int e = 0; e < n; ++e)
        for (int d = 0; d < 3; ++d)
          for (int b = 0; b < 8; ++b)
            hex_geom[((size_t)e * 3 + d) * 8 + b] = vertices[((size_t)e * 6 + kCorner[kXml[b]]) * 3 + d];
    }
  }
  if (hex_field) {
    if (field)
      femrt::prism_to_hex(n, field_modes, F, field, 1, hex_field);
    else
      std::fill(hex_field, hex_field + n, 0.0);
  }
  return FEMRT_OK;
}

FEMRT_API int femrt_mesh_upload_prisms(femrt_ctx* c, int n, const int* geom_modes, const int* field_modes,
                                       const float* vertices, const float* geom, const float* field,
                                       femrt_mesh** out)
{
  if (!c || !out || !vertices || n <= 0) return fail(c, FEMRT_EINVAL, "mesh_upload_prisms: bad argument");
  *out = nullptr;
  int ng = 0, nf = 0;
  int rc = femrt_host_prism_to_hex(n, geom_modes, field_modes, vertices, geom, field, &ng, &nf, nullptr, nullptr,
                                   nullptr);
  if (rc == FEMRT_EINVAL) return fail(c, rc, "mesh_upload_prisms: mode counts missing");
  if (rc != FEMRT_OK)
    return fail(c, rc, "mesh_upload_prisms: geometry modes must be 1..7 (at least 2 in one direction) and field modes 1..9");
  const size_t G3 = (size_t)ng * ng * ng, F3 = (size_t)nf * nf * nf;
  std::vector<float> v8((size_t)n * 24);
  std::vector<double> g((size_t)n * 3 * G3), f((size_t)n * F3);
  femrt_host_prism_to_hex(n, geom_modes, field_modes, vertices, geom, field, &ng, &nf, v8.data(), g.data(), f.data());
  std::vector<float> g32(g.begin(), g.end()), f32(f.begin(), f.end());
  rc = femrt_mesh_upload_hex(c, n, ng, nf, v8.data(), g32.data(), f32.data(), nullptr, out);
  if (rc == FEMRT_OK) (*out)->prism = true;
  return rc;
}

FEMRT_API int femrt_host_tri_to_quad(int n, const int* geom_modes, const int* field_modes, const float* vertices,
                                     const float* geom, const float* field, int* ng, int* nf, float* quad_vertices,
                                     double* quad_geom, double* quad_field)
{
  if (n <= 0 || !vertices || !ng || !nf) return FEMRT_EINVAL;
  if ((geom && !geom_modes) || (field && !field_modes)) return FEMRT_EINVAL;
  const int G = geom ? std::max(geom_modes[0], geom_modes[1]) : 2;
  const int F = field ? std::max(field_modes[0], field_modes[1]) : 1;
  if (geom && (std::min(geom_modes[0], geom_modes[1]) < 1 || G < 2 || G > 7)) return FEMRT_EUNSUP;
  if (field && (std::min(field_modes[0], field_modes[1]) < 1 || F > 9)) return FEMRT_EUNSUP;
  *ng = G;
  *nf = F;
  // the flat triangle in collapsed coordinates: the bilinear quad of
  // v0 v1 v2 v2 (corner i + 2 j at (2i-1, 2j-1)), s = +1 the collapsed edge
  static const int kCorner[4] = {0, 1, 2, 2};
  if (quad_vertices)
    for (int e = 0; e < n; ++e)
      for (int q = 0; q < 4; ++q)
        for (int d = 0; d < 3; ++d) quad_vertices[((size_t)e * 4 + q) * 3 + d] = vertices[((size_t)e * 3 + kCorner[q]) * 3 + d];
  if (quad_geom) {
    if (geom) {
      femrt::tri_to_quad(n, geom_modes, G, geom, 3, quad_geom);
    } else {
      for (int e = 0; e < n; ++e)
        for (int d = 0; d < 3; ++d)
          for (int q = 0; q < 4; ++q) quad_geom[((size_t)e * 3 + d) * 4 + q] = vertices[((size_t)e * 3 + kCorner[q]) * 3 + d];
    }
  }
  if (quad_field) {
    if (field)
      femrt::tri_to_quad(n, field_modes, F, field, 1, quad_field);
    else
      std::fill(quad_field, quad_field + n, 0.0);
  }
  return FEMRT_OK;
}

FEMRT_API int femrt_mesh_upload_tris(femrt_ctx* c, int n, const int* geom_modes, const int* field_modes,
                                     const float* vertices, const float* geom, const float* field, femrt_mesh** out)
{
  if (!c || !out || !vertices || n <= 0) return fail(c, FEMRT_EINVAL, "mesh_upload_tris: bad argument");
  *out = nullptr;
  int ng = 0, nf = 0;
  int rc = femrt_host_tri_to_quad(n, geom_modes, field_modes, vertices, geom, field, &ng, &nf, nullptr, nullptr,
                                  nullptr);
  if (rc == FEMRT_EINVAL) return fail(c, rc, "mesh_upload_tris: mode counts missing");
  if (rc != FEMRT_OK)
    return fail(c, rc, "mesh_upload_tris: geometry modes must be 1..7 (at least 2 in one direction) and field modes 1..9");
  std::vector<float> v4((size_t)n * 12);
  std::vector<double> g((size_t)n * 3 * ng * ng), f((size_t)n * nf * nf);
  femrt_host_tri_to_quad(n, geom_modes, field_modes, vertices, geom, field, &ng, &nf, v4.data(), g.data(), f.data());
  std::vector<float> g32(g.begin(), g.end()), f32(f.begin(), f.end());
  return femrt_mesh_upload_quads(c, n, ng, nf, v4.data(), g32.data(), f32.data(), out);
}

FEMRT_API int femrt_mesh_destroy(femrt_mesh* m)
{
  if (!m) return FEMRT_EINVAL;
  (void)hipSetDevice(m->ctx->device);
  (void)hipStreamSynchronize(m->ctx->stream);
  (void)hipFree(m->d_geom);
  (void)hipFree(m->d_field);
  (void)hipFree(m->d_verts);
  (void)hipFree(m->d_boxes);
  (void)hipFree(m->d_adj);
  (void)hipFree(m->d_lip);
  (void)hipFree(m->d_jinv);
  (void)hipFree(m->d_frange);
  (void)hipFree(m->d_nodes);
  (void)hipFree(m->d_prims);
  delete m;
  return FEMRT_OK;
}

FEMRT_API int femrt_mesh_boxes(const femrt_mesh* m, float* out)
{
  if (!m || !out) return FEMRT_EINVAL;
  std::memcpy(out, m->h_boxes.data(), m->h_boxes.size() * 4);
  return FEMRT_OK;
}

static int bvh_depth(const std::vector<float>& nodes, int node, int depth)
{
  int a, b;
  std::memcpy(&a, &nodes[(size_t)node * 8 + 3], 4);
  std::memcpy(&b, &nodes[(size_t)node * 8 + 7], 4);
  if (b < 0) return depth;
  return std::max(bvh_depth(nodes, a, depth + 1), bvh_depth(nodes, b, depth + 1));
}

FEMRT_API int femrt_mesh_build_bvh(femrt_mesh* m, int builder, int* n_nodes)
{
  if (!m) return FEMRT_EINVAL;
  femrt_ctx* c = m->ctx;
  if (builder != FEMRT_BVH_HOST_SAH && builder != FEMRT_BVH_GPU_LBVH)
    return fail(c, FEMRT_EINVAL, "build_bvh: builder must be 0 (host binned SAH) or 1 (GPU LBVH)");
  if (builder == FEMRT_BVH_GPU_LBVH) {
    const int rc = lbvh_build(c, m);
    if (rc != FEMRT_OK) return rc;
    m->bvh_depth = bvh_depth(m->bvh.nodes, 0, 0);
    if (m->bvh_depth >= FEMRT_STACK)
      return fail(c, FEMRT_EUNSUP, "build_bvh: tree deeper than the traversal stack");
    if (n_nodes) *n_nodes = m->n_nodes;
    return FEMRT_OK;
  }
  build_bvh_sah(m->n_elem, m->h_boxes.data(), m->bvh);
  m->n_nodes = (int)(m->bvh.nodes.size() / 8);
  m->bvh_depth = bvh_depth(m->bvh.nodes, 0, 0);
  if (m->bvh_depth >= FEMRT_STACK)
    return fail(c, FEMRT_EUNSUP, "build_bvh: tree deeper than the traversal stack");
  HIPCHK(c, hipSetDevice(c->device));
  (void)hipFree(m->d_nodes);
  (void)hipFree(m->d_prims);
  m->d_nodes = nullptr;
  m->d_prims = nullptr;
  HIPCHK(c, hipMalloc(&m->d_nodes, m->bvh.nodes.size() * 4));
  HIPCHK(c, hipMalloc(&m->d_prims, m->bvh.prims.size() * 4));
  HIPCHK(c, hipMemcpy(m->d_nodes, m->bvh.nodes.data(), m->bvh.nodes.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(m->d_prims, m->bvh.prims.data(), m->bvh.prims.size() * 4, hipMemcpyHostToDevice));
  if (n_nodes) *n_nodes = m->n_nodes;
  return FEMRT_OK;
}

FEMRT_API int femrt_mesh_get_bvh(const femrt_mesh* m, float* nodes8, int* prims)
{
  if (!m) return FEMRT_EINVAL;
  if (m->n_nodes == 0) return fail(m->ctx, FEMRT_ESTATE, "get_bvh: BVH not built");
  if (nodes8) std::memcpy(nodes8, m->bvh.nodes.data(), m->bvh.nodes.size() * 4);
  if (prims) std::memcpy(prims, m->bvh.prims.data(), m->bvh.prims.size() * 4);
  return FEMRT_OK;
}

FEMRT_API int femrt_eval_field(femrt_mesh* m, int n, const int* elem, const float* rst, int precision,
                               void* out)
{
  if (!m || n < 0 || (n > 0 && (!elem || !rst || !out))) return FEMRT_EINVAL;
  femrt_ctx* c = m->ctx;
  if (n == 0) return FEMRT_OK;
  for (int i = 0; i < n; ++i)
    if (elem[i] < 0 || elem[i] >= m->n_elem) return fail(c, FEMRT_EINVAL, "eval_field: element index out of range");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t osz = precision ? sizeof(double) : sizeof(float);
  int* d_e = nullptr;
  float* d_r = nullptr;
  void* d_o = nullptr;
  HIPCHK(c, hipMalloc(&d_e, (size_t)n * 4));
  HIPCHK(c, hipMalloc(&d_r, (size_t)n * 12));
  HIPCHK(c, hipMalloc(&d_o, (size_t)n * 4 * osz));
  HIPCHK(c, hipMemcpyAsync(d_e, elem, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_r, rst, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
  const int blocks = (n + 255) / 256;
  int rc = FEMRT_EUNSUP;
  FEMRT_DISPATCH_NG(m->NG, FEMRT_DISPATCH_NF(m->NF, {
    if (precision)
      rc = timed_launch(c, "k_eval_field_f64", [&] {
        k_eval_field<double, NG_, NF_><<<blocks, 256, 0, c->stream>>>(n, d_e, d_r, m->d_geom, m->d_field, (double*)d_o);
      });
    else
      rc = timed_launch(c, "k_eval_field_f32", [&] {
        k_eval_field<float, NG_, NF_><<<blocks, 256, 0, c->stream>>>(n, d_e, d_r, m->d_geom, m->d_field, (float*)d_o);
      });
  }));
  if (rc == FEMRT_OK) {
    HIPCHK(c, hipMemcpyAsync(out, d_o, (size_t)n * 4 * osz, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  (void)hipFree(d_e);
  (void)hipFree(d_r);
  (void)hipFree(d_o);
  return rc;
}

FEMRT_API int femrt_invert_points(femrt_mesh* m, int n, const int* elem, const float* x, double* rst_out,
                                  int* iters_out)
{
  if (!m || n < 0 || (n > 0 && (!elem || !x || !rst_out))) return FEMRT_EINVAL;
  femrt_ctx* c = m->ctx;
  if (n == 0) return FEMRT_OK;
  for (int i = 0; i < n; ++i)
    if (elem[i] < 0 || elem[i] >= m->n_elem) return fail(c, FEMRT_EINVAL, "invert_points: element index out of range");
  HIPCHK(c, hipSetDevice(c->device));
  int *d_e = nullptr, *d_it = nullptr;
  float* d_x = nullptr;
  double* d_r = nullptr;
  HIPCHK(c, hipMalloc(&d_e, (size_t)n * 4));
  HIPCHK(c, hipMalloc(&d_x, (size_t)n * 12));
  HIPCHK(c, hipMalloc(&d_r, (size_t)n * 24));
  HIPCHK(c, hipMalloc(&d_it, (size_t)n * 4));
  HIPCHK(c, hipMemcpyAsync(d_e, elem, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_x, x, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
  const int blocks = (n + 255) / 256;
  int rc = FEMRT_EUNSUP;
  FEMRT_DISPATCH_NG(m->NG, {
    rc = timed_launch(c, "k_invert", [&] {
      k_invert<NG_><<<blocks, 256, 0, c->stream>>>(n, d_e, d_x, m->d_geom, d_r, d_it);
    });
  });
  if (rc == FEMRT_OK) {
    HIPCHK(c, hipMemcpyAsync(rst_out, d_r, (size_t)n * 24, hipMemcpyDeviceToHost, c->stream));
    if (iters_out) HIPCHK(c, hipMemcpyAsync(iters_out, d_it, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  (void)hipFree(d_e);
  (void)hipFree(d_x);
  (void)hipFree(d_r);
  (void)hipFree(d_it);
  return rc;
}

FEMRT_API void femrt_trace_params_default(femrt_trace_params* p)
{
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->n_iso = 0;
  p->iso_enable = 1;
  p->samples = 16;
  p->guess = FEMRT_GUESS_PLAIN;
  p->t_min = 0.001f;
  p->t_max = 1e30f;
  p->max_march = 4096;
  p->flags = 1;
}

FEMRT_API int femrt_frame_create(femrt_ctx* c, int n, femrt_frame** out)
{
  if (!c || !out || n <= 0) return fail(c, FEMRT_EINVAL, "frame_create: bad argument");
  *out = nullptr;
  HIPCHK(c, hipSetDevice(c->device));
  femrt_frame* f = new femrt_frame();
  f->ctx = c;
  f->n = n;
  const size_t N = n, K = FEMRT_MAX_ISO;
  bool ok = hipMalloc(&f->d_rays, 6 * N * 4) == hipSuccess &&
            hipMalloc(&f->d_hit_elem, N * 4) == hipSuccess &&
            hipMalloc(&f->d_hit_face, N * 4) == hipSuccess &&
            hipMalloc(&f->d_hit_f32, N * 16) == hipSuccess &&
            hipMalloc(&f->d_hit_t, N * 8) == hipSuccess &&
            hipMalloc(&f->d_hit_rst, 3 * N * 8) == hipSuccess &&
            hipMalloc(&f->d_hit_u, N * 8) == hipSuccess &&
            hipMalloc(&f->d_iso_elem, K * N * 4) == hipSuccess &&
            hipMalloc(&f->d_iso_t, K * N * 8) == hipSuccess &&
            hipMalloc(&f->d_iso_rst, 3 * K * N * 8) == hipSuccess &&
            hipMalloc(&f->d_iso_u, K * N * 8) == hipSuccess &&
            hipMalloc(&f->d_iso_f32, K * N * 16) == hipSuccess &&
            hipMalloc(&f->d_iso_rst32, K * N * 16) == hipSuccess &&
            hipMalloc(&f->d_queue_hit, N * 4) == hipSuccess &&
            hipMalloc(&f->d_retrace, N * 4) == hipSuccess &&
            hipMalloc(&f->d_queue_walk, N * 4) == hipSuccess &&
            hipMalloc(&f->d_qbin, N) == hipSuccess &&
            hipMalloc(&f->d_counts, FEMRT_NCOUNTS * 4) == hipSuccess &&
            hipMalloc(&f->d_rgba, N * 4) == hipSuccess &&
            hipMalloc(&f->d_work, FEMRT_WORK_SLOTS * 8) == hipSuccess;
  if (!ok) {
    femrt_frame_destroy(f);
    return fail(c, FEMRT_ENOMEM, "frame_create: hipMalloc failed");
  }
  if (hipMemset(f->d_counts, 0, FEMRT_NCOUNTS * 4) != hipSuccess ||
      hipMemset(f->d_work, 0, FEMRT_WORK_SLOTS * 8) != hipSuccess ||
      hipMemset(f->d_hit_elem, 0xff, N * 4) != hipSuccess ||
      hipMemset(f->d_iso_elem, 0xff, K * N * 4) != hipSuccess) {
    femrt_frame_destroy(f);
    return fail(c, FEMRT_EHIP, "frame_create: hipMemset failed");
  }
  *out = f;
  return FEMRT_OK;
}

FEMRT_API int femrt_frame_destroy(femrt_frame* f)
{
  if (!f) return FEMRT_EINVAL;
  (void)hipSetDevice(f->ctx->device);
  (void)hipStreamSynchronize(f->ctx->stream);
  void* ptrs[] = {f->d_rays, f->d_hit_elem, f->d_hit_face, f->d_hit_f32, f->d_hit_t, f->d_hit_rst,
                  f->d_hit_u, f->d_iso_elem, f->d_iso_t, f->d_iso_rst, f->d_iso_u, f->d_iso_f32,
                  f->d_iso_rst32, f->d_queue_hit, f->d_retrace, f->d_queue_walk, f->d_qbin, f->d_counts, f->d_rgba, f->d_work,
                  f->d_path_tp, f->d_path_acc, f->d_mats, f->d_mlp_g, f->d_walk_si, f->d_walk_sf, f->d_walk_a, f->d_walk_b,
                  f->d_walk_sorted, f->d_walk_reent, f->d_walk_hist, f->d_iso_seg, f->d_iso_pred};
  for (void* p : ptrs) (void)hipFree(p);
  delete f;
  return FEMRT_OK;
}

FEMRT_API int femrt_frame_set_rays(femrt_frame* f, const float* rays)
{
  if (!f || !rays) return FEMRT_EINVAL;
  femrt_ctx* c = f->ctx;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(f->d_rays, rays, 6 * (size_t)f->n * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  f->img_w = f->img_h = 0;
  f->img_spp = 1;
  return FEMRT_OK;
}

FEMRT_API int femrt_frame_get_rays(femrt_frame* f, float* rays)
{
  if (!f || !rays) return FEMRT_EINVAL;
  femrt_ctx* c = f->ctx;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(rays, f->d_rays, 6 * (size_t)f->n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FEMRT_OK;
}

FEMRT_API int femrt_frame_camera_rays(femrt_frame* f, const float* cam, int W, int H)
{
  if (!f || !cam || W <= 0 || H <= 0 || (long long)W * H != f->n)
    return fail(f ? f->ctx : nullptr, FEMRT_EINVAL, "camera_rays: W*H must equal the frame size");
  femrt_ctx* c = f->ctx;
  // Camera basis (raytracer.h:31-48) on the host in double, rounded once.
  const double lf[3] = {cam[0], cam[1], cam[2]}, la[3] = {cam[3], cam[4], cam[5]},
               vu[3] = {cam[6], cam[7], cam[8]};
  const double th = cam[9] * M_PI / 180.0, hh = std::tan(th / 2.0);
  const double vh = 2.0 * hh, vw = ((double)W / H) * vh;
  double w[3] = {lf[0] - la[0], lf[1] - la[1], lf[2] - la[2]};
  const double fd = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  for (double& x : w) x /= fd;
  double u[3] = {vu[1] * w[2] - vu[2] * w[1], vu[2] * w[0] - vu[0] * w[2], vu[0] * w[1] - vu[1] * w[0]};
  const double un = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  for (double& x : u) x /= un;
  const double v[3] = {w[1] * u[2] - w[2] * u[1], w[2] * u[0] - w[0] * u[2], w[0] * u[1] - w[1] * u[0]};
  float hor[3], ver[3], llc[3];
  for (int k = 0; k < 3; ++k) {
    hor[k] = (float)(fd * vw * u[k]);
    ver[k] = (float)(fd * vh * v[k]);
    llc[k] = (float)(lf[k] - fd * vw * u[k] / 2 - fd * vh * v[k] / 2 - fd * w[k]);
  }
  HIPCHK(c, hipSetDevice(c->device));
  const int n = W * H;
  f->img_w = W;
  f->img_h = H;
  f->img_spp = 1;
  return timed_launch(c, "k_raygen", [&] {
    k_raygen<<<(n + 255) / 256, 256, 0, c->stream>>>(f->d_rays, W, H, cam[0], cam[1], cam[2], llc[0],
                                                    llc[1], llc[2], hor[0], hor[1], hor[2], ver[0],
                                                    ver[1], ver[2]);
  });
}

FEMRT_API int femrt_frame_counters(femrt_frame* f, int enable, long long* out)
{
  if (!f) return FEMRT_EINVAL;
  femrt_ctx* c = f->ctx;
  HIPCHK(c, hipSetDevice(c->device));
  if (out) {
    HIPCHK(c, hipMemcpyAsync(out, f->d_work, FEMRT_NCOUNTERS * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  f->counters = enable != 0;
  return FEMRT_OK;
}

FEMRT_API int femrt_frame_stalled_rays(femrt_frame* f, int cap, int* ids)
{
  if (!f || cap < 0 || (cap > 0 && !ids)) return FEMRT_EINVAL;
  femrt_ctx* c = f->ctx;
  HIPCHK(c, hipSetDevice(c->device));
  long long buf[1 + FEMRT_MAX_STALLED_IDS];
  HIPCHK(c, hipMemcpyAsync(buf, f->d_work + FEMRT_NCOUNTERS, sizeof(buf), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int total = (int)buf[0];
  const int kept = std::min(total, FEMRT_MAX_STALLED_IDS);
  for (int i = 0; i < std::min(kept, cap); ++i) ids[i] = (int)buf[1 + i];
  return total;
}

FEMRT_API int femrt_trace(femrt_mesh* m, femrt_frame* f, const femrt_trace_params* p)
{
  if (!m || !f || !p) return FEMRT_EINVAL;
  femrt_ctx* c = m->ctx;
  if (f->ctx != c) return fail(c, FEMRT_EINVAL, "trace: mesh and frame belong to different contexts");
  if (m->n_nodes == 0) return fail(c, FEMRT_ESTATE, "trace: BVH not built");
  if (p->n_iso < 0 || p->n_iso > FEMRT_MAX_ISO || p->samples < 1 || p->samples > 1024)
    return fail(c, FEMRT_EINVAL, "trace: n_iso must be 0..8 and samples 1..1024");
  if (p->guess < 0 || p->guess > FEMRT_GUESS_LEARNED)
    return fail(c, FEMRT_EINVAL, "trace: unknown guess mode");
  const bool learned = p->guess == FEMRT_GUESS_LEARNED;
  if (learned && !m->guess_mlp)
    return fail(c, FEMRT_ESTATE, "trace: learned guesses need a network (femrt_mesh_set_guess_mlp)");
  if (learned && p->samples > FEMRT_MLP_MAX_SAMPLES)
    return fail(c, FEMRT_EUNSUP, "trace: learned guesses support at most 16 samples per segment");
  if (c->capturing && (f->counters || (p->flags & FEMRT_TRACE_BINNED_WALK)))
    return fail(c, FEMRT_EUNSUP, "trace: work counters and the binned walk cannot be captured in a batch");
  if (m->surface && p->iso_enable && p->n_iso > 0)
    return fail(c, FEMRT_EINVAL, "trace: a surface mesh (curved quads) has no isosurfaces (set iso_enable = 0)");
  if (m->prism && p->iso_enable && p->n_iso > 0)
    return fail(c, FEMRT_EINVAL, "trace: a prism mesh is traced for first hits only (set iso_enable = 0)");
  const bool learned_br = (p->flags & FEMRT_TRACE_ISO_LEARNED_BRACKET) != 0;
  const bool want_seg = learned_br || (p->flags & FEMRT_TRACE_SEG_INFO);
  if (learned_br && !m->bracket_mlp)
    return fail(c, FEMRT_ESTATE, "trace: the learned bracket needs a network (femrt_mesh_set_bracket_mlp)");
  if (want_seg && (p->flags & FEMRT_TRACE_BINNED_WALK))
    return fail(c, FEMRT_EUNSUP, "trace: the binned walk records no bracket segments");
  if (want_seg && !f->d_iso_seg) {
    if (c->capturing)
      return fail(c, FEMRT_ESTATE, "trace: run a trace that records segments once before a batch records");
    const size_t slots = (size_t)FEMRT_MAX_ISO * f->n;
    if (hipMalloc(&f->d_iso_seg, slots * 2 * sizeof(float4)) != hipSuccess ||
        hipMalloc(&f->d_iso_pred, slots * sizeof(float2)) != hipSuccess) {
      (void)hipFree(f->d_iso_seg);
      f->d_iso_seg = nullptr;
      f->d_iso_pred = nullptr;
      return fail(c, FEMRT_ENOMEM, "trace: hipMalloc of the bracket segments failed");
    }
  }
  f->seg_valid = false;
  if (!c->capturing) HIPCHK(c, hipSetDevice(c->device));
  const int n = f->n;
  const int n_iso = p->iso_enable ? p->n_iso : 0;
  long long* work = f->counters ? f->d_work : nullptr;
  if (work) {
    HIPCHK(c, hipMemsetAsync(f->d_work, 0, FEMRT_WORK_SLOTS * 8, c->stream));
    HIPCHK(c, hipMemsetAsync(f->d_work + 16, 0xff, 2 * 8, c->stream));  // minima start at the top
  }
  const int blocks = (n + 255) / 256;
  int rc = timed_launch(c, "k_clear_frame", [&] {
    k_clear_frame<<<blocks, 256, 0, c->stream>>>(n, FEMRT_MAX_ISO, f->d_hit_elem, f->d_iso_elem, f->d_counts);
  });
  if (rc) return rc;
  const int dev_cus = c->cus;
  const LaunchCfg& cfg = m->cfg;
  TileOrder order{0, 0, 1, 0, 0};
  if (f->img_w > 0 && (long long)f->img_w * f->img_h * f->img_spp == n) {
    const int sh = cfg.tile == 4 ? 2 : (cfg.tile == 16 ? 4 : 3);
    order = TileOrder{f->img_w, f->img_h, f->img_spp, sh, (f->img_w + (1 << sh) - 1) >> sh};
  }
  const int pblocks = dev_cus * cfg.polish_bpc;
  const int hblock = cfg.hit_block;
  IsoParams ip;
  std::memset(&ip, 0, sizeof(ip));
  for (int k = 0; k < FEMRT_MAX_ISO; ++k) {
    ip.iso[k] = (float)p->iso[k];
    ip.iso64[k] = p->iso[k];
  }
  ip.n_iso = n_iso;
  ip.samples = p->samples;
  ip.guess = p->guess;
  ip.max_march = std::min(p->max_march, 0xFFFF);
  ip.mlp_lead = (p->flags & FEMRT_TRACE_LEARNED_EVERY_SAMPLE) ? p->samples - 1 : 1;
  ip.tmin = p->t_min;
  ip.skip = (p->flags & FEMRT_TRACE_NO_SAMPLE_SKIP) ? 0 : 1;
  ip.endgame = learned_br ? ISO_LEARNED_BISECT : ((p->flags & FEMRT_TRACE_ISO_BISECT) ? ISO_BISECT : ISO_NEWTON);
  ip.qseg_per_xcd = cfg.qseg_per_xcd;
  ip.hermite = (p->flags & FEMRT_TRACE_NO_HERMITE_START) ? 0 : 1;
  ip.jump = (p->flags & FEMRT_TRACE_NO_SKIP_JUMP) ? 0 : FEMRT_JUMP_AFTER;
  TraceMesh TM;
  TM.geom = m->d_geom;
  TM.field = m->d_field;
  TM.boxes = m->d_boxes;
  TM.verts = m->d_verts;
  TM.adj = m->d_adj;
  TM.nodes = m->d_nodes;
  TM.prims = m->d_prims;
  TM.lip = m->d_lip;
  TM.frange = m->d_frange;
  TM.jinv = m->d_jinv;
  TM.surface = m->surface ? 1 : 0;
  rc = FEMRT_EUNSUP;
  const bool flat = p->guess != FEMRT_GUESS_PLAIN && !(p->flags & FEMRT_TRACE_NESTED_FIRST_HIT);
  FEMRT_DISPATCH_NG(m->NG, {
    rc = timed_launch(c, "k_first_hit", [&] {
      const int grid = (order.threads(n) + hblock - 1) / hblock;
      if (flat)
        k_first_hit<NG_, true><<<grid, hblock, 0, c->stream>>>(f->d_rays, n, TM, p->guess, p->t_min, p->t_max,
                                                              order, f->d_hit_elem, f->d_hit_face, f->d_hit_f32,
                                                              f->d_queue_hit, f->d_counts, work);
      else
        k_first_hit<NG_, false><<<grid, hblock, 0, c->stream>>>(f->d_rays, n, TM, p->guess, p->t_min, p->t_max,
                                                               order, f->d_hit_elem, f->d_hit_face, f->d_hit_f32,
                                                               f->d_queue_hit, f->d_counts, work);
    });
    if (rc) break;
    FEMRT_DISPATCH_NF(m->NF, {
      rc = timed_launch(c, "k_polish_hit", [&] {
        k_polish_hit<NG_, NF_><<<pblocks, 256, 0, c->stream>>>(
            f->d_rays, n, TM, p->t_min, f->d_queue_hit, f->d_counts, f->d_hit_elem, f->d_hit_face,
            f->d_hit_f32, f->d_hit_t, f->d_hit_rst, f->d_hit_u, f->d_qbin, f->d_retrace, work);
      });
      if (rc) break;
      rc = timed_launch(c, "k_retrace", [&] {
        k_retrace<NG_, NF_><<<64, 64, 0, c->stream>>>(f->d_rays, n, TM, p->guess, p->t_min, p->t_max,
                                                      f->d_retrace, f->d_counts, f->d_hit_elem, f->d_hit_face,
                                                      f->d_hit_f32, f->d_hit_t, f->d_hit_rst, f->d_hit_u, work);
      });
      if (rc) break;
      if (n_iso > 0 && !(p->flags & FEMRT_TRACE_BINNED_WALK)) {
        rc = timed_launch(c, "k_walk_order", [&] {
          k_walk_order<<<ip.qseg_per_xcd * FEMRT_XCDS, FEMRT_ORDER_THREADS, 0, c->stream>>>(
              f->d_queue_hit, f->d_qbin, f->d_counts, f->d_queue_walk, ip.qseg_per_xcd * FEMRT_XCDS);
        });
        if (rc) break;
      }
      if (n_iso > 0) {
        if (p->flags & FEMRT_TRACE_BINNED_WALK) {
          rc = walk_run(c, m, f, TM, ip, learned, work);
        } else {
          // persistent: lanes refill from the hit queue until it drains
          const int mblocks = dev_cus * cfg.march_bpc;
          if (learned && ip.mlp_lead > 1) {
            // per-lane rows of every-sample network guesses
            const size_t need = (size_t)mblocks * 256 * (FEMRT_MLP_MAX_SAMPLES - 1) * 3;
            if (f->mlp_g_cap < need) {
              if (c->capturing) {
                rc = fail(c, FEMRT_ESTATE, "trace: run every-sample learned traces once before a batch records");
                break;
              }
              (void)hipFree(f->d_mlp_g);
              f->d_mlp_g = nullptr;
              f->mlp_g_cap = 0;
              if (hipMalloc(&f->d_mlp_g, need * sizeof(_Float16)) != hipSuccess) {
                rc = fail(c, FEMRT_ENOMEM, "trace: hipMalloc of the guess rows failed");
                break;
              }
              f->mlp_g_cap = need;
            }
          }
          // the cooperative scan on request, for up to 16 samples per
          // segment and guesses other than the network's (which start
          // samples that have no earlier neighbour in the sequential scan)
          const bool coop = p->samples <= FEMRT_COOP_MAX_SAMPLES && !learned &&
                            (p->flags & FEMRT_TRACE_COOP_SCAN);
          rc = timed_launch(c, "k_march", [&] {
            MarchArgs a;
            a.blocks = mblocks;
            a.stream = c->stream;
            a.nf = m->NF;
            a.learned = learned;
            a.count = work != nullptr;
            a.coop = coop;
            a.rays = f->d_rays;
            a.n = n;
            a.M = TM;
            a.P = ip;
            a.queue = f->d_queue_walk;
            a.counts = f->d_counts;
            a.hit_elem = f->d_hit_elem;
            a.hit_face = f->d_hit_face;
            a.hit_f32 = f->d_hit_f32;
            a.iso_elem = f->d_iso_elem;
            a.iso_f32 = f->d_iso_f32;
            a.iso_rst32 = f->d_iso_rst32;
            a.work = work;
            a.mlp_pack = learned ? m->guess_mlp->d_pack : nullptr;
            a.mlp_g = learned ? f->d_mlp_g : nullptr;
            a.iso_seg = want_seg ? f->d_iso_seg : nullptr;
            march_launch<NG_>(a);
          });
        }
        if (rc) break;
        if (learned_br) {
          rc = bracket_infer_launch(c, m, f, TM, ip, pblocks);
          if (rc) break;
        }
        rc = timed_launch(c, "k_polish_iso", [&] {
          k_polish_iso<NG_, NF_><<<pblocks, 256, 0, c->stream>>>(
              f->d_rays, n, TM, ip, f->d_queue_hit, f->d_counts, f->d_iso_f32, f->d_iso_rst32,
              f->d_iso_elem, f->d_iso_t, f->d_iso_rst, f->d_iso_u, work, f->d_iso_pred);
        });
        if (rc) break;
      }
    });
  });
  if (rc) return rc;
  f->seg_valid = want_seg && n_iso > 0;
  if (p->flags & 1)
    rc = timed_launch(c, "k_shade", [&] {
      k_shade<<<blocks, 256, 0, c->stream>>>(n, n_iso, f->d_rays, f->d_hit_elem, f->d_hit_u,
                                             f->d_iso_elem, f->d_rgba);
    });
  return rc;
}

FEMRT_API int femrt_frame_get_hits(femrt_frame* f, int* he, int* hf, double* ht, double* hr, double* hu)
{
  if (!f) return FEMRT_EINVAL;
  femrt_ctx* c = f->ctx;
  const size_t N = f->n;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<int> own;
  int* e = he;
  if (!e && (hf || ht || hr || hu)) {
    own.resize(N);
    e = own.data();
  }
  if (e) HIPCHK(c, hipMemcpy(e, f->d_hit_elem, N * 4, hipMemcpyDeviceToHost));
  if (hf) HIPCHK(c, hipMemcpy(hf, f->d_hit_face, N * 4, hipMemcpyDeviceToHost));
  if (ht) HIPCHK(c, hipMemcpy(ht, f->d_hit_t, N * 8, hipMemcpyDeviceToHost));
  if (hr) HIPCHK(c, hipMemcpy(hr, f->d_hit_rst, 3 * N * 8, hipMemcpyDeviceToHost));
  if (hu) HIPCHK(c, hipMemcpy(hu, f->d_hit_u, N * 8, hipMemcpyDeviceToHost));
  // rays without a hit report face -1 and NaN (the device buffers keep
  // whatever an earlier trace left there)
  for (size_t r = 0; e && r < N; ++r) {
    if (e[r] >= 0) continue;
    if (hf) hf[r] = -1;
    if (ht) ht[r] = NAN;
    if (hu) hu[r] = NAN;
    if (hr) hr[r] = hr[N + r] = hr[2 * N + r] = NAN;
  }
  return FEMRT_OK;
}

FEMRT_API int femrt_frame_get_iso(femrt_frame* f, int n_iso, int* ie, double* it, double* ir, double* iu)
{
  if (!f || n_iso < 0 || n_iso > FEMRT_MAX_ISO) return FEMRT_EINVAL;
  femrt_ctx* c = f->ctx;
  const size_t N = f->n, K = n_iso;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (!K) return FEMRT_OK;
  std::vector<int> own;
  int* e = ie;
  if (!e && (it || ir || iu)) {
    own.resize(K * N);
    e = own.data();
  }
  if (e) HIPCHK(c, hipMemcpy(e, f->d_iso_elem, K * N * 4, hipMemcpyDeviceToHost));
  if (it) HIPCHK(c, hipMemcpy(it, f->d_iso_t, K * N * 8, hipMemcpyDeviceToHost));
  if (ir) HIPCHK(c, hipMemcpy(ir, f->d_iso_rst, 3 * K * N * 8, hipMemcpyDeviceToHost));
  if (iu) HIPCHK(c, hipMemcpy(iu, f->d_iso_u, K * N * 8, hipMemcpyDeviceToHost));
  // (ray, isovalue) pairs without a crossing report NaN
  for (size_t k = 0; e && k < K; ++k)
    for (size_t r = 0; r < N; ++r) {
      if (e[k * N + r] >= 0) continue;
      if (it) it[k * N + r] = NAN;
      if (iu) iu[k * N + r] = NAN;
      if (ir) ir[(3 * k) * N + r] = ir[(3 * k + 1) * N + r] = ir[(3 * k + 2) * N + r] = NAN;
    }
  return FEMRT_OK;
}

FEMRT_API int femrt_frame_get_rgba(femrt_frame* f, uint32_t* rgba)
{
  if (!f || !rgba) return FEMRT_EINVAL;
  femrt_ctx* c = f->ctx;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpy(rgba, f->d_rgba, (size_t)f->n * 4, hipMemcpyDeviceToHost));
  return FEMRT_OK;
}

FEMRT_API int femrt_host_boxes(int n_elem, int ng, const float* geom, float* boxes)
{
  if (n_elem <= 0 || ng < 1 || ng > 16 || !geom || !boxes) return FEMRT_EINVAL;
  bernstein_boxes(n_elem, ng, geom, boxes);
  return FEMRT_OK;
}

FEMRT_API int femrt_host_field_lipschitz(int n_elem, int nf, const float* field, float* out)
{
  if (n_elem < 0 || nf < 1 || nf > 16 || !field || !out) return FEMRT_EINVAL;
  field_lipschitz(n_elem, nf, field, out);
  return FEMRT_OK;
}

FEMRT_API int femrt_host_geom_jinv_bound(int n_elem, int ng, const float* geom, float* out)
{
  if (n_elem < 0 || ng < 1 || ng > 16 || !geom || !out) return FEMRT_EINVAL;
  geom_jinv_bound(n_elem, ng, geom, out);
  return FEMRT_OK;
}

FEMRT_API int femrt_host_field_range(int n_elem, int nf, const float* field, float* out)
{
  if (n_elem <= 0 || nf < 1 || nf > 16 || !field || !out) return FEMRT_EINVAL;
  field_range(n_elem, nf, field, out);
  return FEMRT_OK;
}

FEMRT_API int femrt_host_build_bvh(int n_elem, const float* boxes, int* n_nodes, float* nodes8,
                                   int* prims)
{
  if (n_elem <= 0 || !boxes || !n_nodes || !nodes8 || !prims) return FEMRT_EINVAL;
  BvhHost b;
  build_bvh_sah(n_elem, boxes, b);
  const size_t nn = b.nodes.size() / 8;
  if (nn > (size_t)2 * n_elem) return FEMRT_EINVAL;
  std::memcpy(nodes8, b.nodes.data(), b.nodes.size() * 4);
  std::memcpy(prims, b.prims.data(), b.prims.size() * 4);
  *n_nodes = (int)nn;
  return FEMRT_OK;
}

FEMRT_API int femrt_gen_structured_hex(const femrt_gen_params* gp, float* vertices, float* geom,
                                       float* field, int* adjacency)
{
  if (!gp) return FEMRT_EINVAL;
  return gen_structured_hex(gp, vertices, geom, field, adjacency);
}

FEMRT_API int femrt_gen_field_waves(int n_elem, int ng, const float* geom, int nf, int n_waves,
                                    uint64_t seed, float* field)
{
  return gen_field_waves(n_elem, ng, geom, nf, n_waves, seed, field);
}

static thread_local std::string g_io_err;

FEMRT_API const char* femrt_io_last_error(void) { return g_io_err.c_str(); }

FEMRT_API int femrt_load_xml_fld(const char* xml_path, const char* fld_path, int flags, int* n_elem,
                                 int* ng, int* nf, float* vertices, float* geom, float* field,
                                 int* adjacency)
{
  g_io_err.clear();
  if (!xml_path || !n_elem || !ng || !nf) {
    g_io_err = "load_xml_fld: null argument";
    return FEMRT_EINVAL;
  }
  return load_xml_fld(xml_path, fld_path, flags, n_elem, ng, nf, vertices, geom, field, adjacency,
                      &g_io_err);
}

FEMRT_API int femrt_save_xml_fld(const char* xml_path, const char* fld_path, int n_elem, int ng,
                                 int nf, const float* vertices, const float* geom,
                                 const float* field)
{
  g_io_err.clear();
  return save_xml_fld(xml_path, fld_path, n_elem, ng, nf, vertices, geom, field, &g_io_err);
}

}  // extern "C"
