// femrt_path.hip -- the shading and bounce stage: jittered camera paths,
// material scattering at the mesh surface, sky and emission, accumulation
// and gamma (raytracer.cu:21-98 scatter, 362-406 ray_color, 410-476
// render_kernel).  Each depth is one geometry trace of the frame's rays
// (femrt_trace with the iso stage off) followed by k_scatter.
//
// Random numbers come from a counter-based generator keyed by
// (seed, path, depth, draw) instead of per-pixel cuRAND states, so a path is
// reproducible independently of the launch shape.  Unit vectors are drawn
// directly (z = 1 - 2 u, phi = 2 pi v) instead of by rejection; normal +
// unit vector is the same cosine-weighted distribution as the reference's
// normal + normalize(random_in_unit_sphere).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "femrt_runtime.hpp"

using namespace femrt;

namespace {

__host__ __device__ inline unsigned long long path_mix(unsigned long long z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Uniform in [0, 1) with 24 random bits.
__host__ __device__ inline float path_rand(unsigned long long seed, unsigned long long path,
                                           unsigned depth, unsigned k)
{
  const unsigned long long h =
      path_mix(seed ^ path_mix(path * 0x9E3779B97F4A7C15ull + ((unsigned long long)depth << 8 | k)));
  return (float)(h >> 40) * (1.0f / 16777216.0f);
}

__device__ inline void unit_vec(float a, float b, float* v)
{
  const float z = 1.f - 2.f * a, phi = 6.28318530717958648f * b;
  const float r = sqrtf(fmaxf(0.f, 1.f - z * z));
  v[0] = r * cosf(phi);
  v[1] = r * sinf(phi);
  v[2] = z;
}

__device__ inline float dot3(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

__device__ inline void normalize3(float* a)
{
  const float l = sqrtf(dot3(a, a));
  if (l > 0.f)
    for (int c = 0; c < 3; ++c) a[c] /= l;
}

struct Mat {
  float color[3], emission[3], roughness, ior;
  int type;
};

}  // namespace

// Pixel (x, y), sample s -> path p = (y W + x) spp + s, ray through
// ((x + u) / W, (y + v) / H) with u, v the path's draws 0 and 1 at depth 0.
__global__ void k_raygen_jitter(float* __restrict__ rays, int W, int H, int spp, float4 o4, float4 ll4,
                                float4 h4, float4 v4, unsigned long long seed)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = W * H * spp;
  if (p >= n) return;
  const int pix = p / spp, y = pix / W, x = pix - y * W;
  const float u = (x + path_rand(seed, p, 0, 0)) / W, v = (y + path_rand(seed, p, 0, 1)) / H;
  float d[3] = {ll4.x + u * h4.x + v * v4.x - o4.x, ll4.y + u * h4.y + v * v4.y - o4.y,
                ll4.z + u * h4.z + v * v4.z - o4.z};
  normalize3(d);
  rays[p] = o4.x;
  rays[n + p] = o4.y;
  rays[2 * n + p] = o4.z;
  rays[3 * n + p] = d[0];
  rays[4 * n + p] = d[1];
  rays[5 * n + p] = d[2];
}

__global__ void k_path_begin(int n, float4* __restrict__ tp, float4* __restrict__ acc)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  tp[p] = make_float4(1.f, 1.f, 1.f, 1.f);
  acc[p] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// One bounce of every live path after a geometry trace.  Hit: surface normal
// from the Jacobian tangents of the hit face at the FP64-polished rst,
// material (face 0 -> material 1, others material 0, raytracer.cu:248),
// scatter; miss: sky (raytracer.cu:398-402).  A path that ends gets the ray
// origin (1e30, 1e30, 1e30), direction +x, which every box test rejects.
template <int NG>
__global__ void k_scatter(int n, float* __restrict__ rays, TraceMesh M, const int* __restrict__ hit_elem,
                          const int* __restrict__ hit_face, const double* __restrict__ hit_t,
                          const double* __restrict__ hit_rst, const Mat* __restrict__ mats, int n_mats,
                          unsigned long long seed, int depth, int last, float4* __restrict__ tp,
                          float4* __restrict__ acc)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float4 T = tp[p];
  if (T.w == 0.f) return;
  auto kill = [&] {
    for (int c = 0; c < 3; ++c) rays[(size_t)c * n + p] = 1e30f;
    rays[(size_t)3 * n + p] = 1.f;
    rays[(size_t)4 * n + p] = 0.f;
    rays[(size_t)5 * n + p] = 0.f;
  };
  float o[3], d[3];
  for (int c = 0; c < 3; ++c) {
    o[c] = rays[(size_t)c * n + p];
    d[c] = rays[(size_t)(3 + c) * n + p];
  }
  float4 A = acc[p];
  const int e = hit_elem[p];
  if (e < 0) {
    const float t = 0.5f * (d[1] + 1.f);
    A.x += T.x * ((1.f - t) + t * 0.5f);
    A.y += T.y * ((1.f - t) + t * 0.7f);
    A.z += T.z * 1.f;
    acc[p] = A;
    tp[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    kill();
    return;
  }
  const int f = hit_face[p];
  const double t = hit_t[p];
  float rst[3];
  for (int c = 0; c < 3; ++c) rst[c] = (float)hit_rst[(size_t)c * n + p];
  float X[3], J[9];
  geom_eval<float, NG>(M.geom + (size_t)e * (NG * NG * NG), rst[0], rst[1], rst[2], X, J);
  // tangents along the face's two free coordinates; J columns are d/dr,
  // d/ds, d/dt.  n_axis points towards increasing axis coordinate for a
  // positively oriented element; outward = +-n_axis by the face side.
  const int fix = face_fix(f);
  const int a = (fix + 1) % 3, b = (fix + 2) % 3;
  float ta[3] = {J[a], J[3 + a], J[6 + a]}, tb[3] = {J[b], J[3 + b], J[6 + b]};
  float nrm[3] = {ta[1] * tb[2] - ta[2] * tb[1], ta[2] * tb[0] - ta[0] * tb[2], ta[0] * tb[1] - ta[1] * tb[0]};
  normalize3(nrm);
  if (face_val(f) < 0.f)
    for (int c = 0; c < 3; ++c) nrm[c] = -nrm[c];
  const bool front = dot3(d, nrm) < 0.f;
  if (!front)
    for (int c = 0; c < 3; ++c) nrm[c] = -nrm[c];
  float po[3];
  for (int c = 0; c < 3; ++c) po[c] = (float)((double)o[c] + t * (double)d[c]);
  const int mi = n_mats > 1 ? (f == 0 ? 1 : 0) : 0;
  Mat m = mats[mi < n_mats ? mi : 0];
  float att[3] = {m.color[0], m.color[1], m.color[2]};
  float sd[3];
  bool scattered = true;
  if (m.type == 1) {  // metal
    float rv[3], us[3];
    const float dn = dot3(d, nrm);
    for (int c = 0; c < 3; ++c) rv[c] = d[c] - 2.f * dn * nrm[c];
    unit_vec(path_rand(seed, p, depth + 1, 0), path_rand(seed, p, depth + 1, 1), us);
    const float rad = cbrtf(path_rand(seed, p, depth + 1, 2));
    for (int c = 0; c < 3; ++c) sd[c] = rv[c] + m.roughness * rad * us[c];
    scattered = dot3(sd, nrm) > 0.f;
  } else if (m.type == 2) {  // dielectric
    att[0] = att[1] = att[2] = 1.f;
    const float eta = front ? 1.f / m.ior : m.ior;
    const float cosv = fminf(-dot3(d, nrm), 1.f);
    const float sinv = sqrtf(fmaxf(0.f, 1.f - cosv * cosv));
    float r0 = (1.f - eta) / (1.f + eta);
    r0 *= r0;
    const float sch = r0 + (1.f - r0) * powf(1.f - cosv, 5.f);
    if (eta * sinv > 1.f || sch > path_rand(seed, p, depth + 1, 0)) {
      for (int c = 0; c < 3; ++c) sd[c] = d[c] + 2.f * cosv * nrm[c];
    } else {
      const float k = 1.f - eta * eta * (1.f - cosv * cosv);
      for (int c = 0; c < 3; ++c) sd[c] = eta * (d[c] + cosv * nrm[c]) - sqrtf(fmaxf(k, 0.f)) * nrm[c];
    }
  } else {  // Lambertian: normal + uniform unit vector
    float us[3];
    unit_vec(path_rand(seed, p, depth + 1, 0), path_rand(seed, p, depth + 1, 1), us);
    for (int c = 0; c < 3; ++c) sd[c] = nrm[c] + us[c];
    if (dot3(sd, sd) < 1e-12f)
      for (int c = 0; c < 3; ++c) sd[c] = nrm[c];
  }
  const bool emits = m.emission[0] != 0.f || m.emission[1] != 0.f || m.emission[2] != 0.f;
  if (!scattered) {
    A.x += T.x * m.emission[0];
    A.y += T.y * m.emission[1];
    A.z += T.z * m.emission[2];
    acc[p] = A;
    tp[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    kill();
    return;
  }
  T.x *= att[0];
  T.y *= att[1];
  T.z *= att[2];
  if (emits || last) {
    // an emitter ends the path with its emission; the depth cap with black
    if (emits) {
      A.x += T.x * m.emission[0];
      A.y += T.y * m.emission[1];
      A.z += T.z * m.emission[2];
      acc[p] = A;
    }
    tp[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    kill();
    return;
  }
  normalize3(sd);
  tp[p] = T;
  for (int c = 0; c < 3; ++c) {
    rays[(size_t)c * n + p] = po[c];
    rays[(size_t)(3 + c) * n + p] = sd[c];
  }
}

// Mean over the pixel's samples, gamma sqrt, clamp, 8 bits
// (raytracer.cu:460-475).
__global__ void k_resolve(int W, int H, int spp, const float4* __restrict__ acc, uint32_t* __restrict__ rgba)
{
  const int px = blockIdx.x * blockDim.x + threadIdx.x;
  if (px >= W * H) return;
  float c[3] = {0.f, 0.f, 0.f};
  for (int s = 0; s < spp; ++s) {
    const float4 a = acc[(size_t)px * spp + s];
    c[0] += a.x;
    c[1] += a.y;
    c[2] += a.z;
  }
  uint32_t out = 0xFF000000u;
  for (int k = 0; k < 3; ++k) {
    const float g = fminf(1.f, sqrtf(fmaxf(c[k] / spp, 0.f)));
    out |= (uint32_t)(255.99f * g) << (8 * k);
  }
  rgba[px] = out;
}

extern "C" {

FEMRT_API int femrt_frame_camera_rays_jitter(femrt_frame* f, const float* cam12, int W, int H, int spp,
                                             unsigned long long seed)
{
  if (!f || !cam12 || W <= 0 || H <= 0 || spp <= 0) return FEMRT_EINVAL;
  femrt_ctx* c = f->ctx;
  if ((long long)W * H * spp != f->n)
    return fail(c, FEMRT_EINVAL, "camera_rays_jitter: frame must hold width*height*spp rays");
  f->img_w = W;
  f->img_h = H;
  f->img_spp = spp;
  const float4 o = make_float4(cam12[0], cam12[1], cam12[2], 0.f);
  const float4 ll = make_float4(cam12[3], cam12[4], cam12[5], 0.f);
  const float4 h = make_float4(cam12[6], cam12[7], cam12[8], 0.f);
  const float4 v = make_float4(cam12[9], cam12[10], cam12[11], 0.f);
  return timed_launch(c, "k_raygen_jitter", [&] {
    k_raygen_jitter<<<(f->n + 255) / 256, 256, 0, c->stream>>>(f->d_rays, W, H, spp, o, ll, h, v, seed);
  });
}

FEMRT_API int femrt_frame_path_begin(femrt_frame* f)
{
  if (!f) return FEMRT_EINVAL;
  femrt_ctx* c = f->ctx;
  if (!f->d_path_tp) {
    if (c->capturing) return fail(c, FEMRT_ESTATE, "path_begin: path buffers must exist before a batch records");
    HIPCHK(c, hipSetDevice(c->device));
    if (hipMalloc(&f->d_path_tp, (size_t)f->n * 16) != hipSuccess ||
        hipMalloc(&f->d_path_acc, (size_t)f->n * 16) != hipSuccess)
      return fail(c, FEMRT_ENOMEM, "path_begin: hipMalloc failed");
  }
  return timed_launch(c, "k_path_begin", [&] {
    k_path_begin<<<(f->n + 255) / 256, 256, 0, c->stream>>>(f->n, f->d_path_tp, f->d_path_acc);
  });
}

// The material table lives in the frame and is uploaded when it changes
// (never while a batch records the stream).
static int set_materials(femrt_ctx* c, femrt_frame* f, const femrt_material* mats, int n_mats)
{
  std::vector<Mat> hm(n_mats);
  for (int i = 0; i < n_mats; ++i) {
    std::memset(&hm[i], 0, sizeof(Mat));
    std::memcpy(hm[i].color, mats[i].color, 12);
    std::memcpy(hm[i].emission, mats[i].emission, 12);
    hm[i].roughness = mats[i].roughness;
    hm[i].ior = mats[i].ior;
    hm[i].type = mats[i].type;
    if (hm[i].type < 0 || hm[i].type > 2) return fail(c, FEMRT_EINVAL, "materials: type must be 0..2");
  }
  const unsigned char* bytes = reinterpret_cast<const unsigned char*>(hm.data());
  const size_t nb = n_mats * sizeof(Mat);
  if (f->h_mats.size() == nb && std::memcmp(f->h_mats.data(), bytes, nb) == 0) return FEMRT_OK;
  if (c->capturing) return fail(c, FEMRT_ESTATE, "materials: set them before a batch is captured");
  HIPCHK(c, hipSetDevice(c->device));
  if (!f->d_mats && hipMalloc(&f->d_mats, 64 * sizeof(Mat)) != hipSuccess)
    return fail(c, FEMRT_ENOMEM, "materials: hipMalloc failed");
  HIPCHK(c, hipMemcpyAsync(f->d_mats, bytes, nb, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  f->h_mats.assign(bytes, bytes + nb);
  return FEMRT_OK;
}

FEMRT_API int femrt_frame_set_materials(femrt_frame* f, const femrt_material* mats, int n_mats)
{
  if (!f || !mats || n_mats < 1 || n_mats > 64) return FEMRT_EINVAL;
  return set_materials(f->ctx, f, mats, n_mats);
}

FEMRT_API int femrt_frame_scatter(femrt_mesh* m, femrt_frame* f, const femrt_material* mats, int n_mats,
                                  unsigned long long seed, int depth, int last)
{
  if (!m || !f || !mats || n_mats < 1 || n_mats > 64 || depth < 0) return FEMRT_EINVAL;
  femrt_ctx* c = f->ctx;
  if (!f->d_path_tp) return fail(c, FEMRT_ESTATE, "scatter: call femrt_frame_path_begin first");
  const int mrc = set_materials(c, f, mats, n_mats);
  if (mrc) return mrc;
  const Mat* dm = static_cast<const Mat*>(f->d_mats);
  TraceMesh TM{m->d_geom, m->d_field, m->d_boxes, m->d_verts, m->d_adj, m->d_nodes, m->d_prims};
  int rc = FEMRT_EUNSUP;
  auto go = [&](auto tag) {
    constexpr int NG_ = decltype(tag)::value;
    rc = timed_launch(c, "k_scatter", [&] {
      k_scatter<NG_><<<(f->n + 255) / 256, 256, 0, c->stream>>>(
          f->n, f->d_rays, TM, f->d_hit_elem, f->d_hit_face, f->d_hit_t, f->d_hit_rst, dm, n_mats, seed,
          depth, last, f->d_path_tp, f->d_path_acc);
    });
  };
  switch (m->NG) {
    case 4: go(std::integral_constant<int, 4>()); break;
    case 5: go(std::integral_constant<int, 5>()); break;
    case 7: go(std::integral_constant<int, 7>()); break;
  }
  return rc;
}

FEMRT_API int femrt_frame_resolve(femrt_frame* f, int W, int H, int spp, uint32_t* rgba)
{
  if (!f || W <= 0 || H <= 0 || spp <= 0) return FEMRT_EINVAL;
  femrt_ctx* c = f->ctx;
  if ((long long)W * H * spp != f->n) return fail(c, FEMRT_EINVAL, "resolve: frame must hold width*height*spp paths");
  if (!f->d_path_acc) return fail(c, FEMRT_ESTATE, "resolve: no paths traced");
  if (rgba && c->capturing) return fail(c, FEMRT_EINVAL, "resolve: no host copy while a batch records");
  const int rc = timed_launch(c, "k_resolve", [&] {
    k_resolve<<<(W * H + 255) / 256, 256, 0, c->stream>>>(W, H, spp, f->d_path_acc, f->d_rgba);
  });
  if (rc || !rgba) return rc;
  HIPCHK(c, hipMemcpyAsync(rgba, f->d_rgba, (size_t)W * H * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FEMRT_OK;
}

FEMRT_API int femrt_render_path(femrt_mesh* m, femrt_frame* f, const float* cam12, int W, int H, int spp,
                                int max_depth, const femrt_material* mats, int n_mats,
                                unsigned long long seed, uint32_t* rgba)
{
  if (!m || !f || max_depth < 1) return FEMRT_EINVAL;
  int rc = femrt_frame_camera_rays_jitter(f, cam12, W, H, spp, seed);
  if (rc == FEMRT_OK) rc = femrt_frame_path_begin(f);
  femrt_trace_params tp;
  femrt_trace_params_default(&tp);
  tp.iso_enable = 0;
  tp.n_iso = 0;
  tp.flags = 0;
  tp.guess = FEMRT_GUESS_WARM;
  for (int depth = 0; depth < max_depth && rc == FEMRT_OK; ++depth) {
    rc = femrt_trace(m, f, &tp);
    if (rc == FEMRT_OK) rc = femrt_frame_scatter(m, f, mats, n_mats, seed, depth, depth == max_depth - 1);
  }
  if (rc == FEMRT_OK) rc = femrt_frame_resolve(f, W, H, spp, rgba);
  return rc;
}

}  // extern "C"
