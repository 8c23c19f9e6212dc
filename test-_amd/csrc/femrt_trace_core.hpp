// femrt_trace_core.hpp -- per-ray bodies of the trace kernels.
//
// Everything here is __host__ __device__ and free of wave-level intrinsics, so
// the kernels in femrt_kernels.hip are thin wrappers around these functions.
// Reference: raytracer.cu:102-254 (traversal + hex branch), high_order.h:
// 189-304, 409-465 (face Newton, nearest face), fem/HexElement.cu:227-272
// (point inversion), fem/RayIsosurface.cu:67-188 (iso search).
#pragma once

#include <float.h>
#include <math.h>
#include <string.h>

#include "../../include/femrt.h"
#include "femrt_math.hpp"

namespace femrt {

// Device-resident mesh as the trace kernels see it.
struct TraceMesh {
  const GeoMode* geom;  // n_elem x NG^3 (x, y, z)
  const float* field;   // n_elem x NF^3
  const float* boxes;   // n_elem x FEMRT_BOX_STRIDE: element, face, face sub-patch boxes
  const float* verts;   // n_elem x 24 corner vertices, XML order
  const int* adj;       // n_elem x 12 (neighbour, neighbour face) or null
  const float4* nodes;  // BVH nodes x 2
  const int* prims;
  const float* lip;     // n_elem: bound of |u(a) - u(b)| / |a - b|_inf, or null
  const float* frange;  // n_elem x 2: bounds of u on the element (widened), or null
  const float* jinv;    // n_elem: bound of |J^-1|_inf over the element (0: none), or null
  int surface;          // 1: a surface mesh (curved quads as face 0 of maps constant in t)
};

struct IsoParams {
  float iso[FEMRT_MAX_ISO];      // FP32 rounding, used by the scan
  double iso64[FEMRT_MAX_ISO];   // exact values, used by the FP64 endgame
  int n_iso;
  int samples;
  int guess;
  int max_march;
  float tmin;
  int mlp_lead;   // leading interior samples of a segment guessed by the network
  int skip;       // certified sample skipping on (see march_samples)
  int endgame;    // FP64 iso endgame: ISO_NEWTON, ISO_BISECT or ISO_LEARNED_BISECT
  int qseg_per_xcd;  // persistent walk: hit-queue segments per XCD (1, 2 or 4)
  int hermite;       // first interior sample from the Hermite interpolant where tangents exist
  int jump;          // after this many segments in a row that cannot cross, the walk jumps
                     // over the rest of such a run by a BVH query (march_advance); 0: never
                     // (FEMRT_TRACE_NO_SKIP_JUMP)
};

// Segments in a row whose field bounds rule out every open isovalue before
// the walk jumps over the rest of the run (a shorter run is walked: a jump's
// BVH query holds its wave's other lanes longer than a few exit searches)
#define FEMRT_JUMP_AFTER 2

// FP64 iso endgames (IsoParams::endgame)
enum { ISO_NEWTON = 0, ISO_BISECT = 1, ISO_LEARNED_BISECT = 2 };

// Work counters accumulated per lane.
// (32-bit per lane: a lane's counts stay far below 2^32; the kernels add them
// to 64-bit totals)
struct Work {
  unsigned face_it = 0, inv_it = 0, field = 0, geo64 = 0, field64 = 0, slab = 0, mlp = 0,
           marched = 0;
  // element-walk diagnostics: exits found by prediction / face search /
  // marching, re-entries through the BVH, redone segments, iso brackets
  unsigned exit_pred = 0, exit_search = 0, exit_march = 0, reentry = 0, redo = 0, brackets = 0;
  unsigned skipped = 0;  // segments whose samples the field bounds rule out
  unsigned samples_skipped = 0;  // scan samples whose signs the bounds certified
  unsigned stalled = 0;  // walks ended stalled (FEMRT_MAX_WALK_SKIPS stalls skipped already)
  unsigned skips = 0;    // stalls skipped past (runs of FEMRT_MAX_MARCHED_EXITS marched exits)
  unsigned jumps = 0;    // runs of elements that cannot cross passed by a BVH query
};

// The same counters compiled away: the kernels of timed runs take NoWork,
// so the counts cost neither registers nor instructions there.  The code
// that counts only ever increments them.
struct NoCount {
  FEMRT_HD NoCount& operator++() { return *this; }
  FEMRT_HD NoCount& operator+=(unsigned) { return *this; }
};
struct NoWork {
  NoCount face_it, inv_it, field, geo64, field64, slab, mlp, marched;
  NoCount exit_pred, exit_search, exit_march, reentry, redo, brackets, skipped, samples_skipped, stalled, skips,
      jumps;
};

// Whether element e can hold a crossing of an isovalue not yet found
// (`found` bits): some such isovalue lies within the element's bounds of u.
// Without bounds every element can.
FEMRT_HD bool elem_may_cross(const TraceMesh& M, const IsoParams& P, int e, unsigned found)
{
  if (!M.frange) return true;
  const float lo = M.frange[2 * (size_t)e], hi = M.frange[2 * (size_t)e + 1];
  for (int k = 0; k < FEMRT_MAX_ISO; ++k)
    if (k < P.n_iso && !((found >> k) & 1u) && P.iso[k] >= lo && P.iso[k] <= hi) return true;
  return false;
}

// FP32 result of the first-hit stage.
struct HitF32 {
  int elem = -1, face = -1;
  float u = 0.f, v = 0.f, t = 0.f;
};

// FP32 result of the iso stage for one isovalue: sample bracket [ta, tb],
// FP32 estimate tm, g at ta, rst estimate; and the bracket's ends as the
// bracket network reads them: seg = (rst at ta, rst at tb, g at ta, g at tb).
struct IsoBracket {
  int elem;
  float ta, tb, tm, ga;
  float rst[3];
  float seg[8];
};

#define FEMRT_STACK 64
// exits in a row a walk may find by the marching fallback (with no exit found
// by prediction or face search in between) before it counts as stalled, and
// the stalls it may skip past before it ends
#define FEMRT_MAX_MARCHED_EXITS 16
#define FEMRT_MAX_WALK_SKIPS 8
#define FEMRT_FACE_ACC 1e-4f
// Boxes per element: [0] element, [1..6] faces (Nektar order), [7 + 9 f +
// 3 b + a] sub-patch (a, b) of face f on the 3 x 3 split of its (u, v) square.
#ifndef FEMRT_NBOX
#define FEMRT_NBOX 61
#define FEMRT_BOX_STRIDE (FEMRT_NBOX * 6)
#endif

FEMRT_HD int as_int(float x)
{
  int i;
  memcpy(&i, &x, 4);
  return i;
}

FEMRT_HD void face_to_rst(int f, float u, float v, float& r, float& s, float& t)
{
  r = (f == 2) ? 1.f : ((f == 4) ? -1.f : u);
  s = (f == 0 || f == 5) ? v : ((f == 1) ? -1.f : ((f == 3) ? 1.f : u));
  t = (f == 0) ? -1.f : ((f == 5) ? 1.f : v);
}

template <typename T>
FEMRT_HD void face_to_rst_t(int f, T u, T v, T* rst)
{
  rst[0] = (f == 2) ? T(1) : ((f == 4) ? T(-1) : u);
  rst[1] = (f == 0 || f == 5) ? v : ((f == 1) ? T(-1) : ((f == 3) ? T(1) : u));
  rst[2] = (f == 0) ? T(-1) : ((f == 5) ? T(1) : v);
}

FEMRT_HD void rst_to_face(int f, const float* rst, float& u, float& v)
{
  u = (f == 2 || f == 4) ? rst[1] : rst[0];
  v = (f == 0 || f == 5) ? rst[1] : rst[2];
}

// Face-constrained Newton on (u, v, tau), FP32, with the reference's update
// rules (high_order.h:213-300): step capped to 0.25 in the inf-norm, tau
// clamped to [tmin, tmax], u, v clamped to [-1.5, 1.5], divergence break
// after 5 iterations when the residual grows 1.5x, at most 50 iterations.
// FaceIter is the iteration's state; face_newton_step is one iteration with
// x(rst) and J already evaluated at the current (u, v).
struct FaceIter {
  int f;
  float u, v, tau, prev, tmin, tmax;
  int it;
};

enum { FN_CONTINUE = 0, FN_ACCEPT = 1, FN_REJECT = 2, FN_FAIL = 3 };

FEMRT_HD void face_iter_begin(FaceIter& F, int f, float u, float v, float tau, float tmin, float tmax)
{
  F.f = f;
  F.u = u;
  F.v = v;
  F.tmin = tmin;
  F.tmax = tmax;
  F.tau = fminf(fmaxf(tau, tmin), tmax);
  F.prev = FLT_MAX;
  F.it = 0;
}

FEMRT_HD int face_newton_step(FaceIter& F, const float* o, const float* d, const float* X,
                              const float* J)
{
  const float oscale = 1.f + fmaxf(fabsf(o[0]), fmaxf(fabsf(o[1]), fabsf(o[2])));
  const bool us = (F.f == 2 || F.f == 4);
  const bool vs = (F.f == 0 || F.f == 5);
  const float R0 = X[0] - fmaf(F.tau, d[0], o[0]);
  const float R1 = X[1] - fmaf(F.tau, d[1], o[1]);
  const float R2 = X[2] - fmaf(F.tau, d[2], o[2]);
  const float res = fmaxf(fabsf(R0), fmaxf(fabsf(R1), fabsf(R2)));
  const float tol = 2e-6f * (oscale + fabsf(F.tau));
  const float A[9] = {us ? J[1] : J[0], vs ? J[1] : J[2], -d[0],
                      us ? J[4] : J[3], vs ? J[4] : J[5], -d[1],
                      us ? J[7] : J[6], vs ? J[7] : J[8], -d[2]};
  const float b[3] = {-R0, -R1, -R2};
  float dl[3];
  const bool solved = solve3<float>(A, b, dl);
  if (res < tol) {
    // converged: one more (free) Newton step with the Jacobian at hand takes
    // the root to the FP32 floor
    if (solved && fmaxf(fabsf(dl[0]), fabsf(dl[1])) < 0.01f) {
      F.u += dl[0];
      F.v += dl[1];
      F.tau += dl[2];
    }
    return (fmaxf(fabsf(F.u), fabsf(F.v)) <= 1.f + FEMRT_FACE_ACC && F.tau >= F.tmin && F.tau <= F.tmax)
               ? FN_ACCEPT
               : FN_REJECT;
  }
  if (F.it > 5 && res > F.prev * 1.5f) return FN_FAIL;
  F.prev = res;
  if (!solved) return FN_FAIL;
  const float mx = fmaxf(fabsf(dl[0]), fmaxf(fabsf(dl[1]), fabsf(dl[2])));
  const float sc = mx > 0.25f ? 0.25f / mx : 1.f;
  F.u = fminf(fmaxf(F.u + dl[0] * sc, -1.5f), 1.5f);
  F.v = fminf(fmaxf(F.v + dl[1] * sc, -1.5f), 1.5f);
  F.tau = fminf(fmaxf(F.tau + dl[2] * sc, F.tmin), F.tmax);
  if (++F.it >= 50) return FN_FAIL;
  return FN_CONTINUE;
}

// Face Newton from one seed.  Returns 1 for an accepted root (max(|u|,|v|)
// <= 1 + 1e-4, tmin <= tau <= tmax), 0 for a converged root outside the
// face, -1 when the iteration did not converge.
template <int NG, typename C>
FEMRT_HD int face_newton(const GeoMode* __restrict__ gc, const float* o, const float* d, int f,
                         float u, float v, float tau, float tmin, float tmax, float& tau_o,
                         float& u_o, float& v_o, C& iters)
{
  FaceIter F;
  face_iter_begin(F, f, u, v, tau, tmin, tmax);
  for (;;) {
    float r, s, t;
    face_to_rst(F.f, F.u, F.v, r, s, t);
    float X[3], J[9];
    geom_eval<float, NG>(gc, r, s, t, X, J);
    ++iters;
    const int st = face_newton_step(F, o, d, X, J);
    if (st == FN_CONTINUE) continue;
    tau_o = F.tau;
    u_o = F.u;
    v_o = F.v;
    return st == FN_ACCEPT ? 1 : (st == FN_REJECT ? 0 : -1);
  }
}

// Projection of X(seed) onto the ray (high_order.h:154-161, 221-225).
template <int NG>
FEMRT_HD float seed_tau(const GeoMode* __restrict__ gc, const float* o, const float* d, int f,
                        float u, float v)
{
  float r, s, t, X[3];
  face_to_rst(f, u, v, r, s, t);
  geom_point<float, NG>(gc, r, s, t, X);
  const float dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  return ((X[0] - o[0]) * d[0] + (X[1] - o[1]) * d[1] + (X[2] - o[2]) * d[2]) / dd;
}

// The reference's seed grid {-0.8,0,0.8}^2, every seed run, nearest accepted
// root kept.
template <int NG, typename C>
FEMRT_HD bool face_root_grid(const GeoMode* __restrict__ gc, const float* o, const float* d, int f,
                             float tmin, float tmax, float& tau_o, float& u_o, float& v_o,
                             C& iters)
{
  bool found = false;
  float best = FLT_MAX;
  for (int sj = 0; sj < 3; ++sj)
    for (int si = 0; si < 3; ++si) {
      const float su = -0.8f + 0.8f * si, sv = -0.8f + 0.8f * sj;
      const float tau0 = seed_tau<NG>(gc, o, d, f, su, sv);
      float tt, uu, vv;
      if (face_newton<NG>(gc, o, d, f, su, sv, tau0, tmin, tmax, tt, uu, vv, iters) == 1 && tt < best) {
        best = tt;
        tau_o = tt;
        u_o = uu;
        v_o = vv;
        found = true;
      }
    }
  return found;
}

// Nearest accepted root of one face seeded from its Bernstein sub-patches:
// the ray is tested against the 3 x 3 sub-patch boxes (each bounds its piece
// of the face exactly), and a face Newton starts at the centre of every
// sub-patch whose box the ray crosses before `tcull`.  A root of the face lies
// in a crossed sub-patch, so this misses no root the seeds can converge to,
// and a single flat seed's failure mode (converging to a crossing of the
// face's polynomial extension outside the element while a crossing inside
// exists) is avoided.
// Whether the ray meets face f at (u, v) within about 17 degrees of the
// face's tangent plane: |n.d| < 0.3 |n| |d|, n = X_u x X_v.
template <int NG>
FEMRT_HD bool grazing_crossing(const GeoMode* __restrict__ gc, const float* d, int f, float u, float v)
{
  float r, s, t, X[3], J[9];
  face_to_rst(f, u, v, r, s, t);
  geom_eval<float, NG>(gc, r, s, t, X, J);
  const int cu = face_u(f) == 1 ? 1 : 0, cv = face_v(f);
  const float a[3] = {J[cu], J[3 + cu], J[6 + cu]}, b[3] = {J[cv], J[3 + cv], J[6 + cv]};
  const float n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  const float nd = n[0] * d[0] + n[1] * d[1] + n[2] * d[2];
  const float nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
  const float dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  return nd * nd < 0.09f * nn * dd;
}

template <int NG, typename W>
FEMRT_HD bool face_root_patches(const GeoMode* __restrict__ gc, const float* __restrict__ bx,
                                const float* o, const float* d, const float* id, int f, float tmin,
                                float tcull, float& tau_o, float& u_o, float& v_o, W& w)
{
  const float* sb = bx + 6 * (7 + 9 * f);
  bool found = false;
  float best = tcull;
  for (int q = 0; q < 9; ++q) {
    float s0, s1;
    ++w.slab;
    if (!slab(o, id, sb + 6 * q, sb + 6 * q + 3, tmin, best, s0, s1)) continue;
    const float uc = -1.f + (2.f * (q % 3) + 1.f) / 3.f;
    const float vc = -1.f + (2.f * (q / 3) + 1.f) / 3.f;
    float tt, uu, vv;
    int rc = face_newton<NG>(gc, o, d, f, uc, vc, 0.5f * (s0 + s1), tmin, 1e30f, tt, uu, vv, w.face_it);
    // converged to a crossing behind the range (the ray enters and leaves the
    // face within this sub-patch): seed again at the far end of the box
    if (rc == 0 && tt < tmin)
      rc = face_newton<NG>(gc, o, d, f, uc, vc, s1, tmin, 1e30f, tt, uu, vv, w.face_it);
    // A ray close to the tangent plane can cross one sub-patch twice, and
    // Newton from the centre may converge to the farther crossing: then the
    // four quarter centres of the sub-patch seed Newton too, from the near
    // end of the box, and a nearer accepted root replaces it.
    if (rc == 1 && tt < best * 1.0001f && grazing_crossing<NG>(gc, d, f, uu, vv)) {
      ++w.face_it;
      for (int k = 0; k < 4; ++k) {
        float t2, u2, v2;
        const float uq = uc + ((k & 1) ? 1.f : -1.f) / 6.f, vq = vc + ((k & 2) ? 1.f : -1.f) / 6.f;
        if (face_newton<NG>(gc, o, d, f, uq, vq, s0, tmin, 1e30f, t2, u2, v2, w.face_it) == 1 &&
            t2 < tt * (1.f - 1e-6f)) {
          tt = t2;
          uu = u2;
          vv = v2;
        }
      }
    }
    if (rc == 1 && tt < best) {
      best = tt;
      tau_o = tt;
      u_o = uu;
      v_o = vv;
      found = true;
    }
  }
  return found;
}

// Nearest accepted root of one face below `tcull`.  PLAIN: the reference's
// seed grid, then the sub-patch seeds when the grid finds nothing.  WARM and
// LEARNED: the sub-patch seeds.
template <int NG, typename W>
FEMRT_HD bool face_root(const GeoMode* __restrict__ gc, const float* __restrict__ bx, const float* o,
                        const float* d, const float* id, int f, float tmin, float tcull, int guess,
                        float& tau_o, float& u_o, float& v_o, W& w)
{
  if (guess == FEMRT_GUESS_PLAIN) {
    float tt, uu, vv;
    if (face_root_grid<NG>(gc, o, d, f, tmin, 1e30f, tt, uu, vv, w.face_it)) {
      if (!(tt < tcull)) return false;
      tau_o = tt;
      u_o = uu;
      v_o = vv;
      return true;
    }
  }
  return face_root_patches<NG>(gc, bx, o, d, id, f, tmin, tcull, tau_o, u_o, v_o, w);
}

FEMRT_HD void load_ray(const float* __restrict__ rays, int n, int r, float* o, float* d)
{
  o[0] = rays[r];
  o[1] = rays[(size_t)n + r];
  o[2] = rays[2 * (size_t)n + r];
  d[0] = rays[3 * (size_t)n + r];
  d[1] = rays[4 * (size_t)n + r];
  d[2] = rays[5 * (size_t)n + r];
}

// ------------------------------------------------------------- first hit

// (element, face) pairs a traversal skips: false hits already ruled out.
struct HitExcl {
  int n = 0;
  int e[4], f[4];
  FEMRT_HD bool has(int el, int fa) const
  {
    for (int i = 0; i < 4; ++i)
      if (i < n && e[i] == el && f[i] == fa) return true;
    return false;
  }
  FEMRT_HD void add(int el, int fa)
  {
    if (n < 4) {
      e[n] = el;
      f[n] = fa;
      ++n;
    }
  }
};

// BVH traversal over element boxes, near child first, short stack; per leaf
// element: element box, then per face its box and the face root.  The
// nearest accepted root over all elements (but `exclude`) and faces is the
// hit.  With `cross` set, only elements that can hold a crossing of an
// isovalue not in `found` take part (elem_may_cross), and only roots where
// the ray enters the element: the search then starts inside or between
// elements, and an element it only clips at a corner can show its exit root
// alone (the grazing entry root missed), which must not pass for an entry.
template <int NG, typename W>
FEMRT_HD HitF32 first_hit_ray(const TraceMesh& M, const float* o, const float* d, int guess,
                              float tmin, float tmax, W& w, int exclude = -1, const HitExcl* excl = nullptr,
                              const IsoParams* cross = nullptr, unsigned found = 0)
{
  HitF32 h;
  float id[3];
  for (int c = 0; c < 3; ++c) id[c] = 1.f / d[c];
  float best_t = tmax;
  int stack[FEMRT_STACK];
  float stack_t[FEMRT_STACK];
  int sp = 0, node = 0;
  float t0, t1;
  {
    const float4 ra = M.nodes[0], rb = M.nodes[1];
    const float rlo[3] = {ra.x, ra.y, ra.z}, rhi[3] = {rb.x, rb.y, rb.z};
    ++w.slab;
    if (!slab(o, id, rlo, rhi, tmin, best_t, t0, t1)) return h;
  }
  bool go = true;
  while (go) {
    const float4 na = M.nodes[2 * node], nb = M.nodes[2 * node + 1];
    const int a = as_int(na.w), b = as_int(nb.w);
    if (b < 0) {
      for (int i = 0; i < -b; ++i) {
        const int e = M.prims[a + i];
        if (e == exclude) continue;
        if (cross && !elem_may_cross(M, *cross, e, found)) continue;
        const float* bx = M.boxes + (size_t)e * FEMRT_BOX_STRIDE;
        float e0, e1;
        ++w.slab;
        if (!slab(o, id, bx, bx + 3, tmin, best_t, e0, e1)) continue;
        const GeoMode* gc = M.geom + (size_t)e * (NG * NG * NG);
        const int n_faces = M.surface ? 1 : 6;   // a surface element is its face 0
        for (int f = 0; f < n_faces; ++f) {
          if (excl && excl->has(e, f)) continue;
          const float* fb = bx + 6 * (1 + f);
          float f0, f1;
          ++w.slab;
          if (!slab(o, id, fb, fb + 3, tmin, best_t, f0, f1)) continue;
          float tt, uu, vv;
          if (face_root<NG>(gc, bx, o, d, id, f, tmin, best_t, guess, tt, uu, vv, w) && tt < best_t) {
            if (cross) {
              // entry: the face's fixed coordinate moves inward along the
              // ray, sign(face) * (J^-1 d)[c] < 0 (faces 2, 3, 5 at +1)
              float rst[3], X[3], J[9], m[3];
              face_to_rst(f, uu, vv, rst[0], rst[1], rst[2]);
              geom_eval<float, NG>(gc, rst[0], rst[1], rst[2], X, J);
              ++w.face_it;
              const int c = (f == 2 || f == 4) ? 0 : ((f == 1 || f == 3) ? 1 : 2);
              const float sg = (f == 2 || f == 3 || f == 5) ? 1.f : -1.f;
              if (!solve3<float>(J, d, m) || !(sg * m[c] < 0.f)) continue;
            }
            best_t = tt;
            h.elem = e;
            h.face = f;
            h.u = uu;
            h.v = vv;
            h.t = tt;
          }
        }
      }
    } else {
      const float4 la = M.nodes[2 * a], lb = M.nodes[2 * a + 1];
      const float4 qa = M.nodes[2 * b], qb = M.nodes[2 * b + 1];
      const float llo[3] = {la.x, la.y, la.z}, lhi[3] = {lb.x, lb.y, lb.z};
      const float qlo[3] = {qa.x, qa.y, qa.z}, qhi[3] = {qb.x, qb.y, qb.z};
      float l0, l1, q0, q1;
      w.slab += 2;
      const bool hl = slab(o, id, llo, lhi, tmin, best_t, l0, l1);
      const bool hq = slab(o, id, qlo, qhi, tmin, best_t, q0, q1);
      if (hl && hq) {
        const bool lfirst = l0 <= q0;
        if (sp < FEMRT_STACK) {
          stack[sp] = lfirst ? b : a;
          stack_t[sp] = lfirst ? q0 : l0;
          ++sp;
        }
        node = lfirst ? a : b;
        continue;
      }
      if (hl) { node = a; continue; }
      if (hq) { node = b; continue; }
    }
    go = false;
    while (sp > 0) {
      --sp;
      if (stack_t[sp] < best_t) {
        node = stack[sp];
        go = true;
        break;
      }
    }
  }
  return h;
}

// first_hit_ray with sub-patch seeds (WARM, LEARNED), the face-Newton work
// flattened.  Each trip of the loop first advances the traversal of the lanes
// that hold no Newton job until each holds one or has finished, then runs one
// Newton iteration on every lane that holds a job: a wave's geometry
// evaluations run together, where in first_hit_ray each lane's whole solve
// diverges from its neighbours' traversal.  Per lane the jobs, their order
// and the culling by the nearest root so far are first_hit_ray's (a face's
// sub-patch seeds culled at the running nearest root, which is the face
// search's own bound there), so the hit is the same.
template <int NG, typename W>
FEMRT_HD HitF32 first_hit_flat(const TraceMesh& M, const float* o, const float* d, float tmin, float tmax, W& w)
{
  HitF32 h;
  float id[3];
  for (int c = 0; c < 3; ++c) id[c] = 1.f / d[c];
  float best_t = tmax;
  int stack[FEMRT_STACK];
  float stack_t[FEMRT_STACK];
  int sp = 0, node = 0;
  {
    const float4 ra = M.nodes[0], rb = M.nodes[1];
    const float rlo[3] = {ra.x, ra.y, ra.z}, rhi[3] = {rb.x, rb.y, rb.z};
    float t0, t1;
    ++w.slab;
    if (!slab(o, id, rlo, rhi, tmin, best_t, t0, t1)) return h;
  }
  const int n_faces = M.surface ? 1 : 6;
  // traversal: the leaf being scanned (prims [la, la + ln), next li; ln = 0:
  // none), its element e (-1: none), face f and sub-patch q (-1: the face's
  // box not tested yet)
  int la = 0, ln = 0, li = 0, e = -1, f = 0, q = -1;
  const GeoMode* gc = M.geom;
  const float* bx = M.boxes;
  bool done = false;
  // the Newton job: stage 0 the sub-patch centre seed, 1 its far-end reseed,
  // 2..5 the four quarter seeds of a grazing crossing
  bool job = false;
  int stage = 0, rc = -1;
  FaceIter F;
  float s0 = 0.f, s1 = 0.f, uc = 0.f, vc = 0.f, tt = 0.f, uu = 0.f, vv = 0.f;
  for (;;) {
    while (!job && !done) {
      if (e >= 0) {
        if (f >= n_faces) {
          e = -1;
          continue;
        }
        if (q < 0) {
          const float* fb = bx + 6 * (1 + f);
          float f0, f1;
          ++w.slab;
          if (slab(o, id, fb, fb + 3, tmin, best_t, f0, f1))
            q = 0;
          else
            ++f;
          continue;
        }
        if (q >= 9) {
          ++f;
          q = -1;
          continue;
        }
        const float* sb = bx + 6 * (7 + 9 * f) + 6 * q;
        const int qq = q++;
        float a0, a1;
        ++w.slab;
        if (!slab(o, id, sb, sb + 3, tmin, best_t, a0, a1)) continue;
        uc = -1.f + (2.f * (qq % 3) + 1.f) / 3.f;
        vc = -1.f + (2.f * (qq / 3) + 1.f) / 3.f;
        s0 = a0;
        s1 = a1;
        face_iter_begin(F, f, uc, vc, 0.5f * (s0 + s1), tmin, 1e30f);
        stage = 0;
        job = true;
        continue;
      }
      if (li < ln) {
        const int ee = M.prims[la + li];
        ++li;
        const float* eb = M.boxes + (size_t)ee * FEMRT_BOX_STRIDE;
        float e0, e1;
        ++w.slab;
        if (!slab(o, id, eb, eb + 3, tmin, best_t, e0, e1)) continue;
        e = ee;
        bx = eb;
        gc = M.geom + (size_t)e * (NG * NG * NG);
        f = 0;
        q = -1;
        continue;
      }
      bool pop = ln > 0;   // a leaf finished
      ln = 0;
      if (!pop) {
        const float4 na = M.nodes[2 * node], nb = M.nodes[2 * node + 1];
        const int a = as_int(na.w), b = as_int(nb.w);
        if (b < 0) {
          la = a;
          ln = -b;
          li = 0;
          continue;
        }
        const float4 ca = M.nodes[2 * a], cb = M.nodes[2 * a + 1];
        const float4 da = M.nodes[2 * b], db = M.nodes[2 * b + 1];
        const float llo[3] = {ca.x, ca.y, ca.z}, lhi[3] = {cb.x, cb.y, cb.z};
        const float qlo[3] = {da.x, da.y, da.z}, qhi[3] = {db.x, db.y, db.z};
        float l0, l1, q0, q1;
        w.slab += 2;
        const bool hl = slab(o, id, llo, lhi, tmin, best_t, l0, l1);
        const bool hq = slab(o, id, qlo, qhi, tmin, best_t, q0, q1);
        if (hl && hq) {
          const bool lfirst = l0 <= q0;
          if (sp < FEMRT_STACK) {
            stack[sp] = lfirst ? b : a;
            stack_t[sp] = lfirst ? q0 : l0;
            ++sp;
          }
          node = lfirst ? a : b;
          continue;
        }
        if (hl || hq) {
          node = hl ? a : b;
          continue;
        }
      }
      done = true;
      while (sp > 0) {
        --sp;
        if (stack_t[sp] < best_t) {
          node = stack[sp];
          done = false;
          break;
        }
      }
    }
    if (!job) break;
    float r, s, t, X[3], J[9];
    face_to_rst(F.f, F.u, F.v, r, s, t);
    geom_eval<float, NG>(gc, r, s, t, X, J);
    ++w.face_it;
    const int st = face_newton_step(F, o, d, X, J);
    if (st == FN_CONTINUE) continue;
    const int res = st == FN_ACCEPT ? 1 : (st == FN_REJECT ? 0 : -1);
    if (stage <= 1) {
      rc = res;
      tt = F.tau;
      uu = F.u;
      vv = F.v;
      // converged to a crossing behind the range: seed again at the box's far end
      if (stage == 0 && rc == 0 && tt < tmin) {
        stage = 1;
        face_iter_begin(F, f, uc, vc, s1, tmin, 1e30f);
        continue;
      }
      if (rc == 1 && tt < best_t * 1.0001f && grazing_crossing<NG>(gc, d, f, uu, vv)) {
        ++w.face_it;
        stage = 2;
        face_iter_begin(F, f, uc - 1.f / 6.f, vc - 1.f / 6.f, s0, tmin, 1e30f);
        continue;
      }
    } else {
      if (res == 1 && F.tau < tt * (1.f - 1e-6f)) {
        tt = F.tau;
        uu = F.u;
        vv = F.v;
      }
      const int k = stage - 1;   // the next quarter seed
      if (k < 4) {
        stage = k + 2;
        face_iter_begin(F, f, uc + ((k & 1) ? 1.f : -1.f) / 6.f, vc + ((k & 2) ? 1.f : -1.f) / 6.f, s0, tmin, 1e30f);
        continue;
      }
    }
    if (rc == 1 && tt < best_t) {
      best_t = tt;
      h.elem = e;
      h.face = f;
      h.u = uu;
      h.v = vv;
      h.t = tt;
    }
    job = false;
  }
  return h;
}

// FP64 face Newton polish of a hit (unconstrained, quadratic convergence from
// the FP32 root; linear near a double root, hence up to 40 iterations).
// Converged when the step drops below 1e-14, or once it is below 1e-10 and
// stops halving (the FP64 rounding floor).  Returns whether it converged to
// a root on the face (max(|u|,|v|) <= 1 + 1e-4) at tau >= tmin: an FP32 root
// of a ray that passes the face within the FP32 residual tolerance without
// meeting it (a grazing near miss) has no FP64 root, and the caller traces
// such a ray again without this face.
template <int NG, typename W>
FEMRT_HD bool polish_hit(const GeoMode* __restrict__ gc, const double* o, const double* d, int f,
                         double& u, double& v, double& tau, double tmin, W& w)
{
  const bool us = (f == 2 || f == 4), vs = (f == 0 || f == 5);
  double prev = 1e300;
  bool conv = false;
  for (int it = 0; it < 40; ++it) {
    double rst[3], X[3], J[9];
    face_to_rst_t<double>(f, u, v, rst);
    geom_eval<double, NG>(gc, rst[0], rst[1], rst[2], X, J);
    ++w.geo64;
    const double R[3] = {X[0] - (o[0] + tau * d[0]), X[1] - (o[1] + tau * d[1]),
                         X[2] - (o[2] + tau * d[2])};
    const double A[9] = {us ? J[1] : J[0], vs ? J[1] : J[2], -d[0],
                         us ? J[4] : J[3], vs ? J[4] : J[5], -d[1],
                         us ? J[7] : J[6], vs ? J[7] : J[8], -d[2]};
    const double b[3] = {-R[0], -R[1], -R[2]};
    double dl[3];
    if (!solve3<double>(A, b, dl)) break;
    u += dl[0];
    v += dl[1];
    tau += dl[2];
    if (!(fabs(u) < 10.0 && fabs(v) < 10.0)) break;
    const double st = fmax(fabs(dl[0]) + fabs(dl[1]), fabs(dl[2]) / (1.0 + fabs(tau)));
    if (st < 1e-14 || (st < 1e-10 && st > 0.5 * prev)) {
      conv = true;
      break;
    }
    prev = st;
  }
  return conv && fmax(fabs(u), fabs(v)) <= 1.0 + FEMRT_FACE_ACC && tau >= tmin;
}

// ------------------------------------------------------ point inversion

// Newton x -> rst in FP32 from the guess in rst (HexElement.cu:227-272:
// undamped, |delta|_inf < 1e-5, <= 50 iterations).  J holds the Jacobian at
// the last evaluated iterate.
template <int NG, typename C>
FEMRT_HD bool invert_f32(const GeoMode* __restrict__ gc, const float* x, float* rst, float* J,
                         C& iters)
{
  for (int it = 0; it < 50; ++it) {
    float X[3];
    geom_eval<float, NG>(gc, rst[0], rst[1], rst[2], X, J);
    ++iters;
    const float b[3] = {X[0] - x[0], X[1] - x[1], X[2] - x[2]};
    float dl[3];
    if (!solve3<float>(J, b, dl)) return false;
    rst[0] -= dl[0];
    rst[1] -= dl[1];
    rst[2] -= dl[2];
    if (!(fabsf(rst[0]) < 1e4f && fabsf(rst[1]) < 1e4f && fabsf(rst[2]) < 1e4f)) return false;
    if (fabsf(dl[0]) < 1e-5f && fabsf(dl[1]) < 1e-5f && fabsf(dl[2]) < 1e-5f) return true;
  }
  return false;
}

// The preimage of x inside the element (unique for a valid element): Newton
// from the guess in rst; when that converges outside the cube (Newton from
// rst = 0 finds spurious preimages of the polynomial map for points near a
// face) or fails, restart from rst = 0 (when zero_retry: the plain fallback
// of a learned guess), then from the 8 octant centres (+-0.5)^3.
template <int NG, typename C>
FEMRT_HD bool invert_in_cube_f32(const GeoMode* __restrict__ gc, const float* x, float* rst, float* J,
                                 C& iters, bool zero_retry = false)
{
  if (invert_f32<NG>(gc, x, rst, J, iters) &&
      fmaxf(fabsf(rst[0]), fmaxf(fabsf(rst[1]), fabsf(rst[2]))) <= 1.001f)
    return true;
  if (zero_retry) {
    rst[0] = rst[1] = rst[2] = 0.f;
    if (invert_f32<NG>(gc, x, rst, J, iters) &&
        fmaxf(fabsf(rst[0]), fmaxf(fabsf(rst[1]), fabsf(rst[2]))) <= 1.001f)
      return true;
  }
  for (int s = 0; s < 8; ++s) {
    rst[0] = (s & 1) ? 0.5f : -0.5f;
    rst[1] = (s & 2) ? 0.5f : -0.5f;
    rst[2] = (s & 4) ? 0.5f : -0.5f;
    if (invert_f32<NG>(gc, x, rst, J, iters) &&
        fmaxf(fabsf(rst[0]), fmaxf(fabsf(rst[1]), fabsf(rst[2]))) <= 1.001f)
      return true;
  }
  return false;
}

// FP64 Newton x -> rst from a guess.  Converged when the step drops below
// 1e-13, or once it is below 1e-9 and stops halving (the FP64 rounding floor
// of x(rst) - x divided by |J|).  On success J is evaluated at the result.
template <int NG, typename C>
FEMRT_HD bool invert_f64(const GeoMode* __restrict__ gc, const double* x, double* rst, double* J,
                         C& iters)
{
  double prev = 1e300;
  for (int it = 0; it < 30; ++it) {
    double X[3];
    geom_eval<double, NG>(gc, rst[0], rst[1], rst[2], X, J);
    ++iters;
    const double b[3] = {X[0] - x[0], X[1] - x[1], X[2] - x[2]};
    double dl[3];
    if (!solve3<double>(J, b, dl)) return false;
    rst[0] -= dl[0];
    rst[1] -= dl[1];
    rst[2] -= dl[2];
    const double st = fmax(fabs(dl[0]), fmax(fabs(dl[1]), fabs(dl[2])));
    if (st < 1e-13 || (st < 1e-9 && st > 0.5 * prev)) {
      geom_eval<double, NG>(gc, rst[0], rst[1], rst[2], X, J);
      ++iters;
      return true;
    }
    prev = st;
  }
  return false;
}

// ------------------------------------------------------------ segment exit

// Exit of the ray from element e: nearest accepted face root with
// t > t_lo on the faces other than the entry face (and on the entry face too
// when include_entry is set).
template <int NG, typename W>
FEMRT_HD bool find_exit_search(const GeoMode* __restrict__ gc, const float* __restrict__ bx,
                               const float* o, const float* d, const float* id, int f_in,
                               float t_lo, int guess, bool include_entry, int& f_out, float& t_out,
                               float& u_out, float& v_out, W& w)
{
  bool found = false;
  float best = FLT_MAX;
  for (int f = 0; f < 6; ++f) {
    if (f == f_in && !include_entry) continue;
    const float* fb = bx + 6 * (1 + f);
    float f0, f1;
    ++w.slab;
    if (!slab(o, id, fb, fb + 3, t_lo, best, f0, f1)) continue;
    float tt, uu, vv;
    if (face_root<NG>(gc, bx, o, d, id, f, t_lo, best, guess, tt, uu, vv, w) && tt < best) {
      best = tt;
      f_out = f;
      t_out = tt;
      u_out = uu;
      v_out = vv;
      found = true;
    }
  }
  return found;
}

// Exit found by following the ray with point inversion: steps of 1/32 of the
// element box extent until a point has no preimage in the element, bisection
// of that inside/outside predicate, then a face Newton on the face whose
// coordinate was crossed, seeded at the bisected point.  Used when the face
// searches find no exit (two crossings of one face inside one sub-patch).
template <int NG, typename W>
FEMRT_HD bool find_exit_marching(const GeoMode* __restrict__ gc, const float* __restrict__ bx,
                                 const float* o, const float* d, const float* id,
                                 const float* rst_in, float t_in, float t_lo, int& f_out,
                                 float& t_out, float& u_out, float& v_out, W& w)
{
  float b0, b1;
  ++w.slab;
  if (!slab(o, id, bx, bx + 3, t_in, 1e30f, b0, b1)) return false;
  const float ext = fmaxf(bx[3] - bx[0], fmaxf(bx[4] - bx[1], bx[5] - bx[2]));
  const float dt = ext / 32.f;
  float ta = t_in, ra[3] = {rst_in[0], rst_in[1], rst_in[2]};
  float tb = -1.f;
  float J[9];
  for (float t = t_lo; t <= b1 + dt; t += dt) {
    const float x[3] = {fmaf(t, d[0], o[0]), fmaf(t, d[1], o[1]), fmaf(t, d[2], o[2])};
    float r[3] = {ra[0], ra[1], ra[2]};
    if (invert_in_cube_f32<NG>(gc, x, r, J, w.inv_it)) {
      ta = t;
      ra[0] = r[0];
      ra[1] = r[1];
      ra[2] = r[2];
    } else {
      tb = t;
      break;
    }
  }
  if (tb < 0.f) return false;
  for (int it = 0; it < 14; ++it) {
    const float tm = 0.5f * (ta + tb);
    const float x[3] = {fmaf(tm, d[0], o[0]), fmaf(tm, d[1], o[1]), fmaf(tm, d[2], o[2])};
    float r[3] = {ra[0], ra[1], ra[2]};
    if (invert_in_cube_f32<NG>(gc, x, r, J, w.inv_it)) {
      ta = tm;
      ra[0] = r[0];
      ra[1] = r[1];
      ra[2] = r[2];
    } else {
      tb = tm;
    }
  }
  // the crossed face: the coordinate of the last inside point nearest +-1
  int c = 0;
  for (int k = 1; k < 3; ++k)
    if (fabsf(ra[k]) > fabsf(ra[c])) c = k;
  const int f = c == 0 ? (ra[0] > 0.f ? 2 : 4) : (c == 1 ? (ra[1] > 0.f ? 3 : 1) : (ra[2] > 0.f ? 5 : 0));
  float su, sv;
  rst_to_face(f, ra, su, sv);
  float tt, uu, vv;
  if (face_newton<NG>(gc, o, d, f, su, sv, ta, t_lo, 1e30f, tt, uu, vv, w.face_it) == 1 &&
      fabsf(tt - ta) <= 4.f * dt) {
    f_out = f;
    t_out = tt;
    u_out = uu;
    v_out = vv;
    return true;
  }
  f_out = f;
  t_out = ta;
  u_out = fminf(fmaxf(su, -1.f), 1.f);
  v_out = fminf(fmaxf(sv, -1.f), 1.f);
  return true;
}

// Exit predicted from the entry point: the reference-space line
// rst_in + s J^{-1}(rst_in) d meets the planes of the cube faces (the entry
// face excluded) at s_f; a face Newton on each crossed face, nearest crossing
// first (at most 3), seeded at the predicted point, gives the exit.  Returns
// false when none converges to an accepted root.
// m_in, m_out (may be null): drst/dt = J^-1 d at the entry and at the exit
// found (the Jacobian of the last iterate), for the scan's interpolated
// guesses.
template <int NG, typename W>
FEMRT_HD bool find_exit_predicted(const GeoMode* __restrict__ gc, const float* o, const float* d,
                                  int f_in, const float* rst_in, float t_in, float t_lo, int& f_out,
                                  float& t_out, float& u_out, float& v_out, W& w,
                                  float* m_in = nullptr, float* m_out = nullptr)
{
  float X[3], J[9], dl[3];
  geom_eval<float, NG>(gc, rst_in[0], rst_in[1], rst_in[2], X, J);
  ++w.face_it;
  if (!solve3<float>(J, d, dl)) return false;
  if (m_in)
    for (int c = 0; c < 3; ++c) m_in[c] = dl[c];
  // crossings of the three planes the line heads for
  float sc[3];
  int fc[3];
  for (int c = 0; c < 3; ++c) {
    sc[c] = FLT_MAX;
    fc[c] = -1;
    if (dl[c] == 0.f) continue;
    const float target = dl[c] > 0.f ? 1.f : -1.f;
    float sv = (target - rst_in[c]) / dl[c];
    const int f = c == 0 ? (target > 0 ? 2 : 4) : (c == 1 ? (target > 0 ? 3 : 1) : (target > 0 ? 5 : 0));
    if (f == f_in) continue;
    if (!(sv > 0.f)) {
      // the entry point lies on this face's plane already (a ray entering
      // at an edge of the entry face, moving outward): the element is only
      // clipped, and the exit through this face follows at once
      if (!(fabsf(rst_in[c]) >= 0.999f)) continue;
      sv = 0.f;
    }
    sc[c] = sv;
    fc[c] = f;
  }
  // the crossed faces nearest first; one face-Newton iteration per trip of
  // the loop, so a lane moves to its next face as soon as one fails instead
  // of the wave finishing each attempt together
  auto next_face = [&](FaceIter& F) -> bool {
    int cb = -1;
    for (int c = 0; c < 3; ++c)
      if (fc[c] >= 0 && (cb < 0 || sc[c] < sc[cb])) cb = c;
    if (cb < 0) return false;
    const int f = fc[cb];
    const float sv = sc[cb];
    fc[cb] = -1;
    float pr[3];
    for (int c = 0; c < 3; ++c) pr[c] = rst_in[c] + sv * dl[c];
    float su, sw;
    rst_to_face(f, pr, su, sw);
    face_iter_begin(F, f, fminf(fmaxf(su, -1.f), 1.f), fminf(fmaxf(sw, -1.f), 1.f), t_in + sv, t_lo, 1e30f);
    return true;
  };
  FaceIter F;
  bool have = next_face(F);
  while (have) {
    float r, s, t;
    face_to_rst(F.f, F.u, F.v, r, s, t);
    geom_eval<float, NG>(gc, r, s, t, X, J);
    ++w.face_it;
    const int st = face_newton_step(F, o, d, X, J);
    if (st == FN_ACCEPT) {
      f_out = F.f;
      t_out = F.tau;
      u_out = F.u;
      v_out = F.v;
      if (m_out && !solve3<float>(J, d, m_out)) m_out[0] = m_out[1] = m_out[2] = 0.f;
      return true;
    }
    if (st != FN_CONTINUE) have = next_face(F);
  }
  return false;
}

// Coupled Newton for the isosurface point on the ray: unknowns (r, s, t) and
// the ray parameter tau, equations x(rst) - (o + tau d) = 0 and
// u(rst) - iso = 0.  The 4x4 system [J -d; g^T 0] is solved by block
// elimination with the cofactors of J: a = J^-1 F_x, b = J^-1 d,
// dtau = (F_u - g.a) / (g.b), drst = a + b dtau (g.b = grad u . d is the
// derivative of u along the ray).  tau stays inside [lo, hi]; returns false
// when an iterate leaves it or the system is singular.
template <typename T, int NG, int NF, typename C>
FEMRT_HD bool iso_newton(const GeoMode* __restrict__ gc, const float* __restrict__ fc, const T* o,
                         const T* d, T iso, T lo, T hi, int max_it, T tol_t, T tol_r, T& tau, T* rst,
                         T& u, C& geo_it, C& field_it)
{
  for (int it = 0; it < max_it; ++it) {
    T X[3], J[9], g[3];
    geom_eval<T, NG>(gc, rst[0], rst[1], rst[2], X, J);
    field_eval<T, NF>(fc, rst[0], rst[1], rst[2], &u, g);
    ++geo_it;
    ++field_it;
    const T Fx[3] = {X[0] - (o[0] + tau * d[0]), X[1] - (o[1] + tau * d[1]),
                     X[2] - (o[2] + tau * d[2])};
    const T Fu = u - iso;
    T a[3], b[3];
    if (!solve3<T>(J, Fx, a) || !solve3<T>(J, d, b)) return false;
    const T gb = g[0] * b[0] + g[1] * b[1] + g[2] * b[2];
    if (!(gb != T(0))) return false;
    // Newton step: [J -d; g^T 0] [drst; dtau] = [Fx; Fu], then subtract
    const T dtau = (Fu - (g[0] * a[0] + g[1] * a[1] + g[2] * a[2])) / gb;
    const T dr[3] = {a[0] + b[0] * dtau, a[1] + b[1] * dtau, a[2] + b[2] * dtau};
    rst[0] -= dr[0];
    rst[1] -= dr[1];
    rst[2] -= dr[2];
    tau -= dtau;
    if (!(tau >= lo && tau <= hi)) return false;
    if (fabs((double)dtau) <= (double)tol_t * (1.0 + fabs((double)tau)) &&
        fmax(fabs((double)dr[0]), fmax(fabs((double)dr[1]), fabs((double)dr[2]))) <= (double)tol_r)
      return true;
  }
  return false;
}

// ------------------------------------------------------------- element walk
//
// The walk of one hit ray through face-adjacent elements, sampling
// g = u - iso at P.samples uniform sub-intervals of each element segment
// [t_in, t_out]; the first sub-interval with g(prev) g(cur) <= 0 of each
// isovalue is refined and recorded (fem/RayIsosurface.cu:67-188 does the
// same scan over the padded box of one element).  It is split in three
// steps per segment -- exit, samples, advance -- so the kernel can run them
// in wave-uniform rounds (and fetch a new ray into a lane whose ray ended)
// with the batched network guesses in between.

// Per-ray state of the walk between rounds.
struct MarchState {
  int e, f_in;            // current element and its entry face
  float t_in, u_in, v_in;
  int f_out;              // exit of the current segment (valid when ok)
  float t_out, u_out, v_out;
  int step;               // segments walked (bits 0-15), exits found by marching in a row
                          // (bits 16-23), stalls skipped (bits 24-30)
  unsigned found;         // isovalues whose bracket has been recorded
  int run;                // segments in a row whose scan the field bounds ruled out
  bool ok;
  bool redo;              // samples showed a wrong exit: search again, entry face included
  bool have_m;            // m_in, m_out hold drst/dt at the entry and exit (predicted exits)
  float m_in[3], m_out[3];
};

// Whether the walk's last FEMRT_MAX_MARCHED_EXITS + 1 exits all came from the
// marching fallback; march_skips(S): stalls skipped so far.
FEMRT_HD bool march_stalled(const MarchState& S) { return ((S.step >> 16) & 0xFF) > FEMRT_MAX_MARCHED_EXITS; }
FEMRT_HD int march_skips(const MarchState& S) { return (S.step >> 24) & 0x7F; }

FEMRT_HD void march_begin(MarchState& S, const HitF32& h)
{
  S.e = h.elem;
  S.f_in = h.face;
  S.t_in = h.t;
  S.u_in = h.u;
  S.v_in = h.v;
  S.f_out = -1;
  S.t_out = S.u_out = S.v_out = 0.f;
  S.step = 0;
  S.found = 0;
  S.run = 0;
  S.ok = false;
  S.redo = false;
  S.have_m = false;
}

// Exit of the current segment.  First pass: predicted exit (engineered and
// learned guesses), then the face search without and with the entry face (a
// ray can leave through the face it entered by: a curved boundary face seen
// at a grazing angle), then marching.  After a redo: the search including the
// entry face, then marching.
template <int NG, typename W>
FEMRT_HD void march_exit(const TraceMesh& M, const IsoParams& P, const float* o, const float* d,
                         const float* id, MarchState& S, W& w)
{
  const GeoMode* gc = M.geom + (size_t)S.e * (NG * NG * NG);
  const float* bx = M.boxes + (size_t)S.e * FEMRT_BOX_STRIDE;
  const float ext = fmaxf(bx[3] - bx[0], fmaxf(bx[4] - bx[1], bx[5] - bx[2]));
  // exits closer than eps (t_lo below) to the entry are not resolved in FP32
  float rst_in[3];
  face_to_rst(S.f_in, S.u_in, S.v_in, rst_in[0], rst_in[1], rst_in[2]);
  bool ok = false;
  S.have_m = false;
  if (march_stalled(S)) {
    // The last exits all came from marching: the ray runs inside a face or
    // along an edge that elements share, within FP32 resolution of it, and
    // the elements on both sides hand it back and forth with hardly any
    // advance.  The entry moves 1e-3 of the element along the ray (its (u, v)
    // on the entry face from the point's preimage), and the walk goes on from
    // there in this element.
    S.step = (S.step & 0xFFFF) | ((march_skips(S) + 1) << 24);
    ++w.skips;
    S.t_in = fmaxf(S.t_in, S.t_in + 1e-3f * ext);
    const float x[3] = {fmaf(S.t_in, d[0], o[0]), fmaf(S.t_in, d[1], o[1]), fmaf(S.t_in, d[2], o[2])};
    float r[3] = {rst_in[0], rst_in[1], rst_in[2]}, J[9];
    if (invert_in_cube_f32<NG>(gc, x, r, J, w.inv_it)) {
      rst_to_face(S.f_in, r, S.u_in, S.v_in);
      S.u_in = fminf(fmaxf(S.u_in, -1.f), 1.f);
      S.v_in = fminf(fmaxf(S.v_in, -1.f), 1.f);
    }
    face_to_rst(S.f_in, S.u_in, S.v_in, rst_in[0], rst_in[1], rst_in[2]);
  }
  const float t_lo = S.t_in + 1e-4f * ext;
  if (!S.redo) {
    ++w.marched;
    if (P.guess != FEMRT_GUESS_PLAIN) {
      ok = find_exit_predicted<NG>(gc, o, d, S.f_in, rst_in, S.t_in, t_lo, S.f_out, S.t_out, S.u_out,
                                   S.v_out, w, S.m_in, S.m_out);
      w.exit_pred += ok;
      S.have_m = ok;
    }
    if (!ok) {
      ok = find_exit_search<NG>(gc, bx, o, d, id, S.f_in, t_lo, P.guess, false, S.f_out, S.t_out,
                                S.u_out, S.v_out, w);
      w.exit_search += ok;
    }
  } else {
    ++w.redo;
  }
  if (!ok) {
    ok = find_exit_search<NG>(gc, bx, o, d, id, S.f_in, t_lo, P.guess, true, S.f_out, S.t_out,
                              S.u_out, S.v_out, w);
    w.exit_search += ok;
  }
  if (ok) {
    S.step &= ~(0xFF << 16);  // the run of marched exits ends
  } else {
    ok = find_exit_marching<NG>(gc, bx, o, d, id, rst_in, S.t_in, t_lo, S.f_out, S.t_out, S.u_out,
                                S.v_out, w);
    w.exit_march += ok;
    if (ok && ((S.step >> 16) & 0xFF) < 0xFF) S.step += 1 << 16;
  }
  S.ok = ok;
}

// Ray parameter of sample j of the current segment (j = 0 .. samples).
FEMRT_HD float march_sample_t(const MarchState& S, int j, int samples)
{
  return j == samples ? S.t_out : S.t_in + (S.t_out - S.t_in) * ((float)j / (float)samples);
}

// Guess sources for the interior samples' point inversions.
struct NoGuess {
  FEMRT_HD bool get(int, float*) const { return false; }
};

// Samples of the current segment (S.ok).  Brackets found are handed to
// sink.put(k, bracket) and marked in S.found.  Returns true when a sample
// has no preimage inside the element on the first pass (the exit is wrong:
// the caller sets S.redo and finds the exit again).
template <int NG, int NF, typename Guess, typename Sink, typename W>
FEMRT_HD bool march_samples(const TraceMesh& M, const IsoParams& P, const float* o, const float* d,
                            MarchState& S, const Guess& guess, Sink& sink, W& w)
{
  const GeoMode* gc = M.geom + (size_t)S.e * (NG * NG * NG);
  const float* fc = M.field + (size_t)S.e * (NF * NF * NF);
  const float lip = M.lip ? M.lip[S.e] : 0.f;
  const unsigned open_all = (1u << P.n_iso) - 1u;
  const int Ms = P.samples;
  if (M.frange) {
    // no isovalue still open lies within the element's bounds of u: every
    // sample of the segment has the same sign for each of them, so the scan
    // can find no crossing here and is skipped
    const float lo = M.frange[2 * (size_t)S.e], hi = M.frange[2 * (size_t)S.e + 1];
    bool any = false;
    for (int k = 0; k < FEMRT_MAX_ISO; ++k)
      if (k < P.n_iso && !((S.found >> k) & 1u) && P.iso[k] >= lo && P.iso[k] <= hi) any = true;
    if (!any) {
      ++w.skipped;
      return false;
    }
  }
  float rst_in[3];
  face_to_rst(S.f_in, S.u_in, S.v_in, rst_in[0], rst_in[1], rst_in[2]);
  // Certified skipping.  Inside the element |rst(t) - rst(t')|_inf <=
  // jinv |d|_inf |t - t'| (jinv bounds |J^-1|_inf there) and |u(a) - u(b)| <=
  // lip |a - b|_inf, so from an interior sample j with |g_k| > lip err + eps
  // for every open isovalue (err bounding the sample's distance to its
  // preimage, eps the FP32 evaluation error of u), the next
  // R = floor((min_k |g_k| - lip err - eps) / cstep) samples keep the signs
  // of sample j, cstep = lip jinv |d|_inf dt.  The scan moves on to the last
  // of them (never past the last interior sample) instead of the next one:
  // every sample where a sign can change is still evaluated next to its
  // predecessor, so the brackets are those of the full scan.
  float cstep = 0.f;
  if (P.skip && M.jinv && lip > 0.f) {
    const float kj = M.jinv[S.e];
    const float dmax = fmaxf(fabsf(d[0]), fmaxf(fabsf(d[1]), fabsf(d[2])));
    if (kj > 0.f) cstep = lip * kj * dmax * ((S.t_out - S.t_in) / (float)Ms) * (1.f + 1e-5f);
  }
  float gprev[FEMRT_MAX_ISO];
  float rprev[3] = {0.f, 0.f, 0.f}, rprev2[3] = {0.f, 0.f, 0.f};
  float tprev = S.t_in, tprev2 = S.t_in;
  bool jumped = false, jumped_prev = false;  // the steps out of and into the last sample skipped samples
  bool have_prev = false;
  int n_prev = 0;  // consecutive samples with a preimage
  unsigned found_here = 0;
  // sign checks of a resolved sample against the previous one; brackets
  auto take = [&](float t, const float* r, float u) {
    for (int k = 0; k < FEMRT_MAX_ISO; ++k) {
      if (k >= P.n_iso || ((S.found | found_here) >> k) & 1u) continue;
      const float g = u - P.iso[k];
      if (have_prev && gprev[k] * g <= 0.f) {
        IsoBracket br;
        br.elem = S.e;
        br.ta = tprev;
        br.tb = t;
        br.ga = gprev[k];
        // the FP64 endgame starts from the secant (regula falsi) point of the
        // bracket, rst interpolated alike; the root's refinement runs there
        // on a compact queue, one bracket per lane, instead of stalling the
        // walk's wave for the few lanes with a bracket
        const float den = gprev[k] - g;
        const float wt = den != 0.f ? fminf(fmaxf(gprev[k] / den, 0.f), 1.f) : 0.5f;
        br.tm = tprev + wt * (t - tprev);
        for (int c = 0; c < 3; ++c) {
          br.rst[c] = rprev[c] + wt * (r[c] - rprev[c]);
          br.seg[c] = rprev[c];
          br.seg[3 + c] = r[c];
        }
        br.seg[6] = gprev[k];
        br.seg[7] = g;
        sink.put(k, br);
        ++w.brackets;
        found_here |= 1u << k;
      }
      gprev[k] = g;
    }
    for (int c = 0; c < 3; ++c) {
      rprev2[c] = rprev[c];
      rprev[c] = r[c];
    }
    tprev2 = tprev;
    tprev = t;
    have_prev = true;
    ++n_prev;
  };
  // the segment ends lie on faces whose (u, v) the face roots give:
  // evaluate there, never by extrapolating past the face
  {
    const float u = field_value<float, NF>(fc, rst_in[0], rst_in[1], rst_in[2]);
    ++w.field;
    take(S.t_in, rst_in, u);
  }
  // Early sign acceptance: the scan needs only the sign of g_k = u - iso_k
  // at each sample.  Once a Newton step st is below 1e-2 the iterate is
  // within about C st^2 << st of the preimage (quadratic convergence), so
  // |u(iterate) - u(preimage)| < lip * st, lip bounding |u(a) - u(b)| /
  // |a - b|_inf on the element; when every open isovalue has |g_k| > 2 lip st
  // the signs are those of the converged sample and the iterate is taken.
  // One attempt per Newton run (a failed one costs a field value).
  //
  // Interior samples 1 .. Ms-1 (point inversion, then the field).  Each trip
  // of the loop is one Newton iteration; a lane moves on to its next sample
  // as soon as the current one is resolved, so the lanes of a wave stay busy
  // while they need different iteration counts.  Per sample, the Newton runs
  // of invert_in_cube_f32 in its order: from the guess, from rst = 0 (after a
  // network guess), from the octant centres; accepted inside the cube only.
  float rst[3] = {rst_in[0], rst_in[1], rst_in[2]};
  int j = 1, stage = 0, its = 0;
  bool zero_retry = false, cert_tried = false;
  auto start_sample = [&]() {
    zero_retry = false;
    if (n_prev < 2 && S.have_m && P.hermite && P.guess != FEMRT_GUESS_PLAIN) {
      // no two samples to extrapolate from, but the predicted exit left
      // drst/dt at both ends: the cubic Hermite interpolant of rst(t)
      // between entry and exit, exact at both ends with their tangents
      const float L = S.t_out - S.t_in;
      const float q = (march_sample_t(S, j, Ms) - S.t_in) / L, q2 = q * q, q3 = q2 * q;
      const float h00 = 2.f * q3 - 3.f * q2 + 1.f, h10 = q3 - 2.f * q2 + q, h01 = 3.f * q2 - 2.f * q3,
                  h11 = q3 - q2;
      float r_out[3];
      face_to_rst(S.f_out, S.u_out, S.v_out, r_out[0], r_out[1], r_out[2]);
      for (int c = 0; c < 3; ++c)
        rst[c] = fminf(fmaxf(h00 * rst_in[c] + h01 * r_out[c] + L * (h10 * S.m_in[c] + h11 * S.m_out[c]), -1.f),
                       1.f);
    } else if (guess.get(j, rst)) {
      zero_retry = true;  // plain Newton from 0 is the fallback of a learned guess
    } else if (P.guess == FEMRT_GUESS_PLAIN) {
      rst[0] = rst[1] = rst[2] = 0.f;
    } else if (n_prev >= 2 && !jumped && !jumped_prev) {
      // warm start extrapolated from the two previous (uniformly spaced)
      // samples: O(dt^2) from the preimage instead of O(dt)
      for (int c = 0; c < 3; ++c) rst[c] = fminf(fmaxf(2.f * rprev[c] - rprev2[c], -1.f), 1.f);
    } else if (n_prev >= 2) {
      // after a certified skip the spacing is uneven: linear in t
      const float tn = march_sample_t(S, j, Ms);
      const float q = (tn - tprev) / (tprev - tprev2);
      for (int c = 0; c < 3; ++c) rst[c] = fminf(fmaxf(rprev[c] + q * (rprev[c] - rprev2[c]), -1.f), 1.f);
    }
    stage = 0;
    its = 0;
    cert_tried = false;
  };
  if (j < Ms) start_sample();
  while (j < Ms) {
    const float t = march_sample_t(S, j, Ms);
    const float x[3] = {fmaf(t, d[0], o[0]), fmaf(t, d[1], o[1]), fmaf(t, d[2], o[2])};
    float X[3], J[9], dl[3];
    geom_eval<float, NG>(gc, rst[0], rst[1], rst[2], X, J);
    ++w.inv_it;
    ++its;
    const float bv[3] = {X[0] - x[0], X[1] - x[1], X[2] - x[2]};
    bool fail_run = !solve3<float>(J, bv, dl), conv = false, try_cert = false;
    float st = 0.f;
    if (!fail_run) {
      rst[0] -= dl[0];
      rst[1] -= dl[1];
      rst[2] -= dl[2];
      if (!(fabsf(rst[0]) < 1e4f && fabsf(rst[1]) < 1e4f && fabsf(rst[2]) < 1e4f)) {
        fail_run = true;
      } else if (fabsf(dl[0]) < 1e-5f && fabsf(dl[1]) < 1e-5f && fabsf(dl[2]) < 1e-5f) {
        conv = true;
      } else {
        st = fmaxf(fabsf(dl[0]), fmaxf(fabsf(dl[1]), fabsf(dl[2])));
        // early sign acceptance: one attempt per Newton run
        try_cert = st < 1e-2f && !cert_tried && lip > 0.f;
        cert_tried = cert_tried || try_cert;
      }
    }
    const bool inside = fmaxf(fabsf(rst[0]), fmaxf(fabsf(rst[1]), fabsf(rst[2]))) <= 1.001f;
    // one field value per trip (a single call site keeps the lanes together):
    // the converged sample inside the element, or the iterate of an attempt
    float u = 0.f;
    if ((conv && inside) || try_cert) {
      u = field_value<float, NF>(fc, rst[0], rst[1], rst[2]);
      ++w.field;
    }
    bool done = conv;
    if (try_cert) {
      // accepted when every open isovalue is clear of the margin 2 lip st
      const float margin = 2.f * lip * st;
      const unsigned open = open_all & ~(S.found | found_here);
      done = true;
      for (int k = 0; k < FEMRT_MAX_ISO; ++k)
        if (k < P.n_iso && ((open >> k) & 1u) && !(fabsf(u - P.iso[k]) > margin)) done = false;
    }
    if (!fail_run && !done && its >= 50) fail_run = true;
    if (done && !inside) {
      done = false;  // a preimage outside the element: next seed
      fail_run = true;
    }
    if (fail_run) {
      ++stage;
      if (stage == 1 && !zero_retry) stage = 2;
      if (stage <= 9) {
        if (stage == 1) {
          rst[0] = rst[1] = rst[2] = 0.f;
        } else {
          const int sd = stage - 2;
          rst[0] = (sd & 1) ? 0.5f : -0.5f;
          rst[1] = (sd & 2) ? 0.5f : -0.5f;
          rst[2] = (sd & 4) ? 0.5f : -0.5f;
        }
        its = 0;
        cert_tried = false;
        continue;
      }
      // no preimage inside the element: the segment end is wrong
      if (!S.redo) return true;
      have_prev = false;
      n_prev = 0;
      jumped = jumped_prev = false;
      rst[0] = rst[1] = rst[2] = 0.f;
      ++j;
      if (j < Ms) start_sample();
      continue;
    }
    if (done) {
      take(t, rst, u);
      int nj = j + 1;
      if (cstep > 0.f) {
        const unsigned open = open_all & ~(S.found | found_here);
        float gl = 3.4e38f;
        for (int k = 0; k < FEMRT_MAX_ISO; ++k)
          if (k < P.n_iso && ((open >> k) & 1u)) gl = fminf(gl, fabsf(u - P.iso[k]));
        // distance to the preimage: 2 st after an accepted step, else within
        // the convergence test's 1e-5 (doubled); FP32 evaluation error of u
        gl -= lip * (conv ? 2e-5f : 2.f * st) + 2e-4f + 1e-4f * lip;
        if (u == u && gl >= 2.f * cstep) {
          nj = j + (int)fminf(gl / cstep, (float)Ms);
          nj = nj < Ms - 1 ? nj : Ms - 1;
          if (nj <= j) nj = j + 1;
          w.samples_skipped += (unsigned)(nj - j - 1);
        }
      }
      jumped_prev = jumped;
      jumped = nj > j + 1;
      j = nj;
      if (j < Ms) start_sample();
    }
  }
  {
    float r[3];
    face_to_rst(S.f_out, S.u_out, S.v_out, r[0], r[1], r[2]);
    const float u = field_value<float, NF>(fc, r[0], r[1], r[2]);
    ++w.field;
    take(S.t_out, r, u);
  }
  S.found |= found_here;
  return false;
}

// --------------------------------------------- cooperative segment scan
//
// The same scan with the samples of a segment evaluated side by side instead
// of one after another: the kernel hands the interior samples of a few
// segments to the lanes of a wave, 16 lanes per segment, so a wave stays busy
// while its rays are at different stages (and when only a few of them are
// left).  Each sample's point inversion starts from the cubic Hermite
// interpolant of rst(t) between the segment's ends, which matches rst and
// drst/dt = J^-1 d at both ends and is O(dt^4) from the preimage (the
// sequential scan extrapolates from the two previous samples instead).  The
// samples are the scan's samples, the first sign change per isovalue the
// bracket, so the brackets are the sequential scan's (ties aside: FP32
// decisions within rounding of zero).

#define FEMRT_COOP_MAX_SAMPLES 16

// What the lanes sampling a segment need to know about it.
struct SegInfo {
  int e;
  unsigned open;      // isovalues still searched
  float lip;          // field Lipschitz bound of the element (0: none)
  float t_in, t_out;
  float r_in[3], r_out[3];  // rst at the ends
  float m_in[3], m_out[3];  // drst/dt at the ends
  float o[3], d[3];
};

// drst/dt = J^-1 d at both ends of the segment, for exits whose search left
// no Jacobian behind.
template <int NG, typename W>
FEMRT_HD void seg_tangents(const GeoMode* __restrict__ gc, const float* d, const float* r_in,
                           const float* r_out, float* m_in, float* m_out, W& w)
{
  float X[3], J[9];
  geom_eval<float, NG>(gc, r_in[0], r_in[1], r_in[2], X, J);
  if (!solve3<float>(J, d, m_in)) m_in[0] = m_in[1] = m_in[2] = 0.f;
  geom_eval<float, NG>(gc, r_out[0], r_out[1], r_out[2], X, J);
  if (!solve3<float>(J, d, m_out)) m_out[0] = m_out[1] = m_out[2] = 0.f;
  w.inv_it += 2;
}

// Segment info of the walk state (after march_exit found the exit).
template <int NG, typename W>
FEMRT_HD void seg_info(const TraceMesh& M, const IsoParams& P, const float* o, const float* d,
                       const MarchState& S, SegInfo& g, W& w)
{
  g.e = S.e;
  g.open = ((1u << P.n_iso) - 1u) & ~S.found;
  g.lip = M.lip ? M.lip[S.e] : 0.f;
  g.t_in = S.t_in;
  g.t_out = S.t_out;
  face_to_rst(S.f_in, S.u_in, S.v_in, g.r_in[0], g.r_in[1], g.r_in[2]);
  face_to_rst(S.f_out, S.u_out, S.v_out, g.r_out[0], g.r_out[1], g.r_out[2]);
  for (int c = 0; c < 3; ++c) {
    g.o[c] = o[c];
    g.d[c] = d[c];
  }
  if (P.guess == FEMRT_GUESS_PLAIN) {
    // no interpolated guesses (the Hermite curve of the bracket's start
    // point then has zero end slopes)
    for (int c = 0; c < 3; ++c) g.m_in[c] = g.m_out[c] = 0.f;
    return;
  }
  if (S.have_m) {
    for (int c = 0; c < 3; ++c) {
      g.m_in[c] = S.m_in[c];
      g.m_out[c] = S.m_out[c];
    }
  } else {
    seg_tangents<NG>(M.geom + (size_t)S.e * (NG * NG * NG), d, g.r_in, g.r_out, g.m_in, g.m_out, w);
  }
}

// Whether the scan of the segment can find a crossing: some isovalue still
// open lies within the element's bounds of u (else the scan is skipped).
FEMRT_HD bool seg_may_cross(const TraceMesh& M, const IsoParams& P, const MarchState& S)
{
  return elem_may_cross(M, P, S.e, S.found);
}

// Ray parameter of sample j of [t_in, t_out] (the sequential scan's
// march_sample_t, bit for bit).
FEMRT_HD float seg_sample_t(const SegInfo& g, int j, int samples)
{
  return j == samples ? g.t_out : g.t_in + (g.t_out - g.t_in) * ((float)j / (float)samples);
}

// Cubic Hermite interpolant of rst at fraction s of the segment, clamped to
// the cube.
FEMRT_HD void seg_hermite(const SegInfo& g, float s, float* rst)
{
  const float h = g.t_out - g.t_in;
  const float s2 = s * s, s3 = s2 * s;
  const float h00 = 2.f * s3 - 3.f * s2 + 1.f, h10 = s3 - 2.f * s2 + s;
  const float h01 = 3.f * s2 - 2.f * s3, h11 = s3 - s2;
  for (int c = 0; c < 3; ++c) {
    const float v = h00 * g.r_in[c] + h10 * h * g.m_in[c] + h01 * g.r_out[c] + h11 * h * g.m_out[c];
    rst[c] = fminf(fmaxf(v, -1.f), 1.f);
  }
}

// Point inversion of one scan sample, one Newton iteration per call
// (sample_job_newton), then the decision once u is known (sample_job_after):
// the sequential scan's per-sample rules -- convergence |delta|_inf < 1e-5
// inside the cube (|rst|_inf <= 1.001), early sign acceptance after a step
// below 1e-2 when every open isovalue clears 2 lip |step|, restarts from 0
// (after a network guess) and the octant centres, 50 iterations per run.
struct SampleJob {
  float rst[3], x[3];
  int stage, its;
  bool cert_tried, zero_retry;
};

struct SampleTrip {
  bool fail, conv, try_cert, inside;
  float st;
};

enum { SJ_CONTINUE = 0, SJ_DONE = 1, SJ_FAIL = 2 };

FEMRT_HD void sample_job_begin(SampleJob& J, const SegInfo& g, int j, int samples, int guess)
{
  const float t = seg_sample_t(g, j, samples);
  for (int c = 0; c < 3; ++c) J.x[c] = fmaf(t, g.d[c], g.o[c]);
  if (guess == FEMRT_GUESS_PLAIN)
    J.rst[0] = J.rst[1] = J.rst[2] = 0.f;
  else
    seg_hermite(g, (float)j / (float)samples, J.rst);
  J.stage = 0;
  J.its = 0;
  J.cert_tried = false;
  J.zero_retry = false;
}

template <int NG, typename C>
FEMRT_HD void sample_job_newton(SampleJob& J, const GeoMode* __restrict__ gc, float lip, SampleTrip& T,
                                C& iters)
{
  float X[3], Jm[9], dl[3];
  geom_eval<float, NG>(gc, J.rst[0], J.rst[1], J.rst[2], X, Jm);
  ++iters;
  ++J.its;
  const float bv[3] = {X[0] - J.x[0], X[1] - J.x[1], X[2] - J.x[2]};
  T.fail = !solve3<float>(Jm, bv, dl);
  T.conv = T.try_cert = false;
  T.st = 0.f;
  if (!T.fail) {
    J.rst[0] -= dl[0];
    J.rst[1] -= dl[1];
    J.rst[2] -= dl[2];
    if (!(fabsf(J.rst[0]) < 1e4f && fabsf(J.rst[1]) < 1e4f && fabsf(J.rst[2]) < 1e4f)) {
      T.fail = true;
    } else if (fabsf(dl[0]) < 1e-5f && fabsf(dl[1]) < 1e-5f && fabsf(dl[2]) < 1e-5f) {
      T.conv = true;
    } else {
      T.st = fmaxf(fabsf(dl[0]), fmaxf(fabsf(dl[1]), fabsf(dl[2])));
      T.try_cert = T.st < 1e-2f && !J.cert_tried && lip > 0.f;
      J.cert_tried = J.cert_tried || T.try_cert;
    }
  }
  T.inside = fmaxf(fabsf(J.rst[0]), fmaxf(fabsf(J.rst[1]), fabsf(J.rst[2]))) <= 1.001f;
}

FEMRT_HD bool sample_trip_needs_u(const SampleTrip& T) { return !T.fail && ((T.conv && T.inside) || T.try_cert); }

// u: the field at J.rst when sample_trip_needs_u.  SJ_DONE: u is the sample's
// value; SJ_FAIL: no preimage inside the element; SJ_CONTINUE: next trip.
FEMRT_HD int sample_job_after(SampleJob& J, const SampleTrip& T, float u, const IsoParams& P, unsigned open,
                              float lip)
{
  bool done = T.conv, fail = T.fail;
  if (T.try_cert) {
    const float margin = 2.f * lip * T.st;
    done = true;
    for (int k = 0; k < FEMRT_MAX_ISO; ++k)
      if (k < P.n_iso && ((open >> k) & 1u) && !(fabsf(u - P.iso[k]) > margin)) done = false;
  }
  if (!fail && !done && J.its >= 50) fail = true;
  if (done && !T.inside) {
    done = false;
    fail = true;
  }
  if (done) return SJ_DONE;
  if (!fail) return SJ_CONTINUE;
  ++J.stage;
  if (J.stage == 1 && !J.zero_retry) J.stage = 2;
  if (J.stage > 9) return SJ_FAIL;
  if (J.stage == 1) {
    J.rst[0] = J.rst[1] = J.rst[2] = 0.f;
  } else {
    const int sd = J.stage - 2;
    J.rst[0] = (sd & 1) ? 0.5f : -0.5f;
    J.rst[1] = (sd & 2) ? 0.5f : -0.5f;
    J.rst[2] = (sd & 4) ? 0.5f : -0.5f;
  }
  J.its = 0;
  J.cert_tried = false;
  return SJ_CONTINUE;
}

// The scan of one segment from its sample values u(j) = U[j * ustride],
// j = 0 .. samples (NaN: no preimage inside the element).  Returns true when
// an interior sample has no preimage on the first pass (the exit is wrong:
// the caller sets S.redo and finds the exit again); else records the first
// sign change per open isovalue as the sequential scan does and marks it in
// S.found.
template <typename Sink, typename W>
FEMRT_HD bool seg_scan(const IsoParams& P, const SegInfo& g, MarchState& S, const float* U, int ustride,
                       Sink& sink, W& w)
{
  const int Ms = P.samples;
  if (!S.redo)
    for (int j = 1; j < Ms; ++j)
      if (!(U[j * ustride] == U[j * ustride])) return true;
  float gprev[FEMRT_MAX_ISO];
  float tprev = 0.f;
  bool have_prev = false;
  unsigned found_here = 0;
  for (int j = 0; j <= Ms; ++j) {
    const float u = U[j * ustride];
    if (!(u == u)) {
      have_prev = false;
      continue;
    }
    const float t = seg_sample_t(g, j, Ms);
    for (int k = 0; k < FEMRT_MAX_ISO; ++k) {
      if (k >= P.n_iso || ((S.found | found_here) >> k) & 1u) continue;
      const float gk = u - P.iso[k];
      if (have_prev && gprev[k] * gk <= 0.f) {
        IsoBracket br;
        br.elem = S.e;
        br.ta = tprev;
        br.tb = t;
        br.ga = gprev[k];
        const float den = gprev[k] - gk;
        const float wt = den != 0.f ? fminf(fmaxf(gprev[k] / den, 0.f), 1.f) : 0.5f;
        br.tm = tprev + wt * (t - tprev);
        // rst at the secant point from the segment's interpolant (the FP64
        // endgame converges from there)
        const float h = g.t_out - g.t_in;
        seg_hermite(g, h > 0.f ? fminf(fmaxf((br.tm - g.t_in) / h, 0.f), 1.f) : 0.f, br.rst);
        // the bracket's ends from the same interpolant (network inputs only)
        seg_hermite(g, h > 0.f ? fminf(fmaxf((tprev - g.t_in) / h, 0.f), 1.f) : 0.f, br.seg);
        seg_hermite(g, h > 0.f ? fminf(fmaxf((t - g.t_in) / h, 0.f), 1.f) : 0.f, br.seg + 3);
        br.seg[6] = gprev[k];
        br.seg[7] = gk;
        sink.put(k, br);
        ++w.brackets;
        found_here |= 1u << k;
      }
      gprev[k] = gk;
    }
    tprev = t;
    have_prev = true;
  }
  S.found |= found_here;
  return false;
}

// Next segment through the face adjacency.  Returns MARCH_DONE when the ray
// is done (every isovalue found or the step cap), MARCH_NEXT when S now holds
// the neighbour's segment, MARCH_REENTER when the ray left the mesh (boundary
// face, no adjacency, or no exit resolved) and must find the next element it
// meets through the BVH (march_reenter).
enum { MARCH_DONE = 0, MARCH_NEXT = 1, MARCH_REENTER = 2 };

FEMRT_HD int march_next(const TraceMesh& M, const IsoParams& P, MarchState& S)
{
  const unsigned all = (1u << P.n_iso) - 1u;
  S.redo = false;
  // a walk whose exits keep coming from the marching fallback, with no exit
  // found by prediction or face search in between, has stalled between
  // elements (each such exit can advance the ray by as little as 1e-4 of the
  // element and costs thousands of inversion iterations)
  const int seg = (S.step & 0xFFFF) + 1;
  S.step = (S.step & ~0xFFFF) | seg;
  if (S.found == all || seg >= P.max_march) return MARCH_DONE;
  // a stalled walk skips ahead in its next element (march_exit); with its
  // skips used up it ends
  if (march_stalled(S) && march_skips(S) >= FEMRT_MAX_WALK_SKIPS) return MARCH_DONE;
  if (S.ok && M.adj) {
    const int ne = M.adj[(size_t)S.e * 12 + S.f_out * 2];
    const int code = M.adj[(size_t)S.e * 12 + S.f_out * 2 + 1];
    if (ne >= 0) {
      // entry point on the neighbour's face: the exit (u, v) in the
      // neighbour face's own coordinates (orientation in the upper bits)
      float u = S.u_out, v = S.v_out;
      if (code & 8) {
        const float tmp = u;
        u = v;
        v = tmp;
      }
      if (code & 16) u = -u;
      if (code & 32) v = -v;
      S.e = ne;
      S.f_in = code & 7;
      S.t_in = S.t_out;
      S.u_in = u;
      S.v_in = v;
      return MARCH_NEXT;
    }
  }
  return MARCH_REENTER;
}

// Re-entry after march_next returned MARCH_REENTER: the nearest element face
// the ray meets beyond the last resolved point, by the same traversal as the
// first hit.  The current element is not excluded -- a ray that leaves
// through a curved boundary face can re-enter the same element through it
// further on -- so the search starts 1e-4 of the element beyond the point
// (as the exit search does), where the exit's own root no longer counts.
// With `jump` (march_advance), the ray has just entered S.e at S.t_in from
// a segment that could not cross, and S.e cannot either: the search skips
// every element that cannot hold a crossing of an open isovalue and starts
// at t_jump (just before S.t_in), so it lands on the entry of the next
// element along the ray that can (or finds none, and the ray has no
// crossing left).  Elements it skips have no root to step over, and the
// last element that could cross was left before the skipped segment began,
// more than the step back before t_jump.
// Returns false when there is none.
template <int NG, typename W>
FEMRT_HD bool march_reenter(const TraceMesh& M, const IsoParams& P, const float* o, const float* d,
                            MarchState& S, W& w, bool jump = false, float t_jump = 0.f)
{
  const float t_from = S.ok ? S.t_out : S.t_in;
  const float* bx = M.boxes + (size_t)S.e * FEMRT_BOX_STRIDE;
  const float ext = fmaxf(bx[3] - bx[0], fmaxf(bx[4] - bx[1], bx[5] - bx[2]));
  ++w.reentry;
  w.jumps += jump ? 1u : 0u;
  const float t_lo = jump ? t_jump : t_from + fmaxf(1e-4f * ext, 1e-6f * (1.f + fabsf(t_from)));
  const HitF32 h2 = first_hit_ray<NG>(M, o, d, P.guess, t_lo, 1e30f, w, -1, nullptr, jump ? &P : nullptr,
                                      S.found);
  if (h2.elem < 0) return false;
  S.e = h2.elem;
  S.f_in = h2.face;
  S.t_in = h2.t;
  S.u_in = h2.u;
  S.v_in = h2.v;
  S.ok = false;
  return true;
}

// march_next, then march_reenter when the ray left the mesh.  A segment
// whose scan the field bounds ruled out, leading into a neighbour whose
// bounds rule it out too, starts a run of elements that cannot hold a
// crossing: instead of walking it element by element (an exit search each),
// march_reenter jumps to the next element along the ray that can (P.jump).
// The elements passed hold no crossing of an open isovalue, so the brackets
// are the walk's.  The jump comes after P.jump such segments in a row.  The
// search starts half the exit floor (1e-4 of the element) before the
// segment's exit, so an element entered right at that point (the neighbour
// only clipped) is not missed; a skipped segment shorter than the floor
// walks on instead.
template <int NG, typename W>
FEMRT_HD bool march_advance(const TraceMesh& M, const IsoParams& P, const float* o, const float* d,
                            MarchState& S, W& w)
{
  bool none_here = P.jump > 0 && M.frange && S.ok && !seg_may_cross(M, P, S);
  S.run = none_here ? S.run + 1 : 0;
  float t_jump = 0.f;
  if (none_here) {
    const float* bx = M.boxes + (size_t)S.e * FEMRT_BOX_STRIDE;
    const float eps = 1e-4f * fmaxf(bx[3] - bx[0], fmaxf(bx[4] - bx[1], bx[5] - bx[2]));
    none_here = S.run >= P.jump && S.t_out - S.t_in > 2.f * eps;
    t_jump = S.t_out - 0.5f * eps;
  }
  const int k = march_next(M, P, S);
  if (k == MARCH_DONE) {
    if (march_stalled(S) && march_skips(S) >= FEMRT_MAX_WALK_SKIPS) ++w.stalled;
    return false;
  }
  bool jump = false;
  if (k == MARCH_NEXT) {
    if (!(none_here && !seg_may_cross(M, P, S))) return true;
    jump = true;
  }
  return march_reenter<NG>(M, P, o, d, S, w, jump, t_jump);
}

struct BracketArray {
  IsoBracket* out;
  FEMRT_HD void put(int k, const IsoBracket& b) { out[k] = b; }
};

// Sample values of one segment for the cooperative scan, one sample after
// another (the kernel spreads them over lanes): U[j], j = 0 .. samples, NaN
// where a sample has no preimage inside the element.
template <int NG, int NF, typename W>
FEMRT_HD void seg_sample_values(const TraceMesh& M, const IsoParams& P, const SegInfo& g, float* U, W& w)
{
  const GeoMode* gc = M.geom + (size_t)g.e * (NG * NG * NG);
  const float* fc = M.field + (size_t)g.e * (NF * NF * NF);
  const int Ms = P.samples;
  U[0] = field_value<float, NF>(fc, g.r_in[0], g.r_in[1], g.r_in[2]);
  U[Ms] = field_value<float, NF>(fc, g.r_out[0], g.r_out[1], g.r_out[2]);
  w.field += 2;
  for (int j = 1; j < Ms; ++j) {
    SampleJob J;
    sample_job_begin(J, g, j, Ms, P.guess);
    for (;;) {
      SampleTrip T;
      sample_job_newton<NG>(J, gc, g.lip, T, w.inv_it);
      float u = 0.f;
      if (sample_trip_needs_u(T)) {
        u = field_value<float, NF>(fc, J.rst[0], J.rst[1], J.rst[2]);
        ++w.field;
      }
      const int rc = sample_job_after(J, T, u, P, g.open, g.lip);
      if (rc == SJ_CONTINUE) continue;
      U[j] = rc == SJ_DONE ? u : NAN;
      break;
    }
  }
}

// march_ray with the cooperative scan's per-segment rules, one lane.
template <int NG, int NF, typename W>
FEMRT_HD unsigned march_ray_coop(const TraceMesh& M, const IsoParams& P, const float* o, const float* d,
                                 HitF32 h, IsoBracket* out, W& w)
{
  if (P.n_iso <= 0 || P.max_march <= 0 || P.samples > FEMRT_COOP_MAX_SAMPLES) return 0;
  float id[3];
  for (int c = 0; c < 3; ++c) id[c] = 1.f / d[c];
  MarchState S;
  march_begin(S, h);
  BracketArray sink{out};
  float U[FEMRT_COOP_MAX_SAMPLES + 1];
  for (;;) {
    march_exit<NG>(M, P, o, d, id, S, w);
    if (S.ok) {
      if (seg_may_cross(M, P, S)) {
        SegInfo g;
        seg_info<NG>(M, P, o, d, S, g, w);
        seg_sample_values<NG, NF>(M, P, g, U, w);
        if (seg_scan(P, g, S, U, 1, sink, w)) {
          S.redo = true;
          continue;
        }
      } else {
        ++w.skipped;
      }
    }
    if (!march_advance<NG>(M, P, o, d, S, w)) break;
  }
  return S.found;
}

// The whole walk of one ray, one segment after another (host harnesses and
// single-lane use).  Returns the mask of isovalues found.
template <int NG, int NF, typename W>
FEMRT_HD unsigned march_ray(const TraceMesh& M, const IsoParams& P, const float* o, const float* d,
                            HitF32 h, IsoBracket* out, W& w)
{
  if (P.n_iso <= 0 || P.max_march <= 0) return 0;
  float id[3];
  for (int c = 0; c < 3; ++c) id[c] = 1.f / d[c];
  MarchState S;
  march_begin(S, h);
  BracketArray sink{out};
  const NoGuess ng;
  for (;;) {
    march_exit<NG>(M, P, o, d, id, S, w);
    if (S.ok && march_samples<NG, NF>(M, P, o, d, S, ng, sink, w)) {
      S.redo = true;
      continue;
    }
    if (!march_advance<NG>(M, P, o, d, S, w)) break;
  }
  return S.found;
}

// FP64 endgame of one isovalue bracket: point inversion in FP64 (warm from
// the FP32 estimate) and safeguarded Newton on g(t) = u - iso with
// g'(t) = grad_world(u) . d, kept inside the sign-change bracket; bisection
// when the Newton step leaves it.  Converged when the step or the bracket is
// below 1e-13 (1 + |t|).
template <int NG, int NF, typename W>
FEMRT_HD bool polish_iso(const TraceMesh& M, const double* o, const double* d, double iso,
                         const IsoBracket& br, double& t_out, double* rst_out, double& u_out,
                         W& w)
{
  const GeoMode* gc = M.geom + (size_t)br.elem * (NG * NG * NG);
  const float* fc = M.field + (size_t)br.elem * (NF * NF * NF);
  {
    double tau = br.tm, r4[3] = {br.rst[0], br.rst[1], br.rst[2]}, u4;
    if (iso_newton<double, NG, NF>(gc, fc, o, d, iso, (double)br.ta, (double)br.tb, 10, 1e-13, 1e-9,
                                   tau, r4, u4, w.geo64, w.field64)) {
      t_out = tau;
      for (int c = 0; c < 3; ++c) rst_out[c] = r4[c];
      u_out = u4;
      return true;
    }
  }
  double lo = br.ta, hi = br.tb, t = br.tm;
  const bool lo_neg = br.ga < 0.f;
  double rst[3] = {br.rst[0], br.rst[1], br.rst[2]}, J[9];
  double u = 0.0, g_last = 0.0, gp_last = 0.0;
  bool ok = false;
  // FP64 preimage of x in the element: warm from the last iterate, else
  // from the bracket's FP32 estimate, else from the octant centres
  auto locate = [&](const double* x) -> bool {
    if (invert_f64<NG>(gc, x, rst, J, w.geo64)) return true;
    for (int s = -1; s < 8; ++s) {
      for (int c = 0; c < 3; ++c)
        rst[c] = s < 0 ? (double)br.rst[c] : (((s >> c) & 1) ? 0.5 : -0.5);
      if (invert_f64<NG>(gc, x, rst, J, w.geo64) &&
          fmax(fabs(rst[0]), fmax(fabs(rst[1]), fabs(rst[2]))) <= 1.001)
        return true;
    }
    return false;
  };
  for (int it = 0; it < 60; ++it) {
    const double x[3] = {o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]};
    if (!locate(x)) {
      ok = false;  // no root is reported from an iterate that is not converged
      break;
    }
    double gref[3], gw[3];
    field_eval<double, NF>(fc, rst[0], rst[1], rst[2], &u, gref);
    ++w.field64;
    grad_world<double>(J, gref, gw);
    const double g = u - iso;
    const double gp = gw[0] * d[0] + gw[1] * d[1] + gw[2] * d[2];
    g_last = g;
    gp_last = gp;
    ok = true;
    t_out = t;
    for (int c = 0; c < 3; ++c) rst_out[c] = rst[c];
    u_out = u;
    if (g == 0.0) break;
    if ((g < 0.0) == lo_neg) lo = t; else hi = t;
    double tn = (gp != 0.0) ? t - g / gp : 0.5 * (lo + hi);
    if (!(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
    const double step = fabs(tn - t);
    t = tn;
    if (step < 1e-13 * (1.0 + fabs(t)) || (hi - lo) < 1e-13 * (1.0 + fabs(t))) {
      const double x2[3] = {o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]};
      if (locate(x2)) {
        u_out = field_value<double, NF>(fc, rst[0], rst[1], rst[2]);
        ++w.field64;
        t_out = t;
        for (int c = 0; c < 3; ++c) rst_out[c] = rst[c];
        g_last = u_out - iso;
      }
      break;
    }
  }
  // a root only where g vanishes to the distance scale: a bracket that the
  // FP32 scan saw change sign at a tangency, but whose FP64 values keep one
  // sign, narrows onto one of its ends with g away from zero -- no crossing
  if (ok && !(g_last == 0.0 || fabs(g_last) <= 1e-9 * (1.0 + fabs(t_out)) * fabs(gp_last))) ok = false;
  return ok;
}

// Tolerances of the bisection endgames: stop when the bracket is narrower
// than ISO_BISECT_TREL (1 + |t|) and |g| at its midpoint is at most
// ISO_BISECT_GTOL, or the bracket is narrower than 1e-13 (1 + |t|).  The
// midpoint is then within 1e-10 relative of the crossing (its reference
// coordinates within |J^-1 d| 1e-10 (1 + |t|), below 1e-7 for elements down
// to 1/64 of the unit cube) and u within 1e-8 of the isovalue: inside the
// parity bar (t to 1e-6 relative, rst to 1e-7, u to 1e-6).
#define ISO_BISECT_TREL 1e-10
#define ISO_BISECT_GTOL 1e-8

// FP64 bisection endgame of one bracket (fem/RayIsosurface.cu:121-171 is
// the reference's Newton-in-bracket / bisection).  Each step evaluates g =
// u - iso at the bracket's midpoint (FP64 point inversion warm from the
// previous step's preimage, then the field).  With `pred` (the bracket
// network's [lo, hi] in ray parameter, learned bisection), the predicted ends
// that lie inside the scan's bracket are evaluated first and each narrows
// the bracket when g there keeps the sign of the end it replaces (only when
// the prediction is narrower than a quarter of the bracket); the crossing
// stays the one the scan bracketed.  The pair count of FP64 field
// evaluations per bracket is what the bracket network saves.
template <int NG, int NF, typename W>
FEMRT_HD bool polish_iso_bisect(const TraceMesh& M, const double* o, const double* d, double iso,
                                const IsoBracket& br, bool use_pred, float pred_lo, float pred_hi,
                                double& t_out, double* rst_out, double& u_out, W& w)
{
  const GeoMode* gc = M.geom + (size_t)br.elem * (NG * NG * NG);
  const float* fc = M.field + (size_t)br.elem * (NF * NF * NF);
  double lo = br.ta, hi = br.tb;
  const bool lo_neg = br.ga < 0.f;
  double rst[3] = {br.rst[0], br.rst[1], br.rst[2]}, J[9];
  auto locate = [&](const double* x) -> bool {
    if (invert_f64<NG>(gc, x, rst, J, w.geo64) && fmax(fabs(rst[0]), fmax(fabs(rst[1]), fabs(rst[2]))) <= 1.001)
      return true;
    for (int s = -1; s < 8; ++s) {
      for (int c = 0; c < 3; ++c)
        rst[c] = s < 0 ? (double)br.rst[c] : (((s >> c) & 1) ? 0.5 : -0.5);
      if (invert_f64<NG>(gc, x, rst, J, w.geo64) &&
          fmax(fabs(rst[0]), fmax(fabs(rst[1]), fabs(rst[2]))) <= 1.001)
        return true;
    }
    return false;
  };
  // g at t; false where t has no preimage in the element
  auto g_at = [&](double t, double& g) -> bool {
    const double x[3] = {o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]};
    if (!locate(x)) return false;
    u_out = field_value<double, NF>(fc, rst[0], rst[1], rst[2]);
    ++w.field64;
    g = u_out - iso;
    t_out = t;
    for (int c = 0; c < 3; ++c) rst_out[c] = rst[c];
    return true;
  };
  double g;
  const double pl = fmin(pred_lo, pred_hi), ph = fmax(pred_lo, pred_hi);
  // the two evaluations at the predicted ends pay off only when they narrow
  // the bracket by more than the two bisection steps they cost
  if (use_pred && ph - pl < 0.25 * (hi - lo)) {
    if (pl > lo && pl < hi && g_at(pl, g)) {
      if (g == 0.0) return true;
      if ((g < 0.0) == lo_neg) lo = pl;
    }
    if (ph > lo && ph < hi && g_at(ph, g)) {
      if (g == 0.0) return true;
      if ((g < 0.0) != lo_neg) hi = ph;
    }
  }
  for (int it = 0; it < 96; ++it) {
    const double t = 0.5 * (lo + hi);
    if (!g_at(t, g)) return false;
    const double scale = 1.0 + fabs(t);
    if (g == 0.0 || ((hi - lo) <= ISO_BISECT_TREL * scale && fabs(g) <= ISO_BISECT_GTOL)) return true;
    // narrowed to the distance scale with g still away from zero: the FP32
    // scan saw a sign change that the FP64 values do not have (a tangency),
    // so there is no crossing, as in the Newton endgame
    if ((hi - lo) <= 1e-13 * scale) return fabs(g) <= ISO_BISECT_GTOL;
    if ((g < 0.0) == lo_neg) lo = t; else hi = t;
  }
  return false;
}

}  // namespace femrt
