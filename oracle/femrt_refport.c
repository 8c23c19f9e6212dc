/*
 * femrt_refport.c -- reference-faithful FP32 CPU restatement of the
 * reference's per-ray hex path, used only as the CPU baseline that bench.py
 * times beside the GPU (cpu_baseline.kind = "port").
 *
 * TEST INFRASTRUCTURE ONLY: never linked into or called by the product.
 *
 * It keeps the reference's cost structure on purpose (paths relative to
 * FEM-AABB/ of the reference tree):
 *   - modal sums as direct triple loops with the 1-D basis recomputed per
 *     term by the Jacobi recurrence (fem/HexElement.cu:8-94,
 *     fem/ModalBasis.cuh:16-70), the Jacobian as 9 separate passes
 *     (HexElement.cu:194-206);
 *   - the element box re-sampled per ray per leaf (raytracer.cu:199,
 *     ComputeHexAabb high_order.h:51-113) and again inside the hex
 *     intersector (high_order.h:420-427);
 *   - point inversion from rst = 0 (HexElement.cu:227-272);
 *   - the iso search over the padded box segment, 16 sub-intervals, Newton/
 *     bisection refinement (fem/RayIsosurface.cu:67-188); when it finds no
 *     root the face intersector runs (raytracer.cu:207-252);
 *   - 6 faces x 9 seeds x <= 50 face-Newton iterations, first converging seed
 *     returned (high_order.h:189-304, 409-465).
 * One isovalue per pass, as the reference supports; the baseline runs one
 * pass per isovalue.  Candidate elements come from a linear scan of the
 * loader-time boxes sorted by entry distance and culled by the running
 * nearest hit, which visits elements in the order a near-first BVH would.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define RP_EXPORT __attribute__((visibility("default")))

typedef struct { float x, y, z; } f3;

static float jacobiP(int n, float a, float b, float x)
{
  if (n == 0) return 1.0f;
  if (n == 1) return 0.5f * ((2.0f + a + b) * x + (a - b));
  float p0 = 1.0f, p1 = 0.5f * ((2.0f + a + b) * x + (a - b)), pk = 0.0f;
  for (int k = 2; k <= n; ++k) {
    const float kf = (float)k, c = 2.0f * kf + a + b;
    const float den = 2.0f * kf * (kf + a + b) * (c - 2.0f);
    pk = ((c - 1.0f) * (c * (c - 2.0f) * x + (a * a - b * b)) * p1 -
          2.0f * (kf + a - 1.0f) * (kf + b - 1.0f) * c * p0) / den;
    p0 = p1;
    p1 = pk;
  }
  return pk;
}

static float modA(unsigned i, float x)
{
  if (i == 0) return 0.5f * (1.0f - x);
  if (i == 1) return 0.5f * (1.0f + x);
  return 0.25f * (1.0f - x * x) * jacobiP((int)i - 2, 1.0f, 1.0f, x);
}

static float modAd(unsigned i, float x)
{
  if (i == 0) return -0.5f;
  if (i == 1) return 0.5f;
  const int n = (int)i - 2;
  const float p = jacobiP(n, 1.0f, 1.0f, x);
  const float dp = n == 0 ? 0.0f : 0.5f * (n + 3.0f) * jacobiP(n - 1, 2.0f, 2.0f, x);
  return -0.5f * x * p + 0.25f * (1.0f - x * x) * dp;
}

/* which: 0 value, 1 d/dr, 2 d/ds, 3 d/dt */
static float modal3(const float* c, int n, f3 rst, int which)
{
  float v = 0.0f;
  int idx = 0;
  for (int k = 0; k < n; ++k) {
    const float tk = which == 3 ? modAd(k, rst.z) : modA(k, rst.z);
    for (int j = 0; j < n; ++j) {
      const float sj = which == 2 ? modAd(j, rst.y) : modA(j, rst.y);
      for (int i = 0; i < n; ++i) {
        const float ri = which == 1 ? modAd(i, rst.x) : modA(i, rst.x);
        v += c[idx++] * ri * sj * tk;
      }
    }
  }
  return v;
}

typedef struct {
  int ng, nf;
  const float* g[3];
  const float* f;
} elem_t;

static f3 ref2world(const elem_t* e, f3 rst)
{
  f3 p = {modal3(e->g[0], e->ng, rst, 0), modal3(e->g[1], e->ng, rst, 0), modal3(e->g[2], e->ng, rst, 0)};
  return p;
}

static void jac(const elem_t* e, f3 rst, float J[9])
{
  for (int c = 0; c < 3; ++c)
    for (int w = 0; w < 3; ++w) J[3 * c + w] = modal3(e->g[c], e->ng, rst, w + 1);
}

static void inv3(const float J[9], float I[9])
{
  const float det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                    J[2] * (J[3] * J[7] - J[4] * J[6]);
  const float id = 1.0f / det;
  I[0] = (J[4] * J[8] - J[5] * J[7]) * id;
  I[1] = (J[2] * J[7] - J[1] * J[8]) * id;
  I[2] = (J[1] * J[5] - J[2] * J[4]) * id;
  I[3] = (J[5] * J[6] - J[3] * J[8]) * id;
  I[4] = (J[0] * J[8] - J[2] * J[6]) * id;
  I[5] = (J[2] * J[3] - J[0] * J[5]) * id;
  I[6] = (J[3] * J[7] - J[4] * J[6]) * id;
  I[7] = (J[1] * J[6] - J[0] * J[7]) * id;
  I[8] = (J[0] * J[4] - J[1] * J[3]) * id;
}

static int w2r(const elem_t* e, f3 x, f3* rst)
{
  *rst = (f3){0, 0, 0};
  for (int it = 0; it < 50; ++it) {
    const f3 m = ref2world(e, *rst);
    const f3 d = {m.x - x.x, m.y - x.y, m.z - x.z};
    float J[9], I[9];
    jac(e, *rst, J);
    inv3(J, I);
    const float a = I[0] * d.x + I[1] * d.y + I[2] * d.z;
    const float b = I[3] * d.x + I[4] * d.y + I[5] * d.z;
    const float c = I[6] * d.x + I[7] * d.y + I[8] * d.z;
    rst->x -= a;
    rst->y -= b;
    rst->z -= c;
    if (fabsf(a) < 1e-5f && fabsf(b) < 1e-5f && fabsf(c) < 1e-5f) return 1;
  }
  return 0;
}

static void box_sampled(const elem_t* e, int S, f3* lo, f3* hi)
{
  const int ord = e->ng;
  if (S < ord + 2) S = ord + 2;
  *lo = (f3){FLT_MAX, FLT_MAX, FLT_MAX};
  *hi = (f3){-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int k = 0; k < S; ++k)
    for (int j = 0; j < S; ++j)
      for (int i = 0; i < S; ++i) {
        const f3 r = {-1.0f + 2.0f * i / (S - 1), -1.0f + 2.0f * j / (S - 1), -1.0f + 2.0f * k / (S - 1)};
        const f3 p = ref2world(e, r);
        lo->x = fminf(lo->x, p.x); lo->y = fminf(lo->y, p.y); lo->z = fminf(lo->z, p.z);
        hi->x = fmaxf(hi->x, p.x); hi->y = fmaxf(hi->y, p.y); hi->z = fmaxf(hi->z, p.z);
      }
  const float px = fmaxf((hi->x - lo->x) * 0.15f, 1e-3f), py = fmaxf((hi->y - lo->y) * 0.15f, 1e-3f),
              pz = fmaxf((hi->z - lo->z) * 0.15f, 1e-3f);
  lo->x -= px; lo->y -= py; lo->z -= pz;
  hi->x += px; hi->y += py; hi->z += pz;
}

static int slab(f3 o, f3 d, f3 lo, f3 hi, float tmin, float tmax, float* te, float* tx)
{
  float a = tmin, b = tmax;
  const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z}, l[3] = {lo.x, lo.y, lo.z},
              h[3] = {hi.x, hi.y, hi.z};
  for (int c = 0; c < 3; ++c) {
    if (fabsf(dd[c]) < 1e-8f) {
      if (oo[c] < l[c] || oo[c] > h[c]) return 0;
      continue;
    }
    const float inv = 1.0f / dd[c];
    float t0 = (l[c] - oo[c]) * inv, t1 = (h[c] - oo[c]) * inv;
    if (t0 > t1) { const float s = t0; t0 = t1; t1 = s; }
    a = fmaxf(a, t0);
    b = fminf(b, t1);
    if (a > b) return 0;
  }
  *te = a;
  *tx = b;
  return b >= a;
}

static const int FIX[6] = {2, 1, 0, 1, 0, 2}, UA[6] = {0, 0, 1, 0, 1, 0}, VA[6] = {1, 2, 2, 2, 2, 1};
static const float FV[6] = {-1, -1, 1, 1, -1, 1};

static int face_hit(const elem_t* e, f3 o, f3 d, int f, float tmin, float tmax, float* th, f3* rh)
{
  const float sv[3] = {-0.8f, 0.0f, 0.8f};
  for (int sj = 0; sj < 3; ++sj)
    for (int si = 0; si < 3; ++si) {
      float rst[3] = {0, 0, 0};
      rst[FIX[f]] = FV[f];
      rst[UA[f]] = sv[si];
      rst[VA[f]] = sv[sj];
      const f3 X0 = ref2world(e, (f3){rst[0], rst[1], rst[2]});
      const float dd = d.x * d.x + d.y * d.y + d.z * d.z;
      float tau = dd > 0 ? ((X0.x - o.x) * d.x + (X0.y - o.y) * d.y + (X0.z - o.z) * d.z) / dd : 0;
      tau = fminf(fmaxf(tau, tmin), tmax);
      float prev = FLT_MAX;
      for (int it = 0; it < 50; ++it) {
        const f3 rv = {rst[0], rst[1], rst[2]};
        const f3 X = ref2world(e, rv);
        const f3 R = {X.x - (o.x + tau * d.x), X.y - (o.y + tau * d.y), X.z - (o.z + tau * d.z)};
        const float res = fmaxf(fabsf(R.x), fmaxf(fabsf(R.y), fabsf(R.z)));
        if (res < 1e-4f) {
          if (rst[UA[f]] >= -1.001f && rst[UA[f]] <= 1.001f && rst[VA[f]] >= -1.001f &&
              rst[VA[f]] <= 1.001f && tau >= tmin && tau <= tmax) {
            *th = tau;
            *rh = rv;
            return 1;
          }
          break;
        }
        if (it > 5 && res > prev * 1.5f) break;
        prev = res;
        float Jf[9], A[9], I[9];
        jac(e, rv, Jf);
        const int u = UA[f], v = VA[f];
        A[0] = Jf[u]; A[1] = Jf[v]; A[2] = -d.x;
        A[3] = Jf[3 + u]; A[4] = Jf[3 + v]; A[5] = -d.y;
        A[6] = Jf[6 + u]; A[7] = Jf[6 + v]; A[8] = -d.z;
        inv3(A, I);
        const float du = -(I[0] * R.x + I[1] * R.y + I[2] * R.z);
        const float dv = -(I[3] * R.x + I[4] * R.y + I[5] * R.z);
        const float dt = -(I[6] * R.x + I[7] * R.y + I[8] * R.z);
        const float mx = fmaxf(fabsf(du), fmaxf(fabsf(dv), fabsf(dt)));
        const float sc = mx > 0.25f ? 0.25f / mx : 1.0f;
        rst[u] += du * sc;
        rst[v] += dv * sc;
        tau = fminf(fmaxf(tau + dt * sc, tmin), tmax);
        rst[u] = fminf(fmaxf(rst[u], -1.5f), 1.5f);
        rst[v] = fminf(fmaxf(rst[v], -1.5f), 1.5f);
        rst[FIX[f]] = FV[f];
      }
    }
  return 0;
}

static int hex_hit(const elem_t* e, f3 o, f3 d, float tmin, float tmax, float* th)
{
  f3 lo, hi;
  box_sampled(e, 3, &lo, &hi);
  float a, b;
  if (!slab(o, d, lo, hi, tmin, tmax, &a, &b)) return 0;
  int any = 0;
  float best = tmax;
  for (int f = 0; f < 6; ++f) {
    float t;
    f3 r;
    if (face_hit(e, o, d, f, tmin, tmax, &t, &r) && t > tmin && t < best) {
      best = t;
      any = 1;
    }
  }
  if (any) *th = best;
  return any;
}

static int eval_along(const elem_t* e, f3 o, f3 d, float t, float iso, float* val, float* der)
{
  const f3 x = {o.x + t * d.x, o.y + t * d.y, o.z + t * d.z};
  f3 rst;
  if (!w2r(e, x, &rst)) return 0;
  const float u = modal3(e->f, e->nf, rst, 0);
  float gr[3] = {modal3(e->f, e->nf, rst, 1), modal3(e->f, e->nf, rst, 2), modal3(e->f, e->nf, rst, 3)};
  float J[9], I[9];
  jac(e, rst, J);
  inv3(J, I);
  const float gx = I[0] * gr[0] + I[3] * gr[1] + I[6] * gr[2];
  const float gy = I[1] * gr[0] + I[4] * gr[1] + I[7] * gr[2];
  const float gz = I[2] * gr[0] + I[5] * gr[1] + I[8] * gr[2];
  *val = u - iso;
  *der = gx * d.x + gy * d.y + gz * d.z;
  return 1;
}

static int iso_find(const elem_t* e, f3 o, f3 d, float ta, float tb, float iso, float* th)
{
  float pt = ta, pv = 0;
  int have = 0;
  for (int i = 0; i <= 16; ++i) {
    const float t = ta + (tb - ta) * ((float)i / 16);
    float v, dv;
    if (!eval_along(e, o, d, t, iso, &v, &dv)) continue;
    if (!have) { pt = t; pv = v; have = 1; continue; }
    if (pv * v <= 0.0f) {
      float lt = pt, lv = pv, rt = t;
      for (int it = 0; it < 20; ++it) {
        const float mt = 0.5f * (lt + rt);
        float mv, md;
        if (!eval_along(e, o, d, mt, iso, &mv, &md)) return 0;
        float ct = mt;
        if (fabsf(md) > 1e-8f) {
          const float nt = mt - mv / md;
          if (nt > lt && nt < rt) ct = nt;
        }
        float cv, cd;
        if (!eval_along(e, o, d, ct, iso, &cv, &cd)) { ct = mt; cv = mv; }
        if (fabsf(cv) < 1e-5f) { *th = ct; return 1; }
        if (lv * cv <= 0.0f) rt = ct; else { lt = ct; lv = cv; }
      }
      *th = 0.5f * (lt + rt);
      return 1;
    }
    pt = t;
    pv = v;
  }
  return 0;
}

typedef struct { float t; int e; } cand_t;

static int cmp_cand(const void* a, const void* b)
{
  const float x = ((const cand_t*)a)->t, y = ((const cand_t*)b)->t;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* Loader-time boxes (ComputeHexAabb with 6 samples per axis,
 * xml_fld_loader.cpp:271-274): out n_elem x 6 floats. */
RP_EXPORT void rp_loader_boxes(int n_elem, int ng, int nf, const float* geom, float* out)
{
  const size_t n3 = (size_t)ng * ng * ng;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n_elem; ++i) {
    elem_t e = {ng, nf, {geom + ((size_t)i * 3) * n3, geom + ((size_t)i * 3 + 1) * n3, geom + ((size_t)i * 3 + 2) * n3}, 0};
    f3 lo, hi;
    box_sampled(&e, 6, &lo, &hi);
    float* o = out + (size_t)i * 6;
    o[0] = lo.x; o[1] = lo.y; o[2] = lo.z; o[3] = hi.x; o[4] = hi.y; o[5] = hi.z;
  }
}

/* Trace rays[idx[i]] for i < n_sel with the reference's per-ray algorithm,
 * once per isovalue.  Writes the nearest t per (ray, pass) (FLT_MAX on miss).
 * Returns the number of rays processed. */
RP_EXPORT int rp_trace(int n_elem, int ng, int nf, const float* geom, const float* field,
                       const float* boxes, const float* rays, int n_rays, const int* idx, int n_sel,
                       const float* isos, int n_iso, float* t_out)
{
  const size_t n3 = (size_t)ng * ng * ng, f3n = (size_t)nf * nf * nf;
#pragma omp parallel for schedule(dynamic, 1)
  for (int s = 0; s < n_sel; ++s) {
    const int r = idx[s];
    const f3 o = {rays[r], rays[(size_t)n_rays + r], rays[2 * (size_t)n_rays + r]};
    const f3 d = {rays[3 * (size_t)n_rays + r], rays[4 * (size_t)n_rays + r], rays[5 * (size_t)n_rays + r]};
    cand_t* cand = (cand_t*)malloc(sizeof(cand_t) * 4096);
    int nc = 0;
    for (int i = 0; i < n_elem && nc < 4096; ++i) {
      const float* b = boxes + (size_t)i * 6;
      float a0, a1;
      if (slab(o, d, (f3){b[0], b[1], b[2]}, (f3){b[3], b[4], b[5]}, 0.001f, FLT_MAX, &a0, &a1))
        cand[nc++] = (cand_t){a0, i};
    }
    qsort(cand, nc, sizeof(cand_t), cmp_cand);
    for (int k = 0; k < (n_iso > 0 ? n_iso : 1); ++k) {
      float rec = FLT_MAX;
      for (int c = 0; c < nc; ++c) {
        if (cand[c].t >= rec) break;
        const int ei = cand[c].e;
        elem_t e = {ng, nf, {geom + ((size_t)ei * 3) * n3, geom + ((size_t)ei * 3 + 1) * n3, geom + ((size_t)ei * 3 + 2) * n3},
                    field + (size_t)ei * f3n};
        f3 lo, hi;
        box_sampled(&e, 4, &lo, &hi);  /* per ray, per leaf (raytracer.cu:199) */
        float te, tx;
        if (!slab(o, d, lo, hi, 0.001f, rec, &te, &tx)) continue;
        float th;
        if (n_iso > 0 && iso_find(&e, o, d, te, tx, isos[k], &th)) {
          if (th > 0.001f && th < rec) rec = th;
        } else if (hex_hit(&e, o, d, 0.001f, rec, &th)) {
          if (th > 0.001f && th < rec) rec = th;
        }
      }
      t_out[(size_t)s * (n_iso > 0 ? n_iso : 1) + k] = rec;
    }
    free(cand);
  }
  return n_sel;
}
