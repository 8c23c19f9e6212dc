/*
 * femrt_oracle.c -- FP64 CPU restatement of the curved-hex ray-tracing method.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: it is
 * compiled into oracle/_build/liboracle.so and loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
 * checker and never as the thing measured or shipped.  The product path
 * (test-_amd/csrc) never links or calls it.
 *
 * Every function restates the reference algorithm in double precision from the
 * mathematics, citing the reference file:line it follows (paths relative to
 * FEM-AABB/ of the reference tree).  Where the reference is ill-defined the
 * oracle fixes a semantics (documented at each function):
 *   - face intersection returns the NEAREST in-bounds root reached from the
 *     reference's 3x3 seed grid (the reference returns the first seed that
 *     converges, high_order.h:213-250);
 *   - the iso search runs over the ray's true segment inside each element
 *     (entry/exit face roots), not over the padded AABB, with the reference's
 *     16 uniform sub-intervals and a bisection endgame to 1e-14;
 *   - isovalue 0.0 is a valid isovalue (the reference gates on iso != 0,
 *     raytracer.cu:207);
 *   - the geometry first hit and the first hit of each isovalue are separate
 *     outputs.
 * Coefficients arrive as float (the product stores FP32 coefficients, as the
 * reference does, raytracer.h:73-90) and are widened exactly to double, so the
 * oracle solves the same problem as the GPU path in exact-ish arithmetic.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#define ORC_EXPORT __attribute__((visibility("default")))

/* Acceptance slack on the free face coordinates (u,v): a root is on the face
 * when max(|u|,|v|) <= 1 + ORC_FACE_ACC.  The reference uses 1.001
 * (high_order.h:243-245). */
#define ORC_FACE_ACC 1e-4
#define ORC_MAXN 12

/* ------------------------------------------------------------------------ */
/* A1  Modified_A 1-D basis: fem/ModalBasis.cuh:16-70                        */
/* ------------------------------------------------------------------------ */

/* Jacobi polynomial P_n^{(a,b)}(x) by the standard three-term recurrence
 * (Karniadakis & Sherwin, Spectral/hp Element Methods, App. A). */
static double jacobi(int n, double a, double b, double x)
{
  if (n == 0) return 1.0;
  double p0 = 1.0;
  double p1 = 0.5 * (a - b + (a + b + 2.0) * x);
  for (int k = 2; k <= n; ++k) {
    const double c = 2.0 * k + a + b;
    const double a1 = 2.0 * k * (k + a + b) * (c - 2.0);
    const double a2 = (c - 1.0) * (a * a - b * b);
    const double a3 = (c - 2.0) * (c - 1.0) * c;
    const double a4 = 2.0 * (k + a - 1.0) * (k + b - 1.0) * c;
    const double p2 = ((a2 + a3 * x) * p1 - a4 * p0) / a1;
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

/* d/dx P_n^{(a,b)} = (n+a+b+1)/2 * P_{n-1}^{(a+1,b+1)}  (ModalBasis.cuh:43-51) */
static double jacobi_d(int n, double a, double b, double x)
{
  if (n == 0) return 0.0;
  return 0.5 * (n + a + b + 1.0) * jacobi(n - 1, a + 1.0, b + 1.0, x);
}

/* phi_0=(1-x)/2, phi_1=(1+x)/2, phi_p=(1-x)/2 (1+x)/2 P_{p-2}^{(1,1)}(x)
 * (ModalBasis.cuh:53-60) */
ORC_EXPORT double orc_modA(int i, double x)
{
  if (i == 0) return 0.5 * (1.0 - x);
  if (i == 1) return 0.5 * (1.0 + x);
  return 0.25 * (1.0 - x) * (1.0 + x) * jacobi(i - 2, 1.0, 1.0, x);
}

/* derivative of phi_i (ModalBasis.cuh:62-70) */
ORC_EXPORT double orc_modA_d(int i, double x)
{
  if (i == 0) return -0.5;
  if (i == 1) return 0.5;
  return -0.5 * x * jacobi(i - 2, 1.0, 1.0, x) +
         0.25 * (1.0 - x) * (1.0 + x) * jacobi_d(i - 2, 1.0, 1.0, x);
}

/* ------------------------------------------------------------------------ */
/* A2/A6  Modal tensor sums, brute force: fem/HexElement.cu:8-94, 280-306;  */
/* coefficient order p (r) fastest: idx = (k*n + j)*n + i                    */
/* (StdHexExp.cpp:516-532).  Deliberately NOT sum-factorised: the GPU path   */
/* is, so this is an independent check of the factorisation.                */
/* ------------------------------------------------------------------------ */
ORC_EXPORT void orc_eval_modal(const float* c, int n, const double* rst, double* out4)
{
  double pr[ORC_MAXN], dpr[ORC_MAXN], ps[ORC_MAXN], dps[ORC_MAXN], pt[ORC_MAXN], dpt[ORC_MAXN];
  for (int i = 0; i < n; ++i) {
    pr[i] = orc_modA(i, rst[0]);
    dpr[i] = orc_modA_d(i, rst[0]);
    ps[i] = orc_modA(i, rst[1]);
    dps[i] = orc_modA_d(i, rst[1]);
    pt[i] = orc_modA(i, rst[2]);
    dpt[i] = orc_modA_d(i, rst[2]);
  }
  double v = 0.0, dr = 0.0, ds = 0.0, dt = 0.0;
  int idx = 0;
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const double cc = (double)c[idx++];
        v += cc * pr[i] * ps[j] * pt[k];
        dr += cc * dpr[i] * ps[j] * pt[k];
        ds += cc * pr[i] * dps[j] * pt[k];
        dt += cc * pr[i] * ps[j] * dpt[k];
      }
  out4[0] = v;
  out4[1] = dr;
  out4[2] = ds;
  out4[3] = dt;
}

/* ------------------------------------------------------------------------ */
/* Mesh as the oracle sees it: uniform mode counts, FP32 coefficients.      */
/* geom: n_elem x 3 x ng^3 (x block, y block, z block per element)          */
/* field: n_elem x nf^3                                                     */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n_elem;
  int ng;
  int nf;
  int pad_;
  const float* geom;
  const float* field;
} orc_mesh_t;

static inline const float* geom_ptr(const orc_mesh_t* m, int e, int comp)
{
  const size_t n3 = (size_t)m->ng * m->ng * m->ng;
  return m->geom + ((size_t)e * 3 + comp) * n3;
}

/* A3/A4  x(rst) and J = dx/drst, row-major (rows x,y,z; cols r,s,t):
 * HexElement.cu:122-139 (ReferenceToWorldHighOrder), :194-206 (JacobianHighOrder) */
ORC_EXPORT void orc_map(const orc_mesh_t* m, int e, const double* rst, double* x3, double* J9)
{
  for (int c = 0; c < 3; ++c) {
    double o[4];
    orc_eval_modal(geom_ptr(m, e, c), m->ng, rst, o);
    x3[c] = o[0];
    if (J9) {
      J9[3 * c + 0] = o[1];
      J9[3 * c + 1] = o[2];
      J9[3 * c + 2] = o[3];
    }
  }
}

static double det3(const double* J)
{
  return J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
         J[2] * (J[3] * J[7] - J[4] * J[6]);
}

/* cofactor inverse (HexElement.cu:208-225); returns 0 on singular J */
static int inv3(const double* J, double* inv)
{
  const double det = det3(J);
  if (!(fabs(det) > 1e-300)) return 0;
  const double id = 1.0 / det;
  inv[0] = (J[4] * J[8] - J[5] * J[7]) * id;
  inv[1] = (J[2] * J[7] - J[1] * J[8]) * id;
  inv[2] = (J[1] * J[5] - J[2] * J[4]) * id;
  inv[3] = (J[5] * J[6] - J[3] * J[8]) * id;
  inv[4] = (J[0] * J[8] - J[2] * J[6]) * id;
  inv[5] = (J[2] * J[3] - J[0] * J[5]) * id;
  inv[6] = (J[3] * J[7] - J[4] * J[6]) * id;
  inv[7] = (J[1] * J[6] - J[0] * J[7]) * id;
  inv[8] = (J[0] * J[4] - J[1] * J[3]) * id;
  return 1;
}

/* A5  point inversion x -> rst: HexElement.cu:227-272.  Newton from a seed
 * (the reference starts at rst = 0, undamped), converged in FP64 to
 * |drst|_inf < 1e-14 or to the rounding floor; when the undamped iteration
 * fails a damped restart (step capped at 0.5 in inf-norm) is tried.  Returns
 * the iteration count or -1. */
static int invert_from(const orc_mesh_t* m, int e, const double* x, const double* seed, double* rst)
{
  for (int attempt = 0; attempt < 2; ++attempt) {
    rst[0] = seed[0];
    rst[1] = seed[1];
    rst[2] = seed[2];
    double prev = 1e300;
    for (int it = 0; it < 80; ++it) {
      double X[3], J[9], iJ[9];
      orc_map(m, e, rst, X, J);
      if (!inv3(J, iJ)) break;
      const double dx = X[0] - x[0], dy = X[1] - x[1], dz = X[2] - x[2];
      double d0 = iJ[0] * dx + iJ[1] * dy + iJ[2] * dz;
      double d1 = iJ[3] * dx + iJ[4] * dy + iJ[5] * dz;
      double d2 = iJ[6] * dx + iJ[7] * dy + iJ[8] * dz;
      if (attempt == 1) {
        const double mx = fmax(fabs(d0), fmax(fabs(d1), fabs(d2)));
        if (mx > 0.5) {
          const double s = 0.5 / mx;
          d0 *= s; d1 *= s; d2 *= s;
        }
      }
      rst[0] -= d0;
      rst[1] -= d1;
      rst[2] -= d2;
      if (!(fabs(rst[0]) < 1e6 && fabs(rst[1]) < 1e6 && fabs(rst[2]) < 1e6)) break;
      /* converged below 1e-14, or at the FP64 rounding floor (step below
       * 1e-10 that no longer halves) */
      const double st = fmax(fabs(d0), fmax(fabs(d1), fabs(d2)));
      if (st < 1e-14 || (st < 1e-10 && st > 0.5 * prev)) return it + 1;
      prev = st;
    }
  }
  return -1;
}

/* The preimage of x in the element: for a valid element (positive Jacobian)
 * the map is one-to-one on the closed cube, so the in-cube preimage is unique.
 * Newton from rst = 0 (the reference's start) can converge to a spurious
 * preimage of the polynomial map outside the cube for points near a face; the
 * oracle then restarts from the 8 octant centres (+-0.5)^3 and returns the
 * first converged preimage with max|rst| <= 1 + 1e-3 (or, if none is inside,
 * the converged one nearest the cube). */
ORC_EXPORT int orc_invert(const orc_mesh_t* m, int e, const double* x, double* rst)
{
  double best[3] = {0, 0, 0}, best_out = 1e300;
  int best_it = -1;
  for (int s = 0; s < 9; ++s) {
    const double seed[3] = {s == 0 ? 0.0 : ((s - 1) & 1 ? 0.5 : -0.5),
                            s == 0 ? 0.0 : ((s - 1) & 2 ? 0.5 : -0.5),
                            s == 0 ? 0.0 : ((s - 1) & 4 ? 0.5 : -0.5)};
    double r[3];
    const int it = invert_from(m, e, x, seed, r);
    if (it < 0) continue;
    const double out = fmax(fabs(r[0]), fmax(fabs(r[1]), fabs(r[2])));
    if (out <= 1.0 + 1e-3) {
      rst[0] = r[0]; rst[1] = r[1]; rst[2] = r[2];
      return it;
    }
    if (out < best_out) {
      best_out = out;
      best[0] = r[0]; best[1] = r[1]; best[2] = r[2];
      best_it = it;
    }
  }
  rst[0] = best[0]; rst[1] = best[1]; rst[2] = best[2];
  return best_it;
}

/* A6  field value, reference gradient and world gradient J^{-T} grad_ref:
 * HexElement.cu:274-330 */
ORC_EXPORT void orc_field(const orc_mesh_t* m, int e, const double* rst, double* u,
                          double* gref, double* gworld)
{
  const size_t n3 = (size_t)m->nf * m->nf * m->nf;
  double o[4];
  orc_eval_modal(m->field + (size_t)e * n3, m->nf, rst, o);
  *u = o[0];
  if (gref) { gref[0] = o[1]; gref[1] = o[2]; gref[2] = o[3]; }
  if (gworld) {
    double X[3], J[9], iJ[9];
    orc_map(m, e, rst, X, J);
    if (!inv3(J, iJ)) { gworld[0] = gworld[1] = gworld[2] = NAN; return; }
    gworld[0] = iJ[0] * o[1] + iJ[3] * o[2] + iJ[6] * o[3];
    gworld[1] = iJ[1] * o[1] + iJ[4] * o[2] + iJ[7] * o[3];
    gworld[2] = iJ[2] * o[1] + iJ[5] * o[2] + iJ[8] * o[3];
  }
}

/* ------------------------------------------------------------------------ */
/* A9  face-constrained Newton: high_order.h:164-304                        */
/* Face order (Nektar): 0:t=-1 1:s=-1 2:r=+1 3:s=+1 4:r=-1 5:t=+1           */
/* ------------------------------------------------------------------------ */
static const int FACE_FIX[6] = {2, 1, 0, 1, 0, 2};
static const double FACE_VAL[6] = {-1, -1, 1, 1, -1, 1};
static const int FACE_U[6] = {0, 0, 1, 0, 1, 0};
static const int FACE_V[6] = {1, 2, 2, 2, 2, 1};

ORC_EXPORT void orc_face_axes(int face, int* fix, double* val, int* ua, int* va)
{
  *fix = FACE_FIX[face];
  *val = FACE_VAL[face];
  *ua = FACE_U[face];
  *va = FACE_V[face];
}

/* Nearest in-bounds root of X(face(u,v)) = o + tau d over the reference's 9
 * seeds {-0.8,0,0.8}^2 (high_order.h:206-211) and a 5x5 seed grid.  Each seed follows the
 * reference iteration: tau0 = projection of X(seed) onto the ray clamped to
 * [tmin,tmax] (:221-225); Newton on (u,v,tau) with the step scaled to
 * |.|_inf <= 0.25 (:288-293), tau clamped to [tmin,tmax] (:294), u,v clamped
 * to [-1.5,1.5] (:297-298), divergence break when iter>5 and the residual
 * grew 1.5x (:253).  Converged when |R|_inf < 1e-12 (the reference stops at
 * 1e-4, :197,:240); then two unconstrained polishing steps.  A converged root
 * is accepted when max(|u|,|v|) <= 1+ORC_FACE_ACC and tmin <= tau <= tmax.
 * Returns 1 and the nearest accepted root; *margin receives
 * 1+ORC_FACE_ACC-max(|u|,|v|) for the returned root, and *nearmiss the
 * smallest |1+ORC_FACE_ACC-max(|u|,|v|)| over converged but rejected roots
 * (both used to classify tie pixels). */
/* Rejected (converged, just outside the face) roots for tie classification. */
typedef struct {
  double t[64], mg[64];
  int n;
} orc_rej_t;

static int face_roots_impl(const orc_mesh_t* m, int e, int face, const double* o, const double* d,
                           double tmin, double tmax, double* t_out, double* rst_out, double* margin,
                           double* nearmiss, double* all_t, int cap, int* n_all, orc_rej_t* rej)
{
  const int fa = FACE_FIX[face], ua = FACE_U[face], va = FACE_V[face];
  const double fv = FACE_VAL[face];
  /* the reference's 3x3 grid {-0.8,0,0.8}^2 followed by a 5x5 grid
   * {-0.9,-0.45,0,0.45,0.9}^2: the denser grid reaches roots the reference's
   * seeds miss on strongly curved faces */
  double seeds[34][2];
  int ns = 0;
  for (int sj = 0; sj < 3; ++sj)
    for (int si = 0; si < 3; ++si) {
      seeds[ns][0] = -0.8 + 0.8 * si;
      seeds[ns][1] = -0.8 + 0.8 * sj;
      ++ns;
    }
  for (int sj = 0; sj < 5; ++sj)
    for (int si = 0; si < 5; ++si) {
      seeds[ns][0] = -0.9 + 0.45 * si;
      seeds[ns][1] = -0.9 + 0.45 * sj;
      ++ns;
    }
  int found = 0;
  double best_t = tmax;
  double nm = 1e30;
  const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  for (int sq = 0; sq < ns; ++sq) {
    {
      double rst[3] = {0, 0, 0};
      rst[fa] = fv;
      rst[ua] = seeds[sq][0];
      rst[va] = seeds[sq][1];
      double X[3], J[9];
      orc_map(m, e, rst, X, NULL);
      double tau = dd > 0 ? ((X[0] - o[0]) * d[0] + (X[1] - o[1]) * d[1] + (X[2] - o[2]) * d[2]) / dd : 0.0;
      tau = fmin(fmax(tau, tmin), tmax);
      double prev = DBL_MAX;
      int conv = 0;
      for (int it = 0; it < 200; ++it) {
        orc_map(m, e, rst, X, J);
        const double R0 = X[0] - (o[0] + tau * d[0]);
        const double R1 = X[1] - (o[1] + tau * d[1]);
        const double R2 = X[2] - (o[2] + tau * d[2]);
        const double res = fmax(fabs(R0), fmax(fabs(R1), fabs(R2)));
        if (res < 1e-12) { conv = 1; break; }
        if (it > 5 && res > prev * 1.5) break;
        prev = res;
        double A[9] = {J[0 + ua], J[0 + va], -d[0], J[3 + ua], J[3 + va], -d[1],
                       J[6 + ua], J[6 + va], -d[2]};
        double iA[9];
        if (!inv3(A, iA)) break;
        double du = -(iA[0] * R0 + iA[1] * R1 + iA[2] * R2);
        double dv = -(iA[3] * R0 + iA[4] * R1 + iA[5] * R2);
        double dt = -(iA[6] * R0 + iA[7] * R1 + iA[8] * R2);
        const double mx = fmax(fabs(du), fmax(fabs(dv), fabs(dt)));
        const double sc = mx > 0.25 ? 0.25 / mx : 1.0;
        rst[ua] += du * sc;
        rst[va] += dv * sc;
        tau += dt * sc;
        tau = fmin(fmax(tau, tmin), tmax);
        rst[ua] = fmin(fmax(rst[ua], -1.5), 1.5);
        rst[va] = fmin(fmax(rst[va], -1.5), 1.5);
      }
      if (!conv) continue;
      /* unconstrained polish */
      for (int it = 0; it < 2; ++it) {
        orc_map(m, e, rst, X, J);
        const double R0 = X[0] - (o[0] + tau * d[0]);
        const double R1 = X[1] - (o[1] + tau * d[1]);
        const double R2 = X[2] - (o[2] + tau * d[2]);
        double A[9] = {J[0 + ua], J[0 + va], -d[0], J[3 + ua], J[3 + va], -d[1],
                       J[6 + ua], J[6 + va], -d[2]};
        double iA[9];
        if (!inv3(A, iA)) break;
        rst[ua] -= iA[0] * R0 + iA[1] * R1 + iA[2] * R2;
        rst[va] -= iA[3] * R0 + iA[4] * R1 + iA[5] * R2;
        tau -= iA[6] * R0 + iA[7] * R1 + iA[8] * R2;
      }
      const double mu = fmax(fabs(rst[ua]), fabs(rst[va]));
      const double mg = 1.0 + ORC_FACE_ACC - mu;
      if (mg >= 0.0 && tau >= tmin && tau <= tmax) {
        if (all_t) {
          int dup = 0;
          for (int q = 0; q < *n_all; ++q)
            if (fabs(all_t[q] - tau) <= 1e-9 * (1.0 + fabs(tau))) dup = 1;
          if (!dup && *n_all < cap) all_t[(*n_all)++] = tau;
        }
        if (!found || tau < best_t) {
          found = 1;
          best_t = tau;
          rst_out[0] = rst[0];
          rst_out[1] = rst[1];
          rst_out[2] = rst[2];
          if (margin) *margin = mg;
        }
      } else if (tau >= tmin && tau <= tmax) {
        if (fabs(mg) < nm) nm = fabs(mg);
        if (rej && fabs(mg) < 1e-3 && rej->n < 64) {
          rej->t[rej->n] = tau;
          rej->mg[rej->n] = fabs(mg);
          ++rej->n;
        }
      }
    }
  }
  if (nearmiss) *nearmiss = nm;
  if (found) *t_out = best_t;
  return found;
}

ORC_EXPORT int orc_face_hit(const orc_mesh_t* m, int e, int face, const double* o,
                            const double* d, double tmin, double tmax, double* t_out,
                            double* rst_out, double* margin, double* nearmiss)
{
  return face_roots_impl(m, e, face, o, d, tmin, tmax, t_out, rst_out, margin, nearmiss, NULL, 0,
                         NULL, NULL);
}

/* ------------------------------------------------------------------------ */
/* A7/A8  conservative element box by dense sampling (independent of the    */
/* product's Bernstein bound) and the slab test (high_order.h:115-152).     */
/* ------------------------------------------------------------------------ */
ORC_EXPORT void orc_elem_box(const orc_mesh_t* m, int e, int samples, double pad_frac, double* lo,
                             double* hi)
{
  /* The image of the reference cube is bounded by the image of its boundary,
   * so the box samples the 6 faces on a samples x samples grid and pads by
   * pad_frac of the largest extent. */
  lo[0] = lo[1] = lo[2] = DBL_MAX;
  hi[0] = hi[1] = hi[2] = -DBL_MAX;
  for (int f = 0; f < 6; ++f)
    for (int j = 0; j < samples; ++j)
      for (int i = 0; i < samples; ++i) {
        double rst[3];
        rst[FACE_FIX[f]] = FACE_VAL[f];
        rst[FACE_U[f]] = -1.0 + 2.0 * i / (samples - 1);
        rst[FACE_V[f]] = -1.0 + 2.0 * j / (samples - 1);
        double X[3];
        orc_map(m, e, rst, X, NULL);
        for (int c = 0; c < 3; ++c) {
          lo[c] = fmin(lo[c], X[c]);
          hi[c] = fmax(hi[c], X[c]);
        }
      }
  double ext = 0.0;
  for (int c = 0; c < 3; ++c) ext = fmax(ext, hi[c] - lo[c]);
  for (int c = 0; c < 3; ++c) {
    lo[c] -= pad_frac * ext + 1e-9;
    hi[c] += pad_frac * ext + 1e-9;
  }
}

ORC_EXPORT int orc_slab(const double* o, const double* d, const double* lo, const double* hi,
                        double tmin, double tmax, double* t0, double* t1)
{
  double a = tmin, b = tmax;
  for (int c = 0; c < 3; ++c) {
    if (fabs(d[c]) < 1e-300) {
      if (o[c] < lo[c] || o[c] > hi[c]) return 0;
      continue;
    }
    const double inv = 1.0 / d[c];
    double ta = (lo[c] - o[c]) * inv, tb = (hi[c] - o[c]) * inv;
    if (ta > tb) { const double s = ta; ta = tb; tb = s; }
    a = fmax(a, ta);
    b = fmin(b, tb);
    if (a > b) return 0;
  }
  *t0 = a;
  *t1 = b;
  return 1;
}

/* ------------------------------------------------------------------------ */
/* A11  isosurface search in one element segment: fem/RayIsosurface.cu:44-188*/
/* g(t) = u(x^{-1}(o+t d)) - iso sampled at M+1 uniform points over the      */
/* segment [ta,tb] (reference: 16 sub-intervals, :78-86); the first          */
/* sub-interval with g(prev)*g(cur) <= 0 (:108) is refined by bisection to   */
/* |dt| <= 1e-14 (t+1) (the reference mixes Newton and bisection, :121-171). */
/* Returns 1 with the root, its rst and the field there.                     */
/* ------------------------------------------------------------------------ */
typedef struct {
  double t;
  double rst[3];
  double u;
  double gw[3];
} orc_iso_root_t;

static int eval_along(const orc_mesh_t* m, int e, const double* o, const double* d, double t,
                      double* rst, double* u)
{
  const double x[3] = {o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]};
  if (orc_invert(m, e, x, rst) < 0) return 0;
  orc_field(m, e, rst, u, NULL, NULL);
  return 1;
}

/* Tie margin of a scan sample at ray parameter t with g = u - iso: the sign
 * of g is not determined by an FP32 scan when |g| is below 1e-5 (evaluation
 * error), or below 4e-6 t |dg/dt| (the sample's position in FP32, from FP32
 * face roots, moves it by a few 1e-7 t).  Returned on one scale: the smaller
 * of |g| and 1e-5 |g| / (4e-6 t |dg/dt|), so a margin below 1e-5 is a tie. */
static double sample_tie_margin(const orc_mesh_t* m, int e, const double* rst, const double* d, double t,
                                double g)
{
  double u, gr[3], gw[3];
  orc_field(m, e, rst, &u, gr, gw);
  const double gp = fabs(gw[0] * d[0] + gw[1] * d[1] + gw[2] * d[2]);
  const double pos = 4e-6 * fabs(t) * gp;
  return pos > 0.0 ? fmin(fabs(g), 1e-5 * fabs(g) / pos) : fabs(g);
}

ORC_EXPORT int orc_iso_segment(const orc_mesh_t* m, int e, const double* o, const double* d,
                               double ta, double tb, int M, double iso, double* out_t,
                               double* out_rst, double* out_u, double* out_minabs)
{
  double prev_g = 0.0, prev_t = ta;
  int have_prev = 0;
  double minabs = 1e30;
  for (int j = 0; j <= M; ++j) {
    const double t = ta + (tb - ta) * ((double)j / M);
    double rst[3], u;
    if (!eval_along(m, e, o, d, t, rst, &u)) continue;
    const double g = u - iso;
    minabs = fmin(minabs, sample_tie_margin(m, e, rst, d, t, g));
    if (have_prev && prev_g * g <= 0.0) {
      /* more than one crossing inside this sample interval (seen by 32
         sub-samples): which of them an endgame converges to is not fixed by
         the scan -- the reference's Newton / bisection in the bracket
         (fem/RayIsosurface.cu:121-171) takes any -- so the result is a tie */
      {
        double pg = prev_g;
        int changes = 0;
        for (int s = 1; s <= 32; ++s) {
          double sr[3], su;
          if (!eval_along(m, e, o, d, prev_t + (t - prev_t) * ((double)s / 32.0), sr, &su)) continue;
          const double sg = su - iso;
          if ((pg < 0.0) != (sg < 0.0)) ++changes;
          pg = sg;
        }
        if (changes > 1) minabs = 0.0;
      }
      double lt = prev_t, lg = prev_g, rt = t;
      for (int it = 0; it < 200 && fabs(rt - lt) > 1e-14 * (fabs(rt) + 1.0); ++it) {
        const double mt = 0.5 * (lt + rt);
        double mr[3], mu;
        if (!eval_along(m, e, o, d, mt, mr, &mu)) break;
        const double mg = mu - iso;
        if (lg * mg <= 0.0) {
          rt = mt;
        } else {
          lt = mt;
          lg = mg;
        }
      }
      const double root = 0.5 * (lt + rt);
      double rr[3], ru;
      if (!eval_along(m, e, o, d, root, rr, &ru)) return 0;
      *out_t = root;
      out_rst[0] = rr[0];
      out_rst[1] = rr[1];
      out_rst[2] = rr[2];
      *out_u = ru;
      if (out_minabs) *out_minabs = minabs;
      return 1;
    }
    prev_g = g;
    prev_t = t;
    have_prev = 1;
  }
  if (out_minabs) *out_minabs = minabs;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* A10/A13  whole-mesh trace of explicit rays, brute force over elements     */
/* (raytracer.cu:102-254 hex branch; high_order.h:409-465).  Candidate       */
/* elements are those whose sampled box (11^3 samples, 10% pad) the ray      */
/* crosses; the BVH is an optimisation the product must show equivalent to   */
/* this.                                                                     */
/* Output per ray (doubles, ORC_GEOM_STRIDE):                                */
/*   [0] elem (-1 miss) [1] face [2] t [3..5] rst [6] u at hit               */
/*   [7] t of nearest candidate on a different (elem,face) (1e30 none)       */
/*   [8] margin of the accepted root  [9] smallest |margin| of a converged    */
/*   root rejected just outside its face, at or before the hit (any root for */
/*   a miss)                                                                 */
/* Per ray and isovalue (ORC_ISO_STRIDE):                                     */
/*   [0] elem (-1 none) [1] t [2..4] rst [5] u [6] smallest tie margin       */
/*   (sample_tie_margin) over the samples of every segment scanned before    */
/*   the root, up to its bracket (all segments when none is found)           */
/* ------------------------------------------------------------------------ */
#define ORC_GEOM_STRIDE 10
#define ORC_ISO_STRIDE 7
#define ORC_MAX_ISO_REC 8  /* isovalues whose tie margins track every scanned segment */

ORC_EXPORT int orc_geom_stride(void) { return ORC_GEOM_STRIDE; }
ORC_EXPORT int orc_iso_stride(void) { return ORC_ISO_STRIDE; }
ORC_EXPORT double orc_face_acc(void) { return ORC_FACE_ACC; }

typedef struct {
  double lo[3], hi[3];
  int ready;
} box_t;

typedef struct {
  double t;
  int e;
} cand_t;

static int cmp_cand(const void* a, const void* b)
{
  const double x = ((const cand_t*)a)->t, y = ((const cand_t*)b)->t;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* Candidate culling (an optimisation only): a coarse box per element from its
 * 8 corner images padded by its full extent on every side, then, lazily, the
 * face-sampled box (9x9 per face, 10% pad).  Candidates are visited by
 * increasing coarse entry distance; once that distance exceeds every result
 * still open (geometry hit, each isovalue) no later element can improve any
 * result, so the scan stops. */
ORC_EXPORT void orc_trace(const orc_mesh_t* m, const float* rays, int nrays, const double* isos,
                          int niso, int M, double tmin, double* out_geom, double* out_iso)
{
  double* coarse = (double*)malloc(sizeof(double) * 6 * (size_t)m->n_elem);
  box_t* boxes = (box_t*)calloc((size_t)m->n_elem, sizeof(box_t));
  for (int e = 0; e < m->n_elem; ++e) {
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int c8 = 0; c8 < 8; ++c8) {
      const double rst[3] = {(c8 & 1) ? 1.0 : -1.0, (c8 & 2) ? 1.0 : -1.0, (c8 & 4) ? 1.0 : -1.0};
      double X[3];
      orc_map(m, e, rst, X, NULL);
      for (int c = 0; c < 3; ++c) {
        lo[c] = fmin(lo[c], X[c]);
        hi[c] = fmax(hi[c], X[c]);
      }
    }
    double ext = 0.0;
    for (int c = 0; c < 3; ++c) ext = fmax(ext, hi[c] - lo[c]);
    for (int c = 0; c < 3; ++c) {
      coarse[6 * (size_t)e + c] = lo[c] - ext;
      coarse[6 * (size_t)e + 3 + c] = hi[c] + ext;
    }
  }

#pragma omp parallel for schedule(dynamic, 1)
  for (int r = 0; r < nrays; ++r) {
    const double o[3] = {rays[0 * (size_t)nrays + r], rays[1 * (size_t)nrays + r],
                         rays[2 * (size_t)nrays + r]};
    const double d[3] = {rays[3 * (size_t)nrays + r], rays[4 * (size_t)nrays + r],
                         rays[5 * (size_t)nrays + r]};
    double* G = out_geom + (size_t)r * ORC_GEOM_STRIDE;
    G[0] = -1; G[1] = -1; G[2] = 1e30; G[3] = G[4] = G[5] = 0; G[6] = 0;
    G[7] = 1e30; G[8] = 0; G[9] = 1e30;
    for (int k = 0; k < niso; ++k) {
      double* I = out_iso + ((size_t)r * niso + k) * ORC_ISO_STRIDE;
      I[0] = -1; I[1] = 1e30; I[2] = I[3] = I[4] = 0; I[5] = 0; I[6] = 1e30;
    }
    int cap = 256, nc = 0;
    cand_t* cand = (cand_t*)malloc(sizeof(cand_t) * cap);
    /* every scanned segment's start and smallest |u - iso| over its samples
     * (up to its bracket), per isovalue: the tie margin of a first crossing
     * is the smallest over all segments before it, since "no crossing in
     * this segment" is a sign decision too */
    enum { SEGREC = 256 };
    double seg_ta[ORC_MAX_ISO_REC][SEGREC], seg_mn[ORC_MAX_ISO_REC][SEGREC], seg_over[ORC_MAX_ISO_REC];
    int nseg[ORC_MAX_ISO_REC];
    for (int k = 0; k < ORC_MAX_ISO_REC; ++k) {
      nseg[k] = 0;
      seg_over[k] = 1e30;
    }
    orc_rej_t rej;
    rej.n = 0;
    for (int e = 0; e < m->n_elem; ++e) {
      double t0, t1;
      if (!orc_slab(o, d, coarse + 6 * (size_t)e, coarse + 6 * (size_t)e + 3, tmin, 1e30, &t0, &t1))
        continue;
      if (nc == cap) {
        cap *= 2;
        cand = (cand_t*)realloc(cand, sizeof(cand_t) * cap);
      }
      cand[nc].t = t0;
      cand[nc].e = e;
      ++nc;
    }
    qsort(cand, nc, sizeof(cand_t), cmp_cand);
    for (int ci = 0; ci < nc; ++ci) {
      /* sound early exit */
      double open = G[7] < 1e30 ? G[7] : G[2];
      for (int k = 0; k < niso; ++k) open = fmax(open, out_iso[((size_t)r * niso + k) * ORC_ISO_STRIDE + 1]);
      if (G[0] >= 0 && cand[ci].t > open) break;
      const int e = cand[ci].e;
      if (!boxes[e].ready) {
        double lo[3], hi[3];
        orc_elem_box(m, e, 9, 0.10, lo, hi);
#pragma omp critical(orc_boxes)
        {
          if (!boxes[e].ready) {
            memcpy(boxes[e].lo, lo, sizeof(lo));
            memcpy(boxes[e].hi, hi, sizeof(hi));
            boxes[e].ready = 1;
          }
        }
      }
      double t0, t1;
      if (!orc_slab(o, d, boxes[e].lo, boxes[e].hi, tmin, 1e30, &t0, &t1)) continue;
      double ft[6], frst[6][3];
      int fok[6];
      double roots[64];
      int nroots = 0;
      for (int f = 0; f < 6; ++f) {
        double mg = 0, nm = 1e30;
        double fr[16];
        int nf = 0;
        fok[f] = face_roots_impl(m, e, f, o, d, tmin, 1e30, &ft[f], frst[f], &mg, &nm, fr, 16, &nf,
                                 &rej);
        for (int q = 0; q < nf; ++q) {
          int dup = 0;
          for (int z = 0; z < nroots; ++z)
            if (fabs(roots[z] - fr[q]) <= 1e-9 * (1.0 + fabs(fr[q]))) dup = 1;
          if (!dup && nroots < 64) roots[nroots++] = fr[q];
        }
        if (!fok[f]) continue;
        if (ft[f] < G[2]) {
          if (G[0] >= 0) G[7] = fmin(G[7], G[2]);
          G[0] = e; G[1] = f; G[2] = ft[f];
          G[3] = frst[f][0]; G[4] = frst[f][1]; G[5] = frst[f][2];
          G[8] = mg;
        } else {
          G[7] = fmin(G[7], ft[f]);
        }
      }
      if (niso == 0 || nroots < 2) continue;
      /* the ray is inside the element between consecutive crossings of its
       * boundary: [r0, r1], [r2, r3], ... (sorted distinct face roots) */
      for (int a2 = 1; a2 < nroots; ++a2)
        for (int b2 = a2; b2 > 0 && roots[b2] < roots[b2 - 1]; --b2) {
          const double tmp = roots[b2];
          roots[b2] = roots[b2 - 1];
          roots[b2 - 1] = tmp;
        }
      /* which gaps between consecutive roots lie inside the element: the
       * midpoint's preimage in the closed cube.  Pairing the sorted roots as
       * (r0, r1), (r2, r3), ... instead goes wrong for rays that run along
       * an edge or inside a face, where a root is found twice or missed;
       * consecutive inside gaps make one segment */
      int inside[64];
      for (int iv = 0; iv + 1 < nroots; ++iv) {
        const double tm = 0.5 * (roots[iv] + roots[iv + 1]);
        const double xm[3] = {o[0] + tm * d[0], o[1] + tm * d[1], o[2] + tm * d[2]};
        double rm[3];
        const int itm = orc_invert(m, e, xm, rm);
        inside[iv] = itm >= 0 && fmax(fabs(rm[0]), fmax(fabs(rm[1]), fabs(rm[2]))) <= 1.0 + 1e-9;
      }
      for (int iv = 0; iv + 1 < nroots; ++iv) {
        if (!inside[iv]) continue;
        int iw = iv;
        while (iw + 2 < nroots && inside[iw + 1]) ++iw;
        const double ta = roots[iv], tb = roots[iw + 1];
        iv = iw;
        for (int k = 0; k < niso; ++k) {
          double* I = out_iso + ((size_t)r * niso + k) * ORC_ISO_STRIDE;
          if (I[0] >= 0 && I[1] <= ta) continue; /* an earlier root already */
          double t, rst[3], u, mn = 1e30;
          const int okk = orc_iso_segment(m, e, o, d, ta, tb, M, isos[k], &t, rst, &u, &mn);
          if (k < ORC_MAX_ISO_REC) {
            if (nseg[k] < SEGREC) {
              seg_ta[k][nseg[k]] = ta;
              seg_mn[k][nseg[k]] = mn;
              ++nseg[k];
            } else {
              seg_over[k] = fmin(seg_over[k], mn); /* (counted as before any root) */
            }
          }
          if (okk && t < I[1]) {
            I[0] = e; I[1] = t; I[2] = rst[0]; I[3] = rst[1]; I[4] = rst[2]; I[5] = u;
            I[6] = mn;
          }
        }
      }
    }
    free(cand);
    for (int k = 0; k < niso && k < ORC_MAX_ISO_REC; ++k) {
      double* I = out_iso + ((size_t)r * niso + k) * ORC_ISO_STRIDE;
      double mn = fmin(I[6], seg_over[k]);
      for (int q = 0; q < nseg[k]; ++q)
        if (seg_ta[k][q] < I[1]) mn = fmin(mn, seg_mn[k][q]);
      I[6] = mn;
    }
    /* near miss: a rejected root at or before the hit (any, for a miss) */
    for (int q = 0; q < rej.n; ++q)
      if (G[0] < 0 || rej.t[q] <= G[2] * (1.0 + 1e-6) + 1e-12) G[9] = fmin(G[9], rej.mg[q]);
    if (G[0] >= 0) {
      const double rst[3] = {G[3], G[4], G[5]};
      double u;
      orc_field(m, (int)G[0], rst, &u, NULL, NULL);
      G[6] = u;
    }
  }
  free(boxes);
  free(coarse);
}

/* ------------------------------------------------------------------------ */
/* A14  pinhole camera (raytracer.h:31-48) with pixel-centre rays            */
/* (raytracer.cu:439-447 without the jitter).  Rays are SoA floats:          */
/* ox[n] oy[n] oz[n] dx[n] dy[n] dz[n]; pixel index = y*W + x, y=0 bottom.   */
/* ------------------------------------------------------------------------ */
ORC_EXPORT void orc_camera_rays(const double* lookfrom, const double* lookat, const double* vup,
                                double vfov_deg, int W, int H, double* rays_out)
{
  const double theta = vfov_deg * 3.14159265358979323846 / 180.0;
  const double h = tan(theta / 2.0);
  const double vh = 2.0 * h;
  const double vw = ((double)W / (double)H) * vh;
  double w[3] = {lookfrom[0] - lookat[0], lookfrom[1] - lookat[1], lookfrom[2] - lookat[2]};
  double fd = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  for (int c = 0; c < 3; ++c) w[c] /= fd;
  double u[3] = {vup[1] * w[2] - vup[2] * w[1], vup[2] * w[0] - vup[0] * w[2],
                 vup[0] * w[1] - vup[1] * w[0]};
  const double un = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  for (int c = 0; c < 3; ++c) u[c] /= un;
  const double v[3] = {w[1] * u[2] - w[2] * u[1], w[2] * u[0] - w[0] * u[2],
                       w[0] * u[1] - w[1] * u[0]};
  double hor[3], ver[3], llc[3];
  for (int c = 0; c < 3; ++c) {
    hor[c] = fd * vw * u[c];
    ver[c] = fd * vh * v[c];
    llc[c] = lookfrom[c] - hor[c] / 2 - ver[c] / 2 - fd * w[c];
  }
  const size_t n = (size_t)W * H;
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const size_t p = (size_t)y * W + x;
      const double s = (x + 0.5) / W, t = (y + 0.5) / H;
      double dir[3];
      for (int c = 0; c < 3; ++c) dir[c] = llc[c] + s * hor[c] + t * ver[c] - lookfrom[c];
      const double l = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
      for (int c = 0; c < 3; ++c) {
        rays_out[c * n + p] = lookfrom[c];
        rays_out[(3 + c) * n + p] = dir[c] / l;
      }
    }
}

/* ------------------------------------------------------------------------ */
/* Curved quadrilateral surface elements (fem/QuadElement.cu).  A quad is   */
/* the 2D Modified_A map x(r,s) = sum_{j,i} G[c][j n + i] phi_i(r) phi_j(s)   */
/* (EvaluateModal2D, QuadElement.cu:22-37: i fastest) with the field u(r,s)  */
/* alike.  Ray intersection (IntersectRayWithQuad, QuadElement.cu:467-527:   */
/* Newton on (r, s, t) from a seed the reference finds by sampling): here    */
/* the NEAREST root over a dense 7x7 seed grid per quad, converged in FP64,  */
/* accepted when max(|r|,|s|) <= 1 + ORC_FACE_ACC and t >= tmin, brute      */
/* force over every quad.                                                    */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n_quad, ng, nf, pad_;
  const float* geom;   /* n x 3 x ng^2 */
  const float* field;  /* n x nf^2 */
} orc_quads_t;

ORC_EXPORT void orc_quad_map(const orc_quads_t* q, int e, double r, double s, double* x, double* xr,
                             double* xs)
{
  const int n = q->ng;
  double pr[ORC_MAXN], ps[ORC_MAXN], dpr[ORC_MAXN], dps[ORC_MAXN];
  for (int i = 0; i < n; ++i) {
    pr[i] = orc_modA(i, r);
    ps[i] = orc_modA(i, s);
    dpr[i] = orc_modA_d(i, r);
    dps[i] = orc_modA_d(i, s);
  }
  for (int c = 0; c < 3; ++c) {
    const float* g = q->geom + ((size_t)e * 3 + c) * n * n;
    double v = 0, vr = 0, vs = 0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const double a = g[j * n + i];
        v += a * pr[i] * ps[j];
        vr += a * dpr[i] * ps[j];
        vs += a * pr[i] * dps[j];
      }
    x[c] = v;
    if (xr) xr[c] = vr;
    if (xs) xs[c] = vs;
  }
}

ORC_EXPORT double orc_quad_field(const orc_quads_t* q, int e, double r, double s)
{
  const int n = q->nf;
  const float* f = q->field + (size_t)e * n * n;
  double u = 0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) u += f[j * n + i] * orc_modA(i, r) * orc_modA(j, s);
  return u;
}

/* Newton on (r, s, t) from one seed, damped to steps of 0.5 in the
 * inf-norm; converged when the residual is below 1e-12 (1 + |x|), then two
 * undamped steps take the root to the FP64 floor.  1 = converged (any r, s). */
static int quad_newton(const orc_quads_t* q, int e, const double* o, const double* d, double r, double s,
                       double t, double* ro, double* so, double* to)
{
  int conv = 0;
  for (int it = 0; it < 200 && conv < 3; ++it) {
    double x[3], xr[3], xs[3];
    orc_quad_map(q, e, r, s, x, xr, xs);
    const double R[3] = {x[0] - o[0] - t * d[0], x[1] - o[1] - t * d[1], x[2] - o[2] - t * d[2]};
    const double res = fmax(fabs(R[0]), fmax(fabs(R[1]), fabs(R[2])));
    const double sx = 1 + fmax(fabs(x[0]), fmax(fabs(x[1]), fabs(x[2])));
    if (conv || res < 1e-12 * sx) ++conv;
    const double A[9] = {xr[0], xs[0], -d[0], xr[1], xs[1], -d[1], xr[2], xs[2], -d[2]};
    double inv[9];
    if (!inv3(A, inv)) return 0;
    const double dr = -(inv[0] * R[0] + inv[1] * R[1] + inv[2] * R[2]);
    const double ds = -(inv[3] * R[0] + inv[4] * R[1] + inv[5] * R[2]);
    const double dt = -(inv[6] * R[0] + inv[7] * R[1] + inv[8] * R[2]);
    const double mx = fmax(fabs(dr), fmax(fabs(ds), fabs(dt)));
    const double sc = (!conv && mx > 0.5) ? 0.5 / mx : 1.0;
    if (conv == 3) break;   /* converged and polished twice */
    r += sc * dr;
    s += sc * ds;
    t += sc * dt;
    if (fabs(r) > 3 || fabs(s) > 3) return 0;
  }
  if (conv < 3) return 0;
  *ro = r;
  *so = s;
  *to = t;
  return 1;
}

/* Per ray (ORC_GEOM_STRIDE doubles, as orc_trace): [0] quad (-1 miss)
 * [1] face (0) [2] t [3..5] (r, s, -1) [6] u [7] t of the nearest root on
 * another quad [8] margin 1 + ORC_FACE_ACC - max(|r|,|s|) of the hit
 * [9] smallest |margin| of a converged root rejected just outside a quad, at
 * or before the hit (any such root for a miss). */
ORC_EXPORT void orc_quad_trace(const orc_quads_t* q, const float* rays, int nrays, double tmin, double* out)
{
#pragma omp parallel for schedule(dynamic, 1)
  for (int ray = 0; ray < nrays; ++ray) {
    const double o[3] = {rays[ray], rays[(size_t)nrays + ray], rays[2 * (size_t)nrays + ray]};
    const double d[3] = {rays[3 * (size_t)nrays + ray], rays[4 * (size_t)nrays + ray],
                         rays[5 * (size_t)nrays + ray]};
    double* G = out + (size_t)ray * ORC_GEOM_STRIDE;
    for (int k = 0; k < ORC_GEOM_STRIDE; ++k) G[k] = 0.0;
    G[0] = -1;
    G[2] = 1e30;
    G[7] = 1e30;
    G[9] = 1e30;
    double best_r = 0, best_s = 0;
    double rej_t[64], rej_m[64];
    int n_rej = 0;
    for (int e = 0; e < q->n_quad; ++e) {
      double bt = 1e30, br_ = 0, bs_ = 0;
      for (int sj = 0; sj < 7; ++sj)
        for (int si = 0; si < 7; ++si) {
          const double r0 = -1 + (2.0 * si + 1) / 7.0, s0 = -1 + (2.0 * sj + 1) / 7.0;
          double x[3];
          orc_quad_map(q, e, r0, s0, x, NULL, NULL);
          const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
          const double t0 = ((x[0] - o[0]) * d[0] + (x[1] - o[1]) * d[1] + (x[2] - o[2]) * d[2]) / dd;
          double r, s, t;
          if (!quad_newton(q, e, o, d, r0, s0, t0, &r, &s, &t) || t < tmin) continue;
          const double mg = 1 + ORC_FACE_ACC - fmax(fabs(r), fabs(s));
          if (mg < 0) {
            if (-mg < 1e-3 && n_rej < 64) {   /* near miss, classified once the hit is known */
              rej_t[n_rej] = t;
              rej_m[n_rej++] = -mg;
            }
            continue;
          }
          if (t < bt) {
            bt = t;
            br_ = r;
            bs_ = s;
          }
        }
      if (bt >= 1e30) continue;
      if (bt < G[2]) {
        if (G[0] >= 0) G[7] = fmin(G[7], G[2]);
        G[0] = e;
        G[2] = bt;
        best_r = br_;
        best_s = bs_;
      } else {
        G[7] = fmin(G[7], bt);
      }
    }
    if (G[0] >= 0) {
      G[1] = 0;
      G[3] = best_r;
      G[4] = best_s;
      G[5] = -1;
      G[6] = orc_quad_field(q, (int)G[0], best_r, best_s);
      G[8] = 1 + ORC_FACE_ACC - fmax(fabs(best_r), fabs(best_s));
    }
    for (int k = 0; k < n_rej; ++k)
      if (G[0] < 0 || rej_t[k] <= G[2] * (1.0 + 1e-6) + 1e-12) G[9] = fmin(G[9], rej_m[k]);
  }
}

/* ------------------------------------------------------------------------ */
/* Prisms (fem/PrismElement.cu).  The map and field are sums over           */
/* Modified_A(i, r) Modified_A(j, s) Modified_B(i, k, t), i slowest, then j, */
/* then k < mz - i (EvaluatePrismModal, PrismElement.cu:16-37), Modified_B   */
/* as fem/ModalBasis.cuh:72-82 defines it; without geometry coefficients the */
/* map is the linear prism of PrismElement.cu:194-226.  Restated directly,  */
/* not through the hex embedding the product traces.                        */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n, pad_;
  int gm[3], fm[3];
  const float* vertices; /* n x 6 x 3 */
  const float* geom;     /* n x 3 x count(gm), or NULL: linear */
  const float* field;    /* n x count(fm), or NULL: zero */
} orc_prisms_t;

/* Jacobi P_n^(a,b) by its explicit sum (Szego 4.3.2):
 * P_n(x) = sum_s C(n+a, n-s) C(n+b, s) ((x-1)/2)^s ((x+1)/2)^(n-s). */
static double orc_binom(double top, int k)
{
  double r = 1.0;
  for (int q = 1; q <= k; ++q) r *= (top - k + q) / q;
  return r;
}

ORC_EXPORT double orc_jacobi(int n, double a, double b, double x)
{
  double s = 0.0;
  for (int q = 0; q <= n; ++q)
    s += orc_binom(n + a, n - q) * orc_binom(n + b, q) * pow(0.5 * (x - 1.0), q) * pow(0.5 * (x + 1.0), n - q);
  return s;
}

ORC_EXPORT double orc_modB(int i, int k, double x)
{
  if (i == 0) return orc_modA(k, x);
  const double w = pow(0.5 * (1.0 - x), i);
  if (k == 0) return w;
  return w * 0.5 * (1.0 + x) * orc_jacobi(k - 1, 2.0 * i - 1.0, 1.0, x);
}

ORC_EXPORT int orc_prism_count(const int* m)
{
  int c = 0;
  for (int i = 0; i < m[0]; ++i) c += m[1] * (m[2] > i ? m[2] - i : 0);
  return c;
}

static double prism_sum(const float* c, const int* m, const double* rst)
{
  double v = 0.0;
  int idx = 0;
  for (int i = 0; i < m[0]; ++i) {
    const double a = orc_modA(i, rst[0]);
    for (int j = 0; j < m[1]; ++j) {
      const double b = orc_modA(j, rst[1]);
      for (int k = 0; k + i < m[2]; ++k) v += c[idx++] * a * b * orc_modB(i, k, rst[2]);
    }
  }
  return v;
}

ORC_EXPORT void orc_prism_map(const orc_prisms_t* P, int e, const double* rst, double* x)
{
  if (P->geom) {
    const int nc = orc_prism_count(P->gm);
    for (int c = 0; c < 3; ++c) x[c] = prism_sum(P->geom + ((size_t)e * 3 + c) * nc, P->gm, rst);
    return;
  }
  const double r = rst[0], s = rst[1], t = rst[2];
  const double w[6] = {0.125 * (1 - r) * (1 - s) * (1 - t), 0.125 * (1 + r) * (1 - s) * (1 - t),
                       0.125 * (1 + r) * (1 + s) * (1 - t), 0.125 * (1 - r) * (1 + s) * (1 - t),
                       0.25 * (1 - s) * (1 + t), 0.25 * (1 + s) * (1 + t)};
  const float* v = P->vertices + (size_t)e * 18;
  for (int c = 0; c < 3; ++c) {
    x[c] = 0.0;
    for (int q = 0; q < 6; ++q) x[c] += w[q] * v[3 * q + c];
  }
}

ORC_EXPORT double orc_prism_field(const orc_prisms_t* P, int e, const double* rst)
{
  if (!P->field) return 0.0;
  return prism_sum(P->field + (size_t)e * orc_prism_count(P->fm), P->fm, rst);
}

/* ------------------------------------------------------------------------ */
/* Curved triangles (fem/TriElement.cu).  High-order: the sum over          */
/* Modified_A(i, r) Modified_B(i, j, s), i slowest, j < my - i              */
/* (EvaluateTriModal, TriElement.cu:46-63).  Without geometry coefficients: */
/* the flat triangle v0 v1 v2 in the same collapsed coordinates,            */
/* barycentric weights of v1, v2 = (1+r)(1-s)/4, (1+s)/2.                   */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n;
  int gm[2], fm[2];
  int pad_;
  const float* vertices; /* n x 3 x 3 */
  const float* geom;     /* n x 3 x count(gm), or NULL */
  const float* field;    /* n x count(fm), or NULL */
} orc_tris_t;

ORC_EXPORT int orc_tri_count(const int* m)
{
  int c = 0;
  for (int i = 0; i < m[0]; ++i) c += m[1] > i ? m[1] - i : 0;
  return c;
}

static double tri_sum(const float* c, const int* m, double r, double s)
{
  double v = 0.0;
  int idx = 0;
  for (int i = 0; i < m[0]; ++i) {
    const double a = orc_modA(i, r);
    for (int j = 0; j + i < m[1]; ++j) v += c[idx++] * a * orc_modB(i, j, s);
  }
  return v;
}

ORC_EXPORT void orc_tri_map(const orc_tris_t* T, int e, double r, double s, double* x)
{
  if (T->geom) {
    const int nc = orc_tri_count(T->gm);
    for (int c = 0; c < 3; ++c) x[c] = tri_sum(T->geom + ((size_t)e * 3 + c) * nc, T->gm, r, s);
    return;
  }
  const double b1 = 0.25 * (1 + r) * (1 - s), b2 = 0.5 * (1 + s), b0 = 1 - b1 - b2;
  const float* v = T->vertices + (size_t)e * 9;
  for (int c = 0; c < 3; ++c) x[c] = b0 * v[c] + b1 * v[3 + c] + b2 * v[6 + c];
}

ORC_EXPORT double orc_tri_field(const orc_tris_t* T, int e, double r, double s)
{
  if (!T->field) return 0.0;
  return tri_sum(T->field + (size_t)e * orc_tri_count(T->fm), T->fm, r, s);
}
