// femrt_math.hpp -- element math shared by the host and the gfx950 kernels.
//
// Modified_A basis tables by recurrence, sum-factorised evaluation of the
// modal geometry map x(r,s,t) with its Jacobian, and of the scalar field with
// its reference gradient.  Templated on the arithmetic type T (float for the
// bulk, double for the polish) and on the per-direction mode counts, so the
// loops unroll and the 1-D tables live in registers.
//
// Reference: fem/ModalBasis.cuh:16-70 (basis), fem/HexElement.cu:8-330 (modal
// sums, map, Jacobian, inverse), StdHexExp.cpp:199-245 (the sum-factorisation
// the reference's own spec describes but its device code does not use).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define FEMRT_HD __host__ __device__ __forceinline__

namespace femrt {

// Modified_A basis of n modes and its derivative at x.
//   phi_0 = (1-x)/2, phi_1 = (1+x)/2, phi_p = (1-x^2)/4 P^{(1,1)}_{p-2}(x).
// P^{(1,1)}_k = A_k x P_{k-1} - B_k P_{k-2} with
//   A_k = (2k+1)(k+1)/(k(k+2)),  B_k = (k+1)/(k+2)   (k >= 2),
// P_0 = 1, P_1 = 2x; derivatives follow by differentiating the recurrence.
template <typename T, int N>
FEMRT_HD void basis1d(T x, T* __restrict__ phi, T* __restrict__ dphi)
{
  phi[0] = T(0.5) - T(0.5) * x;
  dphi[0] = T(-0.5);
  if (N > 1) {
    phi[1] = T(0.5) + T(0.5) * x;
    dphi[1] = T(0.5);
  }
  if (N > 2) {
    const T b = T(0.25) - T(0.25) * x * x;
    const T db = T(-0.5) * x;
    T pm2 = T(1), dpm2 = T(0);
    T pm1 = T(2) * x, dpm1 = T(2);
    phi[2] = b;
    dphi[2] = db;
    if (N > 3) {
      phi[3] = b * pm1;
      dphi[3] = db * pm1 + b * dpm1;
    }
#pragma unroll
    for (int p = 4; p < N; ++p) {
      const int k = p - 2;
      const T A = T((2.0 * k + 1.0) * (k + 1.0) / (k * (k + 2.0)));
      const T B = T((k + 1.0) / (k + 2.0));
      const T pk = A * x * pm1 - B * pm2;
      const T dpk = A * (pm1 + x * dpm1) - B * dpm2;
      phi[p] = b * pk;
      dphi[p] = db * pk + b * dpk;
      pm2 = pm1;
      dpm2 = dpm1;
      pm1 = pk;
      dpm1 = dpk;
    }
  }
}

// Recurrence coefficients of P^{(1,1)}_k, k = 2..15, as constants (the rolled
// loops below would otherwise divide at every step).
FEMRT_HD double jacobi11_A(int k)
{
  constexpr double t[16] = {0.0, 0.0,
                            5.0 * 3.0 / (2.0 * 4.0), 7.0 * 4.0 / (3.0 * 5.0), 9.0 * 5.0 / (4.0 * 6.0),
                            11.0 * 6.0 / (5.0 * 7.0), 13.0 * 7.0 / (6.0 * 8.0), 15.0 * 8.0 / (7.0 * 9.0),
                            17.0 * 9.0 / (8.0 * 10.0), 19.0 * 10.0 / (9.0 * 11.0), 21.0 * 11.0 / (10.0 * 12.0),
                            23.0 * 12.0 / (11.0 * 13.0), 25.0 * 13.0 / (12.0 * 14.0), 27.0 * 14.0 / (13.0 * 15.0),
                            29.0 * 15.0 / (14.0 * 16.0), 31.0 * 16.0 / (15.0 * 17.0)};
  return t[k];
}
FEMRT_HD double jacobi11_B(int k)
{
  constexpr double t[16] = {0.0, 0.0, 3.0 / 4.0, 4.0 / 5.0, 5.0 / 6.0, 6.0 / 7.0, 7.0 / 8.0, 8.0 / 9.0,
                            9.0 / 10.0, 10.0 / 11.0, 11.0 / 12.0, 12.0 / 13.0, 13.0 / 14.0, 14.0 / 15.0,
                            15.0 / 16.0, 16.0 / 17.0};
  return t[k];
}

// Modified_A basis generated one mode at a time (phi_p, phi'_p for
// p = 0, 1, 2, ...), so rolled loops over modes keep only the recurrence state
// in registers instead of whole tables.
template <typename T>
struct BasisIter {
  T x, b, db;
  T P0, P1, dP0, dP1;  // P_{k-1}, P_k of the (1,1) Jacobi recurrence and derivatives
  int p;
  FEMRT_HD void init(T x_)
  {
    x = x_;
    b = T(0.25) - T(0.25) * x * x;
    db = T(-0.5) * x;
    P0 = T(0);
    dP0 = T(0);
    P1 = T(1);
    dP1 = T(0);
    p = 0;
  }
  FEMRT_HD void next(T& phi, T& dphi)
  {
    if (p == 0) {
      phi = T(0.5) - T(0.5) * x;
      dphi = T(-0.5);
    } else if (p == 1) {
      phi = T(0.5) + T(0.5) * x;
      dphi = T(0.5);
    } else {
      if (p >= 3) {
        // advance P_{k} -> P_{k+1}, k = p - 3
        const int k = p - 2;  // index of the polynomial produced
        T Pn, dPn;
        if (k == 1) {
          Pn = T(2) * x;
          dPn = T(2);
        } else {
          const T A = T(jacobi11_A(k)), B = T(jacobi11_B(k));
          Pn = A * x * P1 - B * P0;
          dPn = A * (P1 + x * dP1) - B * dP0;
        }
        P0 = P1;
        dP0 = dP1;
        P1 = Pn;
        dP1 = dPn;
      }
      phi = b * P1;
      dphi = db * P1 + b * dP1;
    }
    ++p;
  }
};

// One geometry mode: the x, y and z coefficients (12 bytes, one dwordx3 load).
struct GeoMode {
  float x, y, z;
};

// x(r,s,t) and J (row-major, rows x,y,z; columns d/dr, d/ds, d/dt).
// c: NG^3 GeoMode for one element, mode (i,j,k) at (k*NG+j)*NG+i.  1035 FMA at NG=5.  The k and j loops stay rolled (the s and
// t basis is generated by the recurrence as they advance) so only NG
// coefficient triples are live at a time.
template <typename T, int NG>
FEMRT_HD void geom_eval(const GeoMode* __restrict__ c, T r, T s, T t, T* __restrict__ X,
                        T* __restrict__ J)
{
  T pr[NG], dpr[NG];
  basis1d<T, NG>(r, pr, dpr);
  T ps[NG], dps[NG];
  basis1d<T, NG>(s, ps, dps);
  T x0 = 0, x1 = 0, x2 = 0;
  T jr0 = 0, jr1 = 0, jr2 = 0, js0 = 0, js1 = 0, js2 = 0, jt0 = 0, jt1 = 0, jt2 = 0;
  BasisIter<T> bt;
  bt.init(t);
#pragma unroll 1
  for (int k = 0; k < NG; ++k) {
    T tk, dtk;
    bt.next(tk, dtk);
    T b0 = 0, b1 = 0, b2 = 0;     // sum_j phi_j(s) a_jk
    T bs0 = 0, bs1 = 0, bs2 = 0;  // sum_j phi'_j(s) a_jk
    T br0 = 0, br1 = 0, br2 = 0;  // sum_j phi_j(s) a'_jk
#pragma unroll
    for (int j = 0; j < NG; ++j) {
      const T sj = ps[j], dsj = dps[j];
      const GeoMode* cj = c + (k * NG + j) * NG;
      T a0 = 0, a1 = 0, a2 = 0, ad0 = 0, ad1 = 0, ad2 = 0;
#pragma unroll
      for (int i = 0; i < NG; ++i) {
        const GeoMode q = cj[i];
        a0 += pr[i] * T(q.x);
        a1 += pr[i] * T(q.y);
        a2 += pr[i] * T(q.z);
        ad0 += dpr[i] * T(q.x);
        ad1 += dpr[i] * T(q.y);
        ad2 += dpr[i] * T(q.z);
      }
      b0 += sj * a0;
      b1 += sj * a1;
      b2 += sj * a2;
      bs0 += dsj * a0;
      bs1 += dsj * a1;
      bs2 += dsj * a2;
      br0 += sj * ad0;
      br1 += sj * ad1;
      br2 += sj * ad2;
    }
    x0 += tk * b0;
    x1 += tk * b1;
    x2 += tk * b2;
    jr0 += tk * br0;
    jr1 += tk * br1;
    jr2 += tk * br2;
    js0 += tk * bs0;
    js1 += tk * bs1;
    js2 += tk * bs2;
    jt0 += dtk * b0;
    jt1 += dtk * b1;
    jt2 += dtk * b2;
  }
  X[0] = x0;
  X[1] = x1;
  X[2] = x2;
  J[0] = jr0; J[1] = js0; J[2] = jt0;
  J[3] = jr1; J[4] = js1; J[5] = jt1;
  J[6] = jr2; J[7] = js2; J[8] = jt2;
}

// x(r,s,t) only (no Jacobian): 3 NG^3 + 3 NG^2 + 3 NG FMA.
template <typename T, int NG>
FEMRT_HD void geom_point(const GeoMode* __restrict__ c, T r, T s, T t, T* __restrict__ X)
{
  T pr[NG], dpr[NG];
  basis1d<T, NG>(r, pr, dpr);
  T x0 = 0, x1 = 0, x2 = 0;
  BasisIter<T> bt;
  bt.init(t);
#pragma unroll 1
  for (int k = 0; k < NG; ++k) {
    T tk, dtk;
    bt.next(tk, dtk);
    T b0 = 0, b1 = 0, b2 = 0;
    BasisIter<T> bs;
    bs.init(s);
#pragma unroll 1
    for (int j = 0; j < NG; ++j) {
      T sj, dsj;
      bs.next(sj, dsj);
      const GeoMode* cj = c + (k * NG + j) * NG;
      T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
      for (int i = 0; i < NG; ++i) {
        const GeoMode q = cj[i];
        a0 += pr[i] * T(q.x);
        a1 += pr[i] * T(q.y);
        a2 += pr[i] * T(q.z);
      }
      b0 += sj * a0;
      b1 += sj * a1;
      b2 += sj * a2;
    }
    x0 += tk * b0;
    x1 += tk * b1;
    x2 += tk * b2;
  }
  X[0] = x0;
  X[1] = x1;
  X[2] = x2;
}

// Field value and reference gradient (du/dr, du/ds, du/dt): 2 NF^3 + 3 NF^2 +
// 4 NF FMA.
template <typename T, int NF>
FEMRT_HD void field_eval(const float* __restrict__ c, T r, T s, T t, T* __restrict__ u,
                         T* __restrict__ g)
{
  T pr[NF], dpr[NF];
  basis1d<T, NF>(r, pr, dpr);
  T v = 0, gr = 0, gs = 0, gt = 0;
  BasisIter<T> bt;
  bt.init(t);
#pragma unroll 1
  for (int k = 0; k < NF; ++k) {
    T tk, dtk;
    bt.next(tk, dtk);
    T b = 0, bs = 0, br = 0;
    BasisIter<T> bsi;
    bsi.init(s);
#pragma unroll 1
    for (int j = 0; j < NF; ++j) {
      T sj, dsj;
      bsi.next(sj, dsj);
      const float* cj = c + (k * NF + j) * NF;
      T a = 0, ad = 0;
#pragma unroll
      for (int i = 0; i < NF; ++i) {
        const T q = T(cj[i]);
        a += pr[i] * q;
        ad += dpr[i] * q;
      }
      b += sj * a;
      bs += dsj * a;
      br += sj * ad;
    }
    v += tk * b;
    gr += tk * br;
    gs += tk * bs;
    gt += dtk * b;
  }
  *u = v;
  g[0] = gr;
  g[1] = gs;
  g[2] = gt;
}

// Field value only: NF^3 + NF^2 + NF FMA.
template <typename T, int NF>
FEMRT_HD T field_value(const float* __restrict__ c, T r, T s, T t)
{
  T pr[NF], dpr[NF];
  basis1d<T, NF>(r, pr, dpr);
  T ps[NF], dps[NF];
  basis1d<T, NF>(s, ps, dps);
  T v = 0;
  BasisIter<T> bt;
  bt.init(t);
#pragma unroll 1
  for (int k = 0; k < NF; ++k) {
    T tk, dtk;
    bt.next(tk, dtk);
    T b = 0;
#pragma unroll
    for (int j = 0; j < NF; ++j) {
      const float* cj = c + (k * NF + j) * NF;
      T a = 0;
#pragma unroll
      for (int i = 0; i < NF; ++i) a += pr[i] * T(cj[i]);
      b += ps[j] * a;
    }
    v += tk * b;
  }
  return v;
}

template <typename T>
FEMRT_HD T det3(const T* J)
{
  return J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
         J[2] * (J[3] * J[7] - J[4] * J[6]);
}

// Solve J d = b by Cramer's rule (the cofactor inverse of HexElement.cu:208-225
// applied to one right-hand side).  Returns false when det is zero/non-finite.
template <typename T>
FEMRT_HD bool solve3(const T* J, const T* b, T* d)
{
  const T c00 = J[4] * J[8] - J[5] * J[7];
  const T c01 = J[5] * J[6] - J[3] * J[8];
  const T c02 = J[3] * J[7] - J[4] * J[6];
  const T det = J[0] * c00 + J[1] * c01 + J[2] * c02;
  if (!(det != T(0)) || !(det == det)) return false;
  const T id = T(1) / det;
  const T c10 = J[2] * J[7] - J[1] * J[8];
  const T c11 = J[0] * J[8] - J[2] * J[6];
  const T c12 = J[1] * J[6] - J[0] * J[7];
  const T c20 = J[1] * J[5] - J[2] * J[4];
  const T c21 = J[2] * J[3] - J[0] * J[5];
  const T c22 = J[0] * J[4] - J[1] * J[3];
  d[0] = (c00 * b[0] + c10 * b[1] + c20 * b[2]) * id;
  d[1] = (c01 * b[0] + c11 * b[1] + c21 * b[2]) * id;
  d[2] = (c02 * b[0] + c12 * b[1] + c22 * b[2]) * id;
  return true;
}

// world gradient = J^{-T} grad_ref (HexElement.cu:308-330)
template <typename T>
FEMRT_HD bool grad_world(const T* J, const T* gref, T* gw)
{
  // Solve J^T gw = gref
  const T JT[9] = {J[0], J[3], J[6], J[1], J[4], J[7], J[2], J[5], J[8]};
  return solve3<T>(JT, gref, gw);
}

// Face table, Nektar order (high_order.h:172-184):
//   face 0: t=-1 (u=r, v=s)   1: s=-1 (r,t)   2: r=+1 (s,t)
//   face 3: s=+1 (r,t)        4: r=-1 (s,t)   5: t=+1 (r,s)
FEMRT_HD int face_fix(int f) { return (f == 0 || f == 5) ? 2 : ((f == 1 || f == 3) ? 1 : 0); }
FEMRT_HD float face_val(int f) { return (f == 0 || f == 1 || f == 4) ? -1.f : 1.f; }
FEMRT_HD int face_u(int f) { return (f == 2 || f == 4) ? 1 : 0; }
FEMRT_HD int face_v(int f) { return (f == 0 || f == 5) ? 1 : 2; }

// Slab test (geometry.h:162-175, high_order.h:115-152).
FEMRT_HD bool slab(const float* o, const float* inv_d, const float* lo, const float* hi, float tmin,
                   float tmax, float& t0, float& t1)
{
  float a = tmin, b = tmax;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float ta = (lo[c] - o[c]) * inv_d[c];
    float tb = (hi[c] - o[c]) * inv_d[c];
    const float mn = fminf(ta, tb), mx = fmaxf(ta, tb);
    a = fmaxf(a, mn);
    b = fminf(b, mx);
  }
  t0 = a;
  t1 = b;
  return a <= b;
}

}  // namespace femrt
