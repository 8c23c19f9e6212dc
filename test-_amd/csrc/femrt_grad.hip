// femrt_grad.hip -- the field's world gradient at the isosurface crossings
// of a trace: the iso normal of the reference's shading, n = +-grad u /
// |grad u| (raytracer.cu:207-216, fem/HexElement.cu:274-330 for the
// gradient).
//
// Two engines with the same result:
//   k_grad_valu<NG,NF>  one crossing per lane, sum-factorised on the vector
//                       ALUs (field value + reference gradient, geometry
//                       Jacobian, J^-T);
//   k_grad_mfma<NG,NF>  crossings binned by element (collected, then sorted
//                       by element id), 16 consecutive crossings per wave
//                       step.  For each element among the 16 its field
//                       coefficients are staged in LDS, and the contraction
//                       over the r modes, D_v[(j,k)][m] = sum_i C[k][j][i]
//                       A_v[m][i] with A_0 = phi_i(r_m) and A_1 =
//                       phi_i'(r_m), runs on the FP32 matrix cores
//                       (v_mfma_f32_16x16x4f32: 16 (j,k) modes x 16
//                       crossings per instruction, 4 r modes deep), and so
//                       does the same contraction of the geometry's three
//                       components for the Jacobian.  The (j, k) sums with
//                       the s and t basis, and J^-T, run on the vector ALUs.  This is the sum factorisation of
//                       StdHexExp's backward transform (StdHexExp.cpp:
//                       199-245), its first stage as a dense product.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "femrt_internal.hpp"
#include "femrt_math.hpp"
#include "femrt_runtime.hpp"

namespace {

using femrt::GeoMode;
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void wave_sync_lds()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__global__ void k_grad_fill(float* __restrict__ out, size_t m)
{
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x)
    out[i] = NAN;
}

// (element, slot) of every crossing of the last trace, slot = k n + r.
__global__ void k_grad_collect(int n, int n_iso, const int* __restrict__ hit_queue, const int* __restrict__ counts,
                               const int* __restrict__ iso_elem, int* __restrict__ keys, int* __restrict__ vals,
                               int* __restrict__ total)
{
  const int count = counts[0] * n_iso;
  const int base = blockIdx.x * blockDim.x;
  for (int q0 = base; q0 < count; q0 += gridDim.x * blockDim.x) {   // block-uniform trip count
    const int q = q0 + threadIdx.x;
    int slot = -1, e = -1;
    if (q < count) {
      const int r = hit_queue[q / n_iso], k = q % n_iso;
      slot = k * n + r;
      e = iso_elem[slot];
    }
    const int at = wave_append(e >= 0, total);
    if (e >= 0) {
      keys[at] = e;
      vals[at] = slot;
    }
  }
}

template <int NG>
__device__ void grad_out(int n, const GeoMode* __restrict__ gc, float r, float s, float t, const float* gref,
                         int slot, float* __restrict__ out)
{
  float X[3], J[9], gw[3];
  femrt::geom_eval<float, NG>(gc, r, s, t, X, J);
  if (!femrt::grad_world<float>(J, gref, gw)) gw[0] = gw[1] = gw[2] = NAN;
  const int k = slot / n, rr = slot - k * n;
  for (int c = 0; c < 3; ++c) out[((size_t)k * 3 + c) * n + rr] = gw[c];
}

template <int NG, int NF>
__global__ __launch_bounds__(256) void k_grad_valu(int n, const GeoMode* __restrict__ geom,
                                                   const float* __restrict__ field, const int* __restrict__ keys,
                                                   const int* __restrict__ vals, int total,
                                                   const double* __restrict__ iso_rst, float* __restrict__ out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int e = keys[i], slot = vals[i];
  const int k = slot / n, r = slot - k * n;
  const float rr = (float)iso_rst[(size_t)k * 3 * n + r];
  const float ss = (float)iso_rst[(size_t)k * 3 * n + n + r];
  const float tt = (float)iso_rst[(size_t)k * 3 * n + 2 * (size_t)n + r];
  float u, gref[3];
  femrt::field_eval<float, NF>(field + (size_t)e * (NF * NF * NF), rr, ss, tt, &u, gref);
  grad_out<NG>(n, geom + (size_t)e * (NG * NG * NG), rr, ss, tt, gref, slot, out);
}

// Per wave step: the 16 sorted crossings [16 w, 16 w + 16) (each lane of the
// first 16 holds one), processed element by element (sorted, so mostly one
// element per step).
template <int NG, int NF>
__global__ __launch_bounds__(256) void k_grad_mfma(int n, const GeoMode* __restrict__ geom,
                                                   const float* __restrict__ field, const int* __restrict__ keys,
                                                   const int* __restrict__ vals, int total,
                                                   const double* __restrict__ iso_rst, float* __restrict__ out)
{
  constexpr int NF2 = NF * NF, NF3 = NF2 * NF;
  constexpr int NT = (NF2 + 15) / 16;  // 16-row tiles of the (j, k) modes
  constexpr int KS = (NF + 3) / 4;     // 4-deep steps over the r modes
  constexpr int NG2 = NG * NG, NG3 = NG2 * NG;
  constexpr int NTG = (NG2 + 15) / 16;
  constexpr int KSG = (NG + 3) / 4;
  __shared__ float s_coef[4][NF3];
  __shared__ float s_geo[4][3][NG3];     // x, y, z modes of the element
  __shared__ float s_tab[4][6][16][NF];  // phi, phi' in r, s, t of each crossing (field)
  __shared__ float s_tabg[4][6][16][NG]; // the same for the geometry's mode count
  __shared__ float s_res[4][16][12];     // u_r, u_s, u_t, then J row-major
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int steps = (total + 15) / 16;
  for (int w = blockIdx.x * 4 + wave; w < steps; w += gridDim.x * 4) {   // wave-uniform
    const int i0 = 16 * w;
    const bool have = lane < 16 && i0 + lane < total;
    const int my_e = have ? keys[i0 + lane] : -1;
    const int my_slot = have ? vals[i0 + lane] : -1;
    float rst[3] = {0.f, 0.f, 0.f};
    if (have) {
      const int k = my_slot / n, r = my_slot - k * n;
      for (int c = 0; c < 3; ++c) rst[c] = (float)iso_rst[(size_t)k * 3 * n + (size_t)c * n + r];
    }
    // basis tables of the 16 crossings: lane (dir, m), dir 0..2 = r, s, t
    {
      const int m = lane & 15, dir = lane >> 4;
      const float x0 = __shfl(rst[0], m), x1 = __shfl(rst[1], m), x2 = __shfl(rst[2], m);
      if (lane < 48) {
        const float x = dir == 0 ? x0 : (dir == 1 ? x1 : x2);
        float ph[NF], dph[NF];
        femrt::basis1d<float, NF>(x, ph, dph);
        for (int i = 0; i < NF; ++i) {
          s_tab[wave][2 * dir][m][i] = ph[i];
          s_tab[wave][2 * dir + 1][m][i] = dph[i];
        }
        float pg[NG], dpg[NG];
        femrt::basis1d<float, NG>(x, pg, dpg);
        for (int i = 0; i < NG; ++i) {
          s_tabg[wave][2 * dir][m][i] = pg[i];
          s_tabg[wave][2 * dir + 1][m][i] = dpg[i];
        }
      }
    }
    unsigned long long todo = __ballot(have);
    while (todo) {   // wave-uniform
      const int lead = __ffsll((long long)todo) - 1;
      const int e = __shfl(my_e, lead);
      const unsigned long long mine = __ballot(have && my_e == e);
      todo &= ~mine;
      const float* fc = field + (size_t)e * NF3;
      for (int q = lane; q < NF3; q += 64) s_coef[wave][q] = fc[q];
      const GeoMode* gc = geom + (size_t)e * NG3;
      for (int q = lane; q < NG3; q += 64) {
        const GeoMode g = gc[q];
        s_geo[wave][0][q] = g.x;
        s_geo[wave][1][q] = g.y;
        s_geo[wave][2][q] = g.z;
      }
      wave_sync_lds();
      // stage 1 on the matrix cores, transposed so that each output column
      // is one crossing: D_v[(j,k)][m] = sum_i C[k][j][i] A_v[m][i].  Lane l
      // supplies A = C[(j,k) = 16 nt + l % 16][i] and B = A_v[m = l % 16][i]
      // at depth i = 4 st + l / 16, and receives D_v rows 4 (l / 16) + q of
      // column l % 16 (zero columns for crossings of other elements).
      const int m = lane & 15, ak = lane >> 4;
      const bool col = (mine >> m) & 1ull;
      float part[4] = {0.f, 0.f, 0.f, 0.f};   // u, du/dr, du/ds, du/dt of crossing m, this lane's rows
      for (int nt = 0; nt < NT; ++nt) {
        const int jka = 16 * nt + m;
        f32x4 d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
        for (int st = 0; st < KS; ++st) {
          const int i = 4 * st + ak;
          const bool ok = i < NF;
          const float av = (ok && jka < NF2) ? s_coef[wave][jka * NF + i] : 0.f;
          const float b0 = (ok && col) ? s_tab[wave][0][m][i] : 0.f;
          const float b1 = (ok && col) ? s_tab[wave][1][m][i] : 0.f;
          d0 = __builtin_amdgcn_mfma_f32_16x16x4f32(av, b0, d0, 0, 0, 0);
          d1 = __builtin_amdgcn_mfma_f32_16x16x4f32(av, b1, d1, 0, 0, 0);
        }
        // stage 2: the (j, k) sums of this lane's rows with crossing m's s
        // and t basis
        for (int q = 0; q < 4; ++q) {
          const int jk = 16 * nt + 4 * ak + q;
          if (jk < NF2) {
            const int j = jk % NF, kk = jk / NF;
            const float bj = s_tab[wave][2][m][j], dbj = s_tab[wave][3][m][j];
            const float ck = s_tab[wave][4][m][kk], dck = s_tab[wave][5][m][kk];
            part[0] += d0[q] * bj * ck;
            part[1] += d1[q] * bj * ck;
            part[2] += d0[q] * dbj * ck;
            part[3] += d0[q] * bj * dck;
          }
        }
      }
      // the geometry's Jacobian alike: for each component c, stage 1 with
      // the r basis and its derivative, stage 2 with the s and t basis:
      // J[c][0] = sum D'_c b c, J[c][1] = sum D_c b' c, J[c][2] = sum D_c b c'
      float jac[9];
      for (int q = 0; q < 9; ++q) jac[q] = 0.f;
      for (int nt = 0; nt < NTG; ++nt) {
        const int jka = 16 * nt + m;
        f32x4 g0[3], g1[3];
        for (int c = 0; c < 3; ++c) {
          g0[c] = f32x4{0.f, 0.f, 0.f, 0.f};
          g1[c] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
        for (int st = 0; st < KSG; ++st) {
          const int i = 4 * st + ak;
          const bool ok = i < NG;
          const float b0 = (ok && col) ? s_tabg[wave][0][m][i] : 0.f;
          const float b1 = (ok && col) ? s_tabg[wave][1][m][i] : 0.f;
          for (int c = 0; c < 3; ++c) {
            const float av = (ok && jka < NG2) ? s_geo[wave][c][jka * NG + i] : 0.f;
            g0[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, b0, g0[c], 0, 0, 0);
            g1[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, b1, g1[c], 0, 0, 0);
          }
        }
        for (int q = 0; q < 4; ++q) {
          const int jk = 16 * nt + 4 * ak + q;
          if (jk < NG2) {
            const int j = jk % NG, kk = jk / NG;
            const float bj = s_tabg[wave][2][m][j], dbj = s_tabg[wave][3][m][j];
            const float ck = s_tabg[wave][4][m][kk], dck = s_tabg[wave][5][m][kk];
            for (int c = 0; c < 3; ++c) {
              jac[3 * c + 0] += g1[c][q] * bj * ck;
              jac[3 * c + 1] += g0[c][q] * dbj * ck;
              jac[3 * c + 2] += g0[c][q] * bj * dck;
            }
          }
        }
      }
      // the four 16-lane groups hold the rows of one crossing's column
      for (int o = 0; o < 4; ++o) {
        part[o] += __shfl_xor(part[o], 16);
        part[o] += __shfl_xor(part[o], 32);
      }
      for (int o = 0; o < 9; ++o) {
        jac[o] += __shfl_xor(jac[o], 16);
        jac[o] += __shfl_xor(jac[o], 32);
      }
      if (lane < 16) {
        for (int o = 0; o < 3; ++o) s_res[wave][lane][o] = part[o + 1];
        for (int o = 0; o < 9; ++o) s_res[wave][lane][3 + o] = jac[o];
      }
      wave_sync_lds();
      if (lane < 16 && ((mine >> lane) & 1ull)) {
        const float gref[3] = {s_res[wave][lane][0], s_res[wave][lane][1], s_res[wave][lane][2]};
        float J[9], gw[3];
        for (int o = 0; o < 9; ++o) J[o] = s_res[wave][lane][3 + o];
        if (!femrt::grad_world<float>(J, gref, gw)) gw[0] = gw[1] = gw[2] = NAN;
        const int k = my_slot / n, rr = my_slot - k * n;
        for (int c = 0; c < 3; ++c) out[((size_t)k * 3 + c) * n + rr] = gw[c];
      }
      wave_sync_lds();  // the next element rewrites the coefficients and results
    }
    wave_sync_lds();    // the next step rewrites the basis tables
  }
}

int bits_for(int n)
{
  int b = 1;
  while (b < 31 && (1 << b) < n) ++b;
  return b;
}

}  // namespace

extern "C" {

FEMRT_API int femrt_frame_iso_gradients(femrt_mesh* m, femrt_frame* f, int n_iso, int engine, float* out)
{
  if (!m || !f || !out || n_iso < 1 || n_iso > FEMRT_MAX_ISO || (engine != FEMRT_GRAD_MFMA && engine != FEMRT_GRAD_VALU))
    return FEMRT_EINVAL;
  femrt_ctx* c = m->ctx;
  if (f->ctx != c) return fail(c, FEMRT_EINVAL, "iso_gradients: mesh and frame belong to different contexts");
  if ((size_t)n_iso * f->n > (size_t)0x7FFFFFFF)
    return fail(c, FEMRT_EUNSUP, "iso_gradients: more than 2^31 (isovalue, ray) slots");
  HIPCHK(c, hipSetDevice(c->device));
  const int n = f->n;
  const size_t nout = (size_t)n_iso * 3 * n;
  const size_t cap = (size_t)n_iso * n;
  float* d_out = nullptr;
  int *d_keys = nullptr, *d_vals = nullptr, *d_keys2 = nullptr, *d_vals2 = nullptr, *d_total = nullptr;
  void* d_tmp = nullptr;
  auto cleanup = [&]() {
    // best-effort frees of the call's temporaries
    (void)hipFree(d_out);
    (void)hipFree(d_keys);
    (void)hipFree(d_vals);
    (void)hipFree(d_keys2);
    (void)hipFree(d_vals2);
    (void)hipFree(d_total);
    (void)hipFree(d_tmp);
  };
  if (hipMalloc(&d_out, nout * 4) != hipSuccess || hipMalloc(&d_keys, cap * 4) != hipSuccess ||
      hipMalloc(&d_vals, cap * 4) != hipSuccess || hipMalloc(&d_total, 4) != hipSuccess) {
    cleanup();
    return fail(c, FEMRT_ENOMEM, "iso_gradients: device allocation failed");
  }
  int rc = FEMRT_OK;
  const int blocks = 4 * c->cus;
  rc = timed_launch(c, "k_grad_collect", [&] {
    k_grad_fill<<<blocks, 256, 0, c->stream>>>(d_out, nout);
    (void)hipMemsetAsync(d_total, 0, 4, c->stream);
    k_grad_collect<<<blocks, 256, 0, c->stream>>>(n, n_iso, f->d_queue_hit, f->d_counts, f->d_iso_elem, d_keys,
                                                  d_vals, d_total);
  });
  int total = 0;
  if (rc == FEMRT_OK && (hipMemcpyAsync(&total, d_total, 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                         hipStreamSynchronize(c->stream) != hipSuccess))
    rc = fail(c, FEMRT_EHIP, "iso_gradients: count download failed");
  if (rc == FEMRT_OK && total > 0) {
    const int* keys = d_keys;
    const int* vals = d_vals;
    if (engine == FEMRT_GRAD_MFMA) {
      // bin by element: sort the (element, slot) pairs by element
      size_t tmp_bytes = 0;
      const int eb = bits_for(m->n_elem);
      if (hipMalloc(&d_keys2, (size_t)total * 4) != hipSuccess || hipMalloc(&d_vals2, (size_t)total * 4) != hipSuccess ||
          rocprim::radix_sort_pairs(nullptr, tmp_bytes, d_keys, d_keys2, d_vals, d_vals2, total, 0, eb, c->stream) !=
              hipSuccess ||
          hipMalloc(&d_tmp, std::max<size_t>(tmp_bytes, 4)) != hipSuccess)
        rc = fail(c, FEMRT_ENOMEM, "iso_gradients: sort storage");
      if (rc == FEMRT_OK)
        rc = timed_launch(c, "k_grad_sort", [&] {
          (void)rocprim::radix_sort_pairs(d_tmp, tmp_bytes, d_keys, d_keys2, d_vals, d_vals2, total, 0, eb, c->stream);
        });
      keys = d_keys2;
      vals = d_vals2;
    }
    if (rc == FEMRT_OK) {
      rc = FEMRT_EUNSUP;
      FEMRT_DISPATCH_NG(m->NG, {
        FEMRT_DISPATCH_NF(m->NF, {
          if (engine == FEMRT_GRAD_MFMA) {
            const int steps = (total + 15) / 16;
            const int gb = std::max(1, std::min((steps + 3) / 4, 8 * c->cus));
            rc = timed_launch(c, "k_grad_mfma", [&] {
              k_grad_mfma<NG_, NF_><<<gb, 256, 0, c->stream>>>(n, m->d_geom, m->d_field, keys, vals, total,
                                                                 f->d_iso_rst, d_out);
            });
          } else {
            rc = timed_launch(c, "k_grad_valu", [&] {
              k_grad_valu<NG_, NF_><<<(total + 255) / 256, 256, 0, c->stream>>>(n, m->d_geom, m->d_field, keys, vals,
                                                                                 total, f->d_iso_rst, d_out);
            });
          }
        });
      });
    }
  }
  if (rc == FEMRT_OK && (hipMemcpyAsync(out, d_out, nout * 4, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                         hipStreamSynchronize(c->stream) != hipSuccess))
    rc = fail(c, FEMRT_EHIP, "iso_gradients: download failed");
  cleanup();
  return rc;
}

}  // extern "C"
