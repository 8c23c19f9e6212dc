// femrt_point.hpp -- point kernels shared by the library's entry points:
// field value + world gradient at (element, rst) points.
// Reference: fem/HexElement.cu:274-330 (EvaluateModalField,
// EvaluateModalGradientWorld).
#pragma once

#include <hip/hip_runtime.h>

#include "femrt_math.hpp"

// out is 4 x n SoA (u, du/dx, du/dy, du/dz); rst is 3 x n SoA.
template <typename T, int NG, int NF>
__global__ void k_eval_field(int n, const int* __restrict__ elem, const float* __restrict__ rst,
                             const femrt::GeoMode* __restrict__ geom, const float* __restrict__ field,
                             T* __restrict__ out)
{
  using namespace femrt;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int e = elem[p];
  const T r = rst[p], s = rst[n + p], t = rst[2 * n + p];
  T u, gref[3], X[3], J[9], gw[3];
  field_eval<T, NF>(field + (size_t)e * (NF * NF * NF), r, s, t, &u, gref);
  geom_eval<T, NG>(geom + (size_t)e * (NG * NG * NG), r, s, t, X, J);
  if (!grad_world<T>(J, gref, gw)) gw[0] = gw[1] = gw[2] = T(NAN);
  out[p] = u;
  out[n + p] = gw[0];
  out[2 * n + p] = gw[1];
  out[3 * n + p] = gw[2];
}
