// femrt_noalloc.hip -- field evaluation at points without any device
// allocation, for callers that cannot allocate (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// femrt_eval_field stages its host inputs and outputs in device buffers it
// allocates per call.  This variant stages them in `scratch`, a file-scope
// device array of fixed size that each GPU holds from the first use until the
// process ends: element ids (4 B), rst (12 B) and the 4 results (up to 32 B)
// per point, FEMRT_NOALLOC_MAX_POINTS points, 3 MiB.  Every call on a GPU
// shares that GPU's array, so calls that overlap on one GPU (two contexts of
// the same device used from two host threads) race; calls on one context are
// ordered by its stream.  Nothing in the array needs to start at zero: each
// call writes the inputs it reads and every output it copies back.
#include <hip/hip_runtime.h>

#include <mutex>

#include "femrt_internal.hpp"
#include "femrt_point.hpp"
#include "femrt_runtime.hpp"

#define FEMRT_NOALLOC_MAX_DEVICES 64
static constexpr size_t kMaxPts = FEMRT_NOALLOC_MAX_POINTS;
// layout: elem int32 [kMaxPts] | rst float [3 kMaxPts] | out (float or double) [4 kMaxPts]
static constexpr size_t kRstOff = 4 * kMaxPts;
static constexpr size_t kOutOff = 16 * kMaxPts;
static constexpr size_t kScratchBytes = 48 * kMaxPts;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each GPU (looked up once per device, with
// that device current).  In host code the array's name is only a stand-in:
// kernels and copies take these addresses.
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[FEMRT_NOALLOC_MAX_DEVICES];

static int scratch_address(femrt_ctx* c, unsigned char** out)
{
  if (c->device < 0 || c->device >= FEMRT_NOALLOC_MAX_DEVICES)
    return fail(c, FEMRT_EUNSUP, "eval_field_noalloc: device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[c->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(c, FEMRT_EHIP, std::string("eval_field_noalloc: hipGetSymbolAddress(scratch): ") +
                                     hipGetErrorString(e));
    g_scratch_ptr[c->device] = p;
  }
  *out = (unsigned char*)g_scratch_ptr[c->device];
  return FEMRT_OK;
}

template <int NG, int NF>
static int launch_eval(hipStream_t st, int blocks, int n, const int* d_e, const float* d_r,
                       const femrt_mesh* m, int precision, void* d_o)
{
  if (precision)
    k_eval_field<double, NG, NF><<<blocks, 256, 0, st>>>(n, d_e, d_r, m->d_geom, m->d_field, (double*)d_o);
  else
    k_eval_field<float, NG, NF><<<blocks, 256, 0, st>>>(n, d_e, d_r, m->d_geom, m->d_field, (float*)d_o);
  return FEMRT_OK;
}

extern "C" {

FEMRT_API int femrt_eval_field_noalloc(femrt_mesh* m, int n, const int* elem, const float* rst,
                                       int precision, void* out)
{
  if (!m || n < 0 || (n > 0 && (!elem || !rst || !out))) return FEMRT_EINVAL;
  femrt_ctx* c = m->ctx;
  if ((size_t)n > kMaxPts)
    return fail(c, FEMRT_EINVAL, "eval_field_noalloc: more points than the 65536 its scratch holds");
  if (n == 0) return FEMRT_OK;
  for (int i = 0; i < n; ++i)
    if (elem[i] < 0 || elem[i] >= m->n_elem)
      return fail(c, FEMRT_EINVAL, "eval_field_noalloc: element index out of range");
  HIPCHK(c, hipSetDevice(c->device));
  unsigned char* base = nullptr;
  if (int rc = scratch_address(c, &base)) return rc;
  int* d_e = (int*)base;
  float* d_r = (float*)(base + kRstOff);
  void* d_o = base + kOutOff;
  const size_t osz = precision ? sizeof(double) : sizeof(float);
  HIPCHK(c, hipMemcpyAsync(d_e, elem, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_r, rst, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
  const int blocks = (n + 255) / 256;
  int rc = FEMRT_EUNSUP;
  // (parenthesised: the dispatch macros take the launch as one argument)
  FEMRT_DISPATCH_NG(m->NG, FEMRT_DISPATCH_NF(m->NF, {
    rc = (launch_eval<NG_, NF_>(c->stream, blocks, n, d_e, d_r, m, precision, d_o));
  }));
  if (rc) return fail(c, rc, "eval_field_noalloc: no kernel for these mode counts");
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(out, d_o, (size_t)n * 4 * osz, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FEMRT_OK;
}

}  // extern "C"
