// femrt_runtime.hpp -- host objects behind the C ABI handles and the launch
// helpers shared by the library's HIP sources.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "femrt_internal.hpp"
#include "femrt_trace_core.hpp"

using namespace femrt;

// ============================================================== host objects

struct TimedLaunch {
  std::string name;
  hipEvent_t a, b;
};

struct femrt_ctx {
  int device = 0;
  int cus = 256;               // compute units of the device
  hipStream_t stream = nullptr;
  std::string err;
  bool timing = false;
  bool capturing = false;      // a batch is recording the stream into a graph
  std::vector<TimedLaunch> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<double, long long>> stats;
};

// Guess network: FP32 master parameters, Adam moments and gradient
// accumulators, and the packed BF16 operands the kernels read.
struct femrt_mlp {
  femrt_ctx* ctx = nullptr;
  int kind = 0;
  float* d_param = nullptr;   // MLP_NPARAM
  float* d_m = nullptr;       // Adam first moment
  float* d_v = nullptr;       // Adam second moment
  float* d_grad = nullptr;    // MLP_NPARAM gradient + [MLP_NPARAM] loss (femrt_mlp_loss_grad)
  float* d_part = nullptr;    // per training block: its partial gradient and loss
  int part_blocks = 0;
  void* d_pack = nullptr;     // PK_BYTES
  long long adam_t = 0;
};

// Launch shapes of the trace kernels (chosen per mesh by the autotuner).
struct LaunchCfg {
  int hit_block = 256;     // k_first_hit / k_raygen threads per block
  int march_bpc = 4;       // persistent walk: blocks of 256 per compute unit
  int polish_bpc = 8;      // polish kernels: blocks of 256 per compute unit
  int tile = 8;            // first-hit ray order: tiles of tile x tile pixels (4, 8, 16)
  int qseg_per_xcd = 1;    // walk hit queue: segments per XCD (1, 2, 4)
};

struct femrt_mesh {
  femrt_ctx* ctx = nullptr;
  LaunchCfg cfg;
  const femrt_mlp* guess_mlp = nullptr;  // network used by FEMRT_GUESS_LEARNED
  const femrt_mlp* bracket_mlp = nullptr;  // network used by FEMRT_TRACE_ISO_LEARNED_BRACKET
  bool surface = false;                // curved quads (femrt_mesh_upload_quads)
  bool prism = false;                  // prisms (femrt_mesh_upload_prisms): first hits only
  int n_elem = 0, ng = 0, nf = 0;      // true mode counts
  int NG = 0, NF = 0;                  // kernel mode counts (padded)
  GeoMode* d_geom = nullptr;           // n_elem x NG^3 (x, y, z)
  float* d_field = nullptr;            // n_elem x NF^3
  float* d_verts = nullptr;            // n_elem x 24
  float* d_boxes = nullptr;            // n_elem x FEMRT_BOX_STRIDE
  int* d_adj = nullptr;                // n_elem x 12
  float* d_lip = nullptr;              // n_elem: Lipschitz bound of the field in rst
  float* d_jinv = nullptr;             // n_elem: bound of |J^-1|_inf over the element (0: none)
  float* d_frange = nullptr;           // n_elem x 2: bounds of the field
  float4* d_nodes = nullptr;           // n_nodes x 2
  int* d_prims = nullptr;
  int n_nodes = 0;
  int bvh_depth = 0;
  std::vector<float> h_boxes;
  BvhHost bvh;
};

struct femrt_frame {
  femrt_ctx* ctx = nullptr;
  int n = 0;
  // image layout of the rays when a camera made them (ray (y*img_w + x)*img_spp
  // + s); 0 when the caller set them.  Traces then visit the rays in 8x8-pixel
  // tiles so that a wave's lanes walk neighbouring elements.
  int img_w = 0, img_h = 0, img_spp = 1;
  float* d_rays = nullptr;
  int* d_hit_elem = nullptr;
  int* d_hit_face = nullptr;
  float4* d_hit_f32 = nullptr;  // u, v, tau, unused
  double* d_hit_t = nullptr;
  double* d_hit_rst = nullptr;  // 3 x n
  double* d_hit_u = nullptr;
  int* d_iso_elem = nullptr;    // MAX_ISO x n
  double* d_iso_t = nullptr;
  double* d_iso_rst = nullptr;  // MAX_ISO x 3 x n
  double* d_iso_u = nullptr;
  float4* d_iso_f32 = nullptr;  // MAX_ISO x n: (t_a, t_b, t*, packed signs)
  float4* d_iso_rst32 = nullptr;  // MAX_ISO x n: rst* (FP32)
  int* d_queue_hit = nullptr;     // n
  int* d_retrace = nullptr;       // n: hit rays traced again (no FP64 root on the FP32 hit's face)
  int* d_queue_walk = nullptr;    // n: the hit queue in walk order (longest chords first per segment)
  unsigned char* d_qbin = nullptr;  // n: chord bin per hit-queue position
  int* d_counts = nullptr;        // FEMRT_NCOUNTS: [0] hits, [2..3] walk
                                  // rounds, [FEMRT_QSEG_CTR..] walk queue segments
  uint32_t* d_rgba = nullptr;
  long long* d_work = nullptr;
  bool counters = false;
  float4* d_path_tp = nullptr;   // per path: throughput rgb, alive flag (allocated by path_begin)
  float4* d_path_acc = nullptr;  // per path: accumulated radiance rgb
  void* d_mats = nullptr;        // scatter materials (64 slots), uploaded when they change
  _Float16* d_mlp_g = nullptr;   // every-sample network guesses, per walk lane
  size_t mlp_g_cap = 0;          // halfs
  // per (isovalue, ray) bracket: its ends as the bracket network reads them
  // (rst at ta and tb, FP32 g at ta and tb: 2 float4) and the network's
  // prediction (t_lo, t_hi); allocated by the first trace that records them
  float4* d_iso_seg = nullptr;   // MAX_ISO x n x 2
  float2* d_iso_pred = nullptr;  // MAX_ISO x n
  bool seg_valid = false;        // the last trace recorded segments
  std::vector<unsigned char> h_mats;
  // element-walk rounds (allocated by the first trace with isovalues)
  int4* d_walk_si = nullptr;     // per ray: element, entry face, step, found | redo << 8
  float4* d_walk_sf = nullptr;   // per ray: t_in, u_in, v_in
  int* d_walk_a = nullptr;       // round lists
  int* d_walk_b = nullptr;
  int* d_walk_sorted = nullptr;
  int* d_walk_reent = nullptr;
  int* d_walk_hist = nullptr;    // n_elem + 1
  int walk_cap = 0, walk_hist_cap = 0;
};

static int fail(femrt_ctx* ctx, int code, const std::string& msg)
{
  if (ctx) ctx->err = msg;
  return code;
}

#define HIPCHK(ctx, call)                                                                   \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail((ctx), FEMRT_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));    \
  } while (0)

static hipEvent_t take_event(femrt_ctx* c)
{
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

static void collect_timing(femrt_ctx* c)
{
  for (auto& t : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
      auto& s = c->stats[t.name];
      s.first += ms;
      s.second += 1;
    }
    c->pool.push_back(t.a);
    c->pool.push_back(t.b);
  }
  c->pending.clear();
}

// Launch wrapper: records HIP events around the launch on the context stream
// when timing is enabled.
template <typename F>
static int timed_launch(femrt_ctx* c, const char* name, F&& launch)
{
  hipEvent_t a = nullptr, b = nullptr;
  if (c->timing && !c->capturing) {
    a = take_event(c);
    b = take_event(c);
    if (a) (void)hipEventRecord(a, c->stream);
  }
  launch();
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, FEMRT_EHIP, std::string("launch ") + name + ": " + hipGetErrorString(e));
  if (c->timing && !c->capturing && a && b) {
    (void)hipEventRecord(b, c->stream);
    c->pending.push_back({name, a, b});
  }
  return FEMRT_OK;
}

// ============================================================ device helpers

__device__ __forceinline__ void add_work(long long* work, int idx, long long v)
{
  if (v) atomicAdd((unsigned long long*)&work[idx], (unsigned long long)v);
}

// first_hit: the first-hit pass keeps its face-Newton iterations and slab
// tests in counters of their own ([30], [31]), so the walk's are its own.
__device__ __forceinline__ void flush_work(long long* work, const Work& w, bool first_hit = false)
{
  if (!work) return;
  add_work(work, first_hit ? 30 : 2, w.face_it);
  add_work(work, 3, w.inv_it);
  add_work(work, 4, w.field);
  add_work(work, 5, w.geo64);
  add_work(work, 6, w.field64);
  add_work(work, first_hit ? 31 : 7, w.slab);
  add_work(work, 8, w.mlp);
  add_work(work, 9, w.marched);
  add_work(work, 10, w.exit_pred);
  add_work(work, 11, w.exit_search);
  add_work(work, 12, w.exit_march);
  add_work(work, 13, w.reentry);
  add_work(work, 14, w.redo);
  add_work(work, 15, w.brackets);
  add_work(work, 27, w.skipped);
  add_work(work, 28, w.samples_skipped);
  add_work(work, 29, w.stalled);
  add_work(work, 32, w.skips);
  add_work(work, 33, w.jumps);
}

// Append `pred` lanes to a queue: one atomic per wave, ballot + prefix count
// for the slot of each lane.  Must be reached by every lane of the wave.
__device__ __forceinline__ int wave_append(bool pred, int* counter)
{
  const unsigned long long mask = __ballot(pred);
  const int lane = __lane_id();
  const int leader = __ffsll((long long)mask) - 1;
  int base = 0;
  if (mask) {
    if (lane == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader);
  }
  return base + __popcll(mask & ((1ull << lane) - 1ull));
}

// Frame queue counters: [0] hits, [2..3] binned-walk lists, [4] rays to
// trace again (FEMRT_RETRACE_CTR), [5] binned-walk start list,
// [FEMRT_QSEG_CTR + s] progress in segment s of the hit queue (persistent walk).
// The walk's hit queue has FEMRT_XCDS x (1, 2 or 4) segments (a launch
// shape): each XCD's band of the queue is cut into that many sub-bands.
#define FEMRT_RETRACE_CTR 4
#define FEMRT_XCDS 8
#define FEMRT_MAX_QSEG 32
#define FEMRT_QSEG_CTR 16
#define FEMRT_NCOUNTS (FEMRT_QSEG_CTR + FEMRT_MAX_QSEG)

// Grid-stride loop over [0, count) split by XCD.  The workgroups of a
// launch are dispatched round-robin over the FEMRT_XCDS (8) XCDs, so block b
// runs on XCD b % 8; the blocks of XCD x stride over the x-th of 8
// contiguous slices of the range.  Over a queue in tile order each slice is
// one band of the image, so each XCD's L2 caches the elements of its own
// band instead of every XCD fetching every element.  A grid of fewer than 8
// blocks strides over the whole range.
struct XcdStride {
  int q, end, step;
  __device__ XcdStride(int count)
  {
    const int b = blockIdx.x, nb = gridDim.x, t = threadIdx.x, bs = blockDim.x;
    if (nb < FEMRT_XCDS) {
      q = b * bs + t;
      end = count;
      step = nb * bs;
      return;
    }
    const int x = b % FEMRT_XCDS, rank = b / FEMRT_XCDS;
    const int nx = (nb - x + FEMRT_XCDS - 1) / FEMRT_XCDS;  // blocks on XCD x
    q = (int)((long long)count * x / FEMRT_XCDS) + rank * bs + t;
    end = (int)((long long)count * (x + 1) / FEMRT_XCDS);
    step = nx * bs;
  }
};

#define FEMRT_DISPATCH_NG(NGV, BODY)      \
  switch (NGV) {                          \
    case 4: { constexpr int NG_ = 4; BODY; } break; \
    case 5: { constexpr int NG_ = 5; BODY; } break; \
    case 7: { constexpr int NG_ = 7; BODY; } break; \
    default: break;                       \
  }
#define FEMRT_DISPATCH_NF(NFV, BODY)      \
  switch (NFV) {                          \
    case 3: { constexpr int NF_ = 3; BODY; } break; \
    case 5: { constexpr int NF_ = 5; BODY; } break; \
    case 7: { constexpr int NF_ = 7; BODY; } break; \
    case 9: { constexpr int NF_ = 9; BODY; } break; \
    default: break;                       \
  }

#define FEMRT_MLP_MAX_SAMPLES 16
// Work buffer of a frame: FEMRT_NCOUNTERS counters, then the number of walks
// that ended stalled and the first FEMRT_MAX_STALLED_IDS of their ray ids.
#define FEMRT_WORK_SLOTS (FEMRT_NCOUNTERS + 1 + FEMRT_MAX_STALLED_IDS)

// Bracket network inference before the FP64 endgame (femrt_bracket.hip):
// d_iso_pred of every (hit ray, isovalue) bracket of the frame.
int bracket_infer_launch(femrt_ctx* c, femrt_mesh* m, femrt_frame* f, const TraceMesh& TM,
                         const IsoParams& ip, int blocks);

// BVH builder 1: linear BVH on the device (femrt_lbvh.hip).  Sets
// m->d_nodes, m->d_prims, m->n_nodes and the host copy m->bvh.
int lbvh_build(femrt_ctx* c, femrt_mesh* m);

// Element walk in element-binned rounds (femrt_walk.hip).
int walk_run(femrt_ctx* c, femrt_mesh* m, femrt_frame* f, const TraceMesh& TM, const IsoParams& ip,
             bool learned, long long* work);
