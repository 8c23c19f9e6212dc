// femrt_batch.hip -- many traces and path-traced renders of meshes of
// uneven size recorded into one HIP graph and replayed with one launch, and
// the autotuner that picks each mesh's kernel launch shapes.
//
// A batch records the context stream (stream capture) while it enqueues each
// job exactly as the eager calls would, so the graph holds the same kernels
// with the same arguments: replaying it reproduces the eager results.  Jobs
// must be capture-safe: no work counters, no binned walk (its rounds need the
// host), materials and path buffers set up beforehand (femrt_batch_add_paths
// does that).
//
// The jobs are recorded on FEMRT_BATCH_BRANCHES streams forked from the
// context stream inside the capture, so the graph holds that many
// independent branches and one job's short or latency-bound kernels (the
// near-miss retrace, the per-depth path kernels) run beside another job's
// large ones.  Jobs on the same frame stay on one branch, in order; meshes
// are only read by a trace, so jobs on one mesh may run side by side.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <string>
#include <vector>

#include "femrt_runtime.hpp"

using namespace femrt;

struct BatchJob {
  int kind = 0;  // 0 trace, 1 path-traced render
  femrt_mesh* mesh = nullptr;
  femrt_frame* frame = nullptr;
  femrt_trace_params tp;
  float cam12[12];
  int W = 0, H = 0, spp = 1, depth = 1;
  std::vector<femrt_material> mats;
  unsigned long long seed = 0;
};

#define FEMRT_BATCH_BRANCHES 4

struct femrt_batch {
  femrt_ctx* ctx = nullptr;
  std::vector<BatchJob> jobs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t branch[FEMRT_BATCH_BRANCHES - 1] = {};  // forked streams (default priority)
};

static void batch_drop_graph(femrt_batch* b)
{
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

static int enqueue_job(const BatchJob& j)
{
  if (j.kind == 0) return femrt_trace(j.mesh, j.frame, &j.tp);
  int rc = femrt_frame_camera_rays_jitter(j.frame, j.cam12, j.W, j.H, j.spp, j.seed);
  if (rc == FEMRT_OK) rc = femrt_frame_path_begin(j.frame);
  femrt_trace_params tp;
  femrt_trace_params_default(&tp);
  tp.iso_enable = 0;
  tp.n_iso = 0;
  tp.flags = 0;
  tp.guess = FEMRT_GUESS_WARM;
  for (int d = 0; d < j.depth && rc == FEMRT_OK; ++d) {
    rc = femrt_trace(j.mesh, j.frame, &tp);
    if (rc == FEMRT_OK)
      rc = femrt_frame_scatter(j.mesh, j.frame, j.mats.data(), (int)j.mats.size(), j.seed, d, d == j.depth - 1);
  }
  if (rc == FEMRT_OK) rc = femrt_frame_resolve(j.frame, j.W, j.H, j.spp, nullptr);
  return rc;
}

extern "C" {

FEMRT_API int femrt_batch_create(femrt_ctx* c, femrt_batch** out)
{
  if (!c || !out) return FEMRT_EINVAL;
  femrt_batch* b = new femrt_batch;
  b->ctx = c;
  *out = b;
  return FEMRT_OK;
}

FEMRT_API int femrt_batch_destroy(femrt_batch* b)
{
  if (!b) return FEMRT_EINVAL;
  (void)hipStreamSynchronize(b->ctx->stream);
  batch_drop_graph(b);
  for (hipStream_t& s : b->branch)
    if (s) (void)hipStreamDestroy(s);
  delete b;
  return FEMRT_OK;
}

FEMRT_API int femrt_batch_add_trace(femrt_batch* b, femrt_mesh* m, femrt_frame* f, const femrt_trace_params* p)
{
  if (!b || !m || !f || !p) return FEMRT_EINVAL;
  femrt_ctx* c = b->ctx;
  if (m->ctx != c || f->ctx != c) return fail(c, FEMRT_EINVAL, "batch_add_trace: objects of another context");
  if (p->flags & FEMRT_TRACE_BINNED_WALK)
    return fail(c, FEMRT_EUNSUP, "batch_add_trace: the binned walk needs the host between rounds");
  if (f->counters) return fail(c, FEMRT_EUNSUP, "batch_add_trace: switch the frame's work counters off");
  BatchJob j;
  j.kind = 0;
  j.mesh = m;
  j.frame = f;
  j.tp = *p;
  b->jobs.push_back(j);
  batch_drop_graph(b);
  return FEMRT_OK;
}

FEMRT_API int femrt_batch_add_paths(femrt_batch* b, femrt_mesh* m, femrt_frame* f, const float* cam12, int W,
                                    int H, int spp, int max_depth, const femrt_material* mats, int n_mats,
                                    unsigned long long seed)
{
  if (!b || !m || !f || !cam12 || !mats || n_mats < 1 || n_mats > 64 || max_depth < 1 || W <= 0 || H <= 0 ||
      spp <= 0)
    return FEMRT_EINVAL;
  femrt_ctx* c = b->ctx;
  if (m->ctx != c || f->ctx != c) return fail(c, FEMRT_EINVAL, "batch_add_paths: objects of another context");
  if ((long long)W * H * spp != f->n) return fail(c, FEMRT_EINVAL, "batch_add_paths: frame must hold W*H*spp rays");
  if (f->counters) return fail(c, FEMRT_EUNSUP, "batch_add_paths: switch the frame's work counters off");
  BatchJob j;
  j.kind = 1;
  j.mesh = m;
  j.frame = f;
  std::memcpy(j.cam12, cam12, sizeof(j.cam12));
  j.W = W;
  j.H = H;
  j.spp = spp;
  j.depth = max_depth;
  j.mats.assign(mats, mats + n_mats);
  j.seed = seed;
  // path buffers and the material table exist before any capture
  int rc = femrt_frame_path_begin(f);
  if (rc == FEMRT_OK) rc = femrt_frame_set_materials(f, mats, n_mats);
  if (rc) return rc;
  b->jobs.push_back(j);
  batch_drop_graph(b);
  return FEMRT_OK;
}

FEMRT_API int femrt_batch_capture(femrt_batch* b)
{
  if (!b) return FEMRT_EINVAL;
  femrt_ctx* c = b->ctx;
  if (b->jobs.empty()) return fail(c, FEMRT_ESTATE, "batch_capture: no jobs");
  batch_drop_graph(b);
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // branch of each job: frames in order of first use, each new frame to the
  // branch with the least estimated work so far (rays x log2 elements)
  const int nb = std::min<int>(FEMRT_BATCH_BRANCHES, (int)b->jobs.size());
  std::vector<int> jb(b->jobs.size(), 0);
  {
    std::vector<double> load(nb, 0.0);
    std::vector<std::pair<const femrt_frame*, int>> seen;
    for (size_t i = 0; i < b->jobs.size(); ++i) {
      const BatchJob& j = b->jobs[i];
      int k = -1;
      for (const auto& fs : seen)
        if (fs.first == j.frame) k = fs.second;
      if (k < 0) {
        k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        seen.push_back({j.frame, k});
      }
      jb[i] = k;
      load[k] += (double)j.frame->n * (1.0 + std::log2(1.0 + j.mesh->n_elem)) * (j.kind == 1 ? j.depth : 1);
    }
  }
  for (int k = 0; k + 1 < nb; ++k)
    if (!b->branch[k]) HIPCHK(c, hipStreamCreateWithFlags(&b->branch[k], hipStreamNonBlocking));
  hipEvent_t ev[FEMRT_BATCH_BRANCHES] = {};
  for (int k = 0; k < nb; ++k) HIPCHK(c, hipEventCreateWithFlags(&ev[k], hipEventDisableTiming));
  auto drop_events = [&] {
    for (hipEvent_t& x : ev)
      if (x) (void)hipEventDestroy(x);
  };
  const hipStream_t main = c->stream;
  if (hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    drop_events();
    return fail(c, FEMRT_EHIP, "batch_capture: begin capture failed");
  }
  c->capturing = true;
  // fork: the branch streams join the capture by waiting on the main stream
  hipError_t fe = hipEventRecord(ev[0], main);
  for (int k = 0; k + 1 < nb && fe == hipSuccess; ++k) fe = hipStreamWaitEvent(b->branch[k], ev[0], 0);
  int rc = fe == hipSuccess ? FEMRT_OK : fail(c, FEMRT_EHIP, "batch_capture: fork failed");
  for (size_t i = 0; i < b->jobs.size() && rc == FEMRT_OK; ++i) {
    c->stream = jb[i] == 0 ? main : b->branch[jb[i] - 1];
    rc = enqueue_job(b->jobs[i]);
  }
  c->stream = main;
  // join: the main stream waits for every branch before the capture ends
  for (int k = 0; k + 1 < nb; ++k) {
    if (hipEventRecord(ev[k + 1], b->branch[k]) != hipSuccess || hipStreamWaitEvent(main, ev[k + 1], 0) != hipSuccess)
      if (rc == FEMRT_OK) rc = fail(c, FEMRT_EHIP, "batch_capture: join failed");
  }
  c->capturing = false;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(main, &g);
  drop_events();
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess || !g) return fail(c, FEMRT_EHIP, std::string("batch_capture: ") + hipGetErrorString(e));
  b->graph = g;
  if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
    batch_drop_graph(b);
    return fail(c, FEMRT_EHIP, "batch_capture: graph instantiation failed");
  }
  return FEMRT_OK;
}

FEMRT_API int femrt_batch_launch(femrt_batch* b)
{
  if (!b) return FEMRT_EINVAL;
  femrt_ctx* c = b->ctx;
  if (!b->exec) return fail(c, FEMRT_ESTATE, "batch_launch: capture the batch first");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipGraphLaunch(b->exec, c->stream));
  return FEMRT_OK;
}

// ------------------------------------------------------------- autotuner

// The launch shape as FEMRT_LAUNCH_CFG_N ints, in the order of femrt.h.
static int* cfg_slot(LaunchCfg& c, int i)
{
  int* slots[FEMRT_LAUNCH_CFG_N] = {&c.hit_block, &c.march_bpc, &c.polish_bpc, &c.tile, &c.qseg_per_xcd};
  return slots[i];
}

static bool cfg_valid(const LaunchCfg& c)
{
  auto one_of = [](int v, std::initializer_list<int> ok) {
    for (int x : ok)
      if (v == x) return true;
    return false;
  };
  return one_of(c.hit_block, {64, 128, 256}) && c.march_bpc >= 1 && c.march_bpc <= 64 && c.polish_bpc >= 1 &&
         c.polish_bpc <= 64 && one_of(c.tile, {4, 8, 16}) && one_of(c.qseg_per_xcd, {1, 2, 4});
}

FEMRT_API int femrt_mesh_get_launch_cfg(const femrt_mesh* m, int* cfg, int n_cfg)
{
  if (!m || !cfg || n_cfg < 1 || n_cfg > FEMRT_LAUNCH_CFG_N) return FEMRT_EINVAL;
  LaunchCfg c = m->cfg;
  for (int i = 0; i < n_cfg; ++i) cfg[i] = *cfg_slot(c, i);
  return FEMRT_OK;
}

FEMRT_API int femrt_mesh_set_launch_cfg(femrt_mesh* m, const int* cfg, int n_cfg)
{
  if (!m || !cfg) return FEMRT_EINVAL;
  if (n_cfg < 1 || n_cfg > FEMRT_LAUNCH_CFG_N)
    return fail(m->ctx, FEMRT_EINVAL, "set_launch_cfg: 1..5 values");
  LaunchCfg c = m->cfg;
  for (int i = 0; i < n_cfg; ++i) *cfg_slot(c, i) = cfg[i];
  if (!cfg_valid(c))
    return fail(m->ctx, FEMRT_EINVAL,
                "set_launch_cfg: block 64/128/256, blocks per CU 1..64, tile 4/8/16, queue segments per XCD 1/2/4");
  m->cfg = c;
  return FEMRT_OK;
}

// Times each trace kernel under candidate launch shapes on the given job
// (one dimension at a time, the others held) and keeps the fastest on the
// mesh.
FEMRT_API int femrt_autotune(femrt_mesh* m, femrt_frame* f, const femrt_trace_params* p, int reps, int* cfg,
                             int n_cfg)
{
  if (!m || !f || !p || reps < 1 || (cfg && (n_cfg < 1 || n_cfg > FEMRT_LAUNCH_CFG_N))) return FEMRT_EINVAL;
  femrt_ctx* c = m->ctx;
  if (c->capturing) return fail(c, FEMRT_ESTATE, "autotune: not while a batch records");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  const bool was_timing = c->timing;
  const bool was_counting = f->counters;
  const auto saved_stats = c->stats;
  f->counters = false;
  // total event time of the given kernels over reps traces, after one
  // untimed trace (a kernel that did not run counts 0; none ran: 1e30)
  auto measure = [&](std::initializer_list<const char*> kernels, double& ms) -> int {
    c->timing = false;
    int rc = femrt_trace(m, f, p);
    if (rc) return rc;
    c->timing = true;
    c->stats.clear();
    for (int r = 0; r < reps && rc == FEMRT_OK; ++r) rc = femrt_trace(m, f, p);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    collect_timing(c);
    if (rc) return rc;
    bool any = false;
    ms = 0.0;
    for (const char* k : kernels)
      if (c->stats.count(k)) {
        ms += c->stats[k].first;
        any = true;
      }
    if (!any) ms = 1e30;
    return FEMRT_OK;
  };
  // one dimension at a time, the current shape measured first; another
  // shape replaces it only when at least 3% faster (timing noise)
  LaunchCfg best = m->cfg;
  int rc = FEMRT_OK;
  auto tune = [&](int LaunchCfg::*field, std::initializer_list<int> cands,
                  std::initializer_list<const char*> kernels) {
    double bt = 0.0;
    std::vector<int> order{best.*field};
    for (int v : cands)
      if (v != best.*field) order.push_back(v);
    for (size_t i = 0; i < order.size() && rc == FEMRT_OK; ++i) {
      m->cfg = best;
      m->cfg.*field = order[i];
      double ms = 0;
      if ((rc = measure(kernels, ms))) break;
      if (i == 0) {
        bt = ms;
      } else if (ms < 0.97 * bt) {
        bt = ms;
        best.*field = order[i];
      }
    }
  };
  const bool iso = p->iso_enable && p->n_iso > 0;
  const bool image = f->img_w > 0 && (long long)f->img_w * f->img_h * f->img_spp == f->n;
  tune(&LaunchCfg::hit_block, {64, 128, 256}, {"k_first_hit"});
  // the tile decides the order in which the first hit, and through the hit
  // queue the walk, visit the rays (rays without an image layout: no tiles)
  if (rc == FEMRT_OK && image)
    tune(&LaunchCfg::tile, {4, 8, 16}, {"k_first_hit", "k_walk_order", "k_march", "k_polish_iso"});
  if (rc == FEMRT_OK && iso) tune(&LaunchCfg::march_bpc, {2, 4, 8, 16}, {"k_march"});
  if (rc == FEMRT_OK && iso) tune(&LaunchCfg::qseg_per_xcd, {1, 2, 4}, {"k_walk_order", "k_march"});
  if (rc == FEMRT_OK)
    tune(&LaunchCfg::polish_bpc, {4, 8, 16}, iso ? std::initializer_list<const char*>{"k_polish_hit", "k_polish_iso"}
                                                : std::initializer_list<const char*>{"k_polish_hit"});
  m->cfg = best;
  c->stats = saved_stats;
  c->timing = was_timing;
  f->counters = was_counting;
  if (rc) return rc;
  if (cfg)
    for (int i = 0; i < n_cfg; ++i) cfg[i] = *cfg_slot(best, i);
  return FEMRT_OK;
}

}  // extern "C"
