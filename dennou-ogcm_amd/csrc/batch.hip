// batch.hip -- many independent ocean-planet instances of uneven truncation and layer
// count advanced together.  The reference runs one model per process and parallelises
// only over layers (`!$omp parallel do` around SpmlUtil calls, SpmlUtil_mod.f90:360,389);
// here instances that share a truncation are packed side by side as columns of one set
// of device buffers, so one contraction launch covers all of them; the truncation groups
// run concurrently on streams of their own (forked from and joined back to the caller's
// stream), and the whole step is captured once into a hipGraph whose groups are parallel
// branches.
//
// Per instance (KM layers) and step, the dynamics-step transform set:
//   4 prognostic fields (vor, div, T, S) spectral -> grid with lon/lat derivatives
//   (xy_w, xy_GradLon_w, xy_GradLat_w; one fused dsht_inv_vgrad_d call) and 6 tendency
//   fields grid -> spectral (w_xy).
// Layout: an instance's columns are contiguous, fields of one layer count apart: prognostic
// field f, layer k of the instance with first layer column col0 is column NPROG col0 + f KM +
// k of its group's spectral and grid buffers; tendency field f is column NTEND col0 + f KM +
// k.  So an instance alone is one inverse call over NPROG KM columns and one forward call
// over NTEND KM columns, and a group is the same two calls over all its instances.
#include <algorithm>
#include <cstring>
#include <vector>

#include "dsht_internal.hpp"

namespace dsht {

constexpr int NPROG = 4, NTEND = 6, NINV = 3;  // NINV grids per prognostic column: value, GradLon, GradLat

struct BatchInstance {
  dsht_plan* plan;
  int km;
  int group;
  int col0;  // first layer column of this instance inside its group (in layers)
};

struct BatchGroup {
  dsht_plan* plan;
  int cols = 0;  // total layers of the group's instances
  double* spec = nullptr;     // (L, NPROG * cols)  LC layout
  double* grid = nullptr;     // NINV x (NPROG * cols, J, I)
  double* tend = nullptr;     // (NTEND * cols, J, I)
  double* spect = nullptr;    // (L, NTEND * cols)  LC layout
  hipStream_t stream = nullptr;  // the group's branch (default priority)
  hipEvent_t done = nullptr;     // recorded on the branch at the end of its step
};

}  // namespace dsht

struct dsht_batch {
  std::vector<dsht::BatchInstance> inst;
  std::vector<dsht::BatchGroup> groups;
  bool finalized = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t fork = nullptr;
  size_t bytes = 0;
};

using namespace dsht;

static int batch_group_step(BatchGroup& g, hipStream_t s) {
  dsht_plan* p = g.plan;
  const long IJ = (long)p->I * p->J;
  const int C = NPROG * g.cols;
  // value and both gradients of every prognostic column in one fused inverse
  double* gk[NINV];
  for (int k = 0; k < NINV; k++) gk[k] = g.grid + (size_t)k * C * IJ;
  if (int rc = dsht_inv_vgrad_d(p, g.spec, C, 1, gk[0], gk[1], gk[2], IJ, C, nullptr, s)) return rc;
  return dsht_fwd_d(p, DSHT_VAL, g.tend, IJ, g.spect, NTEND * g.cols, 1, NTEND * g.cols, nullptr, s);
}

extern "C" {

int dsht_batch_create(dsht_batch** b) {
  if (!b) return fail(DSHT_EARG, "dsht_batch_create: NULL");
  *b = new dsht_batch();
  return 0;
}

int dsht_batch_add(dsht_batch* b, dsht_plan* plan, int km, int* id) {
  if (!b || !plan || km <= 0) return fail(DSHT_EARG, "dsht_batch_add: bad arguments");
  if (b->finalized) return fail(DSHT_ESTATE, "dsht_batch_add: batch already finalized");
  BatchInstance in{plan, km, -1, 0};
  b->inst.push_back(in);
  if (id) *id = (int)b->inst.size() - 1;
  return 0;
}

int dsht_batch_finalize(dsht_batch* b) {
  if (!b) return fail(DSHT_EARG, "NULL batch");
  if (b->finalized) return 0;
  for (auto& in : b->inst) {
    int gi = -1;
    for (size_t k = 0; k < b->groups.size(); k++)
      if (b->groups[k].plan == in.plan) gi = (int)k;
    if (gi < 0) {
      BatchGroup g;
      g.plan = in.plan;
      b->groups.push_back(g);
      gi = (int)b->groups.size() - 1;
    }
    in.group = gi;
    in.col0 = b->groups[gi].cols;
    b->groups[gi].cols += in.km;
  }
  for (auto& g : b->groups) {
    const dsht_plan* p = g.plan;
    const size_t IJ = (size_t)p->I * p->J, L = p->L;
    const size_t n_spec = L * NPROG * g.cols, n_grid = NINV * IJ * NPROG * g.cols;
    const size_t n_tend = IJ * NTEND * g.cols, n_spect = L * NTEND * g.cols;
    DSHT_HIP(hipSetDevice(p->device));
    if (hipMalloc(&g.spec, 8 * n_spec) != hipSuccess || hipMalloc(&g.grid, 8 * n_grid) != hipSuccess ||
        hipMalloc(&g.tend, 8 * n_tend) != hipSuccess || hipMalloc(&g.spect, 8 * n_spect) != hipSuccess) {
      (void)hipGetLastError();
      return fail(DSHT_ENOMEM, "dsht_batch_finalize: device allocation failed");
    }
    DSHT_HIP(hipMemset(g.spec, 0, 8 * n_spec));
    DSHT_HIP(hipMemset(g.tend, 0, 8 * n_tend));
    b->bytes += 8 * (n_spec + n_grid + n_tend + n_spect);
    if (b->groups.size() > 1) {
      DSHT_HIP(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
      DSHT_HIP(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
    }
  }
  if (b->groups.size() > 1) DSHT_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  b->finalized = true;
  return 0;
}

int dsht_batch_info(const dsht_batch* b, int id, int* group, int* col0, int* group_cols, long* bytes) {
  if (!b || !b->finalized) return fail(DSHT_ESTATE, "batch not finalized");
  if (id >= 0) {
    if (id >= (int)b->inst.size()) return fail(DSHT_EARG, "bad instance id %d", id);
    if (group) *group = b->inst[id].group;
    if (col0) *col0 = b->inst[id].col0;
    if (group_cols) *group_cols = b->groups[b->inst[id].group].cols;
  }
  if (bytes) *bytes = (long)b->bytes;
  return 0;
}

/* device buffers of one truncation group: spec (L, 4*cols) LC, grid 3 x (4*cols, J, I),
   tend (6*cols, J, I), spect (L, 6*cols) LC; instance-contiguous columns (prognostic field f,
   layer k of an instance at 4*col0 + f*km + k, tendency field f at 6*col0 + f*km + k) */
int dsht_batch_buffers(const dsht_batch* b, int group, double** spec, double** grid, double** tend,
                       double** spect) {
  if (!b || !b->finalized) return fail(DSHT_ESTATE, "batch not finalized");
  if (group < 0 || group >= (int)b->groups.size()) return fail(DSHT_EARG, "bad group %d", group);
  const BatchGroup& g = b->groups[group];
  if (spec) *spec = g.spec;
  if (grid) *grid = g.grid;
  if (tend) *tend = g.tend;
  if (spect) *spect = g.spect;
  return 0;
}

int dsht_batch_ngroups(const dsht_batch* b) { return b ? (int)b->groups.size() : 0; }

/* one step, every group: with several groups each runs on its own stream, forked from and
   joined back to the caller's stream (under capture: parallel branches of the graph) */
int dsht_batch_step(dsht_batch* b, void* stream) {
  if (!b || !b->finalized) return fail(DSHT_ESTATE, "batch not finalized");
  hipStream_t s = (hipStream_t)stream;
  if (b->groups.size() == 1) return batch_group_step(b->groups[0], s);
  DSHT_HIP(hipEventRecord(b->fork, s));
  for (auto& g : b->groups) {
    DSHT_HIP(hipStreamWaitEvent(g.stream, b->fork, 0));
    if (int rc = batch_group_step(g, g.stream)) return rc;
    DSHT_HIP(hipEventRecord(g.done, g.stream));
  }
  for (auto& g : b->groups) DSHT_HIP(hipStreamWaitEvent(s, g.done, 0));
  return 0;
}

/* the unbatched baseline: each instance's step as its own two calls (its columns are
   contiguous), one instance after another on the caller's stream */
int dsht_batch_step_sequential(dsht_batch* b, void* stream) {
  if (!b || !b->finalized) return fail(DSHT_ESTATE, "batch not finalized");
  hipStream_t s = (hipStream_t)stream;
  for (auto& in : b->inst) {
    BatchGroup& g = b->groups[in.group];
    dsht_plan* p = g.plan;
    const long IJ = (long)p->I * p->J;
    const int C = NPROG * g.cols, CT = NTEND * g.cols;
    const int c = NPROG * in.col0, ct = NTEND * in.col0;
    double* out[NINV];
    for (int k = 0; k < NINV; k++) out[k] = g.grid + (size_t)k * C * IJ + (size_t)c * IJ;
    if (int rc = dsht_inv_vgrad_d(p, g.spec + c, C, 1, out[0], out[1], out[2], IJ, NPROG * in.km, nullptr, s))
      return rc;
    if (int rc = dsht_fwd_d(p, DSHT_VAL, g.tend + (size_t)ct * IJ, IJ, g.spect + ct, CT, 1, NTEND * in.km, nullptr, s))
      return rc;
  }
  return 0;
}

/* capture one batched step into a hipGraph (after one eager warm-up step on the same
   stream, so every workspace and table exists before capture begins) */
int dsht_batch_capture(dsht_batch* b, void* stream) {
  if (!b || !b->finalized) return fail(DSHT_ESTATE, "batch not finalized");
  hipStream_t s = (hipStream_t)stream;
  if (!s) return fail(DSHT_EARG, "dsht_batch_capture needs a non-default stream");
  if (int rc = dsht_batch_step(b, stream)) return rc;
  DSHT_HIP(hipStreamSynchronize(s));
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  DSHT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = dsht_batch_step(b, stream);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(DSHT_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
  b->graph = g;
  DSHT_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return 0;
}

int dsht_batch_launch(dsht_batch* b, void* stream) {
  if (!b || !b->exec) return fail(DSHT_ESTATE, "batch graph not captured");
  DSHT_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
  return 0;
}

int dsht_batch_destroy(dsht_batch* b) {
  if (!b) return 0;
  (void)hipDeviceSynchronize();
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->fork) (void)hipEventDestroy(b->fork);
  for (auto& g : b->groups) {
    if (g.done) (void)hipEventDestroy(g.done);
    if (g.stream) (void)hipStreamDestroy(g.stream);
    if (g.spec) (void)hipFree(g.spec);
    if (g.grid) (void)hipFree(g.grid);
    if (g.tend) (void)hipFree(g.tend);
    if (g.spect) (void)hipFree(g.spect);
  }
  delete b;
  return 0;
}

/* tuning knobs of a plan (set by the autotuner): column chunk of the device path, the
   column block of the Legendre contractions and the FFT kernel family */
int dsht_set_tuning(dsht_plan* p, int key, int value) {
  if (!p) return fail(DSHT_ESTATE, "NULL plan");
  switch (key) {
    case DSHT_TUNE_CHUNK_COLS:
      if (value < 0) return fail(DSHT_EARG, "chunk must be >= 0");
      p->chunk_cols = value;
      return 0;
    case DSHT_TUNE_LEG_CB:
      if (value != 0 && value != 16 && value != 32) return fail(DSHT_EARG, "column block must be 0, 16 or 32");
      p->leg_cb = value;
      return 0;
    case DSHT_TUNE_FFT:
      if (value < 0 || value > 3) return fail(DSHT_EARG, "fft kernel must be 0, 1, 2 or 3");
      p->fft_kind = value;
      return 0;
    case DSHT_TUNE_LEG:
      if (value < 0 || value > 3) return fail(DSHT_EARG, "contraction kernel must be 0, 1, 2 or 3");
      p->leg_kind = value;
      return 0;
    case DSHT_TUNE_FFT_ORDER:
      if (value != 0 && value != 1) return fail(DSHT_EARG, "fft row order must be 0 or 1");
      p->fft_order = value;
      return 0;
  }
  return fail(DSHT_EARG, "unknown tuning key %d", key);
}

int dsht_get_tuning(const dsht_plan* p, int key) {
  if (!p) return -1;
  switch (key) {
    case DSHT_TUNE_CHUNK_COLS: return p->chunk_cols;
    case DSHT_TUNE_LEG_CB: return p->leg_cb;
    case DSHT_TUNE_FFT: return p->fft_kind;
    case DSHT_TUNE_LEG: return p->leg_kind;
    case DSHT_TUNE_FFT_ORDER: return p->fft_order;
  }
  return -1;
}

}  // extern "C"
