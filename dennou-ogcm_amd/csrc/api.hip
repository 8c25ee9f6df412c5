// api.hip -- the C ABI of libdsht (include/dsht.h).
//
// Plan set-up replaces SpmlUtil_Init/wa_Initial (SpmlUtil_mod.f90:192-241); the host
// entry points replace the SpmlUtil/SPMODEL functions the Fortran model calls (the
// Fortran shim dennou-ogcm_amd/fortran/SpmlUtil_mod.f90 binds them); the *_d entry points
// are the device-resident forms used by batched and resident-state callers.
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <unordered_map>

#include "dsht_internal.hpp"

namespace dsht {

static thread_local char g_err[512];

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

// ---------------------------------------------------------------- kernel timers
// Per kernel class: HIP events around each launch (on the launch stream), plus the
// algorithmic bytes and flops of the launch, so callers can report achieved GB/s and
// TFLOP/s without an external profiler.  Named after the reference's rap timers
// (model/dogcm/src/util/misc/ProfUtil_mod.f90:134-210).
struct ProfEvent {
  std::string name;
  hipEvent_t a, b;
  double bytes, flops;
};
static std::atomic<int> g_prof_on{0};
static std::mutex g_prof_mu;
static std::vector<ProfEvent> g_prof_pending;
struct ProfAcc {
  double ms = 0, bytes = 0, flops = 0;
  long count = 0;
};
static std::map<std::string, ProfAcc> g_prof_acc;

ProfScope::ProfScope(const char* name, hipStream_t st, double bytes, double flops) {
  if (!g_prof_on.load(std::memory_order_relaxed)) return;
  ProfEvent e;
  e.name = name;
  e.bytes = bytes;
  e.flops = flops;
  if (hipEventCreate(&e.a) != hipSuccess || hipEventCreate(&e.b) != hipSuccess) return;
  (void)hipEventRecord(e.a, st);
  std::lock_guard<std::mutex> g(g_prof_mu);
  slot = (int)g_prof_pending.size();
  s = st;
  g_prof_pending.push_back(e);
}

ProfScope::~ProfScope() {
  if (slot < 0) return;
  std::lock_guard<std::mutex> g(g_prof_mu);
  (void)hipEventRecord(g_prof_pending[slot].b, s);
}

static void prof_flush() {
  std::lock_guard<std::mutex> g(g_prof_mu);
  for (auto& e : g_prof_pending) {
    float ms = 0.f;
    (void)hipEventSynchronize(e.b);
    (void)hipEventElapsedTime(&ms, e.a, e.b);
    ProfAcc& acc = g_prof_acc[e.name];
    acc.ms += ms;
    acc.bytes += e.bytes;
    acc.flops += e.flops;
    acc.count++;
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  g_prof_pending.clear();
}

// ---------------------------------------------------------------- workspaces
// One scratch buffer per (plan, stream); a call holds its lock while it enqueues, so
// calls on one stream are serialised by the stream and calls on different streams never
// share scratch.
int workspace_acquire(dsht_plan* p, hipStream_t s, size_t bytes, WsHandle& out) {
  Workspace* ws;
  {
    std::lock_guard<std::mutex> g(p->ws_mu);
    auto& slot = p->ws[(void*)s];
    if (!slot) slot.reset(new Workspace());
    ws = slot.get();
  }
  out.lk = std::unique_lock<std::mutex>(ws->mu);
  if (ws->bytes < bytes) {
    if (ws->buf) {
      DSHT_HIP(hipStreamSynchronize(s));
      DSHT_HIP(hipFree(ws->buf));
      ws->buf = nullptr;
      ws->bytes = 0;
    }
    if (hipMalloc(&ws->buf, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return fail(DSHT_ENOMEM, "workspace allocation of %zu bytes failed", bytes);
    }
    ws->bytes = bytes;
  }
  out.buf = ws->buf;
  out.ws = ws;
  return 0;
}

static int chunk_cols(const dsht_plan* p, int ncols) { return chunk_columns(p, ncols, sizeof(double2)); }

static size_t fourier_bytes(const dsht_plan* p, int ch) {
  return (size_t)(p->N + 1) * p->J * fourier_pitch(ch) * sizeof(double2);
}

static int check_plan(const dsht_plan* p) {
  if (!p || p->N < 0) return fail(DSHT_ESTATE, "plan is NULL or destroyed");
  return 0;
}

// forward family: grids -> spectra
static int run_fwd(dsht_plan* p, int op, const double* in0, const double* in1, long ldg, double* out0,
                   double* out1, long sl, long sc, int ncols, const double* diag0, const double* diag1,
                   double scale, const double* cw, hipStream_t s) {
  if (ncols <= 0) return 0;
  const int nin = (op == F_ALPHA || op == F_UV2VORDIV) ? 2 : 1;
  const int ch = chunk_cols(p, ncols);
  const int nchunks = cdiv(ncols, ch);
  const size_t slot_bytes = nin * fourier_bytes(p, ch);
  WsHandle wl;
  int rc = workspace_acquire(p, s, slot_bytes, wl);
  if (rc) return rc;
  for (int c0 = 0; c0 < ncols; c0 += ch) {
    const int nc = ncols - c0 < ch ? ncols - c0 : ch;
    double2* G0 = (double2*)wl.buf;
    double2* G1 = G0 + (size_t)(p->N + 1) * p->J * fourier_pitch(ch);
    const long gp = fourier_pitch(nc);
    if ((rc = launch_fft_fwd(p, in0 + (long)c0 * ldg, ldg, nc, G0, gp, s))) return rc;
    if (nin > 1 && (rc = launch_fft_fwd(p, in1 + (long)c0 * ldg, ldg, nc, G1, gp, s))) return rc;
    LegFwdArgs a;
    a.G0 = G0;
    a.G1 = nin > 1 ? G1 : nullptr;
    a.W0 = out0 + (long)c0 * sc;
    a.W1 = out1 ? out1 + (long)c0 * sc : nullptr;
    a.sl = sl;
    a.sc = sc;
    a.ncols = nc;
    a.gld = fourier_pitch(nc);
    a.gmst = (long)p->J * a.gld;
    a.tabP = p->d_tabFP;
    a.tabH = p->d_tabFH;
    a.offF = p->d_offF;
    a.cw = cw;
    a.diag0 = diag0;
    a.diag1 = diag1;
    a.scale = scale;
    a.N = p->N;
    a.J2 = p->J2;
    a.J2pad = p->J2pad;
    if ((rc = launch_leg_fwd(p, op, a, s))) return rc;
  }
  return 0;
}

// inverse family: spectra -> grids
// fixed/fixed_bytes: a caller-provided scratch (the static one of dsht_inv_noalloc_d); the
// call then takes no workspace and cuts its columns into chunks whose Fourier arrays fit the
// scratch
static int run_inv(dsht_plan* p, int op, const double* in0, const double* in1, long sl, long sc,
                   double* out0, double* out1, long ldg, int ncols, const double* pro, const double* rs,
                   hipStream_t s, double* out2 = nullptr, void* fixed = nullptr, size_t fixed_bytes = 0) {
  if (ncols <= 0) return 0;
  const int nout = inv_nout(op);
  int ch = chunk_cols(p, ncols);
  if (fixed) {
    // whole 16-column groups of every output's Fourier array within the scratch
    const size_t per16 = (size_t)nout * fourier_bytes(p, 16);
    const long fit = (long)(fixed_bytes / per16) * 16;
    if (fit < 16)
      return fail(DSHT_EARG, "the static scratch of %zu bytes cannot hold 16 columns (%zu bytes) of this plan",
                  fixed_bytes, per16);
    if (ch > fit) ch = (int)fit;
  }
  const int nchunks = cdiv(ncols, ch);
  const size_t slot_bytes = nout * fourier_bytes(p, ch);
  WsHandle wl;
  int rc = 0;
  if (fixed) {
    wl.buf = fixed;
  } else if ((rc = workspace_acquire(p, s, slot_bytes, wl))) {
    return rc;
  }
  for (int c0 = 0; c0 < ncols; c0 += ch) {
    const int nc = ncols - c0 < ch ? ncols - c0 : ch;
    double2* G0 = (double2*)wl.buf;
    double2* G1 = G0 + (size_t)(p->N + 1) * p->J * fourier_pitch(ch);
    double2* G2 = G1 + (size_t)(p->N + 1) * p->J * fourier_pitch(ch);
    LegInvArgs a;
    a.W0 = in0 + (long)c0 * sc;
    a.W1 = in1 ? in1 + (long)c0 * sc : nullptr;
    a.sl = sl;
    a.sc = sc;
    // fused value/gradient: GradLon comes from the value's Fourier coefficients through the
    // derivative inverse FFT (i m X_m, 1/cos row scale), so the contraction writes two outputs
    const bool dlon = op == I_VGRAD && out1 && fft_inv_deriv_ok(p);
    a.G0 = out0 || dlon ? G0 : nullptr;
    a.G1 = nout > 1 && out1 && !dlon ? G1 : nullptr;
    a.G2 = nout > 2 && out2 ? G2 : nullptr;
    a.ncols = nc;
    a.gld = fourier_pitch(nc);
    a.gmst = (long)p->J * a.gld;
    a.tabP = p->d_tabIP;
    a.tabH = p->d_tabIH;
    a.offI = p->d_offI;
    a.rs = rs;
    a.pro = pro;
    a.N = p->N;
    a.J = p->J;
    a.J2 = p->J2;
    a.J2pad = p->J2pad;
    if ((rc = launch_leg_inv(p, op, a, s))) return rc;
    hipStream_t fs = s;
    // the field and its longitude derivative as two launches over the same G0: the FP64
    // dual kernel (one read of G0, coefficients held in registers across both passes) ran
    // slower at T170L60 (2 x 125 against 2 x 80 us per step's worth), the FP32 one faster
    if (out0 && (rc = launch_fft_inv(p, G0, a.gld, nc, out0 + (long)c0 * ldg, ldg, fs))) return rc;
    if (dlon && (rc = launch_fft_inv(p, G0, a.gld, nc, out1 + (long)c0 * ldg, ldg, fs, p->d_rs_icos, 1))) return rc;
    if (a.G1 && (rc = launch_fft_inv(p, G1, a.gld, nc, out1 + (long)c0 * ldg, ldg, fs))) return rc;
    if (a.G2 && (rc = launch_fft_inv(p, G2, a.gld, nc, out2 + (long)c0 * ldg, ldg, fs))) return rc;
  }
  return 0;
}

static int check_grid_ptr(const void* x, long ldg, const dsht_plan* p) {
  if (!x) return fail(DSHT_EARG, "grid pointer is NULL");
  if (((uintptr_t)x) % 16) return fail(DSHT_EARG, "grid pointers must be 16-byte aligned");
  if (ldg < (long)p->I * p->J || (ldg & 1)) return fail(DSHT_EARG, "grid column pitch %ld must be even and >= I*J", ldg);
  return 0;
}

static int check_spec(const void* w, long sl, long sc, int ncols, const dsht_plan* p) {
  if (!w) return fail(DSHT_EARG, "spectral pointer is NULL");
  if (sl <= 0 || sc <= 0) return fail(DSHT_EARG, "spectral strides must be positive");
  bool lc = sc == 1 && sl >= ncols, cl = sl == 1 && sc >= p->L;
  if (!(lc || cl || ncols == 1)) return fail(DSHT_EARG, "spectral strides (sl=%ld, sc=%ld) overlap", sl, sc);
  return 0;
}

// ---------------------------------------------------------------- static scratch
// The Fourier scratch of dsht_inv_noalloc_d: a file-scope device array, one copy per GPU,
// held until the process ends.  Host code never passes the array's name to a kernel (on the
// host it is only a stand-in for the symbol); it uses the device address that
// hipGetSymbolAddress returns for the current device, looked up once per device at the
// first dsht_create on it.
constexpr size_t NOALLOC_SCRATCH_BYTES = DSHT_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) double2 scratch[NOALLOC_SCRATCH_BYTES / sizeof(double2)];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[MAX_DEVICES];
static hipError_t g_scratch_err[MAX_DEVICES];
static bool g_scratch_done[MAX_DEVICES];

// with `device` current: resolve this device's copy of the scratch (once per device)
static void scratch_init(int device) {
  if (device < 0 || device >= MAX_DEVICES) return;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_done[device]) return;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    ptr = nullptr;
  }
  g_scratch_ptr[device] = ptr;
  g_scratch_err[device] = e;
  g_scratch_done[device] = true;
}

static int scratch_of(int device, void** out) {
  if (device < 0 || device >= MAX_DEVICES) return fail(DSHT_EDEVICE, "device %d out of range", device);
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_done[device]) return fail(DSHT_ESTATE, "static scratch of device %d not initialised", device);
  if (!g_scratch_ptr[device])
    return fail(DSHT_EDEVICE, "hipGetSymbolAddress of the static scratch failed on device %d: %s", device,
                hipGetErrorString(g_scratch_err[device]));
  *out = g_scratch_ptr[device];
  return 0;
}

// ---------------------------------------------------------------- host contexts
struct HostCtx {
  hipStream_t s = nullptr;
  double* buf[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t cap[4] = {0, 0, 0, 0};
  double* dout = nullptr;  // small outputs
  size_t dcap = 0;
};

struct HostReg {
  std::mutex mu;
  std::vector<HostCtx*> all;
};

static std::atomic<long> g_plan_serial{1};
static std::mutex g_reg_mu;
static std::unordered_map<long, HostReg*> g_reg;

static thread_local std::unordered_map<long, HostCtx*> t_ctx;

}  // namespace dsht

using namespace dsht;

struct PlanExtra {
  long serial;
};

static std::mutex g_extra_mu;
static std::unordered_map<const dsht_plan*, long> g_serial;

static long plan_serial(const dsht_plan* p) {
  std::lock_guard<std::mutex> g(g_extra_mu);
  auto it = g_serial.find(p);
  return it == g_serial.end() ? 0 : it->second;
}

static int host_ctx(dsht_plan* p, HostCtx** out) {
  long id = plan_serial(p);
  if (!id) return fail(DSHT_ESTATE, "plan is not registered");
  auto it = t_ctx.find(id);
  if (it != t_ctx.end()) {
    *out = it->second;
    return 0;
  }
  HostCtx* c = new HostCtx();
  DSHT_HIP(hipSetDevice(p->device));
  DSHT_HIP(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
  {
    std::lock_guard<std::mutex> g(g_reg_mu);
    HostReg*& r = g_reg[id];
    if (!r) r = new HostReg();
    std::lock_guard<std::mutex> g2(r->mu);
    r->all.push_back(c);
  }
  t_ctx[id] = c;
  *out = c;
  return 0;
}

static int ctx_buf(HostCtx* c, int i, size_t n, double** out) {
  if (c->cap[i] < n) {
    if (c->buf[i]) DSHT_HIP(hipFree(c->buf[i]));
    c->buf[i] = nullptr;
    c->cap[i] = 0;
    if (hipMalloc(&c->buf[i], sizeof(double) * (n ? n : 1)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(DSHT_ENOMEM, "host-API device buffer of %zu doubles failed", n);
    }
    c->cap[i] = n;
  }
  *out = c->buf[i];
  return 0;
}

extern "C" {

const char* dsht_last_error(void) { return dsht::g_err; }

int dsht_prof_enable(int on) {
  if (!on) prof_flush();
  g_prof_on.store(on ? 1 : 0);
  return 0;
}

int dsht_prof_reset(void) {
  prof_flush();
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_prof_acc.clear();
  return 0;
}

int dsht_prof_query(const char* name, double* total_ms, long* count, double* bytes, double* flops) {
  if (!name) return fail(DSHT_EARG, "dsht_prof_query: NULL name");
  prof_flush();
  std::lock_guard<std::mutex> g(g_prof_mu);
  auto it = g_prof_acc.find(name);
  ProfAcc z;
  const ProfAcc& acc = it == g_prof_acc.end() ? z : it->second;
  if (total_ms) *total_ms = acc.ms;
  if (count) *count = acc.count;
  if (bytes) *bytes = acc.bytes;
  if (flops) *flops = acc.flops;
  return 0;
}

int dsht_create(int nmax, int im, int jm, int kmax, double radius, dsht_plan** plan) {
  if (!plan) return fail(DSHT_EARG, "dsht_create: plan pointer is NULL");
  *plan = nullptr;
  if (nmax < 0 || im < 2 || jm < 2) return fail(DSHT_EARG, "dsht_create: need nmax >= 0, im >= 2, jm >= 2");
  if (im % 2) return fail(DSHT_EARG, "dsht_create: im = %d must be even", im);
  if (jm % 2) return fail(DSHT_EARG, "dsht_create: jm = %d must be even", jm);
  if (im < 2 * nmax + 2) return fail(DSHT_EARG, "dsht_create: im = %d < 2*nmax+2 = %d", im, 2 * nmax + 2);
  if (jm < nmax + 1) return fail(DSHT_EARG, "dsht_create: jm = %d < nmax+1 = %d", jm, nmax + 1);
  if (!(radius > 0.0) || !std::isfinite(radius)) return fail(DSHT_EARG, "dsht_create: radius must be positive");
  if (kmax < 0) return fail(DSHT_EARG, "dsht_create: kmax must be >= 0");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return fail(DSHT_EDEVICE, "dsht_create: no HIP device available");
  }
  dsht_plan* p = new dsht_plan();
  p->N = nmax;
  p->I = im;
  p->J = jm;
  p->K = kmax;
  p->L = (nmax + 1) * (nmax + 1);
  p->J2 = jm / 2;
  p->J2pad = cdiv(p->J2, KCH) * KCH;
  p->R = radius;
  if (hipGetDevice(&p->device) != hipSuccess) {
    delete p;
    return fail(DSHT_EDEVICE, "dsht_create: hipGetDevice failed");
  }
  scratch_init(p->device);  // the plan's device is current here
  int rc = fft_plan_init(p);
  if (!rc) rc = build_tables(p);
  if (rc) {
    p->N = 0;
    dsht_destroy(p);
    return rc;
  }
  {
    std::lock_guard<std::mutex> g(g_extra_mu);
    g_serial[p] = g_plan_serial++;
  }
  *plan = p;
  return 0;
}

int dsht_destroy(dsht_plan* p) {
  if (!p) return 0;
  long id = plan_serial(p);
  if (id) {
    HostReg* r = nullptr;
    {
      std::lock_guard<std::mutex> g(g_reg_mu);
      auto it = g_reg.find(id);
      if (it != g_reg.end()) {
        r = it->second;
        g_reg.erase(it);
      }
    }
    if (r) {
      for (HostCtx* c : r->all) {
        if (c->s) (void)hipStreamSynchronize(c->s);
        for (int i = 0; i < 4; i++)
          if (c->buf[i]) (void)hipFree(c->buf[i]);
        if (c->dout) (void)hipFree(c->dout);
        if (c->s) (void)hipStreamDestroy(c->s);
        c->s = nullptr;
        for (int i = 0; i < 4; i++) c->buf[i] = nullptr;
        c->dout = nullptr;
        // the owning thread's map still holds the pointer under the dead serial id; it is
        // never looked up again (serials are not reused) so the small struct is leaked.
      }
      delete r;
    }
    std::lock_guard<std::mutex> g(g_extra_mu);
    g_serial.erase(p);
  }
  (void)hipDeviceSynchronize();
  for (auto& kv : p->ws) {
    if (!kv.second) continue;
    if (kv.second->buf) (void)hipFree(kv.second->buf);
  }
  double* ds[] = {p->d_mu, p->d_wq, p->d_wl, p->d_cwv, p->d_cwd, p->d_rs_icos, p->d_rs_alpha, p->d_rs_uv,
                  p->d_pro_psi, p->d_lat_cs, p->d_tabFP, p->d_tabFH, p->d_tabIP, p->d_tabIH};
  for (double* d : ds)
    if (d) (void)hipFree(d);
  if (p->d_offF) (void)hipFree(p->d_offF);
  if (p->d_offI) (void)hipFree(p->d_offI);
  fft_plan_free(p);
  mixed_free(p);
  p->N = -1;
  delete p;
  return 0;
}

int dsht_get_dims(const dsht_plan* p, int* nmax, int* im, int* jm, int* lm) {
  if (int rc = check_plan(p)) return rc;
  if (nmax) *nmax = p->N;
  if (im) *im = p->I;
  if (jm) *jm = p->J;
  if (lm) *lm = p->L;
  return 0;
}

int dsht_set_precision(dsht_plan* p, int prec) {
  if (int rc = check_plan(p)) return rc;
  if (prec < DSHT_PREC_FP64 || prec > DSHT_PREC_FP16X3) return fail(DSHT_EARG, "unknown precision form %d", prec);
  p->prec = prec;
  return 0;
}

int dsht_get_precision(const dsht_plan* p) { return p ? p->prec : -1; }

int dsht_grid_info(const dsht_plan* p, double* x_lon, double* y_lat, double* x_lon_weight,
                   double* y_lat_weight, double* xy_lon, double* xy_lat) {
  if (int rc = check_plan(p)) return rc;
  for (int i = 0; i < p->I; i++) {
    if (x_lon) x_lon[i] = p->lon[i];
    if (x_lon_weight) x_lon_weight[i] = 2.0 * M_PI / p->I;
  }
  for (int j = 0; j < p->J; j++) {
    if (y_lat) y_lat[j] = p->lat[j];
    if (y_lat_weight) y_lat_weight[j] = p->wl[j];
    for (int i = 0; i < p->I; i++) {
      if (xy_lon) xy_lon[(long)j * p->I + i] = p->lon[i];
      if (xy_lat) xy_lat[(long)j * p->I + i] = p->lat[j];
    }
  }
  return 0;
}

int dsht_sig_grid(int kmax, int tmax, double sig_min, double sig_max, double* sig, double* sig_weight) {
  if (kmax <= 0 || tmax <= 0) return fail(DSHT_EARG, "dsht_sig_grid: kmax = %d and tmax = %d must be positive", kmax, tmax);
  if (kmax % 2) return fail(DSHT_EARG, "dsht_sig_grid: kmax = %d must be even", kmax);
  if (tmax > kmax) return fail(DSHT_EARG, "dsht_sig_grid: tmax = %d must be <= kmax = %d", tmax, kmax);
  const double half = 0.5 * (sig_max - sig_min);
  for (int i = 0; i <= kmax; i++) {
    const double th = M_PI * i / kmax;
    if (sig) sig[i] = 0.5 * (sig_max + sig_min) + half * std::cos(th);
    if (!sig_weight) continue;
    // Clenshaw-Curtis: integrate the degree-tmax Chebyshev interpolant exactly; the even
    // modes k integrate to 2/(1-k^2), the last one counted half when tmax = kmax
    double acc = 1.0;
    for (int k = 2; k <= tmax; k += 2) {
      const double c = 2.0 / (1.0 - (double)k * k) * std::cos(k * th);
      acc += (k == kmax) ? 0.5 * c : c;
    }
    double wgt = 2.0 / kmax * acc * half;
    if (i == 0 || i == kmax) wgt *= 0.5;
    sig_weight[i] = wgt;
  }
  return 0;
}

int dsht_nm_l(const dsht_plan* p, int* n_of_l, int* m_of_l) {
  if (int rc = check_plan(p)) return rc;
  for (int n = 0; n <= p->N; n++)
    for (int m = -n; m <= n; m++) {
      int l = n * (n + 1) + m;
      if (n_of_l) n_of_l[l] = n;
      if (m_of_l) m_of_l[l] = m;
    }
  return 0;
}

int dsht_l_nm(const dsht_plan* p, int n, int m) {
  if (!p || n < 0 || n > p->N || m < -n || m > n) return 0;
  return n * (n + 1) + m + 1;
}

int dsht_rn(const dsht_plan* p, double* rn) {
  if (int rc = check_plan(p)) return rc;
  if (!rn) return fail(DSHT_EARG, "dsht_rn: NULL output");
  for (int n = 0; n <= p->N; n++)
    for (int m = -n; m <= n; m++) rn[n * (n + 1) + m] = -(double)n * (n + 1);
  return 0;
}

// ---------------------------------------------------------------- device API
static int kind_fwd_op(int kind) {
  switch (kind) {
    case DSHT_VAL: return F_VAL;
    case DSHT_DIVLAMBDA: return F_DIVLAMBDA;
    case DSHT_DIVMU: return F_DIVMU;
  }
  return -1;
}

static int dispatch_fwd(dsht_plan* p, int op, const double* in0, const double* in1, long ldg, double* out0,
                        double* out1, long sl, long sc, int ncols, const double* diag0, const double* diag1,
                        double scale, const double* cw, hipStream_t s) {
  if (p->prec != DSHT_PREC_FP64)
    return mixed_run_fwd(p, op, in0, in1, ldg, out0, out1, sl, sc, ncols, diag0, diag1, scale, cw, s);
  return run_fwd(p, op, in0, in1, ldg, out0, out1, sl, sc, ncols, diag0, diag1, scale, cw, s);
}

static int dispatch_inv(dsht_plan* p, int op, const double* in0, const double* in1, long sl, long sc,
                        double* out0, double* out1, long ldg, int ncols, const double* pro, const double* rs,
                        hipStream_t s, double* out2 = nullptr) {
  if (p->prec != DSHT_PREC_FP64)
    return mixed_run_inv(p, op, in0, in1, sl, sc, out0, out1, ldg, ncols, pro, rs, s, out2);
  return run_inv(p, op, in0, in1, sl, sc, out0, out1, ldg, ncols, pro, rs, s, out2);
}

int dsht_fwd_d(dsht_plan* p, int kind, const double* xy, long ldg, double* w, long sl, long sc, int ncols,
               const double* diag, void* stream) {
  if (int rc = check_plan(p)) return rc;
  int op = kind_fwd_op(kind);
  if (op < 0) return fail(DSHT_EARG, "dsht_fwd_d: kind %d is not a forward kind", kind);
  if (ncols < 0) return fail(DSHT_EARG, "ncols < 0");
  if (ncols == 0) return 0;
  if (int rc = check_grid_ptr(xy, ldg, p)) return rc;
  if (int rc = check_spec(w, sl, sc, ncols, p)) return rc;
  return dispatch_fwd(p, op, xy, nullptr, ldg, w, nullptr, sl, sc, ncols, diag, nullptr, 1.0,
                      op == F_VAL ? p->d_cwv : p->d_cwd, (hipStream_t)stream);
}

int dsht_inv_d(dsht_plan* p, int kind, const double* w, long sl, long sc, double* xy, long ldg, int ncols,
               const double* diag, void* stream) {
  if (int rc = check_plan(p)) return rc;
  int op;
  const double* rs = nullptr;
  switch (kind) {
    case DSHT_VAL: op = I_VAL; break;
    case DSHT_GRADLAMBDA: op = I_GRADLAMBDA; break;
    case DSHT_GRADMU: op = I_GRADMU; break;
    case DSHT_GRADLON: op = I_GRADLAMBDA; rs = p->d_rs_icos; break;
    case DSHT_GRADLAT: op = I_GRADMU; rs = p->d_rs_icos; break;
    default: return fail(DSHT_EARG, "dsht_inv_d: kind %d is not an inverse kind", kind);
  }
  if (ncols < 0) return fail(DSHT_EARG, "ncols < 0");
  if (ncols == 0) return 0;
  if (int rc = check_grid_ptr(xy, ldg, p)) return rc;
  if (int rc = check_spec(w, sl, sc, ncols, p)) return rc;
  return dispatch_inv(p, op, w, nullptr, sl, sc, xy, nullptr, ldg, ncols, diag, rs, (hipStream_t)stream);
}

int dsht_inv_noalloc_d(dsht_plan* p, int kind, const double* w, long sl, long sc, double* xy, long ldg, int ncols,
                       const double* diag, void* stream) {
  if (int rc = check_plan(p)) return rc;
  int op;
  const double* rs = nullptr;
  switch (kind) {
    case DSHT_VAL: op = I_VAL; break;
    case DSHT_GRADLAMBDA: op = I_GRADLAMBDA; break;
    case DSHT_GRADMU: op = I_GRADMU; break;
    case DSHT_GRADLON: op = I_GRADLAMBDA; rs = p->d_rs_icos; break;
    case DSHT_GRADLAT: op = I_GRADMU; rs = p->d_rs_icos; break;
    default: return fail(DSHT_EARG, "dsht_inv_noalloc_d: kind %d is not an inverse kind", kind);
  }
  if (p->prec != DSHT_PREC_FP64)
    return fail(DSHT_EUNSUPPORTED, "dsht_inv_noalloc_d runs the FP64 form only (the plan is set to form %d)", p->prec);
  if (ncols < 0) return fail(DSHT_EARG, "ncols < 0");
  if (ncols == 0) return 0;
  if (int rc = check_grid_ptr(xy, ldg, p)) return rc;
  if (int rc = check_spec(w, sl, sc, ncols, p)) return rc;
  void* buf = nullptr;
  if (int rc = scratch_of(p->device, &buf)) return rc;
  return run_inv(p, op, w, nullptr, sl, sc, xy, nullptr, ldg, ncols, diag, rs, (hipStream_t)stream, nullptr, buf,
                 NOALLOC_SCRATCH_BYTES);
}

int dsht_inv_vgrad_d(dsht_plan* p, const double* w, long sl, long sc, double* xy, double* xy_lon, double* xy_lat,
                     long ldg, int ncols, const double* diag, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (ncols < 0) return fail(DSHT_EARG, "ncols < 0");
  if (!xy && !xy_lon && !xy_lat) return fail(DSHT_EARG, "dsht_inv_vgrad_d: every output is NULL");
  if (ncols == 0) return 0;
  double* outs[3] = {xy, xy_lon, xy_lat};
  for (double* o : outs)
    if (o)
      if (int rc = check_grid_ptr(o, ldg, p)) return rc;
  if (int rc = check_spec(w, sl, sc, ncols, p)) return rc;
  hipStream_t s = (hipStream_t)stream;
  return dispatch_inv(p, I_VGRAD, w, nullptr, sl, sc, xy, xy_lon, ldg, ncols, diag, p->d_rs_icos, s, xy_lat);
}

// ---- FP32 grid I/O of the split forms: the same calls with float grids
static int check_f32(const dsht_plan* p) {
  if (p->prec == DSHT_PREC_FP64)
    return fail(DSHT_EUNSUPPORTED, "FP32 grid I/O runs the split precision forms (the plan is set to FP64)");
  if (p->fft_kind != 0 || !reg_fft_supported(p->fft.M))
    return fail(DSHT_EUNSUPPORTED, "FP32 grid I/O needs the register FFT (DSHT_TUNE_FFT 0, I/2 with a register kernel)");
  return 0;
}

int dsht_fwd_f32_d(dsht_plan* p, int kind, const float* xy, long ldg, double* w, long sl, long sc, int ncols,
                   const double* diag, void* stream) {
  if (int rc = check_plan(p)) return rc;
  int op = kind_fwd_op(kind);
  if (op < 0) return fail(DSHT_EARG, "dsht_fwd_f32_d: kind %d is not a forward kind", kind);
  if (ncols < 0) return fail(DSHT_EARG, "ncols < 0");
  if (ncols == 0) return 0;
  if (int rc = check_f32(p)) return rc;
  if (int rc = check_grid_ptr(xy, ldg, p)) return rc;
  if (int rc = check_spec(w, sl, sc, ncols, p)) return rc;
  return mixed_run_fwd(p, op, (const double*)xy, nullptr, ldg, w, nullptr, sl, sc, ncols, diag, nullptr, 1.0,
                       op == F_VAL ? p->d_cwv : p->d_cwd, (hipStream_t)stream, true);
}

int dsht_inv_f32_d(dsht_plan* p, int kind, const double* w, long sl, long sc, float* xy, long ldg, int ncols,
                   const double* diag, void* stream) {
  if (int rc = check_plan(p)) return rc;
  int op;
  const double* rs = nullptr;
  switch (kind) {
    case DSHT_VAL: op = I_VAL; break;
    case DSHT_GRADLAMBDA: op = I_GRADLAMBDA; break;
    case DSHT_GRADMU: op = I_GRADMU; break;
    case DSHT_GRADLON: op = I_GRADLAMBDA; rs = p->d_rs_icos; break;
    case DSHT_GRADLAT: op = I_GRADMU; rs = p->d_rs_icos; break;
    default: return fail(DSHT_EARG, "dsht_inv_f32_d: kind %d is not an inverse kind", kind);
  }
  if (ncols < 0) return fail(DSHT_EARG, "ncols < 0");
  if (ncols == 0) return 0;
  if (int rc = check_f32(p)) return rc;
  if (int rc = check_grid_ptr(xy, ldg, p)) return rc;
  if (int rc = check_spec(w, sl, sc, ncols, p)) return rc;
  return mixed_run_inv(p, op, w, nullptr, sl, sc, (double*)xy, nullptr, ldg, ncols, diag, rs, (hipStream_t)stream,
                       nullptr, true);
}

int dsht_inv_vgrad_f32_d(dsht_plan* p, const double* w, long sl, long sc, float* xy, float* xy_lon, float* xy_lat,
                         long ldg, int ncols, const double* diag, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (ncols < 0) return fail(DSHT_EARG, "ncols < 0");
  if (!xy && !xy_lon && !xy_lat) return fail(DSHT_EARG, "dsht_inv_vgrad_f32_d: every output is NULL");
  if (ncols == 0) return 0;
  if (int rc = check_f32(p)) return rc;
  float* outs[3] = {xy, xy_lon, xy_lat};
  for (float* o : outs)
    if (o)
      if (int rc = check_grid_ptr(o, ldg, p)) return rc;
  if (int rc = check_spec(w, sl, sc, ncols, p)) return rc;
  return mixed_run_inv(p, I_VGRAD, w, nullptr, sl, sc, (double*)xy, (double*)xy_lon, ldg, ncols, diag, p->d_rs_icos,
                       (hipStream_t)stream, (double*)xy_lat, true);
}

int dsht_alpha_fwd_d(dsht_plan* p, const double* a, const double* b, long ldg, double* w, long sl, long sc,
                     int ncols, const double* diag, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (ncols < 0) return fail(DSHT_EARG, "ncols < 0");
  if (ncols == 0) return 0;
  if (int rc = check_grid_ptr(a, ldg, p)) return rc;
  if (int rc = check_grid_ptr(b, ldg, p)) return rc;
  if (int rc = check_spec(w, sl, sc, ncols, p)) return rc;
  return dispatch_fwd(p, F_ALPHA, a, b, ldg, w, nullptr, sl, sc, ncols, diag, nullptr, 1.0 / p->R, p->d_cwd,
                      (hipStream_t)stream);
}

int dsht_alpha_inv_d(dsht_plan* p, const double* wa, const double* wb, long sl, long sc, double* xy, long ldg,
                     int ncols, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (ncols < 0) return fail(DSHT_EARG, "ncols < 0");
  if (ncols == 0) return 0;
  if (int rc = check_grid_ptr(xy, ldg, p)) return rc;
  if (int rc = check_spec(wa, sl, sc, ncols, p)) return rc;
  if (int rc = check_spec(wb, sl, sc, ncols, p)) return rc;
  return dispatch_inv(p, I_ALPHA, wa, wb, sl, sc, xy, nullptr, ldg, ncols, nullptr, p->d_rs_alpha,
                      (hipStream_t)stream);
}

int dsht_uv2vordiv_d(dsht_plan* p, const double* ucos, const double* vcos, long ldg, double* vor, double* div,
                     long sl, long sc, int ncols, const double* diag, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (ncols < 0) return fail(DSHT_EARG, "ncols < 0");
  if (ncols == 0) return 0;
  if (int rc = check_grid_ptr(ucos, ldg, p)) return rc;
  if (int rc = check_grid_ptr(vcos, ldg, p)) return rc;
  if (int rc = check_spec(vor, sl, sc, ncols, p)) return rc;
  if (int rc = check_spec(div, sl, sc, ncols, p)) return rc;
  return dispatch_fwd(p, F_UV2VORDIV, ucos, vcos, ldg, vor, div, sl, sc, ncols, diag, diag, 1.0 / p->R, p->d_cwd,
                      (hipStream_t)stream);
}

int dsht_vordiv2uv_d(dsht_plan* p, const double* vor, const double* div, long sl, long sc, double* u, double* v,
                     long ldg, int ncols, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (ncols < 0) return fail(DSHT_EARG, "ncols < 0");
  if (ncols == 0) return 0;
  if (int rc = check_grid_ptr(u, ldg, p)) return rc;
  if (int rc = check_grid_ptr(v, ldg, p)) return rc;
  if (int rc = check_spec(vor, sl, sc, ncols, p)) return rc;
  if (int rc = check_spec(div, sl, sc, ncols, p)) return rc;
  return dispatch_inv(p, I_VORDIV2UV, vor, div, sl, sc, u, v, ldg, ncols, p->d_pro_psi, p->d_rs_uv,
                      (hipStream_t)stream);
}

int dsht_diag_d(dsht_plan* p, const double* d, const double* w, double* out, long sl, long sc, int ncols,
                void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (!d) return fail(DSHT_EARG, "dsht_diag_d: NULL diagonal");
  if (ncols < 0) return fail(DSHT_EARG, "ncols < 0");
  if (ncols == 0) return 0;
  if (int rc = check_spec(w, sl, sc, ncols, p)) return rc;
  if (int rc = check_spec(out, sl, sc, ncols, p)) return rc;
  return launch_diag(d, w, out, sl, sc, p->L, ncols, (hipStream_t)stream);
}

int dsht_spec_energy_d(dsht_plan* p, int kind, const double* w, long sl, long sc, int ncols, double* nma, double* na,
                       void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (kind != DSHT_SPEC_ENERGY && kind != DSHT_SPEC_ENSTROPHY)
    return fail(DSHT_EARG, "dsht_spec_energy_d: kind must be DSHT_SPEC_ENERGY or DSHT_SPEC_ENSTROPHY");
  if (ncols < 0) return fail(DSHT_EARG, "ncols < 0");
  if (!nma && !na) return fail(DSHT_EARG, "dsht_spec_energy_d: both outputs are NULL");
  if (ncols == 0) return 0;
  if (int rc = check_spec(w, sl, sc, ncols, p)) return rc;
  return launch_spec_energy(p, kind == DSHT_SPEC_ENSTROPHY, w, sl, sc, ncols, nma, na, (hipStream_t)stream);
}

int dsht_intlonlat_d(dsht_plan* p, const double* xy, long ldg, double* out, int ncols, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (ncols < 0 || !out) return fail(DSHT_EARG, "bad arguments");
  if (ncols == 0) return 0;
  if (int rc = check_grid_ptr(xy, ldg, p)) return rc;
  return launch_lonlat_sum(p, xy, ldg, ncols, out, 2.0 * M_PI / p->I, (hipStream_t)stream);
}

int dsht_avrlonlat_d(dsht_plan* p, const double* xy, long ldg, double* out, int ncols, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (ncols < 0 || !out) return fail(DSHT_EARG, "bad arguments");
  if (ncols == 0) return 0;
  if (int rc = check_grid_ptr(xy, ldg, p)) return rc;
  return launch_lonlat_sum(p, xy, ldg, ncols, out, 1.0 / (2.0 * p->I), (hipStream_t)stream);
}

int dsht_ya_intlon_d(dsht_plan* p, const double* xy, long ldg, double* out, int ncols, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (ncols < 0 || !out) return fail(DSHT_EARG, "bad arguments");
  if (ncols == 0) return 0;
  if (int rc = check_grid_ptr(xy, ldg, p)) return rc;
  return launch_lon_sum(p, xy, ldg, ncols, out, 2.0 * M_PI / p->I, p->J, (hipStream_t)stream);
}

int dsht_interpolate_d(dsht_plan* p, const double* w, long sl, const double* lon, const double* lat, double* out,
                       int npts, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (npts < 0 || !w || !lon || !lat || !out || sl <= 0) return fail(DSHT_EARG, "bad arguments");
  return launch_interpolate(p, w, sl, lon, lat, out, npts, (hipStream_t)stream);
}

// ---------------------------------------------------------------- vertical column diagnostics
// address ranges of device arrays, for the no-aliasing checks of the column kernels
// (their outputs are written while inputs are still read; the kernels take restrict pointers)
struct Span {
  uintptr_t lo, hi;
};
static Span span_of(const double* x, long n) {
  const uintptr_t a = (uintptr_t)x;
  return Span{a, a + (uintptr_t)(x && n > 0 ? n : 0) * sizeof(double)};
}
static bool overlaps(const Span& a, const Span& b) { return a.lo < b.hi && b.lo < a.hi && a.hi > a.lo && b.hi > b.lo; }

// fails unless the output span is disjoint from every input span
static int check_disjoint(const char* what, Span out, const Span* ins, const char* const* names, int n) {
  for (int i = 0; i < n; i++)
    if (overlaps(out, ins[i]))
      return fail(DSHT_EARG, "the output %s overlaps the input %s (in-place calls are not supported)", what, names[i]);
  return 0;
}

static int check_col(const dsht_plan* p, int km, const double* cdk, long ldg) {
  if (int rc = check_plan(p)) return rc;
  if (km < 1) return fail(DSHT_EARG, "km must be >= 1");
  if (!cdk) return fail(DSHT_EARG, "layer thickness pointer is NULL");
  if (ldg < (long)p->I * p->J) return fail(DSHT_EARG, "grid column pitch %ld must be >= I*J", ldg);
  return 0;
}

int dsht_col_omega_d(dsht_plan* p, const double* div, const double* h, long ldg, int km, const double* cdk,
                     double* omg, long ldo, void* stream) {
  if (int rc = check_col(p, km, cdk, ldg)) return rc;
  if (!div || !h || !omg) return fail(DSHT_EARG, "grid pointer is NULL");
  if (ldo < (long)p->I * p->J) return fail(DSHT_EARG, "output column pitch %ld must be >= I*J", ldo);
  const long IJ = (long)p->I * p->J;
  const Span ins[] = {span_of(div, (km - 1) * ldg + IJ), span_of(h, (km - 1) * ldg + IJ), span_of(cdk, km)};
  static const char* const names[] = {"div", "h", "cdk"};
  if (int rc = check_disjoint("omega", span_of(omg, km * ldo + IJ), ins, names, 3)) return rc;
  return launch_col_omega(div, h, ldg, km, cdk, omg, ldo, p->I * p->J, (hipStream_t)stream);
}

int dsht_col_hydpres_d(dsht_plan* p, const double* dens, const double* h, long ldg, int km, const double* cdk,
                       double grav, double* out, long ldo, void* stream) {
  if (int rc = check_col(p, km, cdk, ldg)) return rc;
  if (!dens || !h || !out) return fail(DSHT_EARG, "grid pointer is NULL");
  if (ldo < (long)p->I * p->J) return fail(DSHT_EARG, "output column pitch %ld must be >= I*J", ldo);
  const long IJ = (long)p->I * p->J;
  const Span ins[] = {span_of(dens, (km - 1) * ldg + IJ), span_of(h, (km - 1) * ldg + IJ), span_of(cdk, km)};
  static const char* const names[] = {"dens", "h", "cdk"};
  if (int rc = check_disjoint("pressure", span_of(out, (km - 1) * ldo + IJ), ins, names, 3)) return rc;
  return launch_col_hydpres(dens, h, ldg, km, cdk, grav, out, ldo, p->I * p->J, (hipStream_t)stream);
}

int dsht_col_uvbarot_d(dsht_plan* p, const double* u, const double* v, const double* h, long ldg, int km,
                       const double* cdk, double* ub, double* vb, void* stream) {
  if (int rc = check_col(p, km, cdk, ldg)) return rc;
  if (!u || !v || !h) return fail(DSHT_EARG, "grid pointer is NULL");
  if (!ub && !vb) return fail(DSHT_EARG, "both outputs are NULL");
  const long IJ = (long)p->I * p->J;
  const Span ins[] = {span_of(u, (km - 1) * ldg + IJ), span_of(v, (km - 1) * ldg + IJ),
                      span_of(h, (km - 1) * ldg + IJ), span_of(cdk, km)};
  static const char* const names[] = {"u", "v", "h", "cdk"};
  if (int rc = check_disjoint("ub", span_of(ub, IJ), ins, names, 4)) return rc;
  if (int rc = check_disjoint("vb", span_of(vb, IJ), ins, names, 4)) return rc;
  if (overlaps(span_of(ub, IJ), span_of(vb, IJ))) return fail(DSHT_EARG, "the outputs ub and vb overlap");
  return launch_col_uvbarot(u, v, h, ldg, km, cdk, ub, vb, p->I * p->J, (hipStream_t)stream);
}

int dsht_col_quick_d(dsht_plan* p, const double* omg, long ldo, const double* trc, long ldg, int km,
                     const double* cdk, const double* rfdk, int linfree, double* flx, long ldf, void* stream) {
  if (int rc = check_col(p, km, cdk, ldg)) return rc;
  if (km < 2) return fail(DSHT_EARG, "the QUICK flux needs km >= 2");
  if (!omg || !trc || !flx || !rfdk) return fail(DSHT_EARG, "pointer is NULL");
  if (ldo < (long)p->I * p->J || ldf < (long)p->I * p->J) return fail(DSHT_EARG, "half-level pitch must be >= I*J");
  const long IJ = (long)p->I * p->J;
  const Span ins[] = {span_of(omg, km * ldo + IJ), span_of(trc, (km - 1) * ldg + IJ), span_of(cdk, km),
                      span_of(rfdk, km + 1)};
  static const char* const names[] = {"omega", "tracer", "cdk", "rfdk"};
  if (int rc = check_disjoint("flux", span_of(flx, km * ldf + IJ), ins, names, 4)) return rc;
  return launch_col_quick(omg, ldo, trc, ldg, km, cdk, rfdk, linfree ? 1 : 0, flx, ldf, p->I * p->J,
                          (hipStream_t)stream);
}

static int col_vdiff(dsht_plan* p, const double* q0, const double* rhs, const double* kv, const double* e3, long ldg,
                     int km, const double* cdk, const double* fdk, double dt, double alpha, int sfc_bc, int btm_bc,
                     const double* bfric, double* qa, long ldo, double* work, long lwork, void* stream) {
  if (int rc = check_col(p, km, cdk, ldg)) return rc;
  if (km < 2) return fail(DSHT_EARG, "the vertical diffusion solve needs km >= 2");
  if (!q0 || !rhs || !kv || !e3 || !qa || !fdk || !work) return fail(DSHT_EARG, "pointer is NULL");
  if (ldo < (long)p->I * p->J) return fail(DSHT_EARG, "output column pitch %ld must be >= I*J", ldo);
  if ((sfc_bc != 'D' && sfc_bc != 'N') || (btm_bc != 'D' && btm_bc != 'N'))
    return fail(DSHT_EARG, "boundary types must be 'D' or 'N'");
  const long need = 3l * km * p->I * p->J;
  if (lwork < need) return fail(DSHT_EARG, "workspace of %ld doubles is below 3*km*I*J = %ld", lwork, need);
  // the solve writes qa (and the workspace) while it still reads the inputs: no output may
  // share memory with an input
  const long IJ = (long)p->I * p->J;
  const Span out_qa = span_of(qa, (km - 1) * ldo + IJ), out_w = span_of(work, lwork);
  const Span ins[] = {span_of(q0, (km - 1) * ldg + IJ), span_of(rhs, (km - 1) * ldg + IJ),
                      span_of(kv, (km - 1) * ldg + IJ), span_of(e3, (km - 1) * ldg + IJ),
                      span_of(bfric, bfric ? IJ : 0), span_of(cdk, km), span_of(fdk, km + 1)};
  static const char* const names[] = {"q0", "rhs", "kv", "e3", "bfric", "cdk", "fdk"};
  if (int rc = check_disjoint("qa", out_qa, ins, names, 7)) return rc;
  if (int rc = check_disjoint("workspace", out_w, ins, names, 7)) return rc;
  if (overlaps(out_qa, out_w)) return fail(DSHT_EARG, "the output qa overlaps the workspace");
  return launch_col_vdiff(q0, rhs, kv, e3, ldg, km, cdk, fdk, dt, alpha, sfc_bc == 'D', btm_bc == 'D', bfric, qa, ldo,
                          work, p->I * p->J, (hipStream_t)stream);
}

int dsht_col_vdiff_d(dsht_plan* p, const double* q0, const double* rhs, const double* kv, const double* e3, long ldg,
                     int km, const double* cdk, const double* fdk, double dt, double alpha, int sfc_bc, int btm_bc,
                     double* qa, long ldo, double* work, long lwork, void* stream) {
  return col_vdiff(p, q0, rhs, kv, e3, ldg, km, cdk, fdk, dt, alpha, sfc_bc, btm_bc, nullptr, qa, ldo, work, lwork,
                   stream);
}

int dsht_col_vdiff_mom_d(dsht_plan* p, const double* q0, const double* rhs, const double* kv, const double* e3,
                         long ldg, int km, const double* cdk, const double* fdk, double dt, double alpha, int sfc_bc,
                         int btm_bc, const double* bfric, double* qa, long ldo, double* work, long lwork,
                         void* stream) {
  if (!bfric) return fail(DSHT_EARG, "bottom friction coefficient pointer is NULL");
  return col_vdiff(p, q0, rhs, kv, e3, ldg, km, cdk, fdk, dt, alpha, sfc_bc, btm_bc, bfric, qa, ldo, work, lwork,
                   stream);
}

int dsht_coef_d(dsht_plan* p, int which, const double* params, int nparams, double* out, void* stream) {
  if (int rc = check_plan(p)) return rc;
  if (which < DSHT_COEF_LAPLA || which > DSHT_COEF_BEDIFF) return fail(DSHT_EARG, "unknown coefficient table %d", which);
  if (!out || nparams < 0 || nparams > 8) return fail(DSHT_EARG, "bad arguments");
  static const int need[] = {0, 0, 3, 3, 0, 2, 5};
  if (nparams < need[which]) return fail(DSHT_EARG, "coefficient table %d needs %d parameters", which, need[which]);
  return launch_coef(p, which, params, nparams, out, (hipStream_t)stream);
}

// ---------------------------------------------------------------- host API
// copy helpers
static int h2d(HostCtx* c, double* d, const double* h, size_t n) {
  DSHT_HIP(hipMemcpyAsync(d, h, sizeof(double) * n, hipMemcpyHostToDevice, c->s));
  return 0;
}
static int d2h_sync(HostCtx* c, double* h, const double* d, size_t n) {
  DSHT_HIP(hipMemcpyAsync(h, d, sizeof(double) * n, hipMemcpyDeviceToHost, c->s));
  DSHT_HIP(hipStreamSynchronize(c->s));
  return 0;
}

#define HOST_PROLOGUE(p, ncols)                                 \
  if (int rc_ = check_plan(p)) return rc_;                      \
  if ((ncols) < 0) return fail(DSHT_EARG, "ncols < 0");         \
  if ((ncols) == 0) return 0;                                   \
  HostCtx* c;                                                   \
  if (int rc_ = host_ctx(p, &c)) return rc_;                    \
  DSHT_HIP(hipSetDevice(p->device));                            \
  const size_t IJ = (size_t)p->I * p->J, LL = (size_t)p->L;     \
  (void)IJ;                                                     \
  (void)LL;

int dsht_w_xy(dsht_plan* p, int kind, const double* xy, double* w, int ncols) {
  HOST_PROLOGUE(p, ncols);
  if (!xy || !w) return fail(DSHT_EARG, "NULL array");
  double *dg, *dw;
  int rc;
  if ((rc = ctx_buf(c, 0, IJ * ncols, &dg)) || (rc = ctx_buf(c, 1, LL * ncols, &dw))) return rc;
  if ((rc = h2d(c, dg, xy, IJ * ncols))) return rc;
  if ((rc = dsht_fwd_d(p, kind, dg, (long)IJ, dw, 1, (long)LL, ncols, nullptr, c->s))) return rc;
  return d2h_sync(c, w, dw, LL * ncols);
}

int dsht_xy_w(dsht_plan* p, int kind, const double* w, double* xy, int ncols) {
  HOST_PROLOGUE(p, ncols);
  if (!xy || !w) return fail(DSHT_EARG, "NULL array");
  double *dg, *dw;
  int rc;
  if ((rc = ctx_buf(c, 0, IJ * ncols, &dg)) || (rc = ctx_buf(c, 1, LL * ncols, &dw))) return rc;
  if ((rc = h2d(c, dw, w, LL * ncols))) return rc;
  if ((rc = dsht_inv_d(p, kind, dw, 1, (long)LL, dg, (long)IJ, ncols, nullptr, c->s))) return rc;
  return d2h_sync(c, xy, dg, IJ * ncols);
}

int dsht_xy_vgrad_w(dsht_plan* p, const double* w, double* xy, double* xy_lon, double* xy_lat, int ncols) {
  HOST_PROLOGUE(p, ncols);
  if (!w || (!xy && !xy_lon && !xy_lat)) return fail(DSHT_EARG, "NULL array");
  double *dg0, *dg1, *dg2, *dw;
  int rc;
  if ((rc = ctx_buf(c, 0, IJ * ncols, &dg0)) || (rc = ctx_buf(c, 2, IJ * ncols, &dg1)) ||
      (rc = ctx_buf(c, 3, IJ * ncols, &dg2)) || (rc = ctx_buf(c, 1, LL * ncols, &dw)))
    return rc;
  if ((rc = h2d(c, dw, w, LL * ncols))) return rc;
  if ((rc = dsht_inv_vgrad_d(p, dw, 1, (long)LL, xy ? dg0 : nullptr, xy_lon ? dg1 : nullptr, xy_lat ? dg2 : nullptr,
                             (long)IJ, ncols, nullptr, c->s)))
    return rc;
  if (xy) DSHT_HIP(hipMemcpyAsync(xy, dg0, sizeof(double) * IJ * ncols, hipMemcpyDeviceToHost, c->s));
  if (xy_lon) DSHT_HIP(hipMemcpyAsync(xy_lon, dg1, sizeof(double) * IJ * ncols, hipMemcpyDeviceToHost, c->s));
  if (xy_lat) DSHT_HIP(hipMemcpyAsync(xy_lat, dg2, sizeof(double) * IJ * ncols, hipMemcpyDeviceToHost, c->s));
  DSHT_HIP(hipStreamSynchronize(c->s));
  return 0;
}

int dsht_w_alpha_xy(dsht_plan* p, const double* a, const double* b, double* w, int ncols) {
  HOST_PROLOGUE(p, ncols);
  if (!a || !b || !w) return fail(DSHT_EARG, "NULL array");
  double *da, *db, *dw;
  int rc;
  if ((rc = ctx_buf(c, 0, IJ * ncols, &da)) || (rc = ctx_buf(c, 2, IJ * ncols, &db)) ||
      (rc = ctx_buf(c, 1, LL * ncols, &dw)))
    return rc;
  if ((rc = h2d(c, da, a, IJ * ncols)) || (rc = h2d(c, db, b, IJ * ncols))) return rc;
  if ((rc = dsht_alpha_fwd_d(p, da, db, (long)IJ, dw, 1, (long)LL, ncols, nullptr, c->s))) return rc;
  return d2h_sync(c, w, dw, LL * ncols);
}

int dsht_xy_alpha_w(dsht_plan* p, const double* wa, const double* wb, double* xy, int ncols) {
  HOST_PROLOGUE(p, ncols);
  if (!wa || !wb || !xy) return fail(DSHT_EARG, "NULL array");
  double *dg, *dwa, *dwb;
  int rc;
  if ((rc = ctx_buf(c, 0, IJ * ncols, &dg)) || (rc = ctx_buf(c, 1, LL * ncols, &dwa)) ||
      (rc = ctx_buf(c, 3, LL * ncols, &dwb)))
    return rc;
  if ((rc = h2d(c, dwa, wa, LL * ncols)) || (rc = h2d(c, dwb, wb, LL * ncols))) return rc;
  if ((rc = dsht_alpha_inv_d(p, dwa, dwb, 1, (long)LL, dg, (long)IJ, ncols, c->s))) return rc;
  return d2h_sync(c, xy, dg, IJ * ncols);
}

int dsht_uvcoslat2vordiv(dsht_plan* p, const double* ucos, const double* vcos, double* vor, double* div,
                         int ncols) {
  HOST_PROLOGUE(p, ncols);
  if (!ucos || !vcos || !vor || !div) return fail(DSHT_EARG, "NULL array");
  double *du, *dv, *dz, *dd;
  int rc;
  if ((rc = ctx_buf(c, 0, IJ * ncols, &du)) || (rc = ctx_buf(c, 2, IJ * ncols, &dv)) ||
      (rc = ctx_buf(c, 1, LL * ncols, &dz)) || (rc = ctx_buf(c, 3, LL * ncols, &dd)))
    return rc;
  if ((rc = h2d(c, du, ucos, IJ * ncols)) || (rc = h2d(c, dv, vcos, IJ * ncols))) return rc;
  if ((rc = dsht_uv2vordiv_d(p, du, dv, (long)IJ, dz, dd, 1, (long)LL, ncols, nullptr, c->s))) return rc;
  DSHT_HIP(hipMemcpyAsync(vor, dz, sizeof(double) * LL * ncols, hipMemcpyDeviceToHost, c->s));
  return d2h_sync(c, div, dd, LL * ncols);
}

int dsht_vordiv2uv(dsht_plan* p, const double* vor, const double* div, double* u, double* v, int ncols) {
  HOST_PROLOGUE(p, ncols);
  if (!u || !v || !vor || !div) return fail(DSHT_EARG, "NULL array");
  double *du, *dv, *dz, *dd;
  int rc;
  if ((rc = ctx_buf(c, 0, IJ * ncols, &du)) || (rc = ctx_buf(c, 2, IJ * ncols, &dv)) ||
      (rc = ctx_buf(c, 1, LL * ncols, &dz)) || (rc = ctx_buf(c, 3, LL * ncols, &dd)))
    return rc;
  if ((rc = h2d(c, dz, vor, LL * ncols)) || (rc = h2d(c, dd, div, LL * ncols))) return rc;
  if ((rc = dsht_vordiv2uv_d(p, dz, dd, 1, (long)LL, du, dv, (long)IJ, ncols, c->s))) return rc;
  DSHT_HIP(hipMemcpyAsync(u, du, sizeof(double) * IJ * ncols, hipMemcpyDeviceToHost, c->s));
  return d2h_sync(c, v, dv, IJ * ncols);
}

int dsht_w_diag_w(dsht_plan* p, const double* d, const double* w, double* out, int ncols) {
  HOST_PROLOGUE(p, ncols);
  if (!d || !w || !out) return fail(DSHT_EARG, "NULL array");
  double *dd, *dw;
  int rc;
  if ((rc = ctx_buf(c, 2, LL, &dd)) || (rc = ctx_buf(c, 1, LL * ncols, &dw))) return rc;
  if ((rc = h2d(c, dd, d, LL)) || (rc = h2d(c, dw, w, LL * ncols))) return rc;
  if ((rc = launch_diag(dd, dw, dw, 1, (long)LL, p->L, ncols, c->s))) return rc;
  return d2h_sync(c, out, dw, LL * ncols);
}

int dsht_xy_diag_xy(dsht_plan* p, const double* d, const double* xy, double* out, int ncols) {
  HOST_PROLOGUE(p, ncols);
  if (!d || !xy || !out) return fail(DSHT_EARG, "NULL array");
  double *dg, *dw, *dd;
  int rc;
  if ((rc = ctx_buf(c, 0, IJ * ncols, &dg)) || (rc = ctx_buf(c, 1, LL * ncols, &dw)) ||
      (rc = ctx_buf(c, 2, LL, &dd)))
    return rc;
  if ((rc = h2d(c, dd, d, LL)) || (rc = h2d(c, dg, xy, IJ * ncols))) return rc;
  if ((rc = dsht_fwd_d(p, DSHT_VAL, dg, (long)IJ, dw, 1, (long)LL, ncols, dd, c->s))) return rc;
  if ((rc = dsht_inv_d(p, DSHT_VAL, dw, 1, (long)LL, dg, (long)IJ, ncols, nullptr, c->s))) return rc;
  return d2h_sync(c, out, dg, IJ * ncols);
}

static int host_sum(dsht_plan* p, const double* xy, double* out, int ncols, int what) {
  HOST_PROLOGUE(p, ncols);
  if (!xy || !out) return fail(DSHT_EARG, "NULL array");
  double *dg, *dout;
  int rc;
  size_t nout = what >= 2 ? (size_t)p->J * ncols : (size_t)ncols;
  if ((rc = ctx_buf(c, 0, IJ * ncols, &dg)) || (rc = ctx_buf(c, 1, nout, &dout))) return rc;
  if ((rc = h2d(c, dg, xy, IJ * ncols))) return rc;
  switch (what) {
    case 0: rc = dsht_intlonlat_d(p, dg, (long)IJ, dout, ncols, c->s); break;
    case 1: rc = dsht_avrlonlat_d(p, dg, (long)IJ, dout, ncols, c->s); break;
    case 2: rc = dsht_ya_intlon_d(p, dg, (long)IJ, dout, ncols, c->s); break;
    default: rc = launch_lon_sum(p, dg, (long)IJ, ncols, dout, 1.0 / p->I, p->J, c->s); break;
  }
  if (rc) return rc;
  return d2h_sync(c, out, dout, nout);
}

int dsht_intlonlat(dsht_plan* p, const double* xy, double* out, int ncols) { return host_sum(p, xy, out, ncols, 0); }
int dsht_avrlonlat(dsht_plan* p, const double* xy, double* out, int ncols) { return host_sum(p, xy, out, ncols, 1); }
int dsht_ya_intlon(dsht_plan* p, const double* xy, double* out, int ncols) { return host_sum(p, xy, out, ncols, 2); }
int dsht_ya_avrlon(dsht_plan* p, const double* xy, double* out, int ncols) { return host_sum(p, xy, out, ncols, 3); }

int dsht_interpolate(dsht_plan* p, const double* w, const double* lon, const double* lat, double* out, int npts) {
  HOST_PROLOGUE(p, npts);
  if (!w || !lon || !lat || !out) return fail(DSHT_EARG, "NULL array");
  double *dw, *dl;
  int rc;
  if ((rc = ctx_buf(c, 1, LL, &dw)) || (rc = ctx_buf(c, 0, 3 * (size_t)npts, &dl))) return rc;
  if ((rc = h2d(c, dw, w, LL)) || (rc = h2d(c, dl, lon, npts)) || (rc = h2d(c, dl + npts, lat, npts))) return rc;
  if ((rc = launch_interpolate(p, dw, 1, dl, dl + npts, dl + 2 * npts, npts, c->s))) return rc;
  return d2h_sync(c, out, dl + 2 * npts, npts);
}

// ---- vertical column kernels on host arrays: the km layers of a field are km consecutive
// I x J grids (Fortran xyz(IM, JM, KS:KE)); staged through the calling thread's device buffers
static int col_stage(HostCtx* c, int i, const double* const* src, const size_t* n, int nsrc, double** dst) {
  size_t tot = 0;
  for (int k = 0; k < nsrc; k++) tot += n[k];
  if (int rc = ctx_buf(c, i, tot, dst)) return rc;
  size_t off = 0;
  for (int k = 0; k < nsrc; k++) {
    if (src[k])
      if (int rc = h2d(c, *dst + off, src[k], n[k])) return rc;
    off += n[k];
  }
  return 0;
}

int dsht_col_omega(dsht_plan* p, const double* div, const double* h, int km, const double* cdk, double* omg) {
  HOST_PROLOGUE(p, km);
  if (!div || !h || !cdk || !omg) return fail(DSHT_EARG, "NULL array");
  double *din, *dout, *dk;
  const double* in[] = {div, h};
  const size_t nin[] = {IJ * km, IJ * km};
  const size_t nk[] = {(size_t)km};
  int rc;
  if ((rc = col_stage(c, 0, in, nin, 2, &din)) || (rc = col_stage(c, 2, &cdk, nk, 1, &dk)) ||
      (rc = ctx_buf(c, 1, IJ * (km + 1), &dout)))
    return rc;
  if ((rc = dsht_col_omega_d(p, din, din + IJ * km, (long)IJ, km, dk, dout, (long)IJ, c->s))) return rc;
  return d2h_sync(c, omg, dout, IJ * (km + 1));
}

int dsht_col_hydpres(dsht_plan* p, const double* dens, const double* h, int km, const double* cdk, double grav,
                     double* out) {
  HOST_PROLOGUE(p, km);
  if (!dens || !h || !cdk || !out) return fail(DSHT_EARG, "NULL array");
  double *din, *dout, *dk;
  const double* in[] = {dens, h};
  const size_t nin[] = {IJ * km, IJ * km};
  const size_t nk[] = {(size_t)km};
  int rc;
  if ((rc = col_stage(c, 0, in, nin, 2, &din)) || (rc = col_stage(c, 2, &cdk, nk, 1, &dk)) ||
      (rc = ctx_buf(c, 1, IJ * km, &dout)))
    return rc;
  if ((rc = dsht_col_hydpres_d(p, din, din + IJ * km, (long)IJ, km, dk, grav, dout, (long)IJ, c->s))) return rc;
  return d2h_sync(c, out, dout, IJ * km);
}

int dsht_col_uvbarot(dsht_plan* p, const double* u, const double* v, const double* h, int km, const double* cdk,
                     double* ub, double* vb) {
  HOST_PROLOGUE(p, km);
  if (!u || !v || !h || !cdk || !ub || !vb) return fail(DSHT_EARG, "NULL array");
  double *din, *dout, *dk;
  const double* in[] = {u, v, h};
  const size_t nin[] = {IJ * km, IJ * km, IJ * km};
  const size_t nk[] = {(size_t)km};
  int rc;
  if ((rc = col_stage(c, 0, in, nin, 3, &din)) || (rc = col_stage(c, 2, &cdk, nk, 1, &dk)) ||
      (rc = ctx_buf(c, 1, 2 * IJ, &dout)))
    return rc;
  if ((rc = dsht_col_uvbarot_d(p, din, din + IJ * km, din + 2 * IJ * km, (long)IJ, km, dk, dout, dout + IJ, c->s)))
    return rc;
  DSHT_HIP(hipMemcpyAsync(ub, dout, sizeof(double) * IJ, hipMemcpyDeviceToHost, c->s));
  return d2h_sync(c, vb, dout + IJ, IJ);
}

int dsht_col_quick(dsht_plan* p, const double* omg, const double* trc, int km, const double* cdk, const double* rfdk,
                   int linfree, double* flx) {
  HOST_PROLOGUE(p, km);
  if (!omg || !trc || !cdk || !rfdk || !flx) return fail(DSHT_EARG, "NULL array");
  double *din, *dout, *dk;
  const double* in[] = {omg, trc};
  const size_t nin[] = {IJ * (km + 1), IJ * km};
  const double* kin[] = {cdk, rfdk};
  const size_t nk[] = {(size_t)km, (size_t)km + 1};
  int rc;
  if ((rc = col_stage(c, 0, in, nin, 2, &din)) || (rc = col_stage(c, 2, kin, nk, 2, &dk)) ||
      (rc = ctx_buf(c, 1, IJ * (km + 1), &dout)))
    return rc;
  if ((rc = dsht_col_quick_d(p, din, (long)IJ, din + IJ * (km + 1), (long)IJ, km, dk, dk + km, linfree, dout, (long)IJ,
                             c->s)))
    return rc;
  return d2h_sync(c, flx, dout, IJ * (km + 1));
}

int dsht_col_vdiff(dsht_plan* p, const double* q0, const double* rhs, const double* kv, const double* e3, int km,
                   const double* cdk, const double* fdk, double dt, double alpha, int sfc_bc, int btm_bc,
                   const double* bfric, double* qa) {
  HOST_PROLOGUE(p, km);
  if (!q0 || !rhs || !kv || !e3 || !cdk || !fdk || !qa) return fail(DSHT_EARG, "NULL array");
  double *din, *dout, *dk, *dw;
  const double* in[] = {q0, rhs, kv, e3};
  const size_t nin[] = {IJ * km, IJ * km, IJ * km, IJ * km};
  const double* kin[] = {cdk, fdk, bfric};
  const size_t nk[] = {(size_t)km, (size_t)km + 1, bfric ? IJ : 0};
  int rc;
  if ((rc = col_stage(c, 0, in, nin, 4, &din)) || (rc = col_stage(c, 2, kin, nk, 3, &dk)) ||
      (rc = ctx_buf(c, 1, IJ * km, &dout)) || (rc = ctx_buf(c, 3, 3 * IJ * km, &dw)))
    return rc;
  const double* dq = din;
  const double* dr = din + IJ * km;
  const double* dkv = din + 2 * IJ * km;
  const double* de = din + 3 * IJ * km;
  const long lw = (long)(3 * IJ * km);
  rc = bfric ? dsht_col_vdiff_mom_d(p, dq, dr, dkv, de, (long)IJ, km, dk, dk + km, dt, alpha, sfc_bc, btm_bc,
                                    dk + 2 * km + 1, dout, (long)IJ, dw, lw, c->s)
             : dsht_col_vdiff_d(p, dq, dr, dkv, de, (long)IJ, km, dk, dk + km, dt, alpha, sfc_bc, btm_bc, dout,
                                (long)IJ, dw, lw, c->s);
  if (rc) return rc;
  return d2h_sync(c, qa, dout, IJ * km);
}

// ---------------------------------------------------------------- Redi/GM mixing
// (LPhys_RediGM_hspm_vfvm_AddMixingTerm, LPhys_RediGM_hspm_vfvm_mod.f90:218-372)
int dsht_redigm_default_config(dsht_redigm_config* c) {
  if (!c) return fail(DSHT_EARG, "NULL config");
  c->kappa_redi = 800.0;
  c->kappa_gm = 800.0;
  c->slope_max = 4e-3;
  c->sd = 1e-3;
  c->interior_taper = DSHT_TAPER_DM95;
  c->pbl_taper = DSHT_TAPER_LDD95;
  c->omega = 2.0 * 3.1415926535897932 / (60.0 * 60.0 * 23.9345);
  c->ref_dens = 1027.0;
  c->grav = 9.79764;
  return 0;
}

static int redigm_params(const dsht_redigm_config* c, RediParams* r) {
  if (!c) return fail(DSHT_EARG, "NULL config");
  if (c->interior_taper != DSHT_TAPER_DM95 && c->interior_taper != DSHT_TAPER_GKW91)
    return fail(DSHT_EARG, "interior tapering must be DSHT_TAPER_DM95 or DSHT_TAPER_GKW91");
  if (c->pbl_taper != DSHT_TAPER_LDD95)
    return fail(DSHT_EUNSUPPORTED,
                "boundary-layer tapering must be DSHT_TAPER_LDD95 (with the linear form the reference "
                "multiplies the slopes by an undefined factor)");
  r->kappa_redi = c->kappa_redi;
  r->kappa_gm = c->kappa_gm;
  r->slope_max = c->slope_max;
  r->sd = c->sd;
  r->omega = c->omega;
  r->ref_dens = c->ref_dens;
  r->grav = c->grav;
  r->gkw91 = c->interior_taper == DSHT_TAPER_GKW91;
  return 0;
}

// checks shared by the column part and the whole tendency; ldgr: pitch of the gradients
static int redigm_check(dsht_plan* p, const dsht_redigm_config* cfg, const double* ptemp, const double* salt,
                        const double* h, const double* z, long ldg, int km, const double* topo, const double* cdk,
                        const double* rfdk, RediParams* r) {
  if (int rc = check_col(p, km, cdk, ldg)) return rc;
  if (km < 2) return fail(DSHT_EARG, "the Redi/GM mixing needs km >= 2");
  if (km > 1024) return fail(DSHT_EARG, "the Redi/GM mixing supports at most 1024 layers");
  if (!ptemp || !salt || !h || !z || !topo || !rfdk) return fail(DSHT_EARG, "pointer is NULL");
  return redigm_params(cfg, r);
}

int dsht_col_redigm_d(dsht_plan* p, const dsht_redigm_config* cfg, const double* ptemp, const double* salt,
                      const double* h, const double* z, long ldg, int km, const double* topo, const double* cdk,
                      const double* rfdk, const double* grad_lon, const double* grad_lat, double* flx_lon,
                      double* flx_lat, double* vdiv, long ldo, void* stream) {
  RediParams r;
  if (int rc = redigm_check(p, cfg, ptemp, salt, h, z, ldg, km, topo, cdk, rfdk, &r)) return rc;
  if (!grad_lon || !grad_lat || !flx_lon || !flx_lat || !vdiv) return fail(DSHT_EARG, "pointer is NULL");
  if (ldo < (long)p->I * p->J) return fail(DSHT_EARG, "output column pitch %ld must be >= I*J", ldo);
  const long IJ = (long)p->I * p->J;
  const long nl = (km - 1) * ldg + IJ, n2 = (2l * km - 1) * ldg + IJ, no = (2l * km - 1) * ldo + IJ;
  const Span ins[] = {span_of(ptemp, nl),   span_of(salt, nl),     span_of(h, nl),   span_of(z, nl),
                      span_of(topo, IJ),    span_of(cdk, km),      span_of(rfdk, km + 1),
                      span_of(grad_lon, n2), span_of(grad_lat, n2)};
  static const char* const names[] = {"ptemp", "salt", "h", "z", "topo", "cdk", "rfdk", "grad_lon", "grad_lat"};
  const Span outs[] = {span_of(flx_lon, no), span_of(flx_lat, no), span_of(vdiv, no)};
  static const char* const onames[] = {"flx_lon", "flx_lat", "vdiv"};
  for (int i = 0; i < 3; i++) {
    if (int rc = check_disjoint(onames[i], outs[i], ins, names, 9)) return rc;
    for (int k = i + 1; k < 3; k++)
      if (overlaps(outs[i], outs[k])) return fail(DSHT_EARG, "the outputs %s and %s overlap", onames[i], onames[k]);
  }
  return launch_redigm_col(p, r, ptemp, salt, h, z, ldg, km, topo, cdk, rfdk, grad_lon, grad_lat, ldg, flx_lon,
                           flx_lat, vdiv, ldo, (hipStream_t)stream);
}

long dsht_redigm_work_size(const dsht_plan* p, int km) {
  if (check_plan(p)) return -1;
  if (km < 2) {
    fail(DSHT_EARG, "the Redi/GM mixing needs km >= 2");
    return -1;
  }
  return 10l * km * p->I * p->J + 2l * km * p->L;
}

int dsht_redigm_d(dsht_plan* p, const dsht_redigm_config* cfg, const double* ptemp, const double* salt,
                  const double* h, const double* z, long ldg, int km, const double* topo, const double* cdk,
                  const double* rfdk, double* ptemp_rhs, double* salt_rhs, long ldo, double* work, long lwork,
                  void* stream) {
  RediParams r;
  if (int rc = redigm_check(p, cfg, ptemp, salt, h, z, ldg, km, topo, cdk, rfdk, &r)) return rc;
  if (!ptemp_rhs || !salt_rhs || !work) return fail(DSHT_EARG, "pointer is NULL");
  if (ldo < (long)p->I * p->J) return fail(DSHT_EARG, "output column pitch %ld must be >= I*J", ldo);
  const long need = dsht_redigm_work_size(p, km);
  if (lwork < need) return fail(DSHT_EARG, "workspace of %ld doubles is below %ld", lwork, need);
  if (((uintptr_t)work) % 16) return fail(DSHT_EARG, "the workspace must be 16-byte aligned");
  const long IJ = (long)p->I * p->J, nl = (km - 1) * ldg + IJ, no = (km - 1) * ldo + IJ;
  const Span ins[] = {span_of(ptemp, nl), span_of(salt, nl),  span_of(h, nl),       span_of(z, nl),
                      span_of(topo, IJ),  span_of(cdk, km),   span_of(rfdk, km + 1)};
  static const char* const names[] = {"ptemp", "salt", "h", "z", "topo", "cdk", "rfdk"};
  const Span out_t = span_of(ptemp_rhs, no), out_s = span_of(salt_rhs, no), out_w = span_of(work, lwork);
  if (int rc = check_disjoint("ptemp_rhs", out_t, ins, names, 7)) return rc;
  if (int rc = check_disjoint("salt_rhs", out_s, ins, names, 7)) return rc;
  if (int rc = check_disjoint("workspace", out_w, ins, names, 7)) return rc;
  if (overlaps(out_t, out_s) || overlaps(out_t, out_w) || overlaps(out_s, out_w))
    return fail(DSHT_EARG, "ptemp_rhs, salt_rhs and the workspace must not overlap");
  hipStream_t s = (hipStream_t)stream;
  const long n2 = 2l * km;
  double* glon = work;  // 2 km grids each, pitch I*J (temperature layers, then salinity)
  double* glat = glon + n2 * IJ;
  double* flon = glat + n2 * IJ;
  double* flat = flon + n2 * IJ;
  double* vdv = flat + n2 * IJ;
  double* w = vdv + n2 * IJ;  // 2 km spectra, coefficient-major (sl = 2 km)
  int rc;
  // calc_GradTRC (:431-442): w_xy of each layer, xy_GradLon_w and xy_GradLat_w
  if (salt == ptemp + (long)km * ldg) {  // both tracers in one array (xyza_TRC): one call
    if ((rc = dsht_fwd_d(p, DSHT_VAL, ptemp, ldg, w, n2, 1, (int)n2, nullptr, s))) return rc;
  } else {
    if ((rc = dsht_fwd_d(p, DSHT_VAL, ptemp, ldg, w, n2, 1, km, nullptr, s))) return rc;
    if ((rc = dsht_fwd_d(p, DSHT_VAL, salt, ldg, w + km, n2, 1, km, nullptr, s))) return rc;
  }
  if ((rc = dsht_inv_vgrad_d(p, w, n2, 1, nullptr, glon, glat, IJ, (int)n2, nullptr, s))) return rc;
  if ((rc = launch_redigm_col(p, r, ptemp, salt, h, z, ldg, km, topo, cdk, rfdk, glon, glat, IJ, flon, flat, vdv,
                              IJ, s)))
    return rc;
  // append_Redi_GM_RHS (:340-346): xy_w(w_AlphaOptr_xy(FLon cos, FLat cos)) into glon
  if ((rc = dsht_alpha_fwd_d(p, flon, flat, IJ, w, n2, 1, (int)n2, nullptr, s))) return rc;
  if ((rc = dsht_inv_d(p, DSHT_VAL, w, n2, 1, glon, IJ, (int)n2, nullptr, s))) return rc;
  return launch_redigm_rhs(p, ptemp_rhs, salt_rhs, ldo, glon, vdv, IJ, km, s);
}

int dsht_redigm(dsht_plan* p, const dsht_redigm_config* cfg, const double* ptemp, const double* salt,
                const double* h, const double* z, int km, const double* topo, const double* cdk, const double* rfdk,
                double* ptemp_rhs, double* salt_rhs) {
  HOST_PROLOGUE(p, km);
  if (!ptemp || !salt || !h || !z || !topo || !cdk || !rfdk || !ptemp_rhs || !salt_rhs)
    return fail(DSHT_EARG, "NULL array");
  const long lw = dsht_redigm_work_size(p, km);
  if (lw < 0) return DSHT_EARG;
  double *din, *drhs, *dk, *dw;
  const double* in[] = {ptemp, salt, h, z, topo};
  const size_t nin[] = {IJ * km, IJ * km, IJ * km, IJ * km, IJ};
  const double* rin[] = {ptemp_rhs, salt_rhs};
  const size_t nr[] = {IJ * km, IJ * km};
  const double* kin[] = {cdk, rfdk};
  const size_t nk[] = {(size_t)km, (size_t)km + 1};
  int rc;
  if ((rc = col_stage(c, 0, in, nin, 5, &din)) || (rc = col_stage(c, 1, rin, nr, 2, &drhs)) ||
      (rc = col_stage(c, 2, kin, nk, 2, &dk)) || (rc = ctx_buf(c, 3, (size_t)lw, &dw)))
    return rc;
  if ((rc = dsht_redigm_d(p, cfg, din, din + IJ * km, din + 2 * IJ * km, din + 3 * IJ * km, (long)IJ, km,
                          din + 4 * IJ * km, dk, dk + km, drhs, drhs + IJ * km, (long)IJ, dw, lw, c->s)))
    return rc;
  DSHT_HIP(hipMemcpyAsync(ptemp_rhs, drhs, sizeof(double) * IJ * km, hipMemcpyDeviceToHost, c->s));
  return d2h_sync(c, salt_rhs, drhs + IJ * km, IJ * km);
}

int dsht_spec_energy(dsht_plan* p, int kind, const double* wa, int ncols, double* nma, double* na) {
  HOST_PROLOGUE(p, ncols);
  if (!wa || (!nma && !na)) return fail(DSHT_EARG, "NULL array");
  const size_t n1 = (size_t)p->N + 1, nm = n1 * (2 * (size_t)p->N + 1) * ncols, nn = n1 * ncols;
  double *dw, *dout;
  int rc;
  if ((rc = ctx_buf(c, 1, LL * ncols, &dw)) || (rc = ctx_buf(c, 0, nm + nn, &dout))) return rc;
  if ((rc = h2d(c, dw, wa, LL * ncols))) return rc;
  if ((rc = dsht_spec_energy_d(p, kind, dw, 1, (long)LL, ncols, nma ? dout : nullptr, na ? dout + nm : nullptr,
                               c->s)))
    return rc;
  if (nma && na) DSHT_HIP(hipMemcpyAsync(nma, dout, sizeof(double) * nm, hipMemcpyDeviceToHost, c->s));
  return nma && !na ? d2h_sync(c, nma, dout, nm) : d2h_sync(c, na, dout + nm, nn);
}

int dsht_coef(const dsht_plan* pc, int which, const double* params, int nparams, double* out) {
  dsht_plan* p = const_cast<dsht_plan*>(pc);
  HOST_PROLOGUE(p, 1);
  if (!out) return fail(DSHT_EARG, "NULL array");
  double* dd;
  int rc;
  if ((rc = ctx_buf(c, 1, LL, &dd))) return rc;
  if ((rc = dsht_coef_d(p, which, params, nparams, dd, c->s))) return rc;
  return d2h_sync(c, out, dd, LL);
}

}  // extern "C"
