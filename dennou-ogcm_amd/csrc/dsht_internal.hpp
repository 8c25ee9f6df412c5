// dsht_internal.hpp -- shared definitions of libdsht (MI355X spherical-harmonic transforms).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dsht.h"

namespace dsht {

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------- geometry
// K-chunk of the forward contraction (Gauss latitudes of one hemisphere) and of
// the inverse contraction (total wavenumbers of one parity).
constexpr int KCH = 16;
constexpr int TILE = 16;  // MFMA output tile edge

__host__ __device__ inline int cdiv(int a, int b) { return (a + b - 1) / b; }
__host__ __device__ inline long cdivl(long a, long b) { return (a + b - 1) / b; }

// number of total wavenumbers n in [m, N] with (n - m) of parity p
__host__ __device__ inline int n_par(int N, int m, int p) {
  int cnt = N - m + 1;
  return p == 0 ? (cnt + 1) / 2 : cnt / 2;
}

// ---------------------------------------------------------------- ops
enum FwdOp { F_VAL = 0, F_DIVLAMBDA = 1, F_DIVMU = 2, F_ALPHA = 3, F_UV2VORDIV = 4 };
// I_VGRAD: value, GradLon and GradLat of one field from one contraction pass (P and H
// share the data; GradLon is i m times the P result)
enum InvOp { I_VAL = 0, I_GRADLAMBDA = 1, I_GRADMU = 2, I_ALPHA = 3, I_VORDIV2UV = 4, I_VGRAD = 5 };

// inputs, outputs and tables of each operator
template <int OP>
struct FwdTraits {
  static constexpr int NIN = (OP == F_ALPHA || OP == F_UV2VORDIV) ? 2 : 1;
  static constexpr int NOUT = (OP == F_UV2VORDIV) ? 2 : 1;
  static constexpr bool USE_P = (OP != F_DIVMU);
  static constexpr bool USE_H = (OP == F_DIVMU || OP == F_ALPHA || OP == F_UV2VORDIV);
};

template <int OP>
struct InvTraits {
  static constexpr int NIN = (OP == I_ALPHA || OP == I_VORDIV2UV) ? 2 : 1;
  static constexpr int NOUT = (OP == I_VORDIV2UV) ? 2 : (OP == I_VGRAD ? 3 : 1);
  static constexpr bool USE_P = (OP != I_GRADMU);
  static constexpr bool USE_H = (OP == I_GRADMU || OP == I_ALPHA || OP == I_VORDIV2UV || OP == I_VGRAD);
};
// runtime counterparts (host side)
inline int inv_nin(int op) { return (op == I_ALPHA || op == I_VORDIV2UV) ? 2 : 1; }
inline int inv_nout(int op) { return op == I_VORDIV2UV ? 2 : (op == I_VGRAD ? 3 : 1); }


// Arguments of the forward Legendre contraction (Fourier -> spectral).
struct LegFwdArgs {
  const double2* G0;  // Fourier coefficients [m][j*gld + c] (complex), input 0
  const double2* G1;  // input 1 (two-input ops)
  double* W0;         // spectral output 0 (element (l, c) at l*sl + c*sc)
  double* W1;         // spectral output 1
  long sl, sc;
  int ncols;          // columns in this call
  long gld;           // Fourier row pitch (columns per latitude row in G)
  long gmst;          // Fourier m-stride (J * gld)
  const double* tabP; // forward-swizzled P table
  const double* tabH; // forward-swizzled H table
  const long* offF;   // per-m offsets into the forward tables
  const double* cw;   // per-latitude weights of the north half (J2pad)
  const double* diag0;  // per-l epilogue multipliers (nullable)
  const double* diag1;
  double scale;       // extra scalar factor (1/R for the H operator)
  int N, J2, J2pad;
};

// Arguments of the inverse Legendre contraction (spectral -> Fourier).
struct LegInvArgs {
  const double* W0;  // spectral input 0
  const double* W1;  // spectral input 1
  long sl, sc;
  double2* G0;       // Fourier output 0 [m][j*gld + c]
  double2* G1;       // Fourier output 1
  double2* G2 = nullptr;  // Fourier output 2 (I_VGRAD); outputs of I_VGRAD may be NULL (skipped)
  int ncols;
  long gld, gmst;
  const double* tabP;  // inverse-swizzled P table
  const double* tabH;  // inverse-swizzled H table
  const long* offI;    // per-m offsets into the inverse tables
  const double* rs;    // per-latitude output row scale (J), nullable
  const double* pro;   // per-l prologue multipliers applied to both inputs (nullable)
  int N, J, J2, J2pad;
};

// ---------------------------------------------------------------- FFT plan
struct FftPlan {
  int M = 0;          // complex length I/2
  int nrad = 0;
  int rad[24] = {0};
  double2* twM = nullptr;  // exp(-2 pi i k / M), k < M  (device)
  double2* twI = nullptr;  // exp(-2 pi i m / I), m < M  (device)
  float2* twM32 = nullptr;
  float2* twI32 = nullptr;
};

struct FftArgs {
  int I, M, N, J;
  int ncols;
  long ldg;       // grid column pitch (elements)
  long gld, gmst; // Fourier layout: [m][j*gld + c]
  int nrad;
  int rad[24];
  const double2* twM;
  const double2* twI;
  int rows_per_block;
  const double* rowscale;  // optional per-latitude scale applied after the inverse (nullable)
  int threads;             // workgroup size of the register kernels (128 or 256)
  int fp32;                // FP32 arithmetic (FP32 Fourier path of the split forms)
  int order;               // row-group order of the register kernels (0 packed, 1 latitude-major)
  int deriv;               // inverse: transform i m X_m (longitude derivative) instead of X_m
  void* xy2;               // inverse, dual form: the field into xy and its longitude derivative
                           // (times rowscale) into xy2 from one read of the coefficients (nullable)
};

// ---------------------------------------------------------------- workspace
struct Workspace {
  void* buf = nullptr;
  size_t bytes = 0;
  std::mutex mu;
};

struct HostCtx;  // per-thread host-API context (api.hip)

// a locked scratch buffer of the (plan, stream) pair; unlocks when destroyed
struct WsHandle {
  void* buf = nullptr;
  Workspace* ws = nullptr;
  std::unique_lock<std::mutex> lk;
};

// ---------------------------------------------------------------- plan
}  // namespace dsht

struct dsht_plan {
  int N = 0, I = 0, J = 0, K = 0, L = 0, J2 = 0, J2pad = 0;
  double R = 1.0;
  int device = 0;
  int prec = DSHT_PREC_FP64;
  // host copies of the grid (computed on the device)
  std::vector<double> mu, wq, wl, lat, lon;
  // device grid vectors
  double* d_mu = nullptr;
  double* d_wq = nullptr;     // quadrature weights, sum 1 (J)
  double* d_wl = nullptr;     // y_Lat_Weight, sum 2 (J)
  double* d_cwv = nullptr;    // north-half weights wq (J2pad)
  double* d_cwd = nullptr;    // north-half weights wq/(1-mu^2) (J2pad)
  double* d_rs_icos = nullptr;    // 1/cos(phi)       (J)
  double* d_rs_alpha = nullptr;   // 1/(R cos^2 phi)  (J)
  double* d_rs_uv = nullptr;      // 1/(R cos phi)    (J)
  double* d_pro_psi = nullptr;    // R^2 * LaplaInv = -R^2/(n(n+1)), 0 at n=0 (L)
  double* d_lat_cs = nullptr;     // cos(lat_j) (J), then |sin(lat_j)| (J), of the host lat
  // Legendre tables (FP64), swizzled into MFMA fragment order
  double* d_tabFP = nullptr;
  double* d_tabFH = nullptr;
  double* d_tabIP = nullptr;
  double* d_tabIH = nullptr;
  long* d_offF = nullptr;
  long* d_offI = nullptr;
  std::vector<long> offF, offI;
  long sizeF = 0, sizeI = 0;
  dsht::FftPlan fft;
  // workspaces keyed by stream
  std::mutex ws_mu;
  std::map<void*, std::unique_ptr<dsht::Workspace>> ws;
  // mixed-precision tables (built lazily)
  void* mixed = nullptr;
  int chunk_cols = 0;  // column chunk of the device path (0 = auto)
  int leg_cb = 0;      // column block of the contractions (0 = auto, 16, 32)
  int fft_kind = 0;    // 0 = register two-pass FFT (256-thread blocks; FP32 arithmetic in the
                       // split forms' inverse FFT), 1 = LDS Stockham, 2 = register FFT in
                       // 128-thread blocks, 3 = register FFT with FP64 arithmetic in every form
  int fft_order = 1;   // register FFT row-group order: 0 packed, 1 latitude-major (default)
  int leg_kind = 0;    // contractions.  FP64: 0 = LDS-staged inverse, fragment-stream forward;
                       // 1/3 = LDS-staged; 2 = per-wave fragment streams.
                       // Split-form inverse: 0 = per form (whole-K LDS staging for BF16X6 and
                       // FP16X3, packed data panels otherwise), 2 = packed panels, 3 = whole-K
};

namespace dsht {
// Columns per chunk of one device call.  The Fourier scratch of one output stays within
// CHUNK_FOURIER_BYTES, and a call that needs several chunks gets equal ones (rounded up to
// the 16-column MFMA tile).  Whole-call chunks measure faster than MALL-sized ones: at
// T170L60 the FP64 step takes 0.84 ms with 512 MiB chunks against 1.0 ms with 48 MiB
// chunks (fewer, larger FFT and contraction launches).
constexpr long CHUNK_FOURIER_BYTES = 512l << 20;
// Fourier row pitch of a chunk of nc columns: a multiple of the 16-column group, so every
// 16-column run of a latitude row starts on a 128-byte line (FP32) or 256-byte pair of lines
inline long fourier_pitch(int nc) { return ((long)nc + 15) / 16 * 16; }
inline int chunk_columns(const dsht_plan* p, int ncols, long fourier_elem_bytes) {
  if (p->chunk_cols > 0) return p->chunk_cols < ncols ? p->chunk_cols : ncols;
  const long per = (long)(p->N + 1) * p->J * fourier_elem_bytes;
  long cap = CHUNK_FOURIER_BYTES / per;
  cap = cap < 16 ? 16 : cap / 16 * 16;
  if (ncols <= cap) return ncols;
  const long nch = (ncols + cap - 1) / cap;
  const long c = ((ncols + nch - 1) / nch + 15) / 16 * 16;
  return (int)(c < ncols ? c : ncols);
}

// error handling
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);
#define DSHT_HIP(call)                                                            \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return ::dsht::fail(DSHT_EDEVICE, "%s:%d %s: %s", __FILE__, __LINE__, #call, \
                          hipGetErrorString(e_));                                 \
  } while (0)

// api.hip: per-stream scratch
int workspace_acquire(dsht_plan* p, hipStream_t s, size_t bytes, WsHandle& out);
// tables.hip
int build_tables(dsht_plan* p);
// fft.hip
int fft_plan_init(dsht_plan* p);
void fft_plan_free(dsht_plan* p);
int launch_fft_fwd(const dsht_plan* p, const double* xy, long ldg, int ncols, double2* G, long gld,
                   hipStream_t s);
// rowscale: per-latitude factor on the output (nullable); deriv: transform i m X_m (the
// longitude derivative of the field), register kernels only (fft_inv_deriv_ok)
bool fft_inv_deriv_ok(const dsht_plan* p);
// xy2 (dual form, register kernels only): xy receives the field and xy2 its longitude
// derivative times rowscale, from one read of the coefficients
int launch_fft_inv(const dsht_plan* p, const double2* G, long gld, int ncols, double* xy, long ldg,
                   hipStream_t s, const double* rowscale = nullptr, int deriv = 0, double* xy2 = nullptr);
// FP32 Fourier intermediates of the split-precision forms; colmax (nullable) receives the
// per-column max |G| as float bits
bool reg_fft_supported(int M);
template <typename GT, typename XT>
int reg_fft_fwd(const XT* xy, const FftArgs& a, GT* G, unsigned* colmax, hipStream_t s);
template <typename GT, typename XT>
int reg_fft_inv(const GT* G, const FftArgs& a, XT* xy, hipStream_t s);
// legendre_f64d.hip: FP64 contractions as per-wave fragment streams (DSHT_TUNE_LEG = 2)
int launch_leg_inv_direct(const dsht_plan* p, int op, const LegInvArgs& a, hipStream_t s);
int launch_leg_fwd_direct(const dsht_plan* p, int op, const LegFwdArgs& a, hipStream_t s);
int launch_fft_fwd32(const dsht_plan* p, const double* xy, long ldg, int ncols, float2* G, long gld,
                     unsigned* colmax, hipStream_t s);
int launch_fft_inv32(const dsht_plan* p, const float2* G, long gld, int ncols, double* xy, long ldg,
                     hipStream_t s, const double* rowscale = nullptr, int deriv = 0, double* xy2 = nullptr);
// FP32 grids (the split forms' FP32 grid I/O; register kernels only)
int launch_fft_fwd32f(const dsht_plan* p, const float* xy, long ldg, int ncols, float2* G, long gld,
                      unsigned* colmax, hipStream_t s);
int launch_fft_inv32f(const dsht_plan* p, const float2* G, long gld, int ncols, float* xy, long ldg,
                      hipStream_t s, const double* rowscale = nullptr, int deriv = 0, float* xy2 = nullptr);
// legendre_f64.hip
int launch_leg_fwd(const dsht_plan* p, int op, const LegFwdArgs& a, hipStream_t s);
int launch_leg_inv(const dsht_plan* p, int op, const LegInvArgs& a, hipStream_t s);
// ops.hip
int launch_diag(const double* d, const double* w, double* out, long sl, long sc, int L, int ncols,
                hipStream_t s);
// vertical column diagnostics (column.hip)
int launch_col_omega(const double* div, const double* h, long ldg, int km, const double* cdk, double* omg, long ldo,
                     int npts, hipStream_t s);
int launch_col_hydpres(const double* dens, const double* h, long ldg, int km, const double* cdk, double grav,
                       double* p, long ldo, int npts, hipStream_t s);
int launch_col_uvbarot(const double* u, const double* v, const double* h, long ldg, int km, const double* cdk,
                       double* ub, double* vb, int npts, hipStream_t s);
int launch_col_quick(const double* omg, long ldo, const double* trc, long ldg, int km, const double* cdk,
                     const double* rfdk, int linfree, double* flx, long ldf, int npts, hipStream_t s);
int launch_col_vdiff(const double* q0, const double* rhs, const double* kv, const double* e3, long ldg, int km,
                     const double* cdk, const double* fdk, double dt, double alpha, int sfc_dirichlet,
                     int btm_dirichlet, const double* bfric, double* qa, long ldo, double* work, int npts,
                     hipStream_t s);
// redigm.hip: the grid-space part of the Redi/GM mixing and the tracer update
struct RediParams {
  double kappa_redi, kappa_gm, slope_max, sd, omega, ref_dens, grav;
  int gkw91;  // interior tapering GKW91 (else DM95)
};
int launch_redigm_col(const dsht_plan* p, const RediParams& c, const double* pt, const double* sa, const double* h,
                      const double* z, long ldg, int km, const double* topo, const double* cdk, const double* rfdk,
                      const double* glon, const double* glat, long ldgr, double* flon, double* flat, double* vdiv,
                      long ldo, hipStream_t s);
int launch_redigm_rhs(const dsht_plan* p, double* rt, double* rs, long ldo, const double* hdiv, const double* vdiv,
                      long ldw, int km, hipStream_t s);
int launch_spec_energy(const dsht_plan* p, int ens, const double* w, long sl, long sc, int ncols, double* nma,
                       double* na, hipStream_t s);
int launch_lonlat_sum(const dsht_plan* p, const double* xy, long ldg, int ncols, double* out,
                      double factor, hipStream_t s);
int launch_lon_sum(const dsht_plan* p, const double* xy, long ldg, int ncols, double* out,
                   double factor, long out_ld, hipStream_t s);
int launch_interpolate(const dsht_plan* p, const double* w, long sl, const double* lon,
                       const double* lat, double* out, int npts, hipStream_t s);
int launch_coef(const dsht_plan* p, int which, const double* params, int nparams, double* out,
                hipStream_t s);
int launch_transpose(const double* in, long ld_in, double* out, long ld_out, int rows, int cols,
                     hipStream_t s);
// api.hip: kernel timers (HIP events), enabled by dsht_prof_enable
struct ProfScope {
  int slot = -1;
  hipStream_t s = nullptr;
  ProfScope(const char* name, hipStream_t s, double bytes, double flops);
  ~ProfScope();
};
// legendre_mixed.hip: split-precision forms
void mixed_free(dsht_plan* p);
// grid_f32: the grid pointers (in0/in1 forward, out0-out2 inverse) hold float
int mixed_run_fwd(dsht_plan* p, int op, const double* in0, const double* in1, long ldg, double* out0,
                  double* out1, long sl, long sc, int ncols, const double* diag0, const double* diag1,
                  double scale, const double* cw, hipStream_t s, bool grid_f32 = false);
// outputs of I_VGRAD: out0 value, out1 GradLon, out2 GradLat (any may be NULL)
int mixed_run_inv(dsht_plan* p, int op, const double* in0, const double* in1, long sl, long sc,
                  double* out0, double* out1, long ldg, int ncols, const double* pro, const double* rs,
                  hipStream_t s, double* out2 = nullptr, bool grid_f32 = false);
}  // namespace dsht
