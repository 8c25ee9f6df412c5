// legendre_mixed.hip -- split-precision Legendre contractions on CDNA4 16x16x32 MFMA.
//
// Same operators as legendre_f64.hip (the Legendre half of SPMODEL w_xy / xy_w and their
// derivative forms, doc/description/ogcm/space_discretization.tex:199-268), with the
// contraction run in 16-bit MFMA with FP32 accumulation:
//   BF16X3   x = x1 + x2 (bf16), products x1y1 + x1y2 + x2y1
//   BF16X6   x = x1 + x2 + x3 (bf16), products of total order <= 2 (six)
//   FP16     x = x1 (fp16), per-m power-of-two table scale, per-column data scale
//   FP16X3   x = x1 + x2 (fp16), scaled as FP16, products x1y1 + x1y2 + x2y1
// The associated-Legendre tables are built in FP64 (tables.hip) and only split here; the
// Fourier intermediates are FP32; grid and spectral arrays stay FP64 at the interface.
// This file builds the split tables (pre-split and pre-swizzled so that a lane reads its
// 8-element fragment with one 16-byte load) and drives the calls: FFT stage, column maxima
// of the scaled forms, and the contraction kernels of legendre_mx2.hip.
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "legendre_mixed.hpp"

namespace dsht {

// per-m table scale: 2^-e with e = ceil(log2(max |P|, |H|)) over the m block
__global__ void k_mx_scale(int N, int J2pad, const long* offF, const double* tP, const double* tH, double* minv,
                           int scaled) {
  __shared__ double red[256];
  const int m = blockIdx.x;
  const int nE = n_par(N, m, 0), nO = n_par(N, m, 1);
  const long len = (long)(cdiv(nE, TILE) + cdiv(nO, TILE)) * TILE * J2pad;
  double mx = 0.0;
  for (long e = threadIdx.x; e < len; e += blockDim.x) {
    mx = fmax(mx, fabs(tP[offF[m] + e]));
    mx = fmax(mx, fabs(tH[offF[m] + e]));
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double v = red[0];
    int e = 0;
    if (scaled && v > 0.0) e = (int)ceil(log2(v));
    minv[m] = ldexp(1.0, e);
  }
}

// One thread per element of the 16-bit swizzled tables.  Forward: per m, row tiles as in
// the FP64 tables, K = latitude in steps of 32; inverse: per m and parity, latitude tiles,
// K = wavenumber index in steps of 32.  Values are read from the FP64 swizzled tables.
template <typename T>
__global__ void k_mx_tables(int N, int J2, int J2pad, int J2p32, int nterm, const long* offF64,
                            const long* offI64, const long* offF, const long* offI, const double* FP,
                            const double* FH, const double* IP, const double* IH, const double* minv,
                            T* fP0, T* fP1, T* fP2, T* fH0, T* fH1, T* fH2, T* iP0, T* iP1, T* iP2, T* iH0,
                            T* iH1, T* iH2, int forward, double tscale) {
  const int m = blockIdx.y;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nE = n_par(N, m, 0), nO = n_par(N, m, 1);
  const int tE = cdiv(nE, TILE), tO = cdiv(nO, TILE);
  const double sc = 1.0 / minv[m];
  double vP = 0.0, vH = 0.0;
  long dst;
  if (forward) {
    const long len = (long)(tE + tO) * TILE * J2p32;
    if (e >= len) return;
    // e = ((t * KS + ks) * 64 + lane) * 8 + el
    const int el = (int)(e % 8), lane = (int)((e / 8) % 64);
    const long tk = e / 512;
    const int KS = J2p32 / KM;
    const int ks = (int)(tk % KS), t = (int)(tk / KS);
    const int row = lane % 16, jn = ks * KM + 8 * (lane / 16) + el;
    const int p = t < tE ? 0 : 1, k = (p ? t - tE : t) * TILE + row;
    if (jn < J2 && k < (p ? nO : nE)) {
      long fi = offF64[m] + ((long)(p ? tE : 0) + k / TILE) * (J2pad / 4) * 64 + (long)(jn / 4) * 64 + (k % TILE) +
                16 * (jn % 4);
      vP = FP[fi];
      vH = FH[fi];
    }
    dst = offF[m] + e;
  } else {
    const int kq0 = kpad32(nE), kq1 = kpad32(nO);
    const int RTj = J2p32 / TILE;
    const long len = (long)J2p32 * (kq0 + kq1);
    if (e >= len) return;
    const long half0 = (long)J2p32 * kq0;
    const int p = e < half0 ? 0 : 1;
    const long ee = p ? e - half0 : e;
    const int kq = p ? kq1 : kq0;
    const int KS = kq / KM;
    // ee = ((tj * KS + ks) * 64 + lane) * 8 + el
    const int el = (int)(ee % 8), lane = (int)((ee / 8) % 64);
    const long tk = ee / 512;
    const int ks = (int)(tk % KS), tj = (int)(tk / KS);
    (void)RTj;
    const int jn = tj * TILE + lane % 16, k = ks * KM + 8 * (lane / 16) + el;
    const int np = p ? nO : nE;
    if (jn < J2 && k < np) {
      const int kp0 = cdiv(nE, KCH) * KCH, kp1 = cdiv(nO, KCH) * KCH;
      const int kp = p ? kp1 : kp0;
      long ii = offI64[m] + (p ? (long)J2pad * kp0 : 0) + (long)(jn / TILE) * (kp / 4) * 64 + (long)(k / 4) * 64 +
                (jn % TILE) + 16 * (k % 4);
      vP = IP[ii];
      vH = IH[ii];
    }
    dst = offI[m] + e;
  }
  T* oP[3] = {forward ? fP0 : iP0, forward ? fP1 : iP1, forward ? fP2 : iP2};
  T* oH[3] = {forward ? fH0 : iH0, forward ? fH1 : iH1, forward ? fH2 : iH2};
  vP *= sc;
  vH *= sc;
  for (int t = 0; t < nterm; t++) {
    T hp = (T)vP, hh = (T)vH;
    oP[t][dst] = hp;
    oH[t][dst] = hh;
    vP = (vP - (double)hp) * tscale;
    vH = (vH - (double)hh) * tscale;
  }
}

__global__ void k_f32_copy(const double* in, float* out, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}

static int build_mixed(dsht_plan* p, int family) {
  MixedState* st = (MixedState*)p->mixed;
  if (!st) {
    st = new MixedState();
    p->mixed = st;
  }
  MixedTables& t = st->tab[family];
  if (t.nterm) return 0;
  const int N = p->N;
  t.J2p32 = kpad32(p->J2);
  std::vector<long> offF(N + 2), offI(N + 2);
  long f = 0, i = 0;
  for (int m = 0; m <= N; m++) {
    offF[m] = f;
    offI[m] = i;
    int nE = n_par(N, m, 0), nO = n_par(N, m, 1);
    f += (long)(cdiv(nE, TILE) + cdiv(nO, TILE)) * TILE * t.J2p32;
    i += (long)t.J2p32 * (kpad32(nE) + kpad32(nO));
  }
  offF[N + 1] = f;
  offI[N + 1] = i;
  t.sizeF = f;
  t.sizeI = i;
  {
    const long per = (long)t.J2p32 * KM;
    std::vector<int> cbase(N + 2), chunk_m;
    for (int m = 0; m <= N + 1; m++) cbase[m] = (int)(offI[m] / per);
    for (int m = 0; m <= N; m++)
      for (int g = cbase[m]; g < cbase[m + 1]; g++) chunk_m.push_back(m);
    t.nchunks = (int)chunk_m.size();
    DSHT_HIP(hipMalloc(&t.d_cbase, sizeof(int) * (N + 2)));
    DSHT_HIP(hipMalloc(&t.d_chunk_m, sizeof(int) * (chunk_m.size() + 1)));
    DSHT_HIP(hipMemcpy(t.d_cbase, cbase.data(), sizeof(int) * (N + 2), hipMemcpyHostToDevice));
    DSHT_HIP(hipMemcpy(t.d_chunk_m, chunk_m.data(), sizeof(int) * chunk_m.size(), hipMemcpyHostToDevice));
  }
  DSHT_HIP(hipMalloc(&t.d_offF, sizeof(long) * (N + 2)));
  DSHT_HIP(hipMalloc(&t.d_offI, sizeof(long) * (N + 2)));
  DSHT_HIP(hipMemcpy(t.d_offF, offF.data(), sizeof(long) * (N + 2), hipMemcpyHostToDevice));
  DSHT_HIP(hipMemcpy(t.d_offI, offI.data(), sizeof(long) * (N + 2), hipMemcpyHostToDevice));
  DSHT_HIP(hipMalloc(&t.d_minv, sizeof(double) * (N + 1)));
  for (int k = 0; k < 3; k++) {
    DSHT_HIP(hipMalloc(&t.fP[k], 2 * (f ? f : 1)));
    DSHT_HIP(hipMalloc(&t.fH[k], 2 * (f ? f : 1)));
    DSHT_HIP(hipMalloc(&t.iP[k], 2 * (i ? i : 1)));
    DSHT_HIP(hipMalloc(&t.iH[k], 2 * (i ? i : 1)));
    DSHT_HIP(hipMemset(t.fP[k], 0, 2 * f));
    DSHT_HIP(hipMemset(t.fH[k], 0, 2 * f));
    DSHT_HIP(hipMemset(t.iP[k], 0, 2 * i));
    DSHT_HIP(hipMemset(t.iH[k], 0, 2 * i));
  }
  hipLaunchKernelGGL(k_mx_scale, dim3(N + 1), dim3(256), 0, 0, N, p->J2pad, p->d_offF, p->d_tabFP, p->d_tabFH,
                     t.d_minv, family == 1 ? 1 : 0);
  DSHT_HIP(hipGetLastError());
  long maxF = 0, maxI = 0;
  for (int m = 0; m <= N; m++) {
    maxF = std::max(maxF, offF[m + 1] - offF[m]);
    maxI = std::max(maxI, offI[m + 1] - offI[m]);
  }
  for (int fwd = 1; fwd >= 0; fwd--) {
    long mx = fwd ? maxF : maxI;
    dim3 grid((unsigned)cdivl(mx, 256), N + 1);
    if (family == 0) {
      typedef __bf16 T;
      hipLaunchKernelGGL((k_mx_tables<T>), grid, dim3(256), 0, 0, N, p->J2, p->J2pad, t.J2p32, 3, p->d_offF,
                         p->d_offI, t.d_offF, t.d_offI, p->d_tabFP, p->d_tabFH, p->d_tabIP, p->d_tabIH, t.d_minv,
                         (T*)t.fP[0], (T*)t.fP[1], (T*)t.fP[2], (T*)t.fH[0], (T*)t.fH[1], (T*)t.fH[2],
                         (T*)t.iP[0], (T*)t.iP[1], (T*)t.iP[2], (T*)t.iH[0], (T*)t.iH[1], (T*)t.iH[2], fwd, 1.0);
    } else {
      typedef _Float16 T;
      hipLaunchKernelGGL((k_mx_tables<T>), grid, dim3(256), 0, 0, N, p->J2, p->J2pad, t.J2p32, 3, p->d_offF,
                         p->d_offI, t.d_offF, t.d_offI, p->d_tabFP, p->d_tabFH, p->d_tabIP, p->d_tabIH, t.d_minv,
                         (T*)t.fP[0], (T*)t.fP[1], (T*)t.fP[2], (T*)t.fH[0], (T*)t.fH[1], (T*)t.fH[2],
                         (T*)t.iP[0], (T*)t.iP[1], (T*)t.iP[2], (T*)t.iH[0], (T*)t.iH[1], (T*)t.iH[2], fwd, 2048.0);
    }
    DSHT_HIP(hipGetLastError());
  }
  DSHT_HIP(hipMalloc(&t.d_cwv, sizeof(float) * t.J2p32));
  DSHT_HIP(hipMalloc(&t.d_cwd, sizeof(float) * t.J2p32));
  DSHT_HIP(hipMemset(t.d_cwv, 0, sizeof(float) * t.J2p32));
  DSHT_HIP(hipMemset(t.d_cwd, 0, sizeof(float) * t.J2p32));
  hipLaunchKernelGGL(k_f32_copy, dim3(cdiv(p->J2, 256)), dim3(256), 0, 0, p->d_cwv, t.d_cwv, p->J2);
  hipLaunchKernelGGL(k_f32_copy, dim3(cdiv(p->J2, 256)), dim3(256), 0, 0, p->d_cwd, t.d_cwd, p->J2);
  DSHT_HIP(hipGetLastError());
  DSHT_HIP(hipDeviceSynchronize());
  t.nterm = 3;
  t.form_family = family;
  return 0;
}

void mixed_free(dsht_plan* p) {
  MixedState* st = (MixedState*)p->mixed;
  if (!st) return;
  for (auto& t : st->tab) {
    for (int k = 0; k < 3; k++) {
      if (t.fP[k]) (void)hipFree(t.fP[k]);
      if (t.fH[k]) (void)hipFree(t.fH[k]);
      if (t.iP[k]) (void)hipFree(t.iP[k]);
      if (t.iH[k]) (void)hipFree(t.iH[k]);
    }
    if (t.d_offF) (void)hipFree(t.d_offF);
    if (t.d_offI) (void)hipFree(t.d_offI);
    if (t.d_minv) (void)hipFree(t.d_minv);
    if (t.d_cwv) (void)hipFree(t.d_cwv);
    if (t.d_cwd) (void)hipFree(t.d_cwd);
    if (t.d_cbase) (void)hipFree(t.d_cbase);
    if (t.d_chunk_m) (void)hipFree(t.d_chunk_m);
  }
  delete st;
  p->mixed = nullptr;
}

// ---------------------------------------------------------------- helper kernels
__global__ void k_zero_u32(unsigned* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------- launchers
static int form_of(int prec) {
  switch (prec) {
    case DSHT_PREC_BF16X3: return 0;
    case DSHT_PREC_BF16X6: return 1;
    case DSHT_PREC_FP16: return 2;
    case DSHT_PREC_FP16X3: return 3;
  }
  return -1;
}

// ---------------------------------------------------------------- drivers
int mixed_chunk_cols(const dsht_plan* p, int ncols) { return chunk_columns(p, ncols, sizeof(float2)); }

static float weight_bound(const dsht_plan* p, bool div) {
  double mx = 0.0;
  for (int j = p->J2; j < p->J; j++) {
    double w = p->wq[j];
    if (div) w /= (1.0 - p->mu[j]) * (1.0 + p->mu[j]);
    mx = std::max(mx, w);
  }
  return (float)(2.0 * mx);
}

int mixed_run_fwd(dsht_plan* p, int op, const double* in0, const double* in1, long ldg, double* out0,
                  double* out1, long sl, long sc, int ncols, const double* diag0, const double* diag1,
                  double scale, const double* cw, hipStream_t s, bool grid_f32) {
  if (ncols <= 0) return 0;
  const int form = form_of(p->prec);
  if (form < 0) return fail(DSHT_EUNSUPPORTED, "precision form %d not available", p->prec);
  if (p->leg_kind == 1) return fail(DSHT_EUNSUPPORTED, "DSHT_TUNE_LEG=1 is retired for the split forms");
  const int family = MxFormFamily(form);
  if (int rc = build_mixed(p, family)) return rc;
  MixedTables& t = ((MixedState*)p->mixed)->tab[family];
  const int nin = (op == F_ALPHA || op == F_UV2VORDIV) ? 2 : 1;
  const bool scaled = form >= 2;
  const int ch = mixed_chunk_cols(p, ncols);
  const size_t gsz = (size_t)(p->N + 1) * p->J * fourier_pitch(ch);
  const size_t cmb = (2 * (size_t)ch * sizeof(unsigned) + 255) / 256 * 256;
  WsHandle wh;
  if (int rc = workspace_acquire(p, s, nin * gsz * sizeof(float2) + cmb, wh)) return rc;  // Fourier inputs + column maxima
  const bool vdiv = cw != p->d_cwv;
  const float bound = weight_bound(p, vdiv);
  float2* G0 = (float2*)wh.buf;
  float2* G1 = G0 + gsz;
  unsigned* colmax = (unsigned*)(G0 + nin * gsz);
  for (int c0 = 0; c0 < ncols; c0 += ch) {
    const int nc = ncols - c0 < ch ? ncols - c0 : ch;
    if (scaled) {
      hipLaunchKernelGGL(k_zero_u32, dim3(cdiv(2 * nc, 256)), dim3(256), 0, s, colmax, 2 * nc);
      DSHT_HIP(hipGetLastError());
    }
    const long gp = fourier_pitch(nc);
    int rc;
    if (grid_f32) {
      const float *f0 = (const float*)in0, *f1 = (const float*)in1;
      rc = launch_fft_fwd32f(p, f0 + (long)c0 * ldg, ldg, nc, G0, gp, scaled ? colmax : nullptr, s);
      if (!rc && nin > 1) rc = launch_fft_fwd32f(p, f1 + (long)c0 * ldg, ldg, nc, G1, gp, scaled ? colmax + nc : nullptr, s);
    } else {
      rc = launch_fft_fwd32(p, in0 + (long)c0 * ldg, ldg, nc, G0, gp, scaled ? colmax : nullptr, s);
      if (!rc && nin > 1) rc = launch_fft_fwd32(p, in1 + (long)c0 * ldg, ldg, nc, G1, gp, scaled ? colmax + nc : nullptr, s);
    }
    if (rc) return rc;
    double by, fl;
    mixed_work(p, op, true, nc, &by, &fl);
    ProfScope ps("leg_fwd", s, by, fl);
    Mx2FwdJob j;
    j.G0 = G0;
    j.G1 = nin > 1 ? G1 : nullptr;
    j.gld = gp;
    j.gmst = (long)p->J * gp;
    j.ncols = nc;
    j.W0 = out0 + (long)c0 * sc;
    j.W1 = out1 ? out1 + (long)c0 * sc : nullptr;
    j.sl = sl;
    j.sc = sc;
    j.cw = vdiv ? t.d_cwd : t.d_cwv;
    j.colmax = colmax;
    j.bound = bound;
    j.diag0 = diag0;
    j.diag1 = diag1;
    j.scale = scale;
    j.staged = p->leg_kind == 3;
    if ((rc = mx2_fwd(p, t, form, op, j, s))) return rc;
  }
  return 0;
}

// whole-K LDS staging for the inverse: per form by default (measured in the T170L60 step with
// FP32 grids: BF16X6 0.360 against 0.366 ms per step packed, FP16X3 0.360 against 0.377; BF16X3
// and FP16 contractions alone are faster packed), forced by DSHT_TUNE_LEG = 3, never where its
// LDS panel does not fit
static bool inv_staged(const dsht_plan* p, int form, int op) {
  const bool want = p->leg_kind == 3 || (p->leg_kind == 0 && (form == 1 || form == 3));
  return want && mx2_inv_staged_lds(p, form, op) <= MX2_STAGED_LDS_MAX;
}

int mixed_run_inv(dsht_plan* p, int op, const double* in0, const double* in1, long sl, long sc,
                  double* out0, double* out1, long ldg, int ncols, const double* pro, const double* rs,
                  hipStream_t s, double* out2, bool grid_f32) {
  if (ncols <= 0) return 0;
  const int form = form_of(p->prec);
  if (form < 0) return fail(DSHT_EUNSUPPORTED, "precision form %d not available", p->prec);
  if (p->leg_kind == 1) return fail(DSHT_EUNSUPPORTED, "DSHT_TUNE_LEG=1 is retired for the split forms");
  const int family = MxFormFamily(form);
  if (int rc = build_mixed(p, family)) return rc;
  MixedTables& t = ((MixedState*)p->mixed)->tab[family];
  const int nin = inv_nin(op);
  const int nout = inv_nout(op);
  const bool scaled = form >= 2;
  const bool staged = inv_staged(p, form, op);
  const int ch = mixed_chunk_cols(p, ncols);
  const size_t gsz = (size_t)(p->N + 1) * p->J * fourier_pitch(ch);
  const size_t slot_bytes = nout * gsz * sizeof(float2);  // Fourier outputs of one chunk
  const size_t cmb = (2 * (size_t)ch * sizeof(unsigned) + 255) / 256 * 256;
  const size_t pkb = staged ? 0 : mx2_inv_pack_bytes(t, form, op, ch);
  WsHandle wh;
  if (int rc = workspace_acquire(p, s, slot_bytes + cmb + pkb, wh)) return rc;
  float2* G0 = (float2*)wh.buf;
  float2* G1 = G0 + gsz;
  float2* G2 = G1 + gsz;
  unsigned* colmax = (unsigned*)((char*)wh.buf + slot_bytes);
  uint16_t* pack = (uint16_t*)((char*)colmax + cmb);
  for (int c0 = 0; c0 < ncols; c0 += ch) {
    const int nc = ncols - c0 < ch ? ncols - c0 : ch;
    // fused value/gradient: GradLon through the derivative inverse FFT of the value's
    // coefficients (i m X_m, 1/cos row scale); the contraction writes two outputs
    const bool dlon = op == I_VGRAD && out1 && fft_inv_deriv_ok(p);
    float2* O0 = out0 || dlon ? G0 : nullptr;
    float2* O1 = nout > 1 && out1 && !dlon ? G1 : nullptr;
    float2* O2 = nout > 2 && out2 ? G2 : nullptr;
    const double* w0 = in0 + (long)c0 * sc;
    const double* w1 = in1 ? in1 + (long)c0 * sc : nullptr;
    if (scaled) {
      hipLaunchKernelGGL(k_zero_u32, dim3(cdiv(2 * nc, 256)), dim3(256), 0, s, colmax, 2 * nc);
      DSHT_HIP(hipGetLastError());
      if (int rc = launch_spec_colmax(w0, sl, sc, p->L, nc, pro, colmax, s)) return rc;
      if (nin > 1)
        if (int rc = launch_spec_colmax(w1, sl, sc, p->L, nc, pro, colmax + nc, s)) return rc;
    }
    int rc;
    {
      double by, fl;
      mixed_work(p, op, false, nc, &by, &fl, (O0 ? 1 : 0) + (O1 ? 1 : 0) + (O2 ? 1 : 0));
      ProfScope ps("leg_inv", s, by, fl);
      Mx2InvJob j;
      j.W0 = w0;
      j.W1 = w1;
      j.sl = sl;
      j.sc = sc;
      j.ncols = nc;
      j.pro = pro;
      j.rs = rs;
      j.colmax = colmax;
      j.pack = pack;
      j.staged = staged;
      j.G0 = O0;
      j.G1 = O1;
      j.G2 = O2;
      j.gld = fourier_pitch(nc);
      j.gmst = (long)p->J * j.gld;
      rc = mx2_inv(p, t, form, op, j, s);
    }
    if (rc) return rc;
    const long gp = fourier_pitch(nc);
    // one inverse FFT per output; the field and its longitude derivative (fused value/gradient
    // call) come from one read of G0
    auto inv = [&](const float2* G, double* out, const double* rsc, int deriv, double* out2d) -> int {
      if (grid_f32) {
        float* o = (float*)out + (long)c0 * ldg;
        float* o2 = out2d ? (float*)out2d + (long)c0 * ldg : nullptr;
        return launch_fft_inv32f(p, G, gp, nc, o, ldg, s, rsc, deriv, o2);
      }
      return launch_fft_inv32(p, G, gp, nc, out + (long)c0 * ldg, ldg, s, rsc, deriv,
                              out2d ? out2d + (long)c0 * ldg : nullptr);
    };
    if (out0 && dlon) {
      if ((rc = inv(G0, out0, p->d_rs_icos, 0, out1))) return rc;
    } else {
      if (out0 && (rc = inv(G0, out0, nullptr, 0, nullptr))) return rc;
      if (dlon && (rc = inv(G0, out1, p->d_rs_icos, 1, nullptr))) return rc;
    }
    if (O1 && (rc = inv(G1, out1, nullptr, 0, nullptr))) return rc;
    if (O2 && (rc = inv(G2, out2, nullptr, 0, nullptr))) return rc;
  }
  return 0;
}

}  // namespace dsht
