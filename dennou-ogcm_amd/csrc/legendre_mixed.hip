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
// Data are split once when staged to LDS (k-contiguous per column so a lane's 8-element
// fragment is one 16-byte LDS read); table fragments are pre-split, pre-swizzled and read
// with one 16-byte global load per lane.
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

// ---------------------------------------------------------------- kernel arguments
struct MxFwdArgs {
  const float2* G0;
  const float2* G1;
  double* W0;
  double* W1;
  long sl, sc;
  int ncols;
  long gld, gmst;
  const uint16_t* tP[3];
  const uint16_t* tH[3];
  const long* off;
  const float* cw;
  const double* minv;
  const unsigned* colmax;  // [input][column] max |G| (float bits); scaled forms only
  float bound;             // staged |value| <= bound * colmax (2 max c_j)
  const double* diag0;
  const double* diag1;
  double scale;
  int N, J2, J2p32;
};

struct MxInvArgs {
  const double* W0;
  const double* W1;
  long sl, sc;
  float2* G0;
  float2* G1;
  int ncols;
  long gld, gmst;
  const uint16_t* tP[3];
  const uint16_t* tH[3];
  const long* off;
  const double* minv;
  const unsigned* colmax;  // [input][column] max |pro w| (float bits); scaled forms only
  const double* rs;
  const double* pro;
  int N, J2, J2p32;
};

// common power-of-two scale of a column over all inputs, so staged |values| <= 1
template <int NIN>
__device__ __forceinline__ float col_scale(const unsigned* colmax, int ncols, int c, float bound) {
  float mx = 0.f;
#pragma unroll
  for (int i = 0; i < NIN; i++) mx = fmaxf(mx, __uint_as_float(colmax[(long)i * ncols + c]));
  mx *= bound;
  return mx > 0.f ? exp2f(-ceilf(log2f(mx))) : 1.f;
}

// NP products of split table fragments with split data fragments read from LDS
template <int FORM>
__device__ __forceinline__ float4_t mx_prod(const typename MxForm<FORM>::V8 (&A)[MxForm<FORM>::NT],
                                            const typename MxForm<FORM>::T* b, int tstride) {
  using F = MxForm<FORM>;
  using V8 = typename F::V8;
  V8 B[F::NT];
#pragma unroll
  for (int t = 0; t < F::NT; t++) B[t] = *reinterpret_cast<const V8*>(b + t * tstride);
  if constexpr (F::BF) {
    float4_t acc = float4_t{0, 0, 0, 0};
#pragma unroll
    for (int pr = 0; pr < F::NP; pr++) acc = mfma16(A[PA[pr]], B[PB[pr]], acc);
    return acc;
  } else {
    // accumulate each product order separately, then weight by 2^(-11 order)
    float4_t acc[3] = {float4_t{0, 0, 0, 0}, float4_t{0, 0, 0, 0}, float4_t{0, 0, 0, 0}};
#pragma unroll
    for (int pr = 0; pr < F::NP; pr++) acc[PA[pr] + PB[pr]] = mfma16(A[PA[pr]], B[PB[pr]], acc[PA[pr] + PB[pr]]);
    constexpr float D = TermScale<FORM>::DOWN;
    return acc[0] + D * (acc[1] + D * acc[2]);
  }
}

// ---------------------------------------------------------------- forward
// Wave w owns row tiles t = tbase + w + WAVES*tt.  Each chunk is one 32-wide K step over
// north latitudes, staged as [input][S re, S im, A re, A im][term][column][k].
template <int OP, int FORM, int CB, int WAVES, int MAXT>
__global__ __launch_bounds__(WAVES * 64) void k_mx_fwd(MxFwdArgs a) {
  using Tr = FwdTraits<OP>;
  using F = MxForm<FORM>;
  using T = typename F::T;
  using V8 = typename F::V8;
  constexpr int NIN = Tr::NIN, NOUT = Tr::NOUT, NCT = CB / TILE, NTERM = F::NT;
  constexpr int KP = KM + 8;  // k pitch (elements) per column, keeps 16-byte alignment
  constexpr int TS = CB * KP;  // term stride
  constexpr bool UP = Tr::USE_P, UH = Tr::USE_H;
  __shared__ __attribute__((aligned(16))) T sB[NIN][4][NTERM * TS];

  const int m = blockIdx.y;
  const int c0 = blockIdx.x * CB;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int N = a.N, J2 = a.J2;
  const int nE = n_par(N, m, 0), nO = n_par(N, m, 1);
  const int tE = (nE + TILE - 1) / TILE, tO = (nO + TILE - 1) / TILE;
  const int RT = tE + tO;
  const int KS = a.J2p32 / KM;
  const long toff = a.off[m];
  const float2* Gin[2] = {a.G0 + (long)m * a.gmst, a.G1 ? a.G1 + (long)m * a.gmst : nullptr};
  const float dm = (float)m;
  // staging role: thread -> (input, column, group of 8 latitudes)
  constexpr int PER_IN = CB * (KM / 8);
  static_assert(NIN * PER_IN <= WAVES * 64, "staging roles exceed the workgroup");
  const int si = tid / PER_IN, sr = tid % PER_IN;
  const int scol = sr % CB, sg = sr / CB;
  const bool stager = tid < NIN * PER_IN;
  float dsc = 1.f;
  if constexpr (F::SCALED) {
    if (stager && c0 + scol < a.ncols) dsc = col_scale<NIN>(a.colmax, a.ncols, c0 + scol, a.bound);
  }

  for (int tbase = blockIdx.z * WAVES * MAXT; tbase < RT; tbase += gridDim.z * WAVES * MAXT) {
    float4_t acc[NOUT][MAXT][NCT][2];
#pragma unroll
    for (int o = 0; o < NOUT; o++)
#pragma unroll
      for (int t = 0; t < MAXT; t++)
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) acc[o][t][ct][0] = acc[o][t][ct][1] = float4_t{0, 0, 0, 0};

    float2 gN[8], gS[8];
    auto load_b = [&](int ks) {
#pragma unroll
      for (int e = 0; e < 8; e++) gN[e] = gS[e] = make_float2(0.f, 0.f);
      if (stager) {
        const int c = c0 + scol;
#pragma unroll
        for (int e = 0; e < 8; e++) {
          const int jn = ks * KM + sg * 8 + e;
          if (jn < J2 && c < a.ncols) {
            gN[e] = Gin[si][(long)(J2 + jn) * a.gld + c];
            gS[e] = Gin[si][(long)(J2 - 1 - jn) * a.gld + c];
          }
        }
      }
    };
    auto stage = [&](int ks) {
      if (!stager) return;
      T part[4][NTERM][8];
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const float cw = a.cw[ks * KM + sg * 8 + e] * dsc;
        const float v[4] = {cw * (gN[e].x + gS[e].x), cw * (gN[e].y + gS[e].y), cw * (gN[e].x - gS[e].x),
                            cw * (gN[e].y - gS[e].y)};
#pragma unroll
        for (int q = 0; q < 4; q++) {
          T tmp[NTERM];
          split_to<FORM>(v[q], tmp);
#pragma unroll
          for (int t = 0; t < NTERM; t++) part[q][t][e] = tmp[t];
        }
      }
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int t = 0; t < NTERM; t++) {
          V8 v;
#pragma unroll
          for (int e = 0; e < 8; e++) v[e] = part[q][t][e];
          *reinterpret_cast<V8*>(&sB[si][q][t * TS + scol * KP + sg * 8]) = v;
        }
    };
    V8 aP[2][MAXT][NTERM], aH[2][MAXT][NTERM];
    auto load_a = [&](int ks, int slot) {
#pragma unroll
      for (int tt = 0; tt < MAXT; tt++) {
        const int t = tbase + w + WAVES * tt;
        const long o = toff + ((long)t * KS + ks) * 512 + lane * 8;
#pragma unroll
        for (int k = 0; k < NTERM; k++) {
          if constexpr (UP) aP[slot][tt][k] = t < RT ? *reinterpret_cast<const V8*>(a.tP[k] + o) : V8{};
          if constexpr (UH) aH[slot][tt][k] = t < RT ? *reinterpret_cast<const V8*>(a.tH[k] + o) : V8{};
        }
      }
    };
    auto compute = [&](int slot) {
      const int kq = 8 * (lane >> 4), col = lane & 15;
#pragma unroll
      for (int tt = 0; tt < MAXT; tt++) {
        const int t = tbase + w + WAVES * tt;
        if (t >= RT) continue;
        // P pairs with sym (quantities 0,1) for even rows and anti (2,3) for odd; H the opposite
        const int qP = t < tE ? 0 : 2, qH = 2 - qP;
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) {
          const int cc = (ct * TILE + col) * KP + kq;
#pragma unroll
          for (int r = 0; r < 2; r++) {
            // i m z: re <- -m z.im, im <- m z.re
            const float rf = r == 0 ? -dm : dm;
            if constexpr (OP == F_VAL) {
              acc[0][tt][ct][r] += mx_prod<FORM>(aP[slot][tt], &sB[0][qP + r][cc], TS);
            } else if constexpr (OP == F_DIVLAMBDA) {
              acc[0][tt][ct][r] += rf * mx_prod<FORM>(aP[slot][tt], &sB[0][qP + 1 - r][cc], TS);
            } else if constexpr (OP == F_DIVMU) {
              acc[0][tt][ct][r] -= mx_prod<FORM>(aH[slot][tt], &sB[0][qH + r][cc], TS);
            } else if constexpr (OP == F_ALPHA) {  // i m P.A - H.B
              acc[0][tt][ct][r] += rf * mx_prod<FORM>(aP[slot][tt], &sB[0][qP + 1 - r][cc], TS);
              acc[0][tt][ct][r] -= mx_prod<FORM>(aH[slot][tt], &sB[1][qH + r][cc], TS);
            } else {  // F_UV2VORDIV: vor = i m P.V + H.U ; div = i m P.U - H.V
              acc[0][tt][ct][r] += rf * mx_prod<FORM>(aP[slot][tt], &sB[1][qP + 1 - r][cc], TS);
              acc[0][tt][ct][r] += mx_prod<FORM>(aH[slot][tt], &sB[0][qH + r][cc], TS);
              acc[1][tt][ct][r] += rf * mx_prod<FORM>(aP[slot][tt], &sB[0][qP + 1 - r][cc], TS);
              acc[1][tt][ct][r] -= mx_prod<FORM>(aH[slot][tt], &sB[1][qH + r][cc], TS);
            }
          }
        }
      }
    };

    load_b(0);
    load_a(0, 0);
    for (int ks = 0; ks < KS; ks += 2) {
      stage(ks);
      __syncthreads();
      if (ks + 1 < KS) {
        load_b(ks + 1);
        load_a(ks + 1, 1);
      }
      compute(0);
      __syncthreads();
      if (ks + 1 >= KS) break;
      stage(ks + 1);
      __syncthreads();
      if (ks + 2 < KS) {
        load_b(ks + 2);
        load_a(ks + 2, 0);
      }
      compute(1);
      __syncthreads();
    }

    // ---- epilogue (FP64): table and data unscale, diagonal, scatter to SPMODEL order
    const double mi = a.minv[m];
#pragma unroll
    for (int tt = 0; tt < MAXT; tt++) {
      const int t = tbase + w + WAVES * tt;
      if (t >= RT) continue;
      const int p = t < tE ? 0 : 1;
      const int tp = p ? t - tE : t;
#pragma unroll
      for (int ct = 0; ct < NCT; ct++) {
        const int c = c0 + ct * TILE + (lane & 15);
        if (c >= a.ncols) continue;
        double u = mi;
        if constexpr (F::SCALED) u = mi / (double)col_scale<NIN>(a.colmax, a.ncols, c, a.bound);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int k = tp * TILE + acc_row16(lane, i);
          const int n = m + p + 2 * k;
          if (n > N) continue;
          const long lre = (long)n * (n + 1) + m, lim = (long)n * (n + 1) - m;
#pragma unroll
          for (int o = 0; o < NOUT; o++) {
            const double* dg = o == 0 ? a.diag0 : a.diag1;
            double* W = o == 0 ? a.W0 : a.W1;
            W[lre * a.sl + (long)c * a.sc] = (double)acc[o][tt][ct][0][i] * u * a.scale * (dg ? dg[lre] : 1.0);
            if (m > 0)
              W[lim * a.sl + (long)c * a.sc] = (double)acc[o][tt][ct][1][i] * u * a.scale * (dg ? dg[lim] : 1.0);
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------- inverse
// Wave w owns latitude tiles t = tbase + w + WAVES*tt; the K loop runs over the total
// wavenumbers of parity 0 then parity 1, 32 per chunk; data staged as
// [input][re, im][term][column][k].
template <int OP, int FORM, int CB, int WAVES, int MAXT>
__global__ __launch_bounds__(WAVES * 64) void k_mx_inv(MxInvArgs a) {
  using Tr = InvTraits<OP>;
  using F = MxForm<FORM>;
  using T = typename F::T;
  using V8 = typename F::V8;
  constexpr int NIN = Tr::NIN, NOUT = Tr::NOUT, NCT = CB / TILE, NTERM = F::NT;
  constexpr int KP = KM + 8;
  constexpr int TS = CB * KP;
  constexpr bool UP = Tr::USE_P, UH = Tr::USE_H;
  __shared__ __attribute__((aligned(16))) T sB[NIN][2][NTERM * TS];

  const int m = blockIdx.y;
  const int c0 = blockIdx.x * CB;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int N = a.N, J2 = a.J2;
  const int nE = n_par(N, m, 0), nO = n_par(N, m, 1);
  const int kq0 = kpad32(nE), kq1 = kpad32(nO);
  const int nch0 = kq0 / KM, nch = nch0 + kq1 / KM;
  const int RT = a.J2p32 / TILE;
  const long toff = a.off[m];
  const double* Win[2] = {a.W0, a.W1};
  const float dm = (float)m;
  constexpr int PER_IN = CB * (KM / 8);
  static_assert(NIN * PER_IN <= WAVES * 64, "staging roles exceed the workgroup");
  const int si = tid / PER_IN, sr = tid % PER_IN;
  const int scol = sr % CB, sg = sr / CB;
  const bool stager = tid < NIN * PER_IN;
  float dsc = 1.f;
  if constexpr (F::SCALED) {
    if (stager && c0 + scol < a.ncols) dsc = col_scale<NIN>(a.colmax, a.ncols, c0 + scol, 1.f);
  }

  for (int tbase = blockIdx.z * WAVES * MAXT; tbase < RT; tbase += gridDim.z * WAVES * MAXT) {
    float4_t acc[NOUT][MAXT][NCT][2][2];  // [out][tile][ct][re/im][sym/anti]
#pragma unroll
    for (int o = 0; o < NOUT; o++)
#pragma unroll
      for (int t = 0; t < MAXT; t++)
#pragma unroll
        for (int ct = 0; ct < NCT; ct++)
#pragma unroll
          for (int r = 0; r < 2; r++) acc[o][t][ct][r][0] = acc[o][t][ct][r][1] = float4_t{0, 0, 0, 0};

    double bre[8], bim[8];
    auto load_b = [&](int ch) {
      const int p = ch < nch0 ? 0 : 1;
      const int k0 = (p ? ch - nch0 : ch) * KM;
      const int np = p ? nO : nE;
#pragma unroll
      for (int e = 0; e < 8; e++) bre[e] = bim[e] = 0.0;
      if (stager) {
        const int c = c0 + scol;
#pragma unroll
        for (int e = 0; e < 8; e++) {
          const int k = k0 + sg * 8 + e;
          if (k < np && c < a.ncols) {
            const int n = m + p + 2 * k;
            const long lre = (long)n * (n + 1) + m, lim = (long)n * (n + 1) - m;
            const double pr = a.pro ? a.pro[lre] : 1.0;
            bre[e] = pr * Win[si][lre * a.sl + (long)c * a.sc];
            if (m > 0) bim[e] = pr * Win[si][lim * a.sl + (long)c * a.sc];
          }
        }
      }
    };
    auto stage = [&]() {
      if (!stager) return;
      T pre[NTERM][8], pim[NTERM][8];
#pragma unroll
      for (int e = 0; e < 8; e++) {
        T t1[NTERM], t2[NTERM];
        split_to<FORM>((float)(bre[e] * dsc), t1);
        split_to<FORM>((float)(bim[e] * dsc), t2);
#pragma unroll
        for (int t = 0; t < NTERM; t++) {
          pre[t][e] = t1[t];
          pim[t][e] = t2[t];
        }
      }
#pragma unroll
      for (int t = 0; t < NTERM; t++) {
        V8 v1, v2;
#pragma unroll
        for (int e = 0; e < 8; e++) {
          v1[e] = pre[t][e];
          v2[e] = pim[t][e];
        }
        *reinterpret_cast<V8*>(&sB[si][0][t * TS + scol * KP + sg * 8]) = v1;
        *reinterpret_cast<V8*>(&sB[si][1][t * TS + scol * KP + sg * 8]) = v2;
      }
    };
    V8 aP[2][MAXT][NTERM], aH[2][MAXT][NTERM];
    auto load_a = [&](int ch, int slot) {
      const int p = ch < nch0 ? 0 : 1;
      const int ks = p ? ch - nch0 : ch;
      const int KSp = (p ? kq1 : kq0) / KM;
      const long pbase = toff + (p ? (long)a.J2p32 * kq0 : 0);
#pragma unroll
      for (int tt = 0; tt < MAXT; tt++) {
        const int t = tbase + w + WAVES * tt;
        const long o = pbase + ((long)t * KSp + ks) * 512 + lane * 8;
#pragma unroll
        for (int k = 0; k < NTERM; k++) {
          if constexpr (UP) aP[slot][tt][k] = t < RT ? *reinterpret_cast<const V8*>(a.tP[k] + o) : V8{};
          if constexpr (UH) aH[slot][tt][k] = t < RT ? *reinterpret_cast<const V8*>(a.tH[k] + o) : V8{};
        }
      }
    };
    // P-terms land in sym for even parity and anti for odd; H-terms the opposite
    auto compute_p = [&](auto PARITY, int slot) {
      constexpr int sP = decltype(PARITY)::value, sH = 1 - sP;
      const int kq = 8 * (lane >> 4), col = lane & 15;
#pragma unroll
      for (int tt = 0; tt < MAXT; tt++) {
        const int t = tbase + w + WAVES * tt;
        if (t >= RT) continue;
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) {
          const int cc = (ct * TILE + col) * KP + kq;
#pragma unroll
          for (int r = 0; r < 2; r++) {
            const float rf = r == 0 ? -dm : dm;
            if constexpr (OP == I_VAL || OP == I_GRADLAMBDA) {
              acc[0][tt][ct][r][sP] += mx_prod<FORM>(aP[slot][tt], &sB[0][r][cc], TS);
            } else if constexpr (OP == I_GRADMU) {
              acc[0][tt][ct][r][sH] += mx_prod<FORM>(aH[slot][tt], &sB[0][r][cc], TS);
            } else if constexpr (OP == I_ALPHA) {  // i m P.A + H.B
              acc[0][tt][ct][r][sP] += rf * mx_prod<FORM>(aP[slot][tt], &sB[0][1 - r][cc], TS);
              acc[0][tt][ct][r][sH] += mx_prod<FORM>(aH[slot][tt], &sB[1][r][cc], TS);
            } else {  // I_VORDIV2UV: U = i m P.chi - H.psi ; V = i m P.psi + H.chi
              acc[0][tt][ct][r][sP] += rf * mx_prod<FORM>(aP[slot][tt], &sB[1][1 - r][cc], TS);
              acc[0][tt][ct][r][sH] -= mx_prod<FORM>(aH[slot][tt], &sB[0][r][cc], TS);
              acc[1][tt][ct][r][sP] += rf * mx_prod<FORM>(aP[slot][tt], &sB[0][1 - r][cc], TS);
              acc[1][tt][ct][r][sH] += mx_prod<FORM>(aH[slot][tt], &sB[1][r][cc], TS);
            }
          }
        }
      }
    };
    auto compute = [&](int ch, int slot) {
      if (ch < nch0)
        compute_p(std::integral_constant<int, 0>(), slot);
      else
        compute_p(std::integral_constant<int, 1>(), slot);
    };

    load_b(0);
    load_a(0, 0);
    for (int ch = 0; ch < nch; ch += 2) {
      stage();
      __syncthreads();
      if (ch + 1 < nch) {
        load_b(ch + 1);
        load_a(ch + 1, 1);
      }
      compute(ch, 0);
      __syncthreads();
      if (ch + 1 >= nch) break;
      stage();
      __syncthreads();
      if (ch + 2 < nch) {
        load_b(ch + 2);
        load_a(ch + 2, 0);
      }
      compute(ch + 1, 1);
      __syncthreads();
    }

    // ---- epilogue: north = sym + anti, south = sym - anti; unscale, row scale, FP32 store
    const double mi = a.minv[m];
#pragma unroll
    for (int tt = 0; tt < MAXT; tt++) {
      const int t = tbase + w + WAVES * tt;
      if (t >= RT) continue;
#pragma unroll
      for (int ct = 0; ct < NCT; ct++) {
        const int c = c0 + ct * TILE + (lane & 15);
        if (c >= a.ncols) continue;
        double u = mi;
        if constexpr (F::SCALED) u = mi / (double)col_scale<NIN>(a.colmax, a.ncols, c, 1.f);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int jr = t * TILE + acc_row16(lane, i);
          if (jr >= J2) continue;
          const int jN = J2 + jr, jS = J2 - 1 - jr;
          const double sN = u * (a.rs ? a.rs[jN] : 1.0), sS = u * (a.rs ? a.rs[jS] : 1.0);
#pragma unroll
          for (int o = 0; o < NOUT; o++) {
            double nre = (double)acc[o][tt][ct][0][0][i] + (double)acc[o][tt][ct][0][1][i];
            double nim = (double)acc[o][tt][ct][1][0][i] + (double)acc[o][tt][ct][1][1][i];
            double sre = (double)acc[o][tt][ct][0][0][i] - (double)acc[o][tt][ct][0][1][i];
            double sim = (double)acc[o][tt][ct][1][0][i] - (double)acc[o][tt][ct][1][1][i];
            if constexpr (OP == I_GRADLAMBDA) {
              double t1 = -(double)m * nim;
              nim = (double)m * nre;
              nre = t1;
              t1 = -(double)m * sim;
              sim = (double)m * sre;
              sre = t1;
            }
            float2* G = (o == 0 ? a.G0 : a.G1) + (long)m * a.gmst;
            G[(long)jN * a.gld + c] = make_float2((float)(nre * sN), m ? (float)(nim * sN) : 0.f);
            G[(long)jS * a.gld + c] = make_float2((float)(sre * sS), m ? (float)(sim * sS) : 0.f);
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------- helper kernels
__global__ void k_zero_u32(unsigned* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// per-column max |pro[l] w(l, c)| over all l (scaled inverse forms); one block per column
__global__ __launch_bounds__(256) void k_spec_colmax(const double* __restrict__ w, long sl, long sc, int L,
                                                     const double* __restrict__ pro, unsigned* __restrict__ out) {
  __shared__ float red[256];
  const int c = blockIdx.x;
  float mx = 0.f;
  for (int l = threadIdx.x; l < L; l += blockDim.x)
    mx = fmaxf(mx, (float)fabs((pro ? pro[l] : 1.0) * w[(long)l * sl + (long)c * sc]));
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = __float_as_uint(red[0]);
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

// one row tile per wave; the row tiles of heavy wavenumbers are spread over grid z
template <int OP, int FORM>
static int mx_fwd_launch(const dsht_plan* p, const MxFwdArgs& a, hipStream_t s) {
  const int RT0 = cdiv(n_par(p->N, 0, 0), TILE) + cdiv(n_par(p->N, 0, 1), TILE);
  constexpr int CB = FwdTraits<OP>::NOUT == 2 ? 16 : 32;
  dim3 grid(cdiv(a.ncols, CB), p->N + 1, cdiv(RT0, 8));
  hipLaunchKernelGGL((k_mx_fwd<OP, FORM, CB, 8, 1>), grid, dim3(512), 0, s, a);
  DSHT_HIP(hipGetLastError());
  return 0;
}

template <int OP, int FORM>
static int mx_inv_launch(const dsht_plan* p, const MxInvArgs& a, hipStream_t s) {
  const int RT = a.J2p32 / TILE;
  constexpr int CB = InvTraits<OP>::NOUT == 2 ? 16 : 32;
  dim3 grid(cdiv(a.ncols, CB), p->N + 1, cdiv(RT, 8));
  hipLaunchKernelGGL((k_mx_inv<OP, FORM, CB, 8, 1>), grid, dim3(512), 0, s, a);
  DSHT_HIP(hipGetLastError());
  return 0;
}

template <int FORM>
static int mx_fwd_op(const dsht_plan* p, int op, const MxFwdArgs& a, hipStream_t s) {
  switch (op) {
    case F_VAL: return mx_fwd_launch<F_VAL, FORM>(p, a, s);
    case F_DIVLAMBDA: return mx_fwd_launch<F_DIVLAMBDA, FORM>(p, a, s);
    case F_DIVMU: return mx_fwd_launch<F_DIVMU, FORM>(p, a, s);
    case F_ALPHA: return mx_fwd_launch<F_ALPHA, FORM>(p, a, s);
    case F_UV2VORDIV: return mx_fwd_launch<F_UV2VORDIV, FORM>(p, a, s);
  }
  return fail(DSHT_EARG, "bad forward op %d", op);
}

template <int FORM>
static int mx_inv_op(const dsht_plan* p, int op, const MxInvArgs& a, hipStream_t s) {
  switch (op) {
    case I_VAL: return mx_inv_launch<I_VAL, FORM>(p, a, s);
    case I_GRADLAMBDA: return mx_inv_launch<I_GRADLAMBDA, FORM>(p, a, s);
    case I_GRADMU: return mx_inv_launch<I_GRADMU, FORM>(p, a, s);
    case I_ALPHA: return mx_inv_launch<I_ALPHA, FORM>(p, a, s);
    case I_VORDIV2UV: return mx_inv_launch<I_VORDIV2UV, FORM>(p, a, s);
  }
  return fail(DSHT_EARG, "bad inverse op %d", op);
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
  const int family = MxFormFamily(form);
  if (int rc = build_mixed(p, family)) return rc;
  MixedTables& t = ((MixedState*)p->mixed)->tab[family];
  const int nin = (op == F_ALPHA || op == F_UV2VORDIV) ? 2 : 1;
  const bool scaled = form >= 2;
  // the first (LDS-staged) split design keeps the BF16 forms only: its FP16 product-order
  // weighting loses up to ~2e-5 relative on zero-mean data (the fragment-stream kernels
  // hold the FP16 forms to their bounds)
  if (scaled && p->leg_kind == 1)
    return fail(DSHT_EUNSUPPORTED, "DSHT_TUNE_LEG=1 (first split design) supports the BF16 forms only");
  const int ch = mixed_chunk_cols(p, ncols);
  const int nchunks = cdiv(ncols, ch);
  const size_t gsz = (size_t)(p->N + 1) * p->J * fourier_pitch(ch);
  const size_t cmb = (2 * (size_t)ch * sizeof(unsigned) + 255) / 256 * 256;
  const size_t slot_bytes = nin * gsz * sizeof(float2) + cmb;  // Fourier inputs + column maxima
  WsHandle wh;
  if (int rc = workspace_acquire(p, s, (p->pipeline && nchunks > 1 ? 2 : 1) * slot_bytes, wh)) return rc;
  ChunkPipe cp;
  if (int rc = pipe_begin(wh, s, p, nchunks, cp)) return rc;
  const bool vdiv = cw != p->d_cwv;
  const float bound = weight_bound(p, vdiv);
  for (int k = 0, c0 = 0; c0 < ncols; k++, c0 += ch) {
    const int nc = ncols - c0 < ch ? ncols - c0 : ch;
    const int slot = cp.on ? (k & 1) : 0;
    float2* G0 = (float2*)((char*)wh.buf + slot * slot_bytes);
    float2* G1 = G0 + gsz;
    unsigned* colmax = (unsigned*)(G0 + nin * gsz);
    hipStream_t fs = s;  // FFT stage
    if (cp.on) {
      fs = cp.aux;
      if (k >= 2) DSHT_HIP(hipStreamWaitEvent(fs, cp.evL[slot], 0));  // slot free again
    }
    if (scaled) {
      hipLaunchKernelGGL(k_zero_u32, dim3(cdiv(2 * nc, 256)), dim3(256), 0, fs, colmax, 2 * nc);
      DSHT_HIP(hipGetLastError());
    }
    const long gp = fourier_pitch(nc);
    int rc;
    if (grid_f32) {
      const float *f0 = (const float*)in0, *f1 = (const float*)in1;
      rc = launch_fft_fwd32f(p, f0 + (long)c0 * ldg, ldg, nc, G0, gp, scaled ? colmax : nullptr, fs);
      if (!rc && nin > 1) rc = launch_fft_fwd32f(p, f1 + (long)c0 * ldg, ldg, nc, G1, gp, scaled ? colmax + nc : nullptr, fs);
    } else {
      rc = launch_fft_fwd32(p, in0 + (long)c0 * ldg, ldg, nc, G0, gp, scaled ? colmax : nullptr, fs);
      if (!rc && nin > 1) rc = launch_fft_fwd32(p, in1 + (long)c0 * ldg, ldg, nc, G1, gp, scaled ? colmax + nc : nullptr, fs);
    }
    if (rc) return rc;
    if (cp.on && (rc = pipe_after(cp.evF[slot], fs, s))) return rc;
    MxFwdArgs a;
    a.G0 = G0;
    a.G1 = nin > 1 ? G1 : nullptr;
    a.W0 = out0 + (long)c0 * sc;
    a.W1 = out1 ? out1 + (long)c0 * sc : nullptr;
    a.sl = sl;
    a.sc = sc;
    a.ncols = nc;
    a.gld = fourier_pitch(nc);
    a.gmst = (long)p->J * a.gld;
    for (int q = 0; q < 3; q++) {
      a.tP[q] = t.fP[q];
      a.tH[q] = t.fH[q];
    }
    a.off = t.d_offF;
    a.cw = vdiv ? t.d_cwd : t.d_cwv;
    a.minv = t.d_minv;
    a.colmax = colmax;
    a.bound = bound;
    a.diag0 = diag0;
    a.diag1 = diag1;
    a.scale = scale;
    a.N = p->N;
    a.J2 = p->J2;
    a.J2p32 = t.J2p32;
    {
      double by, fl;
      mixed_work(p, op, true, nc, &by, &fl);
      ProfScope ps("leg_fwd", s, by, fl);
      if (p->leg_kind != 1) {
        Mx2FwdJob j;
        j.G0 = G0;
        j.G1 = nin > 1 ? G1 : nullptr;
        j.gld = fourier_pitch(nc);
        j.gmst = (long)p->J * j.gld;
        j.ncols = nc;
        j.W0 = a.W0;
        j.W1 = a.W1;
        j.sl = sl;
        j.sc = sc;
        j.cw = a.cw;
        j.colmax = colmax;
        j.bound = bound;
        j.diag0 = diag0;
        j.diag1 = diag1;
        j.scale = scale;
        j.shared = p->leg_kind != 2;  // 2: the per-wave fragment-stream forward
        rc = mx2_fwd(p, t, form, op, j, s);
      } else {
        switch (form) {
          case 0: rc = mx_fwd_op<0>(p, op, a, s); break;
          case 1: rc = mx_fwd_op<1>(p, op, a, s); break;
          case 2: rc = mx_fwd_op<2>(p, op, a, s); break;
          default: rc = mx_fwd_op<3>(p, op, a, s); break;
        }
      }
    }
    if (rc) return rc;
    if (cp.on) DSHT_HIP(hipEventRecord(cp.evL[slot], s));
  }
  return pipe_join(cp);
}

int mixed_run_inv(dsht_plan* p, int op, const double* in0, const double* in1, long sl, long sc,
                  double* out0, double* out1, long ldg, int ncols, const double* pro, const double* rs,
                  hipStream_t s, double* out2, bool grid_f32) {
  if (ncols <= 0) return 0;
  const int form = form_of(p->prec);
  if (form < 0) return fail(DSHT_EUNSUPPORTED, "precision form %d not available", p->prec);
  const int family = MxFormFamily(form);
  if (int rc = build_mixed(p, family)) return rc;
  MixedTables& t = ((MixedState*)p->mixed)->tab[family];
  const int nin = inv_nin(op);
  const int nout = inv_nout(op);
  if (op == I_VGRAD && p->leg_kind == 1) return fail(DSHT_EUNSUPPORTED, "fused value/gradient inverse needs DSHT_TUNE_LEG != 1");
  const bool scaled = form >= 2;
  // the first (LDS-staged) split design keeps the BF16 forms only: its FP16 product-order
  // weighting loses up to ~2e-5 relative on zero-mean data (the fragment-stream kernels
  // hold the FP16 forms to their bounds)
  if (scaled && p->leg_kind == 1)
    return fail(DSHT_EUNSUPPORTED, "DSHT_TUNE_LEG=1 (first split design) supports the BF16 forms only");
  const int ch = mixed_chunk_cols(p, ncols);
  const int nchunks = cdiv(ncols, ch);
  const size_t gsz = (size_t)(p->N + 1) * p->J * fourier_pitch(ch);
  const size_t slot_bytes = nout * gsz * sizeof(float2);  // Fourier outputs of one chunk
  const int nslot = p->pipeline && nchunks > 1 ? 2 : 1;
  const size_t cmb = (2 * (size_t)ch * sizeof(unsigned) + 255) / 256 * 256;
  const size_t pkb = p->leg_kind == 0 ? mx2_inv_pack_bytes(t, form, op, ch) : 0;
  WsHandle wh;
  if (int rc = workspace_acquire(p, s, nslot * slot_bytes + cmb + pkb, wh)) return rc;
  ChunkPipe cp;
  if (int rc = pipe_begin(wh, s, p, nchunks, cp)) return rc;
  // column maxima and data panels are used by the contraction stage only (caller's stream)
  unsigned* colmax = (unsigned*)((char*)wh.buf + nslot * slot_bytes);
  uint16_t* pack = (uint16_t*)((char*)colmax + cmb);
  for (int k = 0, c0 = 0; c0 < ncols; k++, c0 += ch) {
    const int nc = ncols - c0 < ch ? ncols - c0 : ch;
    const int slot = cp.on ? (k & 1) : 0;
    float2* G0 = (float2*)((char*)wh.buf + slot * slot_bytes);
    float2* G1 = G0 + gsz;
    float2* G2 = G1 + gsz;
    // fused value/gradient: GradLon through the derivative inverse FFT of the value's
    // coefficients (i m X_m, 1/cos row scale); the contraction writes two outputs
    const bool dlon = op == I_VGRAD && out1 && fft_inv_deriv_ok(p);
    float2* O0 = out0 || dlon ? G0 : nullptr;
    float2* O1 = nout > 1 && out1 && !dlon ? G1 : nullptr;
    float2* O2 = nout > 2 && out2 ? G2 : nullptr;
    const double* w0 = in0 + (long)c0 * sc;
    const double* w1 = in1 ? in1 + (long)c0 * sc : nullptr;
    if (cp.on && k >= 2) DSHT_HIP(hipStreamWaitEvent(s, cp.evF[slot], 0));  // slot free again
    if (scaled) {
      hipLaunchKernelGGL(k_zero_u32, dim3(cdiv(2 * nc, 256)), dim3(256), 0, s, colmax, 2 * nc);
      DSHT_HIP(hipGetLastError());
      if (int rc = launch_spec_colmax(w0, sl, sc, p->L, nc, pro, colmax, s)) return rc;
      if (nin > 1)
        if (int rc = launch_spec_colmax(w1, sl, sc, p->L, nc, pro, colmax + nc, s)) return rc;
    }
    MxInvArgs a;
    a.W0 = w0;
    a.W1 = w1;
    a.sl = sl;
    a.sc = sc;
    a.G0 = G0;
    a.G1 = nout > 1 ? G1 : nullptr;
    a.ncols = nc;
    a.gld = fourier_pitch(nc);
    a.gmst = (long)p->J * a.gld;
    for (int q = 0; q < 3; q++) {
      a.tP[q] = t.iP[q];
      a.tH[q] = t.iH[q];
    }
    a.off = t.d_offI;
    a.minv = t.d_minv;
    a.colmax = colmax;
    a.rs = rs;
    a.pro = pro;
    a.N = p->N;
    a.J2 = p->J2;
    a.J2p32 = t.J2p32;
    int rc;
    {
      double by, fl;
      mixed_work(p, op, false, nc, &by, &fl, (O0 ? 1 : 0) + (O1 ? 1 : 0) + (O2 ? 1 : 0));
      ProfScope ps("leg_inv", s, by, fl);
      if (p->leg_kind != 1) {
        Mx2InvJob j;
        j.W0 = w0;
        j.W1 = w1;
        j.sl = sl;
        j.sc = sc;
        j.ncols = nc;
        j.pro = pro;
        j.rs = rs;
        j.colmax = colmax;
        j.pack = p->leg_kind == 0 ? pack : nullptr;
        j.staged = p->leg_kind == 3;
        j.G0 = O0;
        j.G1 = O1;
        j.G2 = O2;
        j.gld = fourier_pitch(nc);
        j.gmst = (long)p->J * j.gld;
        rc = mx2_inv(p, t, form, op, j, s);
      } else {
        switch (form) {
          case 0: rc = mx_inv_op<0>(p, op, a, s); break;
          case 1: rc = mx_inv_op<1>(p, op, a, s); break;
          case 2: rc = mx_inv_op<2>(p, op, a, s); break;
          default: rc = mx_inv_op<3>(p, op, a, s); break;
        }
      }
    }
    if (rc) return rc;
    hipStream_t fs = s;  // FFT stage
    if (cp.on) {
      fs = cp.aux;
      if ((rc = pipe_after(cp.evL[slot], s, fs))) return rc;
    }
    const long gp = fourier_pitch(nc);
    // one inverse FFT per output; the field and its longitude derivative (fused value/gradient
    // call) come from one read of G0
    auto inv = [&](const float2* G, double* out, const double* rsc, int deriv, double* out2d) -> int {
      if (grid_f32) {
        float* o = (float*)out + (long)c0 * ldg;
        float* o2 = out2d ? (float*)out2d + (long)c0 * ldg : nullptr;
        return launch_fft_inv32f(p, G, gp, nc, o, ldg, fs, rsc, deriv, o2);
      }
      return launch_fft_inv32(p, G, gp, nc, out + (long)c0 * ldg, ldg, fs, rsc, deriv,
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
    if (cp.on) DSHT_HIP(hipEventRecord(cp.evF[slot], fs));
  }
  return pipe_join(cp);
}

}  // namespace dsht
