// legendre_mx2.hip -- split-precision Legendre contractions on 16x16x32 MFMA.
//
// The Legendre half of SPMODEL w_xy / xy_w and their derivative forms
// (doc/description/ogcm/space_discretization.tex:199-268) in the split forms of
// legendre_mixed.hip.  Table fragments are pre-split and pre-swizzled (one 16-byte load per
// lane per term).  Two inverse kernels:
//  * packed: k_mx2_ipack first packs the spectra into split-term panels in MFMA B-fragment
//    order (prefactor, column scale and FP16 term scaling applied there); every wave then
//    streams table and panel fragments from memory, with no LDS and no barrier in its K loop;
//  * whole-K staged: a block covers one wavenumber, a column group and every latitude tile;
//    it gathers and splits its spectra once into LDS behind one barrier, and its waves read
//    the data fragments from there (no round trip of packed panels through memory).
// Forward: each block stages the symmetric and antisymmetric latitude combinations of its
// columns, split, through a double-buffered LDS panel, one K step at a time.
// Products of one output accumulate in one MFMA chain per (table, input) group; the i*m
// factors, signs, north/south recombination and unscaling happen once in the epilogue.
// Blocks are placed so that all blocks of one zonal wavenumber run on one XCD (shared
// table and data panels stay in that XCD's L2) and wavenumbers alternate across XCDs.
#include <algorithm>
#include <cmath>

#include "legendre_mixed.hpp"

namespace dsht {

namespace {

// B-operand terms and product list of each form.  FP16 forms keep the whole sum scaled by
// 2^11 (undone in the epilogue) so that no product needs a separate weight:
//   FP16    b -> fp16(2^11 b);                           products (a0, b0)
//   FP16X3  b -> {2^11 h, fp16(2^11 (b - h)), h}, h = fp16(b); products (a0,b0) (a0,b1) (a1,b2)
template <int FORM>
struct Mx2 {
  static constexpr int NTA = FORM == 1 ? 3 : (FORM == 2 ? 1 : 2);
  static constexpr int NTB = FORM == 1 ? 3 : (FORM == 2 ? 1 : (FORM == 3 ? 3 : 2));
  static constexpr int NP = FORM == 1 ? 6 : (FORM == 2 ? 1 : 3);
  static constexpr double DSCALE = (FORM >= 2) ? 1.0 / 2048.0 : 1.0;
  __host__ __device__ static constexpr int pa(int i) {
    return FORM == 3 ? (i == 2 ? 1 : 0) : (i == 2 || i == 4 ? 1 : (i == 5 ? 2 : 0));
  }
  __host__ __device__ static constexpr int pb(int i) {
    return FORM == 3 ? i : (i == 1 || i == 4 ? 1 : (i == 3 ? 2 : 0));
  }
};

template <int FORM>
__device__ __forceinline__ void bsplit(float x, typename MxForm<FORM>::T* out) {
  using T = typename MxForm<FORM>::T;
  if constexpr (FORM == 0 || FORM == 1) {
#pragma unroll
    for (int t = 0; t < Mx2<FORM>::NTB; t++) {
      const T h = (T)x;
      out[t] = h;
      x -= (float)h;
    }
  } else if constexpr (FORM == 2) {
    out[0] = (T)(2048.0f * x);
  } else {
    const T h = (T)x;
    out[0] = (T)(2048.0f * (float)h);
    out[1] = (T)(2048.0f * (x - (float)h));
    out[2] = h;
  }
}

// operator groups: each is one (table, input) product stream; table 0 = P, 1 = H
// I_VGRAD: (P, s) and (H, s)
__host__ __device__ constexpr int inv_ng(int op) {
  return (op == I_ALPHA || op == I_VGRAD) ? 2 : (op == I_VORDIV2UV ? 4 : 1);
}
__host__ __device__ constexpr int inv_tab(int op, int g) {
  return op == I_GRADMU ? 1
                        : ((op == I_ALPHA || op == I_VGRAD) ? (g == 1 ? 1 : 0) : (op == I_VORDIV2UV ? (g >= 2 ? 1 : 0) : 0));
}
__host__ __device__ constexpr int inv_in(int op, int g) {
  return op == I_ALPHA ? g : (op == I_VORDIV2UV ? (g & 1) : 0);
}
// forward UV2VORDIV groups: (P,V) (H,U) (P,U) (H,V)
__host__ __device__ constexpr int fwd_ng(int op) { return op == F_ALPHA ? 2 : (op == F_UV2VORDIV ? 4 : 1); }
__host__ __device__ constexpr int fwd_tab(int op, int g) {
  return op == F_DIVMU ? 1 : (op == F_ALPHA ? (g == 1 ? 1 : 0) : (op == F_UV2VORDIV ? (g & 1) : 0));
}
__host__ __device__ constexpr int fwd_in(int op, int g) {
  return op == F_ALPHA ? g : (op == F_UV2VORDIV ? ((g == 0 || g == 3) ? 1 : 0) : 0);
}

// block -> (m, sub-block): wavenumbers m = x, x+8, ... run on XCD x (blocks are dispatched
// to the 8 XCDs round robin), the CBLK blocks of one m are consecutive on that XCD
__device__ __forceinline__ bool xcd_decode(int N, int cblk, int& m, int& sub) {
  const int b = blockIdx.x;
  const int idx = b >> 3;
  m = (b & 7) + 8 * (idx / cblk);
  sub = idx % cblk;
  return m <= N;
}

__device__ __forceinline__ float colscale_of(float mx, float bound) {
  mx *= bound;
  return mx > 0.f ? exp2f(-ceilf(log2f(mx))) : 1.f;
}
template <int NIN>
__device__ __forceinline__ float colscale(const unsigned* colmax, int ncols, int c, float bound) {
  float mx = 0.f;
#pragma unroll
  for (int i = 0; i < NIN; i++) mx = fmaxf(mx, __uint_as_float(colmax[(long)i * ncols + c]));
  return colscale_of(mx, bound);
}

struct Mx2InvArgs {
  const double* W0;
  const double* W1;
  long sl, sc;
  int ncols, CT;
  const double* pro;
  const double* rs;
  const unsigned* colmax;
  uint16_t* pack;
  float2* G0;
  float2* G1;
  float2* G2;  // I_VGRAD GradLat; outputs of I_VGRAD may be NULL (not stored)
  long gld, gmst;
  const uint16_t* tP[3];
  const uint16_t* tH[3];
  const long* off;
  const int* cbase;
  const int* chunk_m;
  int nchunks;
  const double* minv;
  int N, J2, J2p32, RT;
  int CG, RG;  // column groups and row groups per wavenumber
};

// ---------------------------------------------------------------- inverse data panels
// panel (chunk g, column tile ct, input i): [re, im][term][lane][8], lane = c + 16*(k/8).
// Threads of one (chunk, input) run along the columns (64 consecutive columns per wave, so
// every gathered coefficient row is read in 512-byte runs); thread (k-group q, column c)
// fills fragment lane c%16 + 16 q of tile c/16 with the 8 wavenumbers of its group.
template <int FORM, int NIN>
__global__ __launch_bounds__(256) void k_mx2_ipack(Mx2InvArgs a) {
  using F = MxForm<FORM>;
  using T = typename F::T;
  using V8 = typename F::V8;
  constexpr int NTB = Mx2<FORM>::NTB;
  const int CC = a.CT * TILE;  // padded columns
  const long per = 4l * CC;    // threads per (chunk, input)
  const long gt = (long)blockIdx.x * 256 + threadIdx.x;
  if (gt >= per * a.nchunks * NIN) return;
  const int gi = (int)(gt / per), rem = (int)(gt % per);
  const int i = gi % NIN, g = gi / NIN;
  const int q = rem / CC, c = rem % CC, ct = c / TILE;
  const int m = a.chunk_m[g], ch = g - a.cbase[m];
  const int nE = n_par(a.N, m, 0), nO = n_par(a.N, m, 1);
  const int nch0 = kpad32(nE) / KM;
  const int p = ch >= nch0 ? 1 : 0, ks = p ? ch - nch0 : ch, np = p ? nO : nE;
  const double* W = i ? a.W1 : a.W0;
  float dsc = 1.f;
  if constexpr (F::SCALED) {
    if (c < a.ncols) dsc = colscale<NIN>(a.colmax, a.ncols, c, 1.f);
  }
  // all 16 loads are issued before any is used: unconditional (a padding element reads
  // element 0) with the predicate and the prefactor applied afterwards
  double vre[8], vim[8], pr[8];
  bool ok[8];
  long lp[8];
#pragma unroll
  for (int e = 0; e < 8; e++) {
    const int k = ks * KM + 8 * q + e;
    const int n = m + p + 2 * k;
    ok[e] = k < np && c < a.ncols;
    const long l0 = ok[e] ? (long)n * (n + 1) : 0, co = ok[e] ? (long)c * a.sc : 0;
    lp[e] = ok[e] ? l0 + m : 0;
    // the spectra are read once here: non-temporal
    vre[e] = __builtin_nontemporal_load(W + (ok[e] ? l0 + m : 0) * a.sl + co);
    vim[e] = __builtin_nontemporal_load(W + (ok[e] ? l0 - m : 0) * a.sl + co);
    pr[e] = 1.0;
  }
  if (a.pro) {
#pragma unroll
    for (int e = 0; e < 8; e++) pr[e] = a.pro[lp[e]];
  }
#pragma unroll
  for (int e = 0; e < 8; e++) {
    vre[e] = ok[e] ? pr[e] * vre[e] : 0.0;
    vim[e] = ok[e] && m > 0 ? pr[e] * vim[e] : 0.0;
  }
  T tre[NTB][8], tim[NTB][8];
#pragma unroll
  for (int e = 0; e < 8; e++) {
    T t1[NTB], t2[NTB];
    bsplit<FORM>((float)vre[e] * dsc, t1);
    bsplit<FORM>((float)vim[e] * dsc, t2);
#pragma unroll
    for (int t = 0; t < NTB; t++) {
      tre[t][e] = t1[t];
      tim[t][e] = t2[t];
    }
  }
  const int lane = (c % TILE) + TILE * q;
  uint16_t* dst = a.pack + ((((long)g * a.CT + ct) * NIN + i) * 2 * NTB) * 512 + lane * 8;
#pragma unroll
  for (int t = 0; t < NTB; t++) {
    V8 v1, v2;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      v1[e] = tre[t][e];
      v2[e] = tim[t][e];
    }
    *reinterpret_cast<V8*>(dst + (long)t * 512) = v1;
    *reinterpret_cast<V8*>(dst + (long)(NTB + t) * 512) = v2;
  }
}

// ---- inverse epilogue: recombine north/south, i*m factors, unscale, row scale, FP32 store
template <int OP, int FORM, int NCT, int MAXT>
__device__ __forceinline__ void inv_epilogue(const Mx2InvArgs& a,
                                             const float4_t (&acc)[inv_ng(OP)][MAXT][NCT][2][2], int m,
                                             int t0, int ct0, int lane) {
  using F = MxForm<FORM>;
  using X = Mx2<FORM>;
  constexpr int NG = inv_ng(OP), NIN = InvTraits<OP>::NIN, NOUT = InvTraits<OP>::NOUT;
  const double mi = a.minv[m] * X::DSCALE;
  const double dm = (double)m;
  // The lane's row scales (north and south) and column scales are loaded before any is used,
  // so that they cost one round trip rather than one per row (an out-of-range row or column
  // reads entry 0; its results are not stored).
  double rsN[MAXT][4], rsS[MAXT][4];
#pragma unroll
  for (int tt = 0; tt < MAXT; tt++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int jr = (t0 + tt) * TILE + acc_row16(lane, i);
      const bool ok = jr < a.J2;
      rsN[tt][i] = rsS[tt][i] = 1.0;
      if (a.rs) {
        rsN[tt][i] = a.rs[ok ? a.J2 + jr : 0];
        rsS[tt][i] = a.rs[ok ? a.J2 - 1 - jr : 0];
      }
    }
  double us[NCT];
#pragma unroll
  for (int ct = 0; ct < NCT; ct++) us[ct] = mi;
  if constexpr (F::SCALED) {
    unsigned cm[NCT][NIN];
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
      const int c = (ct0 + ct) * TILE + (lane & 15);
#pragma unroll
      for (int i = 0; i < NIN; i++) cm[ct][i] = a.colmax[(long)i * a.ncols + (c < a.ncols ? c : 0)];
    }
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
      float mx = 0.f;
#pragma unroll
      for (int i = 0; i < NIN; i++) mx = fmaxf(mx, __uint_as_float(cm[ct][i]));
      us[ct] = mi / (double)colscale_of(mx, 1.f);
    }
  }
#pragma unroll
  for (int tt = 0; tt < MAXT; tt++) {
    const int t = t0 + tt;
    if (t >= a.RT) continue;
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
      const int c = (ct0 + ct) * TILE + (lane & 15);
      if (c >= a.ncols) continue;
      const double u = us[ct];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int jr = t * TILE + acc_row16(lane, i);
        if (jr >= a.J2) continue;
        const int jN = a.J2 + jr, jS = a.J2 - 1 - jr;
        // per group: north and south complex values
        double nr[NG], ni[NG], sr[NG], si[NG];
#pragma unroll
        for (int gi = 0; gi < NG; gi++) {
          const double a0 = acc[gi][tt][ct][0][0][i], a1 = acc[gi][tt][ct][0][1][i];
          const double b0 = acc[gi][tt][ct][1][0][i], b1 = acc[gi][tt][ct][1][1][i];
          nr[gi] = a0 + a1;
          sr[gi] = a0 - a1;
          ni[gi] = b0 + b1;
          si[gi] = b0 - b1;
        }
        double onr[NOUT], oni[NOUT], osr[NOUT], osi[NOUT];
        if constexpr (OP == I_VAL || OP == I_GRADMU) {
          onr[0] = nr[0], oni[0] = ni[0], osr[0] = sr[0], osi[0] = si[0];
        } else if constexpr (OP == I_GRADLAMBDA) {  // i m X
          onr[0] = -dm * ni[0], oni[0] = dm * nr[0], osr[0] = -dm * si[0], osi[0] = dm * sr[0];
        } else if constexpr (OP == I_VGRAD) {  // X, i m X, H-group
          onr[0] = nr[0], oni[0] = ni[0], osr[0] = sr[0], osi[0] = si[0];
          onr[1] = -dm * ni[0], oni[1] = dm * nr[0], osr[1] = -dm * si[0], osi[1] = dm * sr[0];
          onr[2] = nr[1], oni[2] = ni[1], osr[2] = sr[1], osi[2] = si[1];
        } else if constexpr (OP == I_ALPHA) {  // i m P.A + H.B
          onr[0] = -dm * ni[0] + nr[1], oni[0] = dm * nr[0] + ni[1];
          osr[0] = -dm * si[0] + sr[1], osi[0] = dm * sr[0] + si[1];
        } else {  // groups (P,psi) (P,chi) (H,psi) (H,chi): U = i m P.chi - H.psi, V = i m P.psi + H.chi
          onr[0] = -dm * ni[1] - nr[2], oni[0] = dm * nr[1] - ni[2];
          osr[0] = -dm * si[1] - sr[2], osi[0] = dm * sr[1] - si[2];
          onr[1] = -dm * ni[0] + nr[3], oni[1] = dm * nr[0] + ni[3];
          osr[1] = -dm * si[0] + sr[3], osi[1] = dm * sr[0] + si[3];
        }
        const double sN = u * rsN[tt][i], sS = u * rsS[tt][i];
#pragma unroll
        for (int o = 0; o < NOUT; o++) {
          float2* Gb = o == 0 ? a.G0 : (o == 1 ? a.G1 : a.G2);
          if (OP == I_VGRAD && !Gb) continue;
          // the value output of I_VGRAD carries no row scale
          const double fN = (OP == I_VGRAD && o == 0) ? u : sN, fS = (OP == I_VGRAD && o == 0) ? u : sS;
          float2* G = Gb + (long)m * a.gmst;
          G[(long)jN * a.gld + c] = make_float2((float)(onr[o] * fN), m ? (float)(oni[o] * fN) : 0.f);
          G[(long)jS * a.gld + c] = make_float2((float)(osr[o] * fS), m ? (float)(osi[o] * fS) : 0.f);
        }
      }
    }
  }
}

// ---------------------------------------------------------------- inverse contraction
// block = 4 waves of one wavenumber m and one group of NCT column tiles; wave w owns the
// MAXT latitude tiles starting at (rg*4 + w)*MAXT.  K runs over the parity-0 chunks, then
// the parity-1 chunks; P of parity p accumulates into sym (p = 0) or anti (p = 1), H the
// opposite; north = sym + anti, south = sym - anti.  Table and packed data fragments stream
// from memory (one 16-byte load per lane and term); a tile or column tile past the last one
// repeats the last (in bounds; its results are not stored).
template <int OP, int FORM, int NCT, int MAXT>
__global__ __launch_bounds__(256, 2) void k_mx2_inv(Mx2InvArgs a) {
  using F = MxForm<FORM>;
  using X = Mx2<FORM>;
  using V8 = typename F::V8;
  constexpr int NG = inv_ng(OP), NIN = InvTraits<OP>::NIN;
  constexpr int NTA = X::NTA, NTB = X::NTB;
  constexpr bool UP = InvTraits<OP>::USE_P, UH = InvTraits<OP>::USE_H;
  int m, sub;
  if (!xcd_decode(a.N, a.CG * a.RG, m, sub)) return;
  const int cg = sub % a.CG, rg = sub / a.CG;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int t0 = (rg * 4 + w) * MAXT;
  if (t0 >= a.RT) return;
  const int nE = n_par(a.N, m, 0), nO = n_par(a.N, m, 1);
  const int kq0 = kpad32(nE), kq1 = kpad32(nO);
  const int nch0 = kq0 / KM, nch1 = kq1 / KM;
  const long toff = a.off[m];
  const int gc0 = a.cbase[m];
  const int ct0 = cg * NCT;

  float4_t acc[NG][MAXT][NCT][2][2];
#pragma unroll
  for (int g = 0; g < NG; g++)
#pragma unroll
    for (int tt = 0; tt < MAXT; tt++)
#pragma unroll
      for (int ct = 0; ct < NCT; ct++)
#pragma unroll
        for (int r = 0; r < 2; r++) acc[g][tt][ct][r][0] = acc[g][tt][ct][r][1] = float4_t{0, 0, 0, 0};

  auto run = [&](auto PAR) {
    constexpr int PP = decltype(PAR)::value;
    const int nch = PP ? nch1 : nch0;
    const long pbase = toff + (PP ? (long)a.J2p32 * kq0 : 0);
    for (int ks = 0; ks < nch; ks++) {
      V8 A[2][MAXT][NTA];
#pragma unroll
      for (int tt = 0; tt < MAXT; tt++) {
        const int t = t0 + tt < a.RT ? t0 + tt : a.RT - 1;
        const long o = pbase + ((long)t * nch + ks) * 512 + lane * 8;
#pragma unroll
        for (int k = 0; k < NTA; k++) {
          if constexpr (UP) A[0][tt][k] = *reinterpret_cast<const V8*>(a.tP[k] + o);
          if constexpr (UH) A[1][tt][k] = *reinterpret_cast<const V8*>(a.tH[k] + o);
        }
      }
      V8 B[NIN][NCT][2][NTB];
      const long g = gc0 + (PP ? nch0 : 0) + ks;
#pragma unroll
      for (int i = 0; i < NIN; i++)
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) {
          const int ctg = ct0 + ct < a.CT ? ct0 + ct : a.CT - 1;
          const uint16_t* src = a.pack + (((g * a.CT + ctg) * NIN + i) * 2 * NTB) * 512 + lane * 8;
#pragma unroll
          for (int r = 0; r < 2; r++)
#pragma unroll
            for (int t = 0; t < NTB; t++) B[i][ct][r][t] = *reinterpret_cast<const V8*>(src + (long)(r * NTB + t) * 512);
        }
#pragma unroll
      for (int gi = 0; gi < NG; gi++) {
        const int tb = inv_tab(OP, gi), in = inv_in(OP, gi);
        const int sidx = tb == 0 ? PP : 1 - PP;
#pragma unroll
        for (int tt = 0; tt < MAXT; tt++)
#pragma unroll
          for (int ct = 0; ct < NCT; ct++)
#pragma unroll
            for (int r = 0; r < 2; r++)
#pragma unroll
              for (int pr = 0; pr < X::NP; pr++)
                acc[gi][tt][ct][r][sidx] =
                    mfma16(A[tb][tt][X::pa(pr)], B[in][ct][r][X::pb(pr)], acc[gi][tt][ct][r][sidx]);
      }
    }
  };
  run(std::integral_constant<int, 0>());
  run(std::integral_constant<int, 1>());

  inv_epilogue<OP, FORM, NCT, MAXT>(a, acc, m, t0, ct0, lane);
}

// Inverse with the whole K range of the block's spectra staged once in LDS.  A block covers
// one wavenumber m, NCT column tiles and every latitude tile (wave w owns MAXT of them), so
// each FP64 coefficient is read from memory once in the whole launch: the block's threads
// gather the coefficients of all its K chunks (consecutive threads on consecutive columns),
// split them into their terms and store them as B fragments in LDS behind one barrier; the
// waves then run their K loops (parity 0 chunks, then parity 1) with table fragments from
// memory and data fragments from LDS, and no further barrier.  Same accumulators and
// epilogue as k_mx2_inv.
extern __shared__ __attribute__((aligned(16))) unsigned char mx2_lds[];

template <int OP, int FORM, int NCT>
__host__ __device__ constexpr int wk_chunk_elems() {
  return InvTraits<OP>::NIN * NCT * 2 * Mx2<FORM>::NTB * 512;
}

template <int OP, int FORM, int NCT, int MAXT>
__global__ __launch_bounds__(256, 2) void k_mx2_inv_wk(Mx2InvArgs a) {
  using F = MxForm<FORM>;
  using X = Mx2<FORM>;
  using T = typename F::T;
  using V8 = typename F::V8;
  constexpr int NG = inv_ng(OP), NIN = InvTraits<OP>::NIN;
  constexpr int NTA = X::NTA, NTB = X::NTB;
  constexpr bool UP = InvTraits<OP>::USE_P, UH = InvTraits<OP>::USE_H;
  constexpr int CB = NCT * TILE;             // columns of the block
  constexpr int IPC = NIN * CB * 4 * 2;      // staging items per chunk: (input, column, k-group, re/im)
  constexpr int CHUNK = wk_chunk_elems<OP, FORM, NCT>();
  T* sB = reinterpret_cast<T*>(mx2_lds);
  int m, sub;
  if (!xcd_decode(a.N, a.CG * a.RG, m, sub)) return;  // uniform over the block
  const int cg = sub % a.CG, rg = sub / a.CG;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int t0 = (rg * 4 + w) * MAXT;
  const int nE = n_par(a.N, m, 0), nO = n_par(a.N, m, 1);
  const int kq0 = kpad32(nE), kq1 = kpad32(nO);
  const int nch0 = kq0 / KM, nch1 = kq1 / KM, nch = nch0 + nch1;
  const long toff = a.off[m];
  const int ct0 = cg * NCT;

  // ---- staging: item it -> chunk it / IPC; within the chunk (column cc, k-group kg, re/im r,
  // input i) with cc fastest, so a wave's loads of one coefficient row are contiguous
  {
    const int cc = tid % CB;  // IPC and 256 are multiples of CB: a thread keeps its column
    const int c = ct0 * TILE + cc;
    const bool cok = c < a.ncols;
    float dsc = 1.f;
    if constexpr (F::SCALED) {
      if (cok) dsc = colscale<NIN>(a.colmax, a.ncols, c, 1.f);
    }
    const long coff = (long)c * a.sc;
    const int total = nch * IPC;
    constexpr int U = 2;  // items per thread in flight (deeper prefetch measured no faster)
    for (int base = tid; base < total; base += 256 * U) {
      // All loads of a round are issued before any is used: they are unconditional (a
      // padding element reads element 0, which always exists) and the predicate is applied
      // afterwards.  A load whose value is used inside the branch that guards it is waited
      // for inside that branch, which serializes the round trips.
      double v[U][8], pr[U][8];
      bool ok[U][8];
      long lp[U][8];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int it = base + 256 * u;
        const int ch = it / IPC, rest = (it % IPC) / CB;
        const int kg = rest & 3, r = (rest >> 2) & 1, i = rest >> 3;
        const int p = ch >= nch0 ? 1 : 0, ks = p ? ch - nch0 : ch, np = p ? nO : nE;
        const double* W = i ? a.W1 : a.W0;
#pragma unroll
        for (int e = 0; e < 8; e++) {
          const int k = ks * KM + 8 * kg + e;
          const int n = m + p + 2 * k;
          const long l0 = (long)n * (n + 1);
          ok[u][e] = it < total && cok && k < np && (r == 0 || m > 0);
          lp[u][e] = ok[u][e] ? l0 + m : 0;
          v[u][e] = __builtin_nontemporal_load(W + (ok[u][e] ? (l0 + (r ? -m : m)) * a.sl + coff : 0));
          pr[u][e] = 1.0;
        }
      }
      if (a.pro) {
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
          for (int e = 0; e < 8; e++) pr[u][e] = a.pro[lp[u][e]];
      }
#pragma unroll
      for (int u = 0; u < U; u++)
#pragma unroll
        for (int e = 0; e < 8; e++) v[u][e] = ok[u][e] ? pr[u][e] * v[u][e] : 0.0;
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int it = base + 256 * u;
        if (it >= total) continue;
        const int ch = it / IPC, rest = (it % IPC) / CB;
        const int kg = rest & 3, r = (rest >> 2) & 1, i = rest >> 3;
        T tt[NTB][8];
#pragma unroll
        for (int e = 0; e < 8; e++) {
          T tmp[NTB];
          bsplit<FORM>((float)v[u][e] * dsc, tmp);
#pragma unroll
          for (int t = 0; t < NTB; t++) tt[t][e] = tmp[t];
        }
        const int fl = (cc % TILE) + TILE * kg;
        T* dst = sB + (long)ch * CHUNK + ((((i * NCT + cc / TILE) * 2 + r) * NTB) * 512) + fl * 8;
#pragma unroll
        for (int t = 0; t < NTB; t++) {
          V8 vv;
#pragma unroll
          for (int e = 0; e < 8; e++) vv[e] = tt[t][e];
          *reinterpret_cast<V8*>(dst + t * 512) = vv;
        }
      }
    }
  }
  __syncthreads();
  if (t0 >= a.RT) return;  // no barrier follows

  float4_t acc[NG][MAXT][NCT][2][2];
#pragma unroll
  for (int g = 0; g < NG; g++)
#pragma unroll
    for (int tt = 0; tt < MAXT; tt++)
#pragma unroll
      for (int ct = 0; ct < NCT; ct++)
#pragma unroll
        for (int r = 0; r < 2; r++) acc[g][tt][ct][r][0] = acc[g][tt][ct][r][1] = float4_t{0, 0, 0, 0};

  auto run = [&](auto PAR) {
    constexpr int PP = decltype(PAR)::value;
    const int nchp = PP ? nch1 : nch0;
    const long pbase = toff + (PP ? (long)a.J2p32 * kq0 : 0);
    for (int ks = 0; ks < nchp; ks++) {
      V8 A[2][MAXT][NTA];
#pragma unroll
      for (int tt = 0; tt < MAXT; tt++) {
        const int t = t0 + tt;
        const long o = pbase + ((long)t * nchp + ks) * 512 + lane * 8;
#pragma unroll
        for (int k = 0; k < NTA; k++) {
          if constexpr (UP) A[0][tt][k] = t < a.RT ? *reinterpret_cast<const V8*>(a.tP[k] + o) : V8{};
          if constexpr (UH) A[1][tt][k] = t < a.RT ? *reinterpret_cast<const V8*>(a.tH[k] + o) : V8{};
        }
      }
      const T* panel = sB + (long)(PP ? nch0 + ks : ks) * CHUNK + lane * 8;
#pragma unroll
      for (int ct = 0; ct < NCT; ct++)
#pragma unroll
        for (int r = 0; r < 2; r++) {
          V8 B[NIN][NTB];
#pragma unroll
          for (int i = 0; i < NIN; i++)
#pragma unroll
            for (int t = 0; t < NTB; t++)
              B[i][t] = *reinterpret_cast<const V8*>(panel + (((i * NCT + ct) * 2 + r) * NTB + t) * 512);
#pragma unroll
          for (int gi = 0; gi < NG; gi++) {
            const int tb = inv_tab(OP, gi), in = inv_in(OP, gi);
            const int sidx = tb == 0 ? PP : 1 - PP;
#pragma unroll
            for (int tt = 0; tt < MAXT; tt++)
#pragma unroll
              for (int pr = 0; pr < X::NP; pr++)
                acc[gi][tt][ct][r][sidx] = mfma16(A[tb][tt][X::pa(pr)], B[in][X::pb(pr)], acc[gi][tt][ct][r][sidx]);
          }
        }
    }
  };
  run(std::integral_constant<int, 0>());
  run(std::integral_constant<int, 1>());
  inv_epilogue<OP, FORM, NCT, MAXT>(a, acc, m, t0, ct0, lane);
}

// ---------------------------------------------------------------- forward contraction
struct Mx2FwdArgs {
  const float2* G0;
  const float2* G1;
  long gld, gmst;
  int ncols, CT;
  double* W0;
  double* W1;
  long sl, sc;
  const float* cw;
  const unsigned* colmax;
  float bound;
  const double* diag0;
  const double* diag1;
  double scale;
  const uint16_t* tP[3];
  const uint16_t* tH[3];
  const long* off;
  const double* minv;
  int N, J2, J2p32;
};

// Forward with the split data shared through LDS.  The block's 256 threads form the symmetric
// and antisymmetric combinations of one 32-latitude K step once (two latitudes of one column
// each), split them and store both combinations' B fragments to a double-buffered LDS panel
// (one barrier per K step).  K is the outer loop: each wave keeps the accumulators of up to
// NSW slots (slot s = round*4*NSW + w + 4*i; a slot is MAXT 16-row tiles of one parity) and
// reads, per slot, the combination its parity needs (P with sym at parity 0, with anti at
// parity 1; H the opposite).
//
// KSC > 0: the number of K steps (J2p32 / 32) as a constant, so that the K loop unrolls into
// straight-line code.  Then one-group operators, in a round where no wave has a second slot,
// request their slot's table fragments of the next K step right after issuing this step's
// products: the next step's products find them loaded, and the compiler's waits (exact in
// straight-line code) leave the younger Fourier loads in flight.
template <int OP, int FORM, int MAXT, int NSW, int NCT, int KSC = 0>
__global__ __launch_bounds__(256, fwd_ng(OP) == 1 ? 2 : 1) void k_mx2_fwd_sh(Mx2FwdArgs a) {
  using F = MxForm<FORM>;
  using X = Mx2<FORM>;
  using T = typename F::T;
  using V8 = typename F::V8;
  constexpr int NG = fwd_ng(OP), NIN = FwdTraits<OP>::NIN, NOUT = FwdTraits<OP>::NOUT;
  constexpr int NTA = X::NTA, NTB = X::NTB;
  constexpr bool UP = FwdTraits<OP>::USE_P, UH = FwdTraits<OP>::USE_H;
  constexpr int FRAG = 64 * 8;                    // elements of one B fragment (64 lanes x 8)
  constexpr int PANEL = NIN * 2 * 2 * NTB * FRAG;  // [input][sym, anti][re, im][term] fragments
  __shared__ __attribute__((aligned(16))) T sB[2][NCT * PANEL];  // [buffer][column tile][panel]
  int m, cg;
  if (!xcd_decode(a.N, cdiv(a.CT, NCT), m, cg)) return;  // uniform over the block
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int nE = n_par(a.N, m, 0), nO = n_par(a.N, m, 1);
  const int tE = cdiv(nE, TILE), tO = cdiv(nO, TILE);
  const int sE = cdiv(tE, MAXT), sO = cdiv(tO, MAXT);
  const int KS = KSC ? KSC : a.J2p32 / KM;
  const long toff = a.off[m];
  const float2* Gm[2] = {a.G0 + (long)m * a.gmst, NIN > 1 ? a.G1 + (long)m * a.gmst : nullptr};
  const double dm = (double)m;
  // staging role: column scol of each of the block's NCT tiles, latitudes 2*jp and 2*jp + 1
  // of the K step; they are elements e0, e0 + 1 of fragment lane fl
  const int scol = tid & 15, jp = tid >> 4;
  const int fl = scol + 16 * (jp >> 2), e0 = 2 * (jp & 3);
  int cs[NCT];
  bool csok[NCT];
  float sdsc[NCT];
#pragma unroll
  for (int t = 0; t < NCT; t++) {
    cs[t] = (cg * NCT + t) * TILE + scol;
    csok[t] = cs[t] < a.ncols;
    sdsc[t] = 1.f;
    if constexpr (F::SCALED) {
      if (csok[t]) sdsc[t] = colscale<NIN>(a.colmax, a.ncols, cs[t], a.bound);
    }
  }

  // staging is split in two so that the next K step's loads are in flight during this
  // step's MFMAs: fetch() loads the north/south values, stage() combines, splits and stores.
  // fetch() loads unconditionally (a latitude past the last one or a column past the last
  // one reads latitude 0 of column 0) and stage() zeroes those values: with no load inside
  // a branch, a wait for the table fragments does not also wait for these loads.
  float2 gN[NCT][NIN][2], gS[NCT][NIN][2];
  float gw[2];
  bool gok[NCT][2];
  auto fetch = [&](int ks) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int jn = ks * KM + 2 * jp + h;
      gw[h] = a.cw[jn];
#pragma unroll
      for (int t = 0; t < NCT; t++) {
        gok[t][h] = jn < a.J2 && csok[t];
        const long oN = gok[t][h] ? (long)(a.J2 + jn) * a.gld + cs[t] : 0;
        const long oS = gok[t][h] ? (long)(a.J2 - 1 - jn) * a.gld + cs[t] : 0;
#pragma unroll
        for (int i = 0; i < NIN; i++) {
          gN[t][i][h] = Gm[i][oN];
          gS[t][i][h] = Gm[i][oS];
        }
      }
    }
  };
  auto stage = [&](int buf, int) {
#pragma unroll
    for (int t = 0; t < NCT; t++)
#pragma unroll
      for (int i = 0; i < NIN; i++) {
        float v[2][4];  // [latitude][sym re, sym im, anti re, anti im]
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const float2 z = make_float2(0.f, 0.f);
          const float2 gn = gok[t][h] ? gN[t][i][h] : z, gs = gok[t][h] ? gS[t][i][h] : z;
          const float cw = gw[h] * sdsc[t];
          v[h][0] = cw * (gn.x + gs.x);
          v[h][1] = cw * (gn.y + gs.y);
          v[h][2] = cw * (gn.x - gs.x);
          v[h][3] = cw * (gn.y - gs.y);
        }
#pragma unroll
        for (int qr = 0; qr < 4; qr++) {  // qr = 2*q + r
          T t0[NTB], t1[NTB];
          bsplit<FORM>(v[0][qr], t0);
          bsplit<FORM>(v[1][qr], t1);
#pragma unroll
          for (int k = 0; k < NTB; k++) {
            T* dst = sB[buf] + t * PANEL + ((i * 4 + qr) * NTB + k) * FRAG + fl * 8 + e0;
            dst[0] = t0[k];
            dst[1] = t1[k];
          }
        }
      }
  };


  const int nslot = sE + sO;
  for (int rbase = 0; rbase < nslot; rbase += 4 * NSW) {
    float4_t acc[NSW][NG][MAXT][NCT][2];
#pragma unroll
    for (int q = 0; q < NSW; q++)
#pragma unroll
      for (int g = 0; g < NG; g++)
#pragma unroll
        for (int tt = 0; tt < MAXT; tt++)
#pragma unroll
          for (int t = 0; t < NCT; t++) acc[q][g][tt][t][0] = acc[q][g][tt][t][1] = float4_t{0, 0, 0, 0};
    fetch(0);
    stage(0, 0);
    __syncthreads();
    // one-group operators request the next K step's values after the first slot's table
    // fragments (below); the multi-group ones, whose accumulators leave no registers for
    // both, request them first
    constexpr bool LATE = NG == 1;
    const bool pre = LATE && KSC > 0 && nslot - rbase <= 4;  // uniform over the block
    auto load_tables = [&](int ks, int sidx, V8 (&A)[2][MAXT][NTA]) {
      const int PP = sidx >= sE ? 1 : 0;
      const int tp0 = (PP ? sidx - sE : sidx) * MAXT, tpar = PP ? tO : tE;
#pragma unroll
      for (int tt = 0; tt < MAXT; tt++) {
        // a tile past the slot's last one repeats the last (its results are not stored)
        const int tp = tp0 + tt < tpar ? tp0 + tt : tpar - 1;
        const long o = toff + ((long)(PP ? tE + tp : tp) * KS + ks) * 512 + lane * 8;
#pragma unroll
        for (int k = 0; k < NTA; k++) {
          if constexpr (UP) A[0][tt][k] = *reinterpret_cast<const V8*>(a.tP[k] + o);
          if constexpr (UH) A[1][tt][k] = *reinterpret_cast<const V8*>(a.tH[k] + o);
        }
      }
    };
    V8 A0k[2][MAXT][NTA];  // slot 0's fragments kept across K steps (KSC > 0)
    if (pre && rbase + w < nslot) load_tables(0, rbase + w, A0k);
#pragma unroll(KSC ? KSC : 1)
    for (int ks = 0; ks < KS; ks++) {
      const int buf = ks & 1;
      if (!LATE && ks + 1 < KS) fetch(ks + 1);
#pragma unroll
      for (int q = 0; q < NSW; q++) {
        const int sidx = rbase + w + 4 * q;
        const bool act = sidx < nslot;  // uniform over the wave
        const int PP = sidx >= sE ? 1 : 0;
        auto products = [&](const V8 (&A)[2][MAXT][NTA]) {
#pragma unroll
          for (int gi = 0; gi < NG; gi++) {
            const int tb = fwd_tab(OP, gi), in = fwd_in(OP, gi);
            const int cq = tb == 0 ? PP : 1 - PP;  // 0 = sym, 1 = anti
#pragma unroll
            for (int t = 0; t < NCT; t++)
#pragma unroll
              for (int r = 0; r < 2; r++) {
                V8 B[NTB];
#pragma unroll
                for (int k = 0; k < NTB; k++)
                  B[k] = *reinterpret_cast<const V8*>(sB[buf] + t * PANEL + ((in * 4 + 2 * cq + r) * NTB + k) * FRAG +
                                                      lane * 8);
#pragma unroll
                for (int tt = 0; tt < MAXT; tt++)
#pragma unroll
                  for (int pr = 0; pr < X::NP; pr++)
                    acc[q][gi][tt][t][r] = mfma16(A[tb][tt][X::pa(pr)], B[X::pb(pr)], acc[q][gi][tt][t][r]);
              }
          }
        };
        if constexpr (KSC > 0) {
          // slot 0's fragments live across K steps; with pre they were requested a step ahead
          V8 A1[2][MAXT][NTA];
          auto& A = q == 0 ? A0k : A1;
          if (act && !(q == 0 && pre)) load_tables(ks, sidx, A);
          // requested after the first slot's table fragments, that slot's products wait for
          // its tables only
          if (LATE && q == 0 && ks + 1 < KS) fetch(ks + 1);
          if (!act) continue;
          products(A);
          if (q == 0 && pre && ks + 1 < KS) load_tables(ks + 1, sidx, A0k);
        } else {
          V8 A[2][MAXT][NTA];
          if (act) load_tables(ks, sidx, A);
          if (LATE && q == 0 && ks + 1 < KS) fetch(ks + 1);
          if (!act) continue;
          products(A);
        }
      }
      if (ks + 1 < KS) stage(buf ^ 1, ks + 1);
      __syncthreads();  // buf ^ 1 is complete; buf is restaged only after this barrier
    }
    // ---- epilogue (FP64), per slot and column tile: combine, unscale, diagonal, scatter
#pragma unroll
    for (int t = 0; t < NCT; t++) {
      const int c = (cg * NCT + t) * TILE + (lane & 15);
      if (c >= a.ncols) continue;
      double u = a.minv[m] * X::DSCALE * a.scale;
      if constexpr (F::SCALED) u /= (double)colscale<NIN>(a.colmax, a.ncols, c, a.bound);
#pragma unroll
      for (int q = 0; q < NSW; q++) {
        const int sidx = rbase + w + 4 * q;
        if (sidx >= nslot) continue;
        const int PP = sidx >= sE ? 1 : 0;
        const int tp0 = (PP ? sidx - sE : sidx) * MAXT, tpar = PP ? tO : tE;
#pragma unroll
        for (int tt = 0; tt < MAXT; tt++) {
          const int tp = tp0 + tt;
          if (tp >= tpar) continue;
#pragma unroll
          for (int i = 0; i < 4; i++) {
            const int k = tp * TILE + acc_row16(lane, i);
            const int n = m + PP + 2 * k;
            if (n > a.N) continue;
            double xr[NG], xi[NG];
#pragma unroll
            for (int gi = 0; gi < NG; gi++) {
              xr[gi] = acc[q][gi][tt][t][0][i];
              xi[gi] = acc[q][gi][tt][t][1][i];
            }
            double orr[NOUT], oii[NOUT];
            if constexpr (OP == F_VAL) {
              orr[0] = xr[0], oii[0] = xi[0];
            } else if constexpr (OP == F_DIVLAMBDA) {  // i m P.G
              orr[0] = -dm * xi[0], oii[0] = dm * xr[0];
            } else if constexpr (OP == F_DIVMU) {  // -H.G
              orr[0] = -xr[0], oii[0] = -xi[0];
            } else if constexpr (OP == F_ALPHA) {  // i m P.A - H.B
              orr[0] = -dm * xi[0] - xr[1], oii[0] = dm * xr[0] - xi[1];
            } else {  // vor = i m P.V + H.U ; div = i m P.U - H.V
              orr[0] = -dm * xi[0] + xr[1], oii[0] = dm * xr[0] + xi[1];
              orr[1] = -dm * xi[2] - xr[3], oii[1] = dm * xr[2] - xi[3];
            }
            const long lre = (long)n * (n + 1) + m, lim = (long)n * (n + 1) - m;
#pragma unroll
            for (int o = 0; o < NOUT; o++) {
              const double* dg = o == 0 ? a.diag0 : a.diag1;
              double* W = o == 0 ? a.W0 : a.W1;
              __builtin_nontemporal_store(orr[o] * u * (dg ? dg[lre] : 1.0), W + lre * a.sl + (long)c * a.sc);
              if (m > 0) __builtin_nontemporal_store(oii[o] * u * (dg ? dg[lim] : 1.0), W + lim * a.sl + (long)c * a.sc);
            }
          }
        }
      }
    }
    if (rbase + 4 * NSW < nslot) __syncthreads();  // the next round restages buffer 0
  }
}

// Forward with every latitude of the block's columns staged once in LDS.  A block covers one
// wavenumber m and NCT column tiles; its threads load the north and south Fourier values of all
// its K steps (an item is 8 latitudes of one column), form the symmetric and antisymmetric
// combinations, split them and store them as B fragments in LDS behind one barrier.  The row
// tiles of both parities are then dealt round robin to the waves; each wave walks (tile, K
// step) pairs with the table fragments of the next pair in flight, reads its data fragments
// from LDS (the combination its tile's parity needs) and runs the epilogue when a tile's K
// loop ends, with no further barrier.  Same tables and epilogue formulas as k_mx2_fwd_sh.
template <int OP, int FORM, int NCT>
__host__ __device__ constexpr int fwk_panel_elems() {  // one K step: [tile][input][sym, anti][re, im][term]
  return NCT * FwdTraits<OP>::NIN * 2 * 2 * Mx2<FORM>::NTB * 512;
}

template <int OP, int FORM, int NCT>
__global__ __launch_bounds__(256, 2) void k_mx2_fwd_wk(Mx2FwdArgs a) {
  using F = MxForm<FORM>;
  using X = Mx2<FORM>;
  using T = typename F::T;
  using V8 = typename F::V8;
  constexpr int NG = fwd_ng(OP), NIN = FwdTraits<OP>::NIN, NOUT = FwdTraits<OP>::NOUT;
  constexpr int NTA = X::NTA, NTB = X::NTB;
  constexpr bool UP = FwdTraits<OP>::USE_P, UH = FwdTraits<OP>::USE_H;
  constexpr int CB = NCT * TILE;
  constexpr int PANEL = fwk_panel_elems<OP, FORM, NCT>();
  T* sB = reinterpret_cast<T*>(mx2_lds);
  int m, cg;
  if (!xcd_decode(a.N, cdiv(a.CT, NCT), m, cg)) return;  // uniform over the block
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int nE = n_par(a.N, m, 0), nO = n_par(a.N, m, 1);
  const int tE = cdiv(nE, TILE), tO = cdiv(nO, TILE);
  const int KS = a.J2p32 / KM;
  const long toff = a.off[m];
  const float2* Gm[2] = {a.G0 + (long)m * a.gmst, NIN > 1 ? a.G1 + (long)m * a.gmst : nullptr};
  const double dm = (double)m;

  // ---- staging: item it -> K step it / IPS; within it (column cc, k-group kg, input i), cc fastest
  {
    const int cc = tid % CB;  // 256 is a multiple of CB: a thread keeps its column
    const int c = cg * CB + cc;
    const bool cok = c < a.ncols;
    float dsc = 1.f;
    if constexpr (F::SCALED) {
      if (cok) dsc = colscale<NIN>(a.colmax, a.ncols, c, a.bound);
    }
    constexpr int IPS = NIN * CB * 4;
    const int total = KS * IPS;
    for (int it = tid; it < total; it += 256) {
      const int ks = it / IPS, rest = (it % IPS) / CB;
      const int kg = rest & 3, i = rest >> 2;
      const float2* g = Gm[i] + c;
      float2 gn[8], gs[8];
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int jn = ks * KM + 8 * kg + e;
        gn[e] = gs[e] = make_float2(0.f, 0.f);
        if (cok && jn < a.J2) {
          gn[e] = g[(long)(a.J2 + jn) * a.gld];
          gs[e] = g[(long)(a.J2 - 1 - jn) * a.gld];
        }
      }
      T ts[2][2][NTB][8];  // [sym, anti][re, im][term][latitude]
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const float cw = a.cw[ks * KM + 8 * kg + e] * dsc;
        const float v[2][2] = {{cw * (gn[e].x + gs[e].x), cw * (gn[e].y + gs[e].y)},
                               {cw * (gn[e].x - gs[e].x), cw * (gn[e].y - gs[e].y)}};
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
          for (int r = 0; r < 2; r++) {
            T tmp[NTB];
            bsplit<FORM>(v[q][r], tmp);
#pragma unroll
            for (int t = 0; t < NTB; t++) ts[q][r][t][e] = tmp[t];
          }
      }
      T* dst = sB + (long)ks * PANEL + (((cc / TILE) * NIN + i) * 4 * NTB) * 512 + ((cc % TILE) + TILE * kg) * 8;
#pragma unroll
      for (int q = 0; q < 2; q++)
#pragma unroll
        for (int r = 0; r < 2; r++)
#pragma unroll
          for (int t = 0; t < NTB; t++) {
            V8 vv;
#pragma unroll
            for (int e = 0; e < 8; e++) vv[e] = ts[q][r][t][e];
            *reinterpret_cast<V8*>(dst + ((q * 2 + r) * NTB + t) * 512) = vv;
          }
    }
  }
  __syncthreads();

  // ---- contraction: wave w takes row tiles w, w + 4, ... (both parities, tile index over
  // parity 0 then parity 1 as in the table); (tile, K step) pairs s = q KS + ks, q its q-th tile
  const int ntile = tE + tO;
  const int mytiles = w < ntile ? (ntile - w + 3) / 4 : 0;
  const int nseq = mytiles * KS;
  float4_t acc[NG][NCT][2];
#pragma unroll
  for (int g = 0; g < NG; g++)
#pragma unroll
    for (int t = 0; t < NCT; t++) acc[g][t][0] = acc[g][t][1] = float4_t{0, 0, 0, 0};
  auto loadA = [&](int sq, V8 (&A)[2][NTA]) {
    const int tile = w + 4 * (sq / KS), ks = sq % KS;
    const long o = toff + ((long)tile * KS + ks) * 512 + lane * 8;
#pragma unroll
    for (int k = 0; k < NTA; k++) {
      if constexpr (UP) A[0][k] = *reinterpret_cast<const V8*>(a.tP[k] + o);
      if constexpr (UH) A[1][k] = *reinterpret_cast<const V8*>(a.tH[k] + o);
    }
  };
  auto epilogue = [&](int tile) {
    const int PP = tile >= tE ? 1 : 0, tp = PP ? tile - tE : tile;
#pragma unroll
    for (int t = 0; t < NCT; t++) {
      const int c = (cg * NCT + t) * TILE + (lane & 15);
      if (c >= a.ncols) continue;
      double u = a.minv[m] * X::DSCALE * a.scale;
      if constexpr (F::SCALED) u /= (double)colscale<NIN>(a.colmax, a.ncols, c, a.bound);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int k = tp * TILE + acc_row16(lane, i);
        const int n = m + PP + 2 * k;
        if (n > a.N) continue;
        double xr[NG], xi[NG];
#pragma unroll
        for (int gi = 0; gi < NG; gi++) {
          xr[gi] = acc[gi][t][0][i];
          xi[gi] = acc[gi][t][1][i];
        }
        double orr[NOUT], oii[NOUT];
        if constexpr (OP == F_VAL) {
          orr[0] = xr[0], oii[0] = xi[0];
        } else if constexpr (OP == F_DIVLAMBDA) {  // i m P.G
          orr[0] = -dm * xi[0], oii[0] = dm * xr[0];
        } else if constexpr (OP == F_DIVMU) {  // -H.G
          orr[0] = -xr[0], oii[0] = -xi[0];
        } else if constexpr (OP == F_ALPHA) {  // i m P.A - H.B
          orr[0] = -dm * xi[0] - xr[1], oii[0] = dm * xr[0] - xi[1];
        } else {  // vor = i m P.V + H.U ; div = i m P.U - H.V
          orr[0] = -dm * xi[0] + xr[1], oii[0] = dm * xr[0] + xi[1];
          orr[1] = -dm * xi[2] - xr[3], oii[1] = dm * xr[2] - xi[3];
        }
        const long lre = (long)n * (n + 1) + m, lim = (long)n * (n + 1) - m;
#pragma unroll
        for (int o = 0; o < NOUT; o++) {
          const double* dg = o == 0 ? a.diag0 : a.diag1;
          double* W = o == 0 ? a.W0 : a.W1;
          __builtin_nontemporal_store(orr[o] * u * (dg ? dg[lre] : 1.0), W + lre * a.sl + (long)c * a.sc);
          if (m > 0) __builtin_nontemporal_store(oii[o] * u * (dg ? dg[lim] : 1.0), W + lim * a.sl + (long)c * a.sc);
        }
      }
    }
#pragma unroll
    for (int g = 0; g < NG; g++)
#pragma unroll
      for (int t = 0; t < NCT; t++) acc[g][t][0] = acc[g][t][1] = float4_t{0, 0, 0, 0};
  };
  auto step = [&](int sq, const V8 (&A)[2][NTA]) {
    const int tile = w + 4 * (sq / KS), ks = sq % KS;
    const int PP = tile >= tE ? 1 : 0;
    const T* pan = sB + (long)ks * PANEL + lane * 8;
#pragma unroll
    for (int gi = 0; gi < NG; gi++) {
      const int tb = fwd_tab(OP, gi), in = fwd_in(OP, gi);
      const int cq = tb == 0 ? PP : 1 - PP;  // 0 = sym, 1 = anti
#pragma unroll
      for (int t = 0; t < NCT; t++)
#pragma unroll
        for (int r = 0; r < 2; r++) {
          V8 B[NTB];
#pragma unroll
          for (int k = 0; k < NTB; k++)
            B[k] = *reinterpret_cast<const V8*>(pan + (((t * NIN + in) * 4 + 2 * cq + r) * NTB + k) * 512);
#pragma unroll
          for (int pr = 0; pr < X::NP; pr++)
            acc[gi][t][r] = mfma16(A[tb][X::pa(pr)], B[X::pb(pr)], acc[gi][t][r]);
        }
    }
    if (ks == KS - 1) epilogue(tile);
  };
  V8 A0[2][NTA], A1[2][NTA];
  if (nseq > 0) loadA(0, A0);
  for (int sq = 0; sq < nseq; sq += 2) {
    if (sq + 1 < nseq) loadA(sq + 1, A1);
    step(sq, A0);
    if (sq + 1 < nseq) {
      if (sq + 2 < nseq) loadA(sq + 2, A0);
      step(sq + 1, A1);
    }
  }
}

// per-column max |pro[l] w(l, c)| (float bits), threads along columns
__global__ __launch_bounds__(256) void k_spec_colmax2(const double* __restrict__ w, long sl, long sc, int L,
                                                      int ncols, const double* __restrict__ pro,
                                                      unsigned* __restrict__ out) {
  __shared__ float red[256];
  const int tid = threadIdx.x;
  const int c = blockIdx.x * 64 + (tid & 63);
  float mx[4] = {0.f, 0.f, 0.f, 0.f};
  if (c < ncols) {
    const int step = 4 * gridDim.y;
    int l = blockIdx.y * 4 + (tid >> 6);
    for (; l + 3 * step < L; l += 4 * step) {  // four independent loads in flight
#pragma unroll
      for (int u = 0; u < 4; u++) {
        double v = w[(long)(l + u * step) * sl + (long)c * sc];
        if (pro) v *= pro[l + u * step];
        mx[u] = fmaxf(mx[u], fabsf((float)v));
      }
    }
    for (; l < L; l += step) {
      double v = w[(long)l * sl + (long)c * sc];
      if (pro) v *= pro[l];
      mx[0] = fmaxf(mx[0], fabsf((float)v));
    }
  }
  red[tid] = fmaxf(fmaxf(mx[0], mx[1]), fmaxf(mx[2], mx[3]));
  __syncthreads();
  if (tid < 64) {
    const float m2 = fmaxf(fmaxf(red[tid], red[tid + 64]), fmaxf(red[tid + 128], red[tid + 192]));
    if (c < ncols) atomicMax(out + c, __float_as_uint(m2));
  }
}


// ---------------------------------------------------------------- launchers
// tile shapes per operator: one-group operators take 2x2 tiles per wave, the two- and
// four-group ones fewer so that their accumulators stay within the register budget
// (NCT: column tiles per wave; the plan's column-block knob selects 1 tile for 16, else 2)
// column tiles of a whole-K staged inverse block: one for two-input and four-group
// operators, else the requested count (two by default, one with the 16-column knob)
template <int OP>
constexpr int wk_nct(int req) {
  return (InvTraits<OP>::NIN > 1 || inv_ng(OP) == 4) ? 1 : req;
}
// LDS of the whole-K staged inverse: every K chunk of wavenumber 0 (the most) for the
// block's column tiles
template <int OP, int FORM, int WNCT = wk_nct<OP>(2)>
size_t staged_lds(const dsht_plan* p) {
  const int nchmax = (kpad32(n_par(p->N, 0, 0)) + kpad32(n_par(p->N, 0, 1))) / KM;
  return (size_t)nchmax * wk_chunk_elems<OP, FORM, WNCT>() * sizeof(typename MxForm<FORM>::T);
}

template <int OP, int NCT_>
struct InvShape {
  // packed inverse: two latitude tiles per wave (one row group covers T170's 8 tiles) where
  // the accumulators allow
  static constexpr int NCT = inv_ng(OP) == 4 ? 1 : NCT_;
  static constexpr int MAXT = (inv_ng(OP) == 1 || (inv_ng(OP) == 2 && InvTraits<OP>::NIN == 1)) ? 2 : 1;
};
template <int OP>
struct FwdShape {
  static constexpr int MAXT = fwd_ng(OP) == 1 ? 2 : 1;
  // slots per wave and round of the LDS-shared form (accumulators NSW * NG * MAXT * 2 float4)
  static constexpr int NSW = fwd_ng(OP) == 2 ? 3 : 2;
};

template <int OP, int FORM, int NCT_>
int inv_launch(const dsht_plan* p, const MixedTables& t, const Mx2InvJob& j, hipStream_t s) {
  constexpr int NIN = InvTraits<OP>::NIN;
  Mx2InvArgs a;
  a.W0 = j.W0;
  a.W1 = j.W1;
  a.sl = j.sl;
  a.sc = j.sc;
  a.ncols = j.ncols;
  a.CT = cdiv(j.ncols, TILE);
  a.pro = j.pro;
  a.rs = j.rs;
  a.colmax = j.colmax;
  a.pack = j.pack;
  a.G0 = j.G0;
  a.G1 = j.G1;
  a.G2 = j.G2;
  a.gld = j.gld;
  a.gmst = j.gmst;
  for (int k = 0; k < 3; k++) {
    a.tP[k] = t.iP[k];
    a.tH[k] = t.iH[k];
  }
  a.off = t.d_offI;
  a.cbase = t.d_cbase;
  a.chunk_m = t.d_chunk_m;
  a.nchunks = t.nchunks;
  a.minv = t.d_minv;
  a.N = p->N;
  a.J2 = p->J2;
  a.J2p32 = t.J2p32;
  a.RT = t.J2p32 / TILE;
  constexpr int NCT = InvShape<OP, NCT_>::NCT, MAXT = InvShape<OP, NCT_>::MAXT;
  a.CG = cdiv(a.CT, NCT);
  a.RG = cdiv(a.RT, 4 * MAXT);
  const long nb = 8l * cdiv(p->N + 1, 8) * a.CG * a.RG;
  if (j.staged) {
    // whole-K staging: every latitude tile in one block where 4 waves x 2 tiles cover them
    constexpr int WNCT = wk_nct<OP>(NCT_);
    const int wmaxt = a.RT <= 4 ? 1 : 2;
    const size_t lds = staged_lds<OP, FORM, WNCT>(p);
    a.CG = cdiv(a.CT, WNCT);
    a.RG = cdiv(a.RT, 4 * wmaxt);
    const long nbw = 8l * cdiv(p->N + 1, 8) * a.CG * a.RG;
    auto k = wmaxt == 1 ? k_mx2_inv_wk<OP, FORM, WNCT, 1> : k_mx2_inv_wk<OP, FORM, WNCT, 2>;
    if (lds > 64 * 1024) {  // above the default limit; once per kernel and process
      static const hipError_t attr =
          hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)MX2_STAGED_LDS_MAX);
      if (attr != hipSuccess) return fail(DSHT_EDEVICE, "LDS size attribute: %s", hipGetErrorString(attr));
    }
    hipLaunchKernelGGL(k, dim3((unsigned)nbw), dim3(256), lds, s, a);
  } else {
    const long pt = 4l * a.CT * TILE * t.nchunks * NIN;  // packing threads
    hipLaunchKernelGGL((k_mx2_ipack<FORM, NIN>), dim3((unsigned)cdivl(pt, 256)), dim3(256), 0, s, a);
    DSHT_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_mx2_inv<OP, FORM, NCT, MAXT>), dim3((unsigned)nb), dim3(256), 0, s, a);
  }
  DSHT_HIP(hipGetLastError());
  return 0;
}

template <int OP, int FORM>
int fwd_launch(const dsht_plan* p, const MixedTables& t, const Mx2FwdJob& j, hipStream_t s) {
  Mx2FwdArgs a;
  a.G0 = j.G0;
  a.G1 = j.G1;
  a.gld = j.gld;
  a.gmst = j.gmst;
  a.ncols = j.ncols;
  a.CT = cdiv(j.ncols, TILE);
  a.W0 = j.W0;
  a.W1 = j.W1;
  a.sl = j.sl;
  a.sc = j.sc;
  a.cw = j.cw;
  a.colmax = j.colmax;
  a.bound = j.bound;
  a.diag0 = j.diag0;
  a.diag1 = j.diag1;
  a.scale = j.scale;
  for (int k = 0; k < 3; k++) {
    a.tP[k] = t.fP[k];
    a.tH[k] = t.fH[k];
  }
  a.off = t.d_offF;
  a.minv = t.d_minv;
  a.N = p->N;
  a.J2 = p->J2;
  a.J2p32 = t.J2p32;
  // NCT column tiles per block.  Two tiles halve the table fetches per column but double the
  // accumulators; measured at T170 they pay only for the single-group bf16x6 form (6
  // products per data fragment: forward 101 -> 84 us).  The column-block knob forces 1 (16)
  // or 2 (32) tiles.
  if (j.staged) {  // whole-K staged forward: one column tile per block (two with leg_cb = 32)
    const int KS = a.J2p32 / KM;
    const bool two = p->leg_cb == 32;
    const size_t lds = (size_t)KS * sizeof(typename MxForm<FORM>::T) *
                       (two ? fwk_panel_elems<OP, FORM, 2>() : fwk_panel_elems<OP, FORM, 1>());
    if (lds <= MX2_STAGED_LDS_MAX) {
      const long nbw = 8l * cdiv(p->N + 1, 8) * cdiv(a.CT, two ? 2 : 1);
      auto k = two ? k_mx2_fwd_wk<OP, FORM, 2> : k_mx2_fwd_wk<OP, FORM, 1>;
      if (lds > 64 * 1024) {  // above the default limit; once per kernel and process
        static const hipError_t attr =
            hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)MX2_STAGED_LDS_MAX);
        if (attr != hipSuccess) return fail(DSHT_EDEVICE, "LDS size attribute: %s", hipGetErrorString(attr));
      }
      hipLaunchKernelGGL(k, dim3((unsigned)nbw), dim3(256), lds, s, a);
      DSHT_HIP(hipGetLastError());
      return 0;
    }
  }
  const int fnct = p->leg_cb == 16 ? 1 : (p->leg_cb == 32 ? 2 : (fwd_ng(OP) == 1 && Mx2<FORM>::NP == 6 ? 2 : 1));
  const long nb = 8l * cdiv(p->N + 1, 8) * cdiv(a.CT, fnct);
  // four K steps (a 256-latitude grid such as T170's): the unrolled form
  const bool ks4 = fwd_ng(OP) == 1 && a.J2p32 / KM == 4;
  if (fnct == 2 && ks4)
    hipLaunchKernelGGL((k_mx2_fwd_sh<OP, FORM, FwdShape<OP>::MAXT, FwdShape<OP>::NSW, 2, fwd_ng(OP) == 1 ? 4 : 0>),
                       dim3((unsigned)nb), dim3(256), 0, s, a);
  else if (fnct == 2)
    hipLaunchKernelGGL((k_mx2_fwd_sh<OP, FORM, FwdShape<OP>::MAXT, FwdShape<OP>::NSW, 2>), dim3((unsigned)nb), dim3(256), 0,
                       s, a);
  else if (ks4)
    hipLaunchKernelGGL((k_mx2_fwd_sh<OP, FORM, FwdShape<OP>::MAXT, FwdShape<OP>::NSW, 1, fwd_ng(OP) == 1 ? 4 : 0>),
                       dim3((unsigned)nb), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((k_mx2_fwd_sh<OP, FORM, FwdShape<OP>::MAXT, FwdShape<OP>::NSW, 1>), dim3((unsigned)nb), dim3(256), 0,
                       s, a);
  DSHT_HIP(hipGetLastError());
  return 0;
}

template <int FORM, int NCT>
int inv_form_nct(const dsht_plan* p, const MixedTables& t, int op, const Mx2InvJob& j, hipStream_t s) {
  switch (op) {
    case I_VAL: return inv_launch<I_VAL, FORM, NCT>(p, t, j, s);
    case I_GRADLAMBDA: return inv_launch<I_GRADLAMBDA, FORM, NCT>(p, t, j, s);
    case I_GRADMU: return inv_launch<I_GRADMU, FORM, NCT>(p, t, j, s);
    case I_ALPHA: return inv_launch<I_ALPHA, FORM, NCT>(p, t, j, s);
    case I_VORDIV2UV: return inv_launch<I_VORDIV2UV, FORM, NCT>(p, t, j, s);
    case I_VGRAD: return inv_launch<I_VGRAD, FORM, NCT>(p, t, j, s);
  }
  return fail(DSHT_EARG, "bad inverse op %d", op);
}

template <int FORM>
int inv_form(const dsht_plan* p, const MixedTables& t, int op, const Mx2InvJob& j, hipStream_t s) {
  return p->leg_cb == 16 ? inv_form_nct<FORM, 1>(p, t, op, j, s) : inv_form_nct<FORM, 2>(p, t, op, j, s);
}

template <int FORM>
int fwd_form(const dsht_plan* p, const MixedTables& t, int op, const Mx2FwdJob& j, hipStream_t s) {
  switch (op) {
    case F_VAL: return fwd_launch<F_VAL, FORM>(p, t, j, s);
    case F_DIVLAMBDA: return fwd_launch<F_DIVLAMBDA, FORM>(p, t, j, s);
    case F_DIVMU: return fwd_launch<F_DIVMU, FORM>(p, t, j, s);
    case F_ALPHA: return fwd_launch<F_ALPHA, FORM>(p, t, j, s);
    case F_UV2VORDIV: return fwd_launch<F_UV2VORDIV, FORM>(p, t, j, s);
  }
  return fail(DSHT_EARG, "bad forward op %d", op);
}

}  // namespace

template <int FORM>
static size_t staged_lds_form(const dsht_plan* p, int op) {
  switch (op) {
    case I_VAL: return staged_lds<I_VAL, FORM>(p);
    case I_GRADLAMBDA: return staged_lds<I_GRADLAMBDA, FORM>(p);
    case I_GRADMU: return staged_lds<I_GRADMU, FORM>(p);
    case I_ALPHA: return staged_lds<I_ALPHA, FORM>(p);
    case I_VORDIV2UV: return staged_lds<I_VORDIV2UV, FORM>(p);
    default: return staged_lds<I_VGRAD, FORM>(p);
  }
}

size_t mx2_inv_staged_lds(const dsht_plan* p, int form, int op) {
  switch (form) {
    case 0: return staged_lds_form<0>(p, op);
    case 1: return staged_lds_form<1>(p, op);
    case 2: return staged_lds_form<2>(p, op);
    default: return staged_lds_form<3>(p, op);
  }
}

size_t mx2_inv_pack_bytes(const MixedTables& t, int form, int op, int ncols) {
  const int ntb = form == 1 ? 3 : (form == 2 ? 1 : (form == 3 ? 3 : 2));
  const int nin = (op == I_ALPHA || op == I_VORDIV2UV) ? 2 : 1;
  return (size_t)t.nchunks * cdiv(ncols, TILE) * nin * 2 * ntb * 512 * sizeof(uint16_t);
}

int mx2_inv(const dsht_plan* p, const MixedTables& t, int form, int op, const Mx2InvJob& j, hipStream_t s) {
  switch (form) {
    case 0: return inv_form<0>(p, t, op, j, s);
    case 1: return inv_form<1>(p, t, op, j, s);
    case 2: return inv_form<2>(p, t, op, j, s);
    case 3: return inv_form<3>(p, t, op, j, s);
  }
  return fail(DSHT_EARG, "bad form %d", form);
}

int mx2_fwd(const dsht_plan* p, const MixedTables& t, int form, int op, const Mx2FwdJob& j, hipStream_t s) {
  switch (form) {
    case 0: return fwd_form<0>(p, t, op, j, s);
    case 1: return fwd_form<1>(p, t, op, j, s);
    case 2: return fwd_form<2>(p, t, op, j, s);
    case 3: return fwd_form<3>(p, t, op, j, s);
  }
  return fail(DSHT_EARG, "bad form %d", form);
}

int launch_spec_colmax(const double* w, long sl, long sc, int L, int ncols, const double* pro, unsigned* out,
                       hipStream_t s) {
  // enough workgroups to cover the chip: ~16 rows of l per thread
  dim3 grid(cdiv(ncols, 64), std::max(1, std::min(1024, cdiv(L, 64))));
  hipLaunchKernelGGL(k_spec_colmax2, grid, dim3(256), 0, s, w, sl, sc, L, ncols, pro, out);
  DSHT_HIP(hipGetLastError());
  return 0;
}

}  // namespace dsht
