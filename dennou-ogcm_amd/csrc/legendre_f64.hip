// legendre_f64.hip -- FP64 Legendre contractions on CDNA4 MFMA (v_mfma_f64_16x16x4f64).
//
// The Legendre half of ISPACK's SNPACK/SJPACK transforms behind SPMODEL's w_xy, xy_w
// and their derivative forms (doc/description/ogcm/space_discretization.tex:199-268):
//
//   forward  s_n^m = sum_j c_j G^m(mu_j) T_n^m(mu_j),  T in {P, H}, c_j = w_j or w_j/(1-mu_j^2)
//   inverse  G^m(mu_j) = sum_n s_n^m T_n^m(mu_j)
//
// Even/odd split: P_n^m(-mu) = (-1)^{n-m} P_n^m(mu) and H_n^m has the opposite parity,
// so each zonal wavenumber m is two dense contractions over one hemisphere:
//   forward  rows n of parity p use G(north) +/- G(south),
//   inverse  sym/anti accumulators give G(north) = sym + anti, G(south) = sym - anti.
// One workgroup owns one m and a block of CB complex columns (layers x fields); its four
// waves share the staged data tile in LDS and each owns a set of 16-row MFMA tiles.  The
// table operand comes straight from global memory (L2/MALL resident), pre-swizzled so
// that lane l of a wave reads element l of a 64-double fragment.
//
// Fused operator epilogues (SpmlUtil_mod.f90:297-473):
//   F_DIVLAMBDA  i m * (P . cG)            F_DIVMU   -(H . cG)
//   F_ALPHA      (i m P . cA - H . cB)/R   (w_AlphaOptr_xy)
//   F_UV2VORDIV  vor = (i m P.cV + H.cU)/R, div = (i m P.cU - H.cV)/R (calc_UVCosLat2VorDiv)
//   I_GRADLAMBDA i m P . s                 I_GRADMU  H . s
//   I_ALPHA      (i m P.sA + H.sB)/(R cos^2)  (xy_AlphaOptr_w)
//   I_VORDIV2UV  psi, chi = -R^2/(n(n+1)) (vor, div);
//                U = (i m P.chi - H.psi)/(R cos), V = (i m P.psi + H.chi)/(R cos)  (calc_VorDiv2UV)
//   I_VGRAD      P.s, i m P.s / cos, H.s / cos from one pass (xy_w, xy_GradLon_w, xy_GradLat_w)
// plus an optional per-l diagonal (filters, implicit diffusion, Helmholtz) on the spectral side.
#include <type_traits>

#include "dsht_internal.hpp"

namespace dsht {

__device__ __forceinline__ double4_t mfma(double a, double b, double4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
// v_mfma_f64_16x16x4f64 lane layout: A[r][k] at lane r + 16k, B[k][c] at lane c + 16k,
// result element i of lane l holds D[l/16 + 4i][l%16].
__device__ __forceinline__ int acc_row(int lane, int i) { return (lane >> 4) + 4 * i; }

// ------------------------------------------------------------------ block placement
// 1-D grid of 8 * ceil((N+1)/8) * CB-column blocks.  Workgroups are dispatched to the 8 XCDs
// round robin, so block b runs on XCD b mod 8: wavenumber m = b mod 8 + 8k takes the
// consecutive blocks of XCD m mod 8, and every column block of one m shares that XCD's L2
// copy of its table panel (the table is fetched from HBM/MALL once per m, not once per
// column block).  Wavenumbers alternate over the XCDs, which spreads the uneven per-m work.
__device__ __forceinline__ bool xcd_block(int N, int ncols, int CB, int& m, int& cb) {
  const int nblk = (ncols + CB - 1) / CB;
  const int b = blockIdx.x, idx = b >> 3;
  m = (b & 7) + 8 * (idx / nblk);
  cb = idx % nblk;
  return m <= N;
}
static inline dim3 xcd_grid(int N, int ncols, int CB) { return dim3(8u * (unsigned)cdiv(N + 1, 8) * cdiv(ncols, CB)); }

// ------------------------------------------------------------------ forward
// WAVES waves per workgroup; wave w owns row tiles t = tbase + w + WAVES*tt (tt < MAXT).
// The K loop (north-hemisphere latitudes) runs in chunks of KC and is software pipelined:
// the Fourier values of chunk k+1 and its table fragments are loaded into registers while
// the MFMAs of chunk k run (two table-fragment register sets used alternately), then the
// data are staged to LDS behind one barrier.  Tables are padded to multiples of KCH = 16
// only; partial chunks are predicated per 4-wide K step.
template <int OP, int CB, int WAVES, int MAXT, int KC>
__global__ __launch_bounds__(WAVES * 64) void k_leg_fwd(LegFwdArgs a) {
  using Tr = FwdTraits<OP>;
  constexpr int NT = WAVES * 64;
  constexpr int NIN = Tr::NIN, NOUT = Tr::NOUT, NCT = CB / TILE;
  constexpr int PITCH = CB + 1;  // LDS row pitch (doubles)
  constexpr int EPT = (KC * CB + NT - 1) / NT;  // staged elements per thread
  constexpr int NKS = KC / 4;
  constexpr bool UP = Tr::USE_P, UH = Tr::USE_H;
  // [input][S re, S im, A re, A im][KC][PITCH]
  __shared__ double sB[NIN][4][KC * PITCH];

  int m, cb;
  if (!xcd_block(a.N, a.ncols, CB, m, cb)) return;  // uniform over the workgroup
  const int c0 = cb * CB;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int N = a.N, J2 = a.J2;
  const int nE = n_par(N, m, 0), nO = n_par(N, m, 1);
  const int tE = (nE + TILE - 1) / TILE, tO = (nO + TILE - 1) / TILE;
  const int RT = tE + tO;
  const int KS = a.J2pad / 4;  // valid 4-wide K steps
  const int nch = (a.J2pad + KC - 1) / KC;
  const long kstepsF = a.J2pad / 4;
  const double* tabP = a.tabP + a.offF[m];
  const double* tabH = a.tabH + a.offF[m];
  const double2* Gin[2] = {a.G0 + (long)m * a.gmst, a.G1 ? a.G1 + (long)m * a.gmst : nullptr};
  const double dm = (double)m;

  for (int tbase = 0; tbase < RT; tbase += WAVES * MAXT) {
    double4_t acc[NOUT][MAXT][NCT][2];
#pragma unroll
    for (int o = 0; o < NOUT; o++)
#pragma unroll
      for (int t = 0; t < MAXT; t++)
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) acc[o][t][ct][0] = acc[o][t][ct][1] = double4_t{0, 0, 0, 0};

    double2 gN[NIN][EPT], gS[NIN][EPT];
    double aP0[MAXT][NKS], aH0[MAXT][NKS], aP1[MAXT][NKS], aH1[MAXT][NKS];

    auto load_b = [&](int k0) {
#pragma unroll
      for (int i = 0; i < NIN; i++)
#pragma unroll
        for (int q = 0; q < EPT; q++) {
          const int e = tid + NT * q;
          const int kk = e / CB, cc = e % CB;
          const int jn = k0 + kk, c = c0 + cc;
          gN[i][q] = gS[i][q] = make_double2(0.0, 0.0);
          if (e < KC * CB && jn < J2 && c < a.ncols) {
            gN[i][q] = Gin[i][(long)(J2 + jn) * a.gld + c];
            gS[i][q] = Gin[i][(long)(J2 - 1 - jn) * a.gld + c];
          }
        }
    };
    auto load_a = [&](int k0, double (&xP)[MAXT][NKS], double (&xH)[MAXT][NKS]) {
#pragma unroll
      for (int tt = 0; tt < MAXT; tt++) {
        const int t = tbase + w + WAVES * tt;
#pragma unroll
        for (int ks = 0; ks < NKS; ks++) {
          const int kg = k0 / 4 + ks;
          const long fo = (long)t * kstepsF * 64 + (long)kg * 64 + lane;
          const bool ok = t < RT && kg < KS;
          xP[tt][ks] = (UP && ok) ? tabP[fo] : 0.0;
          xH[tt][ks] = (UH && ok) ? tabH[fo] : 0.0;
        }
      }
    };
    auto stage = [&](int k0) {
#pragma unroll
      for (int i = 0; i < NIN; i++)
#pragma unroll
        for (int q = 0; q < EPT; q++) {
          const int e = tid + NT * q;
          if (e < KC * CB) {
            const int kk = e / CB, cc = e % CB;
            const double cw = (k0 + kk < a.J2pad) ? a.cw[k0 + kk] : 0.0;
            sB[i][0][kk * PITCH + cc] = cw * (gN[i][q].x + gS[i][q].x);
            sB[i][1][kk * PITCH + cc] = cw * (gN[i][q].y + gS[i][q].y);
            sB[i][2][kk * PITCH + cc] = cw * (gN[i][q].x - gS[i][q].x);
            sB[i][3][kk * PITCH + cc] = cw * (gN[i][q].y - gS[i][q].y);
          }
        }
    };
    auto compute = [&](int k0, const double (&xPa)[MAXT][NKS], const double (&xHa)[MAXT][NKS]) {
      const int kvalid = KS - k0 / 4;
#pragma unroll
      for (int ks = 0; ks < NKS; ks++) {
        if (ks < kvalid) {
        const int krow = ks * 4 + (lane >> 4);
        double b[NIN][2][NCT][2];
#pragma unroll
        for (int i = 0; i < NIN; i++)
#pragma unroll
          for (int ct = 0; ct < NCT; ct++) {
            const int col = ct * TILE + (lane & 15);
            b[i][0][ct][0] = sB[i][0][krow * PITCH + col];
            b[i][0][ct][1] = sB[i][1][krow * PITCH + col];
            b[i][1][ct][0] = sB[i][2][krow * PITCH + col];
            b[i][1][ct][1] = sB[i][3][krow * PITCH + col];
          }
#pragma unroll
        for (int tt = 0; tt < MAXT; tt++) {
          const int t = tbase + w + WAVES * tt;
          if (t >= RT) continue;
          const int p = t < tE ? 0 : 1;
          // P pairs with (sym for even rows, anti for odd); H with the opposite
          const int gP = p, gH = 1 - p;
          const double xP = xPa[tt][ks], xH = xHa[tt][ks];
#pragma unroll
          for (int ct = 0; ct < NCT; ct++) {
            if constexpr (OP == F_VAL || OP == F_DIVLAMBDA) {
              acc[0][tt][ct][0] = mfma(xP, b[0][gP][ct][0], acc[0][tt][ct][0]);
              acc[0][tt][ct][1] = mfma(xP, b[0][gP][ct][1], acc[0][tt][ct][1]);
            } else if constexpr (OP == F_DIVMU) {
              acc[0][tt][ct][0] = mfma(xH, b[0][gH][ct][0], acc[0][tt][ct][0]);
              acc[0][tt][ct][1] = mfma(xH, b[0][gH][ct][1], acc[0][tt][ct][1]);
            } else if constexpr (OP == F_ALPHA) {
              // i m P.A - H.B :  (i m z) = (-m z.im, m z.re)
              acc[0][tt][ct][0] = mfma(xP, -dm * b[0][gP][ct][1], acc[0][tt][ct][0]);
              acc[0][tt][ct][1] = mfma(xP, dm * b[0][gP][ct][0], acc[0][tt][ct][1]);
              acc[0][tt][ct][0] = mfma(-xH, b[1][gH][ct][0], acc[0][tt][ct][0]);
              acc[0][tt][ct][1] = mfma(-xH, b[1][gH][ct][1], acc[0][tt][ct][1]);
            } else {  // F_UV2VORDIV, inputs U cos (0), V cos (1)
              // vor = i m P.V + H.U
              acc[0][tt][ct][0] = mfma(xP, -dm * b[1][gP][ct][1], acc[0][tt][ct][0]);
              acc[0][tt][ct][1] = mfma(xP, dm * b[1][gP][ct][0], acc[0][tt][ct][1]);
              acc[0][tt][ct][0] = mfma(xH, b[0][gH][ct][0], acc[0][tt][ct][0]);
              acc[0][tt][ct][1] = mfma(xH, b[0][gH][ct][1], acc[0][tt][ct][1]);
              // div = i m P.U - H.V
              acc[1][tt][ct][0] = mfma(xP, -dm * b[0][gP][ct][1], acc[1][tt][ct][0]);
              acc[1][tt][ct][1] = mfma(xP, dm * b[0][gP][ct][0], acc[1][tt][ct][1]);
              acc[1][tt][ct][0] = mfma(-xH, b[1][gH][ct][0], acc[1][tt][ct][0]);
              acc[1][tt][ct][1] = mfma(-xH, b[1][gH][ct][1], acc[1][tt][ct][1]);
            }
          }
        }
        }
      }
    };

    load_b(0);
    load_a(0, aP0, aH0);
    for (int ch = 0; ch < nch; ch += 2) {
      const int k0 = ch * KC;
      stage(k0);
      __syncthreads();
      if (ch + 1 < nch) {
        load_b(k0 + KC);
        load_a(k0 + KC, aP1, aH1);
      }
      compute(k0, aP0, aH0);
      __syncthreads();
      if (ch + 1 >= nch) break;
      stage(k0 + KC);
      __syncthreads();
      if (ch + 2 < nch) {
        load_b(k0 + 2 * KC);
        load_a(k0 + 2 * KC, aP0, aH0);
      }
      compute(k0 + KC, aP1, aH1);
      __syncthreads();
    }

    // ---- epilogue: rotation/sign, scale, diagonal, scatter to SPMODEL order
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
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int k = tp * TILE + acc_row(lane, i);
          const int n = m + p + 2 * k;
          if (n > N) continue;
          const long lre = (long)n * (n + 1) + m, lim = (long)n * (n + 1) - m;
#pragma unroll
          for (int o = 0; o < NOUT; o++) {
            double re = acc[o][tt][ct][0][i], im = acc[o][tt][ct][1][i];
            if constexpr (OP == F_DIVLAMBDA) {
              double r2 = -dm * im;
              im = dm * re;
              re = r2;
            } else if constexpr (OP == F_DIVMU) {
              re = -re;
              im = -im;
            }
            const double* dg = o == 0 ? a.diag0 : a.diag1;
            double* W = o == 0 ? a.W0 : a.W1;
            const double sre = a.scale * (dg ? dg[lre] : 1.0);
            __builtin_nontemporal_store(re * sre, W + lre * a.sl + (long)c * a.sc);  // written once
            if (m > 0) {
              const double sim = a.scale * (dg ? dg[lim] : 1.0);
              __builtin_nontemporal_store(im * sim, W + lim * a.sl + (long)c * a.sc);
            }
          }
        }
      }
    }
  }
}

// ------------------------------------------------------------------ inverse
// Wave w owns latitude tiles t = tbase + w + WAVES*tt; the K loop runs over the total
// wavenumbers of parity 0 then parity 1 in chunks of KC (never mixing parities in one
// chunk), pipelined like the forward.
template <int OP, int CB, int WAVES, int MAXT, int KC>
__global__ __launch_bounds__(WAVES * 64) void k_leg_inv(LegInvArgs a) {
  using Tr = InvTraits<OP>;
  constexpr int NT = WAVES * 64;
  constexpr int NIN = Tr::NIN, NOUT = Tr::NOUT, NCT = CB / TILE;
  constexpr int NACC = OP == I_VGRAD ? 2 : NOUT;  // I_VGRAD: P and H products of one input
  constexpr int PITCH = CB + 1;
  constexpr int EPT = (KC * CB + NT - 1) / NT;
  constexpr int NKS = KC / 4;
  constexpr bool UP = Tr::USE_P, UH = Tr::USE_H;
  __shared__ double sB[NIN][2][KC * PITCH];  // [input][re, im][KC][PITCH]

  int m, cb;
  if (!xcd_block(a.N, a.ncols, CB, m, cb)) return;  // uniform over the workgroup
  const int c0 = cb * CB;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int N = a.N, J2 = a.J2;
  const int nE = n_par(N, m, 0), nO = n_par(N, m, 1);
  const int kp0 = (nE + KCH - 1) / KCH * KCH, kp1 = (nO + KCH - 1) / KCH * KCH;
  const int nch0 = (kp0 + KC - 1) / KC, nch = nch0 + (kp1 + KC - 1) / KC;
  const int RT = a.J2pad / TILE;
  const double* tabP = a.tabP + a.offI[m];
  const double* tabH = a.tabH + a.offI[m];
  const double* Win[2] = {a.W0, a.W1};
  const double dm = (double)m;

  for (int tbase = 0; tbase < RT; tbase += WAVES * MAXT) {
    // [output][tile][ct][re/im][sym/anti]
    double4_t acc[NACC][MAXT][NCT][2][2];
#pragma unroll
    for (int o = 0; o < NACC; o++)
#pragma unroll
      for (int t = 0; t < MAXT; t++)
#pragma unroll
        for (int ct = 0; ct < NCT; ct++)
#pragma unroll
          for (int r = 0; r < 2; r++) acc[o][t][ct][r][0] = acc[o][t][ct][r][1] = double4_t{0, 0, 0, 0};

    double bre[NIN][EPT], bim[NIN][EPT], bpr[EPT];
    bool bok[EPT];
    double aP0[MAXT][NKS], aH0[MAXT][NKS], aP1[MAXT][NKS], aH1[MAXT][NKS];
    auto chunk_pos = [&](int ch, int& p, int& k0) {
      p = ch < nch0 ? 0 : 1;
      k0 = (p ? ch - nch0 : ch) * KC;
    };
    // The next chunk's loads are issued before this chunk's products and used after them,
    // so none of them may be used inside the branch that guards it (such a load is waited
    // for there): load_b loads unconditionally (a padding element reads element 0) and
    // stage() applies the predicate and the prefactor; load_a clamps its index (padding
    // tiles and K steps are never multiplied or stored).
    auto load_b = [&](int ch) {
      int p, k0;
      chunk_pos(ch, p, k0);
      const int np = p ? nO : nE;
#pragma unroll
      for (int q = 0; q < EPT; q++) {
        const int e = tid + NT * q;
        const int kk = e / CB, cc = e % CB;
        const int k = k0 + kk, c = c0 + cc;
        const int n = m + p + 2 * k;
        bok[q] = e < KC * CB && k < np && c < a.ncols;
        const long l0 = bok[q] ? (long)n * (n + 1) : 0;
        const long lre = bok[q] ? l0 + m : 0, lim = bok[q] ? l0 - m : 0;
        const long co = bok[q] ? (long)c * a.sc : 0;
        bpr[q] = 1.0;
        if (a.pro) bpr[q] = a.pro[lre];
#pragma unroll
        for (int i = 0; i < NIN; i++) {
          // each block reads its spectra once: non-temporal
          bre[i][q] = __builtin_nontemporal_load(Win[i] + lre * a.sl + co);
          bim[i][q] = __builtin_nontemporal_load(Win[i] + lim * a.sl + co);
        }
      }
    };
    auto load_a = [&](int ch, double (&xP)[MAXT][NKS], double (&xH)[MAXT][NKS]) {
      int p, k0;
      chunk_pos(ch, p, k0);
      const int kp = p ? kp1 : kp0;
      const long pbase = p ? (long)a.J2pad * kp0 : 0;
#pragma unroll
      for (int tt = 0; tt < MAXT; tt++) {
        const int t = tbase + w + WAVES * tt;
#pragma unroll
        for (int ks = 0; ks < NKS; ks++) {
          const int kg = k0 / 4 + ks;
          const bool ok = t < RT && kg < kp / 4;
          const long io = ok ? pbase + (long)t * (kp / 4) * 64 + (long)kg * 64 + lane : lane;
          xP[tt][ks] = UP ? tabP[io] : 0.0;
          xH[tt][ks] = UH ? tabH[io] : 0.0;
        }
      }
    };
    auto stage = [&]() {
#pragma unroll
      for (int i = 0; i < NIN; i++)
#pragma unroll
        for (int q = 0; q < EPT; q++) {
          const int e = tid + NT * q;
          if (e < KC * CB) {
            const int kk = e / CB, cc = e % CB;
            sB[i][0][kk * PITCH + cc] = bok[q] ? bpr[q] * bre[i][q] : 0.0;
            sB[i][1][kk * PITCH + cc] = bok[q] && m > 0 ? bpr[q] * bim[i][q] : 0.0;
          }
        }
    };
    // P-terms land in sym for even parity, anti for odd; H-terms the opposite.  The
    // parity is uniform per chunk and a template argument, so accumulator selection is static.
    auto compute_p = [&](auto PARITY, int kvalid, const double (&xPa)[MAXT][NKS],
                         const double (&xHa)[MAXT][NKS]) {
      constexpr int sP = decltype(PARITY)::value, sH = 1 - decltype(PARITY)::value;
#pragma unroll
      for (int ks = 0; ks < NKS; ks++) {
        if (ks < kvalid) {
        const int krow = ks * 4 + (lane >> 4);
        double b[NIN][NCT][2];
#pragma unroll
        for (int i = 0; i < NIN; i++)
#pragma unroll
          for (int ct = 0; ct < NCT; ct++) {
            const int col = ct * TILE + (lane & 15);
            b[i][ct][0] = sB[i][0][krow * PITCH + col];
            b[i][ct][1] = sB[i][1][krow * PITCH + col];
          }
#pragma unroll
        for (int tt = 0; tt < MAXT; tt++) {
          const int t = tbase + w + WAVES * tt;
          if (t >= RT) continue;
          const double xP = xPa[tt][ks], xH = xHa[tt][ks];
#pragma unroll
          for (int ct = 0; ct < NCT; ct++) {
            if constexpr (OP == I_VAL || OP == I_GRADLAMBDA) {
#pragma unroll
              for (int r = 0; r < 2; r++) acc[0][tt][ct][r][sP] = mfma(xP, b[0][ct][r], acc[0][tt][ct][r][sP]);
            } else if constexpr (OP == I_GRADMU) {
#pragma unroll
              for (int r = 0; r < 2; r++) acc[0][tt][ct][r][sH] = mfma(xH, b[0][ct][r], acc[0][tt][ct][r][sH]);
            } else if constexpr (OP == I_VGRAD) {
#pragma unroll
              for (int r = 0; r < 2; r++) {
                acc[0][tt][ct][r][sP] = mfma(xP, b[0][ct][r], acc[0][tt][ct][r][sP]);
                acc[1][tt][ct][r][sH] = mfma(xH, b[0][ct][r], acc[1][tt][ct][r][sH]);
              }
            } else if constexpr (OP == I_ALPHA) {
              // i m P.A + H.B
              acc[0][tt][ct][0][sP] = mfma(xP, -dm * b[0][ct][1], acc[0][tt][ct][0][sP]);
              acc[0][tt][ct][1][sP] = mfma(xP, dm * b[0][ct][0], acc[0][tt][ct][1][sP]);
              acc[0][tt][ct][0][sH] = mfma(xH, b[1][ct][0], acc[0][tt][ct][0][sH]);
              acc[0][tt][ct][1][sH] = mfma(xH, b[1][ct][1], acc[0][tt][ct][1][sH]);
            } else {  // I_VORDIV2UV: inputs psi (0), chi (1) after the prologue
              // U = i m P.chi - H.psi
              acc[0][tt][ct][0][sP] = mfma(xP, -dm * b[1][ct][1], acc[0][tt][ct][0][sP]);
              acc[0][tt][ct][1][sP] = mfma(xP, dm * b[1][ct][0], acc[0][tt][ct][1][sP]);
              acc[0][tt][ct][0][sH] = mfma(-xH, b[0][ct][0], acc[0][tt][ct][0][sH]);
              acc[0][tt][ct][1][sH] = mfma(-xH, b[0][ct][1], acc[0][tt][ct][1][sH]);
              // V = i m P.psi + H.chi
              acc[1][tt][ct][0][sP] = mfma(xP, -dm * b[0][ct][1], acc[1][tt][ct][0][sP]);
              acc[1][tt][ct][1][sP] = mfma(xP, dm * b[0][ct][0], acc[1][tt][ct][1][sP]);
              acc[1][tt][ct][0][sH] = mfma(xH, b[1][ct][0], acc[1][tt][ct][0][sH]);
              acc[1][tt][ct][1][sH] = mfma(xH, b[1][ct][1], acc[1][tt][ct][1][sH]);
            }
          }
        }
        }
      }
    };
    auto compute = [&](int ch, const double (&xPa)[MAXT][NKS], const double (&xHa)[MAXT][NKS]) {
      int p, k0;
      chunk_pos(ch, p, k0);
      const int kvalid = (p ? kp1 : kp0) / 4 - k0 / 4;
      if (p == 0)
        compute_p(std::integral_constant<int, 0>(), kvalid, xPa, xHa);
      else
        compute_p(std::integral_constant<int, 1>(), kvalid, xPa, xHa);
    };

    load_b(0);
    load_a(0, aP0, aH0);
    for (int ch = 0; ch < nch; ch += 2) {
      stage();
      __syncthreads();
      if (ch + 1 < nch) {
        load_b(ch + 1);
        load_a(ch + 1, aP1, aH1);
      }
      compute(ch, aP0, aH0);
      __syncthreads();
      if (ch + 1 >= nch) break;
      stage();
      __syncthreads();
      if (ch + 2 < nch) {
        load_b(ch + 2);
        load_a(ch + 2, aP0, aH0);
      }
      compute(ch + 1, aP1, aH1);
      __syncthreads();
    }

    // ---- epilogue: north = sym + anti, south = sym - anti, row scale
#pragma unroll
    for (int tt = 0; tt < MAXT; tt++) {
      const int t = tbase + w + WAVES * tt;
      if (t >= RT) continue;
#pragma unroll
      for (int ct = 0; ct < NCT; ct++) {
        const int c = c0 + ct * TILE + (lane & 15);
        if (c >= a.ncols) continue;
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int jr = t * TILE + acc_row(lane, i);
          if (jr >= J2) continue;
          const int jN = J2 + jr, jS = J2 - 1 - jr;
          const double sN = a.rs ? a.rs[jN] : 1.0, sS = a.rs ? a.rs[jS] : 1.0;
#pragma unroll
          for (int o = 0; o < NOUT; o++) {
            // I_VGRAD: output 0 = P result (no row scale), 1 = i m P result, 2 = H result
            const int ai = OP == I_VGRAD ? (o == 2 ? 1 : 0) : o;
            const bool deriv = OP == I_GRADLAMBDA || (OP == I_VGRAD && o == 1);
            const bool rowsc = !(OP == I_VGRAD && o == 0);
            double2* Gb = o == 0 ? a.G0 : (o == 1 ? a.G1 : a.G2);
            if (OP == I_VGRAD && !Gb) continue;
            double nre = acc[ai][tt][ct][0][0][i] + acc[ai][tt][ct][0][1][i];
            double nim = acc[ai][tt][ct][1][0][i] + acc[ai][tt][ct][1][1][i];
            double sre = acc[ai][tt][ct][0][0][i] - acc[ai][tt][ct][0][1][i];
            double sim = acc[ai][tt][ct][1][0][i] - acc[ai][tt][ct][1][1][i];
            if (deriv) {
              double t1 = -dm * nim;
              nim = dm * nre;
              nre = t1;
              t1 = -dm * sim;
              sim = dm * sre;
              sre = t1;
            }
            const double fN = rowsc ? sN : 1.0, fS = rowsc ? sS : 1.0;
            double2* G = Gb + (long)m * a.gmst;
            G[(long)jN * a.gld + c] = make_double2(nre * fN, m ? nim * fN : 0.0);
            G[(long)jS * a.gld + c] = make_double2(sre * fS, m ? sim * fS : 0.0);
          }
        }
      }
    }
  }
}

// ------------------------------------------------------------------ launchers
template <int OP, int CB, int WAVES, int MAXT, int KC>
static int fwd_launch(const dsht_plan* p, const LegFwdArgs& a, hipStream_t s) {
  const dim3 grid = xcd_grid(p->N, a.ncols, CB);
  hipLaunchKernelGGL((k_leg_fwd<OP, CB, WAVES, MAXT, KC>), grid, dim3(WAVES * 64), 0, s, a);
  DSHT_HIP(hipGetLastError());
  return 0;
}

template <int OP, int CB, int WAVES, int MAXT, int KC>
static int inv_launch(const dsht_plan* p, const LegInvArgs& a, hipStream_t s) {
  const dim3 grid = xcd_grid(p->N, a.ncols, CB);
  hipLaunchKernelGGL((k_leg_inv<OP, CB, WAVES, MAXT, KC>), grid, dim3(WAVES * 64), 0, s, a);
  DSHT_HIP(hipGetLastError());
  return 0;
}

// forward: row tiles of the heaviest wavenumber (m = 0) decide waves and tiles per wave
template <int OP, int CB>
static int fwd_pick(const dsht_plan* p, const LegFwdArgs& a, hipStream_t s) {
  const int RT0 = cdiv(n_par(p->N, 0, 0), TILE) + cdiv(n_par(p->N, 0, 1), TILE);
  // two-input forms keep 16-column blocks when a wave owns several tiles (register budget)
  constexpr bool TWO = FwdTraits<OP>::NIN == 2;
  constexpr int CB1 = FwdTraits<OP>::NOUT == 2 ? 16 : CB, CB2 = TWO ? 16 : CB;
  constexpr int KC2 = TWO ? 16 : 32;
  if (RT0 <= 4) return fwd_launch<OP, CB1, 4, 1, 32>(p, a, s);
  if (RT0 <= 8) return fwd_launch<OP, CB1, 8, 1, 32>(p, a, s);
  if (RT0 <= 16) return fwd_launch<OP, CB2, 8, 2, KC2>(p, a, s);
  return fwd_launch<OP, CB2, 8, 3, 16>(p, a, s);
}

template <int OP, int CB>
static int inv_pick(const dsht_plan* p, const LegInvArgs& a, hipStream_t s) {
  const int RT = a.J2pad / TILE;
  // both tables: half the K chunk; the fused value/gradient form with 32-column blocks holds
  // 3 outputs' epilogue beside its accumulators, so its K chunk is halved again
  constexpr int KC = (InvTraits<OP>::USE_P && InvTraits<OP>::USE_H) ? (OP == I_VGRAD && CB == 32 ? 16 : 32) : 64;
  constexpr int CB1 = InvTraits<OP>::NOUT == 2 ? 16 : CB;
  if (RT <= 2) return inv_launch<OP, CB1, 2, 1, KC>(p, a, s);
  if (RT <= 4) return inv_launch<OP, CB1, 4, 1, KC>(p, a, s);
  return inv_launch<OP, CB1, 8, 1, KC>(p, a, s);
}

// Algorithmic work of one contraction launch: every (input, table) term costs
// 2 J T flops per column (even/odd split, re and im), T = (N+1)(N+2)/2; bytes are the
// Fourier and spectral arrays touched once plus one read of each table used.
static void leg_work(const dsht_plan* p, int nin, int nout, int nterm, int ntab, bool fwd, int ncols,
                     double* bytes, double* flops) {
  double T = 0.5 * (p->N + 1.0) * (p->N + 2.0);
  double four = (double)(p->N + 1) * p->J * 16.0 * ncols, spec = (double)p->L * 8.0 * ncols;
  double tab = 8.0 * (fwd ? p->sizeF : p->sizeI) * ntab;
  *bytes = fwd ? nin * four + nout * spec + tab : nin * spec + nout * four + tab;
  *flops = 2.0 * p->J * T * ncols * nterm;
}

int launch_leg_fwd(const dsht_plan* p, int op, const LegFwdArgs& a, hipStream_t s) {
  const bool wide = p->leg_cb ? p->leg_cb == 32 : a.ncols > 16;
  double by, fl;
  static const int nin[] = {1, 1, 1, 2, 2}, nout[] = {1, 1, 1, 1, 2}, nt[] = {1, 1, 1, 2, 4}, ntab[] = {1, 1, 1, 2, 2};
  leg_work(p, nin[op], nout[op], nt[op], ntab[op], true, a.ncols, &by, &fl);
  ProfScope ps("leg_fwd", s, by, fl);
  // forward: the per-wave fragment-stream kernel by default (measured 165 against 174 us at
  // T170 x 360); DSHT_TUNE_LEG = 1 or 3 selects the LDS-staged kernel
  if (p->leg_kind == 0 || p->leg_kind == 2) return launch_leg_fwd_direct(p, op, a, s);
  switch (op) {
    case F_VAL: return wide ? fwd_pick<F_VAL, 32>(p, a, s) : fwd_pick<F_VAL, 16>(p, a, s);
    case F_DIVLAMBDA: return wide ? fwd_pick<F_DIVLAMBDA, 32>(p, a, s) : fwd_pick<F_DIVLAMBDA, 16>(p, a, s);
    case F_DIVMU: return wide ? fwd_pick<F_DIVMU, 32>(p, a, s) : fwd_pick<F_DIVMU, 16>(p, a, s);
    case F_ALPHA: return wide ? fwd_pick<F_ALPHA, 32>(p, a, s) : fwd_pick<F_ALPHA, 16>(p, a, s);
    case F_UV2VORDIV: return wide ? fwd_pick<F_UV2VORDIV, 32>(p, a, s) : fwd_pick<F_UV2VORDIV, 16>(p, a, s);
  }
  return fail(DSHT_EARG, "launch_leg_fwd: bad op %d", op);
}

int launch_leg_inv(const dsht_plan* p, int op, const LegInvArgs& a, hipStream_t s) {
  // the fused value/gradient form defaults to 16-column blocks (its 32-column form needs the
  // shorter K chunk to fit its registers; measured 185 against 191 us at T170 x 240)
  const bool wide = p->leg_cb ? p->leg_cb == 32 : (op != I_VGRAD && a.ncols > 16);
  double by, fl;
  static const int nin[] = {1, 1, 1, 2, 2, 1}, nout[] = {1, 1, 1, 1, 2, 3}, nt[] = {1, 1, 1, 2, 4, 2},
                   ntab[] = {1, 1, 1, 2, 2, 2};
  if (op < 0 || op > I_VGRAD) return fail(DSHT_EARG, "launch_leg_inv: bad op %d", op);
  // the fused value/gradient form counts the Fourier arrays it writes (GradLon may come
  // from the derivative FFT instead)
  const int nw = op == I_VGRAD ? (a.G0 ? 1 : 0) + (a.G1 ? 1 : 0) + (a.G2 ? 1 : 0) : nout[op];
  leg_work(p, nin[op], nw, nt[op], ntab[op], false, a.ncols, &by, &fl);
  ProfScope ps("leg_inv", s, by, fl);
  if (p->leg_kind == 2) return launch_leg_inv_direct(p, op, a, s);
  switch (op) {
    case I_VAL: return wide ? inv_pick<I_VAL, 32>(p, a, s) : inv_pick<I_VAL, 16>(p, a, s);
    case I_GRADLAMBDA: return wide ? inv_pick<I_GRADLAMBDA, 32>(p, a, s) : inv_pick<I_GRADLAMBDA, 16>(p, a, s);
    case I_GRADMU: return wide ? inv_pick<I_GRADMU, 32>(p, a, s) : inv_pick<I_GRADMU, 16>(p, a, s);
    case I_ALPHA: return wide ? inv_pick<I_ALPHA, 32>(p, a, s) : inv_pick<I_ALPHA, 16>(p, a, s);
    case I_VORDIV2UV: return wide ? inv_pick<I_VORDIV2UV, 32>(p, a, s) : inv_pick<I_VORDIV2UV, 16>(p, a, s);
    case I_VGRAD: return wide ? inv_pick<I_VGRAD, 32>(p, a, s) : inv_pick<I_VGRAD, 16>(p, a, s);
  }
  return fail(DSHT_EARG, "launch_leg_inv: bad op %d", op);
}

}  // namespace dsht
