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
// plus an optional per-l diagonal (filters, implicit diffusion, Helmholtz) on the spectral side.
#include "dsht_internal.hpp"

namespace dsht {

__device__ __forceinline__ double4_t mfma(double a, double b, double4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
// v_mfma_f64_16x16x4f64 lane layout: A[r][k] at lane r + 16k, B[k][c] at lane c + 16k,
// result element i of lane l holds D[l/16 + 4i][l%16].
__device__ __forceinline__ int acc_row(int lane, int i) { return (lane >> 4) + 4 * i; }

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
  static constexpr int NOUT = (OP == I_VORDIV2UV) ? 2 : 1;
  static constexpr bool USE_P = (OP != I_GRADMU);
  static constexpr bool USE_H = (OP == I_GRADMU || OP == I_ALPHA || OP == I_VORDIV2UV);
};

// ------------------------------------------------------------------ forward
template <int OP, int CB, int MAXT>
__global__ __launch_bounds__(256) void k_leg_fwd(LegFwdArgs a) {
  using Tr = FwdTraits<OP>;
  constexpr int NIN = Tr::NIN, NOUT = Tr::NOUT, NCT = CB / TILE;
  constexpr int PITCH = CB + 1;  // LDS row pitch (doubles)
  // [input][S re, S im, A re, A im][KCH][PITCH]
  __shared__ double sB[NIN][4][KCH * PITCH];

  const int m = blockIdx.y;
  const int c0 = blockIdx.x * CB;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int N = a.N, J2 = a.J2;
  const int nE = n_par(N, m, 0), nO = n_par(N, m, 1);
  const int tE = (nE + TILE - 1) / TILE, tO = (nO + TILE - 1) / TILE;
  const int RT = tE + tO;
  const long kstepsF = a.J2pad / 4;
  const double* tabP = a.tabP + a.offF[m];
  const double* tabH = a.tabH + a.offF[m];
  const double2* Gin[2] = {a.G0 + (long)m * a.gmst, a.G1 ? a.G1 + (long)m * a.gmst : nullptr};
  const double dm = (double)m;

  for (int tbase = 0; tbase < RT; tbase += 4 * MAXT) {
    double4_t acc[NOUT][MAXT][NCT][2];
#pragma unroll
    for (int o = 0; o < NOUT; o++)
#pragma unroll
      for (int t = 0; t < MAXT; t++)
#pragma unroll
        for (int ct = 0; ct < NCT; ct++) acc[o][t][ct][0] = acc[o][t][ct][1] = double4_t{0, 0, 0, 0};

    for (int k0 = 0; k0 < a.J2pad; k0 += KCH) {
      // ---- stage G(north) +/- G(south), weighted, into LDS
#pragma unroll
      for (int i = 0; i < NIN; i++) {
        for (int e = tid; e < KCH * CB; e += 256) {
          int kk = e / CB, cc = e % CB;
          int jn = k0 + kk, c = c0 + cc;
          double2 gn = make_double2(0.0, 0.0), gs = gn;
          if (jn < J2 && c < a.ncols) {
            gn = Gin[i][(long)(J2 + jn) * a.gld + c];
            gs = Gin[i][(long)(J2 - 1 - jn) * a.gld + c];
          }
          double cw = a.cw[jn];
          sB[i][0][kk * PITCH + cc] = cw * (gn.x + gs.x);
          sB[i][1][kk * PITCH + cc] = cw * (gn.y + gs.y);
          sB[i][2][kk * PITCH + cc] = cw * (gn.x - gs.x);
          sB[i][3][kk * PITCH + cc] = cw * (gn.y - gs.y);
        }
      }
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < KCH / 4; ks++) {
        const int krow = ks * 4 + (lane >> 4);
        // B fragments: [input][sym/anti][ct][re/im]
        double b[NIN][2][NCT][2];
#pragma unroll
        for (int i = 0; i < NIN; i++)
#pragma unroll
          for (int ct = 0; ct < NCT; ct++) {
            int col = ct * TILE + (lane & 15);
            b[i][0][ct][0] = sB[i][0][krow * PITCH + col];
            b[i][0][ct][1] = sB[i][1][krow * PITCH + col];
            b[i][1][ct][0] = sB[i][2][krow * PITCH + col];
            b[i][1][ct][1] = sB[i][3][krow * PITCH + col];
          }
#pragma unroll
        for (int tt = 0; tt < MAXT; tt++) {
          const int t = tbase + tt * 4 + w;
          if (t >= RT) continue;
          const int p = t < tE ? 0 : 1;
          const long fo = (long)t * kstepsF * 64 + (long)(k0 / 4 + ks) * 64 + lane;
          // P pairs with (sym for even rows, anti for odd); H with the opposite
          const int gP = p, gH = 1 - p;
          if constexpr (OP == F_VAL || OP == F_DIVLAMBDA) {
            double aP = tabP[fo];
#pragma unroll
            for (int ct = 0; ct < NCT; ct++) {
              acc[0][tt][ct][0] = mfma(aP, b[0][gP][ct][0], acc[0][tt][ct][0]);
              acc[0][tt][ct][1] = mfma(aP, b[0][gP][ct][1], acc[0][tt][ct][1]);
            }
          } else if constexpr (OP == F_DIVMU) {
            double aH = tabH[fo];
#pragma unroll
            for (int ct = 0; ct < NCT; ct++) {
              acc[0][tt][ct][0] = mfma(aH, b[0][gH][ct][0], acc[0][tt][ct][0]);
              acc[0][tt][ct][1] = mfma(aH, b[0][gH][ct][1], acc[0][tt][ct][1]);
            }
          } else if constexpr (OP == F_ALPHA) {
            // i m P.A - H.B :  (i m z) = (-m z.im, m z.re)
            double aP = tabP[fo], aHn = -tabH[fo];
#pragma unroll
            for (int ct = 0; ct < NCT; ct++) {
              acc[0][tt][ct][0] = mfma(aP, -dm * b[0][gP][ct][1], acc[0][tt][ct][0]);
              acc[0][tt][ct][1] = mfma(aP, dm * b[0][gP][ct][0], acc[0][tt][ct][1]);
              acc[0][tt][ct][0] = mfma(aHn, b[1][gH][ct][0], acc[0][tt][ct][0]);
              acc[0][tt][ct][1] = mfma(aHn, b[1][gH][ct][1], acc[0][tt][ct][1]);
            }
          } else {  // F_UV2VORDIV, inputs U cos (0), V cos (1)
            double aP = tabP[fo], aH = tabH[fo];
#pragma unroll
            for (int ct = 0; ct < NCT; ct++) {
              // vor = i m P.V + H.U
              acc[0][tt][ct][0] = mfma(aP, -dm * b[1][gP][ct][1], acc[0][tt][ct][0]);
              acc[0][tt][ct][1] = mfma(aP, dm * b[1][gP][ct][0], acc[0][tt][ct][1]);
              acc[0][tt][ct][0] = mfma(aH, b[0][gH][ct][0], acc[0][tt][ct][0]);
              acc[0][tt][ct][1] = mfma(aH, b[0][gH][ct][1], acc[0][tt][ct][1]);
              // div = i m P.U - H.V
              acc[1][tt][ct][0] = mfma(aP, -dm * b[0][gP][ct][1], acc[1][tt][ct][0]);
              acc[1][tt][ct][1] = mfma(aP, dm * b[0][gP][ct][0], acc[1][tt][ct][1]);
              acc[1][tt][ct][0] = mfma(-aH, b[1][gH][ct][0], acc[1][tt][ct][0]);
              acc[1][tt][ct][1] = mfma(-aH, b[1][gH][ct][1], acc[1][tt][ct][1]);
            }
          }
        }
      }
      __syncthreads();
    }

    // ---- epilogue: rotation/sign, scale, diagonal, scatter to SPMODEL order
#pragma unroll
    for (int tt = 0; tt < MAXT; tt++) {
      const int t = tbase + tt * 4 + w;
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
            double sre = a.scale * (dg ? dg[lre] : 1.0);
            W[lre * a.sl + (long)c * a.sc] = re * sre;
            if (m > 0) {
              double sim = a.scale * (dg ? dg[lim] : 1.0);
              W[lim * a.sl + (long)c * a.sc] = im * sim;
            }
          }
        }
      }
    }
  }
}

// ------------------------------------------------------------------ inverse
template <int OP, int CB, int MAXT>
__global__ __launch_bounds__(256) void k_leg_inv(LegInvArgs a) {
  using Tr = InvTraits<OP>;
  constexpr int NIN = Tr::NIN, NOUT = Tr::NOUT, NCT = CB / TILE;
  constexpr int PITCH = CB + 1;
  __shared__ double sB[NIN][2][KCH * PITCH];  // [input][re, im][KCH][PITCH]

  const int m = blockIdx.y;
  const int c0 = blockIdx.x * CB;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int N = a.N, J2 = a.J2;
  const int nE = n_par(N, m, 0), nO = n_par(N, m, 1);
  const int kp0 = (nE + KCH - 1) / KCH * KCH, kp1 = (nO + KCH - 1) / KCH * KCH;
  const int RT = a.J2pad / TILE;
  const double* Win[2] = {a.W0, a.W1};
  const double dm = (double)m;

  for (int tbase = 0; tbase < RT; tbase += 4 * MAXT) {
    // [output][tile][ct][re/im][sym/anti]
    double4_t acc[NOUT][MAXT][NCT][2][2];
#pragma unroll
    for (int o = 0; o < NOUT; o++)
#pragma unroll
      for (int t = 0; t < MAXT; t++)
#pragma unroll
        for (int ct = 0; ct < NCT; ct++)
#pragma unroll
          for (int r = 0; r < 2; r++) acc[o][t][ct][r][0] = acc[o][t][ct][r][1] = double4_t{0, 0, 0, 0};

    for (int p = 0; p < 2; p++) {
      const int np = p ? nO : nE;
      const int kp = p ? kp1 : kp0;
      const long pbase = p ? (long)a.J2pad * kp0 : 0;
      for (int k0 = 0; k0 < kp; k0 += KCH) {
#pragma unroll
        for (int i = 0; i < NIN; i++) {
          for (int e = tid; e < KCH * CB; e += 256) {
            int kk = e / CB, cc = e % CB;
            int k = k0 + kk, c = c0 + cc;
            double re = 0.0, im = 0.0;
            if (k < np && c < a.ncols) {
              int n = m + p + 2 * k;
              long lre = (long)n * (n + 1) + m, lim = (long)n * (n + 1) - m;
              double pr = a.pro ? a.pro[lre] : 1.0;
              re = pr * Win[i][lre * a.sl + (long)c * a.sc];
              if (m > 0) im = pr * Win[i][lim * a.sl + (long)c * a.sc];
            }
            sB[i][0][kk * PITCH + cc] = re;
            sB[i][1][kk * PITCH + cc] = im;
          }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < KCH / 4; ks++) {
          const int krow = ks * 4 + (lane >> 4);
          double b[NIN][NCT][2];
#pragma unroll
          for (int i = 0; i < NIN; i++)
#pragma unroll
            for (int ct = 0; ct < NCT; ct++) {
              int col = ct * TILE + (lane & 15);
              b[i][ct][0] = sB[i][0][krow * PITCH + col];
              b[i][ct][1] = sB[i][1][krow * PITCH + col];
            }
          // P-terms land in sym for even parity, anti for odd; H-terms the opposite
          const int sP = p, sH = 1 - p;
#pragma unroll
          for (int tt = 0; tt < MAXT; tt++) {
            const int t = tbase + tt * 4 + w;
            if (t >= RT) continue;
            const long io = pbase + (long)t * (kp / 4) * 64 + (long)(k0 / 4 + ks) * 64 + lane;
            if constexpr (OP == I_VAL || OP == I_GRADLAMBDA) {
              double aP = a.tabP[a.offI[m] + io];
#pragma unroll
              for (int ct = 0; ct < NCT; ct++)
#pragma unroll
                for (int r = 0; r < 2; r++) acc[0][tt][ct][r][sP] = mfma(aP, b[0][ct][r], acc[0][tt][ct][r][sP]);
            } else if constexpr (OP == I_GRADMU) {
              double aH = a.tabH[a.offI[m] + io];
#pragma unroll
              for (int ct = 0; ct < NCT; ct++)
#pragma unroll
                for (int r = 0; r < 2; r++) acc[0][tt][ct][r][sH] = mfma(aH, b[0][ct][r], acc[0][tt][ct][r][sH]);
            } else if constexpr (OP == I_ALPHA) {
              // i m P.A + H.B
              double aP = a.tabP[a.offI[m] + io], aH = a.tabH[a.offI[m] + io];
#pragma unroll
              for (int ct = 0; ct < NCT; ct++) {
                acc[0][tt][ct][0][sP] = mfma(aP, -dm * b[0][ct][1], acc[0][tt][ct][0][sP]);
                acc[0][tt][ct][1][sP] = mfma(aP, dm * b[0][ct][0], acc[0][tt][ct][1][sP]);
                acc[0][tt][ct][0][sH] = mfma(aH, b[1][ct][0], acc[0][tt][ct][0][sH]);
                acc[0][tt][ct][1][sH] = mfma(aH, b[1][ct][1], acc[0][tt][ct][1][sH]);
              }
            } else {  // I_VORDIV2UV: inputs psi (0), chi (1) after the prologue
              double aP = a.tabP[a.offI[m] + io], aH = a.tabH[a.offI[m] + io];
#pragma unroll
              for (int ct = 0; ct < NCT; ct++) {
                // U = i m P.chi - H.psi
                acc[0][tt][ct][0][sP] = mfma(aP, -dm * b[1][ct][1], acc[0][tt][ct][0][sP]);
                acc[0][tt][ct][1][sP] = mfma(aP, dm * b[1][ct][0], acc[0][tt][ct][1][sP]);
                acc[0][tt][ct][0][sH] = mfma(-aH, b[0][ct][0], acc[0][tt][ct][0][sH]);
                acc[0][tt][ct][1][sH] = mfma(-aH, b[0][ct][1], acc[0][tt][ct][1][sH]);
                // V = i m P.psi + H.chi
                acc[1][tt][ct][0][sP] = mfma(aP, -dm * b[0][ct][1], acc[1][tt][ct][0][sP]);
                acc[1][tt][ct][1][sP] = mfma(aP, dm * b[0][ct][0], acc[1][tt][ct][1][sP]);
                acc[1][tt][ct][0][sH] = mfma(aH, b[1][ct][0], acc[1][tt][ct][0][sH]);
                acc[1][tt][ct][1][sH] = mfma(aH, b[1][ct][1], acc[1][tt][ct][1][sH]);
              }
            }
          }
        }
        __syncthreads();
      }
    }

    // ---- epilogue: north = sym + anti, south = sym - anti, row scale
#pragma unroll
    for (int tt = 0; tt < MAXT; tt++) {
      const int t = tbase + tt * 4 + w;
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
            double nre = acc[o][tt][ct][0][0][i] + acc[o][tt][ct][0][1][i];
            double nim = acc[o][tt][ct][1][0][i] + acc[o][tt][ct][1][1][i];
            double sre = acc[o][tt][ct][0][0][i] - acc[o][tt][ct][0][1][i];
            double sim = acc[o][tt][ct][1][0][i] - acc[o][tt][ct][1][1][i];
            if constexpr (OP == I_GRADLAMBDA) {
              double t1 = -dm * nim;
              nim = dm * nre;
              nre = t1;
              t1 = -dm * sim;
              sim = dm * sre;
              sre = t1;
            }
            double2* G = (o == 0 ? a.G0 : a.G1) + (long)m * a.gmst;
            G[(long)jN * a.gld + c] = make_double2(nre * sN, m ? nim * sN : 0.0);
            G[(long)jS * a.gld + c] = make_double2(sre * sS, m ? sim * sS : 0.0);
          }
        }
      }
    }
  }
}

// ------------------------------------------------------------------ launchers
template <int OP, int CB>
static int fwd_launch(const dsht_plan* p, const LegFwdArgs& a, hipStream_t s) {
  dim3 grid(cdiv(a.ncols, CB), p->N + 1);
  hipLaunchKernelGGL((k_leg_fwd<OP, CB, 3>), grid, dim3(256), 0, s, a);
  DSHT_HIP(hipGetLastError());
  return 0;
}

template <int OP, int CB>
static int inv_launch(const dsht_plan* p, const LegInvArgs& a, hipStream_t s) {
  dim3 grid(cdiv(a.ncols, CB), p->N + 1);
  hipLaunchKernelGGL((k_leg_inv<OP, CB, 2>), grid, dim3(256), 0, s, a);
  DSHT_HIP(hipGetLastError());
  return 0;
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
  const bool wide = a.ncols > 16;
  double by, fl;
  static const int nin[] = {1, 1, 1, 2, 2}, nout[] = {1, 1, 1, 1, 2}, nt[] = {1, 1, 1, 2, 4}, ntab[] = {1, 1, 1, 2, 2};
  leg_work(p, nin[op], nout[op], nt[op], ntab[op], true, a.ncols, &by, &fl);
  ProfScope ps("leg_fwd", s, by, fl);
  switch (op) {
    case F_VAL: return wide ? fwd_launch<F_VAL, 32>(p, a, s) : fwd_launch<F_VAL, 16>(p, a, s);
    case F_DIVLAMBDA:
      return wide ? fwd_launch<F_DIVLAMBDA, 32>(p, a, s) : fwd_launch<F_DIVLAMBDA, 16>(p, a, s);
    case F_DIVMU: return wide ? fwd_launch<F_DIVMU, 32>(p, a, s) : fwd_launch<F_DIVMU, 16>(p, a, s);
    case F_ALPHA: return fwd_launch<F_ALPHA, 16>(p, a, s);
    case F_UV2VORDIV: return fwd_launch<F_UV2VORDIV, 16>(p, a, s);
  }
  return fail(DSHT_EARG, "launch_leg_fwd: bad op %d", op);
}

int launch_leg_inv(const dsht_plan* p, int op, const LegInvArgs& a, hipStream_t s) {
  const bool wide = a.ncols > 16;
  double by, fl;
  static const int nin[] = {1, 1, 1, 2, 2}, nout[] = {1, 1, 1, 1, 2}, nt[] = {1, 1, 1, 2, 4}, ntab[] = {1, 1, 1, 2, 2};
  leg_work(p, nin[op], nout[op], nt[op], ntab[op], false, a.ncols, &by, &fl);
  ProfScope ps("leg_inv", s, by, fl);
  switch (op) {
    case I_VAL: return wide ? inv_launch<I_VAL, 32>(p, a, s) : inv_launch<I_VAL, 16>(p, a, s);
    case I_GRADLAMBDA:
      return wide ? inv_launch<I_GRADLAMBDA, 32>(p, a, s) : inv_launch<I_GRADLAMBDA, 16>(p, a, s);
    case I_GRADMU: return wide ? inv_launch<I_GRADMU, 32>(p, a, s) : inv_launch<I_GRADMU, 16>(p, a, s);
    case I_ALPHA: return inv_launch<I_ALPHA, 16>(p, a, s);
    case I_VORDIV2UV: return inv_launch<I_VORDIV2UV, 16>(p, a, s);
  }
  return fail(DSHT_EARG, "launch_leg_inv: bad op %d", op);
}

}  // namespace dsht
