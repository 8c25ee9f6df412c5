// legendre_f64d.hip -- FP64 Legendre contractions as per-wave fragment streams.
//
// Same operators and tables as legendre_f64.hip (the Legendre half of SPMODEL w_xy / xy_w
// and their derivative forms, doc/description/ogcm/space_discretization.tex:199-268), on
// v_mfma_f64_16x16x4f64.  In FP64 a B fragment is one double per lane, and lane
// (c, k) = (l%16, l/16) of a 16-column tile reads 16 consecutive columns of one
// coefficient row (LC spectra) or one latitude row (Fourier intermediate).  So every wave
// loads both operands straight from memory with coalesced 128-byte runs: no LDS, no
// barrier, no data staging.  Loads run U K-steps ahead of the MFMAs in two register slots.
// Blocks of one zonal wavenumber share an XCD (tables and data stay in its L2) and
// wavenumbers alternate over the XCDs.
#include <type_traits>

#include "dsht_internal.hpp"

namespace dsht {

namespace {

__device__ __forceinline__ double4_t mfma64(double a, double b, double4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
// lane layout: A[r][k] at lane r + 16k, B[k][c] at lane c + 16k, D[l/16 + 4i][l%16]
__device__ __forceinline__ int acc_row64(int lane, int i) { return (lane >> 4) + 4 * i; }

// block b runs wavenumber m = b mod 8 + 8 * (b / 8 / cblk) (XCD = b mod 8) and sub-block
// (b / 8) mod cblk
__device__ __forceinline__ bool xcd_decode64(int N, int cblk, int& m, int& sub) {
  const int b = blockIdx.x;
  const int idx = b >> 3;
  m = (b & 7) + 8 * (idx / cblk);
  sub = idx % cblk;
  return m <= N;
}


// ---------------------------------------------------------------- inverse
// block = 4 waves of one m and NCT column tiles; wave w owns MAXT latitude tiles from
// (rg*4 + w)*MAXT.  P of parity p accumulates into sym (p = 0) or anti (p = 1), H the other;
// the i*m factors of ALPHA/VORDIV2UV are folded into the data (exact in FP64 up to one
// rounding, as the LDS-staged kernels do), GRADLAMBDA's into the epilogue.
// U: K steps (4 wavenumbers each) per register slot
template <int OP, int NCT, int MAXT, int U>
__global__ __launch_bounds__(256) void k_f64_inv(LegInvArgs a, int CG, int RG) {
  using Tr = InvTraits<OP>;
  constexpr int NIN = Tr::NIN, NOUT = Tr::NOUT;
  constexpr int NACC = OP == I_VGRAD ? 2 : NOUT;  // I_VGRAD: P and H products of one input
  constexpr bool UP = Tr::USE_P, UH = Tr::USE_H;
  int m, sub;
  if (!xcd_decode64(a.N, CG * RG, m, sub)) return;
  const int cg = sub % CG, rg = sub / CG;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int RT = a.J2pad / TILE;
  const int t0 = (rg * 4 + w) * MAXT;
  if (t0 >= RT) return;
  const int nE = n_par(a.N, m, 0), nO = n_par(a.N, m, 1);
  const int kp0 = cdiv(nE, KCH) * KCH, kp1 = cdiv(nO, KCH) * KCH;
  const double* tabP = a.tabP + a.offI[m];
  const double* tabH = a.tabH + a.offI[m];
  const int ct0 = cg * NCT, krow = lane >> 4, col = lane & 15;
  const double dm = (double)m;
  const double* Win[2] = {a.W0, a.W1};

  double4_t acc[NACC][MAXT][NCT][2][2];
#pragma unroll
  for (int o = 0; o < NACC; o++)
#pragma unroll
    for (int tt = 0; tt < MAXT; tt++)
#pragma unroll
      for (int ct = 0; ct < NCT; ct++)
#pragma unroll
        for (int r = 0; r < 2; r++) acc[o][tt][ct][r][0] = acc[o][tt][ct][r][1] = double4_t{0, 0, 0, 0};

  auto run = [&](auto PAR) {
    constexpr int PP = decltype(PAR)::value, sP = PP, sH = 1 - PP;
    const int kp = PP ? kp1 : kp0, np = PP ? nO : nE;
    const long pbase = PP ? (long)a.J2pad * kp0 : 0;
    const int nks = kp / 4;
    if (nks == 0) return;
    double xa[2][U][2][MAXT];         // [slot][step][P/H][tile]
    double xb[2][U][NIN][NCT][2];     // [slot][step][input][tile][re/im]
    auto load = [&](int ks0, int slot) {
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int ks = ks0 + u;
        const bool ok = ks < nks;
#pragma unroll
        for (int tt = 0; tt < MAXT; tt++) {
          const int t = t0 + tt;
          const long io = pbase + (long)t * nks * 64 + (long)ks * 64 + lane;
          const bool okt = ok && t < RT;
          xa[slot][u][0][tt] = (UP && okt) ? tabP[io] : 0.0;
          xa[slot][u][1][tt] = (UH && okt) ? tabH[io] : 0.0;
        }
        const int k = ks * 4 + krow;
        const int n = m + PP + 2 * k;
        const long l0 = (long)n * (n + 1);
        const double pr = (ok && k < np && a.pro) ? a.pro[l0 + m] : 1.0;
#pragma unroll
        for (int i = 0; i < NIN; i++)
#pragma unroll
          for (int ct = 0; ct < NCT; ct++) {
            const int c = (ct0 + ct) * TILE + col;
            double re = 0.0, im = 0.0;
            if (ok && k < np && c < a.ncols) {
              re = pr * Win[i][(l0 + m) * a.sl + (long)c * a.sc];
              if (m > 0) im = pr * Win[i][(l0 - m) * a.sl + (long)c * a.sc];
            }
            xb[slot][u][i][ct][0] = re;
            xb[slot][u][i][ct][1] = im;
          }
      }
    };
    auto comp = [&](int slot, int ks0) {
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (ks0 + u < nks) {  // predicated, so the register slots stay statically indexed
#pragma unroll
        for (int tt = 0; tt < MAXT; tt++) {
          const double xP = xa[slot][u][0][tt], xH = xa[slot][u][1][tt];
#pragma unroll
          for (int ct = 0; ct < NCT; ct++) {
            const double(&b0)[2] = xb[slot][u][0][ct];
            if constexpr (OP == I_VAL || OP == I_GRADLAMBDA) {
#pragma unroll
              for (int r = 0; r < 2; r++) acc[0][tt][ct][r][sP] = mfma64(xP, b0[r], acc[0][tt][ct][r][sP]);
            } else if constexpr (OP == I_GRADMU) {
#pragma unroll
              for (int r = 0; r < 2; r++) acc[0][tt][ct][r][sH] = mfma64(xH, b0[r], acc[0][tt][ct][r][sH]);
            } else if constexpr (OP == I_VGRAD) {
#pragma unroll
              for (int r = 0; r < 2; r++) {
                acc[0][tt][ct][r][sP] = mfma64(xP, b0[r], acc[0][tt][ct][r][sP]);
                acc[NACC - 1][tt][ct][r][sH] = mfma64(xH, b0[r], acc[NACC - 1][tt][ct][r][sH]);
              }
            } else {
              const double(&b1)[2] = xb[slot][u][NIN - 1][ct];
              if constexpr (OP == I_ALPHA) {  // i m P.A + H.B
                acc[0][tt][ct][0][sP] = mfma64(xP, -dm * b0[1], acc[0][tt][ct][0][sP]);
                acc[0][tt][ct][1][sP] = mfma64(xP, dm * b0[0], acc[0][tt][ct][1][sP]);
                acc[0][tt][ct][0][sH] = mfma64(xH, b1[0], acc[0][tt][ct][0][sH]);
                acc[0][tt][ct][1][sH] = mfma64(xH, b1[1], acc[0][tt][ct][1][sH]);
              } else {  // I_VORDIV2UV (psi = input 0, chi = input 1): U = i m P.chi - H.psi
                acc[0][tt][ct][0][sP] = mfma64(xP, -dm * b1[1], acc[0][tt][ct][0][sP]);
                acc[0][tt][ct][1][sP] = mfma64(xP, dm * b1[0], acc[0][tt][ct][1][sP]);
                acc[0][tt][ct][0][sH] = mfma64(-xH, b0[0], acc[0][tt][ct][0][sH]);
                acc[0][tt][ct][1][sH] = mfma64(-xH, b0[1], acc[0][tt][ct][1][sH]);
                // V = i m P.psi + H.chi
                acc[NOUT - 1][tt][ct][0][sP] = mfma64(xP, -dm * b0[1], acc[NOUT - 1][tt][ct][0][sP]);
                acc[NOUT - 1][tt][ct][1][sP] = mfma64(xP, dm * b0[0], acc[NOUT - 1][tt][ct][1][sP]);
                acc[NOUT - 1][tt][ct][0][sH] = mfma64(xH, b1[0], acc[NOUT - 1][tt][ct][0][sH]);
                acc[NOUT - 1][tt][ct][1][sH] = mfma64(xH, b1[1], acc[NOUT - 1][tt][ct][1][sH]);
              }
            }
          }
        }
        }
      }
    };
    load(0, 0);
    for (int ks0 = 0; ks0 < nks; ks0 += 2 * U) {
      if (ks0 + U < nks) load(ks0 + U, 1);
      comp(0, ks0);
      if (ks0 + U >= nks) break;
      if (ks0 + 2 * U < nks) load(ks0 + 2 * U, 0);
      comp(1, ks0 + U);
    }
  };
  run(std::integral_constant<int, 0>());
  run(std::integral_constant<int, 1>());

  // ---- epilogue: north = sym + anti, south = sym - anti, row scale, FP64 store
#pragma unroll
  for (int tt = 0; tt < MAXT; tt++) {
    const int t = t0 + tt;
    if (t >= RT) continue;
#pragma unroll
    for (int ct = 0; ct < NCT; ct++) {
      const int c = (ct0 + ct) * TILE + col;
      if (c >= a.ncols) continue;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int jr = t * TILE + acc_row64(lane, i);
        if (jr >= a.J2) continue;
        const int jN = a.J2 + jr, jS = a.J2 - 1 - jr;
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

// ---------------------------------------------------------------- forward
// block = 4 waves of one m and one column tile; the spectral row tiles of each parity are
// cut into slots of MAXT tiles taken by the waves in turn.  Rows of parity 0 pair P with
// c_j (G_N + G_S) and H with c_j (G_N - G_S); parity 1 the other way round.
// U: K steps (4 latitudes each) per register slot
template <int OP, int MAXT, int U>
__global__ __launch_bounds__(256) void k_f64_fwd(LegFwdArgs a, int CT) {
  using Tr = FwdTraits<OP>;
  constexpr int NIN = Tr::NIN, NOUT = Tr::NOUT;
  constexpr bool UP = Tr::USE_P, UH = Tr::USE_H;
  int m, ct;
  if (!xcd_decode64(a.N, CT, m, ct)) return;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nE = n_par(a.N, m, 0), nO = n_par(a.N, m, 1);
  const int tE = cdiv(nE, TILE), tO = cdiv(nO, TILE);
  const int sE = cdiv(tE, MAXT), sO = cdiv(tO, MAXT);
  const int nks = a.J2pad / 4;
  const double* tabP = a.tabP + a.offF[m];
  const double* tabH = a.tabH + a.offF[m];
  const int krow = lane >> 4, col = lane & 15;
  const int c = ct * TILE + col;
  const bool cok = c < a.ncols;
  const double2* Gm[2] = {a.G0 + (long)m * a.gmst, NIN > 1 ? a.G1 + (long)m * a.gmst : nullptr};
  const double dm = (double)m;

  auto slot = [&](auto PAR, int tp0) {
    constexpr int PP = decltype(PAR)::value;
    const int tpar = PP ? tO : tE;
    double4_t acc[NOUT][MAXT][2];
#pragma unroll
    for (int o = 0; o < NOUT; o++)
#pragma unroll
      for (int tt = 0; tt < MAXT; tt++) acc[o][tt][0] = acc[o][tt][1] = double4_t{0, 0, 0, 0};
    double xa[2][U][2][MAXT];     // [slot][step][P/H][tile]
    double2 gv[2][U][NIN][2];     // [slot][step][input][north/south]
    double cwv[2][U];
    auto load = [&](int ks0, int s) {
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int ks = ks0 + u;
        const bool ok = ks < nks;
#pragma unroll
        for (int tt = 0; tt < MAXT; tt++) {
          const int tp = tp0 + tt;
          const bool okt = ok && tp < tpar;
          const long fo = (long)(PP ? tE + tp : tp) * nks * 64 + (long)ks * 64 + lane;
          xa[s][u][0][tt] = (UP && okt) ? tabP[fo] : 0.0;
          xa[s][u][1][tt] = (UH && okt) ? tabH[fo] : 0.0;
        }
        const int jn = ks * 4 + krow;
        const bool okj = ok && jn < a.J2 && cok;
        cwv[s][u] = okj ? a.cw[jn] : 0.0;
#pragma unroll
        for (int i = 0; i < NIN; i++) {
          gv[s][u][i][0] = okj ? Gm[i][(long)(a.J2 + jn) * a.gld + c] : make_double2(0.0, 0.0);
          gv[s][u][i][1] = okj ? Gm[i][(long)(a.J2 - 1 - jn) * a.gld + c] : make_double2(0.0, 0.0);
        }
      }
    };
    auto comp = [&](int s, int ks0) {
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (ks0 + u < nks) {
        // b[input][q][re/im]: q = 0 pairs with P, q = 1 with H
        double b[NIN][2][2];
#pragma unroll
        for (int i = 0; i < NIN; i++) {
          const double2 gn = gv[s][u][i][0], gs = gv[s][u][i][1];
          const double cw = cwv[s][u];
          const double symr = cw * (gn.x + gs.x), symi = cw * (gn.y + gs.y);
          const double antr = cw * (gn.x - gs.x), anti = cw * (gn.y - gs.y);
          b[i][0][0] = PP ? antr : symr;
          b[i][0][1] = PP ? anti : symi;
          b[i][1][0] = PP ? symr : antr;
          b[i][1][1] = PP ? symi : anti;
        }
#pragma unroll
        for (int tt = 0; tt < MAXT; tt++) {
          const double xP = xa[s][u][0][tt], xH = xa[s][u][1][tt];
          if constexpr (OP == F_VAL || OP == F_DIVLAMBDA) {
            acc[0][tt][0] = mfma64(xP, b[0][0][0], acc[0][tt][0]);
            acc[0][tt][1] = mfma64(xP, b[0][0][1], acc[0][tt][1]);
          } else if constexpr (OP == F_DIVMU) {
            acc[0][tt][0] = mfma64(xH, b[0][1][0], acc[0][tt][0]);
            acc[0][tt][1] = mfma64(xH, b[0][1][1], acc[0][tt][1]);
          } else if constexpr (OP == F_ALPHA) {  // i m P.A - H.B
            acc[0][tt][0] = mfma64(xP, -dm * b[0][0][1], acc[0][tt][0]);
            acc[0][tt][1] = mfma64(xP, dm * b[0][0][0], acc[0][tt][1]);
            acc[0][tt][0] = mfma64(-xH, b[NIN - 1][1][0], acc[0][tt][0]);
            acc[0][tt][1] = mfma64(-xH, b[NIN - 1][1][1], acc[0][tt][1]);
          } else {  // F_UV2VORDIV, inputs U cos (0), V cos (1): vor = i m P.V + H.U
            acc[0][tt][0] = mfma64(xP, -dm * b[NIN - 1][0][1], acc[0][tt][0]);
            acc[0][tt][1] = mfma64(xP, dm * b[NIN - 1][0][0], acc[0][tt][1]);
            acc[0][tt][0] = mfma64(xH, b[0][1][0], acc[0][tt][0]);
            acc[0][tt][1] = mfma64(xH, b[0][1][1], acc[0][tt][1]);
            // div = i m P.U - H.V
            acc[NOUT - 1][tt][0] = mfma64(xP, -dm * b[0][0][1], acc[NOUT - 1][tt][0]);
            acc[NOUT - 1][tt][1] = mfma64(xP, dm * b[0][0][0], acc[NOUT - 1][tt][1]);
            acc[NOUT - 1][tt][0] = mfma64(-xH, b[NIN - 1][1][0], acc[NOUT - 1][tt][0]);
            acc[NOUT - 1][tt][1] = mfma64(-xH, b[NIN - 1][1][1], acc[NOUT - 1][tt][1]);
          }
        }
        }
      }
    };
    load(0, 0);
    for (int ks0 = 0; ks0 < nks; ks0 += 2 * U) {
      if (ks0 + U < nks) load(ks0 + U, 1);
      comp(0, ks0);
      if (ks0 + U >= nks) break;
      if (ks0 + 2 * U < nks) load(ks0 + 2 * U, 0);
      comp(1, ks0 + U);
    }
    // ---- epilogue: rotation/sign, scale, diagonal, scatter to SPMODEL order
    if (!cok) return;
#pragma unroll
    for (int tt = 0; tt < MAXT; tt++) {
      const int tp = tp0 + tt;
      if (tp >= tpar) continue;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int k = tp * TILE + acc_row64(lane, i);
        const int n = m + PP + 2 * k;
        if (n > a.N) continue;
        const long lre = (long)n * (n + 1) + m, lim = (long)n * (n + 1) - m;
#pragma unroll
        for (int o = 0; o < NOUT; o++) {
          double re = acc[o][tt][0][i], im = acc[o][tt][1][i];
          if constexpr (OP == F_DIVLAMBDA) {
            const double r2 = -dm * im;
            im = dm * re;
            re = r2;
          } else if constexpr (OP == F_DIVMU) {
            re = -re;
            im = -im;
          }
          const double* dg = o == 0 ? a.diag0 : a.diag1;
          double* W = o == 0 ? a.W0 : a.W1;
          W[lre * a.sl + (long)c * a.sc] = re * a.scale * (dg ? dg[lre] : 1.0);
          if (m > 0) W[lim * a.sl + (long)c * a.sc] = im * a.scale * (dg ? dg[lim] : 1.0);
        }
      }
    }
  };

  for (int s = w; s < sE + sO; s += 4) {
    if (s < sE)
      slot(std::integral_constant<int, 0>(), s * MAXT);
    else
      slot(std::integral_constant<int, 1>(), (s - sE) * MAXT);
  }
}

#ifndef F64D_U
#define F64D_U 2
#endif
// tile shapes: one-output operators 2 row tiles x 2 column tiles per wave (inverse) and
// 2 row tiles per wave (forward); the two-output ones 1 x 1 to bound the accumulators
template <int OP>
int inv_go(const dsht_plan* p, const LegInvArgs& a, hipStream_t s, bool narrow) {
  constexpr bool TWO = InvTraits<OP>::NOUT == 2 || InvTraits<OP>::NIN == 2;
  const int RT = a.J2pad / TILE;
  auto go = [&](auto NCTc, auto MAXTc) {
    constexpr int NCT = decltype(NCTc)::value, MAXT = decltype(MAXTc)::value;
    const int CT = cdiv(a.ncols, TILE), CG = cdiv(CT, NCT), RG = cdiv(RT, 4 * MAXT);
    const long nb = 8l * cdiv(p->N + 1, 8) * CG * RG;
    hipLaunchKernelGGL((k_f64_inv<OP, NCT, MAXT, F64D_U>), dim3((unsigned)nb), dim3(256), 0, s, a, CG, RG);
    DSHT_HIP(hipGetLastError());
    return 0;
  };
  if constexpr (TWO) {
    return go(std::integral_constant<int, 1>(), std::integral_constant<int, 1>());
  } else if constexpr (OP == I_VGRAD) {  // two accumulator groups: 1 row tile x 2 column tiles
    if (narrow) return go(std::integral_constant<int, 1>(), std::integral_constant<int, 1>());
    return go(std::integral_constant<int, 2>(), std::integral_constant<int, 1>());
  } else {
    if (narrow) return go(std::integral_constant<int, 1>(), std::integral_constant<int, 1>());
    return go(std::integral_constant<int, 2>(), std::integral_constant<int, 2>());
  }
}

template <int OP>
int fwd_go(const dsht_plan* p, const LegFwdArgs& a, hipStream_t s) {
  constexpr int MAXT = (FwdTraits<OP>::NOUT == 2 || FwdTraits<OP>::NIN == 2) ? 1 : 2;
  const int CT = cdiv(a.ncols, TILE);
  const long nb = 8l * cdiv(p->N + 1, 8) * CT;
  hipLaunchKernelGGL((k_f64_fwd<OP, MAXT, F64D_U>), dim3((unsigned)nb), dim3(256), 0, s, a, CT);
  DSHT_HIP(hipGetLastError());
  return 0;
}

}  // namespace

int launch_leg_inv_direct(const dsht_plan* p, int op, const LegInvArgs& a, hipStream_t s) {
  const bool narrow = p->leg_cb == 16;
  switch (op) {
    case I_VAL: return inv_go<I_VAL>(p, a, s, narrow);
    case I_GRADLAMBDA: return inv_go<I_GRADLAMBDA>(p, a, s, narrow);
    case I_GRADMU: return inv_go<I_GRADMU>(p, a, s, narrow);
    case I_ALPHA: return inv_go<I_ALPHA>(p, a, s, narrow);
    case I_VORDIV2UV: return inv_go<I_VORDIV2UV>(p, a, s, narrow);
    case I_VGRAD: return inv_go<I_VGRAD>(p, a, s, narrow);
  }
  return fail(DSHT_EARG, "launch_leg_inv_direct: bad op %d", op);
}

int launch_leg_fwd_direct(const dsht_plan* p, int op, const LegFwdArgs& a, hipStream_t s) {
  switch (op) {
    case F_VAL: return fwd_go<F_VAL>(p, a, s);
    case F_DIVLAMBDA: return fwd_go<F_DIVLAMBDA>(p, a, s);
    case F_DIVMU: return fwd_go<F_DIVMU>(p, a, s);
    case F_ALPHA: return fwd_go<F_ALPHA>(p, a, s);
    case F_UV2VORDIV: return fwd_go<F_UV2VORDIV>(p, a, s);
  }
  return fail(DSHT_EARG, "launch_leg_fwd_direct: bad op %d", op);
}

}  // namespace dsht
