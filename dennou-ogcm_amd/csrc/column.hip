// column.hip -- vertical column diagnostics on device grids, between the inverse and the
// forward transforms of a dynamics step.
//   HBEDiagnose_OMG     (model/dogcm/src/dynamics/hspm_vfvm/HBEDiagnose_hspm_vfvm_mod.f90:90-143)
//   HBEDiagnose_UVBarot (same file:221-255)
//   HBEDiagnose_HydPres (same file:290-325)
//   calc_ADVFlxZ_QUICK  (model/dogcm/src/dynamics/hspm_vfvm/HBEBaroc_hspm_vfvm_mod.f90:233-317)
//   calc_VDiffTRCEq     (model/dogcm/src/physics/DOGCM_Phys_hspm_vfvm_mod.f90:752-861)
//   calc_VDiffMOMEq     (same file:865-975)
//
// Layout: the km layers of one field are km consecutive device grid columns (column pitch
// ldg doubles, layer 0 on top = the reference's KS), each an I x J grid with longitude
// fastest, as the transforms read and write them.  Half levels (the reference's r
// arrays, KS-1..KE) are km+1 columns, half level k lying on top of layer k.
//
// One thread per grid point walks its column in the reference's order: loads of CU layers
// are issued together (adjacent threads read adjacent points of one layer, 512-byte runs
// per wave), then the running sum advances through them.  Every product and sum is the
// reference's, in its order, with no contraction into FMAs, so the results equal an FP64
// CPU evaluation of the reference loops bit for bit.  The kernels are bound by HBM: each
// input and output value is touched once, through non-temporal loads and stores (with
// cached access they ran at 5.2 TB/s on T170L60 grids, non-temporal at 5.8-6.0 TB/s).
#include "dsht_internal.hpp"

namespace dsht {

namespace {

constexpr int CU = 8;  // layers whose loads are in flight together
constexpr int VDIFF_LR = 10;  // finished rows of the vertical solve kept in LDS per thread (60 KB)

// every grid value is read or written once per call: non-temporal
__device__ __forceinline__ double ntl(const double* x) { return __builtin_nontemporal_load(x); }
__device__ __forceinline__ void nts(double* x, double v) { __builtin_nontemporal_store(v, x); }

// omg(km) = 0; omg(k) = omg(k+1) - div(k) * (h(k) * cdk(k))          (:123-131)
__global__ __launch_bounds__(256) void k_col_omega(const double* __restrict__ div, const double* __restrict__ h,
                                                   long ldg, int km, const double* __restrict__ cdk,
                                                   double* __restrict__ omg, long ldo, int npts) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= npts) return;
  double r = 0.0;
  nts(omg + (long)km * ldo + i, r);
  int k = km - 1;
  for (; k >= CU - 1; k -= CU) {
    double d[CU], t[CU];
#pragma unroll
    for (int u = 0; u < CU; u++) {
      d[u] = ntl(div + (long)(k - u) * ldg + i);
      t[u] = ntl(h + (long)(k - u) * ldg + i);
    }
#pragma unroll
    for (int u = 0; u < CU; u++) {
      const double dz = t[u] * cdk[k - u];
      const double dom = d[u] * dz;
      r = r - dom;
      nts(omg + (long)(k - u) * ldo + i, r);
    }
  }
  for (; k >= 0; k--) {
    const double dz = ntl(h + (long)k * ldg + i) * cdk[k];
    const double dom = ntl(div + (long)k * ldg + i) * dz;
    r = r - dom;
    nts(omg + (long)k * ldo + i, r);
  }
}

// d(k) = grav * h(k) * dens(k) * cdk(k); r(-1) = 0, r(k) = r(k-1) + d(k);
// p(k) = 0.5 * (r(k-1) + r(k))                                        (:315-321)
__global__ __launch_bounds__(256) void k_col_hydpres(const double* __restrict__ dens, const double* __restrict__ h,
                                                     long ldg, int km, const double* __restrict__ cdk, double grav,
                                                     double* __restrict__ p, long ldo, int npts) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= npts) return;
  double r = 0.0;
  int k = 0;
  for (; k + CU <= km; k += CU) {
    double e[CU], t[CU];
#pragma unroll
    for (int u = 0; u < CU; u++) {
      e[u] = ntl(dens + (long)(k + u) * ldg + i);
      t[u] = ntl(h + (long)(k + u) * ldg + i);
    }
#pragma unroll
    for (int u = 0; u < CU; u++) {
      const double gh = grav * t[u];
      const double ghe = gh * e[u];
      const double d = ghe * cdk[k + u];
      const double rn = r + d;
      nts(p + (long)(k + u) * ldo + i, 0.5 * (r + rn));
      r = rn;
    }
  }
  for (; k < km; k++) {
    const double gh = grav * ntl(h + (long)k * ldg + i);
    const double ghe = gh * ntl(dens + (long)k * ldg + i);
    const double d = ghe * cdk[k];
    const double rn = r + d;
    nts(p + (long)k * ldo + i, 0.5 * (r + rn));
    r = rn;
  }
}

// f(k) = h(k) * cdk(k); ub = sum_k u(k) f(k) / sum_k f(k), vb likewise  (:246-251)
__global__ __launch_bounds__(256) void k_col_uvbarot(const double* __restrict__ u, const double* __restrict__ v,
                                                     const double* __restrict__ h, long ldg, int km,
                                                     const double* __restrict__ cdk, double* __restrict__ ub,
                                                     double* __restrict__ vb, int npts) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= npts) return;
  double dep = 0.0, su = 0.0, sv = 0.0;
  int k = 0;
  for (; k + CU <= km; k += CU) {
    double a[CU], b[CU], t[CU];
#pragma unroll
    for (int q = 0; q < CU; q++) {
      a[q] = ntl(u + (long)(k + q) * ldg + i);
      b[q] = ntl(v + (long)(k + q) * ldg + i);
      t[q] = ntl(h + (long)(k + q) * ldg + i);
    }
#pragma unroll
    for (int q = 0; q < CU; q++) {
      const double f = t[q] * cdk[k + q];
      const double fu = a[q] * f, fv = b[q] * f;
      dep = dep + f;
      su = su + fu;
      sv = sv + fv;
    }
  }
  for (; k < km; k++) {
    const double f = ntl(h + (long)k * ldg + i) * cdk[k];
    const double fu = ntl(u + (long)k * ldg + i) * f, fv = ntl(v + (long)k * ldg + i) * f;
    dep = dep + f;
    su = su + fu;
    sv = sv + fv;
  }
  if (ub) nts(ub + i, su / dep);
  if (vb) nts(vb + i, sv / dep);
}

// QUICK curvature term of layer l from its neighbours (HBEBaroc_hspm_vfvm_mod.f90:274-277):
// ((T[l-1] - T[l]) rfdk[l] - (T[l] - T[l+1]) rfdk[l+1]) (rfdk[l] + rfdk[l+1])
__device__ __forceinline__ double quick_diff(double ta, double tb, double tc, double ra, double rb) {
#pragma clang fp contract(off)
  const double x = (ta - tb) * ra;
  const double y = (tb - tc) * rb;
  return (x - y) * (ra + rb);
}

// Vertical advective flux at the km+1 half levels (half level h on top of layer h):
//   h = 0      0, or omg[0] T[0] with a linear free surface                (:298, 307-312)
//   h = 1, km-1  omg[h] (m1 T[h-1] + m2 T[h])                                (:300-301)
//   2 <= h <= km-2  omg[h] ((m1 T[h-1] + m2 T[h]) - dfac (D[h-1] + D[h]))
//                   + |omg[h]| dfac (D[h-1] - D[h])                          (:286-291)
//   h = km     0
// with m1 = cdk[h]/(cdk[h-1]+cdk[h]), m2 = cdk[h-1]/(cdk[h-1]+cdk[h]), dfac = 0.125 cdk[h] cdk[h-1]
// (:264-266) and D the curvature term above.  The block computes m1, m2 and dfac of every
// half level once into LDS (3 (km+1) doubles of dynamic LDS); each thread keeps a
// three-layer window of the tracer and the previous curvature term in registers, so each
// grid value is loaded once.
__global__ __launch_bounds__(256) void k_col_quick(const double* __restrict__ omg, long ldo,
                                                   const double* __restrict__ trc, long ldg, int km,
                                                   const double* __restrict__ cdk, const double* __restrict__ rfdk,
                                                   int linfree, double* __restrict__ flx, long ldf, int npts) {
#pragma clang fp contract(off)
  extern __shared__ double wts[];
  double* sm1 = wts;
  double* sm2 = wts + (km + 1);
  double* sdf = wts + 2 * (km + 1);
  for (int h = 1 + (int)threadIdx.x; h < km; h += 256) {
    const double s = cdk[h - 1] + cdk[h];
    sm1[h] = cdk[h] / s;
    sm2[h] = cdk[h - 1] / s;
    sdf[h] = (0.125 * cdk[h]) * cdk[h - 1];
  }
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= npts) return;
  const double* T = trc + i;
  const double* W = omg + i;
  double* F = flx + i;
  double b = ntl(T), c = ntl(T + ldg);  // T[h-1], T[h] entering h = 1
  nts(F, linfree ? ntl(W) * b : 0.0);
  {
    const double x = sm1[1] * b, y = sm2[1] * c;
    nts(F + ldf, ntl(W + ldo) * (x + y));
  }
  if (km >= 4) {
    double d = ntl(T + 2 * ldg);  // T[2]
    double dprev = quick_diff(b, c, d, rfdk[1], rfdk[2]);
    b = c;
    c = d;  // entering h = 2: b = T[1], c = T[2]
    // interior half levels h = 2 .. km-2: groups of CU with their loads issued together
    auto step = [&](int h, double tn, double w) {
      const double dcur = quick_diff(b, c, tn, rfdk[h], rfdk[h + 1]);
      const double df = sdf[h];
      const double x = sm1[h] * b, y = sm2[h] * c;
      const double q = dprev + dcur;
      const double e = df * q;
      const double g = w * ((x + y) - e);
      const double aw = fabs(w) * df;
      const double j = aw * (dprev - dcur);
      nts(F + (long)h * ldf, g + j);
      dprev = dcur;
      b = c;
      c = tn;
    };
    int h = 2;
    for (; h + CU - 1 <= km - 2; h += CU) {
      double tn[CU], w[CU];
#pragma unroll
      for (int u = 0; u < CU; u++) {
        tn[u] = ntl(T + (long)(h + u + 1) * ldg);
        w[u] = ntl(W + (long)(h + u) * ldo);
      }
#pragma unroll
      for (int u = 0; u < CU; u++) step(h + u, tn[u], w[u]);
    }
    for (; h <= km - 2; h++) step(h, ntl(T + (long)(h + 1) * ldg), ntl(W + (long)h * ldo));
  } else if (km == 3) {
    b = c;
    c = ntl(T + 2 * ldg);
  }
  if (km >= 3) {  // h = km - 1 with b = T[km-2], c = T[km-1]
    const int h = km - 1;
    const double x = sm1[h] * b, y = sm2[h] * c;
    nts(F + (long)h * ldf, ntl(W + (long)h * ldo) * (x + y));
  }
  nts(F + (long)km * ldf, 0.0);
}

// Implicit vertical diffusion of a tracer increment, one column per thread
// (calc_VDiffTRCEq, DOGCM_Phys_hspm_vfvm_mod.f90:800-858): the tridiagonal system
//   Al x[k-1] + Ad x[k] + Au x[k+1] = dt rhs[k] / e3[k],   qa[k] = q0[k] + x[k]
// with coef = dt/(e3 cdk), the half-level coefficients r (:809-823)
//   r[0] = kv[0]/(fdk[0] e3[0]),  r[km] = kv[km-1]/(fdk[km] e3[km-1]),
//   r[h] = (m1 kv[h-1] + m2 kv[h]) / (fdk[h] (m1 e3[h-1] + m2 e3[h])),  0 < h < km,
// rows Al = -alpha coef r[k], Ad = 1 + alpha coef (r[k] + r[k+1]), Au = -alpha coef r[k+1]
// (:825-830), and Dirichlet ('D', the boundary r doubled) or Neumann ('N', dropped) rows at
// the surface and the bottom (:836-852).  The solve is LAPACK DGTSV (Gaussian elimination
// with partial pivoting by row interchange, then back substitution; LAPACK 3.x
// SRC/dgtsv.f, the NRHS = 1 path), in its order of operations.  Each finished row of the
// forward sweep (d, du and its right-hand side) waits for the backward sweep in one of three
// places: the last RR rows in registers (slots fixed by unrolled loops), the LR rows before
// them in LDS (row-major over the block's threads), the rest in `work` (d, du; pitch npts;
// 3 km values per point reserved) and qa (the right-hand side).  The grid has one thread per
// point and few waves per SIMD, so the registers cost no occupancy.  The second superdiagonal
// is nonzero only after a row interchange, so up to 64 layers a bit mask in a register
// records the interchanges and only those rows store it.  The backward sweep writes
// qa = q0 + x.  Both sweeps keep the loads of CU rows in flight together.  A singular pivot
// (impossible with kv >= 0 and positive thicknesses) is not reported.
// MOM selects calc_VDiffMOMEq (:865-975): right-hand side dt rhs[k] (no division by e3) and,
// for a Dirichlet bottom, Ad = 1 + alpha coef (bfric + r[km-1]) with the bottom friction
// coefficient bfric of the grid point (:963).
struct VdiffRow {
  double al, ad, au, b;
};

template <bool MOM, bool MASK, int RR, int LR>
__global__ __launch_bounds__(256) void k_col_vdiff(const double* __restrict__ q0, const double* __restrict__ rhs,
                                                   const double* __restrict__ kv, const double* __restrict__ e3,
                                                   long ldg, int km, const double* __restrict__ cdk,
                                                   const double* __restrict__ fdk, double dt, double alpha,
                                                   int sfc_dirichlet, int btm_dirichlet,
                                                   const double* __restrict__ bfric, double* __restrict__ qa,
                                                   long ldo, double* __restrict__ work, int npts) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= npts) return;
  const double bf = (MOM && btm_dirichlet) ? ntl(bfric + i) : 0.0;
  const long nw = npts;
  double* wd = work + i;                   // d of finished rows
  double* wu = work + (long)km * nw + i;   // du
  double* w2 = work + 2l * km * nw + i;    // second superdiagonal
  // MASK (km <= 64): bit j set where step j interchanged rows; the second superdiagonal is
  // nonzero only there, so only those rows store it and the backward sweep reads only those
  unsigned long long piv = 0ull;
  // half-level coefficient r[h] from the layers above (h-1) and below (h)
  auto rhalf = [&](int h, double kva, double e3a, double kvb, double e3b) {
    const double s = cdk[h - 1] + cdk[h];
    const double m1 = cdk[h] / s, m2 = cdk[h - 1] / s;
    const double num = m1 * kva + m2 * kvb;
    const double den = fdk[h] * (m1 * e3a + m2 * e3b);
    return num / den;
  };
  // layer k's values and its lower half-level coefficient: carried one layer ahead
  double kv_c = ntl(kv + i), e3_c = ntl(e3 + i);
  double r_up = kv_c / (fdk[0] * e3_c);  // r[0]
  auto row = [&](int k, double r_top, double r_bot, double e3k, double rk) {
    VdiffRow o;
    const double coef = dt / (e3k * cdk[k]);
    const double ac = alpha * coef;
    const double bt = dt * rk;
    o.b = MOM ? bt : bt / e3k;
    if (k == 0) {
      o.al = 0.0;
      o.ad = sfc_dirichlet ? 1.0 + ac * (r_bot + 2.0 * r_top) : 1.0 + ac * r_bot;
      o.au = -(ac * r_bot);
    } else if (k == km - 1) {
      o.al = -(ac * r_top);
      if (MOM)
        o.ad = btm_dirichlet ? 1.0 + ac * (bf + r_top) : 1.0 + ac * r_top;
      else
        o.ad = btm_dirichlet ? 1.0 + ac * (2.0 * r_bot + r_top) : 1.0 + ac * r_top;
      o.au = 0.0;
    } else {
      o.al = -(ac * r_top);
      o.ad = 1.0 + ac * (r_top + r_bot);
      o.au = -(ac * r_bot);
    }
    return o;
  };
  // builds row k from rhs[k] and layer k+1 (kv_n, e3_n: unused for the last row)
  auto build_row = [&](int k, double rk, double kv_n, double e3_n) {
    double r_bot;
    if (k + 1 < km) {
      r_bot = rhalf(k + 1, kv_c, e3_c, kv_n, e3_n);
      const VdiffRow o = row(k, r_up, r_bot, e3_c, rk);
      kv_c = kv_n;
      e3_c = e3_n;
      r_up = r_bot;
      return o;
    }
    r_bot = kv_c / (fdk[km] * e3_c);
    return row(k, r_up, r_bot, e3_c, rk);
  };
  auto next_row = [&](int k) {
    const int kn = k + 1 < km ? k + 1 : km - 1;
    return build_row(k, ntl(rhs + (long)k * ldg + i), ntl(kv + (long)kn * ldg + i), ntl(e3 + (long)kn * ldg + i));
  };

  // ---- forward elimination (DGTSV): row j current (d, du, b), row j+1 pending (dl = its Al)
  VdiffRow r0 = next_row(0);
  double d = r0.ad, du = r0.au, b = r0.b;
  VdiffRow r1 = next_row(1);
  double dl = r1.al, dn = r1.ad, dun = r1.au, bn = r1.b;
  // RR > 0: the finished rows of the last RR steps before the final one (j = J1 .. km-3) stay
  // in registers (d, du, b); the LR rows before them (j = J0 .. J1-1) in LDS, row-major over
  // the block's threads; only the rows above J0 make the workspace round trip.  The rare
  // second superdiagonal of an interchange goes to the workspace in every tier.
  const int J1 = km - 2 - RR;
  const int J0 = J1 - LR > 0 ? J1 - LR : 0;
  double rd[RR > 0 ? RR : 1], ru[RR > 0 ? RR : 1], rb[RR > 0 ? RR : 1];
  __shared__ double sd[LR > 0 ? LR : 1][256], su[LR > 0 ? LR : 1][256], sb[LR > 0 ? LR : 1][256];
  const int tx = threadIdx.x;
  // step j with row j+2 = r2; the finished row j goes to register slot reg (reg >= 0), to
  // LDS (J0 <= j < J1) or to the workspace
  auto eliminate = [&](int j, const VdiffRow& r2, int reg) {
    double du2;
    const bool keep = fabs(d) >= fabs(dl);  // no interchange
    if (keep) {
      const double fact = dl / d;
      dn = dn - fact * du;
      bn = bn - fact * b;
      du2 = 0.0;
    } else {
      const double fact = d / dl;
      d = dl;
      const double temp = dn;
      dn = du - fact * temp;
      du2 = dun;
      dun = -fact * du2;
      du = temp;
      const double tb = b;
      b = bn;
      bn = tb - fact * bn;
    }
    if (RR > 0 && reg >= 0) {
      rd[reg] = d;
      ru[reg] = du;
      rb[reg] = b;
    } else if (LR > 0 && j >= J0) {
      sd[j - J0][tx] = d;
      su[j - J0][tx] = du;
      sb[j - J0][tx] = b;
    } else {
      wd[(long)j * nw] = d;
      wu[(long)j * nw] = du;
      qa[(long)j * ldo + i] = b;
    }
    if (!MASK || !keep) w2[(long)j * nw] = du2;
    if (MASK && !keep) piv |= 1ull << j;
    d = dn, du = dun, b = bn;
    dl = r2.al, dn = r2.ad, dun = r2.au, bn = r2.b;
  };
  int j = 0;
  for (; j + CU - 1 <= J1 - 1; j += CU) {  // rows j+2 .. j+CU+1 loaded together
    double pr[CU], pk[CU], pe[CU];
#pragma unroll
    for (int u = 0; u < CU; u++) {
      const int k = j + u + 2, kn = k + 1 < km ? k + 1 : km - 1;
      pr[u] = ntl(rhs + (long)k * ldg + i);
      pk[u] = ntl(kv + (long)kn * ldg + i);
      pe[u] = ntl(e3 + (long)kn * ldg + i);
    }
#pragma unroll
    for (int u = 0; u < CU; u++) eliminate(j + u, build_row(j + u + 2, pr[u], pk[u], pe[u]), -1);
  }
  for (; j < J1; j++) eliminate(j, next_row(j + 2), -1);
  if constexpr (RR > 0) {  // the register rows, unrolled so the slots are registers
#pragma unroll
    for (int u0 = 0; u0 < RR; u0 += CU) {
      double pr[CU], pk[CU], pe[CU];
#pragma unroll
      for (int v = 0; v < CU; v++) {
        if (u0 + v < RR) {
          const int k = J1 + u0 + v + 2, kn = k + 1 < km ? k + 1 : km - 1;
          pr[v] = ntl(rhs + (long)k * ldg + i);
          pk[v] = ntl(kv + (long)kn * ldg + i);
          pe[v] = ntl(e3 + (long)kn * ldg + i);
        }
      }
#pragma unroll
      for (int v = 0; v < CU; v++)
        if (u0 + v < RR) eliminate(J1 + u0 + v, build_row(J1 + u0 + v + 2, pr[v], pk[v], pe[v]), u0 + v);
    }
  }
  {  // last elimination step, rows km-2 and km-1
    if (fabs(d) >= fabs(dl)) {
      const double fact = dl / d;
      dn = dn - fact * du;
      bn = bn - fact * b;
    } else {
      const double fact = d / dl;
      d = dl;
      const double temp = dn;
      dn = du - fact * temp;
      du = temp;
      const double tb = b;
      b = bn;
      bn = tb - fact * bn;
    }
  }
  // ---- back substitution
  const double x1 = bn / dn;  // row km-1
  const double x0 = (b - du * x1) / d;  // row km-2
  nts(qa + (long)(km - 1) * ldo + i, ntl(q0 + (long)(km - 1) * ldg + i) + x1);
  nts(qa + (long)(km - 2) * ldo + i, ntl(q0 + (long)(km - 2) * ldg + i) + x0);
  double xa = x0, xb = x1;  // x[j+1], x[j+2]
  auto back = [&](int jj, double dj, double uj, double vj, double bj, double qj) {
    const double t = bj - uj * xa;
    const double x = (t - vj * xb) / dj;
    nts(qa + (long)jj * ldo + i, qj + x);
    xb = xa;
    xa = x;
  };
  if constexpr (RR > 0) {  // the register rows, in reverse
#pragma unroll
    for (int u0 = RR - 1; u0 >= 0; u0 -= CU) {
      double pv[CU], pq[CU];
#pragma unroll
      for (int v = 0; v < CU; v++) {
        if (u0 - v >= 0) {
          const long r = J1 + u0 - v;
          pv[v] = (!MASK || ((piv >> r) & 1ull)) ? w2[r * nw] : 0.0;
          pq[v] = ntl(q0 + r * ldg + i);
        }
      }
#pragma unroll
      for (int v = 0; v < CU; v++)
        if (u0 - v >= 0) back(J1 + u0 - v, rd[u0 - v], ru[u0 - v], pv[v], rb[u0 - v], pq[v]);
    }
  }
  if constexpr (LR > 0) {  // the LDS rows, in reverse
    int jl = J1 - 1;
    for (; jl - (CU - 1) >= J0; jl -= CU) {
      double pv[CU], pq[CU];
#pragma unroll
      for (int u = 0; u < CU; u++) {
        const long r = jl - u;
        pv[u] = (!MASK || ((piv >> r) & 1ull)) ? w2[r * nw] : 0.0;
        pq[u] = ntl(q0 + r * ldg + i);
      }
#pragma unroll
      for (int u = 0; u < CU; u++) {
        const int r = jl - u;
        back(r, sd[r - J0][tx], su[r - J0][tx], pv[u], sb[r - J0][tx], pq[u]);
      }
    }
    for (; jl >= J0; jl--)
      back(jl, sd[jl - J0][tx], su[jl - J0][tx], (!MASK || ((piv >> jl) & 1ull)) ? w2[(long)jl * nw] : 0.0,
           sb[jl - J0][tx], ntl(q0 + (long)jl * ldg + i));
  }
  int jb = (LR > 0 ? J0 : J1) - 1;
  for (; jb - (CU - 1) >= 0; jb -= CU) {  // rows jb .. jb-CU+1 loaded together
    double pd[CU], pu[CU], pv[CU], pb[CU], pq[CU];
#pragma unroll
    for (int u = 0; u < CU; u++) {
      const long r = jb - u;
      pd[u] = wd[r * nw];
      pu[u] = wu[r * nw];
      pv[u] = (!MASK || ((piv >> r) & 1ull)) ? w2[r * nw] : 0.0;
      pb[u] = qa[r * ldo + i];
      pq[u] = ntl(q0 + r * ldg + i);
    }
#pragma unroll
    for (int u = 0; u < CU; u++) back(jb - u, pd[u], pu[u], pv[u], pb[u], pq[u]);
  }
  for (; jb >= 0; jb--)
    back(jb, wd[(long)jb * nw], wu[(long)jb * nw], (!MASK || ((piv >> jb) & 1ull)) ? w2[(long)jb * nw] : 0.0,
         qa[(long)jb * ldo + i], ntl(q0 + (long)jb * ldg + i));
}

int col_grid(int npts) { return (npts + 255) / 256; }

}  // namespace

int launch_col_omega(const double* div, const double* h, long ldg, int km, const double* cdk, double* omg, long ldo,
                     int npts, hipStream_t s) {
  hipLaunchKernelGGL(k_col_omega, dim3(col_grid(npts)), dim3(256), 0, s, div, h, ldg, km, cdk, omg, ldo, npts);
  DSHT_HIP(hipGetLastError());
  return 0;
}

int launch_col_hydpres(const double* dens, const double* h, long ldg, int km, const double* cdk, double grav,
                       double* p, long ldo, int npts, hipStream_t s) {
  hipLaunchKernelGGL(k_col_hydpres, dim3(col_grid(npts)), dim3(256), 0, s, dens, h, ldg, km, cdk, grav, p, ldo,
                     npts);
  DSHT_HIP(hipGetLastError());
  return 0;
}

int launch_col_uvbarot(const double* u, const double* v, const double* h, long ldg, int km, const double* cdk,
                       double* ub, double* vb, int npts, hipStream_t s) {
  hipLaunchKernelGGL(k_col_uvbarot, dim3(col_grid(npts)), dim3(256), 0, s, u, v, h, ldg, km, cdk, ub, vb, npts);
  DSHT_HIP(hipGetLastError());
  return 0;
}

int launch_col_quick(const double* omg, long ldo, const double* trc, long ldg, int km, const double* cdk,
                     const double* rfdk, int linfree, double* flx, long ldf, int npts, hipStream_t s) {
  const size_t lds = 3 * (size_t)(km + 1) * sizeof(double);
  if (lds > 64 * 1024) return fail(DSHT_EARG, "the QUICK flux supports at most 2729 layers");
  hipLaunchKernelGGL(k_col_quick, dim3(col_grid(npts)), dim3(256), lds, s, omg, ldo, trc, ldg, km, cdk, rfdk,
                     linfree, flx, ldf, npts);
  DSHT_HIP(hipGetLastError());
  return 0;
}

int launch_col_vdiff(const double* q0, const double* rhs, const double* kv, const double* e3, long ldg, int km,
                     const double* cdk, const double* fdk, double dt, double alpha, int sfc_dirichlet,
                     int btm_dirichlet, const double* bfric, double* qa, long ldo, double* work, int npts,
                     hipStream_t s) {
  // the last finished rows stay in registers (32 with the interchange mask, 24 without) and
  // the VDIFF_LR before them in LDS when the column is long enough for the register rows
  constexpr int RM = 32, RN = 24;
  auto k = bfric ? (km <= 64 ? (km >= RM + 2 ? k_col_vdiff<true, true, RM, VDIFF_LR> : k_col_vdiff<true, true, 0, 0>)
                             : k_col_vdiff<true, false, RN, VDIFF_LR>)
                 : (km <= 64 ? (km >= RM + 2 ? k_col_vdiff<false, true, RM, VDIFF_LR> : k_col_vdiff<false, true, 0, 0>)
                             : k_col_vdiff<false, false, RN, VDIFF_LR>);
  hipLaunchKernelGGL(k, dim3(col_grid(npts)), dim3(256), 0, s, q0, rhs, kv, e3, ldg, km, cdk, fdk, dt, alpha,
                     sfc_dirichlet, btm_dirichlet, bfric, qa, ldo, work, npts);
  DSHT_HIP(hipGetLastError());
  return 0;
}

}  // namespace dsht
