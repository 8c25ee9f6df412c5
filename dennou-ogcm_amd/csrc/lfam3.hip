// lfam3.hip -- a device-resident LF-AM3 dynamics step of dogcm for many independent runs of
// one truncation and layer count (DOGCM_TInt_LFAM3_advance_Dyn,
// model/dogcm/src/tint/DOGCM_TInt_LFAM3_mod.f90:281-880).
//
// Configuration: hydrostatic Boussinesq dynamics on the hspm_vfvm grid with a rigid lid
// (one barotropic iteration, :147-170, 685-774), a flat bottom (xyz_H = D, xyz_Z = D z_CK,
// DOGCM_Admin_Grid_mod.f90:376-388), the linear equation of state (EOS_Linear_mod.f90) and
// no physics forcing.  Per layer a step makes the reference's 42 forward and 28 inverse
// transform units (VorDivDiag x3, MOMBarocRHS x2, HTRCRHS x2 over two tracers, ImplUV,
// ImplTRC) and per run 9 forward and 9 inverse 2-D units of the barotropic mode; every
// transform runs over all layers of all active runs in one library call.  The grid-space
// work between the transforms is fused into the kernels below; the column work (omega,
// hydrostatic pressure, barotropic velocity, QUICK flux, the implicit vertical solves) uses
// the library's column kernels.
//
// Layout (FP64 device memory): run r, layer k of a layered field is grid column r*K + k;
// tracers are run-major with temperature then salinity: column (2r + t)*K + k; 2-D fields
// one column per run; half-level fields (K+1) levels per run (per run and tracer for the
// vertical fluxes).  Spectra of internal transforms use the fast "LC" layout.
#include <cmath>
#include <cstring>
#include <vector>

#include "dsht_internal.hpp"

namespace dsht {
namespace {

constexpr int LF_THREADS = 256;

inline dim3 lf_grid(long n) {
  long b = (n + LF_THREADS - 1) / LF_THREADS;
  if (b > 1l << 20) b = 1l << 20;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

#define LF_LOOP(n) for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < (n); g += (long)gridDim.x * blockDim.x)
// 2-D walk over (column c, point p) of ncols columns of npts points, g = c npts + p: the column
// is uniform over a block (grid y), so its layer and run come from scalar arithmetic, and
// the latitude row of p is one 32-bit division
#define LF_CP(ncols, npts)                                                                                \
  for (long c = blockIdx.y; c < (ncols); c += gridDim.y)                                                 \
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < (npts); p += (long)gridDim.x * blockDim.x)
#define LF_ROW(p, I) ((int)((unsigned)(p) / (unsigned)(I)))

inline dim3 lf_grid2(long ncols, long npts) {
  long bx = (npts + LF_THREADS - 1) / LF_THREADS;
  if (bx > 4096) bx = 4096;
  long by = ncols < 65535 ? ncols : 65535;
  return dim3((unsigned)(bx < 1 ? 1 : bx), (unsigned)(by < 1 ? 1 : by));
}

// x*cos(phi_j), y*cos(phi_j) (the calc_UVCosLat2VorDiv inputs, HBEDiagnose_hspm_vfvm_mod.f90:278-281)
__global__ void k_lf_uvcos(const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ cs,
                           long npts, int I, long ncols, double* __restrict__ ucos, double* __restrict__ vcos) {
  LF_CP(ncols, npts) {
    const long g = c * npts + p;
    const double w = cs[LF_ROW(p, I)];
    ucos[g] = u[g] * w;
    vcos[g] = v[g] * w;
  }
}

// GeoPot = Grav * D * z_CK(k) (DOGCM_TInt_LFAM3_mod.f90:460)
__global__ void k_lf_geo(double* __restrict__ geo, long npts, int K, const double* __restrict__ zck, double depth,
                         double grav, long ncols) {
  LF_CP(ncols, npts) {
    const int k = (int)(c % K);
    geo[c * npts + p] = grav * (depth * zck[k]);
  }
}

__global__ void k_lf_fill(double* __restrict__ x, double v, long n) {
  LF_LOOP(n) x[g] = v;
}

// EOS_Linear_Eval(theta, S, p = -RefDens GeoPot) for every layer of every run
struct EosArgs {
  double ref_dens, ptemp_dep, press_dep, salt_dep, ref_temp, ref_salt, depth, grav;
};
__global__ void k_lf_eos(const double* __restrict__ ts, long npts, int K, const double* __restrict__ zck, EosArgs e,
                         long ncols, double* __restrict__ dens) {
  LF_CP(ncols, npts) {
    const long g = c * npts + p;
    const int k = (int)(c % K);
    const long r = c / K;
    const double t = ts[((2 * r) * K + k) * npts + p];
    const double s = ts[((2 * r + 1) * K + k) * npts + p];
    const double geo = e.grav * (e.depth * zck[k]);
    const double pr = -e.ref_dens * geo;
    dens[g] = e.ref_dens * (e.ptemp_dep * (t - e.ref_temp) + e.press_dep * pr + e.salt_dep * (s - e.ref_salt));
  }
}

// The grid-space part of HBEBaroc_MOMRHS_VorDivForm (HBEBaroc_hspm_vfvm_mod.f90:366-420):
// the vertical advection of momentum at the half levels, the absolute-vorticity fluxes A
// and B (times cos phi) and the Bernoulli function C.  Writes -A (so that one
// calc_UVCosLat2VorDiv call with (U cos, V cos) = (B, -A) gives -H(A, B) and H(B, -A)), B
// and C.  The Coriolis terms are f (cn x_N + cb x_B) (:474-481, 565-570).
struct MomArgs {
  const double *u, *v, *omg, *vor, *h, *hyd, *sp, *dens, *glon, *glat, *un, *ub, *vn, *vb;
  const double *two_omega, *mu, *cs, *rfdk;
  double ref_dens, radius, cn, cb;
  long npts;
  int I, K;
  double *am, *bg, *cg;
};
__global__ void k_lf_mom(MomArgs a, long ncols) {
  LF_CP(ncols, a.npts) {
    const long g = c * a.npts + p;
    const int k = (int)(c % a.K);
    const long r = c / a.K;
    const int j = LF_ROW(p, a.I);
    const long np = a.npts;
    const double u = a.u[g], v = a.v[g];
    const double* om = a.omg + r * (a.K + 1) * np + p;  // half level h at om[h*np]
    double vau_a = 0.0, vav_a = 0.0, vau_b = 0.0, vav_b = 0.0;
    if (k > 0) {  // half level k, between layers k-1 and k
      vau_a = om[(long)k * np] * (a.u[g - np] - u) * a.rfdk[k];
      vav_a = om[(long)k * np] * (a.v[g - np] - v) * a.rfdk[k];
    }
    if (k < a.K - 1) {  // half level k+1, between layers k and k+1
      vau_b = om[(long)(k + 1) * np] * (u - a.u[g + np]) * a.rfdk[k + 1];
      vav_b = om[(long)(k + 1) * np] * (v - a.v[g + np]) * a.rfdk[k + 1];
    }
    const double f = a.two_omega[r] * a.mu[j];
    const double coru = f * (a.cn * a.un[g] + a.cb * a.ub[g]);
    const double corv = f * (a.cn * a.vn[g] + a.cb * a.vb[g]);
    const double h = a.h[g], vor = a.vor[g], dens = a.dens[g], cs = a.cs[j];
    const double rr = a.ref_dens * a.radius;
    double A = vor * u + coru + 0.5 * (vav_b + vav_a) / h;
    A = cs * (A + dens * a.glat[g] / rr);
    double B = vor * v + corv - 0.5 * (vau_b + vau_a) / h;
    B = cs * (B + dens * a.glon[g] / rr);
    const double pres = a.hyd[g] + a.sp[r * np + p];
    a.am[g] = -A;
    a.bg[g] = B;
    a.cg[g] = (0.5 * (u * u + v * v) + pres / a.ref_dens) / (a.radius * a.radius);
  }
}

// w_out[l, c] += w_in[l, c] * s  (s = 1, or 1/(RefDens R^2) for the barotropic pressure term)
__global__ void k_lf_spec_axpy(double* __restrict__ wout, long slo, const double* __restrict__ win, long sli, int L,
                               int ncols, double s, int divide) {
  LF_CP(L, ncols) {
    const double x = win[c * sli + p];
    wout[c * slo + p] += divide ? x / s : x;
  }
}

// predictor momentum (DOGCM_TInt_LFAM3_mod.f90:495-517): optional implicit Coriolis with
// CoriImplFac = 0.5 (2 dt) f, then U = 5/6 dt U_RHS + (2 U_N + U_B)/3
__global__ void k_lf_pred_uv(const double* __restrict__ urhs, const double* __restrict__ vrhs,
                             const double* __restrict__ un, const double* __restrict__ ub,
                             const double* __restrict__ vn, const double* __restrict__ vb,
                             const double* __restrict__ two_omega, const double* __restrict__ mu, long npts, int I,
                             int K, double half2dt, double c56dt, int implicit, long ncols, double* __restrict__ uo,
                             double* __restrict__ vo) {
  LF_CP(ncols, npts) {
    const long g = c * npts + p;
    const long r = c / K;
    const int j = LF_ROW(p, I);
    double ur = urhs[g], vr = vrhs[g];
    if (implicit) {
      const double cif = half2dt * (two_omega[r] * mu[j]);
      const double ut = ur;
      ur = (ut + cif * vr) / (1.0 + cif * cif);
      vr = (vr - cif * ut) / (1.0 + cif * cif);
    }
    uo[g] = c56dt * ur + (2.0 * un[g] + ub[g]) / 3.0;
    vo[g] = c56dt * vr + (2.0 * vn[g] + vb[g]) / 3.0;
  }
}

// x = x - xbar + xbar_ref for every layer (the barotropic part of x replaced, :520-528, 783-793)
__global__ void k_lf_barot_shift(double* __restrict__ u, double* __restrict__ v, const double* __restrict__ ubt,
                                 const double* __restrict__ vbt, const double* __restrict__ ubr,
                                 const double* __restrict__ vbr, long npts, int K, long ncols) {
  LF_CP(ncols, npts) {
    const long g = c * npts + p;
    const long q = (c / K) * npts + p;
    u[g] = u[g] - ubt[q] + ubr[q];
    v[g] = v[g] - vbt[q] + vbr[q];
  }
}

// The grid-space part of HBEBaroc_HTRCRHS (HBEBaroc_hspm_vfvm_mod.f90:162-172): the
// horizontal flux (U, V) H cos(phi) T and the vertical flux divergence
// -(F(k-1) - F(k)) / z_CDK(k), for both tracers of every run
__global__ void k_lf_htrc_grid(const double* __restrict__ ts, const double* __restrict__ u,
                               const double* __restrict__ v, const double* __restrict__ h,
                               const double* __restrict__ cs, const double* __restrict__ flx,
                               const double* __restrict__ rcdk, long npts, int I, int K, long ncols,
                               double* __restrict__ fu, double* __restrict__ fv, double* __restrict__ vf) {
  LF_CP(ncols, npts) {
    const long g = c * npts + p;
    const int k = (int)(c % K);
    const long rt = c / K;  // 2 r + t
    const long r = rt / 2;
    const long q = (r * K + k) * npts + p;
    const int j = LF_ROW(p, I);
    const double tmp = h[q] * cs[j] * ts[g];
    fu[g] = u[q] * tmp;
    fv[g] = v[q] * tmp;
    const double* fl = flx + rt * (K + 1) * npts + p;
    vf[g] = -(fl[(long)k * npts] - fl[(long)(k + 1) * npts]) * rcdk[k];
  }
}

// w = AdvFilter (-w_alpha + w_vert) (HBEBaroc_hspm_vfvm_mod.f90:163-171), in place over w_alpha
__global__ void k_lf_htrc_spec(double* __restrict__ wa, const double* __restrict__ wv, const double* __restrict__ filt,
                               int L, int ncols) {
  LF_CP(L, ncols) {
    const long g = c * ncols + p;
    wa[g] = filt[c] * (-wa[g] + wv[g]);
  }
}

// predictor tracers (:535-541): ((H (2 T_N + T_B))/3 + 5/6 dt HTRC_RHS) / H
__global__ void k_lf_pred_trc(const double* __restrict__ tn, const double* __restrict__ tb,
                              const double* __restrict__ ht, const double* __restrict__ h, long npts, int K,
                              double c56dt, long ncols, double* __restrict__ to) {
  LF_CP(ncols, npts) {
    const long g = c * npts + p;
    const long q = ((c / K) / 2 * K + c % K) * npts + p;
    const double hh = h[q] + 0.0;
    to[g] = (hh * (2.0 * tn[g] + tb[g]) / 3.0 + c56dt * ht[g]) / hh;
  }
}

// U_RHS = (U_A - U_N)/dt after the implicit solves (:615-619)
__global__ void k_lf_rhs_diff(const double* __restrict__ ua, const double* __restrict__ un,
                              const double* __restrict__ va, const double* __restrict__ vn, double dt, long n,
                              double* __restrict__ ur, double* __restrict__ vr) {
  LF_LOOP(n) {
    ur[g] = (ua[g] - un[g]) / dt;
    vr[g] = (va[g] - vn[g]) / dt;
  }
}

// barotropic forcing of the baroclinic tendencies (:626-630): UBarocForce - CoriVBarot,
// VBarocForce + CoriUBarot, CoriXBarot = f (cn x_N + cb x_B)
__global__ void k_lf_baroc_force(double* __restrict__ ubf, double* __restrict__ vbf, const double* __restrict__ ubn,
                                 const double* __restrict__ vbn, const double* __restrict__ ubb,
                                 const double* __restrict__ vbb, const double* __restrict__ two_omega,
                                 const double* __restrict__ mu, long npts, int I, double cn, double cb, long nruns) {
  LF_CP(nruns, npts) {
    const long g = c * npts + p;
    const double f = two_omega[c] * mu[LF_ROW(p, I)];
    const double coru = f * (cn * ubn[g] + cb * ubb[g]);
    const double corv = f * (cn * vbn[g] + cb * vbb[g]);
    ubf[g] = ubf[g] - corv;
    vbf[g] = vbf[g] + coru;
  }
}

// implicit Coriolis with CoriImplFac = alpha_ dt f and U_A = U_N + dt U_RHS (:632-646); also the
// barotropic velocity step of the sub-cycle (:712-721) when the inputs are 2-D
__global__ void k_lf_cori_step(const double* __restrict__ ur_in, const double* __restrict__ vr_in,
                               const double* __restrict__ u0, const double* __restrict__ v0,
                               const double* __restrict__ two_omega, const double* __restrict__ mu, long npts, int I,
                               int K, double adt, double dt, int form, long ncols, double* __restrict__ uo,
                               double* __restrict__ vo) {
  LF_CP(ncols, npts) {
    const long g = c * npts + p;
    const double cif = adt * (two_omega[c / K] * mu[LF_ROW(p, I)]);
    const double ur = ur_in[g], vr = vr_in[g];
    if (form == 0) {  // baroclinic: the tendency is replaced, then U_A = U_N + dt U_RHS
      const double u2 = (ur + cif * vr) / (1.0 + cif * cif);
      const double v2 = (vr - cif * ur) / (1.0 + cif * cif);
      uo[g] = u0[g] + dt * u2;
      vo[g] = v0[g] + dt * v2;
    } else {  // barotropic sub-step: U + dts (U_RHS + c V_RHS)/(1 + c^2)
      uo[g] = u0[g] + dt * (ur + cif * vr) / (1.0 + cif * cif);
      vo[g] = v0[g] + dt * (vr - cif * ur) / (1.0 + cif * cif);
    }
  }
}

// HBEBarot_MOMRHS grid part (HBEBarot_hspm_vfvm_mod.f90:197-201): -A and B with
// A = cos (f U - VBarocForce), B = cos (f V + UBarocForce); and SfcPres - SfcPres_N
__global__ void k_lf_barot_rhs_grid(const double* __restrict__ ub, const double* __restrict__ vb,
                                    const double* __restrict__ ubf, const double* __restrict__ vbf,
                                    const double* __restrict__ sp, const double* __restrict__ spn,
                                    const double* __restrict__ two_omega, const double* __restrict__ mu,
                                    const double* __restrict__ cs, long npts, int I, long nruns, double* __restrict__ am,
                                    double* __restrict__ bg, double* __restrict__ spm) {
  LF_CP(nruns, npts) {
    const long g = c * npts + p;
    const int j = LF_ROW(p, I);
    const double f = two_omega[c] * mu[j];
    const double coru = f * ub[g], corv = f * vb[g];
    am[g] = -(cs[j] * (coru - vbf[g]));
    bg[g] = cs[j] * (corv + ubf[g]);
    spm[g] = sp[g] - spn[g];
  }
}

// HBEBarot_Update_LinFreeSfc spectral part with vareps1 = 0 (:283-297): w_Tmp = -w_Div,
// w_DSfcPres = RefDens w_Coef w_Tmp / (PresTAvgCoefA dts), w_DDiv = n(n+1)/R^2 w_Coef w_Tmp
__global__ void k_lf_barot_proj(const double* __restrict__ wdiv, long sld, const double* __restrict__ coef,
                                const double* __restrict__ e, double ref_dens, double den, int L, int ncols,
                                double* __restrict__ wdsp, double* __restrict__ wddiv) {
  LF_CP(L, ncols) {
    const long g = c * ncols + p;
    const double wt = -wdiv[c * sld + p];
    wdsp[g] = ref_dens * coef[c] * wt / den;
    wddiv[g] = e[c] * coef[c] * wt;
  }
}

__global__ void k_lf_add2(const double* __restrict__ a, const double* __restrict__ da, const double* __restrict__ b,
                          const double* __restrict__ db, const double* __restrict__ c, const double* __restrict__ dc,
                          long n, double* __restrict__ ao, double* __restrict__ bo, double* __restrict__ co) {
  LF_LOOP(n) {
    ao[g] = a[g] + da[g];
    bo[g] = b[g] + db[g];
    co[g] = c[g] + dc[g];
  }
}

// full velocity and the time-centred fields of the tracer update (:783-795):
// U_A = U_A - UBarot + UBarot_A; U = (5 U_A + 8 U_N - U_B)/12; H likewise
__global__ void k_lf_final_uv(double* __restrict__ ua, double* __restrict__ va, const double* __restrict__ ubt,
                              const double* __restrict__ vbt, const double* __restrict__ uba,
                              const double* __restrict__ vba, const double* __restrict__ un,
                              const double* __restrict__ ub, const double* __restrict__ vn,
                              const double* __restrict__ vb, const double* __restrict__ h, long npts, int K, long ncols,
                              double* __restrict__ uo, double* __restrict__ vo, double* __restrict__ hm) {
  LF_CP(ncols, npts) {
    const long g = c * npts + p;
    const long q = (c / K) * npts + p;
    const double a = ua[g] - ubt[q] + uba[q];
    const double b = va[g] - vbt[q] + vba[q];
    ua[g] = a;
    va[g] = b;
    uo[g] = (5.0 * a + 8.0 * un[g] - ub[g]) / 12.0;
    vo[g] = (5.0 * b + 8.0 * vn[g] - vb[g]) / 12.0;
    const double hh = h[g];
    hm[g] = (5.0 * hh + 8.0 * hh - hh) / 12.0;
  }
}

// T_A = H_0 T_A / H_A (DOGCM_Phys_hspm_vfvm_mod.f90:641)
__global__ void k_lf_trc_final(double* __restrict__ ta, const double* __restrict__ h, long npts, int K, long ncols) {
  LF_CP(ncols, npts) {
    const long g = c * npts + p;
    const long q = ((c / K) / 2 * K + c % K) * npts + p;
    ta[g] = h[q] * ta[g] / h[q];
  }
}

// time levels: B <- N, N <- A
__global__ void k_lf_rotate(double* __restrict__ b, double* __restrict__ nn, const double* __restrict__ a, long n) {
  LF_LOOP(n) {
    b[g] = nn[g];
    nn[g] = a[g];
  }
}

}  // namespace
}  // namespace dsht

using namespace dsht;

struct dsht_lfam3 {
  dsht_plan* p = nullptr;
  dsht_lfam3_config cfg;
  int K = 0, R = 0, L = 0, I = 0, J = 0;
  long np = 0;
  std::vector<double> cdk, fdk, rcdk, rfdk, zck;
  std::vector<double*> allocs;
  // constants
  double *d_cdk = nullptr, *d_fdk = nullptr, *d_rcdk = nullptr, *d_rfdk = nullptr, *d_zck = nullptr;
  double *d_mu = nullptr, *d_cs = nullptr, *d_two_omega = nullptr;
  double *d_dmom = nullptr, *d_dtrc = nullptr, *d_advf = nullptr, *d_nn1 = nullptr, *d_hcoef = nullptr,
         *d_e = nullptr;
  double *H = nullptr, *KV = nullptr, *KD = nullptr, *BF = nullptr;
  // state: levels B (0), N (1), A (2)
  double *U[3] = {}, *V[3] = {}, *TS[3] = {}, *SP[3] = {};
  // work
  double *ucos = nullptr, *vcos = nullptr, *VD = nullptr, *OMG = nullptr, *Dens = nullptr, *Hyd = nullptr;
  double *Geo = nullptr, *GLon = nullptr, *GLat = nullptr, *Am = nullptr, *Bg = nullptr, *Cg = nullptr;
  double *URHS = nullptr, *VRHS = nullptr, *Upr = nullptr, *Vpr = nullptr, *Hm = nullptr;
  double *TSpr = nullptr, *HT = nullptr, *FLX = nullptr, *FU = nullptr, *FV = nullptr, *VF = nullptr;
  double *b2[20] = {};  // 2-D barotropic work, one grid per run each
  double *wVD = nullptr, *wGeo = nullptr, *wC = nullptr, *w1 = nullptr, *w2 = nullptr;
  double *wB = nullptr, *wC2 = nullptr, *wDSP = nullptr, *wDDiv = nullptr, *wZero = nullptr;
  double* vwork = nullptr;
  long lvwork = 0;
};

namespace {

enum B2 { UBN, VBN, UBB, VBB, UBT, VBT, UBF, VBF, UBR, VBR, UBAR, VBAR, UBA, VBA, DU, DV, DSP, SPM, AM2, BG2 };

int lf_alloc(dsht_lfam3* s, double** out, size_t n) {
  void* ptr = nullptr;
  if (hipMalloc(&ptr, sizeof(double) * (n ? n : 1)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(DSHT_ENOMEM, "LF-AM3 pool: device allocation of %zu doubles failed", n);
  }
  s->allocs.push_back((double*)ptr);
  *out = (double*)ptr;
  return 0;
}

int lf_upload(dsht_lfam3* s, double** out, const std::vector<double>& v) {
  if (int rc = lf_alloc(s, out, v.size())) return rc;
  DSHT_HIP(hipMemcpy(*out, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
  return 0;
}

void lf_free(dsht_lfam3* s) {
  (void)hipDeviceSynchronize();
  for (double* d : s->allocs) (void)hipFree(d);
  s->allocs.clear();
}

#define LF_LAUNCH(kernel, n, s, ...)                                                     \
  do {                                                                                   \
    hipLaunchKernelGGL(kernel, lf_grid(n), dim3(LF_THREADS), 0, s, __VA_ARGS__);         \
    DSHT_HIP(hipGetLastError());                                                         \
  } while (0)
// kernels walking (column, point) pairs with LF_CP
#define LF_LAUNCH2(kernel, ncols, npts, s, ...)                                          \
  do {                                                                                   \
    hipLaunchKernelGGL(kernel, lf_grid2(ncols, npts), dim3(LF_THREADS), 0, s, __VA_ARGS__); \
    DSHT_HIP(hipGetLastError());                                                         \
  } while (0)

// ---- the reference's diagnostics on every active run (column kernels)
int lf_uvbarot(dsht_lfam3* s, const double* u, const double* v, double* ub, double* vb, int nr, hipStream_t st) {
  const long np = s->np, K = s->K;
  for (int r = 0; r < nr; r++)
    if (int rc = launch_col_uvbarot(u + r * K * np, v + r * K * np, s->H + r * K * np, np, s->K, s->d_cdk,
                                    ub + r * np, vb + r * np, (int)np, st))
      return rc;
  return 0;
}

int lf_omega(dsht_lfam3* s, const double* div, int nr, hipStream_t st) {
  const long np = s->np, K = s->K;
  for (int r = 0; r < nr; r++)
    if (int rc = launch_col_omega(div + r * K * np, s->H + r * K * np, np, s->K, s->d_cdk, s->OMG + r * (K + 1) * np,
                                  np, (int)np, st))
      return rc;
  return 0;
}

int lf_hydpres(dsht_lfam3* s, int nr, hipStream_t st) {
  const long np = s->np, K = s->K;
  for (int r = 0; r < nr; r++)
    if (int rc = launch_col_hydpres(s->Dens + r * K * np, s->H + r * K * np, np, s->K, s->d_cdk, s->cfg.grav,
                                    s->Hyd + r * K * np, np, (int)np, st))
      return rc;
  return 0;
}

// HBEDiagnose_VorDiv (HBEDiagnose_hspm_vfvm_mod.f90:259-286) over every layer of the active
// runs: vorticity into VD[0, nc), divergence into VD[nc, 2 nc)
int lf_vordiv(dsht_lfam3* s, const double* u, const double* v, int nc, hipStream_t st) {
  dsht_plan* p = s->p;
  LF_LAUNCH2(k_lf_uvcos, nc, s->np, st, u, v, s->d_cs, s->np, s->I, (long)nc, s->ucos, s->vcos);
  if (int rc = dsht_uv2vordiv_d(p, s->ucos, s->vcos, s->np, s->wVD, s->wVD + nc, 2l * nc, 1, nc, nullptr, st))
    return rc;
  return dsht_inv_d(p, DSHT_VAL, s->wVD, 2l * nc, 1, s->VD, s->np, 2 * nc, nullptr, st);
}

// HBEBaroc_MOMRHS_VorDivForm (HBEBaroc_hspm_vfvm_mod.f90:321-443) -> URHS, VRHS
int lf_mom_rhs(dsht_lfam3* s, const double* u, const double* v, const double* cor_un, const double* cor_ub,
               const double* cor_vn, const double* cor_vb, double cn, double cb, const double* sp, int nc,
               hipStream_t st) {
  dsht_plan* p = s->p;
  const long n = (long)nc * s->np;
  // w_GeoPot = w_xy(GeoPot); xy_GradLat_w, xy_GradLon_w of it
  if (int rc = dsht_fwd_d(p, DSHT_VAL, s->Geo, s->np, s->wGeo, nc, 1, nc, nullptr, st)) return rc;
  if (int rc = dsht_inv_vgrad_d(p, s->wGeo, nc, 1, nullptr, s->GLon, s->GLat, s->np, nc, nullptr, st)) return rc;
  MomArgs a;
  a.u = u;
  a.v = v;
  a.omg = s->OMG;
  a.vor = s->VD;
  a.h = s->H;
  a.hyd = s->Hyd;
  a.sp = sp;
  a.dens = s->Dens;
  a.glon = s->GLon;
  a.glat = s->GLat;
  a.un = cor_un;
  a.ub = cor_ub;
  a.vn = cor_vn;
  a.vb = cor_vb;
  a.two_omega = s->d_two_omega;
  a.mu = s->d_mu;
  a.cs = s->d_cs;
  a.rfdk = s->d_rfdk;
  a.ref_dens = s->cfg.ref_dens;
  a.radius = p->R;
  a.cn = cn;
  a.cb = cb;
  a.npts = s->np;
  a.I = s->I;
  a.K = s->K;
  a.am = s->Am;
  a.bg = s->Bg;
  a.cg = s->Cg;
  LF_LAUNCH2(k_lf_mom, nc, s->np, st, a, (long)nc);
  // (w_Vor_RHS, w_Div_RHS) = (-H(A, B), H(B, -A)) = calc_UVCosLat2VorDiv(B, -A); then
  // w_Div_RHS - w_Lapla_w(w_xy(C)) = + n(n+1) w_xy(C)
  if (int rc = dsht_uv2vordiv_d(p, s->Bg, s->Am, s->np, s->wVD, s->wVD + nc, 2l * nc, 1, nc, nullptr, st)) return rc;
  if (int rc = dsht_fwd_d(p, DSHT_VAL, s->Cg, s->np, s->wC, nc, 1, nc, s->d_nn1, st)) return rc;
  LF_LAUNCH2(k_lf_spec_axpy, s->L, nc, st, s->wVD + nc, 2l * nc, s->wC, (long)nc, s->L, nc, 1.0, 0);
  return dsht_vordiv2uv_d(p, s->wVD, s->wVD + nc, 2l * nc, 1, s->URHS, s->VRHS, s->np, nc, st);
}

// HBEBaroc_HTRCRHS (HBEBaroc_hspm_vfvm_mod.f90:119-202) for both tracers of the active runs
int lf_htrc(dsht_lfam3* s, const double* ts, const double* u, const double* v, const double* h, int nr,
            hipStream_t st) {
  dsht_plan* p = s->p;
  const long np = s->np, K = s->K;
  const int nt = 2 * nr * s->K;
  for (int rt = 0; rt < 2 * nr; rt++) {
    const int r = rt / 2;
    if (int rc = launch_col_quick(s->OMG + r * (K + 1) * np, np, ts + rt * K * np, np, s->K, s->d_cdk, s->d_rfdk, 0,
                                  s->FLX + rt * (K + 1) * np, np, (int)np, st))
      return rc;
  }
  LF_LAUNCH2(k_lf_htrc_grid, nt, np, st, ts, u, v, h, s->d_cs, s->FLX, s->d_rcdk, np, s->I, s->K, (long)nt, s->FU, s->FV,
             s->VF);
  if (int rc = dsht_alpha_fwd_d(p, s->FU, s->FV, np, s->w1, nt, 1, nt, nullptr, st)) return rc;
  if (int rc = dsht_fwd_d(p, DSHT_VAL, s->VF, np, s->w2, nt, 1, nt, nullptr, st)) return rc;
  LF_LAUNCH2(k_lf_htrc_spec, s->L, nt, st, s->w1, s->w2, s->d_advf, s->L, nt);
  return dsht_inv_d(p, DSHT_VAL, s->w1, nt, 1, s->HT, np, nt, nullptr, st);
}

}  // namespace

extern "C" {

int dsht_lfam3_default_config(dsht_lfam3_config* c) {
  if (!c) return fail(DSHT_EARG, "dsht_lfam3_default_config: NULL");
  std::memset(c, 0, sizeof(*c));
  c->km = 20;
  c->max_runs = 1;
  c->sub_cycles = 1;
  c->numdiff_ord = 4;
  c->dt = 3600.0;
  c->depth = 5000.0;
  c->grav = 9.79764;
  c->ref_dens = 1027.0;
  c->alpha = 0.51;
  c->lambda = 0.5;
  c->hvisc = 3e5;
  c->hdiff = 0.0;
  c->numdiff = 1e12;
  c->vvisc = 1e-3;
  c->vdiff = 3e-5;
  c->btm_frict = 0.0;
  c->mom_sfc_bc = 'N';
  c->mom_btm_bc = 'D';
  c->trc_sfc_bc = 'N';
  c->trc_btm_bc = 'N';
  c->ref_temp = 283.0;
  c->ref_salt = 35.0;
  c->beta_t = 1.67e-4;
  c->beta_s = 0.78e-3;
  c->cp0 = 3986.0;
  c->cs0 = 1490.0;
  return 0;
}

int dsht_lfam3_create(dsht_plan* p, const dsht_lfam3_config* cfg, dsht_lfam3** out) {
  if (!out) return fail(DSHT_EARG, "dsht_lfam3_create: NULL output");
  *out = nullptr;
  if (!p || p->N < 0) return fail(DSHT_ESTATE, "plan is NULL or destroyed");
  if (!cfg) return fail(DSHT_EARG, "dsht_lfam3_create: NULL config");
  const dsht_lfam3_config& c = *cfg;
  if (c.km < 2 || c.max_runs < 1) return fail(DSHT_EARG, "LF-AM3 pool needs km >= 2 and max_runs >= 1");
  if (!(c.dt > 0) || !(c.depth > 0) || !(c.lambda > 0) || c.sub_cycles < 1)
    return fail(DSHT_EARG, "LF-AM3 pool needs dt, depth, lambda > 0 and sub_cycles >= 1");
  auto bc_ok = [](int b) { return b == 'D' || b == 'N'; };
  if (!bc_ok(c.mom_sfc_bc) || !bc_ok(c.mom_btm_bc) || !bc_ok(c.trc_sfc_bc) || !bc_ok(c.trc_btm_bc))
    return fail(DSHT_EARG, "boundary types must be 'D' or 'N'");
  if (c.km % 2) return fail(DSHT_EARG, "km = %d must be even (the Chebyshev-Lobatto vertical grid)", c.km);
  dsht_lfam3* s = new dsht_lfam3();
  s->p = p;
  s->cfg = c;
  s->K = c.km;
  s->R = c.max_runs;
  s->L = p->L;
  s->I = p->I;
  s->J = p->J;
  s->np = (long)p->I * p->J;
  const int K = s->K, R = s->R, L = s->L, J = s->J;
  const long np = s->np, NC = (long)R * K;
  // vertical grid: faces g_Sig (at_module_omp.f90:598-601 on [-1, 0]), centres, thicknesses
  // and centre spacings (DOGCM_GaussSpmVFvmGrid_mod.f90:392-410)
  std::vector<double> fk(K + 1), ck(K + 2);
  if (int rc = dsht_sig_grid(K, K, -1.0, 0.0, fk.data(), nullptr)) {
    delete s;
    return rc;
  }
  s->cdk.resize(K);
  s->zck.resize(K);
  s->rcdk.resize(K);
  s->fdk.resize(K + 1);
  s->rfdk.resize(K + 1);
  for (int k = 0; k < K; k++) {
    s->cdk[k] = fk[k] - fk[k + 1];
    ck[k + 1] = 0.5 * (fk[k] + fk[k + 1]);
  }
  ck[0] = fk[0] + 0.5 * s->cdk[0];
  ck[K + 1] = fk[K] - 0.5 * s->cdk[K - 1];
  for (int k = 0; k <= K; k++) {
    s->fdk[k] = ck[k] - ck[k + 1];
    s->rfdk[k] = 1.0 / s->fdk[k];
  }
  for (int k = 0; k < K; k++) {
    s->zck[k] = ck[k + 1];
    s->rcdk[k] = 1.0 / s->cdk[k];
  }
  std::vector<double> cs(J), nn1(L), e(L), hcoef(L);
  for (int j = 0; j < J; j++) cs[j] = std::sqrt((1.0 - p->mu[j]) * (1.0 + p->mu[j]));
  for (int n = 0; n <= p->N; n++)
    for (int m = -n; m <= n; m++) {
      const int l = n * (n + 1) + m;
      nn1[l] = (double)n * (n + 1.0);
      e[l] = (double)n * (n + 1.0) / (p->R * p->R);
      hcoef[l] = l == 0 ? 0.0 : 1.0 / (e[l] + 0.0);  // w_Coef, HBEBarot_hspm_vfvm_mod.f90:281-290
    }
  int rc = 0;
  auto A = [&](double** d, size_t n) {
    if (!rc) rc = lf_alloc(s, d, n);
  };
  auto U = [&](double** d, const std::vector<double>& v) {
    if (!rc) rc = lf_upload(s, d, v);
  };
  U(&s->d_cdk, s->cdk);
  U(&s->d_fdk, s->fdk);
  U(&s->d_rcdk, s->rcdk);
  U(&s->d_rfdk, s->rfdk);
  U(&s->d_zck, s->zck);
  U(&s->d_mu, p->mu);
  U(&s->d_cs, cs);
  U(&s->d_nn1, nn1);
  U(&s->d_e, e);
  U(&s->d_hcoef, hcoef);
  U(&s->d_two_omega, std::vector<double>(R, 2.0 * 7.292115e-5));
  A(&s->d_dmom, L);
  A(&s->d_dtrc, L);
  A(&s->d_advf, L);
  A(&s->H, NC * np);
  A(&s->KV, NC * np);
  A(&s->KD, NC * np);
  A(&s->BF, R * np);
  for (int lv = 0; lv < 3; lv++) {
    A(&s->U[lv], NC * np);
    A(&s->V[lv], NC * np);
    A(&s->TS[lv], 2 * NC * np);
    A(&s->SP[lv], R * np);
  }
  A(&s->ucos, NC * np);
  A(&s->vcos, NC * np);
  A(&s->VD, 2 * NC * np);
  A(&s->OMG, R * (K + 1) * np);
  A(&s->Dens, NC * np);
  A(&s->Hyd, NC * np);
  A(&s->Geo, NC * np);
  A(&s->GLon, NC * np);
  A(&s->GLat, NC * np);
  A(&s->Am, NC * np);
  A(&s->Bg, NC * np);
  A(&s->Cg, NC * np);
  A(&s->URHS, NC * np);
  A(&s->VRHS, NC * np);
  A(&s->Upr, NC * np);
  A(&s->Vpr, NC * np);
  A(&s->Hm, NC * np);
  A(&s->TSpr, 2 * NC * np);
  A(&s->HT, 2 * NC * np);
  A(&s->FLX, 2 * R * (K + 1) * np);
  A(&s->FU, 2 * NC * np);
  A(&s->FV, 2 * NC * np);
  A(&s->VF, 2 * NC * np);
  for (auto& b : s->b2) A(&b, R * np);
  A(&s->wVD, (size_t)L * 2 * NC);
  A(&s->wGeo, (size_t)L * NC);
  A(&s->wC, (size_t)L * NC);
  A(&s->w1, (size_t)L * 2 * NC);
  A(&s->w2, (size_t)L * 2 * NC);
  A(&s->wB, (size_t)L * 2 * R);
  A(&s->wC2, (size_t)L * R);
  A(&s->wDSP, (size_t)L * R);
  A(&s->wDDiv, (size_t)L * R);
  A(&s->wZero, (size_t)L * R);
  s->lvwork = 3l * K * np;
  A(&s->vwork, s->lvwork);
  if (rc) {
    lf_free(s);
    delete s;
    return rc;
  }
  hipStream_t st = nullptr;
  auto fill = [&](double* d, double v, long n) -> int {
    hipLaunchKernelGGL(k_lf_fill, lf_grid(n), dim3(LF_THREADS), 0, st, d, v, n);
    DSHT_HIP(hipGetLastError());
    return 0;
  };
  // constant fields: xyz_H = D, the vertical viscosity/diffusivity, bottom friction; the
  // state and every work array start at zero
  if (!rc) rc = fill(s->H, c.depth, NC * np);
  if (!rc) rc = fill(s->KV, c.vvisc, NC * np);
  if (!rc) rc = fill(s->KD, c.vdiff, NC * np);
  if (!rc) rc = fill(s->BF, c.btm_frict, R * np);
  if (!rc) rc = fill(s->wZero, 0.0, (long)L * R);
  for (int lv = 0; lv < 3 && !rc; lv++) {
    rc = fill(s->U[lv], 0.0, NC * np);
    if (!rc) rc = fill(s->V[lv], 0.0, NC * np);
    if (!rc) rc = fill(s->TS[lv], 0.0, 2 * NC * np);
    if (!rc) rc = fill(s->SP[lv], 0.0, R * np);
  }
  if (!rc) {
    hipLaunchKernelGGL(k_lf_geo, lf_grid2(NC, np), dim3(LF_THREADS), 0, st, s->Geo, np, K, s->d_zck, c.depth, c.grav,
                       (long)NC);
    if (hipGetLastError() != hipSuccess) rc = fail(DSHT_EDEVICE, "LF-AM3 pool: geopotential fill failed");
  }
  // lateral viscosity / diffusion factors 1/(1 - dt w_HViscCoefH), 1/(1 - dt w_HDiffCoefH)
  // (LPhys_DIFF_spm_mod.f90:148-154) and the advection filter (HBEBaroc_hspm_vfvm_mod.f90:83-95)
  const double pm[5] = {c.dt, c.hvisc, c.numdiff, (double)c.numdiff_ord, 1.0};
  const double pt[5] = {c.dt, c.hdiff, c.numdiff, (double)c.numdiff_ord, 0.0};
  if (!rc) rc = dsht_coef_d(p, DSHT_COEF_BEDIFF, pm, 5, s->d_dmom, st);
  if (!rc) rc = dsht_coef_d(p, DSHT_COEF_BEDIFF, pt, 5, s->d_dtrc, st);
  if (!rc) rc = dsht_coef_d(p, DSHT_COEF_ADVFILTER, nullptr, 0, s->d_advf, st);
  if (!rc && hipDeviceSynchronize() != hipSuccess) rc = fail(DSHT_EDEVICE, "LF-AM3 pool: set-up failed");
  if (rc) {
    lf_free(s);
    delete s;
    return rc;
  }
  *out = s;
  return 0;
}

int dsht_lfam3_destroy(dsht_lfam3* s) {
  if (!s) return 0;
  lf_free(s);
  delete s;
  return 0;
}

int dsht_lfam3_set_rotation(dsht_lfam3* s, int run, double omega) {
  if (!s) return fail(DSHT_ESTATE, "LF-AM3 pool is NULL");
  if (run < 0 || run >= s->R) return fail(DSHT_EARG, "run %d out of range [0, %d)", run, s->R);
  const double two = 2.0 * omega;
  DSHT_HIP(hipMemcpy(s->d_two_omega + run, &two, sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

int dsht_lfam3_state(dsht_lfam3* s, int field, int level, double** ptr, long* ncols) {
  if (!s) return fail(DSHT_ESTATE, "LF-AM3 pool is NULL");
  if (level < DSHT_LF_B || level > DSHT_LF_A) return fail(DSHT_EARG, "bad time level %d", level);
  if (!ptr) return fail(DSHT_EARG, "NULL output");
  const long NC = (long)s->R * s->K;
  double* d = nullptr;
  long n = 0;
  switch (field) {
    case DSHT_LF_U: d = s->U[level]; n = NC; break;
    case DSHT_LF_V: d = s->V[level]; n = NC; break;
    case DSHT_LF_TRC: d = s->TS[level]; n = 2 * NC; break;
    case DSHT_LF_SFCPRES: d = s->SP[level]; n = s->R; break;
    default: return fail(DSHT_EARG, "bad field %d", field);
  }
  *ptr = d;
  if (ncols) *ncols = n;
  return 0;
}

int dsht_lfam3_units(const dsht_lfam3* s, int nruns, long* fwd_units, long* inv_units) {
  if (!s) return fail(DSHT_ESTATE, "LF-AM3 pool is NULL");
  if (fwd_units) *fwd_units = (42l * s->K + 9) * nruns;
  if (inv_units) *inv_units = (28l * s->K + 9) * nruns;
  return 0;
}

int dsht_lfam3_copy_run(dsht_lfam3* s, int dst, int src, void* stream) {
  if (!s) return fail(DSHT_ESTATE, "LF-AM3 pool is NULL");
  if (dst < 0 || src < 0 || dst >= s->R || src >= s->R) return fail(DSHT_EARG, "run index out of range");
  if (dst == src) return 0;
  hipStream_t st = (hipStream_t)stream;
  const size_t lay = (size_t)s->K * s->np * sizeof(double), two = (size_t)s->np * sizeof(double);
  for (int lv = 0; lv < 2; lv++) {
    DSHT_HIP(hipMemcpyAsync((char*)s->U[lv] + dst * lay, (char*)s->U[lv] + src * lay, lay, hipMemcpyDeviceToDevice, st));
    DSHT_HIP(hipMemcpyAsync((char*)s->V[lv] + dst * lay, (char*)s->V[lv] + src * lay, lay, hipMemcpyDeviceToDevice, st));
    DSHT_HIP(hipMemcpyAsync((char*)s->TS[lv] + 2 * dst * lay, (char*)s->TS[lv] + 2 * src * lay, 2 * lay,
                            hipMemcpyDeviceToDevice, st));
    DSHT_HIP(hipMemcpyAsync((char*)s->SP[lv] + dst * two, (char*)s->SP[lv] + src * two, two, hipMemcpyDeviceToDevice,
                            st));
  }
  DSHT_HIP(hipMemcpyAsync(s->d_two_omega + dst, s->d_two_omega + src, sizeof(double), hipMemcpyDeviceToDevice, st));
  return 0;
}

int dsht_lfam3_step(dsht_lfam3* s, int nr, void* stream) {
  if (!s) return fail(DSHT_ESTATE, "LF-AM3 pool is NULL");
  if (nr < 0 || nr > s->R) return fail(DSHT_EARG, "nruns = %d out of range [0, %d]", nr, s->R);
  if (nr == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  dsht_plan* p = s->p;
  const dsht_lfam3_config& c = s->cfg;
  const int K = s->K, I = s->I, L = s->L;
  const long np = s->np;
  const int nc = nr * K, nt = 2 * nc;
  const long n = (long)nc * np, ntr = (long)nt * np, n2 = (long)nr * np;
  const double dt = c.dt;
  const double gamma = 1.0 - 2.0 * c.alpha;
  const double c56dt = 5.0 / 6.0 * dt;
  double* const* b2 = s->b2;
  double *UB = s->U[0], *UN = s->U[1], *UA = s->U[2];
  double *VB = s->V[0], *VN = s->V[1], *VA = s->V[2];
  double *TB = s->TS[0], *TN = s->TS[1], *TA = s->TS[2];
  double *PN = s->SP[1], *PA = s->SP[2];
  int rc;
  // ---- preparation: barotropic velocities of levels N and B
  if ((rc = lf_uvbarot(s, UN, VN, b2[UBN], b2[VBN], nr, st))) return rc;
  if ((rc = lf_uvbarot(s, UB, VB, b2[UBB], b2[VBB], nr, st))) return rc;
  // ---- predictor
  if ((rc = lf_vordiv(s, UN, VN, nc, st))) return rc;
  if ((rc = lf_omega(s, s->VD + (long)nc * np, nr, st))) return rc;  // H_A = H_B: omega from H_B
  EosArgs ea;
  ea.ref_dens = c.ref_dens;
  ea.ptemp_dep = -c.beta_t;
  ea.press_dep = 1.0 / (c.ref_dens * (c.cs0 * c.cs0)) - c.ref_temp * (c.beta_t * c.beta_t) / (c.ref_dens * c.cp0);
  ea.salt_dep = c.beta_s;
  ea.ref_temp = c.ref_temp;
  ea.ref_salt = c.ref_salt;
  ea.depth = c.depth;
  ea.grav = c.grav;
  LF_LAUNCH2(k_lf_eos, nc, np, st, TN, np, K, s->d_zck, ea, (long)nc, s->Dens);
  if ((rc = lf_hydpres(s, nr, st))) return rc;
  if ((rc = lf_mom_rhs(s, UN, VN, UN, UB, VN, VB, gamma, 1.0 - gamma, PN, nc, st))) return rc;
  LF_LAUNCH2(k_lf_pred_uv, nc, np, st, s->URHS, s->VRHS, UN, UB, VN, VB, s->d_two_omega, s->d_mu, np, I, K,
             0.5 * (2.0 * dt), c56dt, c.alpha > 0.0 ? 1 : 0, (long)nc, s->Upr, s->Vpr);
  if ((rc = lf_uvbarot(s, s->Upr, s->Vpr, b2[UBT], b2[VBT], nr, st))) return rc;
  LF_LAUNCH2(k_lf_barot_shift, nc, np, st, s->Upr, s->Vpr, b2[UBT], b2[VBT], b2[UBN], b2[VBN], np, K, (long)nc);
  if ((rc = lf_htrc(s, TN, UN, VN, s->H, nr, st))) return rc;
  LF_LAUNCH2(k_lf_pred_trc, nt, np, st, TN, TB, s->HT, s->H, np, K, c56dt, (long)nt, s->TSpr);
  // ---- corrector
  if ((rc = lf_vordiv(s, s->Upr, s->Vpr, nc, st))) return rc;
  if ((rc = lf_omega(s, s->VD + (long)nc * np, nr, st))) return rc;
  LF_LAUNCH2(k_lf_eos, nc, np, st, s->TSpr, np, K, s->d_zck, ea, (long)nc, s->Dens);
  if ((rc = lf_hydpres(s, nr, st))) return rc;
  const double alpha_ = 0.5, gamma_ = 0.5;
  const double cn = gamma_ + alpha_, cb = 1.0 - alpha_ - gamma_;
  if ((rc = lf_mom_rhs(s, s->Upr, s->Vpr, UN, UB, VN, VB, cn, cb, PN, nc, st))) return rc;
  // ImplUV (DOGCM_Phys_hspm_vfvm_mod.f90:666-748): implicit vertical viscosity, then
  // backward-Euler lateral viscosity on vorticity and divergence
  for (int r = 0; r < nr; r++) {
    const long o = (long)r * K * np;
    if ((rc = launch_col_vdiff(UN + o, s->URHS + o, s->KV + o, s->H + o, np, K, s->d_cdk, s->d_fdk, dt, c.lambda,
                               c.mom_sfc_bc == 'D', c.mom_btm_bc == 'D', s->BF + r * np, UA + o, np, s->vwork,
                               (int)np, st)))
      return rc;
    if ((rc = launch_col_vdiff(VN + o, s->VRHS + o, s->KV + o, s->H + o, np, K, s->d_cdk, s->d_fdk, dt, c.lambda,
                               c.mom_sfc_bc == 'D', c.mom_btm_bc == 'D', s->BF + r * np, VA + o, np, s->vwork,
                               (int)np, st)))
      return rc;
  }
  LF_LAUNCH2(k_lf_uvcos, nc, np, st, UA, VA, s->d_cs, np, I, (long)nc, s->ucos, s->vcos);
  if ((rc = dsht_uv2vordiv_d(p, s->ucos, s->vcos, np, s->wVD, s->wVD + nc, 2l * nc, 1, nc, s->d_dmom, st))) return rc;
  if ((rc = dsht_vordiv2uv_d(p, s->wVD, s->wVD + nc, 2l * nc, 1, UA, VA, np, nc, st))) return rc;
  LF_LAUNCH(k_lf_rhs_diff, n, st, UA, UN, VA, VN, dt, n, s->URHS, s->VRHS);
  if ((rc = lf_uvbarot(s, s->URHS, s->VRHS, b2[UBF], b2[VBF], nr, st))) return rc;
  LF_LAUNCH2(k_lf_baroc_force, nr, np, st, b2[UBF], b2[VBF], b2[UBN], b2[VBN], b2[UBB], b2[VBB], s->d_two_omega,
             s->d_mu, np, I, cn, cb, (long)nr);
  LF_LAUNCH2(k_lf_cori_step, nc, np, st, s->URHS, s->VRHS, UN, VN, s->d_two_omega, s->d_mu, np, I, K, alpha_ * dt, dt, 0,
             (long)nc, UA, VA);
  // ---- barotropic mode: one iteration with dts = dt / SubCycleNum (rigid lid)
  const double dts = dt / c.sub_cycles;
  LF_LAUNCH2(k_lf_barot_rhs_grid, nr, np, st, b2[UBN], b2[VBN], b2[UBF], b2[VBF], PN, PN, s->d_two_omega, s->d_mu,
             s->d_cs, np, I, (long)nr, b2[AM2], b2[BG2], b2[SPM]);
  if ((rc = dsht_uv2vordiv_d(p, b2[BG2], b2[AM2], np, s->wB, s->wB + nr, 2l * nr, 1, nr, nullptr, st))) return rc;
  if ((rc = dsht_fwd_d(p, DSHT_VAL, b2[SPM], np, s->wC2, nr, 1, nr, s->d_nn1, st))) return rc;
  LF_LAUNCH2(k_lf_spec_axpy, L, nr, st, s->wB + nr, 2l * nr, s->wC2, (long)nr, L, nr, c.ref_dens * (p->R * p->R), 1);
  if ((rc = dsht_vordiv2uv_d(p, s->wB, s->wB + nr, 2l * nr, 1, b2[UBR], b2[VBR], np, nr, st))) return rc;
  LF_LAUNCH2(k_lf_cori_step, nr, np, st, b2[UBR], b2[VBR], b2[UBN], b2[VBN], s->d_two_omega, s->d_mu, np, I, 1,
             alpha_ * dts, dts, 1, (long)nr, b2[UBAR], b2[VBAR]);
  // HBEBarot_Update_LinFreeSfc, rigid lid: project onto the divergence-free part
  LF_LAUNCH2(k_lf_uvcos, nr, np, st, b2[UBAR], b2[VBAR], s->d_cs, np, I, (long)nr, b2[AM2], b2[BG2]);
  if ((rc = dsht_uv2vordiv_d(p, b2[AM2], b2[BG2], np, s->wB, s->wB + nr, 2l * nr, 1, nr, nullptr, st))) return rc;
  LF_LAUNCH2(k_lf_barot_proj, L, nr, st, s->wB + nr, 2l * nr, s->d_hcoef, s->d_e, c.ref_dens, alpha_ * dts, L, nr,
             s->wDSP, s->wDDiv);
  if ((rc = dsht_inv_d(p, DSHT_VAL, s->wDSP, nr, 1, b2[DSP], np, nr, nullptr, st))) return rc;
  if ((rc = dsht_vordiv2uv_d(p, s->wZero, s->wDDiv, nr, 1, b2[DU], b2[DV], np, nr, st))) return rc;
  LF_LAUNCH(k_lf_add2, n2, st, b2[UBAR], b2[DU], b2[VBAR], b2[DV], PN, b2[DSP], n2, b2[UBA], b2[VBA], PA);
  // ---- full velocity, time-centred fields, tracers
  if ((rc = lf_uvbarot(s, UA, VA, b2[UBT], b2[VBT], nr, st))) return rc;
  LF_LAUNCH2(k_lf_final_uv, nc, np, st, UA, VA, b2[UBT], b2[VBT], b2[UBA], b2[VBA], UN, UB, VN, VB, s->H, np, K, (long)nc,
             s->Upr, s->Vpr, s->Hm);
  if ((rc = lf_vordiv(s, s->Upr, s->Vpr, nc, st))) return rc;
  if ((rc = lf_omega(s, s->VD + (long)nc * np, nr, st))) return rc;  // from H_N
  if ((rc = lf_htrc(s, s->TSpr, s->Upr, s->Vpr, s->Hm, nr, st))) return rc;
  // ImplTRC (DOGCM_Phys_hspm_vfvm_mod.f90:544-662)
  for (int rt = 0; rt < 2 * nr; rt++) {
    const long o = (long)rt * K * np, ol = (long)(rt / 2) * K * np;
    if ((rc = launch_col_vdiff(TN + o, s->HT + o, s->KD + ol, s->H + ol, np, K, s->d_cdk, s->d_fdk, dt, c.lambda,
                               c.trc_sfc_bc == 'D', c.trc_btm_bc == 'D', nullptr, TA + o, np, s->vwork, (int)np, st)))
      return rc;
  }
  if ((rc = dsht_fwd_d(p, DSHT_VAL, TA, np, s->w1, nt, 1, nt, s->d_dtrc, st))) return rc;
  if ((rc = dsht_inv_d(p, DSHT_VAL, s->w1, nt, 1, TA, np, nt, nullptr, st))) return rc;
  LF_LAUNCH2(k_lf_trc_final, nt, np, st, TA, s->H, np, K, (long)nt);
  // ---- time levels
  LF_LAUNCH(k_lf_rotate, n, st, UB, UN, UA, n);
  LF_LAUNCH(k_lf_rotate, n, st, VB, VN, VA, n);
  LF_LAUNCH(k_lf_rotate, ntr, st, TB, TN, TA, ntr);
  LF_LAUNCH(k_lf_rotate, n2, st, s->SP[0], PN, PA, n2);
  return 0;
}

}  // extern "C"
