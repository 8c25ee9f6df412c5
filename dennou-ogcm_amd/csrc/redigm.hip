// redigm.hip -- the grid-space part of dogcm's Redi/GM isopycnal mixing of the tracers
// (LPhys_RediGM_hspm_vfvm_AddMixingTerm, model/dogcm/src/physics/lphys/
// LPhys_RediGM_hspm_vfvm_mod.f90:218-372) on device grids.
//
// One thread per grid point walks its column.  A first pass finds the two LDD95 tapering
// depths (the surface and bottom scans of TaperingFunc_LDD95, LPhys_RediGMHelper_mod.f90:
// 392-428): it evaluates the slopes of the top layers down to the first layer deeper than
// the eddy displacement and of the bottom layers up to the first one higher than it above
// the bottom, and stops there, as the reference's scans exit.  The second pass walks every
// layer once with a one-layer window: vertical differences on the half levels (calc_GradTRC
// :412-424), the JM95 expansion coefficients and the slopes (calc_IsoNeutralSlope_new
// :745-775), the taper (DM95 or GKW91 times LDD95), the Redi and skew fluxes of both tracers
// on the layer and on the half level above it (calc_IsopycDiffFlux :497-553, calc_SkewFlux
// :607-676), and writes per tracer (FLonRedi + FLonGM) cos(lat), (FLatRedi + FLatGM)
// cos(lat) and the vertical flux divergence (FSig above - FSig below) RCDK / H of the layer
// above (append_Redi_GM_RHS :340-366).  Every expression keeps the reference's order of
// operations and no multiply-add is contracted, so the results equal the reference loops in
// FP64 up to the last bit of tanh and sin.
//
// Bound: HBM.  Algorithmic bytes per call: the 8 input layer fields (ptemp, salt, h, z, the
// four gradients) and the 6 output layer fields once: (14 km + 1) I J 8 bytes; the first
// pass rereads the few layers of the scans.

#include <hip/hip_runtime.h>

#include <cmath>

#include "dsht_internal.hpp"

#pragma clang fp contract(off)

namespace dsht {

namespace {

// EOS_JM95_alpha_beta (EOS_JM95_mod.f90:284-330): the McDougall (1987) polynomials of
// alpha/beta and beta in theta [degC], S - 35 [psu] and p [dbar]
constexpr double AB000 = 0.665157e-1, AB100 = 0.170907e-1, AB200 = -0.203814e-3, AB300 = 0.298357e-5,
                 AB400 = -0.255019e-7, AB010 = 0.378110e-2, AB110 = -0.846960e-4, AB011 = -0.164759e-6,
                 AB012 = -0.251520e-11, AB020 = -0.678662e-5, AB001 = 0.380374e-4, AB101 = -0.933746e-6,
                 AB201 = 0.791325e-8, AB202 = 0.512857e-12, AB003 = -0.302285e-13;
constexpr double B000 = 0.785567e-3, B100 = -0.301985e-5, B200 = 0.555579e-7, B300 = -0.415613e-9,
                 B010 = -0.356603e-6, B110 = 0.788212e-8, B011 = 0.408195e-10, B012 = -0.602281e-15,
                 B020 = 0.515032e-8, B001 = -0.121555e-7, B101 = 0.192867e-9, B201 = -0.213127e-11,
                 B002 = 0.176621e-12, B102 = -0.175379e-14, B003 = 0.121551e-17;
constexpr double PI_REF = 3.1415926535897932;  // DOGCM_Admin_Constants_mod PI

struct RediArgs {
  const double *pt, *sa, *h, *z, *topo, *cdk, *rfdk, *glon, *glat, *lat_cs;
  double *flon, *flat, *vdiv;
  long ldg, ldgr, ldo, npts;  // pitches: inputs, gradients, outputs
  int km, I, J, gkw91;
  double R, rg, kr, kg, smax, sd, two_omega;
};

// EOSDriver_alpha_beta with p = -RefDens*Grav*z [Pa] (calc_IsoNeutralSlope_new :740-743):
// K2degC(theta), Pa2bar(p)*10 (EOSDriver_mod.f90), then the JM95 polynomials
__device__ inline void jm95_alpha_beta(double theta, double s, double z, double rg, double& al, double& be) {
  const double t = theta - 273.15;
  const double p = (1e-05 * (-(rg * z))) * 10.0;
  const double sm = s - 35.0;
  const double adb = ((AB000 + t * (AB100 + t * (AB200 + t * (AB300 + t * AB400)))) +
                      sm * (((AB010 + AB110 * t) + p * (AB011 + AB012 * p)) + sm * AB020)) +
                     p * ((AB001 + t * (AB101 + t * (AB201 + p * AB202))) + (p * p) * AB003);
  be = ((B000 + t * (B100 + t * (B200 + t * B300))) + sm * (((B010 + B110 * t) + p * (B011 + B012 * p)) + sm * B020)) +
       p * ((B001 + t * ((B101 + t * B201) + p * B102)) + p * (B002 + p * B003));
  al = adb * be;
}

// DSig of x on the half levels above (da) and below (db) layer k, from x at k-1, k, k+1
// (the top and bottom half levels copy their neighbours, calc_GradTRC :420-424)
__device__ inline void dsig_pair(int k, int km, const double* rfdk, double xm, double xc, double xp, double& da,
                                 double& db) {
  const double up = k > 0 ? (xm - xc) * rfdk[k] : 0.0;
  const double dn = k + 1 < km ? (xc - xp) * rfdk[k + 1] : 0.0;
  da = k > 0 ? up : dn;
  db = k + 1 < km ? dn : up;
}

// everything of one layer that the fluxes of the layer and of the half level above need
struct Lay {
  double h, tap, slon, slat;
  double gl[2], ga[2];  // horizontal gradients / RPlanet (temperature, salinity)
  double da[2], db[2];  // DSig above and below the layer
};

// the slopes of layer k (calc_IsoNeutralSlope_new :745-775)
__device__ inline void layer_slopes(const RediArgs& a, int k, double tm, double tc, double tp, double sm, double sc,
                                    double sp, double z, Lay& L) {
  dsig_pair(k, a.km, a.rfdk, tm, tc, tp, L.da[0], L.db[0]);
  dsig_pair(k, a.km, a.rfdk, sm, sc, sp, L.da[1], L.db[1]);
  double al, be;
  jm95_alpha_beta(tc, sc, z, a.rg, al, be);
  const double den = (1e-10 + (al * 0.5) * (L.da[0] + L.db[0])) - (be * 0.5) * (L.da[1] + L.db[1]);
  const double fac = -(L.h / den);
  L.slon = fac * (al * L.gl[0] - be * L.gl[1]);
  L.slat = fac * (al * L.ga[0] - be * L.ga[1]);
}

// the loads of one layer: h, z and the four gradients of layer k, and ptemp/salt of the next
// layer in the walk's direction (k+1 down, k-1 up; unused past the end)
struct Pre {
  double h, z, g0, a0, g1, a1, tn, sn;
};

constexpr int RING = 2;  // layers whose loads are in flight ahead of the one being computed

constexpr int REDI_LDS_LAYERS = 1024;  // layers whose half-level weights the block keeps in LDS

__global__ __launch_bounds__(256, 2) void k_redigm_col(RediArgs a) {
  // per layer k: m1 = cdk[k]/(cdk[k-1]+cdk[k]), m2 = cdk[k-1]/(cdk[k-1]+cdk[k]) (r_m1, r_m2 of
  // the half level on top of layer k, calc_IsopycDiffFlux :491-494) and z_RCDK = 1/cdk[k]
  // (DOGCM_Admin_Grid_mod.f90:315), each the same IEEE division as in the loop
  __shared__ double wm1[REDI_LDS_LAYERS], wm2[REDI_LDS_LAYERS], wrc[REDI_LDS_LAYERS];
  const int km = a.km;
  for (int k = threadIdx.x; k < km; k += 256) {
    const double c1 = a.cdk[k];
    wrc[k] = 1.0 / c1;
    if (k > 0) {
      const double c0 = a.cdk[k - 1];
      wm1[k] = c1 / (c0 + c1);
      wm2[k] = c0 / (c0 + c1);
    }
  }
  __syncthreads();
  const long q = (long)blockIdx.x * 256 + threadIdx.x;
  if (q >= a.npts) return;
  const int j = (int)(q / a.I);
  const double cs = a.lat_cs[j], asn = a.lat_cs[a.J + j];
  const double topo = a.topo[q];
  const long TS = (long)km * a.ldgr;  // salinity block of the gradient inputs
  // dir = +1: the next layer is k+1 (downward walk); -1: k-1 (upward)
  auto fetch = [&](int k, int dir, Pre& f) {
    const long g = (long)k * a.ldg + q, gr = (long)k * a.ldgr + q;
    f.h = a.h[g];
    f.z = a.z[g];
    f.g0 = a.glon[gr];
    f.a0 = a.glat[gr];
    f.g1 = a.glon[TS + gr];
    f.a1 = a.glat[TS + gr];
    const int kn = k + dir;
    const bool in = kn >= 0 && kn < km;
    f.tn = in ? a.pt[g + dir * a.ldg] : 0.0;
    f.sn = in ? a.sa[g + dir * a.ldg] : 0.0;
  };
  // layer k's slopes from its loads and ptemp/salt at k-1, k, k+1
  auto slopes_of = [&](int k, const Pre& f, double tm, double tc, double tp, double sm, double sc, double sp,
                       Lay& L) {
    L.h = f.h;
    L.gl[0] = f.g0 / a.R;
    L.ga[0] = f.a0 / a.R;
    L.gl[1] = f.g1 / a.R;
    L.ga[1] = f.a1 / a.R;
    layer_slopes(a, k, tm, tc, tp, sm, sc, sp, f.z, L);
  };

  // ---- pass 1: the LDD95 scans (TaperingFunc_LDD95 :392-428), each stopping at its first hit
  double disph = 2.0 / (a.two_omega * asn);
  disph = fmin(fmax(15e3, disph), 100e3);
  const double dzmax = a.smax * disph;
  int ks = -1, kb = km;
  double tmps = 0.0, tmpb = 0.0;
  for (int dir = 1; dir >= -1; dir -= 2) {
    const int k0 = dir > 0 ? 0 : km - 1;
    Pre ring[RING];
#pragma unroll
    for (int u = 0; u < RING; u++)
      if (u < km) fetch(k0 + dir * u, dir, ring[u]);
    double xc = a.pt[(long)k0 * a.ldg + q], yc = a.sa[(long)k0 * a.ldg + q];
    double xb = xc, yb = yc;  // the layer behind in the walk (k0 itself at the start)
    double mx = 0.0;
    bool done = false;
    for (int i0 = 0; i0 < km && !done; i0 += RING) {
#pragma unroll
      for (int u = 0; u < RING; u++) {
        const int i = i0 + u;
        if (done || i >= km) continue;
        const int k = k0 + dir * i;
        const Pre c = ring[u];
        if (i + RING < km) fetch(k0 + dir * (i + RING), dir, ring[u]);
        const bool end = i + 1 == km;
        const double xn = end ? xc : c.tn, yn = end ? yc : c.sn;
        Lay L;
        if (dir > 0)
          slopes_of(k, c, xb, xc, xn, yb, yc, yn, L);
        else
          slopes_of(k, c, xn, xc, xb, yn, yc, yb, L);
        const double sab = sqrt(L.slon * L.slon + L.slat * L.slat);
        if (sab > mx) mx = sab;
        const bool hit = dir > 0 ? (0.0 - c.z) > dzmax : (c.z - (-topo)) > dzmax;
        if (hit) {
          if (dir > 0) {
            ks = k;
            tmps = fmin(mx, a.smax) * disph;
          } else {
            kb = k;
            tmpb = fmin(mx, a.smax) * disph;
          }
          done = true;
        }
        xb = xc, yb = yc, xc = xn, yc = yn;
      }
    }
  }

  // ---- pass 2: every layer, the half level above it, and the divergence of the layer above
  const double kr = a.kr, kg = a.kg;
  Lay P;                                          // layer k-1
  double pl[2] = {0.0, 0.0}, pa[2] = {0.0, 0.0};  // psi of layer k-1 (GM)
  double fs_up[2] = {0.0, 0.0};                   // FSig on the half level above layer k-1
  double tc = a.pt[q], sc = a.sa[q], tm = tc, sm = sc;
  const long oS = (long)km * a.ldo;
  auto body = [&](int k, const Pre& f) {
    const bool last = k + 1 == km;
    const double tp = last ? tc : f.tn, sp = last ? sc : f.sn;
    Lay C;
    slopes_of(k, f, tm, tc, tp, sm, sc, sp, C);
    const double z = f.z;
    // taper (prepare_SlopeTapering: T1 T2)
    const double mag2 = C.slon * C.slon + C.slat * C.slat;
    const double t1 =
        a.gkw91 ? fmin(1.0, (a.smax * a.smax) / mag2) : 0.5 * (1.0 + tanh((a.smax - sqrt(mag2)) / a.sd));
    double r = 1.0;
    if (k <= ks) r = fmin((0.0 - z) / (tmps + 1e-13), 1.0);
    if (k >= kb) r = fmin((z - (-topo)) / (tmpb + 1e-13), 1.0);
    // r = 1 gives 0.5 (1 + sin(pi/2)) = 1 exactly; sin only where the ratio is below 1
    const double t2 = r == 1.0 ? 1.0 : 0.5 * (1.0 + sin(PI_REF * (r - 0.5)));
    C.tap = t1 * t2;
    const long o = (long)k * a.ldo + q;
    double cl[2], ca[2];
#pragma unroll
    for (int x = 0; x < 2; x++) {
      const double dz = (0.5 * (C.da[x] + C.db[x])) / C.h;
      const double flr = kr * (C.gl[x] + (C.tap * C.slon) * dz);
      const double far_ = kr * (C.ga[x] + (C.tap * C.slat) * dz);
      cl[x] = -((kg * C.tap) * C.slat);  // Psi(LON)
      ca[x] = (kg * C.tap) * C.slon;     // Psi(LAT)
      const double flg = -(ca[x] * dz);
      const double fag = cl[x] * dz;
      a.flon[x * oS + o] = (flr + flg) * cs;
      a.flat[x * oS + o] = (far_ + fag) * cs;
    }
    if (k > 0) {  // the half level between layers k-1 (up) and k (lo)
      const double m1 = wm1[k], m2 = wm2[k];
      const double sl = (m1 * P.tap) * P.slon + (m2 * C.tap) * C.slon;
      const double sa_ = (m1 * P.tap) * P.slat + (m2 * C.tap) * C.slat;
      const double hden = m1 * P.h + m2 * C.h;
      const double rc = wrc[k - 1];
#pragma unroll
      for (int x = 0; x < 2; x++) {
        const double glon = m1 * P.gl[x] + m2 * C.gl[x];
        const double glat = m1 * P.ga[x] + m2 * C.ga[x];
        const double fr = kr * ((sl * glon + sa_ * glat) + ((sl * sl + sa_ * sa_) * C.da[x]) / hden);
        const double pln = m1 * pl[x] + m2 * cl[x];
        const double pat = m1 * pa[x] + m2 * ca[x];
        const double fg = pat * glon - pln * glat;
        const double fs = fr + fg;
        a.vdiv[x * oS + (long)(k - 1) * a.ldo + q] = ((fs_up[x] - fs) * rc) / P.h;
        fs_up[x] = fs;
      }
    }
    P = C;
    pl[0] = cl[0], pl[1] = cl[1], pa[0] = ca[0], pa[1] = ca[1];
    tm = tc, sm = sc, tc = tp, sc = sp;
  };
  {
    Pre ring[RING];
#pragma unroll
    for (int u = 0; u < RING; u++)
      if (u < km) fetch(u, 1, ring[u]);
    for (int k0 = 0; k0 < km; k0 += RING) {
#pragma unroll
      for (int u = 0; u < RING; u++) {
        const int k = k0 + u;
        if (k < km) {
          const Pre c = ring[u];
          if (k + RING < km) fetch(k + RING, 1, ring[u]);
          body(k, c);
        }
      }
    }
  }
  {  // the bottom layer: FSig below it is zero
    const double rc = wrc[km - 1];
    for (int x = 0; x < 2; x++)
      a.vdiv[x * oS + (long)(km - 1) * a.ldo + q] = ((fs_up[x] - 0.0) * rc) / P.h;
  }
}

// rhs[c] = (rhs[c] + hdiv[c]) + vdiv[c] over the 2 km layer grids of both tracers
__global__ __launch_bounds__(256) void k_redigm_rhs(double* __restrict__ rt, double* __restrict__ rs, long ldo,
                                                    const double* __restrict__ hdiv, const double* __restrict__ vdiv,
                                                    long ldw, int km, long npts) {
  const long q = (long)blockIdx.x * 256 + threadIdx.x;
  const int c = blockIdx.y;
  if (q >= npts) return;
  double* r = c < km ? rt + (long)c * ldo : rs + (long)(c - km) * ldo;
  const long w = (long)c * ldw + q;
  r[q] = (r[q] + hdiv[w]) + vdiv[w];
}

}  // namespace

int launch_redigm_col(const dsht_plan* p, const RediParams& c, const double* pt, const double* sa, const double* h,
                      const double* z, long ldg, int km, const double* topo, const double* cdk, const double* rfdk,
                      const double* glon, const double* glat, long ldgr, double* flon, double* flat, double* vdiv,
                      long ldo, hipStream_t s) {
  RediArgs a;
  a.pt = pt, a.sa = sa, a.h = h, a.z = z, a.topo = topo, a.cdk = cdk, a.rfdk = rfdk, a.glon = glon, a.glat = glat;
  a.lat_cs = p->d_lat_cs;
  a.flon = flon, a.flat = flat, a.vdiv = vdiv;
  a.ldg = ldg, a.ldgr = ldgr, a.ldo = ldo, a.npts = (long)p->I * p->J;
  a.km = km, a.I = p->I, a.J = p->J, a.gkw91 = c.gkw91;
  a.R = p->R, a.rg = c.ref_dens * c.grav, a.kr = c.kappa_redi, a.kg = c.kappa_gm, a.smax = c.slope_max, a.sd = c.sd;
  a.two_omega = 2.0 * c.omega;
  const double bytes = (14.0 * km + 1.0) * (double)a.npts * 8.0;
  ProfScope prof("redigm_col", s, bytes, 0.0);
  hipLaunchKernelGGL(k_redigm_col, dim3((unsigned)((a.npts + 255) / 256)), dim3(256), 0, s, a);
  DSHT_HIP(hipGetLastError());
  return 0;
}

int launch_redigm_rhs(const dsht_plan* p, double* rt, double* rs, long ldo, const double* hdiv, const double* vdiv,
                      long ldw, int km, hipStream_t s) {
  const long npts = (long)p->I * p->J;
  const double bytes = 4.0 * 2.0 * km * (double)npts * 8.0;
  ProfScope prof("redigm_rhs", s, bytes, 0.0);
  hipLaunchKernelGGL(k_redigm_rhs, dim3((unsigned)((npts + 255) / 256), 2 * km), dim3(256), 0, s, rt, rs, ldo, hdiv,
                     vdiv, ldw, km, npts);
  DSHT_HIP(hipGetLastError());
  return 0;
}

}  // namespace dsht
