// column.hip -- vertical column diagnostics on device grids, between the inverse and the
// forward transforms of a dynamics step.
//   HBEDiagnose_OMG     (model/dogcm/src/dynamics/hspm_vfvm/HBEDiagnose_hspm_vfvm_mod.f90:90-143)
//   HBEDiagnose_UVBarot (same file:221-255)
//   HBEDiagnose_HydPres (same file:290-325)
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

}  // namespace dsht
