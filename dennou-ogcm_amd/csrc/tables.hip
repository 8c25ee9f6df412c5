// tables.hip -- FP64 setup on the device: Gauss-Legendre nodes and weights, the
// normalised associated Legendre functions P_n^m and H_n^m = (1-mu^2) dP_n^m/dmu,
// swizzled into MFMA fragment order for the forward and inverse contractions.
//
// Replaces the table set-up inside wa_Initial (SPMODEL/ISPACK, called from
// model/dogcm/src/util/math/SpmlUtil_mod.f90:218-223).  Normalisation follows
// doc/description/ogcm/space_discretization.tex:160-170 (int P^2 dmu = 2) and the
// m = 0 recurrence of model/dogcm/src/physics/lphys/LPhys_DIFF_spm_mod.f90:497-507.
#include <cmath>
#include <cstdio>

#include "dsht_internal.hpp"

namespace dsht {

// Newton iteration on P_J for root k (descending), weights 2/((1-x^2) P_J'(x)^2).
__global__ void k_gauss(int J, double* mu, double* wl) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= J / 2) return;
  double x = cospi((k + 0.75) / (J + 0.5));
  double p0 = 1.0, p1 = x, dp = 1.0;
  for (int it = 0; it < 64; it++) {
    p0 = 1.0;
    p1 = x;
    for (int n = 2; n <= J; n++) {
      double p2 = ((2 * n - 1) * x * p1 - (n - 1) * p0) / n;
      p0 = p1;
      p1 = p2;
    }
    dp = J * (x * p1 - p0) / (x * x - 1.0);
    double dx = p1 / dp;
    x -= dx;
    if (fabs(dx) <= 1e-17) break;
  }
  p0 = 1.0;
  p1 = x;
  for (int n = 2; n <= J; n++) {
    double p2 = ((2 * n - 1) * x * p1 - (n - 1) * p0) / n;
    p0 = p1;
    p1 = p2;
  }
  dp = J * (x * p1 - p0) / (x * x - 1.0);
  double w = 2.0 / ((1.0 - x) * (1.0 + x) * dp * dp);
  mu[J - 1 - k] = x;
  mu[k] = -x;
  wl[J - 1 - k] = w;
  wl[k] = w;
}

// derived per-latitude vectors
__global__ void k_grid_vectors(int J, int J2, int J2pad, double R, const double* mu, const double* wl,
                               double* wq, double* cwv, double* cwd, double* rs_icos,
                               double* rs_alpha, double* rs_uv) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < J) {
    double c2 = (1.0 - mu[j]) * (1.0 + mu[j]);
    double c = sqrt(c2);
    wq[j] = 0.5 * wl[j];
    rs_icos[j] = 1.0 / c;
    rs_alpha[j] = 1.0 / (R * c2);
    rs_uv[j] = 1.0 / (R * c);
  }
  if (j < J2pad) {
    if (j < J2) {
      double x = mu[J2 + j];
      double w = 0.5 * wl[J2 + j];
      cwv[j] = w;
      cwd[j] = w / ((1.0 - x) * (1.0 + x));
    } else {
      cwv[j] = 0.0;
      cwd[j] = 0.0;
    }
  }
}

__global__ void k_pro_psi(int N, int L, double R, double* pro) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= L) return;
  int n = (int)floor(sqrt((double)l));
  while (n * (n + 1) + n < l) n++;
  while (n * n > l) n--;
  pro[l] = n == 0 ? 0.0 : -R * R / ((double)n * (n + 1));
}

__device__ inline double eps_nm(int n, int m) {
  double d = (double)(4 * n * n - 1);
  return n <= 0 ? 0.0 : sqrt((double)(n * n - m * m) / d);
}

// One thread per (m, north latitude jn): upward recurrence in n with an extended
// exponent (value = x * 2^e) so sin^m(theta) at polar nodes does not underflow.
__global__ void k_legendre(int N, int J2, int J2pad, const double* mu, const long* offF,
                           const long* offI, double* tFP, double* tFH, double* tIP, double* tIH) {
  int jn = blockIdx.x * blockDim.x + threadIdx.x;
  int m = blockIdx.y;
  if (jn >= J2 || m > N) return;
  double x = mu[J2 + jn];
  double s = sqrt((1.0 - x) * (1.0 + x));
  double pm = 1.0;
  int e = 0;
  for (int k = 1; k <= m; k++) {
    pm *= sqrt((2.0 * k + 1.0) / (2.0 * k)) * s;
    if (pm < 0x1p-400) {
      pm *= 0x1p400;
      e -= 400;
    }
  }
  int nE = n_par(N, m, 0), nO = n_par(N, m, 1);
  int tE = cdiv(nE, TILE), tO = cdiv(nO, TILE);
  int kp0 = cdiv(nE, KCH) * KCH, kp1 = cdiv(nO, KCH) * KCH;
  (void)tO;
  long fbase = offF[m], ibase = offI[m];
  long kstepsF = J2pad / 4;
  double pprev = 0.0, pcur = pm;
  for (int n = m; n <= N; n++) {
    double pnext = (x * pcur - eps_nm(n, m) * pprev) / eps_nm(n + 1, m);
    double Pv = ldexp(pcur, e);
    double Hv = ldexp(-(double)n * eps_nm(n + 1, m) * pnext + (double)(n + 1) * eps_nm(n, m) * pprev, e);
    int p = (n - m) & 1;
    int k = (n - m - p) / 2;  // row index within parity
    // forward swizzle: tile (p ? tE : 0) + k/16, row k%16, K index jn
    long fi = fbase + ((long)(p ? tE : 0) + k / TILE) * kstepsF * 64 + (long)(jn / 4) * 64 +
              (k % TILE) + 16 * (jn % 4);
    tFP[fi] = Pv;
    tFH[fi] = Hv;
    // inverse swizzle: row tile jn/16, row jn%16, K index k
    int kp = p ? kp1 : kp0;
    long ii = ibase + (p ? (long)J2pad * kp0 : 0) + (long)(jn / TILE) * (kp / 4) * 64 +
              (long)(k / 4) * 64 + (jn % TILE) + 16 * (k % 4);
    tIP[ii] = Pv;
    tIH[ii] = Hv;
    pprev = pcur;
    pcur = pnext;
    if (e < 0 && fabs(pcur) > 0x1p400) {
      pcur *= 0x1p-400;
      pprev *= 0x1p-400;
      e += 400;
    }
  }
}

int build_tables(dsht_plan* p) {
  const int N = p->N, J = p->J, J2 = p->J2, J2pad = p->J2pad, L = p->L;
  auto alloc = [&](double** d, size_t n) -> int {
    DSHT_HIP(hipMalloc(d, sizeof(double) * (n ? n : 1)));
    return 0;
  };
  int rc = 0;
  if ((rc = alloc(&p->d_mu, J)) || (rc = alloc(&p->d_wq, J)) || (rc = alloc(&p->d_wl, J)) ||
      (rc = alloc(&p->d_cwv, J2pad)) || (rc = alloc(&p->d_cwd, J2pad)) ||
      (rc = alloc(&p->d_rs_icos, J)) || (rc = alloc(&p->d_rs_alpha, J)) ||
      (rc = alloc(&p->d_rs_uv, J)) || (rc = alloc(&p->d_pro_psi, L)))
    return rc;
  // table offsets
  p->offF.assign(N + 2, 0);
  p->offI.assign(N + 2, 0);
  long f = 0, i = 0;
  for (int m = 0; m <= N; m++) {
    p->offF[m] = f;
    p->offI[m] = i;
    int nE = n_par(N, m, 0), nO = n_par(N, m, 1);
    f += (long)(cdiv(nE, TILE) + cdiv(nO, TILE)) * TILE * J2pad;
    i += (long)J2pad * (cdiv(nE, KCH) * KCH + cdiv(nO, KCH) * KCH);
  }
  p->offF[N + 1] = f;
  p->offI[N + 1] = i;
  p->sizeF = f;
  p->sizeI = i;
  if ((rc = alloc(&p->d_tabFP, f)) || (rc = alloc(&p->d_tabFH, f)) || (rc = alloc(&p->d_tabIP, i)) ||
      (rc = alloc(&p->d_tabIH, i)))
    return rc;
  DSHT_HIP(hipMalloc(&p->d_offF, sizeof(long) * (N + 2)));
  DSHT_HIP(hipMalloc(&p->d_offI, sizeof(long) * (N + 2)));
  DSHT_HIP(hipMemcpy(p->d_offF, p->offF.data(), sizeof(long) * (N + 2), hipMemcpyHostToDevice));
  DSHT_HIP(hipMemcpy(p->d_offI, p->offI.data(), sizeof(long) * (N + 2), hipMemcpyHostToDevice));
  DSHT_HIP(hipMemset(p->d_tabFP, 0, sizeof(double) * f));
  DSHT_HIP(hipMemset(p->d_tabFH, 0, sizeof(double) * f));
  DSHT_HIP(hipMemset(p->d_tabIP, 0, sizeof(double) * i));
  DSHT_HIP(hipMemset(p->d_tabIH, 0, sizeof(double) * i));

  hipLaunchKernelGGL(k_gauss, dim3(cdiv(J / 2, 64)), dim3(64), 0, 0, J, p->d_mu, p->d_wl);
  DSHT_HIP(hipGetLastError());
  int nb = cdiv(J > J2pad ? J : J2pad, 128);
  hipLaunchKernelGGL(k_grid_vectors, dim3(nb), dim3(128), 0, 0, J, J2, J2pad, p->R, p->d_mu,
                     p->d_wl, p->d_wq, p->d_cwv, p->d_cwd, p->d_rs_icos, p->d_rs_alpha, p->d_rs_uv);
  DSHT_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_pro_psi, dim3(cdiv(L, 256)), dim3(256), 0, 0, N, L, p->R, p->d_pro_psi);
  DSHT_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_legendre, dim3(cdiv(J2, 64), N + 1), dim3(64), 0, 0, N, J2, J2pad, p->d_mu,
                     p->d_offF, p->d_offI, p->d_tabFP, p->d_tabFH, p->d_tabIP, p->d_tabIH);
  DSHT_HIP(hipGetLastError());
  DSHT_HIP(hipDeviceSynchronize());

  // host copies of the grid description
  p->mu.resize(J);
  p->wl.resize(J);
  p->wq.resize(J);
  p->lat.resize(J);
  p->lon.resize(p->I);
  DSHT_HIP(hipMemcpy(p->mu.data(), p->d_mu, sizeof(double) * J, hipMemcpyDeviceToHost));
  DSHT_HIP(hipMemcpy(p->wl.data(), p->d_wl, sizeof(double) * J, hipMemcpyDeviceToHost));
  DSHT_HIP(hipMemcpy(p->wq.data(), p->d_wq, sizeof(double) * J, hipMemcpyDeviceToHost));
  for (int j = 0; j < J; j++) p->lat[j] = std::asin(p->mu[j]);
  for (int k = 0; k < p->I; k++) p->lon[k] = 2.0 * M_PI * k / p->I;
  // xy_CosLat = cos(xy_Lat) (SpmlUtil_mod.f90:232) and |sin(xy_Lat)| as the reference
  // evaluates them on its latitudes
  std::vector<double> cs(2 * (size_t)J);
  for (int j = 0; j < J; j++) {
    cs[j] = std::cos(p->lat[j]);
    cs[J + j] = std::fabs(std::sin(p->lat[j]));
  }
  if ((rc = alloc(&p->d_lat_cs, 2 * (size_t)J))) return rc;
  DSHT_HIP(hipMemcpy(p->d_lat_cs, cs.data(), sizeof(double) * 2 * J, hipMemcpyHostToDevice));
  return 0;
}

}  // namespace dsht
