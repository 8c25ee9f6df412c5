// ops.hip -- diagonal spectral operators, FP64 global sums, point interpolation and
// coefficient tables.
//   diagonal multipliers: w_Lapla_w / w_LaplaInv_w (SPMODEL, cascaded at
//     SpmlUtil_mod.f90:117), w_HDiffCoefH / w_HViscCoefH (LPhys_DIFF_spm_mod.f90:140-167,
//     629-634), w_AdvFilter (HBEBaroc_hspm_vfvm_mod.f90:83-95), the Helmholtz w_Coef
//     (HBEBarot_hspm_vfvm_mod.f90:286-299), the implicit factor 1/(1-dt c)
//     (DOGCM_Phys_hspm_vfvm_mod.f90:640, 741-742)
//   sums: IntLonLat_xy, AvrLonLat_xy, ya_IntLon_xya (SpmlUtil_mod.f90:123-124)
//   Interpolate_w (SpmlUtil_mod.f90:127)
//   energy and enstrophy spectra of a streamfunction (SpmlUtil_mod.f90:131-132)
#include <cmath>

#include "dsht_internal.hpp"

namespace dsht {

__global__ void k_diag(const double* __restrict__ d, const double* __restrict__ w, double* __restrict__ out,
                       long sl, long sc, int L, int ncols) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long n = (long)L * ncols;
  if (e >= n) return;
  // iterate in the memory order of the layout
  long l, c;
  if (sc == 1) {
    l = e / ncols;
    c = e % ncols;
  } else {
    c = e / L;
    l = e % L;
  }
  long k = l * sl + c * sc;
  out[k] = d[l] * w[k];
}

int launch_diag(const double* d, const double* w, double* out, long sl, long sc, int L, int ncols,
                hipStream_t s) {
  long n = (long)L * ncols;
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_diag, dim3((unsigned)cdivl(n, 256)), dim3(256), 0, s, d, w, out, sl, sc, L, ncols);
  DSHT_HIP(hipGetLastError());
  return 0;
}

// Energy and enstrophy spectra of a streamfunction (nma_/na_EnergyFromStreamfunc_wa and
// nma_/na_EnstrophyFromStreamfunc_wa, cascaded from SPMODEL at SpmlUtil_mod.f90:131-132).
// With q = sum s_n^m P_n^m e^{i m lambda} (P normalised to 2) the global mean of q^2 is
// sum_n [w(n,0)^2 + 2 sum_{m>0} (w(n,m)^2 + w(n,-m)^2)], so on the unit sphere
//   E(n,+-m) = (1/2) n(n+1) c_m w(n,+-m)^2,   Z(n,+-m) = (1/2) n^2 (n+1)^2 c_m w(n,+-m)^2,
// c_0 = 1, c_m = 2 otherwise; summed over (n, m) they are the global means of |grad psi|^2/2
// and (lapla psi)^2/2.  nma(n, m, c) at n + (N+1) ((m+N) + (2N+1) c), zero where |m| > n;
// na(n, c) = sum over m = -n..n (ascending) at n + (N+1) c.  One thread per (n, column).
__global__ void k_spec_energy(const double* __restrict__ w, long sl, long sc, int N, int ncols, int ens,
                              double* __restrict__ nma, double* __restrict__ na) {
#pragma clang fp contract(off)
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)(N + 1) * ncols) return;
  const int n = (int)(e % (N + 1)), c = (int)(e / (N + 1));
  const double nn1 = (double)n * (double)(n + 1);
  const double f = ens ? 0.5 * nn1 * nn1 : 0.5 * nn1;
  double sum = 0.0;
  for (int m = -N; m <= N; m++) {
    double v = 0.0;
    if (m >= -n && m <= n) {
      const double x = w[((long)n * (n + 1) + m) * sl + (long)c * sc];
      v = ((x * x) * (m == 0 ? 1.0 : 2.0)) * f;
      sum += v;
    }
    if (nma) nma[n + (long)(N + 1) * ((m + N) + (long)(2 * N + 1) * c)] = v;
  }
  if (na) na[n + (long)(N + 1) * c] = sum;
}

int launch_spec_energy(const dsht_plan* p, int ens, const double* w, long sl, long sc, int ncols, double* nma,
                       double* na, hipStream_t s) {
  const long n = (long)(p->N + 1) * ncols;
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_spec_energy, dim3((unsigned)cdivl(n, 256)), dim3(256), 0, s, w, sl, sc, p->N, ncols, ens, nma,
                     na);
  DSHT_HIP(hipGetLastError());
  return 0;
}

__device__ double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i];
  __syncthreads();
  return s;
}

// sum_j wl_j sum_i x_ij * factor, one block per column (fixed order: deterministic)
__global__ __launch_bounds__(256) void k_lonlat_sum(const double* __restrict__ xy, long ldg, int I, int J,
                                                    const double* __restrict__ wl, double factor,
                                                    double* __restrict__ out) {
  __shared__ double red[4];
  const double* x = xy + (long)blockIdx.x * ldg;
  double acc = 0.0;
  for (int j = 0; j < J; j++) {
    double r = 0.0;
    for (int i = threadIdx.x; i < I; i += blockDim.x) r += x[(long)j * I + i];
    acc += wl[j] * r;
  }
  double s = block_sum(acc, red);
  if (threadIdx.x == 0) out[blockIdx.x] = s * factor;
}

int launch_lonlat_sum(const dsht_plan* p, const double* xy, long ldg, int ncols, double* out,
                      double factor, hipStream_t s) {
  if (ncols <= 0) return 0;
  hipLaunchKernelGGL(k_lonlat_sum, dim3(ncols), dim3(256), 0, s, xy, ldg, p->I, p->J, p->d_wl, factor, out);
  DSHT_HIP(hipGetLastError());
  return 0;
}

// per (j, column) longitude sum, one wave per row; out[c*out_ld + j]
__global__ __launch_bounds__(256) void k_lon_sum(const double* __restrict__ xy, long ldg, int I, int J, int ncols,
                                                 double factor, double* __restrict__ out, long out_ld) {
  long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (row >= (long)J * ncols) return;
  int c = (int)(row / J), j = (int)(row % J);
  const double* x = xy + (long)c * ldg + (long)j * I;
  double r = 0.0;
  for (int i = lane; i < I; i += 64) r += x[i];
  for (int off = 32; off > 0; off >>= 1) r += __shfl_down(r, off, 64);
  if (lane == 0) out[(long)c * out_ld + j] = r * factor;
}

int launch_lon_sum(const dsht_plan* p, const double* xy, long ldg, int ncols, double* out,
                   double factor, long out_ld, hipStream_t s) {
  long rows = (long)p->J * ncols;
  if (rows == 0) return 0;
  hipLaunchKernelGGL(k_lon_sum, dim3((unsigned)cdivl(rows, 4)), dim3(256), 0, s, xy, ldg, p->I, p->J, ncols,
                     factor, out, out_ld);
  DSHT_HIP(hipGetLastError());
  return 0;
}

__device__ inline double eps_nm2(int n, int m) {
  return n <= 0 ? 0.0 : sqrt((double)(n * n - m * m) / (double)(4 * n * n - 1));
}

// q(lon, lat) = sum_n s_n^0 P_n^0 + 2 sum_{m>0} Re(s_n^m P_n^m e^{i m lon}); one block per point,
// threads over m, FP64 recurrence with an extended exponent.
__global__ __launch_bounds__(256) void k_interp(const double* __restrict__ w, long sl, int N,
                                                const double* __restrict__ lon, const double* __restrict__ lat,
                                                double* __restrict__ out) {
  __shared__ double red[4];
  const int k = blockIdx.x;
  const double x = sin(lat[k]), sn = sqrt((1.0 - x) * (1.0 + x)), lam = lon[k];
  double acc = 0.0;
  for (int m = threadIdx.x; m <= N; m += blockDim.x) {
    double pm = 1.0;
    int e = 0;
    for (int q = 1; q <= m; q++) {
      pm *= sqrt((2.0 * q + 1.0) / (2.0 * q)) * sn;
      if (pm < 0x1p-400) { pm *= 0x1p400; e -= 400; }
    }
    double pprev = 0.0, pcur = pm, gr = 0.0, gi = 0.0;
    for (int n = m; n <= N; n++) {
      double pv = ldexp(pcur, e);
      gr += pv * w[((long)n * (n + 1) + m) * sl];
      if (m > 0) gi += pv * w[((long)n * (n + 1) - m) * sl];
      double pnext = (x * pcur - eps_nm2(n, m) * pprev) / eps_nm2(n + 1, m);
      pprev = pcur;
      pcur = pnext;
      if (e < 0 && fabs(pcur) > 0x1p400) { pcur *= 0x1p-400; pprev *= 0x1p-400; e += 400; }
    }
    if (m == 0) acc += gr;
    else {
      double sm, cm;
      sincos((double)m * lam, &sm, &cm);
      acc += 2.0 * (gr * cm - gi * sm);
    }
  }
  double s = block_sum(acc, red);
  if (threadIdx.x == 0) out[k] = s;
}

int launch_interpolate(const dsht_plan* p, const double* w, long sl, const double* lon,
                       const double* lat, double* out, int npts, hipStream_t s) {
  if (npts <= 0) return 0;
  hipLaunchKernelGGL(k_interp, dim3(npts), dim3(256), 0, s, w, sl, p->N, lon, lat, out);
  DSHT_HIP(hipGetLastError());
  return 0;
}

// coefficient tables; params are copied by value
struct CoefParams {
  double v[8];
};

__global__ void k_coef(int which, int N, int L, double R, CoefParams cp, double* out) {
  int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= L) return;
  int n = (int)floor(sqrt((double)l));
  while (n * (n + 1) + n < l) n++;
  while (n * n > l) n--;
  const double nn = (double)n * (n + 1);
  const double e = nn / (R * R);
  double v = 0.0;
  switch (which) {
    case DSHT_COEF_LAPLA: v = -nn; break;
    case DSHT_COEF_LAPLAINV: v = n == 0 ? 0.0 : -1.0 / nn; break;
    case DSHT_COEF_HDIFF: {
      int ord2 = (int)cp.v[2] / 2;
      v = -(cp.v[0] * e + cp.v[1] * pow(e, ord2));
      break;
    }
    case DSHT_COEF_HVISC: {
      int ord2 = (int)cp.v[2] / 2;
      double r2 = 2.0 / (R * R);
      v = -(cp.v[0] * (e - r2) + cp.v[1] * (pow(e, ord2) - pow(r2, ord2)));
      break;
    }
    case DSHT_COEF_ADVFILTER: {
      int Nc = (int)(0.66 * N);
      if (n < Nc) v = 1.0;
      else {
        double r = (double)(n - Nc) / (double)(N - Nc);
        v = exp(-1.0 * r * r * r * r);
      }
      break;
    }
    case DSHT_COEF_HELMHOLTZ: v = n == 0 ? cp.v[1] : 1.0 / (e + cp.v[0]); break;
    case DSHT_COEF_BEDIFF: {
      // {dt, kappa|nu, numdiff, ord, momentum}
      int ord2 = (int)cp.v[3] / 2;
      double c;
      if (cp.v[4] != 0.0) {
        double r2 = 2.0 / (R * R);
        c = -(cp.v[1] * (e - r2) + cp.v[2] * (pow(e, ord2) - pow(r2, ord2)));
      } else {
        c = -(cp.v[1] * e + cp.v[2] * pow(e, ord2));
      }
      v = 1.0 / (1.0 - cp.v[0] * c);
      break;
    }
  }
  out[l] = v;
}

int launch_coef(const dsht_plan* p, int which, const double* params, int nparams, double* out,
                hipStream_t s) {
  CoefParams cp;
  for (int i = 0; i < 8; i++) cp.v[i] = (params && i < nparams) ? params[i] : 0.0;
  // default numerical-diffusion coefficient, LPhys_DIFF_spm_mod.f90:629-634:
  // NumDiffCoefH = (N(N+1)/R^2)^(-ord/2) / tau when the caller passes numdiff = -1 and tau
  auto dflt = [&](int inum, int iord, int itau) {
    if (cp.v[inum] == -1.0 && nparams > itau) {
      int ord2 = (int)cp.v[iord] / 2;
      double e = (double)p->N * (p->N + 1) / (p->R * p->R);
      cp.v[inum] = pow(e, -ord2) / cp.v[itau];
    }
  };
  if (which == DSHT_COEF_HDIFF || which == DSHT_COEF_HVISC) dflt(1, 2, 3);
  if (which == DSHT_COEF_BEDIFF) dflt(2, 3, 5);
  hipLaunchKernelGGL(k_coef, dim3(cdiv(p->L, 256)), dim3(256), 0, s, which, p->N, p->L, p->R, cp, out);
  DSHT_HIP(hipGetLastError());
  return 0;
}

__global__ void k_transpose(const double* __restrict__ in, long ld_in, double* __restrict__ out, long ld_out,
                            int rows, int cols) {
  __shared__ double t[32][33];
  int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    int r = by + k, c = bx + threadIdx.x;
    if (r < rows && c < cols) t[k][threadIdx.x] = in[(long)r * ld_in + c];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    int c = bx + k, r = by + threadIdx.x;
    if (r < rows && c < cols) out[(long)c * ld_out + r] = t[threadIdx.x][k];
  }
}

int launch_transpose(const double* in, long ld_in, double* out, long ld_out, int rows, int cols,
                     hipStream_t s) {
  if (rows <= 0 || cols <= 0) return 0;
  hipLaunchKernelGGL(k_transpose, dim3(cdiv(cols, 32), cdiv(rows, 32)), dim3(32, 8), 0, s, in, ld_in, out,
                     ld_out, rows, cols);
  DSHT_HIP(hipGetLastError());
  return 0;
}

}  // namespace dsht
