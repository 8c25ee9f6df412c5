/*
 * dsht_oracle.c -- FP64 CPU restatement of dogcm's spherical-harmonic transform path.
 *
 * TEST INFRASTRUCTURE ONLY (parity oracle and timed CPU baseline; see dsht_oracle.h).
 *
 * The arithmetic of the reference path lives in SPMODEL >= 0.8.0 (wa_module) and
 * ISPACK 1.0.4 (SNPACK/SJPACK Legendre transforms, FTPACK FFT), neither of which is
 * vendored in the reference tree (doc/INSTALL.rd:51-54).  This file restates the
 * published mathematics the reference documents:
 *   - grid, transforms, derivative transforms: doc/description/ogcm/space_discretization.tex:17-384
 *   - SpmlUtil wrappers (R and cos(phi) factors): model/dogcm/src/util/math/SpmlUtil_mod.f90:297-473
 *   - m = 0 Legendre recurrence and normalisation: model/dogcm/src/physics/lphys/LPhys_DIFF_spm_mod.f90:497-507
 *   - diffusion coefficients: LPhys_DIFF_spm_mod.f90:140-167, 629-634
 *   - advection filter: model/dogcm/src/dynamics/hspm_vfvm/HBEBaroc_hspm_vfvm_mod.f90:83-95
 *   - Helmholtz coefficient: model/dogcm/src/dynamics/hspm_vfvm/HBEBarot_hspm_vfvm_mod.f90:286-305
 * Legendre sums use the even/odd split P_n^m(-mu) = (-1)^{n-m} P_n^m(mu), the same
 * algorithm class as ISPACK; the tests pin it against an unsplit brute-force sum.
 */
#include "dsht_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#undef I /* complex.h defines I; this file spells the unit _Complex_I */
typedef double complex cplx;

struct orc {
  int N, I, J, L, M; /* M = I/2 */
  double R;
  double *mu, *wq, *wl, *lat, *lon, *coslat;
  double *P, *H; /* blocks per m: [n-m][j] */
  long *offP, *offH;
  int nrad, rad[40];
  cplx* tw;  /* exp(-2 pi i t / M), t = 0..M-1 */
  cplx* twI; /* exp(-2 pi i m / I), m = 0..M   */
};

static __thread char g_err[256];
static void set_err(const char* s) { snprintf(g_err, sizeof g_err, "%s", s); }
const char* orc_last_error(void) { return g_err; }

int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

static inline int lidx(int n, int m) { return n * (n + 1) + m; }

/* ---------------------------------------------------------------- Gauss nodes */
/* Roots of P_J by Newton iteration in extended precision; weights 2/((1-x^2) P_J'^2).
   Ordered -1 < mu_1 < ... < mu_J < 1 (space_discretization.tex:33-38). */
static void gauss_nodes(int J, double* mu, double* wl) {
  for (int k = 0; k < J / 2; k++) {
    long double x = cosl(3.14159265358979323846264338327950288L * (k + 0.75L) / (J + 0.5L));
    long double p0 = 0, p1 = 0, dp = 0;
    for (int it = 0; it < 100; it++) {
      p0 = 1.0L;
      p1 = x;
      for (int n = 2; n <= J; n++) {
        long double p2 = ((2 * n - 1) * x * p1 - (n - 1) * p0) / n;
        p0 = p1;
        p1 = p2;
      }
      dp = J * (x * p1 - p0) / (x * x - 1.0L);
      long double dx = p1 / dp;
      x -= dx;
      if (fabsl(dx) < 1e-19L) break;
    }
    p0 = 1.0L;
    p1 = x;
    for (int n = 2; n <= J; n++) {
      long double p2 = ((2 * n - 1) * x * p1 - (n - 1) * p0) / n;
      p0 = p1;
      p1 = p2;
    }
    dp = J * (x * p1 - p0) / (x * x - 1.0L);
    long double w = 2.0L / ((1.0L - x * x) * dp * dp);
    mu[J - 1 - k] = (double)x;
    mu[k] = -(double)x;
    wl[J - 1 - k] = (double)w;
    wl[k] = (double)w;
  }
}

/* ------------------------------------------------------ associated Legendre */
/* Normalised P_n^m (int P^2 dmu = 2, positive sectoral) for n = m..N+1 and
   H_n^m = (1-mu^2) dP/dmu = -n eps_{n+1} P_{n+1} + (n+1) eps_n P_{n-1}
   (space_discretization.tex:236-245, with eps_{m,n} = sqrt((n^2-m^2)/(4n^2-1))).
   Extended precision keeps sin^m(theta) representable at T170 polar nodes. */
static inline long double epsl_(int n, int m) {
  return sqrtl((long double)(n * n - m * m) / (long double)(4 * n * n - 1));
}

static void legendre_column(int N, int m, long double mu, long double* p /* N+2-m */) {
  long double s2 = 1.0L - mu * mu;
  long double s = sqrtl(s2);
  long double pmm = 1.0L;
  for (int k = 1; k <= m; k++) pmm *= sqrtl((2.0L * k + 1.0L) / (2.0L * k)) * s;
  p[0] = pmm;
  if (m + 1 <= N + 1) p[1] = sqrtl(2.0L * m + 3.0L) * mu * pmm;
  for (int n = m + 2; n <= N + 1; n++)
    p[n - m] = (mu * p[n - 1 - m] - epsl_(n - 1, m) * p[n - 2 - m]) / epsl_(n, m);
}

static void build_tables(orc_t* h) {
  int N = h->N, J = h->J;
  long tp = 0, th = 0;
  for (int m = 0; m <= N; m++) {
    h->offP[m] = tp;
    h->offH[m] = th;
    tp += (long)(N + 2 - m) * J;
    th += (long)(N + 1 - m) * J;
  }
  h->P = (double*)malloc(sizeof(double) * tp);
  h->H = (double*)malloc(sizeof(double) * th);
#pragma omp parallel
  {
    long double* p = (long double*)malloc(sizeof(long double) * (N + 3));
#pragma omp for schedule(dynamic)
    for (int m = 0; m <= N; m++) {
      for (int j = 0; j < J; j++) {
        legendre_column(N, m, (long double)h->mu[j], p);
        for (int n = m; n <= N + 1; n++) h->P[h->offP[m] + (long)(n - m) * J + j] = (double)p[n - m];
        for (int n = m; n <= N; n++) {
          long double hv = -(long double)n * epsl_(n + 1, m) * p[n + 1 - m];
          if (n - 1 >= m) hv += (long double)(n + 1) * epsl_(n, m) * p[n - 1 - m];
          h->H[h->offH[m] + (long)(n - m) * J + j] = (double)hv;
        }
      }
    }
    free(p);
  }
}

/* ------------------------------------------------------------------- FFT */
/* Mixed-radix (4, 2, 3, 5) decimation-in-time complex DFT of length M with a
   master twiddle table; the real transforms of length I = 2M use the standard
   half-length packing z_k = x_{2k} + i x_{2k+1}. */
static int factorize(int M, int* rad) {
  int n = 0;
  while (M % 4 == 0) { rad[n++] = 4; M /= 4; }
  while (M % 2 == 0) { rad[n++] = 2; M /= 2; }
  while (M % 3 == 0) { rad[n++] = 3; M /= 3; }
  while (M % 5 == 0) { rad[n++] = 5; M /= 5; }
  return M == 1 ? n : -1;
}

/* forward (sign -1) via tw, inverse (sign +1) via conj(tw) */
static void dft_rec(const orc_t* h, int n, const cplx* x, int stride, cplx* out,
                    const int* rad, int tstep, int inv) {
  if (n == 1) { out[0] = x[0]; return; }
  int r = rad[0], m = n / r;
  for (int p = 0; p < r; p++)
    dft_rec(h, m, x + (long)p * stride, stride * r, out + (long)p * m, rad + 1, tstep * r, inv);
  cplx t[5], y[5];
  int Mtot = h->M;
  for (int k = 0; k < m; k++) {
    for (int p = 0; p < r; p++) {
      long e = ((long)p * k * tstep) % Mtot;
      cplx w = inv ? conj(h->tw[e]) : h->tw[e];
      t[p] = w * out[p * m + k];
    }
    for (int q = 0; q < r; q++) {
      cplx acc = 0;
      for (int p = 0; p < r; p++) {
        long e = ((long)p * q * m * tstep) % Mtot;
        cplx w = inv ? conj(h->tw[e]) : h->tw[e];
        acc += w * t[p];
      }
      y[q] = acc;
    }
    for (int q = 0; q < r; q++) out[q * m + k] = y[q];
  }
}

/* one real row x[I] -> G[m], m = 0..N, G^m = (1/I) sum_i x_i e^{-i m lambda_i} */
static void rfft_row(const orc_t* h, const double* x, cplx* G, cplx* z, cplx* Z) {
  int M = h->M, I = h->I;
  for (int k = 0; k < M; k++) z[k] = x[2 * k] + x[2 * k + 1] * _Complex_I;
  dft_rec(h, M, z, 1, Z, h->rad, 1, 0);
  for (int m = 0; m <= h->N; m++) {
    cplx a = Z[m % M], b = conj(Z[(M - m) % M]);
    cplx E = 0.5 * (a + b), O = -0.5 * _Complex_I * (a - b);
    G[m] = (E + h->twI[m] * O) / (double)I;
  }
}

/* G[m], m = 0..N (G^m = 0 above N) -> x_i = G^0 + 2 Re sum_{m>=1} G^m e^{i m lambda_i} */
static void irfft_row(const orc_t* h, const cplx* G, double* x, cplx* z, cplx* Z) {
  int M = h->M, N = h->N;
  for (int m = 0; m < M; m++) {
    /* X[m] and conj(X[M-m]); N < M so the Nyquist coefficient X[M] is zero */
    cplx Xm = (m <= N) ? G[m] : 0;
    int mm = M - m;
    cplx Xc = (m >= 1 && mm <= N) ? conj(G[mm]) : 0;
    cplx A = Xm + Xc;
    cplx B = (Xm - Xc) * conj(h->twI[m]);
    Z[m] = A + _Complex_I * B;
  }
  dft_rec(h, M, Z, 1, z, h->rad, 1, 1);
  for (int k = 0; k < M; k++) {
    x[2 * k] = creal(z[k]);
    x[2 * k + 1] = cimag(z[k]);
  }
}

void orc_fft_fwd(const orc_t* h, const double* x, double* G, int nrow) {
#pragma omp parallel
  {
    cplx* z = (cplx*)malloc(sizeof(cplx) * h->M * 2);
#pragma omp for
    for (int r = 0; r < nrow; r++) rfft_row(h, x + (long)r * h->I, (cplx*)G + (long)r * (h->N + 1), z, z + h->M);
    free(z);
  }
}

void orc_fft_inv(const orc_t* h, const double* G, double* x, int nrow) {
#pragma omp parallel
  {
    cplx* z = (cplx*)malloc(sizeof(cplx) * h->M * 2);
#pragma omp for
    for (int r = 0; r < nrow; r++) irfft_row(h, (const cplx*)G + (long)r * (h->N + 1), x + (long)r * h->I, z, z + h->M);
    free(z);
  }
}

/* ------------------------------------------------------------- create */
orc_t* orc_create(int nmax, int im, int jm, double radius) {
  if (nmax < 0 || im < 2 || jm < 2) { set_err("orc_create: nmax >= 0, im >= 2, jm >= 2 required"); return NULL; }
  if (im % 2) { set_err("orc_create: im must be even"); return NULL; }
  if (jm % 2) { set_err("orc_create: jm must be even"); return NULL; }
  if (im < 2 * nmax + 2) { set_err("orc_create: im must be >= 2*nmax+2"); return NULL; }
  if (jm < nmax + 1) { set_err("orc_create: jm must be >= nmax+1"); return NULL; }
  orc_t* h = (orc_t*)calloc(1, sizeof(orc_t));
  h->N = nmax; h->I = im; h->J = jm; h->L = (nmax + 1) * (nmax + 1); h->M = im / 2; h->R = radius;
  h->nrad = factorize(h->M, h->rad);
  if (h->nrad < 0) { set_err("orc_create: im/2 must factor into 2, 3, 5"); free(h); return NULL; }
  int J = jm, I = im;
  h->mu = malloc(sizeof(double) * J); h->wq = malloc(sizeof(double) * J);
  h->wl = malloc(sizeof(double) * J); h->lat = malloc(sizeof(double) * J);
  h->coslat = malloc(sizeof(double) * J); h->lon = malloc(sizeof(double) * I);
  gauss_nodes(J, h->mu, h->wl);
  for (int j = 0; j < J; j++) {
    h->wq[j] = 0.5 * h->wl[j];
    h->lat[j] = asin(h->mu[j]);
    h->coslat[j] = sqrt((1.0 - h->mu[j]) * (1.0 + h->mu[j]));
  }
  for (int i = 0; i < I; i++) h->lon[i] = 2.0 * M_PI * i / I;
  h->offP = malloc(sizeof(long) * (nmax + 1)); h->offH = malloc(sizeof(long) * (nmax + 1));
  build_tables(h);
  h->tw = malloc(sizeof(cplx) * h->M);
  for (int t = 0; t < h->M; t++) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * t / h->M;
    h->tw[t] = (double)cosl(a) + (double)sinl(a) * _Complex_I;
  }
  h->twI = malloc(sizeof(cplx) * (h->M + 1));
  for (int m = 0; m <= h->M; m++) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * m / I;
    h->twI[m] = (double)cosl(a) + (double)sinl(a) * _Complex_I;
  }
  return h;
}

void orc_destroy(orc_t* h) {
  if (!h) return;
  free(h->mu); free(h->wq); free(h->wl); free(h->lat); free(h->coslat); free(h->lon);
  free(h->P); free(h->H); free(h->offP); free(h->offH); free(h->tw); free(h->twI);
  free(h);
}

void orc_grid(const orc_t* h, double* mu, double* wq, double* lat, double* lon,
              double* y_lat_weight, double* x_lon_weight) {
  for (int j = 0; j < h->J; j++) {
    if (mu) mu[j] = h->mu[j];
    if (wq) wq[j] = h->wq[j];
    if (lat) lat[j] = h->lat[j];
    if (y_lat_weight) y_lat_weight[j] = h->wl[j];
  }
  for (int i = 0; i < h->I; i++) {
    if (lon) lon[i] = h->lon[i];
    if (x_lon_weight) x_lon_weight[i] = 2.0 * M_PI / h->I;
  }
}

int orc_legendre(const orc_t* h, int m, int n, double* p_out, double* h_out) {
  if (m < 0 || m > h->N || n < m || n > h->N + 1) { set_err("orc_legendre: bad (n,m)"); return -1; }
  for (int j = 0; j < h->J; j++) {
    if (p_out) p_out[j] = h->P[h->offP[m] + (long)(n - m) * h->J + j];
    if (h_out) h_out[j] = n <= h->N ? h->H[h->offH[m] + (long)(n - m) * h->J + j] : 0.0;
  }
  return 0;
}

/* ------------------------------------------------------------ transforms */
/* forward, one column.  VAL: s = sum_j w_j G P;  DIVLAMBDA: s = i m sum_j w_j/(1-mu^2) G P;
   DIVMU: s = -sum_j w_j/(1-mu^2) G H (integration by parts in mu, tex:229-268). */
static void fwd_column(const orc_t* h, int kind, const double* x, double* w, cplx* G, cplx* work) {
  int N = h->N, J = h->J, I = h->I, J2 = J / 2;
  cplx* z = work;              /* M */
  cplx* Z = work + h->M;       /* M */
  double* gsr = (double*)(work + 2 * h->M); /* J2 each: sym re/im, anti re/im, weights */
  double* gsi = gsr + J2;
  double* gar = gsi + J2;
  double* gai = gar + J2;
  double* cw = gai + J2;
  for (int j = 0; j < J; j++) rfft_row(h, x + (long)j * I, G + (long)j * (N + 1), z, Z);
  for (int q = 0; q < J2; q++) {
    int jn = J2 + q;
    cw[q] = (kind == ORC_VAL) ? h->wq[jn] : h->wq[jn] / ((1.0 - h->mu[jn]) * (1.0 + h->mu[jn]));
  }
  for (int m = 0; m <= N; m++) {
    for (int q = 0; q < J2; q++) {
      int jn = J2 + q, js = J2 - 1 - q;
      cplx a = G[(long)jn * (N + 1) + m], b = G[(long)js * (N + 1) + m];
      gsr[q] = cw[q] * (creal(a) + creal(b));
      gsi[q] = cw[q] * (cimag(a) + cimag(b));
      gar[q] = cw[q] * (creal(a) - creal(b));
      gai[q] = cw[q] * (cimag(a) - cimag(b));
    }
    const double* T = (kind == ORC_DIVMU) ? h->H + h->offH[m] : h->P + h->offP[m];
    for (int n = m; n <= N; n++) {
      const double* t = T + (long)(n - m) * J + J2;
      int odd = (n - m) & 1;
      /* P has parity (-1)^{n-m}; H has the opposite parity */
      int use_sym = (kind == ORC_DIVMU) ? odd : !odd;
      const double* gr = use_sym ? gsr : gar;
      const double* gi = use_sym ? gsi : gai;
      double ar = 0, ai = 0;
      for (int q = 0; q < J2; q++) { ar += t[q] * gr[q]; ai += t[q] * gi[q]; }
      cplx s = ar + ai * _Complex_I;
      if (kind == ORC_DIVLAMBDA) s *= (double)m * _Complex_I;
      if (kind == ORC_DIVMU) s = -s;
      w[lidx(n, m)] = creal(s);
      if (m > 0) w[lidx(n, -m)] = cimag(s);
    }
  }
}

/* inverse, one column.  VAL/GRADLAMBDA use P (GRADLAMBDA times i m), GRADMU uses H;
   GRADLON/GRADLAT divide by cos(phi) (SPMODEL xy_GradLon_w / xy_GradLat_w). */
static void inv_column(const orc_t* h, int kind, const double* w, double* x, cplx* G, cplx* work) {
  int N = h->N, J = h->J, I = h->I, J2 = J / 2;
  cplx* z = work;
  cplx* Z = work + h->M;
  double* Er = (double*)(work + 2 * h->M); /* J2 each */
  double* Ei = Er + J2;
  double* Or = Ei + J2;
  double* Oi = Or + J2;
  int useH = (kind == ORC_GRADMU || kind == ORC_GRADLAT);
  int useIm = (kind == ORC_GRADLAMBDA || kind == ORC_GRADLON);
  for (int m = 0; m <= N; m++) {
    for (int q = 0; q < J2; q++) Er[q] = Ei[q] = Or[q] = Oi[q] = 0;
    const double* T = useH ? h->H + h->offH[m] : h->P + h->offP[m];
    /* four wavenumbers per pass (two of each parity) to halve accumulator traffic */
    int n = m;
    for (; n + 3 <= N; n += 4) {
      double s0r = w[lidx(n, m)], s1r = w[lidx(n + 1, m)], s2r = w[lidx(n + 2, m)], s3r = w[lidx(n + 3, m)];
      double s0i = m > 0 ? w[lidx(n, -m)] : 0.0, s1i = m > 0 ? w[lidx(n + 1, -m)] : 0.0;
      double s2i = m > 0 ? w[lidx(n + 2, -m)] : 0.0, s3i = m > 0 ? w[lidx(n + 3, -m)] : 0.0;
      const double* t0 = T + (long)(n - m) * J + J2;
      const double* t1 = t0 + J;
      const double* t2 = t1 + J;
      const double* t3 = t2 + J;
      double* ar = ((n - m) & 1) ? Or : Er;
      double* ai = ((n - m) & 1) ? Oi : Ei;
      double* br = ((n - m) & 1) ? Er : Or;
      double* bi = ((n - m) & 1) ? Ei : Oi;
      for (int q = 0; q < J2; q++) {
        ar[q] += t0[q] * s0r + t2[q] * s2r;
        ai[q] += t0[q] * s0i + t2[q] * s2i;
        br[q] += t1[q] * s1r + t3[q] * s3r;
        bi[q] += t1[q] * s1i + t3[q] * s3i;
      }
    }
    for (; n <= N; n++) {
      double sr = w[lidx(n, m)], si = m > 0 ? w[lidx(n, -m)] : 0.0;
      const double* t = T + (long)(n - m) * J + J2;
      double* ar = ((n - m) & 1) ? Or : Er;
      double* ai = ((n - m) & 1) ? Oi : Ei;
      for (int q = 0; q < J2; q++) { ar[q] += t[q] * sr; ai[q] += t[q] * si; }
    }
    cplx f = useIm ? (double)m * _Complex_I : 1.0;
    for (int q = 0; q < J2; q++) {
      int jn = J2 + q, js = J2 - 1 - q;
      cplx E = Er[q] + Ei[q] * _Complex_I, O = Or[q] + Oi[q] * _Complex_I;
      G[(long)jn * (N + 1) + m] = f * (E + O);
      G[(long)js * (N + 1) + m] = useH ? f * (O - E) : f * (E - O);
    }
  }
  for (int j = 0; j < J; j++) {
    irfft_row(h, G + (long)j * (N + 1), x + (long)j * I, z, Z);
    if (kind == ORC_GRADLON || kind == ORC_GRADLAT)
      for (int i = 0; i < I; i++) x[(long)j * I + i] /= h->coslat[j];
  }
}

static size_t work_elems(const orc_t* h) { return (size_t)2 * h->M + (size_t)3 * h->J + 8; } /* >= 2M cplx + 5 J/2 doubles */

int orc_fwd(const orc_t* h, int kind, const double* xy, double* w, int ncols) {
  if (kind != ORC_VAL && kind != ORC_DIVLAMBDA && kind != ORC_DIVMU) { set_err("orc_fwd: bad kind"); return -1; }
  long IJ = (long)h->I * h->J;
#pragma omp parallel
  {
    cplx* G = (cplx*)malloc(sizeof(cplx) * (size_t)h->J * (h->N + 1));
    cplx* work = (cplx*)malloc(sizeof(cplx) * work_elems(h));
#pragma omp for schedule(dynamic)
    for (int c = 0; c < ncols; c++) fwd_column(h, kind, xy + c * IJ, w + (long)c * h->L, G, work);
    free(G); free(work);
  }
  return 0;
}

int orc_inv(const orc_t* h, int kind, const double* w, double* xy, int ncols) {
  if (kind != ORC_VAL && (kind < ORC_GRADLAMBDA || kind > ORC_GRADLAT)) { set_err("orc_inv: bad kind"); return -1; }
  long IJ = (long)h->I * h->J;
#pragma omp parallel
  {
    cplx* G = (cplx*)malloc(sizeof(cplx) * (size_t)h->J * (h->N + 1));
    cplx* work = (cplx*)malloc(sizeof(cplx) * work_elems(h));
#pragma omp for schedule(dynamic)
    for (int c = 0; c < ncols; c++) inv_column(h, kind, w + (long)c * h->L, xy + c * IJ, G, work);
    free(G); free(work);
  }
  return 0;
}

/* ---------------------------------------------- SpmlUtil compositions */
/* w_AlphaOptr_xy = (w_DivLambda_xy(A) + w_DivMu_xy(B)) / R   (SpmlUtil_mod.f90:398-412) */
int orc_alpha_fwd(const orc_t* h, const double* a, const double* b, double* w, int ncols) {
  long n = (long)h->L * ncols;
  double* t = (double*)malloc(sizeof(double) * n);
  orc_fwd(h, ORC_DIVLAMBDA, a, w, ncols);
  orc_fwd(h, ORC_DIVMU, b, t, ncols);
  for (long i = 0; i < n; i++) w[i] = (w[i] + t[i]) / h->R;
  free(t);
  return 0;
}

/* xy_AlphaOptr_w = (xy_GradLambda_w(A) + xy_GradMu_w(B)) / (R cos^2 phi)  (SpmlUtil_mod.f90:414-428) */
int orc_alpha_inv(const orc_t* h, const double* wa, const double* wb, double* xy, int ncols) {
  long IJ = (long)h->I * h->J, n = IJ * ncols;
  double* t = (double*)malloc(sizeof(double) * n);
  orc_inv(h, ORC_GRADLAMBDA, wa, xy, ncols);
  orc_inv(h, ORC_GRADMU, wb, t, ncols);
  for (long i = 0; i < n; i++) {
    int j = (int)((i % IJ) / h->I);
    double c2 = (1.0 - h->mu[j]) * (1.0 + h->mu[j]);
    xy[i] = (xy[i] + t[i]) / (h->R * c2);
  }
  free(t);
  return 0;
}

/* calc_UVCosLat2VorDiv: vor = H(V cos, -U cos), div = H(U cos, V cos)  (SpmlUtil_mod.f90:299-314) */
int orc_uv2vordiv(const orc_t* h, const double* ucos, const double* vcos, double* vor,
                  double* div, int ncols) {
  long n = (long)h->I * h->J * ncols;
  double* mu_ = (double*)malloc(sizeof(double) * n);
  for (long i = 0; i < n; i++) mu_[i] = -ucos[i];
  orc_alpha_fwd(h, vcos, mu_, vor, ncols);
  orc_alpha_fwd(h, ucos, vcos, div, ncols);
  free(mu_);
  return 0;
}

void orc_lapla(const orc_t* h, const double* w, double* out, int ncols) {
  for (int c = 0; c < ncols; c++)
    for (int n = 0; n <= h->N; n++)
      for (int m = -n; m <= n; m++) {
        long k = (long)c * h->L + lidx(n, m);
        out[k] = -(double)n * (n + 1) * w[k];
      }
}

void orc_laplainv(const orc_t* h, const double* w, double* out, int ncols) {
  for (int c = 0; c < ncols; c++)
    for (int n = 0; n <= h->N; n++)
      for (int m = -n; m <= n; m++) {
        long k = (long)c * h->L + lidx(n, m);
        out[k] = n == 0 ? 0.0 : -w[k] / ((double)n * (n + 1));
      }
}

/* calc_VorDiv2UV: psi = R^2 LaplaInv(vor), chi = R^2 LaplaInv(div),
   U = cos(phi) H^(chi, -psi), V = cos(phi) H^(psi, chi)  (SpmlUtil_mod.f90:316-340) */
int orc_vordiv2uv(const orc_t* h, const double* vor, const double* div, double* u, double* v,
                  int ncols) {
  long nl = (long)h->L * ncols, IJ = (long)h->I * h->J;
  double* psi = malloc(sizeof(double) * nl);
  double* chi = malloc(sizeof(double) * nl);
  double* mpsi = malloc(sizeof(double) * nl);
  orc_laplainv(h, vor, psi, ncols);
  orc_laplainv(h, div, chi, ncols);
  double R2 = h->R * h->R;
  for (long i = 0; i < nl; i++) { psi[i] *= R2; chi[i] *= R2; mpsi[i] = -psi[i]; }
  orc_alpha_inv(h, chi, mpsi, u, ncols);
  orc_alpha_inv(h, psi, chi, v, ncols);
  for (long i = 0; i < IJ * ncols; i++) {
    int j = (int)((i % IJ) / h->I);
    u[i] *= h->coslat[j];
    v[i] *= h->coslat[j];
  }
  free(psi); free(chi); free(mpsi);
  return 0;
}

/* ------------------------------------------------------------ integrals */
/* IntLonLat_xy = sum_ij xy * (2 pi / I) * y_Lat_Weight_j (y_Lat_Weight sums to 2) */
void orc_intlonlat(const orc_t* h, const double* xy, double* out, int ncols) {
  long IJ = (long)h->I * h->J;
  for (int c = 0; c < ncols; c++) {
    long double acc = 0;
    for (int j = 0; j < h->J; j++) {
      long double r = 0;
      for (int i = 0; i < h->I; i++) r += xy[c * IJ + (long)j * h->I + i];
      acc += r * h->wl[j];
    }
    out[c] = (double)(acc * (2.0L * 3.14159265358979323846264338327950288L / h->I));
  }
}

void orc_avrlonlat(const orc_t* h, const double* xy, double* out, int ncols) {
  long IJ = (long)h->I * h->J;
  for (int c = 0; c < ncols; c++) {
    long double acc = 0;
    for (int j = 0; j < h->J; j++) {
      long double r = 0;
      for (int i = 0; i < h->I; i++) r += xy[c * IJ + (long)j * h->I + i];
      acc += r * h->wl[j];
    }
    out[c] = (double)(acc / (2.0L * h->I));
  }
}

void orc_ya_intlon(const orc_t* h, const double* xy, double* out, int ncols) {
  for (long r = 0; r < (long)ncols * h->J; r++) {
    long double acc = 0;
    for (int i = 0; i < h->I; i++) acc += xy[r * h->I + i];
    out[r] = (double)(acc * (2.0L * 3.14159265358979323846264338327950288L / h->I));
  }
}

void orc_ya_avrlon(const orc_t* h, const double* xy, double* out, int ncols) {
  for (long r = 0; r < (long)ncols * h->J; r++) {
    long double acc = 0;
    for (int i = 0; i < h->I; i++) acc += xy[r * h->I + i];
    out[r] = (double)(acc / h->I);
  }
}

/* Interpolate_w: q(lon, lat) = sum s_n^m P_n^m(sin lat) e^{i m lon} */
void orc_interpolate(const orc_t* h, const double* w, const double* lon, const double* lat,
                     double* out, int npts) {
  int N = h->N;
  long double* p = malloc(sizeof(long double) * (N + 3));
  for (int k = 0; k < npts; k++) {
    long double mu = sinl((long double)lat[k]);
    long double acc = 0;
    for (int m = 0; m <= N; m++) {
      legendre_column(N, m, mu, p);
      long double gr = 0, gi = 0;
      for (int n = m; n <= N; n++) {
        gr += p[n - m] * w[lidx(n, m)];
        if (m > 0) gi += p[n - m] * w[lidx(n, -m)];
      }
      if (m == 0) acc += gr;
      else acc += 2.0L * (gr * cosl((long double)m * lon[k]) - gi * sinl((long double)m * lon[k]));
    }
    out[k] = (double)acc;
  }
  free(p);
}

/* ------------------------------------------------------ diagonal tables */
/* w_HDiffCoefH(l) = -( kappa (-lambda_l) + nu_num (-lambda_l)^(ord/2) ), lambda_l = rn(l,1)/R^2
   (LPhys_DIFF_spm_mod.f90:144-151); ord/2 is Fortran integer division. */
void orc_coef_hdiff(const orc_t* h, double kappa, double numdiff, int ord, double* c) {
  int p = ord / 2;
  for (int n = 0; n <= h->N; n++) {
    double e = (double)n * (n + 1) / (h->R * h->R);
    double v = -(kappa * e + numdiff * pow(e, p));
    for (int m = -n; m <= n; m++) c[lidx(n, m)] = v;
  }
}

/* w_HViscCoefH(l) = -( nu (-lambda_l - 2/R^2) + nu_num ((-lambda_l)^(ord/2) - (2/R^2)^(ord/2)) )
   (LPhys_DIFF_spm_mod.f90:150-153) */
void orc_coef_hvisc(const orc_t* h, double nu, double numdiff, int ord, double* c) {
  int p = ord / 2;
  double r2 = 2.0 / (h->R * h->R);
  for (int n = 0; n <= h->N; n++) {
    double e = (double)n * (n + 1) / (h->R * h->R);
    double v = -(nu * (e - r2) + numdiff * (pow(e, p) - pow(r2, p)));
    for (int m = -n; m <= n; m++) c[lidx(n, m)] = v;
  }
}

/* w_AdvFilter: 1 for n < Nc, exp(-((n-Nc)/(N-Nc))^4) otherwise; Nc = int(0.66 N)
   (HBEBaroc_hspm_vfvm_mod.f90:83-95; Fortran real->integer assignment truncates) */
void orc_coef_advfilter(const orc_t* h, double* f) {
  int N = h->N;
  int Nc = (int)(0.66 * N);
  for (int n = 0; n <= N; n++) {
    double v = 1.0;
    if (n >= Nc) {
      double r = (double)(n - Nc) / (double)(N - Nc);
      v = exp(-1.0 * r * r * r * r);
    }
    for (int m = -n; m <= n; m++) f[lidx(n, m)] = v;
  }
}

/* w_Coef(1) = c0 ; w_Coef(l) = 1/(n(n+1)/R^2 + diag), l >= 2  (HBEBarot_hspm_vfvm_mod.f90:286-299) */
void orc_coef_helmholtz(const orc_t* h, double diag, double c0, double* c) {
  c[0] = c0;
  for (int n = 1; n <= h->N; n++) {
    double v = 1.0 / ((double)n * (n + 1) / (h->R * h->R) + diag);
    for (int m = -n; m <= n; m++) c[lidx(n, m)] = v;
  }
}

void orc_diag(const orc_t* h, const double* d, double* w, int ncols) {
  for (int c = 0; c < ncols; c++)
    for (int l = 0; l < h->L; l++) w[(long)c * h->L + l] *= d[l];
}

/* ---------------------------------------------------------------- vertical column diagnostics
 * HBEDiagnose_OMG, _HydPres, _UVBarot (model/dogcm/src/dynamics/hspm_vfvm/
 * HBEDiagnose_hspm_vfvm_mod.f90:118-137, 312-322, 244-252): per grid point, the reference's
 * loop over layers with its products and sums in its order.  fp-contract is off so that no
 * product is fused into an add (the Fortran statements round every product). */
#define ORC_NOFMA __attribute__((optimize("fp-contract=off")))

ORC_NOFMA void orc_col_omega(const double* div, const double* h, int km, const double* cdk, double* omg,
                             long npts) {
  for (long i = 0; i < npts; i++) {
    double r = 0.0;
    omg[(long)km * npts + i] = r; /* r_OMG(KE) = 0 */
    for (int k = km - 1; k >= 0; k--) {
      const double dz = h[(long)k * npts + i] * cdk[k];   /* z_DZ = H * CDK */
      const double dom = div[(long)k * npts + i] * dz;    /* z_DOMG = Div * z_DZ */
      r = r - dom;                                        /* r_OMG(k) = r_OMG(k+1) - z_DOMG(k+1) */
      omg[(long)k * npts + i] = r;
    }
  }
}

ORC_NOFMA void orc_col_hydpres(const double* dens, const double* h, int km, const double* cdk, double grav,
                               double* p, long npts) {
  for (long i = 0; i < npts; i++) {
    double r = 0.0; /* r_HydPres(KS-1) = 0 */
    for (int k = 0; k < km; k++) {
      const double d = ((grav * h[(long)k * npts + i]) * dens[(long)k * npts + i]) * cdk[k];
      const double rn = r + d;
      p[(long)k * npts + i] = 0.5 * (r + rn);
      r = rn;
    }
  }
}

ORC_NOFMA void orc_col_uvbarot(const double* u, const double* v, const double* h, int km, const double* cdk,
                               double* ub, double* vb, long npts) {
  for (long i = 0; i < npts; i++) {
    double dep = 0.0, su = 0.0, sv = 0.0;
    for (int k = 0; k < km; k++) {
      const double f = h[(long)k * npts + i] * cdk[k]; /* z_Fac */
      dep = dep + f;
      su = su + u[(long)k * npts + i] * f;
      sv = sv + v[(long)k * npts + i] * f;
    }
    ub[i] = su / dep;
    vb[i] = sv / dep;
  }
}

/* calc_ADVFlxZ_QUICK, HBEBaroc_hspm_vfvm_mod.f90:233-317, in the reference's structure: the
 * interpolation weights per half level (:263-267), the curvature array of the interior layers
 * (:271-279), the interior fluxes (:283-292), then the boundary half levels (:296-302,
 * 307-312).  0-based: layer l = k - KS, half level h = k - (KS-1). */
ORC_NOFMA void orc_col_quick(const double* omg, const double* trc, int km, const double* cdk,
                             const double* rfdk, int linfree, double* flx, long npts) {
  double* m1 = calloc((size_t)km + 1, sizeof(double));
  double* m2 = calloc((size_t)km + 1, sizeof(double));
  double* dfac = calloc((size_t)km + 1, sizeof(double));
  double* diff = calloc((size_t)km, sizeof(double));
  for (int h = 1; h <= km - 1; h++) { /* k = KS..KE-1 */
    m1[h] = cdk[h] / (cdk[h - 1] + cdk[h]);
    m2[h] = cdk[h - 1] / (cdk[h - 1] + cdk[h]);
    dfac[h] = 0.125 * cdk[h] * cdk[h - 1];
  }
  for (long i = 0; i < npts; i++) {
#define TR(l) trc[(long)(l) * npts + i]
#define OM(h) omg[(long)(h) * npts + i]
    for (int l = 1; l <= km - 2; l++) /* k = KS+1..KE-1 */
      diff[l] = ((TR(l - 1) - TR(l)) * rfdk[l] - (TR(l) - TR(l + 1)) * rfdk[l + 1]) * (rfdk[l] + rfdk[l + 1]);
    for (int h = 2; h <= km - 2; h++) /* k = KS+1..KE-2 */
      flx[(long)h * npts + i] = OM(h) * ((m1[h] * TR(h - 1) + m2[h] * TR(h)) - dfac[h] * (diff[h - 1] + diff[h])) +
                                fabs(OM(h)) * dfac[h] * (diff[h - 1] - diff[h]);
    flx[i] = 0.0;
    flx[(long)km * npts + i] = 0.0;
    flx[npts + i] = OM(1) * (m1[1] * TR(0) + m2[1] * TR(1));
    flx[(long)(km - 1) * npts + i] = OM(km - 1) * (m1[km - 1] * TR(km - 2) + m2[km - 1] * TR(km - 1));
    if (linfree) flx[i] = OM(0) * TR(0);
#undef TR
#undef OM
  }
  free(m1);
  free(m2);
  free(dfac);
  free(diff);
}

/* LAPACK DGTSV (LAPACK 3.x, SRC/dgtsv.f), the NRHS = 1 path: Gaussian elimination with
 * partial pivoting by row interchange (the interchange fills the second superdiagonal, kept
 * in dl), then back substitution.  0-based restatement of the published routine. */
ORC_NOFMA int orc_dgtsv(int n, double* dl, double* d, double* du, double* b) {
  int swaps = 0;
  for (int i = 0; i <= n - 3; i++) {
    if (fabs(d[i]) >= fabs(dl[i])) {
      if (d[i] == 0.0) return -(i + 1);
      const double fact = dl[i] / d[i];
      d[i + 1] = d[i + 1] - fact * du[i];
      b[i + 1] = b[i + 1] - fact * b[i];
      dl[i] = 0.0;
    } else {
      const double fact = d[i] / dl[i];
      d[i] = dl[i];
      double temp = d[i + 1];
      d[i + 1] = du[i] - fact * temp;
      dl[i] = du[i + 1];
      du[i + 1] = -fact * dl[i];
      du[i] = temp;
      temp = b[i];
      b[i] = b[i + 1];
      b[i + 1] = temp - fact * b[i + 1];
      swaps++;
    }
  }
  if (n > 1) {
    const int i = n - 2;
    if (fabs(d[i]) >= fabs(dl[i])) {
      if (d[i] == 0.0) return -(i + 1);
      const double fact = dl[i] / d[i];
      d[i + 1] = d[i + 1] - fact * du[i];
      b[i + 1] = b[i + 1] - fact * b[i];
    } else {
      const double fact = d[i] / dl[i];
      d[i] = dl[i];
      const double temp = d[i + 1];
      d[i + 1] = du[i] - fact * temp;
      du[i] = temp;
      const double tb = b[i];
      b[i] = b[i + 1];
      b[i + 1] = tb - fact * b[i + 1];
      swaps++;
    }
  }
  if (d[n - 1] == 0.0) return -n;
  b[n - 1] = b[n - 1] / d[n - 1];
  if (n > 1) b[n - 2] = (b[n - 2] - du[n - 2] * b[n - 1]) / d[n - 2];
  for (int i = n - 3; i >= 0; i--) b[i] = (b[i] - du[i] * b[i + 1] - dl[i] * b[i + 2]) / d[i];
  return swaps;
}

/* calc_VDiffTRCEq, DOGCM_Phys_hspm_vfvm_mod.f90:800-858, per grid point in the reference's
 * order: coef (:809), half-level coefficients r (:811-823), interior rows (:825-830),
 * right-hand side (:831), boundary rows (:836-852), DGTSV (:855), q0 + x (:857).  Layer k
 * (0-based) is the reference's KS+k; half level h is KS-1+h. */
static long col_vdiff_any(const double* q0, const double* rhs, const double* kv, const double* e3, int km,
                          const double* cdk, const double* fdk, double dt, double alpha, int sfc_bc, int btm_bc,
                          const double* bfric, double* qa, long npts);

long orc_col_vdiff(const double* q0, const double* rhs, const double* kv, const double* e3, int km,
                   const double* cdk, const double* fdk, double dt, double alpha, int sfc_bc, int btm_bc,
                   double* qa, long npts) {
  return col_vdiff_any(q0, rhs, kv, e3, km, cdk, fdk, dt, alpha, sfc_bc, btm_bc, NULL, qa, npts);
}

/* calc_VDiffMOMEq (DOGCM_Phys_hspm_vfvm_mod.f90:865-975): the same system with b = dt*rhs (:944)
 * and the Dirichlet bottom row 1 + alpha*coef*(xy_BtmFrictCoef + r(KE-1)) (:963). */
long orc_col_vdiff_mom(const double* q0, const double* rhs, const double* kv, const double* e3, int km,
                       const double* cdk, const double* fdk, double dt, double alpha, int sfc_bc, int btm_bc,
                       const double* bfric, double* qa, long npts) {
  return col_vdiff_any(q0, rhs, kv, e3, km, cdk, fdk, dt, alpha, sfc_bc, btm_bc, bfric, qa, npts);
}

ORC_NOFMA static long col_vdiff_any(const double* q0, const double* rhs, const double* kv, const double* e3, int km,
                                    const double* cdk, const double* fdk, double dt, double alpha, int sfc_bc,
                                    int btm_bc, const double* bfric, double* qa, long npts) {
  double* coef = malloc(sizeof(double) * (size_t)km);
  double* r = malloc(sizeof(double) * (size_t)(km + 1));
  double* al = malloc(sizeof(double) * (size_t)km);
  double* ad = malloc(sizeof(double) * (size_t)km);
  double* au = malloc(sizeof(double) * (size_t)km);
  double* b = malloc(sizeof(double) * (size_t)km);
  long swaps = 0;
  for (long i = 0; i < npts; i++) {
#define KV(k) kv[(long)(k) * npts + i]
#define E3(k) e3[(long)(k) * npts + i]
    for (int k = 0; k < km; k++) coef[k] = dt / (E3(k) * cdk[k]);
    for (int k = 0; k < km - 1; k++) { /* reference k = KS..KE-1: half level k+1 */
      const double m1 = cdk[k + 1] / (cdk[k] + cdk[k + 1]);
      const double m2 = cdk[k] / (cdk[k] + cdk[k + 1]);
      r[k + 1] = (m1 * KV(k) + m2 * KV(k + 1)) / (fdk[k + 1] * (m1 * E3(k) + m2 * E3(k + 1)));
    }
    r[0] = KV(0) / (fdk[0] * E3(0));
    r[km] = KV(km - 1) / (fdk[km] * E3(km - 1));
    for (int k = 1; k <= km - 2; k++) { /* r(k-1) -> r[k], r(k) -> r[k+1] */
      al[k] = -(alpha * coef[k] * r[k]);
      ad[k] = 1.0 + alpha * coef[k] * (r[k] + r[k + 1]);
      au[k] = -(alpha * coef[k] * r[k + 1]);
    }
    for (int k = 0; k < km; k++) b[k] = bfric ? dt * rhs[(long)k * npts + i] : dt * rhs[(long)k * npts + i] / E3(k);
    if (sfc_bc == 'D') {
      ad[0] = 1.0 + alpha * coef[0] * (r[1] + 2.0 * r[0]);
      au[0] = -(alpha * coef[0] * r[1]);
    } else {
      ad[0] = 1.0 + alpha * coef[0] * (r[1]);
      au[0] = -(alpha * coef[0] * r[1]);
    }
    const int n = km - 1;
    if (btm_bc == 'D') {
      al[n] = -(alpha * coef[n] * r[n]);
      ad[n] = bfric ? 1.0 + alpha * coef[n] * (bfric[i] + r[n]) : 1.0 + alpha * coef[n] * (2.0 * r[n + 1] + r[n]);
    } else {
      al[n] = -(alpha * coef[n] * r[n]);
      ad[n] = 1.0 + alpha * coef[n] * (+r[n]);
    }
    /* DGTSV(N, 1, Al(KS+1:KE), Ad(KS:KE), Au(KS:KE-1), b, N, info) */
    const int sw = orc_dgtsv(km, al + 1, ad, au, b);
    if (sw > 0) swaps += sw;
    for (int k = 0; k < km; k++) qa[(long)k * npts + i] = q0[(long)k * npts + i] + b[k];
#undef KV
#undef E3
  }
  free(coef);
  free(r);
  free(al);
  free(ad);
  free(au);
  free(b);
  return swaps;
}
