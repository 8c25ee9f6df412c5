// fft.hip -- longitude real FFTs (the FTPACK part of ISPACK that SPMODEL's w_xy /
// xy_w call; doc/description/ogcm/space_discretization.tex:188-196 and 217-219).
//
// Forward:  G^m(mu_j) = (1/I) sum_i q_ij e^{-i m lambda_i},  m = 0..N
// Inverse:  q_ij = G^0 + 2 Re sum_{m=1}^{N} G^m e^{i m lambda_i}
//
// A real row of length I = 2M is packed as z_k = x_{2k} + i x_{2k+1} and run through a
// mixed-radix (4, 2, 3, 5) Stockham complex FFT of length M held in LDS (ping-pong
// buffers, one group of threads per row, several rows per workgroup).  Rows are
// enumerated r = j*ncols + c so that the Fourier output G[m][r] of consecutive rows is
// contiguous: a workgroup writes runs of rows_per_block complex values per m.
#include <cmath>

#include "dsht_internal.hpp"

namespace dsht {

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cmulc(double2 a, double2 b) {  // a * conj(b)
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// multiply by -i (SIGN = -1) or +i (SIGN = +1)
template <int SIGN>
__device__ inline double2 mul_i(double2 a) {
  return SIGN < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}

// in-register radix-R DFT, y_q = sum_p v_p exp(SIGN 2 pi i p q / R)
template <int SIGN>
__device__ inline void dft_small(double2* v, int R) {
  if (R == 2) {
    double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if (R == 4) {
    double2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    double2 t2 = cadd(v[1], v[3]), t3 = mul_i<SIGN>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[1] = cadd(t1, t3);
    v[2] = csub(t0, t2);
    v[3] = csub(t1, t3);
  } else if (R == 3) {
    const double s3 = 0.86602540378443864676;
    double2 t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    double2 a = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
    double2 b = mul_i<SIGN>(make_double2(s3 * t2.x, s3 * t2.y));
    v[0] = cadd(v[0], t1);
    v[1] = cadd(a, b);
    v[2] = csub(a, b);
  } else {  // R == 5
    const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
    const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
    double2 t1 = cadd(v[1], v[4]), t4 = csub(v[1], v[4]);
    double2 t2 = cadd(v[2], v[3]), t3 = csub(v[2], v[3]);
    double2 a1 = make_double2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
    double2 a2 = make_double2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
    double2 b1 = mul_i<SIGN>(make_double2(s1 * t4.x + s2 * t3.x, s1 * t4.y + s2 * t3.y));
    double2 b2 = mul_i<SIGN>(make_double2(s2 * t4.x - s1 * t3.x, s2 * t4.y - s1 * t3.y));
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
  }
}

// Stockham stages over all rows of the block; returns the buffer holding the result.
template <int SIGN>
__device__ double2* stockham(double2* src, double2* dst, int M, int nrad, const int* rad,
                             const double2* __restrict__ twM, int rpb, int tpr, int row, int tr) {
  int Ns = 1;
  for (int s = 0; s < nrad; s++) {
    const int R = rad[s];
    const int MR = M / R;
    const int tstride = M / (Ns * R);
    if (row < rpb) {
      double2* S = src + (long)row * M;
      double2* D = dst + (long)row * M;
      for (int jj = tr; jj < MR; jj += tpr) {
        int k = jj % Ns;
        double2 v[5];
        for (int r = 0; r < R; r++) {
          double2 x = S[jj + r * MR];
          if (r && k) {
            double2 w = twM[(k * r * tstride) % M];
            x = SIGN < 0 ? cmul(x, w) : cmulc(x, w);
          }
          v[r] = x;
        }
        dft_small<SIGN>(v, R);
        int d = (jj / Ns) * Ns * R + k;
        for (int r = 0; r < R; r++) D[d + r * Ns] = v[r];
      }
    }
    __syncthreads();
    double2* t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  return src;
}

__global__ __launch_bounds__(256) void k_fft_fwd(const double* __restrict__ xy, double2* __restrict__ G,
                                                 FftArgs a) {
  extern __shared__ double2 lds[];
  const int M = a.M, rpb = a.rows_per_block;
  const int tpr = blockDim.x / rpb;
  const int row = threadIdx.x / tpr, tr = threadIdx.x % tpr;
  const long nrows = (long)a.J * a.ncols;
  const long r0 = (long)blockIdx.x * rpb;
  double2* A = lds;
  double2* B = lds + (long)rpb * M;
  // load rows: z_k = x_{2k} + i x_{2k+1}
  if (row < rpb) {
    long r = r0 + row;
    if (r < nrows) {
      int j = (int)(r / a.ncols), c = (int)(r % a.ncols);
      const double2* x = reinterpret_cast<const double2*>(xy + c * a.ldg + (long)j * a.I);
      for (int k = tr; k < M; k += tpr) A[(long)row * M + k] = x[k];
    } else {
      for (int k = tr; k < M; k += tpr) A[(long)row * M + k] = make_double2(0.0, 0.0);
    }
  }
  __syncthreads();
  double2* Z = stockham<-1>(A, B, M, a.nrad, a.rad, a.twM, rpb, tpr, row, tr);
  // real post-processing and coalesced store: idx -> (m, row) with row fastest
  const double inv_i = 1.0 / a.I;
  const int N = a.N;
  for (int idx = threadIdx.x; idx < (N + 1) * rpb; idx += blockDim.x) {
    int m = idx / rpb, rr = idx % rpb;
    long r = r0 + rr;
    if (r >= nrows) continue;
    double2 zm = Z[(long)rr * M + m];
    double2 zc = Z[(long)rr * M + (M - m) % M];
    zc.y = -zc.y;
    double2 E = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
    double2 O = make_double2(0.5 * (zm.y - zc.y), -0.5 * (zm.x - zc.x));  // -i (zm - zc) / 2
    double2 g = cadd(E, cmul(a.twI[m], O));
    int j = (int)(r / a.ncols), c = (int)(r % a.ncols);
    G[(long)m * a.gmst + (long)j * a.gld + c] = make_double2(g.x * inv_i, g.y * inv_i);
  }
}

__global__ __launch_bounds__(256) void k_fft_inv(const double2* __restrict__ G, double* __restrict__ xy,
                                                 FftArgs a) {
  extern __shared__ double2 lds[];
  const int M = a.M, rpb = a.rows_per_block, N = a.N;
  const int tpr = blockDim.x / rpb;
  const int row = threadIdx.x / tpr, tr = threadIdx.x % tpr;
  const long nrows = (long)a.J * a.ncols;
  const long r0 = (long)blockIdx.x * rpb;
  double2* X = lds;
  double2* Zb = lds + (long)rpb * M;
  // gather X[m] (m <= N) coalesced over rows; zero above N
  for (int idx = threadIdx.x; idx < M * rpb; idx += blockDim.x) {
    int m = idx / rpb, rr = idx % rpb;
    long r = r0 + rr;
    double2 g = make_double2(0.0, 0.0);
    if (m <= N && r < nrows) {
      int j = (int)(r / a.ncols), c = (int)(r % a.ncols);
      g = G[(long)m * a.gmst + (long)j * a.gld + c];
      if (m == 0) g.y = 0.0;
    }
    X[(long)rr * M + m] = g;
  }
  __syncthreads();
  // Z_m = A_m + i B_m, A = X_m + conj(X_{M-m}), B = (X_m - conj(X_{M-m})) e^{2 pi i m / I}
  if (row < rpb) {
    for (int m = tr; m < M; m += tpr) {
      double2 xm = X[(long)row * M + m];
      double2 xc = m ? X[(long)row * M + (M - m)] : make_double2(0.0, 0.0);
      xc.y = -xc.y;
      double2 A = cadd(xm, xc);
      double2 Bv = cmulc(csub(xm, xc), a.twI[m]);
      Zb[(long)row * M + m] = make_double2(A.x - Bv.y, A.y + Bv.x);
    }
  }
  __syncthreads();
  double2* z = stockham<+1>(Zb, X, M, a.nrad, a.rad, a.twM, rpb, tpr, row, tr);
  if (row < rpb) {
    long r = r0 + row;
    if (r < nrows) {
      int j = (int)(r / a.ncols), c = (int)(r % a.ncols);
      double s = a.rowscale ? a.rowscale[j] : 1.0;
      double2* x = reinterpret_cast<double2*>(xy + c * a.ldg + (long)j * a.I);
      for (int k = tr; k < M; k += tpr) {
        double2 v = z[(long)row * M + k];
        x[k] = make_double2(v.x * s, v.y * s);
      }
    }
  }
}

__global__ void k_twiddles(int M, int I, double2* twM, double2* twI) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < M) {
    double s, c;
    sincospi(-2.0 * k / M, &s, &c);
    twM[k] = make_double2(c, s);
    sincospi(-2.0 * k / I, &s, &c);
    twI[k] = make_double2(c, s);
  }
}

int fft_plan_init(dsht_plan* p) {
  FftPlan& f = p->fft;
  f.M = p->I / 2;
  int M = f.M, n = 0;
  while (M % 4 == 0) { f.rad[n++] = 4; M /= 4; }
  while (M % 2 == 0) { f.rad[n++] = 2; M /= 2; }
  while (M % 3 == 0) { f.rad[n++] = 3; M /= 3; }
  while (M % 5 == 0) { f.rad[n++] = 5; M /= 5; }
  if (M != 1 || n > 24) return fail(DSHT_EARG, "dsht_create: im/2 = %d must be a product of 2, 3, 5", f.M);
  f.nrad = n;
  DSHT_HIP(hipMalloc(&f.twM, sizeof(double2) * f.M));
  DSHT_HIP(hipMalloc(&f.twI, sizeof(double2) * f.M));
  hipLaunchKernelGGL(k_twiddles, dim3(cdiv(f.M, 256)), dim3(256), 0, 0, f.M, p->I, f.twM, f.twI);
  DSHT_HIP(hipGetLastError());
  return 0;
}

void fft_plan_free(dsht_plan* p) {
  if (p->fft.twM) (void)hipFree(p->fft.twM);
  if (p->fft.twI) (void)hipFree(p->fft.twI);
  p->fft.twM = p->fft.twI = nullptr;
}

static int fft_rows_per_block(int M) {
  // two ping-pong buffers of rpb rows of M complex doubles within 64 KiB
  int rpb = 16;
  while (rpb > 1 && 2L * rpb * M * (long)sizeof(double2) > 65536) rpb /= 2;
  return rpb;
}

static FftArgs fft_args(const dsht_plan* p, long ldg, int ncols, long gld) {
  FftArgs a;
  a.I = p->I;
  a.M = p->fft.M;
  a.N = p->N;
  a.J = p->J;
  a.ncols = ncols;
  a.ldg = ldg;
  a.gld = gld;
  a.gmst = (long)p->J * gld;
  a.nrad = p->fft.nrad;
  for (int i = 0; i < 24; i++) a.rad[i] = p->fft.rad[i];
  a.twM = p->fft.twM;
  a.twI = p->fft.twI;
  a.rows_per_block = fft_rows_per_block(p->fft.M);
  a.rowscale = nullptr;
  return a;
}

int launch_fft_fwd(const dsht_plan* p, const double* xy, long ldg, int ncols, double2* G, long gld,
                   hipStream_t s) {
  FftArgs a = fft_args(p, ldg, ncols, gld);
  long nrows = (long)p->J * ncols;
  int nb = (int)cdivl(nrows, a.rows_per_block);
  size_t lds = 2ul * a.rows_per_block * a.M * sizeof(double2);
  hipLaunchKernelGGL(k_fft_fwd, dim3(nb), dim3(256), lds, s, xy, G, a);
  DSHT_HIP(hipGetLastError());
  return 0;
}

int launch_fft_inv(const dsht_plan* p, const double2* G, long gld, int ncols, double* xy, long ldg,
                   hipStream_t s) {
  FftArgs a = fft_args(p, ldg, ncols, gld);
  long nrows = (long)p->J * ncols;
  int nb = (int)cdivl(nrows, a.rows_per_block);
  size_t lds = 2ul * a.rows_per_block * a.M * sizeof(double2);
  hipLaunchKernelGGL(k_fft_inv, dim3(nb), dim3(256), lds, s, G, xy, a);
  DSHT_HIP(hipGetLastError());
  return 0;
}

}  // namespace dsht
