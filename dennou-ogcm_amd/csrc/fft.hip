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

// Fourier coefficients are stored as double2 (FP64 form) or float2 (split-precision forms);
// arithmetic is FP64 either way.
__device__ inline double2 load_g(const double2* p) { return *p; }
__device__ inline double2 load_g(const float2* p) {
  float2 v = *p;
  return make_double2((double)v.x, (double)v.y);
}
__device__ inline void store_g(double2* p, double2 v) { *p = v; }
__device__ inline void store_g(float2* p, double2 v) { *p = make_float2((float)v.x, (float)v.y); }

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

template <typename GT>
__global__ __launch_bounds__(256) void k_fft_fwd(const double* __restrict__ xy, GT* __restrict__ G, FftArgs a,
                                                 unsigned* __restrict__ colmax) {
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
    g = make_double2(g.x * inv_i, g.y * inv_i);
    store_g(&G[(long)m * a.gmst + (long)j * a.gld + c], g);
    if (colmax) atomicMax(colmax + c, __float_as_uint((float)fmax(fabs(g.x), fabs(g.y))));
  }
}

template <typename GT>
__global__ __launch_bounds__(256) void k_fft_inv(const GT* __restrict__ G, double* __restrict__ xy, FftArgs a) {
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
      g = load_g(&G[(long)m * a.gmst + (long)j * a.gld + c]);
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

// ---------------------------------------------------------------- power-of-two lengths
// Compile-time radix plan (radix 4 stages, one radix-2 stage when log2 M is odd): index
// arithmetic is shifts and masks, twiddles come from LDS, and ROWS rows share each
// workgroup so the Fourier stores of consecutive rows form ROWS*16-byte runs per m.
// Stages run in place: each thread reads all of its butterflies, the block synchronises,
// then it writes them.
template <int M>
struct P2Plan {
  static constexpr int LOG2 = __builtin_ctz(M);
  static constexpr int N4 = LOG2 / 2;         // radix-4 stages
  static constexpr bool HAS2 = (LOG2 & 1) != 0;  // trailing radix-2 stage
};

template <int M, int ROWS, int R, int NS, int SIGN>
__device__ __forceinline__ void p2_stage(double2* __restrict__ buf, const double2* __restrict__ tw, int tid) {
  constexpr int MR = M / R;
  constexpr int NB = ROWS * MR;               // butterflies in the block
  constexpr int PER = (NB + 255) / 256;       // per thread
  constexpr int TS = M / (NS * R);            // twiddle stride
  double2 v[PER][R];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int b = tid + 256 * q;
    if (NB % 256 == 0 || b < NB) {
      const int row = b / MR, jj = b % MR, k = jj % NS;
      const double2* S = buf + row * M;
#pragma unroll
      for (int r = 0; r < R; r++) {
        double2 x = S[jj + r * MR];
        if (r) {
          double2 w = tw[k * r * TS];
          x = SIGN < 0 ? cmul(x, w) : cmulc(x, w);
        }
        v[q][r] = x;
      }
      dft_small<SIGN>(v[q], R);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int b = tid + 256 * q;
    if (NB % 256 == 0 || b < NB) {
      const int row = b / MR, jj = b % MR, k = jj % NS;
      double2* D = buf + row * M;
      const int d = (jj / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) D[d + r * NS] = v[q][r];
    }
  }
  __syncthreads();
}

template <int M, int ROWS, int S4, int SIGN>
__device__ __forceinline__ void p2_stages4(double2* buf, const double2* tw, int tid) {
  if constexpr (S4 < P2Plan<M>::N4) {
    constexpr int NS = 1 << (2 * S4);
    p2_stage<M, ROWS, 4, NS, SIGN>(buf, tw, tid);
    p2_stages4<M, ROWS, S4 + 1, SIGN>(buf, tw, tid);
  }
}

template <int M, int ROWS, int SIGN>
__device__ __forceinline__ void p2_fft(double2* buf, const double2* tw, int tid) {
  p2_stages4<M, ROWS, 0, SIGN>(buf, tw, tid);
  if constexpr (P2Plan<M>::HAS2) p2_stage<M, ROWS, 2, (1 << (2 * P2Plan<M>::N4)), SIGN>(buf, tw, tid);
}

template <int M, int ROWS, typename GT>
__global__ __launch_bounds__(256) void k_fft_fwd_p2(const double* __restrict__ xy, GT* __restrict__ G,
                                                    FftArgs a, unsigned* __restrict__ colmax) {
  extern __shared__ double2 lds[];
  __shared__ unsigned smax[ROWS];
  double2* buf = lds;                  // ROWS * M
  double2* tw = lds + ROWS * M;        // M
  double2* twi = tw + M;               // M
  const int tid = threadIdx.x;
  for (int k = tid; k < M; k += 256) {
    tw[k] = a.twM[k];
    twi[k] = a.twI[k];
  }
  const long nrows = (long)a.J * a.ncols;
  const int N = a.N;
  const double inv_i = 1.0 / (2 * M);
  for (long r0 = (long)blockIdx.x * ROWS; r0 < nrows; r0 += (long)gridDim.x * ROWS) {
    __syncthreads();
    // load ROWS rows (packed as complex) with every thread
    for (int e = tid; e < ROWS * M; e += 256) {
      const int rr = e / M, k = e % M;
      const long r = r0 + rr;
      double2 z = make_double2(0.0, 0.0);
      if (r < nrows) {
        const int j = (int)(r / a.ncols), c = (int)(r % a.ncols);
        z = reinterpret_cast<const double2*>(xy + c * a.ldg + (long)j * (2 * M))[k];
      }
      buf[e] = z;
    }
    __syncthreads();
    if (colmax && tid < ROWS) smax[tid] = 0u;
    p2_fft<M, ROWS, -1>(buf, tw, tid);
    // each thread keeps one row (256 is a multiple of ROWS)
    float gmax = 0.f;
    for (int idx = tid; idx < (N + 1) * ROWS; idx += 256) {
      const int m = idx / ROWS, rr = idx % ROWS;
      const long r = r0 + rr;
      if (r >= nrows) continue;
      double2 zm = buf[rr * M + m];
      double2 zc = buf[rr * M + ((M - m) & (M - 1))];
      zc.y = -zc.y;
      double2 E = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
      double2 O = make_double2(0.5 * (zm.y - zc.y), -0.5 * (zm.x - zc.x));
      double2 g = cadd(E, cmul(twi[m], O));
      const int j = (int)(r / a.ncols), c = (int)(r % a.ncols);
      g = make_double2(g.x * inv_i, g.y * inv_i);
      store_g(&G[(long)m * a.gmst + (long)j * a.gld + c], g);
      gmax = fmaxf(gmax, (float)fmax(fabs(g.x), fabs(g.y)));
    }
    if (colmax) {  // per-column magnitude for the scaled FP16 contraction
      atomicMax(&smax[tid % ROWS], __float_as_uint(gmax));
      __syncthreads();
      if (tid < ROWS && r0 + tid < nrows) atomicMax(colmax + (int)((r0 + tid) % a.ncols), smax[tid]);
    }
  }
}

template <int M, int ROWS, typename GT>
__global__ __launch_bounds__(256) void k_fft_inv_p2(const GT* __restrict__ G, double* __restrict__ xy,
                                                    FftArgs a) {
  extern __shared__ double2 lds[];
  double2* buf = lds;
  double2* tw = lds + ROWS * M;
  double2* twi = tw + M;
  const int tid = threadIdx.x;
  for (int k = tid; k < M; k += 256) {
    tw[k] = a.twM[k];
    twi[k] = a.twI[k];
  }
  const long nrows = (long)a.J * a.ncols;
  const int N = a.N;
  for (long r0 = (long)blockIdx.x * ROWS; r0 < nrows; r0 += (long)gridDim.x * ROWS) {
    __syncthreads();
    for (int idx = tid; idx < M * ROWS; idx += 256) {
      const int m = idx / ROWS, rr = idx % ROWS;
      const long r = r0 + rr;
      double2 g = make_double2(0.0, 0.0);
      if (m <= N && r < nrows) {
        const int j = (int)(r / a.ncols), c = (int)(r % a.ncols);
        g = load_g(&G[(long)m * a.gmst + (long)j * a.gld + c]);
        if (m == 0) g.y = 0.0;
      }
      buf[rr * M + m] = g;
    }
    __syncthreads();
    // Z_m = A + iB in place: the thread owning (m, M-m) writes both
    for (int idx = tid; idx < ROWS * (M / 2 + 1); idx += 256) {
      const int rr = idx / (M / 2 + 1), m = idx % (M / 2 + 1);
      double2* X = buf + rr * M;
      const int mc = (M - m) & (M - 1);
      double2 xm = X[m], xo = X[mc];
      // Z_m from (X_m, X_{M-m}); Z_{M-m} from (X_{M-m}, X_m); X_M = 0 for m = 0
      double2 c1 = m ? make_double2(xo.x, -xo.y) : make_double2(0.0, 0.0);
      double2 A1 = cadd(xm, c1), B1 = cmulc(csub(xm, c1), twi[m]);
      double2 z1 = make_double2(A1.x - B1.y, A1.y + B1.x);
      if (m != 0 && m != M / 2) {
        double2 c2 = make_double2(xm.x, -xm.y);
        double2 A2 = cadd(xo, c2), B2 = cmulc(csub(xo, c2), twi[mc]);
        X[mc] = make_double2(A2.x - B2.y, A2.y + B2.x);
      }
      X[m] = z1;
    }
    __syncthreads();
    p2_fft<M, ROWS, +1>(buf, tw, tid);
    for (int e = tid; e < ROWS * M; e += 256) {
      const int rr = e / M, k = e % M;
      const long r = r0 + rr;
      if (r >= nrows) continue;
      const int j = (int)(r / a.ncols), c = (int)(r % a.ncols);
      const double s = a.rowscale ? a.rowscale[j] : 1.0;
      double2 v = buf[e];
      reinterpret_cast<double2*>(xy + c * a.ldg + (long)j * (2 * M))[k] = make_double2(v.x * s, v.y * s);
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
  a.threads = p->fft_kind == 2 ? 128 : 256;
  a.fp32 = 0;
  a.order = p->fft_order;
  a.deriv = 0;
  a.xy2 = nullptr;
  return a;
}

// rows per workgroup of the power-of-two kernels: ROWS*M complex doubles of LDS <= 64 KiB
template <int M>
struct P2Rows {
  static constexpr int ROWS = (M >= 512) ? 8 : (M >= 256 ? 16 : (M >= 128 ? 16 : 32));
};

template <int M, typename GT>
static int p2_fwd(const dsht_plan* p, const double* xy, const FftArgs& a, GT* G, unsigned* colmax,
                  hipStream_t s) {
  constexpr int ROWS = P2Rows<M>::ROWS;
  long nrows = (long)p->J * a.ncols;
  long nb = cdivl(nrows, ROWS);
  if (nb > 4096) nb = 4096;
  size_t lds = (size_t)(ROWS * M + 2 * M) * sizeof(double2);
  hipLaunchKernelGGL((k_fft_fwd_p2<M, ROWS, GT>), dim3((unsigned)nb), dim3(256), lds, s, xy, G, a, colmax);
  DSHT_HIP(hipGetLastError());
  return 0;
}

template <int M, typename GT>
static int p2_inv(const dsht_plan* p, const GT* G, const FftArgs& a, double* xy, hipStream_t s) {
  constexpr int ROWS = P2Rows<M>::ROWS;
  long nrows = (long)p->J * a.ncols;
  long nb = cdivl(nrows, ROWS);
  if (nb > 4096) nb = 4096;
  size_t lds = (size_t)(ROWS * M + 2 * M) * sizeof(double2);
  hipLaunchKernelGGL((k_fft_inv_p2<M, ROWS, GT>), dim3((unsigned)nb), dim3(256), lds, s, G, xy, a);
  DSHT_HIP(hipGetLastError());
  return 0;
}

template <typename GT, typename XT>
static int fft_fwd_any(const dsht_plan* p, const XT* xy, long ldg, int ncols, GT* G, long gld,
                       unsigned* colmax, hipStream_t s) {
  FftArgs a = fft_args(p, ldg, ncols, gld);
  // forward: FP32 arithmetic for the split forms' FP32 Fourier coefficients; the register
  // kernel transforms each row relative to its first sample (FP64), so a large zonal mean
  // does not cost the m > 0 coefficients their relative precision
  a.fp32 = (sizeof(GT) == 8 && p->fft_kind == 0) ? 1 : 0;
  long nrows = (long)p->J * ncols;
  ProfScope ps("fft_fwd", s, (double)nrows * ((double)sizeof(XT) * p->I + (double)sizeof(GT) * (p->N + 1)),
               (double)nrows * 2.5 * p->I * std::log2((double)p->I));
  if (p->fft_kind != 1 && reg_fft_supported(a.M)) return reg_fft_fwd<GT, XT>(xy, a, G, colmax, s);
  if constexpr (sizeof(XT) == 4) {
    return fail(DSHT_EUNSUPPORTED, "FP32 grids need a register FFT kernel (DSHT_TUNE_FFT 0)");
  } else {
    switch (a.M) {
      case 16: return p2_fwd<16>(p, xy, a, G, colmax, s);
      case 32: return p2_fwd<32>(p, xy, a, G, colmax, s);
      case 64: return p2_fwd<64>(p, xy, a, G, colmax, s);
      case 128: return p2_fwd<128>(p, xy, a, G, colmax, s);
      case 256: return p2_fwd<256>(p, xy, a, G, colmax, s);
      case 512: return p2_fwd<512>(p, xy, a, G, colmax, s);
    }
    int nb = (int)cdivl(nrows, a.rows_per_block);
    size_t lds = 2ul * a.rows_per_block * a.M * sizeof(double2);
    hipLaunchKernelGGL(k_fft_fwd<GT>, dim3(nb), dim3(256), lds, s, xy, G, a, colmax);
    DSHT_HIP(hipGetLastError());
    return 0;
  }
}

bool fft_inv_deriv_ok(const dsht_plan* p) { return p->fft_kind != 1 && reg_fft_supported(p->fft.M); }

template <typename GT, typename XT>
static int fft_inv_any(const dsht_plan* p, const GT* G, long gld, int ncols, XT* xy, long ldg,
                       hipStream_t s, const double* rowscale, int deriv, XT* xy2) {
  FftArgs a = fft_args(p, ldg, ncols, gld);
  a.rowscale = rowscale;
  a.deriv = deriv;
  a.xy2 = xy2;
  if ((deriv || xy2) && !fft_inv_deriv_ok(p))
    return fail(DSHT_EUNSUPPORTED, "derivative inverse FFT needs a register kernel");
  // inverse: FP32 arithmetic on the split forms' FP32 Fourier coefficients (each output point
  // is a sum of already-FP32 terms; the error stays at FP32 rounding of the output scale)
  a.fp32 = (sizeof(GT) == 8 && p->fft_kind == 0) ? 1 : 0;
  long nrows = (long)p->J * ncols;
  ProfScope ps("fft_inv", s, (double)nrows * ((xy2 ? 2.0 : 1.0) * sizeof(XT) * p->I + (double)sizeof(GT) * (p->N + 1)),
               (double)nrows * (xy2 ? 2.0 : 1.0) * 2.5 * p->I * std::log2((double)p->I));
  if (p->fft_kind != 1 && reg_fft_supported(a.M)) return reg_fft_inv<GT, XT>(G, a, xy, s);
  if constexpr (sizeof(XT) == 4) {
    return fail(DSHT_EUNSUPPORTED, "FP32 grids need a register FFT kernel (DSHT_TUNE_FFT 0)");
  } else {
    switch (a.M) {
      case 16: return p2_inv<16>(p, G, a, xy, s);
      case 32: return p2_inv<32>(p, G, a, xy, s);
      case 64: return p2_inv<64>(p, G, a, xy, s);
      case 128: return p2_inv<128>(p, G, a, xy, s);
      case 256: return p2_inv<256>(p, G, a, xy, s);
      case 512: return p2_inv<512>(p, G, a, xy, s);
    }
    int nb = (int)cdivl(nrows, a.rows_per_block);
    size_t lds = 2ul * a.rows_per_block * a.M * sizeof(double2);
    hipLaunchKernelGGL(k_fft_inv<GT>, dim3(nb), dim3(256), lds, s, G, xy, a);
    DSHT_HIP(hipGetLastError());
    return 0;
  }
}

int launch_fft_fwd(const dsht_plan* p, const double* xy, long ldg, int ncols, double2* G, long gld,
                   hipStream_t s) {
  return fft_fwd_any<double2, double>(p, xy, ldg, ncols, G, gld, nullptr, s);
}

int launch_fft_inv(const dsht_plan* p, const double2* G, long gld, int ncols, double* xy, long ldg,
                   hipStream_t s, const double* rowscale, int deriv, double* xy2) {
  return fft_inv_any<double2, double>(p, G, gld, ncols, xy, ldg, s, rowscale, deriv, xy2);
}

int launch_fft_fwd32(const dsht_plan* p, const double* xy, long ldg, int ncols, float2* G, long gld,
                     unsigned* colmax, hipStream_t s) {
  return fft_fwd_any<float2, double>(p, xy, ldg, ncols, G, gld, colmax, s);
}

int launch_fft_inv32(const dsht_plan* p, const float2* G, long gld, int ncols, double* xy, long ldg,
                     hipStream_t s, const double* rowscale, int deriv, double* xy2) {
  return fft_inv_any<float2, double>(p, G, gld, ncols, xy, ldg, s, rowscale, deriv, xy2);
}

int launch_fft_fwd32f(const dsht_plan* p, const float* xy, long ldg, int ncols, float2* G, long gld,
                      unsigned* colmax, hipStream_t s) {
  return fft_fwd_any<float2, float>(p, xy, ldg, ncols, G, gld, colmax, s);
}

int launch_fft_inv32f(const dsht_plan* p, const float2* G, long gld, int ncols, float* xy, long ldg,
                      hipStream_t s, const double* rowscale, int deriv, float* xy2) {
  return fft_inv_any<float2, float>(p, G, gld, ncols, xy, ldg, s, rowscale, deriv, xy2);
}

}  // namespace dsht
