// fft_reg.hip -- register-resident longitude FFTs for the grid lengths in use
// (I = 2M with M = R1 * R2, both factors small).  Same transform as fft.hip:
//
// Forward:  G^m(mu_j) = (1/I) sum_i q_ij e^{-i m lambda_i},  m = 0..N
// Inverse:  q_ij = G^0 + 2 Re sum_{m=1}^{N} G^m e^{i m lambda_i}
//
// The complex length-M FFT of a packed row z_k = x_{2k} + i x_{2k+1} runs as two passes
// of in-register DFTs (index n = n1 + R1 n2, k = k2 + R2 k1):
//   pass 1, thread n1 of the row:  Y[n1][k2] = W_M^{n1 k2} sum_{n2} z[n1 + R1 n2] W_R2^{n2 k2}
//   pass 2, thread k2 of the row:  Z[k2 + R2 k1] = sum_{n1} Y[n1][k2] W_R1^{n1 k1}
// so a row crosses LDS once between the passes (plus once more for the real/half-complex
// split), against 2 log4 M times for a radix-4 Stockham chain.  Small DFTs are fully
// unrolled with compile-time twiddles.  The grid side is read or written straight from
// registers in 256-byte runs; the Fourier side goes through LDS so that consecutive
// threads take consecutive rows (r = j*ncols + c) and each m is stored as one run.
// LDS rows are padded to an odd number of complex values so that both the per-row and
// the across-row access patterns are free of bank conflicts.
#include <cmath>
#include <type_traits>

#include "dsht_internal.hpp"

namespace dsht {

namespace {

// complex values in FP64 (double2) or FP32 (float2) arithmetic
template <typename V>
struct Cx;
template <>
struct Cx<double2> {
  using S = double;
  __device__ static double2 mk(double x, double y) { return make_double2(x, y); }
};
template <>
struct Cx<float2> {
  using S = float;
  __device__ static float2 mk(float x, float y) { return make_float2(x, y); }
};
template <typename V>
__device__ __forceinline__ V mk(typename Cx<V>::S x, typename Cx<V>::S y) {
  return Cx<V>::mk(x, y);
}
template <typename V, typename W>
__device__ __forceinline__ V cvt(W a) {  // between double2 and float2
  using S = typename Cx<V>::S;
  return Cx<V>::mk((S)a.x, (S)a.y);
}
template <typename V>
__device__ __forceinline__ V rcmul(V a, V b) {
  return mk<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename V>
__device__ __forceinline__ V rcmulc(V a, V b) {  // a * conj(b)
  return mk<V>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename V>
__device__ __forceinline__ V radd(V a, V b) {
  return mk<V>(a.x + b.x, a.y + b.y);
}
template <typename V>
__device__ __forceinline__ V rsub(V a, V b) {
  return mk<V>(a.x - b.x, a.y - b.y);
}
template <int SIGN, typename V>  // multiply by SIGN * i
__device__ __forceinline__ V rmul_i(V a) {
  return SIGN < 0 ? mk<V>(a.y, -a.x) : mk<V>(-a.y, a.x);
}

// ---- compile-time twiddles: cos and sin of 2 pi e / R by reduction to [0, pi/4] and
// Taylor series (terms to x^27, exact to rounding for |x| <= pi/4)
constexpr double ct_sinp(double x) {
  double x2 = x * x, t = x, s = x;
  for (int k = 1; k < 14; k++) {
    t *= -x2 / ((2.0 * k) * (2.0 * k + 1.0));
    s += t;
  }
  return s;
}
constexpr double ct_cosp(double x) {
  double x2 = x * x, t = 1.0, s = 1.0;
  for (int k = 1; k < 14; k++) {
    t *= -x2 / ((2.0 * k - 1.0) * (2.0 * k));
    s += t;
  }
  return s;
}
constexpr double ct_cs(int e, int R, bool want_sin) {
  e = ((e % R) + R) % R;
  const int q = (4 * e) / R, r = 4 * e - q * R;  // angle = (pi/2)(q + r/R), 0 <= r < R
  const double h = 1.57079632679489661923;
  double c = 1.0, s = 0.0;
  if (2 * r <= R) {
    const double x = h * r / R;
    c = ct_cosp(x);
    s = ct_sinp(x);
  } else {
    const double x = h * (R - r) / R;
    c = ct_sinp(x);
    s = ct_cosp(x);
  }
  for (int i = 0; i < q; i++) {
    const double t = c;
    c = -s;
    s = t;
  }
  return want_sin ? s : c;
}

// x * exp(SIGN 2 pi i E / R), E a compile-time constant
template <int E, int R, int SIGN, typename V>
__device__ __forceinline__ V twc(V x) {
  using S = typename Cx<V>::S;
  constexpr int e = ((E % R) + R) % R;
  if constexpr (e == 0) {
    return x;
  } else if constexpr (2 * e == R) {
    return mk<V>(-x.x, -x.y);
  } else if constexpr (4 * e == R) {
    return rmul_i<SIGN>(x);
  } else if constexpr (4 * e == 3 * R) {
    return rmul_i<-SIGN>(x);
  } else {
    constexpr double c = ct_cs(e, R, false), s = SIGN * ct_cs(e, R, true);
    return mk<V>(x.x * (S)c - x.y * (S)s, x.x * (S)s + x.y * (S)c);
  }
}

template <int I, int N, typename F>
__device__ __forceinline__ void sfor_impl(F& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    sfor_impl<I + 1, N>(f);
  }
}
template <int N, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  sfor_impl<0, N>(f);
}

template <int R>
struct Split {
  static constexpr int P = (R % 4 == 0 && R > 4)   ? 4
                           : (R % 2 == 0 && R > 2) ? 2
                           : (R % 3 == 0 && R > 3) ? 3
                           : (R % 5 == 0 && R > 5) ? 5
                                                   : 1;
};

// in-register DFT y_q = sum_p x_p exp(SIGN 2 pi i p q / R), natural order in and out
template <int R, int SIGN, typename V>
__device__ __forceinline__ void dft(V* x) {
  using S = typename Cx<V>::S;
  if constexpr (R == 1) {
    return;
  } else if constexpr (R == 2) {
    const V a = x[0], b = x[1];
    x[0] = radd(a, b);
    x[1] = rsub(a, b);
  } else if constexpr (R == 4) {
    const V t0 = radd(x[0], x[2]), t1 = rsub(x[0], x[2]);
    const V t2 = radd(x[1], x[3]), t3 = rmul_i<SIGN>(rsub(x[1], x[3]));
    x[0] = radd(t0, t2);
    x[1] = radd(t1, t3);
    x[2] = rsub(t0, t2);
    x[3] = rsub(t1, t3);
  } else if constexpr (R == 3) {
    const S s3 = (S)0.86602540378443864676, h = (S)0.5;
    const V t1 = radd(x[1], x[2]), t2 = rsub(x[1], x[2]);
    const V a = mk<V>(x[0].x - h * t1.x, x[0].y - h * t1.y);
    const V b = rmul_i<SIGN>(mk<V>(s3 * t2.x, s3 * t2.y));
    x[0] = radd(x[0], t1);
    x[1] = radd(a, b);
    x[2] = rsub(a, b);
  } else if constexpr (R == 5) {
    const S c1 = (S)0.30901699437494742410, c2 = (S)-0.80901699437494742410;
    const S s1 = (S)0.95105651629515357212, s2 = (S)0.58778525229247312917;
    const V t1 = radd(x[1], x[4]), t4 = rsub(x[1], x[4]);
    const V t2 = radd(x[2], x[3]), t3 = rsub(x[2], x[3]);
    const V a1 = mk<V>(x[0].x + c1 * t1.x + c2 * t2.x, x[0].y + c1 * t1.y + c2 * t2.y);
    const V a2 = mk<V>(x[0].x + c2 * t1.x + c1 * t2.x, x[0].y + c2 * t1.y + c1 * t2.y);
    const V b1 = rmul_i<SIGN>(mk<V>(s1 * t4.x + s2 * t3.x, s1 * t4.y + s2 * t3.y));
    const V b2 = rmul_i<SIGN>(mk<V>(s2 * t4.x - s1 * t3.x, s2 * t4.y - s1 * t3.y));
    x[0] = radd(x[0], radd(t1, t2));
    x[1] = radd(a1, b1);
    x[4] = rsub(a1, b1);
    x[2] = radd(a2, b2);
    x[3] = rsub(a2, b2);
  } else {
    // R = P * Q: n = p + P q, k = k1 + Q k2
    constexpr int P = Split<R>::P, Q = R / P;
    static_assert(P > 1, "radix must factor into 2, 3, 4, 5");
    V y[R];
    sfor<P>([&](auto pc) {
      constexpr int p = decltype(pc)::value;
      V t[Q];
#pragma unroll
      for (int q = 0; q < Q; q++) t[q] = x[p + P * q];
      dft<Q, SIGN>(t);
      sfor<Q>([&](auto kc) {
        constexpr int k1 = decltype(kc)::value;
        y[p * Q + k1] = twc<p * k1, R, SIGN>(t[k1]);
      });
    });
#pragma unroll
    for (int k1 = 0; k1 < Q; k1++) {
      V s[P];
#pragma unroll
      for (int p = 0; p < P; p++) s[p] = y[p * Q + k1];
      dft<P, SIGN>(s);
#pragma unroll
      for (int k2 = 0; k2 < P; k2++) x[k1 + Q * k2] = s[k2];
    }
  }
}

constexpr int odd_up(int x) { return x | 1; }

template <typename GT>
__device__ __forceinline__ GT zero_of();
template <>
__device__ __forceinline__ double2 zero_of<double2>() { return make_double2(0.0, 0.0); }
template <>
__device__ __forceinline__ float2 zero_of<float2>() { return make_float2(0.f, 0.f); }

template <int R1, int R2, int NT, typename V>
struct RegFft {
  static constexpr int M = R1 * R2;
  static constexpr int THREADS = NT;
  static constexpr int ROWS = THREADS / R1;                      // rows per workgroup
  static constexpr int PY = (R2 % 2 == 0) ? R2 + 1 : R2 + 2;     // padded pass-1 stride (odd)
  static constexpr int SR = odd_up(R1 * PY > M + 1 ? R1 * PY : M + 1);  // padded row (odd)
  static constexpr size_t LDS = (size_t)(ROWS * SR + 2 * M) * sizeof(V);
};

// Row groups.  A group is ROWS rows (latitude j, column c) that share one workgroup pass;
// its rows are consecutive columns of one latitude, so the Fourier side (columns fastest)
// is stored and loaded in ROWS-column runs.  Group order (a.order):
//   0  packed: row r = g*ROWS + rr of the latitude-major row list (j = r / ncols); a group
//      may straddle two latitudes;
//   1  latitude-major (default): group g takes latitude g mod J and columns (g / J)*ROWS + rr,
//      so every group's Fourier run is ROWS columns of one latitude (whole lines with the
//      16-column row pitch) and the groups in flight spread over all latitudes.  At T170L60
//      (bf16x6) this takes the step from 0.61 to 0.58 ms.
// (32-bit index arithmetic: J * ncols < 2^31 rows per call)
__device__ __forceinline__ bool map_row(const FftArgs& a, long g, int rr, int rows, int& j, int& c) {
  const int gi = (int)g;
  if (a.order) {
    j = gi % a.J;
    c = (gi / a.J) * rows + rr;
    return c < a.ncols;
  }
  const int r = gi * rows + rr;
  j = r / a.ncols;
  c = r - j * a.ncols;
  return j < a.J;
}
__host__ __device__ inline long row_groups(const FftArgs& a, int rows) {
  return a.order ? (long)a.J * ((a.ncols + rows - 1) / rows) : ((long)a.J * a.ncols + rows - 1) / rows;
}

}  // namespace

// grid rows are streamed once per transform: non-temporal loads (forward) and stores (inverse)
typedef double dvec2 __attribute__((ext_vector_type(2)));

// waves per SIMD the register kernels are compiled for: FP32 arithmetic fits three
// (<= 168 VGPRs, 4 workgroups of LDS per CU), FP64 two (its LDS rows allow two workgroups)
template <typename V>
constexpr int WPE = sizeof(typename Cx<V>::S) == 4 ? 3 : 2;

typedef float fvec2 __attribute__((ext_vector_type(2)));

// two consecutive grid samples (one packed complex input) as FP64, non-temporal
__device__ __forceinline__ double2 load_pair(const double* x) {
  const dvec2 t = __builtin_nontemporal_load(reinterpret_cast<const dvec2*>(x));
  return make_double2(t.x, t.y);
}
__device__ __forceinline__ double2 load_pair(const float* x) {
  const fvec2 t = __builtin_nontemporal_load(reinterpret_cast<const fvec2*>(x));
  return make_double2((double)t.x, (double)t.y);
}
// the same pair in the grid's own type, converted where it is used: a conversion next to the
// load would wait for the load there
__device__ __forceinline__ dvec2 load_raw(const double* x) {
  return __builtin_nontemporal_load(reinterpret_cast<const dvec2*>(x));
}
__device__ __forceinline__ fvec2 load_raw(const float* x) {
  return __builtin_nontemporal_load(reinterpret_cast<const fvec2*>(x));
}
__device__ __forceinline__ void store_pair(double* x, double a, double b) {
  const dvec2 t = {a, b};
  __builtin_nontemporal_store(t, reinterpret_cast<dvec2*>(x));
}
__device__ __forceinline__ void store_pair(float* x, double a, double b) {
  const fvec2 t = {(float)a, (float)b};
  __builtin_nontemporal_store(t, reinterpret_cast<fvec2*>(x));
}

// V: arithmetic type (double2, or float2 for the FP32 Fourier path of the split forms);
// XT: grid element type (double, or float for the split forms' FP32 grid I/O)
template <int R1, int R2, int NT, typename GT, typename V, typename XT>
__global__ __launch_bounds__(NT, WPE<V>) void k_fft_fwd_r(const XT* __restrict__ xy, GT* __restrict__ G,
                                                  FftArgs a, unsigned* __restrict__ colmax) {
  using C = RegFft<R1, R2, NT, V>;
  using S = typename Cx<V>::S;
  constexpr int M = C::M, ROWS = C::ROWS, PY = C::PY, SR = C::SR;
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  V* const lds = reinterpret_cast<V*>(lds_raw);
  __shared__ unsigned smax[ROWS];
  // FP32 arithmetic transforms x - x0 (x0 = the row's first sample, subtracted in FP64) and
  // adds x0 back to m = 0 in FP64: rounding then scales with the row's zonal variation, not
  // with a large zonal mean (temperature, salinity backgrounds)
  constexpr bool REF = sizeof(typename Cx<V>::S) == 4;
  __shared__ double xref[REF ? ROWS : 1];
  V* buf = lds;              // ROWS * SR
  V* tw2 = lds + ROWS * SR;  // W_M^{n1 k2} at k2 * R1 + n1
  V* twi = tw2 + M;          // W_I^m
  const int tid = threadIdx.x;
  for (int e = tid; e < M; e += C::THREADS) {
    const int k2 = e / R1, n1 = e % R1;
    tw2[e] = cvt<V>(a.twM[(n1 * k2) % M]);
    twi[e] = cvt<V>(a.twI[e]);
  }
  const long ngroups = row_groups(a, ROWS);
  const int N = a.N;
  const S inv_i = (S)(1.0 / (2 * M)), h = (S)0.5;
  const int rr1 = tid / R1, n1 = tid % R1;  // pass-1 role
  const int rr2 = tid / R2, k2 = tid % R2;  // pass-2 role
  // store role: with THREADS a multiple of ROWS every item of a thread is in row tid % ROWS
  constexpr bool FIXR = C::THREADS % ROWS == 0;
  const int rs = tid % ROWS;
  // pass-1 inputs straight from global memory; the next group's rows are loaded into the
  // same registers once the current group's pass-1 output is in LDS
  // FP64 arithmetic loads straight into v; FP32 arithmetic (REF) keeps the samples as loaded,
  // in the grid's type, and converts them at the top of the next group, so the loads stay in
  // flight across the barrier.  REF loads are issued by every thread: a thread whose row is
  // past the last column reads the grid's first row, which is transformed but never stored
  // (rows do not mix in either pass or in the store).
  V v[R2];
  using RV = std::conditional_t<sizeof(XT) == 4, fvec2, dvec2>;
  RV raw[REF ? R2 : 1];
  double x0 = 0.0;  // reference sample of the thread's pass-1 row (REF)
  XT x0n = 0;       // the next group's, as loaded
  auto fetch = [&](long g) {
    int j = 0, c = 0;
    if constexpr (REF) {
      const bool ok = rr1 < ROWS && map_row(a, g, rr1, ROWS, j, c);
      const XT* xr = xy + (ok ? c * a.ldg + (long)j * (2 * M) : 0L);
      x0n = xr[0];
#pragma unroll
      for (int n2 = 0; n2 < R2; n2++) raw[n2] = load_raw(xr + 2 * (n1 + R1 * n2));
    } else if (rr1 < ROWS && map_row(a, g, rr1, ROWS, j, c)) {
      const XT* xr = xy + c * a.ldg + (long)j * (2 * M);
      if constexpr (sizeof(XT) == 8) {
        const double2* x = reinterpret_cast<const double2*>(xr);
#pragma unroll
        for (int n2 = 0; n2 < R2; n2++) v[n2] = cvt<V>(x[n1 + R1 * n2]);
      } else {
#pragma unroll
        for (int n2 = 0; n2 < R2; n2++) v[n2] = cvt<V>(load_pair(xr + 2 * (n1 + R1 * n2)));
      }
    } else {
#pragma unroll
      for (int n2 = 0; n2 < R2; n2++) v[n2] = mk<V>(0, 0);
    }
  };
  const long gstep = gridDim.x;
  if ((long)blockIdx.x < ngroups) fetch(blockIdx.x);
  for (long g0 = blockIdx.x; g0 < ngroups; g0 += gstep) {
    if constexpr (REF) {
      x0 = (double)x0n;
#pragma unroll
      for (int n2 = 0; n2 < R2; n2++) v[n2] = mk<V>((double)raw[n2].x - x0, (double)raw[n2].y - x0);
    }
    dft<R2, -1>(v);
    if (colmax && tid < ROWS) smax[tid] = 0u;
    __syncthreads();  // the previous rows are out of buf; twiddles are in place
    if (rr1 < ROWS) {
      V* Y = buf + rr1 * SR + n1 * PY;
      Y[0] = v[0];
#pragma unroll
      for (int q = 1; q < R2; q++) Y[q] = rcmul(v[q], tw2[q * R1 + n1]);
      if constexpr (REF) {
        if (n1 == 0) xref[rr1] = x0;
      }
    }
    if (g0 + gstep < ngroups) fetch(g0 + gstep);
    __syncthreads();
    V u[R1];
    if (rr2 < ROWS) {
      const V* Y = buf + rr2 * SR + k2;
#pragma unroll
      for (int p = 0; p < R1; p++) u[p] = Y[p * PY];
      dft<R1, -1>(u);
    }
    __syncthreads();
    if (rr2 < ROWS) {
      V* X = buf + rr2 * SR;
#pragma unroll
      for (int k1 = 0; k1 < R1; k1++) X[k2 + R2 * k1] = u[k1];
    }
    __syncthreads();
    // half-complex split of the packed transform, 1/I, store; rows fastest across threads
    int js = 0, cs = 0;
    const bool oks = FIXR && map_row(a, g0, rs, ROWS, js, cs);
    for (int idx = tid; idx < (N + 1) * ROWS; idx += C::THREADS) {
      const int m = idx / ROWS, rr = FIXR ? rs : idx % ROWS;
      int j = js, c = cs;
      if (FIXR ? !oks : !map_row(a, g0, rr, ROWS, j, c)) continue;
      const V* X = buf + rr * SR;
      const V zm = X[m];
      V zc = X[m ? M - m : 0];
      zc.y = -zc.y;
      const V E = mk<V>(h * (zm.x + zc.x), h * (zm.y + zc.y));
      const V O = mk<V>(h * (zm.y - zc.y), -h * (zm.x - zc.x));
      V g = radd(E, rcmul(twi[m], O));
      g = mk<V>(g.x * inv_i, g.y * inv_i);
      if constexpr (REF) {
        if (m == 0) g.x = (S)((double)g.x + xref[rr]);
      }
      G[(long)m * a.gmst + (long)j * a.gld + c] = cvt<GT>(g);
      if (colmax) atomicMax(&smax[rr], __float_as_uint(fmaxf(fabsf((float)g.x), fabsf((float)g.y))));
    }
    if (colmax) {  // per-column magnitude for the scaled FP16 contraction
      __syncthreads();
      int j, c;
      if (tid < ROWS && map_row(a, g0, tid, ROWS, j, c)) atomicMax(colmax + c, smax[tid]);
    }
  }
}

// MODE 0: the field; 1: its longitude derivative (transform i m X_m, times a.rowscale);
// 2 (dual): both from one read of the coefficients -- the field into xy, then the
// derivative, times a.rowscale, into a.xy2 -- the coefficients stay in registers between the
// two passes.  Separate instantiations keep each form's register allocation.  XT: grid
// element type.
template <int R1, int R2, int NT, typename GT, typename V, int MODE, typename XT>
__global__ __launch_bounds__(NT, WPE<V>) void k_fft_inv_r(const GT* __restrict__ G, XT* __restrict__ xy,
                                                  FftArgs a) {
  using C = RegFft<R1, R2, NT, V>;
  using S = typename Cx<V>::S;
  constexpr int M = C::M, ROWS = C::ROWS, PY = C::PY, SR = C::SR;
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  V* const lds = reinterpret_cast<V*>(lds_raw);
  V* buf = lds;
  V* tw2 = lds + ROWS * SR;
  V* twi = tw2 + M;
  const int tid = threadIdx.x;
  for (int e = tid; e < M; e += C::THREADS) {
    const int k2 = e / R1, n1 = e % R1;
    tw2[e] = cvt<V>(a.twM[(n1 * k2) % M]);
    twi[e] = cvt<V>(a.twI[e]);
  }
  const long ngroups = row_groups(a, ROWS);
  const int N = a.N;
  const int rr1 = tid / R1, n1 = tid % R1;
  const int rr2 = tid / R2, k2 = tid % R2;
  // Fourier coefficients of a group of ROWS rows, raw: consecutive threads take consecutive
  // rows, zero above the truncation.  The next group's loads are issued as soon as the
  // current group is in LDS, so they overlap its FFT passes and grid stores.
  constexpr int NL = (M * ROWS + C::THREADS - 1) / C::THREADS;
  GT gr[NL];
  // with THREADS a multiple of ROWS every item idx = tid + q * THREADS is in row tid % ROWS
  constexpr bool FIXR = C::THREADS % ROWS == 0;
  auto fetch = [&](long g) {
    int j0 = 0, c0 = 0;
    const bool ok0 = FIXR && map_row(a, g, tid % ROWS, ROWS, j0, c0);
#pragma unroll
    for (int q = 0; q < NL; q++) {
      const int idx = tid + q * C::THREADS;
      const int m = idx / ROWS;
      gr[q] = cvt<GT>(make_float2(0.f, 0.f));
      int j = j0, c = c0;
      const bool ok = FIXR ? ok0 : map_row(a, g, idx % ROWS, ROWS, j, c);
      if (ok && idx < M * ROWS && m <= N) gr[q] = G[(long)m * a.gmst + (long)j * a.gld + c];
    }
  };
  // prefetching holds the next group in registers through the whole pipeline: only where
  // that fits the register budget (FP32 coefficients, M <= 256)
  constexpr bool PF = sizeof(GT) == 8 && M <= 256;
  constexpr int NPASS = MODE == 2 ? 2 : 1;
  const long gstep = gridDim.x;
  if (PF && (long)blockIdx.x < ngroups) fetch(blockIdx.x);
  for (long g0 = blockIdx.x; g0 < ngroups; g0 += gstep) {
    if (!PF) fetch(g0);
#pragma unroll
  for (int pass = 0; pass < NPASS; pass++) {
    const bool deriv = MODE == 1 || (MODE == 2 && pass == 1);
    __syncthreads();  // the previous rows are out of buf
#pragma unroll
    for (int q = 0; q < NL; q++) {
      const int idx = tid + q * C::THREADS;
      if (idx < M * ROWS) {
        V v = cvt<V>(gr[q]);
        if (idx < ROWS) v.y = (S)0;  // m = 0
        buf[(idx % ROWS) * SR + idx / ROWS] = v;
      }
    }
    if (PF && pass == NPASS - 1 && g0 + gstep < ngroups) fetch(g0 + gstep);
    __syncthreads();
    // packed spectrum Z_m from the half-complex pair (X_m, X_{M-m}); X_M = 0 for m = 0
    for (int idx = tid; idx < ROWS * (M / 2 + 1); idx += C::THREADS) {
      const int m = idx / ROWS, rr = idx % ROWS;
      V* X = buf + rr * SR;
      const int mc = m ? M - m : 0;
      V xm = X[m], xo = X[mc];
      if (deriv) {  // longitude derivative: each X_k becomes i k X_k
        const S fm = (S)m, fc = (S)mc;
        xm = mk<V>(-fm * xm.y, fm * xm.x);
        xo = mk<V>(-fc * xo.y, fc * xo.x);
      }
      const V c1 = m ? mk<V>(xo.x, -xo.y) : mk<V>(0, 0);
      const V A1 = radd(xm, c1), B1 = rcmulc(rsub(xm, c1), twi[m]);
      const V z1 = mk<V>(A1.x - B1.y, A1.y + B1.x);
      if (m != 0 && 2 * m != M) {
        const V c2 = mk<V>(xm.x, -xm.y);
        const V A2 = radd(xo, c2), B2 = rcmulc(rsub(xo, c2), twi[mc]);
        X[mc] = mk<V>(A2.x - B2.y, A2.y + B2.x);
      }
      X[m] = z1;
    }
    __syncthreads();
    V v[R2];
    if (rr1 < ROWS) {
      const V* Z = buf + rr1 * SR + n1;
#pragma unroll
      for (int n2 = 0; n2 < R2; n2++) v[n2] = Z[R1 * n2];
    }
    dft<R2, +1>(v);
    __syncthreads();
    if (rr1 < ROWS) {
      V* Y = buf + rr1 * SR + n1 * PY;
      Y[0] = v[0];
#pragma unroll
      for (int q = 1; q < R2; q++) Y[q] = rcmulc(v[q], tw2[q * R1 + n1]);
    }
    __syncthreads();
    if (rr2 < ROWS) {
      V u[R1];
      const V* Y = buf + rr2 * SR + k2;
#pragma unroll
      for (int p = 0; p < R1; p++) u[p] = Y[p * PY];
      dft<R1, +1>(u);
      int j, c;
      if (map_row(a, g0, rr2, ROWS, j, c)) {
        // dual: the field unscaled into xy, the derivative times the row scale into xy2
        const double s = (a.rowscale && (MODE != 2 || pass == 1)) ? a.rowscale[j] : 1.0;
        XT* out = (MODE == 2 && pass == 1) ? reinterpret_cast<XT*>(a.xy2) : xy;
        XT* x = out + c * a.ldg + (long)j * (2 * M);
#pragma unroll
        for (int k1 = 0; k1 < R1; k1++) store_pair(x + 2 * (k2 + R2 * k1), (double)u[k1].x * s, (double)u[k1].y * s);
      }
    }
  }  // pass
  }
}

// ---------------------------------------------------------------- dispatch
// The register kernels walk row groups in a persistent loop (each workgroup's loads for its
// next group overlap the current one), so the grid is the number of workgroups that are
// resident at once: every workgroup then runs from the start, and no second, partly empty
// round of workgroups follows the first.
static int cu_count() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
  }
  return cus;
}

template <typename K>
static long resident_grid(K kernel, int threads, size_t lds) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds) != hipSuccess || per_cu <= 0) {
    (void)hipGetLastError();
    per_cu = 2;
  }
  return (long)per_cu * cu_count();
}

template <int R1, int R2, int NT, typename GT, typename V, typename XT>
static int reg_fwd_nt(const XT* xy, const FftArgs& a, GT* G, unsigned* colmax, hipStream_t s) {
  using C = RegFft<R1, R2, NT, V>;
  static const long cap = resident_grid(k_fft_fwd_r<R1, R2, NT, GT, V, XT>, C::THREADS, C::LDS);
  long nb = row_groups(a, C::ROWS);
  if (nb > cap) nb = cap;
  hipLaunchKernelGGL((k_fft_fwd_r<R1, R2, NT, GT, V, XT>), dim3((unsigned)nb), dim3(C::THREADS), C::LDS, s, xy, G, a,
                     colmax);
  DSHT_HIP(hipGetLastError());
  return 0;
}

template <int R1, int R2, int NT, typename GT, typename V, int MODE, typename XT>
static int reg_inv_mode(const GT* G, const FftArgs& a, XT* xy, hipStream_t s) {
  using C = RegFft<R1, R2, NT, V>;
  static const long cap = resident_grid(k_fft_inv_r<R1, R2, NT, GT, V, MODE, XT>, C::THREADS, C::LDS);
  long nb = row_groups(a, C::ROWS);
  if (nb > cap) nb = cap;
  hipLaunchKernelGGL((k_fft_inv_r<R1, R2, NT, GT, V, MODE, XT>), dim3((unsigned)nb), dim3(C::THREADS), C::LDS, s, G,
                     xy, a);
  DSHT_HIP(hipGetLastError());
  return 0;
}

template <int R1, int R2, int NT, typename GT, typename V, typename XT>
static int reg_inv_nt(const GT* G, const FftArgs& a, XT* xy, hipStream_t s) {
  if (a.xy2) return reg_inv_mode<R1, R2, NT, GT, V, 2, XT>(G, a, xy, s);
  if (a.deriv) return reg_inv_mode<R1, R2, NT, GT, V, 1, XT>(G, a, xy, s);
  return reg_inv_mode<R1, R2, NT, GT, V, 0, XT>(G, a, xy, s);
}

// workgroup size 256 threads (or 128: half the LDS per workgroup); arithmetic FP64, or
// FP32 for the FP32 Fourier path of the split forms (a.fp32); FP32 grids only with the
// split forms' FP32 path
template <int R1, int R2, typename GT, typename XT>
static int reg_fwd(const XT* xy, const FftArgs& a, GT* G, unsigned* colmax, hipStream_t s) {
  if constexpr (sizeof(XT) == 4) {
    if constexpr (sizeof(GT) == 8) {
      if (a.fp32) return reg_fwd_nt<R1, R2, 256, GT, float2, XT>(xy, a, G, colmax, s);
    }
    return fail(DSHT_EUNSUPPORTED, "FP32 grids need the split forms' FP32 FFT path");
  } else {
    if constexpr (sizeof(GT) == 8) {
      if (a.fp32) return reg_fwd_nt<R1, R2, 256, GT, float2, XT>(xy, a, G, colmax, s);
    }
    if (a.threads == 128 && R1 <= 128) return reg_fwd_nt<R1, R2, 128, GT, double2, XT>(xy, a, G, colmax, s);
    return reg_fwd_nt<R1, R2, 256, GT, double2, XT>(xy, a, G, colmax, s);
  }
}

template <int R1, int R2, typename GT, typename XT>
static int reg_inv(const GT* G, const FftArgs& a, XT* xy, hipStream_t s) {
  if constexpr (sizeof(XT) == 4) {
    if constexpr (sizeof(GT) == 8) {
      if (a.fp32) return reg_inv_nt<R1, R2, 256, GT, float2, XT>(G, a, xy, s);
    }
    return fail(DSHT_EUNSUPPORTED, "FP32 grids need the split forms' FP32 FFT path");
  } else {
    if constexpr (sizeof(GT) == 8) {
      if (a.fp32) return reg_inv_nt<R1, R2, 256, GT, float2, XT>(G, a, xy, s);
    }
    if (a.threads == 128 && R1 <= 128) return reg_inv_nt<R1, R2, 128, GT, double2, XT>(G, a, xy, s);
    return reg_inv_nt<R1, R2, 256, GT, double2, XT>(G, a, xy, s);
  }
}

// factor pairs (R1, R2) with M = R1 * R2 that have register kernels
#define DSHT_REG_FFT_SIZES(X) \
  X(4, 2)                     \
  X(4, 4)                     \
  X(6, 6)                     \
  X(8, 4)                     \
  X(8, 6)                     \
  X(8, 8)                     \
  X(9, 8)                     \
  X(12, 8)                    \
  X(16, 8)                    \
  X(15, 9)                    \
  X(16, 12)                   \
  X(16, 16)                   \
  X(18, 15)

bool reg_fft_supported(int M) {
#define X(a, b) \
  if (M == (a) * (b)) return true;
  DSHT_REG_FFT_SIZES(X)
#undef X
  return false;
}

template <typename GT, typename XT>
int reg_fft_fwd(const XT* xy, const FftArgs& a, GT* G, unsigned* colmax, hipStream_t s) {
#define X(r1, r2) \
  if (a.M == (r1) * (r2)) return reg_fwd<r1, r2, GT, XT>(xy, a, G, colmax, s);
  DSHT_REG_FFT_SIZES(X)
#undef X
  return fail(DSHT_EUNSUPPORTED, "no register FFT for M = %d", a.M);
}

template <typename GT, typename XT>
int reg_fft_inv(const GT* G, const FftArgs& a, XT* xy, hipStream_t s) {
#define X(r1, r2) \
  if (a.M == (r1) * (r2)) return reg_inv<r1, r2, GT, XT>(G, a, xy, s);
  DSHT_REG_FFT_SIZES(X)
#undef X
  return fail(DSHT_EUNSUPPORTED, "no register FFT for M = %d", a.M);
}

template int reg_fft_fwd<double2, double>(const double*, const FftArgs&, double2*, unsigned*, hipStream_t);
template int reg_fft_fwd<float2, double>(const double*, const FftArgs&, float2*, unsigned*, hipStream_t);
template int reg_fft_fwd<float2, float>(const float*, const FftArgs&, float2*, unsigned*, hipStream_t);
template int reg_fft_inv<double2, double>(const double2*, const FftArgs&, double*, hipStream_t);
template int reg_fft_inv<float2, double>(const float2*, const FftArgs&, double*, hipStream_t);
template int reg_fft_inv<float2, float>(const float2*, const FftArgs&, float*, hipStream_t);

}  // namespace dsht
