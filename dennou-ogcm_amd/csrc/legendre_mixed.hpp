// legendre_mixed.hpp -- shared definitions of the split-precision Legendre contractions:
// 16-bit MFMA fragment types, the split forms, and the per-plan split tables.
#pragma once
#include <cstdint>
#include <type_traits>

#include "dsht_internal.hpp"

namespace dsht {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));

constexpr int KM = 32;  // MFMA K of the 16-bit forms (one chunk)

// form index: 0 bf16x3, 1 bf16x6, 2 fp16, 3 fp16x3
template <int FORM>
struct MxForm {
  static constexpr bool BF = (FORM == 0 || FORM == 1);
  static constexpr int NT = FORM == 1 ? 3 : (FORM == 2 ? 1 : 2);  // split terms
  static constexpr int NP = FORM == 1 ? 6 : (FORM == 2 ? 1 : 3);  // products
  static constexpr bool SCALED = !BF;
  using T = typename std::conditional<BF, __bf16, _Float16>::type;
  using V8 = typename std::conditional<BF, bf16x8_t, f16x8_t>::type;
};

__device__ __forceinline__ float4_t mfma16(bf16x8_t a, bf16x8_t b, float4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ float4_t mfma16(f16x8_t a, f16x8_t b, float4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}
// 16x16x32 16-bit MFMA lane layout: A[r][k] at lane r + 16*(k/8), element k%8; B[k][c] at lane
// c + 16*(k/8), element k%8; result element i of lane l holds D[4*(l/16) + i][l%16].
__device__ __forceinline__ int acc_row16(int lane, int i) { return 4 * (lane >> 4) + i; }

// ---------------------------------------------------------------- tables
struct MixedTables {
  int form_family = -1;  // 0: bf16, 1: fp16 (each holds up to 3 terms)
  int nterm = 0;
  uint16_t* fP[3] = {nullptr, nullptr, nullptr};
  uint16_t* fH[3] = {nullptr, nullptr, nullptr};
  uint16_t* iP[3] = {nullptr, nullptr, nullptr};
  uint16_t* iH[3] = {nullptr, nullptr, nullptr};
  long* d_offF = nullptr;  // per-m offsets (elements) of the forward tables
  long* d_offI = nullptr;
  long sizeF = 0, sizeI = 0;
  double* d_minv = nullptr;  // per-m table unscale 2^e (1 for bf16)
  float* d_cwv = nullptr;    // north-half quadrature weights, FP32
  float* d_cwd = nullptr;
  int J2p32 = 0;
  // K chunks of the inverse contraction: chunk g belongs to wavenumber chunk_m[g]; the
  // chunks of m start at cbase[m] (parity 0 first)
  int* d_chunk_m = nullptr;
  int* d_cbase = nullptr;
  int nchunks = 0;
};

struct MixedState {
  MixedTables tab[2];  // [family]
};

// bf16 forms share one table family, fp16 forms the other
static inline int MxFormFamily(int form) { return form <= 1 ? 0 : 1; }

// algorithmic work of one split-precision contraction launch (FP64-equivalent flops;
// Fourier arrays in FP32, spectral in FP64, one read of each 16-bit table term used)
// nwritten: Fourier arrays an inverse writes when it differs from the operator's outputs
// (the fused value/gradient form with GradLon from the derivative FFT); -1 = all
static inline void mixed_work(const dsht_plan* p, int op, bool fwd, int ncols, double* bytes, double* flops,
                              int nwritten = -1) {
  // inverse I_VGRAD (index 5): one input, three outputs, two products
  static const int nin[] = {1, 1, 1, 2, 2, 1}, nout[] = {1, 1, 1, 1, 2, 3}, nt[] = {1, 1, 1, 2, 4, 2};
  double T = 0.5 * (p->N + 1.0) * (p->N + 2.0);
  double four = (double)(p->N + 1) * p->J * 8.0 * ncols, spec = (double)p->L * 8.0 * ncols;
  *bytes = fwd ? nin[op] * four + nout[op] * spec : nin[op] * spec + (nwritten >= 0 ? nwritten : nout[op]) * four;
  *flops = 2.0 * p->J * T * ncols * nt[op];
}

__host__ __device__ inline int kpad32(int n) { return (n + KM - 1) / KM * KM; }

// ---------------------------------------------------------------- direct-fragment kernels
// (legendre_mx2.hip) one column chunk of an inverse contraction
struct Mx2InvJob {
  const double* W0;
  const double* W1;
  long sl, sc;
  int ncols;
  const double* pro;        // per-l prefactor of the data (nullable)
  const double* rs;         // per-latitude output scale (nullable)
  const unsigned* colmax;   // [input][column] max |pro w| (float bits); scaled forms only
  uint16_t* pack;           // packed data panels (workspace, mx2_inv_pack_bytes); unused if staged
  bool staged;              // whole-K staged inverse: split the data once per block into LDS
  float2* G0;
  float2* G1;
  float2* G2 = nullptr;     // third output (I_VGRAD)
  long gld, gmst;
};

struct Mx2FwdJob {
  const float2* G0;
  const float2* G1;
  long gld, gmst;
  int ncols;
  double* W0;
  double* W1;
  long sl, sc;
  const float* cw;          // north-half quadrature weights (FP32, zero padded)
  const unsigned* colmax;   // scaled forms only
  float bound;
  const double* diag0;
  const double* diag1;
  double scale;
  bool staged = false;      // whole-K staged forward (k_mx2_fwd_wk) where its LDS panel fits
};

size_t mx2_inv_pack_bytes(const MixedTables& t, int form, int op, int ncols);
// LDS bytes per block of the whole-K staged inverse (at most MX2_STAGED_LDS_MAX to run)
constexpr size_t MX2_STAGED_LDS_MAX = 160 * 1024;
size_t mx2_inv_staged_lds(const dsht_plan* p, int form, int op);
int mx2_inv(const dsht_plan* p, const MixedTables& t, int form, int op, const Mx2InvJob& j, hipStream_t s);
int mx2_fwd(const dsht_plan* p, const MixedTables& t, int form, int op, const Mx2FwdJob& j, hipStream_t s);
int launch_spec_colmax(const double* w, long sl, long sc, int L, int ncols, const double* pro, unsigned* out,
                       hipStream_t s);

}  // namespace dsht
