/*
 * dsht.h -- C ABI of the MI355X spherical-harmonic transform library (libdsht.so).
 *
 * Drop-in replacement for the horizontal half of dogcm's SpmlUtil_mod
 * (model/dogcm/src/util/math/SpmlUtil_mod.f90:71-135) and the SPMODEL wa_module /
 * ISPACK routines it cascades.  The Fortran module in
 * dennou-ogcm_amd/fortran/SpmlUtil_mod.f90 binds these entry points with
 * ISO_C_BINDING and keeps the reference's function-style API.
 *
 * Conventions
 *   grid      xy(0:I-1,1:J) per column, longitude fastest; lambda_i = 2 pi i / I;
 *             mu_j = sin(lat_j) ascending (south -> north), roots of P_J
 *             (doc/description/ogcm/space_discretization.tex:27-38).
 *   spectral  SPMODEL order l = n(n+1)+m+1 (1-based; 0-based index n(n+1)+m),
 *             w(l_nm(n,m)) = Re s_n^m for m >= 0, w(l_nm(n,-m)) = Im s_n^m for m > 0,
 *             q = sum s_n^m P_n^m(mu) e^{i m lambda}, s_n^{-m} = conj(s_n^m),
 *             int_{-1}^{1} (P_n^m)^2 dmu = 2, no Condon-Shortley phase.
 *   host API  arrays are Fortran column-major: xya(0:I-1,J,ncols), wa(L,ncols).
 *   device API  pointers are device memory; grids: column c at xy + c*ldg;
 *             spectra: element (l, c) at w + l*sl + c*sc  (DSHT layout "LC":
 *             sl = ld >= ncols, sc = 1 is the fast layout; Fortran "CL": sl = 1, sc = ld).
 *   All entry points return 0 on success or a DSHT_E* code; dsht_last_error()
 *   gives the message (thread-local).  The host API is reentrant: each calling
 *   thread gets its own HIP stream and buffers (the reference calls SpmlUtil from
 *   inside `!$omp parallel do` loops, SpmlUtil_mod.f90:360,389,446).
 */
#ifndef DSHT_H
#define DSHT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define DSHT_OK 0
#define DSHT_EARG 1        /* invalid argument (sizes, null pointers, kind)  */
#define DSHT_ENOMEM 2      /* device or host allocation failed               */
#define DSHT_EDEVICE 3     /* HIP runtime error                              */
#define DSHT_ESTATE 4      /* plan not initialised / already destroyed       */
#define DSHT_EUNSUPPORTED 5

/* transform kinds */
#define DSHT_VAL 0        /* w_xy / xy_w                                      */
#define DSHT_DIVLAMBDA 1  /* w_DivLambda_xy : (1/(1-mu^2)) d/dlambda          */
#define DSHT_DIVMU 2      /* w_DivMu_xy     : d/dmu                            */
#define DSHT_GRADLAMBDA 3 /* xy_GradLambda_w: d/dlambda                        */
#define DSHT_GRADMU 4     /* xy_GradMu_w    : (1-mu^2) d/dmu                   */
#define DSHT_GRADLON 5    /* xy_GradLon_w   : (1/cos phi) d/dlambda            */
#define DSHT_GRADLAT 6    /* xy_GradLat_w   : d/dphi                           */

/* Legendre-contraction precision forms */
#define DSHT_PREC_FP64 0    /* FP64 MFMA end to end                          */
#define DSHT_PREC_BF16X3 1  /* 2-term BF16 split, 3 products, FP32 accumulate  */
#define DSHT_PREC_FP16 2    /* FP16 with per-m power-of-two scaling, FP32 acc  */
#define DSHT_PREC_BF16X6 3  /* 3-term BF16 split, 6 products, FP32 accumulate  */
#define DSHT_PREC_FP16X3 4  /* 2-term FP16 split (scaled), 3 products, FP32 acc */

/* diagonal spectral coefficient tables (dsht_coef) */
#define DSHT_COEF_LAPLA 0      /* -n(n+1)                 (w_Lapla_w)                          */
#define DSHT_COEF_LAPLAINV 1   /* -1/(n(n+1)), 0 at n=0    (w_LaplaInv_w)                       */
#define DSHT_COEF_HDIFF 2      /* params {kappa, numdiff, ord}: w_HDiffCoefH, LPhys_DIFF_spm_mod.f90:148-149 */
#define DSHT_COEF_HVISC 3      /* params {nu, numdiff, ord}: w_HViscCoefH, LPhys_DIFF_spm_mod.f90:151-154    */
#define DSHT_COEF_ADVFILTER 4  /* w_AdvFilter, HBEBaroc_hspm_vfvm_mod.f90:83-95                 */
#define DSHT_COEF_HELMHOLTZ 5  /* params {diag, c0}: w_Coef, HBEBarot_hspm_vfvm_mod.f90:290-299 */
#define DSHT_COEF_BEDIFF 6     /* params {dt, kappa|nu, numdiff, ord, momentum(0/1)}:
                                  1/(1 - dt*c_n), DOGCM_Phys_hspm_vfvm_mod.f90:640,741-742 */

typedef struct dsht_plan dsht_plan;

/* ---------------------------------------------------------------- setup
 * dsht_create replaces SpmlUtil_Init -> wa_Initial(nMax,iMax,jMax,kMax+1[,np])
 * (SpmlUtil_mod.f90:192-241).  Builds Gauss nodes/weights and the P_n^m, H_n^m
 * tables in FP64 on the current HIP device.  Requires im even, im >= 2*nmax+2,
 * im/2 a product of 2, 3, 5, jm even, jm >= nmax+1.  kmax is a sizing hint.      */
int dsht_create(int nmax, int im, int jm, int kmax, double radius, dsht_plan** plan);
/* SpmlUtil_Final (SpmlUtil_mod.f90:246-255) */
int dsht_destroy(dsht_plan* plan);
const char* dsht_last_error(void);
int dsht_get_dims(const dsht_plan* plan, int* nmax, int* im, int* jm, int* lm);
int dsht_set_precision(dsht_plan* plan, int prec);
int dsht_get_precision(const dsht_plan* plan);

/* get_SpmlGridInfo (SpmlUtil_mod.f90:271-295) and the cascaded x_Lon, y_Lat,
   x_Lon_Weight, y_Lat_Weight (sum 2), xy_Lon, xy_Lat; any pointer may be NULL.   */
int dsht_grid_info(const dsht_plan* plan, double* x_lon, double* y_lat, double* x_lon_weight,
                   double* y_lat_weight, double* xy_lon, double* xy_lat);
/* The vertical half of get_SpmlGridInfo (SpmlUtil_mod.f90:288,291): g_Sig and
   g_Sig_WEIGHT, the kmax+1 Chebyshev-Lobatto points of [sig_min, sig_max] and their
   Clenshaw-Curtis weights for tmax Chebyshev modes, as at_Initial(kMax, tMax, SigMin,
   SigMax) builds them (at_module_omp.f90:598-619; SpmlUtil_init passes SigMin = -1,
   SigMax = 0, SpmlUtil_mod.f90:173-174,226).  Host-only; needs no plan or device.
   Errors as at_Initial's: kmax, tmax > 0, kmax even, tmax <= kmax.  sig, sig_weight
   hold kmax+1 values each (either may be NULL); sig[0] = sig_max.                 */
int dsht_sig_grid(int kmax, int tmax, double sig_min, double sig_max, double* sig, double* sig_weight);
/* nm_l / l_nm (cascaded, SpmlUtil_mod.f90:120): 1-based l; returns 0 for bad (n,m) */
int dsht_nm_l(const dsht_plan* plan, int* n_of_l, int* m_of_l);
int dsht_l_nm(const dsht_plan* plan, int n, int m);
/* rn(l,1) = -n(n+1) (cascaded rn, used in LPhys_DIFF_spm_mod.f90:144) */
int dsht_rn(const dsht_plan* plan, double* rn);
/* diagonal coefficient table of length L into a host array */
int dsht_coef(const dsht_plan* plan, int which, const double* params, int nparams, double* out);

/* ---------------------------------------------------------------- host API
 * Fortran arrays in host memory; ncols columns (layers).                       */
/* w_xy, w_DivLambda_xy, w_DivMu_xy, wa_xya (kind = DSHT_VAL/DIVLAMBDA/DIVMU)    */
int dsht_w_xy(dsht_plan* plan, int kind, const double* xy, double* w, int ncols);
/* xy_w, xy_GradLambda_w, xy_GradMu_w, xy_GradLon_w, xy_GradLat_w, xya_wa          */
int dsht_xy_w(dsht_plan* plan, int kind, const double* w, double* xy, int ncols);
/* xy_w, xy_GradLon_w and xy_GradLat_w of the same field in one pass (SpmlUtil_mod.f90:414-428;
   the reference callers take both gradients of one field, HBEBaroc_hspm_vfvm_mod.f90:403,416,
   LPhys_RediGM_hspm_vfvm_mod.f90:432-433,437-438).  The contraction reads
   the spectra once for the P and H products.  Any one or two outputs may be NULL.   */
int dsht_xy_vgrad_w(dsht_plan* plan, const double* w, double* xy, double* xy_lon, double* xy_lat,
                    int ncols);
/* w_AlphaOptr_xy / wa_AlphaOptr_xya (SpmlUtil_mod.f90:398-412,453-473)           */
int dsht_w_alpha_xy(dsht_plan* plan, const double* a, const double* b, double* w, int ncols);
/* xy_AlphaOptr_w / xya_AlphaOptr_wa (SpmlUtil_mod.f90:414-451)                   */
int dsht_xy_alpha_w(dsht_plan* plan, const double* wa, const double* wb, double* xy, int ncols);
/* calc_UVCosLat2VorDiv (SpmlUtil_mod.f90:299-314, 342-368)                        */
int dsht_uvcoslat2vordiv(dsht_plan* plan, const double* ucos, const double* vcos, double* vor,
                         double* div, int ncols);
/* calc_VorDiv2UV (SpmlUtil_mod.f90:316-340, 370-396)                              */
int dsht_vordiv2uv(dsht_plan* plan, const double* vor, const double* div, double* u, double* v,
                   int ncols);
/* out = d[l] * w  (w_Lapla_w, w_LaplaInv_w, filters, implicit-diffusion factors) */
int dsht_w_diag_w(dsht_plan* plan, const double* d, const double* w, double* out, int ncols);
/* Forward transform, diagonal epilogue, inverse transform in one call:
   xy_w( d * w_xy(xy) ) -- DOGCM_Phys_hspm_vfvm_mod.f90:636-644 with d = w_Filter/(1-dt*c) */
int dsht_xy_diag_xy(dsht_plan* plan, const double* d, const double* xy, double* out, int ncols);
/* IntLonLat_xy / AvrLonLat_xy per column; ya_IntLon_xya / ya_AvrLon_xya per (j, column) */
int dsht_intlonlat(dsht_plan* plan, const double* xy, double* out, int ncols);
int dsht_avrlonlat(dsht_plan* plan, const double* xy, double* out, int ncols);
int dsht_ya_intlon(dsht_plan* plan, const double* xy, double* out, int ncols);
int dsht_ya_avrlon(dsht_plan* plan, const double* xy, double* out, int ncols);
/* Interpolate_w / a_Interpolate_wa: column w evaluated at npts (lon, lat) points */
int dsht_interpolate(dsht_plan* plan, const double* w, const double* lon, const double* lat,
                     double* out, int npts);

/* Energy and enstrophy spectra of streamfunctions: nma_EnergyFromStreamfunc_wa,
 * na_EnergyFromStreamfunc_wa, nma_EnstrophyFromStreamfunc_wa, na_EnstrophyFromStreamfunc_wa
 * (cascaded from SPMODEL, SpmlUtil_mod.f90:131-132).  On the unit sphere, for the
 * normalisation above: E(n,+-m) = (1/2) n(n+1) c_m w(n,+-m)^2 and Z(n,+-m) =
 * (1/2) n^2 (n+1)^2 c_m w(n,+-m)^2 with c_0 = 1, c_m = 2 (m != 0), so that their sum over
 * (n, m) is the global mean of |grad psi|^2 / 2, resp. (lapla psi)^2 / 2 (divide by R^2, R^4
 * for a sphere of radius R).  nma(0:nmax, -nmax:nmax, ncols), Fortran order, zero where
 * |m| > n; na(0:nmax, ncols) = the sums over m.  Either output may be NULL.  The reference's
 * SPMODEL source is absent: this normalisation is fixed by the Parseval identity above.     */
#define DSHT_SPEC_ENERGY 0
#define DSHT_SPEC_ENSTROPHY 1
int dsht_spec_energy(dsht_plan* plan, int kind, const double* wa, int ncols, double* nma, double* na);

/* ---------------------------------------------------------------- device API
 * Device pointers, stream-ordered on `stream` (a hipStream_t; NULL = default).
 * diag: optional device array of L multipliers fused into the spectral side
 * (forward epilogue / inverse prologue); NULL means 1.                         */
int dsht_fwd_d(dsht_plan* plan, int kind, const double* xy, long ldg, double* w, long sl, long sc,
               int ncols, const double* diag, void* stream);
int dsht_inv_d(dsht_plan* plan, int kind, const double* w, long sl, long sc, double* xy, long ldg,
               int ncols, const double* diag, void* stream);
/* FP32 grid I/O of the split precision forms (no counterpart in the reference, whose grids
 * are real(DP)): dsht_fwd_d, dsht_inv_d and dsht_inv_vgrad_d with float grids (same layout
 * and pitch in elements).  Spectra stay FP64; the Legendre contractions and FFTs run as in
 * the split forms, so only the rounding of the grid values to FP32 is added (relative
 * 6e-8).  DSHT_EUNSUPPORTED in the FP64 form and without the register FFT kernels.         */
int dsht_fwd_f32_d(dsht_plan* plan, int kind, const float* xy, long ldg, double* w, long sl,
                   long sc, int ncols, const double* diag, void* stream);
int dsht_inv_f32_d(dsht_plan* plan, int kind, const double* w, long sl, long sc, float* xy,
                   long ldg, int ncols, const double* diag, void* stream);
int dsht_inv_vgrad_f32_d(dsht_plan* plan, const double* w, long sl, long sc, float* xy,
                         float* xy_lon, float* xy_lat, long ldg, int ncols, const double* diag,
                         void* stream);
/* No counterpart in the reference: dsht_inv_d for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as dsht_inv_d, FP64 form only (other forms:
 * DSHT_EUNSUPPORTED).  It calls no hipMalloc/hipFree: its Fourier scratch is a static
 * device array of DSHT_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends, and
 * columns are processed in chunks that fit it.  A plan whose 16-column Fourier block
 * ((nmax+1)*jm*16*16 bytes) exceeds it is refused with DSHT_EARG.  Every call on one GPU
 * uses the same scratch: calls that overlap in time on one GPU (different streams or
 * threads) race and must be serialised by the caller.                                   */
#define DSHT_NOALLOC_SCRATCH_BYTES (64u << 20)
int dsht_inv_noalloc_d(dsht_plan* plan, int kind, const double* w, long sl, long sc, double* xy,
                       long ldg, int ncols, const double* diag, void* stream);
/* device form of dsht_xy_vgrad_w: value, GradLon and GradLat grids of each column */
int dsht_inv_vgrad_d(dsht_plan* plan, const double* w, long sl, long sc, double* xy, double* xy_lon,
                     double* xy_lat, long ldg, int ncols, const double* diag, void* stream);
int dsht_alpha_fwd_d(dsht_plan* plan, const double* a, const double* b, long ldg, double* w,
                     long sl, long sc, int ncols, const double* diag, void* stream);
int dsht_alpha_inv_d(dsht_plan* plan, const double* wa, const double* wb, long sl, long sc,
                     double* xy, long ldg, int ncols, void* stream);
int dsht_uv2vordiv_d(dsht_plan* plan, const double* ucos, const double* vcos, long ldg,
                     double* vor, double* div, long sl, long sc, int ncols, const double* diag,
                     void* stream);
int dsht_vordiv2uv_d(dsht_plan* plan, const double* vor, const double* div, long sl, long sc,
                     double* u, double* v, long ldg, int ncols, void* stream);
int dsht_diag_d(dsht_plan* plan, const double* d, const double* w, double* out, long sl, long sc,
                int ncols, void* stream);
/* device form of dsht_spec_energy; nma, na device memory                                  */
int dsht_spec_energy_d(dsht_plan* plan, int kind, const double* w, long sl, long sc, int ncols,
                       double* nma, double* na, void* stream);
/* FP64 global sums; out is device memory */
int dsht_intlonlat_d(dsht_plan* plan, const double* xy, long ldg, double* out, int ncols,
                     void* stream);
int dsht_avrlonlat_d(dsht_plan* plan, const double* xy, long ldg, double* out, int ncols,
                     void* stream);
int dsht_ya_intlon_d(dsht_plan* plan, const double* xy, long ldg, double* out, int ncols,
                     void* stream);
int dsht_interpolate_d(dsht_plan* plan, const double* w, long sl, const double* lon,
                       const double* lat, double* out, int npts, void* stream);
/* Vertical column diagnostics between the inverse and the forward transforms, on device
 * grids.  The km layers of one field are km consecutive grid columns with pitch ldg
 * (layer 0 on top: the reference's KS), each an I x J grid; cdk holds the km layer
 * thicknesses in sigma (z_CDK(KS:KE), DOGCM_GaussSpmVFvmGrid_mod.f90:400-402), device
 * memory.  Each result equals the reference loop evaluated in FP64, in its order.
 * Outputs must not overlap any input (the kernels write outputs while inputs are still
 * being read); overlapping calls are refused with DSHT_EARG.
 *
 * HBEDiagnose_OMG (HBEDiagnose_hspm_vfvm_mod.f90:90-143): omg holds the km+1 half levels
 * KS-1..KE (pitch ldo), half level k on top of layer k: omg[km] = 0,
 * omg[k] = omg[k+1] - div[k] * (h[k] * cdk[k]).                                          */
int dsht_col_omega_d(dsht_plan* plan, const double* div, const double* h, long ldg, int km,
                     const double* cdk, double* omg, long ldo, void* stream);
/* HBEDiagnose_HydPres (:290-325): r = 0; per layer r' = r + grav*h*dens*cdk,
 * out[k] = 0.5*(r + r') (km columns, pitch ldo).                                        */
int dsht_col_hydpres_d(dsht_plan* plan, const double* dens, const double* h, long ldg, int km,
                       const double* cdk, double grav, double* out, long ldo, void* stream);
/* HBEDiagnose_UVBarot (:221-255): thickness-weighted layer means of u and v, one I x J
 * grid each; either output may be NULL.                                                 */
int dsht_col_uvbarot_d(dsht_plan* plan, const double* u, const double* v, const double* h,
                       long ldg, int km, const double* cdk, double* ub, double* vb, void* stream);
/* calc_ADVFlxZ_QUICK (HBEBaroc_hspm_vfvm_mod.f90:233-317): the vertical advective flux of a
 * tracer at the km+1 half levels (pitch ldf) from the half-level omg (pitch ldo, as written
 * by dsht_col_omega_d) and the km tracer layers trc (pitch ldg); rfdk holds the km+1
 * reciprocal half-level spacings z_RFDK(KS-1:KE) (device memory).  flx[0] is 0, or
 * omg[0]*trc[0] when linfree is nonzero (KinBCTYPE_LinFreeSurf); flx[km] is 0.  km >= 2. */
int dsht_col_quick_d(dsht_plan* plan, const double* omg, long ldo, const double* trc, long ldg,
                     int km, const double* cdk, const double* rfdk, int linfree, double* flx,
                     long ldf, void* stream);
/* calc_VDiffTRCEq (DOGCM_Phys_hspm_vfvm_mod.f90:752-861): implicit vertical diffusion of a
 * tracer.  Per grid point, solves the tridiagonal system of the reference (coefficients from
 * kv, e3 = xyz_H, z_CDK and the km+1 half-level spacings fdk = z_FDK(KS-1:KE); right-hand
 * side dt*rhs/e3; 'D' or 'N' surface and bottom rows) with LAPACK DGTSV's elimination and
 * back substitution, and writes qa = q0 + x (km layers, pitch ldo).  work: device scratch
 * of lwork >= 3*km*I*J doubles (contents undefined on return).  km >= 2.
 * No aliasing: qa and work must not overlap each other or any input (q0, rhs, kv, e3,
 * bfric, cdk, fdk); an in-place call (qa = q0) is refused with DSHT_EARG.            */
int dsht_col_vdiff_d(dsht_plan* plan, const double* q0, const double* rhs, const double* kv,
                     const double* e3, long ldg, int km, const double* cdk, const double* fdk,
                     double dt, double alpha, int sfc_bc, int btm_bc, double* qa, long ldo,
                     double* work, long lwork, void* stream);
/* calc_VDiffMOMEq (DOGCM_Phys_hspm_vfvm_mod.f90:865-975): the momentum form of the same solve:
 * right-hand side dt*rhs (no division by e3) and, with a Dirichlet bottom, the bottom row
 * 1 + alpha*coef*(bfric + r) with bfric = xy_BtmFrictCoef, one I x J device grid.      */
int dsht_col_vdiff_mom_d(dsht_plan* plan, const double* q0, const double* rhs, const double* kv,
                         const double* e3, long ldg, int km, const double* cdk, const double* fdk,
                         double dt, double alpha, int sfc_bc, int btm_bc, const double* bfric,
                         double* qa, long ldo, double* work, long lwork, void* stream);
/* The column kernels on host arrays (the reference's own calling form): each field is
 * xyz(IM, JM, KS:KE), km consecutive I x J grids; half-level arrays have km+1 grids
 * (xyr(:, :, KS-1:KE)); cdk, rfdk, fdk are host vectors.  Staged through the calling
 * thread's device buffers; results are those of the device entry points above.
 * dsht_col_vdiff runs the momentum form when bfric (an I x J grid) is not NULL.     */
int dsht_col_omega(dsht_plan* plan, const double* div, const double* h, int km,
                   const double* cdk, double* omg);
int dsht_col_hydpres(dsht_plan* plan, const double* dens, const double* h, int km,
                     const double* cdk, double grav, double* out);
int dsht_col_uvbarot(dsht_plan* plan, const double* u, const double* v, const double* h, int km,
                     const double* cdk, double* ub, double* vb);
int dsht_col_quick(dsht_plan* plan, const double* omg, const double* trc, int km,
                   const double* cdk, const double* rfdk, int linfree, double* flx);
int dsht_col_vdiff(dsht_plan* plan, const double* q0, const double* rhs, const double* kv,
                   const double* e3, int km, const double* cdk, const double* fdk, double dt,
                   double alpha, int sfc_bc, int btm_bc, const double* bfric, double* qa);
/* coefficient table written to device memory */
int dsht_coef_d(dsht_plan* plan, int which, const double* params, int nparams, double* out,
                void* stream);

/* ---------------------------------------------------------------- tuning
 * Knobs chosen per (truncation, column count) by the autotuner (dsht/autotune.py):
 * the column chunk of the device path (Fourier scratch per chunk) and the column block
 * of the Legendre contractions.  No counterpart in the reference.                   */
#define DSHT_TUNE_CHUNK_COLS 0 /* 0 = automatic */
#define DSHT_TUNE_LEG_CB 1     /* 0 = automatic, 16 or 32 */
#define DSHT_TUNE_FFT 2        /* 0 = register two-pass kernel where the length has one (256-thread
                                  blocks; FP32 arithmetic in the inverse FFT of the split forms,
                                  whose Fourier coefficients are FP32), 1 = LDS Stockham, 2 = register kernel, 128-thread blocks,
                                  3 = register kernel, FP64 arithmetic in every form */
#define DSHT_TUNE_LEG 3        /* Legendre contractions.  FP64: 0 = LDS-staged inverse with the
                                  fragment-stream forward (default), 1 or 3 = LDS-staged, 2 = per-wave
                                  fragment streams.  Split-form inverse: 0 = per form (default:
                                  whole-K LDS staging for BF16X6 and FP16X3, packed data panels otherwise),
                                  2 = packed data panels, 3 = whole-K LDS staging; 1 (a retired
                                  design) returns DSHT_EUNSUPPORTED.  (Key 4 is retired.) */
#define DSHT_TUNE_FFT_ORDER 5  /* row-group order of the register FFTs: 0 = packed (groups walk
                                  the columns of one latitude first and may straddle two
                                  latitudes), 1 = latitude-major (default: each group is 16
                                  columns of one latitude, the groups in flight spread over all
                                  latitudes) */
int dsht_set_tuning(dsht_plan* plan, int key, int value);
int dsht_get_tuning(const dsht_plan* plan, int key);

/* ---------------------------------------------------------------- batches
 * Many independent instances (planets) of uneven truncation and layer count advanced
 * together.  Instances sharing a plan are packed as columns of one set of device buffers
 * (field f, layer column c at f*cols + c; spectra in the LC layout) so each transform
 * launch covers all of them; one step of every group is captured into a hipGraph.  The
 * reference runs one model per process (OpenMP over layers, SpmlUtil_mod.f90:360-366).
 * Step per instance: 4 fields spectral -> grid with lon/lat derivatives, 6 tendency
 * fields grid -> spectral.                                                            */
typedef struct dsht_batch dsht_batch;
int dsht_batch_create(dsht_batch** batch);
int dsht_batch_add(dsht_batch* batch, dsht_plan* plan, int km, int* instance_id);
int dsht_batch_finalize(dsht_batch* batch);
int dsht_batch_info(const dsht_batch* batch, int instance_id, int* group, int* col0, int* group_cols,
                    long* bytes);
int dsht_batch_ngroups(const dsht_batch* batch);
int dsht_batch_buffers(const dsht_batch* batch, int group, double** spec, double** grid, double** tend,
                       double** spect);
int dsht_batch_step(dsht_batch* batch, void* stream);
int dsht_batch_step_sequential(dsht_batch* batch, void* stream);
int dsht_batch_capture(dsht_batch* batch, void* stream);
int dsht_batch_launch(dsht_batch* batch, void* stream);
int dsht_batch_destroy(dsht_batch* batch);

/* ---------------------------------------------------------------- kernel timers
 * HIP-event timers around every kernel launch of the library, keyed by kernel class
 * ("fft_fwd", "fft_inv", "leg_fwd", "leg_inv", ...), with the algorithmic bytes and
 * flops of each launch.  The counterpart of the reference's ProfUtil rap timers
 * (model/dogcm/src/util/misc/ProfUtil_mod.f90:134-269).  Off by default.           */
int dsht_prof_enable(int on);
int dsht_prof_reset(void);
int dsht_prof_query(const char* name, double* total_ms, long* count, double* bytes, double* flops);

/* ---------------------------------------------------------------- LF-AM3 step
 * A device-resident LF-AM3 dynamics step (DOGCM_TInt_LFAM3_advance_Dyn,
 * model/dogcm/src/tint/DOGCM_TInt_LFAM3_mod.f90:281-880) for up to max_runs independent runs
 * of one plan (truncation) and km layers: rigid lid, flat bottom of depth `depth`, linear
 * equation of state (EOS_Linear_mod.f90), no physics forcing.  Per run and step it makes the
 * reference's 42 forward and 28 inverse transform units per layer plus 9 + 9 barotropic
 * ones (dsht_lfam3_units), each transform over every layer of every active run in one call.
 * State in device memory (dsht_lfam3_state): U, V (km grids per run, run r layer k at grid
 * column r*km + k), tracers (potential temperature [K] then salinity per run: column
 * (2r + t)*km + k), surface pressure (one grid per run), at time levels B and N; a step
 * advances runs [0, nruns) and rotates the levels (B <- N, N <- A).  The vertical grid is
 * the Chebyshev-Lobatto one of dsht_sig_grid(km, km, -1, 0) (km even).  A step allocates
 * nothing itself; the plan's transform workspaces are sized by the first step on a stream.
 * Runs are independent: a run's result does not depend on the other runs in the pool.      */
typedef struct dsht_lfam3 dsht_lfam3;
typedef struct dsht_lfam3_config {
  int km, max_runs, sub_cycles, numdiff_ord;        /* SubCycleNum, NumDiffOrdH            */
  double dt, depth, grav, ref_dens;                 /* DelTime, flat depth, Grav, RefDens  */
  double alpha, lambda;                             /* CoriolisTermACoef, VDiffTermACoef   */
  double hvisc, hdiff, numdiff, vvisc, vdiff, btm_frict; /* hViscCoef, hDiffCoef, NumDiffCoefH,
                                                       vViscCoef, vDiffCoef, xy_BtmFrictCoef */
  int mom_sfc_bc, mom_btm_bc, trc_sfc_bc, trc_btm_bc; /* 'D' or 'N'                      */
  double ref_temp, ref_salt, beta_t, beta_s, cp0, cs0; /* EOS_Linear constants           */
} dsht_lfam3_config;
#define DSHT_LF_U 0
#define DSHT_LF_V 1
#define DSHT_LF_TRC 2
#define DSHT_LF_SFCPRES 3
#define DSHT_LF_B 0
#define DSHT_LF_N 1
#define DSHT_LF_A 2
/* defaults: APEOGCircExp.nml (SemiImplicitScheme_nml, Constants_nml), EOS_Linear_Init     */
int dsht_lfam3_default_config(dsht_lfam3_config* cfg);
int dsht_lfam3_create(dsht_plan* plan, const dsht_lfam3_config* cfg, dsht_lfam3** pool);
int dsht_lfam3_destroy(dsht_lfam3* pool);
/* rotation rate Omega of one run (xy_Cori = 2 Omega sin(lat)); default Earth's          */
int dsht_lfam3_set_rotation(dsht_lfam3* pool, int run, double omega);
/* device pointer and column count of a state array (field DSHT_LF_*, level B, N or A)   */
int dsht_lfam3_state(dsht_lfam3* pool, int field, int level, double** ptr, long* ncols);
int dsht_lfam3_step(dsht_lfam3* pool, int nruns, void* stream);
/* copies run src's state (levels B and N, rotation rate) into slot dst (pool compaction) */
int dsht_lfam3_copy_run(dsht_lfam3* pool, int dst, int src, void* stream);
/* transform units of one step of nruns runs (42 km + 9 forward, 28 km + 9 inverse each)   */
int dsht_lfam3_units(const dsht_lfam3* pool, int nruns, long* fwd_units, long* inv_units);

/* ---------------------------------------------------------------- Redi/GM mixing
 * Isopycnal (Redi) diffusion and Gent-McWilliams skew flux of potential temperature and
 * salinity: LPhys_RediGM_hspm_vfvm_AddMixingTerm (model/dogcm/src/physics/lphys/
 * LPhys_RediGM_hspm_vfvm_mod.f90:218-372).  Slopes from the JM95 expansion coefficients
 * (EOSDriver_alpha_beta, EOS_JM95_mod.f90:272-332) of theta = ptemp [K], S [psu] at
 * p = -ref_dens*grav*z; interior slope tapering DM95 or GKW91 times the LDD95 boundary
 * tapering (LPhys_RediGMHelper_mod.f90:175-431); Redi and skew fluxes
 * (calc_IsopycDiffFlux :460-556, calc_SkewFlux :560-679); then, per tracer,
 * rhs += xy_w(w_AlphaOptr_xy(FLon cos, FLat cos)) and rhs += (FSig above - FSig below)
 * RCDK / H (:308-370).  The plan's radius is RPlanet.  The 'linear' boundary tapering
 * (LinearSlopeTapering) is not supported: with it the reference multiplies by an
 * undefined factor (prepare_SlopeTapering :219-233).                                     */
#define DSHT_TAPER_GKW91 1
#define DSHT_TAPER_DM95 2
#define DSHT_TAPER_LDD95 3
typedef struct dsht_redigm_config {
  double kappa_redi, kappa_gm;  /* DiffCoefRedi, DiffCoefGM [m2/s]                          */
  double slope_max, sd;         /* SlopeMax, Sd (DM95)                                      */
  int interior_taper;           /* DSHT_TAPER_DM95 or DSHT_TAPER_GKW91                      */
  int pbl_taper;                /* DSHT_TAPER_LDD95                                         */
  double omega;                 /* Omega [1/s] (LDD95's eddy displacement c/(2 Omega|sin|)) */
  double ref_dens, grav;        /* RefDens, Grav                                            */
} dsht_redigm_config;
/* defaults: RediGM_nml of APEOGCircExp.nml (800, 800, 4e-3, 1e-3, DM95, LDD95), Earth's
 * Omega (DOGCM_Admin_Constants_mod.f90:158), RefDens 1027, Grav 9.79764                    */
int dsht_redigm_default_config(dsht_redigm_config* cfg);
/* The grid-space part on device grids, given the horizontal gradients: ptemp, salt, h
 * (xyz_H), z (xyz_Z, negative below the surface): km layers each (pitch ldg, layer 0 on top);
 * topo: one I x J grid (xy_Topo, the positive bottom depth); cdk (km) and rfdk (km+1,
 * z_RFDK(KS-1:KE)) device vectors; grad_lon, grad_lat: 2*km grids (pitch ldg), the
 * xy_GradLon_w / xy_GradLat_w of w_xy(ptemp) (layers 0..km-1) then of w_xy(salt), not
 * divided by RPlanet.  Writes 2*km grids each (pitch ldo; temperature layers, then
 * salinity): flx_lon = (FLonRedi + FLonGM) cos(lat), flx_lat likewise (the AlphaOptr
 * inputs) and vdiv = (FSig above - FSig below) RCDK / H.  Outputs must not overlap inputs.
 * 2 <= km <= 1024.                                                                        */
int dsht_col_redigm_d(dsht_plan* plan, const dsht_redigm_config* cfg, const double* ptemp,
                      const double* salt, const double* h, const double* z, long ldg, int km,
                      const double* topo, const double* cdk, const double* rfdk,
                      const double* grad_lon, const double* grad_lat, double* flx_lon,
                      double* flx_lat, double* vdiv, long ldo, void* stream);
/* device doubles of work that dsht_redigm_d needs for km layers                            */
long dsht_redigm_work_size(const dsht_plan* plan, int km);
/* The whole tendency on device grids: the forward transform of ptemp and salt, their
 * GradLon and GradLat (one fused inverse), dsht_col_redigm_d, the AlphaOptr forward of the
 * fluxes and the inverse of the divergence; then ptemp_rhs[k] = (ptemp_rhs[k] + hdiv[k]) +
 * vdiv[k], salt_rhs likewise (km grids each, pitch ldo; they may not overlap the inputs).
 * The transforms run in the plan's precision form.  work: lwork >= dsht_redigm_work_size
 * device doubles (contents undefined on return).  After one call on a stream the call
 * allocates nothing and can be captured in a hipGraph on that stream.                    */
int dsht_redigm_d(dsht_plan* plan, const dsht_redigm_config* cfg, const double* ptemp,
                  const double* salt, const double* h, const double* z, long ldg, int km,
                  const double* topo, const double* cdk, const double* rfdk, double* ptemp_rhs,
                  double* salt_rhs, long ldo, double* work, long lwork, void* stream);
/* host-array form, the reference's calling form: xyz(IM, JM, KS:KE) fields, xy_Topo, host
 * vectors cdk and rfdk; ptemp_rhs and salt_rhs are updated in place                        */
int dsht_redigm(dsht_plan* plan, const dsht_redigm_config* cfg, const double* ptemp,
                const double* salt, const double* h, const double* z, int km, const double* topo,
                const double* cdk, const double* rfdk, double* ptemp_rhs, double* salt_rhs);

#ifdef __cplusplus
}
#endif
#endif /* DSHT_H */
