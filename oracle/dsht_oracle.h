/*
 * dsht_oracle.h -- FP64 CPU restatement of dogcm's spherical-harmonic transform path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity oracle for the HIP library in
 * dennou-ogcm_amd/ and the timed CPU baseline of bench.py.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the product
 * path (libdsht.so, the Fortran SpmlUtil_mod shim, the Python host wrapper) never
 * calls into it.
 *
 * Conventions (the same ones the HIP library implements):
 *   grid     xy(0:I-1, 1:J) per column, longitude fastest: x[c*I*J + j*I + i],
 *            lambda_i = 2*pi*i/I, mu_j ascending (south -> north), roots of P_J
 *            (doc/description/ogcm/space_discretization.tex:27-38).
 *   spectral w(l) per column, l = n*(n+1) + m (0-based; SPMODEL's l_nm(n,m)-1):
 *            w(n,m)  = Re s_n^m  (m >= 0),  w(n,-m) = Im s_n^m (m > 0),
 *            q(lambda,mu) = sum_{n,|m|<=n} s_n^m P_n^m(mu) e^{i m lambda},
 *            s_n^{-m} = conj(s_n^m), P_n^m normalised to int_{-1}^{1} P^2 dmu = 2,
 *            no Condon-Shortley phase (space_discretization.tex:160-170).
 *   forward  G^m(mu_j) = (1/I) sum_i q_ij e^{-i m lambda_i};
 *            s_n^m = sum_j w_j G^m(mu_j) P_n^m(mu_j), sum_j w_j = 1 (tex:188-208).
 *   inverse  G^m(mu_j) = sum_n s_n^m P_n^m(mu_j); q_ij = sum_m G^m e^{i m lambda_i} (tex:211-219).
 */
#ifndef DSHT_ORACLE_H
#define DSHT_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

/* transform kinds (same numbering as include/dsht.h) */
enum {
  ORC_VAL = 0,        /* w_xy / xy_w                                   */
  ORC_DIVLAMBDA = 1,  /* w_DivLambda_xy : (1/(1-mu^2)) d/dlambda        */
  ORC_DIVMU = 2,      /* w_DivMu_xy     : d/dmu                          */
  ORC_GRADLAMBDA = 3, /* xy_GradLambda_w: d/dlambda                      */
  ORC_GRADMU = 4,     /* xy_GradMu_w    : (1-mu^2) d/dmu                 */
  ORC_GRADLON = 5,    /* xy_GradLon_w   : (1/cos phi) d/dlambda          */
  ORC_GRADLAT = 6     /* xy_GradLat_w   : d/dphi                         */
};

const char* orc_last_error(void);
orc_t* orc_create(int nmax, int im, int jm, double radius);
void orc_destroy(orc_t* h);
int orc_set_threads(int n);

/* grid: mu[J], w[J] (sum 1), lat[J], lon[I], y_lat_weight[J] (sum 2), x_lon_weight[I] */
void orc_grid(const orc_t* h, double* mu, double* wq, double* lat, double* lon,
              double* y_lat_weight, double* x_lon_weight);
/* P_n^m(mu_j), n = m..nmax+1 and H_n^m = (1-mu^2) dP/dmu, n = m..nmax, for all j */
int orc_legendre(const orc_t* h, int m, int n, double* p_out, double* h_out);
/* longitude real FFT of nrow rows: G[r*(N+1)+m] (complex, interleaved re/im) */
void orc_fft_fwd(const orc_t* h, const double* x, double* G, int nrow);
void orc_fft_inv(const orc_t* h, const double* G, double* x, int nrow);

/* transforms over ncols columns (grid column stride I*J, spectral column stride L) */
int orc_fwd(const orc_t* h, int kind, const double* xy, double* w, int ncols);
int orc_inv(const orc_t* h, int kind, const double* w, double* xy, int ncols);
int orc_alpha_fwd(const orc_t* h, const double* a, const double* b, double* w, int ncols);
int orc_alpha_inv(const orc_t* h, const double* wa, const double* wb, double* xy, int ncols);
int orc_uv2vordiv(const orc_t* h, const double* ucos, const double* vcos,
                  double* vor, double* div, int ncols);
int orc_vordiv2uv(const orc_t* h, const double* vor, const double* div,
                  double* u, double* v, int ncols);
void orc_lapla(const orc_t* h, const double* w, double* out, int ncols);
void orc_laplainv(const orc_t* h, const double* w, double* out, int ncols);
/* integrals: IntLonLat_xy, AvrLonLat_xy per column; ya_IntLon_xya / ya_AvrLon_xya per row */
void orc_intlonlat(const orc_t* h, const double* xy, double* out, int ncols);
void orc_avrlonlat(const orc_t* h, const double* xy, double* out, int ncols);
void orc_ya_intlon(const orc_t* h, const double* xy, double* out, int ncols);
void orc_ya_avrlon(const orc_t* h, const double* xy, double* out, int ncols);
/* Interpolate_w: value of column w at (lon, lat) points */
void orc_interpolate(const orc_t* h, const double* w, const double* lon,
                     const double* lat, double* out, int npts);
/* diagonal coefficient tables (length L), LPhys_DIFF_spm_mod.f90:140-167,
   HBEBaroc_hspm_vfvm_mod.f90:83-95, HBEBarot_hspm_vfvm_mod.f90:286-305 */
void orc_coef_hdiff(const orc_t* h, double kappa, double numdiff, int ord, double* c);
void orc_coef_hvisc(const orc_t* h, double nu, double numdiff, int ord, double* c);
void orc_coef_advfilter(const orc_t* h, double* f);
void orc_coef_helmholtz(const orc_t* h, double diag, double c0, double* c);
/* w <- d[l] * w  (column-wise) */
void orc_diag(const orc_t* h, const double* d, double* w, int ncols);

/* Vertical column diagnostics (model/dogcm/src/dynamics/hspm_vfvm/HBEDiagnose_hspm_vfvm_mod.f90).
 * Layer k (0 = top, the reference's KS) of a field at x[k*npts + i], i < npts grid points;
 * cdk[k] the layer thickness in sigma.  The reference loops, in their order, with no
 * fused multiply-adds.
 *   OMG     (:90-143)  omg (km+1 half levels): omg[km] = 0, omg[k] = omg[k+1] - div[k]*(h[k]*cdk[k])
 *   HydPres (:290-325) r = 0; r' = r + ((grav*h[k])*dens[k])*cdk[k]; p[k] = 0.5*(r + r')
 *   UVBarot (:221-255) ub = (sum_k u[k]*(h[k]*cdk[k])) / (sum_k h[k]*cdk[k]), vb likewise */
void orc_col_omega(const double* div, const double* h, int km, const double* cdk, double* omg, long npts);
void orc_col_hydpres(const double* dens, const double* h, int km, const double* cdk, double grav,
                     double* p, long npts);
void orc_col_uvbarot(const double* u, const double* v, const double* h, int km, const double* cdk,
                     double* ub, double* vb, long npts);
/* calc_ADVFlxZ_QUICK (model/dogcm/src/dynamics/hspm_vfvm/HBEBaroc_hspm_vfvm_mod.f90:233-317):
 * flux at the km+1 half levels from omg (km+1 half levels), trc (km layers), cdk (km) and
 * rfdk (km+1 reciprocal half-level spacings); linfree selects the linear free surface
 * boundary at the top.  km >= 2. */
void orc_col_quick(const double* omg, const double* trc, int km, const double* cdk, const double* rfdk,
                   int linfree, double* flx, long npts);
/* LAPACK DGTSV, NRHS = 1 (LAPACK 3.x SRC/dgtsv.f): solves the tridiagonal system with
 * subdiagonal dl[n-1], diagonal d[n], superdiagonal du[n-1] in place (b <- x; dl, d, du
 * overwritten as DGTSV does).  Returns the number of row interchanges, or -(i+1) for a
 * zero pivot at row i (DGTSV's INFO). */
int orc_dgtsv(int n, double* dl, double* d, double* du, double* b);
/* calc_VDiffTRCEq (model/dogcm/src/physics/DOGCM_Phys_hspm_vfvm_mod.f90:752-861): qa = q0 + x
 * per grid point, x from the reference's tridiagonal system solved by orc_dgtsv; fdk holds the
 * km+1 half-level spacings, sfc_bc / btm_bc are 'D' or 'N'.  Returns the total number of
 * row interchanges over all points. */
long orc_col_vdiff(const double* q0, const double* rhs, const double* kv, const double* e3, int km,
                   const double* cdk, const double* fdk, double dt, double alpha, int sfc_bc,
                   int btm_bc, double* qa, long npts);
/* calc_VDiffMOMEq (same file:865-975): b = dt*rhs; Dirichlet bottom row 1 + alpha*coef*(bfric + r). */
long orc_col_vdiff_mom(const double* q0, const double* rhs, const double* kv, const double* e3, int km,
                       const double* cdk, const double* fdk, double dt, double alpha, int sfc_bc,
                       int btm_bc, const double* bfric, double* qa, long npts);

#ifdef __cplusplus
}
#endif
#endif
