!-------------------------------------------------------------
! SpmlUtil_mod -- drop-in replacement of the horizontal half of dogcm's SpmlUtil_mod
! (model/dogcm/src/util/math/SpmlUtil_mod.f90:71-135) and of the SPMODEL wa_module names it
! cascades, running on an MI355X through the C ABI of libdsht.so (include/dsht.h).
!
! Same procedure names, argument shapes and meanings as the reference:
!   grid     xy(0:im-1, jm), xya(0:im-1, jm, :)     (longitude fastest, mu ascending)
!   spectral w(lm), wa(lm, :)  with lm = (nm+1)**2, SPMODEL order l = n*(n+1)+m+1
! The vertical grid (g_Sig, g_Sig_WEIGHT: Chebyshev-Lobatto levels of [-1, 0] and their
! Clenshaw-Curtis weights, at_module_omp.f90:598-619) is built at SpmlUtil_Init and returned
! by get_SpmlGridInfo.  The vertical (Chebyshev) operators of the reference module
! (xyz_DSig_xyz, IntSig_*, t_g, g_t, ...) are not part of this module; they stay with the
! caller's vertical code.
! Errors from the library stop the program with the library's message, as the
! reference's MessageNotify('E', ...) does.  Calls are thread safe (OpenMP loops).
!-------------------------------------------------------------
module SpmlUtil_mod
  use iso_c_binding
  implicit none
  private

  integer, parameter :: DP = c_double

  integer(c_int), parameter :: DSHT_VAL = 0, DSHT_DIVLAMBDA = 1, DSHT_DIVMU = 2
  integer(c_int), parameter :: DSHT_GRADLAMBDA = 3, DSHT_GRADMU = 4, DSHT_GRADLON = 5, DSHT_GRADLAT = 6
  integer(c_int), parameter :: DSHT_COEF_LAPLA = 0, DSHT_COEF_LAPLAINV = 1

  ! dsht_redigm_config (include/dsht.h): RediGM_nml parameters and the constants it uses
  type, bind(C) :: redigm_config_t
    real(c_double) :: kappa_redi, kappa_gm, slope_max, sd
    integer(c_int) :: interior_taper, pbl_taper
    real(c_double) :: omega, ref_dens, grav
  end type redigm_config_t

  interface
    integer(c_int) function c_create(nmax, im, jm, kmax, radius, plan) bind(C, name="dsht_create")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nmax, im, jm, kmax
      real(c_double), value :: radius
      type(c_ptr), intent(out) :: plan
    end function
    integer(c_int) function c_destroy(plan) bind(C, name="dsht_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: plan
    end function
    type(c_ptr) function c_last_error() bind(C, name="dsht_last_error")
      import :: c_ptr
    end function
    integer(c_int) function c_grid_info(plan, x_lon, y_lat, x_w, y_w, xy_lon, xy_lat) bind(C, name="dsht_grid_info")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      real(c_double) :: x_lon(*), y_lat(*), x_w(*), y_w(*), xy_lon(*), xy_lat(*)
    end function
    integer(c_int) function c_sig_grid(kmax, tmax, sig_min, sig_max, sig, sig_weight) bind(C, name="dsht_sig_grid")
      import :: c_int, c_double
      integer(c_int), value :: kmax, tmax
      real(c_double), value :: sig_min, sig_max
      real(c_double) :: sig(*), sig_weight(*)
    end function
    integer(c_int) function c_coef(plan, which, params, np, out) bind(C, name="dsht_coef")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      integer(c_int), value :: which, np
      real(c_double) :: params(*), out(*)
    end function
    integer(c_int) function c_w_xy(plan, kind, xy, w, ncols) bind(C, name="dsht_w_xy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      integer(c_int), value :: kind, ncols
      real(c_double), intent(in) :: xy(*)
      real(c_double) :: w(*)
    end function
    integer(c_int) function c_xy_w(plan, kind, w, xy, ncols) bind(C, name="dsht_xy_w")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      integer(c_int), value :: kind, ncols
      real(c_double), intent(in) :: w(*)
      real(c_double) :: xy(*)
    end function
    integer(c_int) function c_xy_vgrad_w(plan, w, xy, xy_lon, xy_lat, ncols) bind(C, name="dsht_xy_vgrad_w")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      real(c_double), intent(in) :: w(*)
      type(c_ptr), value :: xy
      real(c_double) :: xy_lon(*), xy_lat(*)
      integer(c_int), value :: ncols
    end function
    integer(c_int) function c_w_alpha_xy(plan, a, b, w, ncols) bind(C, name="dsht_w_alpha_xy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      integer(c_int), value :: ncols
      real(c_double), intent(in) :: a(*), b(*)
      real(c_double) :: w(*)
    end function
    integer(c_int) function c_xy_alpha_w(plan, wa, wb, xy, ncols) bind(C, name="dsht_xy_alpha_w")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      integer(c_int), value :: ncols
      real(c_double), intent(in) :: wa(*), wb(*)
      real(c_double) :: xy(*)
    end function
    integer(c_int) function c_uv2vordiv(plan, u, v, vor, div, ncols) bind(C, name="dsht_uvcoslat2vordiv")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      integer(c_int), value :: ncols
      real(c_double), intent(in) :: u(*), v(*)
      real(c_double) :: vor(*), div(*)
    end function
    integer(c_int) function c_vordiv2uv(plan, vor, div, u, v, ncols) bind(C, name="dsht_vordiv2uv")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      integer(c_int), value :: ncols
      real(c_double), intent(in) :: vor(*), div(*)
      real(c_double) :: u(*), v(*)
    end function
    integer(c_int) function c_w_diag_w(plan, d, w, out, ncols) bind(C, name="dsht_w_diag_w")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      integer(c_int), value :: ncols
      real(c_double), intent(in) :: d(*), w(*)
      real(c_double) :: out(*)
    end function
    integer(c_int) function c_intlonlat(plan, xy, out, ncols) bind(C, name="dsht_intlonlat")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      integer(c_int), value :: ncols
      real(c_double), intent(in) :: xy(*)
      real(c_double) :: out(*)
    end function
    integer(c_int) function c_avrlonlat(plan, xy, out, ncols) bind(C, name="dsht_avrlonlat")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      integer(c_int), value :: ncols
      real(c_double), intent(in) :: xy(*)
      real(c_double) :: out(*)
    end function
    integer(c_int) function c_ya_intlon(plan, xy, out, ncols) bind(C, name="dsht_ya_intlon")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      integer(c_int), value :: ncols
      real(c_double), intent(in) :: xy(*)
      real(c_double) :: out(*)
    end function
    integer(c_int) function c_ya_avrlon(plan, xy, out, ncols) bind(C, name="dsht_ya_avrlon")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      integer(c_int), value :: ncols
      real(c_double), intent(in) :: xy(*)
      real(c_double) :: out(*)
    end function
    integer(c_int) function c_interpolate(plan, w, lon, lat, out, npts) bind(C, name="dsht_interpolate")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      integer(c_int), value :: npts
      real(c_double), intent(in) :: w(*), lon(*), lat(*)
      real(c_double) :: out(*)
    end function
    ! vertical column kernels (host-array forms)
    integer(c_int) function c_col_omega(plan, div, h, nk, cdk, omg) bind(C, name="dsht_col_omega")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      integer(c_int), value :: nk
      real(c_double), intent(in) :: div(*), h(*), cdk(*)
      real(c_double) :: omg(*)
    end function
    integer(c_int) function c_col_hydpres(plan, dens, h, nk, cdk, grav, out) bind(C, name="dsht_col_hydpres")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      integer(c_int), value :: nk
      real(c_double), intent(in) :: dens(*), h(*), cdk(*)
      real(c_double), value :: grav
      real(c_double) :: out(*)
    end function
    integer(c_int) function c_col_uvbarot(plan, u, v, h, nk, cdk, ub, vb) bind(C, name="dsht_col_uvbarot")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      integer(c_int), value :: nk
      real(c_double), intent(in) :: u(*), v(*), h(*), cdk(*)
      real(c_double) :: ub(*), vb(*)
    end function
    integer(c_int) function c_col_quick(plan, omg, trc, nk, cdk, rfdk, linfree, flx) bind(C, name="dsht_col_quick")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      integer(c_int), value :: nk, linfree
      real(c_double), intent(in) :: omg(*), trc(*), cdk(*), rfdk(*)
      real(c_double) :: flx(*)
    end function
    integer(c_int) function c_col_vdiff(plan, q0, rhs, kv, e3, nk, cdk, fdk, dt, alpha, sfc, btm, bfric, qa) &
         bind(C, name="dsht_col_vdiff")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan
      integer(c_int), value :: nk, sfc, btm
      real(c_double), intent(in) :: q0(*), rhs(*), kv(*), e3(*), cdk(*), fdk(*)
      real(c_double), value :: dt, alpha
      type(c_ptr), value :: bfric
      real(c_double) :: qa(*)
    end function
    integer(c_int) function c_spec_energy(plan, kind, wa, ncols, nma, na) bind(C, name="dsht_spec_energy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: plan, nma, na
      integer(c_int), value :: kind, ncols
      real(c_double), intent(in) :: wa(*)
    end function
    integer(c_int) function c_redigm_default_config(cfg) bind(C, name="dsht_redigm_default_config")
      import :: c_int, redigm_config_t
      type(redigm_config_t) :: cfg
    end function
    integer(c_int) function c_redigm(plan, cfg, pt, sa, h, z, nk, topo, cdk, rfdk, pt_rhs, sa_rhs) &
         bind(C, name="dsht_redigm")
      import :: c_int, c_ptr, c_double, redigm_config_t
      type(c_ptr), value :: plan
      type(redigm_config_t), intent(in) :: cfg
      integer(c_int), value :: nk
      real(c_double), intent(in) :: pt(*), sa(*), h(*), z(*), topo(*), cdk(*), rfdk(*)
      real(c_double) :: pt_rhs(*), sa_rhs(*)
    end function
  end interface

  ! public procedures (SpmlUtil_mod.f90:71-135, horizontal part)
  public :: SpmlUtil_Init, SpmlUtil_Final, isInitialzed, get_SpmlGridInfo
  public :: w_AlphaOptr_xy, xy_AlphaOptr_w, wa_AlphaOptr_xya, xya_AlphaOptr_wa
  public :: calc_VorDiv2UV, calc_UVCosLat2VorDiv
  public :: w_xy, xy_w, wa_xya, xya_wa
  public :: xya_GradLon_wa, xya_GradLambda_wa, xya_GradLat_wa, xya_GradMu_wa
  public :: xy_GradLon_w, xy_GradLambda_w, xy_GradLat_w, xy_GradMu_w
  public :: w_DivLambda_xy, w_DivMu_xy, w_Lapla_w, w_LaplaInv_w
  public :: l_nm, nm_l
  public :: IntLonLat_xy, ya_IntLon_xya, AvrLonLat_xy, ya_AvrLon_xya
  public :: a_Interpolate_wa, Interpolate_w
  public :: nma_EnergyFromStreamfunc_wa, na_EnergyFromStreamfunc_wa
  public :: nma_EnstrophyFromStreamfunc_wa, na_EnstrophyFromStreamfunc_wa
  ! extension without a counterpart in the reference module (see calc_GradLonLat_xya_wa)
  public :: calc_GradLonLat_xya_wa
  ! extensions: the grid-space column kernels of the dynamics and physics modules
  ! (HBEDiagnose_OMG / _HydPres / _UVBarot, calc_ADVFlxZ_QUICK, calc_VDiffTRCEq / MOMEq)
  public :: calc_OMG_xyr, calc_HydPres_xyz, calc_UVBarot_xy, calc_ADVFlxZ_QUICK_xyr, calc_VDiff_xyz
  ! extension: the Redi/GM isopycnal mixing of LPhys_RediGM_hspm_vfvm_AddMixingTerm
  public :: calc_RediGM_xyz

  interface calc_VorDiv2UV
    module procedure calc_VorDiv2UV_w_xy
    module procedure calc_VorDiv2UV_wa_xya
  end interface calc_VorDiv2UV
  interface calc_UVCosLat2VorDiv
    module procedure calc_UVCosLat2VorDiv_xy_w
    module procedure calc_UVCosLat2VorDiv_xya_wa
  end interface calc_UVCosLat2VorDiv

  ! public variables (cascaded grid information and spectral arrays)
  real(DP), public, save, allocatable :: xy_CosLat(:,:)
  real(DP), public, save, allocatable :: x_Lon(:), y_Lat(:), x_Lon_Weight(:), y_Lat_Weight(:)
  real(DP), public, save, allocatable :: xy_Lon(:,:), xy_Lat(:,:)
  real(DP), public, save, allocatable :: rn(:,:)   ! rn(:,1) = -n(n+1); rn(:,2) = n
  ! vertical grid (SpmlUtil_mod.f90:50,107): g_Sig(0:km) from the top (0) to the bottom (-1)
  real(DP), public, save, allocatable :: g_Sig_WEIGHT(:)
  real(DP), save, allocatable :: g_Sig(:)
  real(DP), parameter :: SigMin = -1d0, SigMax = 0d0

  type(c_ptr), save :: plan = c_null_ptr
  integer, save :: im, jm, km, nm, lm, tm
  real(DP), save :: Radius
  logical, save :: initalizedFlag = .false.

contains

  subroutine chk(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (rc == 0) return
    call c_f_pointer(c_last_error(), msg, [512])
    text = ''
    do i = 1, 512
      if (msg(i) == c_null_char) exit
      text(i:i) = msg(i)
    end do
    write(*, '(a)') 'SpmlUtil_mod: ' // where // ': ' // trim(text)
    error stop 1
  end subroutine chk

  ! error stop with a message, for argument checks made on the Fortran side
  subroutine arg_error(where, what)
    character(*), intent(in) :: where, what
    write(*, '(a)') 'SpmlUtil_mod: ' // where // ': ' // what
    error stop 1
  end subroutine arg_error

  subroutine SpmlUtil_Init(iMax, jMax, kMax, nMax, tMax, RPlanet, np)
    integer, intent(in) :: iMax, jMax, kMax, nMax, tMax
    real(DP), intent(in) :: RPlanet
    integer, intent(in), optional :: np
    integer :: n, m, l
    im = iMax; jm = jMax; km = kMax; nm = nMax; tm = tMax
    lm = (nm + 1)**2
    Radius = RPlanet
    call chk(c_create(int(nMax, c_int), int(iMax, c_int), int(jMax, c_int), int(kMax + 1, c_int), &
         & real(RPlanet, c_double), plan), 'SpmlUtil_Init')
    allocate(x_Lon(0:im-1), x_Lon_Weight(0:im-1), y_Lat(jm), y_Lat_Weight(jm))
    allocate(xy_Lon(0:im-1, jm), xy_Lat(0:im-1, jm), xy_CosLat(0:im-1, jm), rn(lm, 2))
    call chk(c_grid_info(plan, x_Lon, y_Lat, x_Lon_Weight, y_Lat_Weight, xy_Lon, xy_Lat), 'get grid')
    ! at_Initial(kMax, tMax, SigMin, SigMax) (SpmlUtil_mod.f90:226)
    allocate(g_Sig(0:km), g_Sig_WEIGHT(0:km))
    call chk(c_sig_grid(int(kMax, c_int), int(tMax, c_int), SigMin, SigMax, g_Sig, g_Sig_WEIGHT), &
         & 'SpmlUtil_Init (vertical grid)')
    xy_CosLat = cos(xy_Lat)
    do n = 0, nm
      do m = -n, n
        l = n*(n + 1) + m + 1
        rn(l, 1) = -dble(n*(n + 1))
        rn(l, 2) = dble(n)
      end do
    end do
    if (present(np)) continue
    initalizedFlag = .true.
  end subroutine SpmlUtil_Init

  subroutine SpmlUtil_Final()
    integer(c_int) :: rc
    if (initalizedFlag) then
      rc = c_destroy(plan)
      plan = c_null_ptr
      deallocate(x_Lon, x_Lon_Weight, y_Lat, y_Lat_Weight, xy_Lon, xy_Lat, xy_CosLat, rn, g_Sig, g_Sig_WEIGHT)
      initalizedFlag = .false.
    end if
  end subroutine SpmlUtil_Final

  logical function isInitialzed()
    isInitialzed = initalizedFlag
  end function isInitialzed

  subroutine get_SpmlGridInfo(x_Lon_, y_Lat_, z_Sig_, xy_Lon_, xy_Lat_, x_Lon_Weight_, y_Lat_Weight_, z_Sig_Weight_)
    real(DP), dimension(0:im-1), intent(out), optional :: x_Lon_, x_Lon_Weight_
    real(DP), dimension(jm), intent(out), optional :: y_Lat_, y_Lat_Weight_
    real(DP), dimension(0:km), intent(out), optional :: z_Sig_, z_Sig_Weight_
    real(DP), dimension(0:im-1, jm), intent(out), optional :: xy_Lon_, xy_Lat_
    if (present(x_Lon_)) x_Lon_ = x_Lon
    if (present(y_Lat_)) y_Lat_ = y_Lat
    if (present(x_Lon_Weight_)) x_Lon_Weight_ = x_Lon_Weight
    if (present(y_Lat_Weight_)) y_Lat_Weight_ = y_Lat_Weight
    if (present(xy_Lon_)) xy_Lon_ = xy_Lon
    if (present(xy_Lat_)) xy_Lat_ = xy_Lat
    if (present(z_Sig_)) z_Sig_ = g_Sig
    if (present(z_Sig_Weight_)) z_Sig_Weight_ = g_Sig_WEIGHT
  end subroutine get_SpmlGridInfo

  ! ---------------------------------------------------- index maps
  integer function l_nm(n, m)
    integer, intent(in) :: n, m
    l_nm = n*(n + 1) + m + 1
  end function l_nm

  function nm_l(l)
    integer, intent(in) :: l
    integer :: nm_l(2)
    nm_l(1) = int(sqrt(dble(l - 1)))
    nm_l(2) = l - 1 - nm_l(1)*(nm_l(1) + 1)
  end function nm_l

  ! ---------------------------------------------------- transforms
  function w_xy(xy)
    real(DP), intent(in) :: xy(0:im-1, jm)
    real(DP) :: w_xy(lm)
    call chk(c_w_xy(plan, DSHT_VAL, xy, w_xy, 1_c_int), 'w_xy')
  end function w_xy

  function xy_w(w)
    real(DP), intent(in) :: w(lm)
    real(DP) :: xy_w(0:im-1, jm)
    call chk(c_xy_w(plan, DSHT_VAL, w, xy_w, 1_c_int), 'xy_w')
  end function xy_w

  function wa_xya(xya)
    real(DP), intent(in) :: xya(:,:,:)
    real(DP) :: wa_xya(lm, size(xya, 3))
    call chk(c_w_xy(plan, DSHT_VAL, xya, wa_xya, int(size(xya, 3), c_int)), 'wa_xya')
  end function wa_xya

  function xya_wa(wa)
    real(DP), intent(in) :: wa(:,:)
    real(DP) :: xya_wa(0:im-1, jm, size(wa, 2))
    call chk(c_xy_w(plan, DSHT_VAL, wa, xya_wa, int(size(wa, 2), c_int)), 'xya_wa')
  end function xya_wa

  function w_DivLambda_xy(xy)
    real(DP), intent(in) :: xy(0:im-1, jm)
    real(DP) :: w_DivLambda_xy(lm)
    call chk(c_w_xy(plan, DSHT_DIVLAMBDA, xy, w_DivLambda_xy, 1_c_int), 'w_DivLambda_xy')
  end function w_DivLambda_xy

  function w_DivMu_xy(xy)
    real(DP), intent(in) :: xy(0:im-1, jm)
    real(DP) :: w_DivMu_xy(lm)
    call chk(c_w_xy(plan, DSHT_DIVMU, xy, w_DivMu_xy, 1_c_int), 'w_DivMu_xy')
  end function w_DivMu_xy

  function xy_GradLambda_w(w)
    real(DP), intent(in) :: w(lm)
    real(DP) :: xy_GradLambda_w(0:im-1, jm)
    call chk(c_xy_w(plan, DSHT_GRADLAMBDA, w, xy_GradLambda_w, 1_c_int), 'xy_GradLambda_w')
  end function xy_GradLambda_w

  function xy_GradMu_w(w)
    real(DP), intent(in) :: w(lm)
    real(DP) :: xy_GradMu_w(0:im-1, jm)
    call chk(c_xy_w(plan, DSHT_GRADMU, w, xy_GradMu_w, 1_c_int), 'xy_GradMu_w')
  end function xy_GradMu_w

  function xy_GradLon_w(w)
    real(DP), intent(in) :: w(lm)
    real(DP) :: xy_GradLon_w(0:im-1, jm)
    call chk(c_xy_w(plan, DSHT_GRADLON, w, xy_GradLon_w, 1_c_int), 'xy_GradLon_w')
  end function xy_GradLon_w

  function xy_GradLat_w(w)
    real(DP), intent(in) :: w(lm)
    real(DP) :: xy_GradLat_w(0:im-1, jm)
    call chk(c_xy_w(plan, DSHT_GRADLAT, w, xy_GradLat_w, 1_c_int), 'xy_GradLat_w')
  end function xy_GradLat_w

  ! Extension: xya_GradLon_wa(wa) and xya_GradLat_wa(wa) of the same field in one fused pass
  ! (one contraction for both gradients).  For callers that take both gradients of one field:
  ! HBEBaroc_hspm_vfvm_mod.f90:403,416, LPhys_RediGM_hspm_vfvm_mod.f90:432-438.
  subroutine calc_GradLonLat_xya_wa(wa, xya_lon, xya_lat)
    real(DP), intent(in) :: wa(:,:)
    real(DP), intent(out) :: xya_lon(0:im-1, jm, size(wa, 2)), xya_lat(0:im-1, jm, size(wa, 2))
    call chk(c_xy_vgrad_w(plan, wa, c_null_ptr, xya_lon, xya_lat, int(size(wa, 2), c_int)), &
             'calc_GradLonLat_xya_wa')
  end subroutine calc_GradLonLat_xya_wa

  function xya_GradLambda_wa(wa)
    real(DP), intent(in) :: wa(:,:)
    real(DP) :: xya_GradLambda_wa(0:im-1, jm, size(wa, 2))
    call chk(c_xy_w(plan, DSHT_GRADLAMBDA, wa, xya_GradLambda_wa, int(size(wa, 2), c_int)), 'xya_GradLambda_wa')
  end function xya_GradLambda_wa

  function xya_GradMu_wa(wa)
    real(DP), intent(in) :: wa(:,:)
    real(DP) :: xya_GradMu_wa(0:im-1, jm, size(wa, 2))
    call chk(c_xy_w(plan, DSHT_GRADMU, wa, xya_GradMu_wa, int(size(wa, 2), c_int)), 'xya_GradMu_wa')
  end function xya_GradMu_wa

  function xya_GradLon_wa(wa)
    real(DP), intent(in) :: wa(:,:)
    real(DP) :: xya_GradLon_wa(0:im-1, jm, size(wa, 2))
    call chk(c_xy_w(plan, DSHT_GRADLON, wa, xya_GradLon_wa, int(size(wa, 2), c_int)), 'xya_GradLon_wa')
  end function xya_GradLon_wa

  function xya_GradLat_wa(wa)
    real(DP), intent(in) :: wa(:,:)
    real(DP) :: xya_GradLat_wa(0:im-1, jm, size(wa, 2))
    call chk(c_xy_w(plan, DSHT_GRADLAT, wa, xya_GradLat_wa, int(size(wa, 2), c_int)), 'xya_GradLat_wa')
  end function xya_GradLat_wa

  function w_Lapla_w(w)
    real(DP), intent(in) :: w(lm)
    real(DP) :: w_Lapla_w(lm)
    real(DP) :: d(lm), p(1)
    p = 0d0
    call chk(c_coef(plan, DSHT_COEF_LAPLA, p, 0_c_int, d), 'w_Lapla_w')
    call chk(c_w_diag_w(plan, d, w, w_Lapla_w, 1_c_int), 'w_Lapla_w')
  end function w_Lapla_w

  function w_LaplaInv_w(w)
    real(DP), intent(in) :: w(lm)
    real(DP) :: w_LaplaInv_w(lm)
    real(DP) :: d(lm), p(1)
    p = 0d0
    call chk(c_coef(plan, DSHT_COEF_LAPLAINV, p, 0_c_int, d), 'w_LaplaInv_w')
    call chk(c_w_diag_w(plan, d, w, w_LaplaInv_w, 1_c_int), 'w_LaplaInv_w')
  end function w_LaplaInv_w

  ! ---------------------------------------------------- vector operators (SpmlUtil_mod.f90:297-473)
  function w_AlphaOptr_xy(xy_A, xy_B)
    real(DP), intent(in) :: xy_A(0:im-1, jm), xy_B(0:im-1, jm)
    real(DP) :: w_AlphaOptr_xy(lm)
    call chk(c_w_alpha_xy(plan, xy_A, xy_B, w_AlphaOptr_xy, 1_c_int), 'w_AlphaOptr_xy')
  end function w_AlphaOptr_xy

  function xy_AlphaOptr_w(w_A, w_B)
    real(DP), intent(in) :: w_A(lm), w_B(lm)
    real(DP) :: xy_AlphaOptr_w(0:im-1, jm)
    call chk(c_xy_alpha_w(plan, w_A, w_B, xy_AlphaOptr_w, 1_c_int), 'xy_AlphaOptr_w')
  end function xy_AlphaOptr_w

  function wa_AlphaOptr_xya(xya_A, xya_B)
    real(DP), intent(in) :: xya_A(:,:,:), xya_B(:,:,:)
    real(DP) :: wa_AlphaOptr_xya(lm, size(xya_A, 3))
    call chk(c_w_alpha_xy(plan, xya_A, xya_B, wa_AlphaOptr_xya, int(size(xya_A, 3), c_int)), 'wa_AlphaOptr_xya')
  end function wa_AlphaOptr_xya

  function xya_AlphaOptr_wa(wa_A, wa_B)
    real(DP), intent(in) :: wa_A(:,:), wa_B(:,:)
    real(DP) :: xya_AlphaOptr_wa(0:im-1, jm, size(wa_A, 2))
    call chk(c_xy_alpha_w(plan, wa_A, wa_B, xya_AlphaOptr_wa, int(size(wa_A, 2), c_int)), 'xya_AlphaOptr_wa')
  end function xya_AlphaOptr_wa

  subroutine calc_UVCosLat2VorDiv_xy_w(xy_UCosLat, xy_VCosLat, w_Vor, w_Div)
    real(DP), intent(in) :: xy_UCosLat(0:im-1, jm), xy_VCosLat(0:im-1, jm)
    real(DP), intent(out) :: w_Vor(lm), w_Div(lm)
    call chk(c_uv2vordiv(plan, xy_UCosLat, xy_VCosLat, w_Vor, w_Div, 1_c_int), 'calc_UVCosLat2VorDiv')
  end subroutine calc_UVCosLat2VorDiv_xy_w

  subroutine calc_UVCosLat2VorDiv_xya_wa(xya_UCosLat, xya_VCosLat, wa_Vor, wa_Div)
    real(DP), intent(in) :: xya_UCosLat(:,:,:), xya_VCosLat(:,:,:)
    real(DP), intent(out) :: wa_Vor(lm, size(xya_UCosLat, 3)), wa_Div(lm, size(xya_UCosLat, 3))
    call chk(c_uv2vordiv(plan, xya_UCosLat, xya_VCosLat, wa_Vor, wa_Div, int(size(xya_UCosLat, 3), c_int)), &
         & 'calc_UVCosLat2VorDiv')
  end subroutine calc_UVCosLat2VorDiv_xya_wa

  subroutine calc_VorDiv2UV_w_xy(w_Vor, w_Div, xy_U, xy_V)
    real(DP), intent(in) :: w_Vor(lm), w_Div(lm)
    real(DP), intent(out) :: xy_U(0:im-1, jm), xy_V(0:im-1, jm)
    call chk(c_vordiv2uv(plan, w_Vor, w_Div, xy_U, xy_V, 1_c_int), 'calc_VorDiv2UV')
  end subroutine calc_VorDiv2UV_w_xy

  subroutine calc_VorDiv2UV_wa_xya(wa_Vor, wa_Div, xya_U, xya_V)
    real(DP), intent(in) :: wa_Vor(:,:), wa_Div(:,:)
    real(DP), intent(out) :: xya_U(0:im-1, jm, size(wa_Vor, 2)), xya_V(0:im-1, jm, size(wa_Vor, 2))
    call chk(c_vordiv2uv(plan, wa_Vor, wa_Div, xya_U, xya_V, int(size(wa_Vor, 2), c_int)), 'calc_VorDiv2UV')
  end subroutine calc_VorDiv2UV_wa_xya

  ! ---------------------------------------------------- integrals and means (FP64 sums)
  real(DP) function IntLonLat_xy(xy)
    real(DP), intent(in) :: xy(0:im-1, jm)
    real(DP) :: out(1)
    call chk(c_intlonlat(plan, xy, out, 1_c_int), 'IntLonLat_xy')
    IntLonLat_xy = out(1)
  end function IntLonLat_xy

  real(DP) function AvrLonLat_xy(xy)
    real(DP), intent(in) :: xy(0:im-1, jm)
    real(DP) :: out(1)
    call chk(c_avrlonlat(plan, xy, out, 1_c_int), 'AvrLonLat_xy')
    AvrLonLat_xy = out(1)
  end function AvrLonLat_xy

  function ya_IntLon_xya(xya)
    real(DP), intent(in) :: xya(:,:,:)
    real(DP) :: ya_IntLon_xya(jm, size(xya, 3))
    call chk(c_ya_intlon(plan, xya, ya_IntLon_xya, int(size(xya, 3), c_int)), 'ya_IntLon_xya')
  end function ya_IntLon_xya

  function ya_AvrLon_xya(xya)
    real(DP), intent(in) :: xya(:,:,:)
    real(DP) :: ya_AvrLon_xya(jm, size(xya, 3))
    call chk(c_ya_avrlon(plan, xya, ya_AvrLon_xya, int(size(xya, 3), c_int)), 'ya_AvrLon_xya')
  end function ya_AvrLon_xya

  ! ---------------------------------------------------- point evaluation
  real(DP) function Interpolate_w(w, alon, alat)
    real(DP), intent(in) :: w(lm), alon, alat
    real(DP) :: lon(1), lat(1), out(1)
    lon(1) = alon; lat(1) = alat
    call chk(c_interpolate(plan, w, lon, lat, out, 1_c_int), 'Interpolate_w')
    Interpolate_w = out(1)
  end function Interpolate_w

  function a_Interpolate_wa(wa, alon, alat)
    real(DP), intent(in) :: wa(:,:), alon, alat
    real(DP) :: a_Interpolate_wa(size(wa, 2))
    integer :: k
    do k = 1, size(wa, 2)
      a_Interpolate_wa(k) = Interpolate_w(wa(:, k), alon, alat)
    end do
  end function a_Interpolate_wa

  ! ---------------------------------------------------------------- vertical column kernels
  ! Fields are xyz(0:im-1, jm, nk) (layer 1 on top = KS), half-level fields xyr(0:im-1, jm, nk+1)
  ! (level 1 = KS-1); z_CDK(nk), z_RFDK(nk+1), z_FDK(nk+1) as in the reference's grid module.

  ! Shape checks of the column wrappers: every field is (0:im-1, jm, nk) without halos, and
  ! the layer vectors have nk (z_CDK) or nk+1 (z_FDK, z_RFDK) entries.  The reference's
  ! KA-sized halo arrays must be passed as their KS:KE (KS-1:KE for half levels) sections.
  subroutine col_shape(where, nk, f, nlev)
    character(*), intent(in) :: where
    integer, intent(in) :: nk, nlev
    real(DP), intent(in) :: f(:,:,:)
    character(len=160) :: msg
    if (size(f, 1) /= im .or. size(f, 2) /= jm .or. size(f, 3) /= nlev) then
      write(msg, '(a,3(i0,1x),a,3(i0,1x))') 'field shape ', size(f, 1), size(f, 2), size(f, 3), &
           & '/= im jm levels ', im, jm, nlev
      call arg_error(where, trim(msg))
    end if
    if (nk < 1) call arg_error(where, 'no layers')
  end subroutine col_shape

  subroutine col_vec(where, name, v, n)
    character(*), intent(in) :: where, name
    real(DP), intent(in) :: v(:)
    integer, intent(in) :: n
    character(len=160) :: msg
    if (size(v) /= n) then
      write(msg, '(a,a,i0,a,i0)') name, ' has ', size(v), ' entries, expected ', n
      call arg_error(where, trim(msg))
    end if
  end subroutine col_vec

  ! HBEDiagnose_OMG (HBEDiagnose_hspm_vfvm_mod.f90:90-143)
  subroutine calc_OMG_xyr(xyr_OMG, xyz_Div, xyz_H, z_CDK)
    real(DP), intent(in) :: xyz_Div(:,:,:), xyz_H(:,:,:), z_CDK(:)
    real(DP), intent(out) :: xyr_OMG(0:im-1, jm, size(xyz_Div, 3) + 1)
    integer :: nk
    nk = size(xyz_Div, 3)
    call col_shape('calc_OMG_xyr', nk, xyz_Div, nk)
    call col_shape('calc_OMG_xyr', nk, xyz_H, nk)
    call col_vec('calc_OMG_xyr', 'z_CDK', z_CDK, nk)
    call chk(c_col_omega(plan, xyz_Div, xyz_H, int(size(xyz_Div, 3), c_int), z_CDK, xyr_OMG), 'calc_OMG_xyr')
  end subroutine calc_OMG_xyr

  ! HBEDiagnose_HydPres (HBEDiagnose_hspm_vfvm_mod.f90:290-325)
  subroutine calc_HydPres_xyz(xyz_HydPres, xyz_DensEdd, xyz_H, z_CDK, Grav)
    real(DP), intent(in) :: xyz_DensEdd(:,:,:), xyz_H(:,:,:), z_CDK(:), Grav
    real(DP), intent(out) :: xyz_HydPres(0:im-1, jm, size(xyz_DensEdd, 3))
    integer :: nk
    nk = size(xyz_DensEdd, 3)
    call col_shape('calc_HydPres_xyz', nk, xyz_DensEdd, nk)
    call col_shape('calc_HydPres_xyz', nk, xyz_H, nk)
    call col_vec('calc_HydPres_xyz', 'z_CDK', z_CDK, nk)
    call chk(c_col_hydpres(plan, xyz_DensEdd, xyz_H, int(size(xyz_DensEdd, 3), c_int), z_CDK, Grav, xyz_HydPres), &
             'calc_HydPres_xyz')
  end subroutine calc_HydPres_xyz

  ! HBEDiagnose_UVBarot (HBEDiagnose_hspm_vfvm_mod.f90:221-255)
  subroutine calc_UVBarot_xy(xy_UBarot, xy_VBarot, xyz_U, xyz_V, xyz_H, z_CDK)
    real(DP), intent(in) :: xyz_U(:,:,:), xyz_V(:,:,:), xyz_H(:,:,:), z_CDK(:)
    real(DP), intent(out) :: xy_UBarot(0:im-1, jm), xy_VBarot(0:im-1, jm)
    integer :: nk
    nk = size(xyz_U, 3)
    call col_shape('calc_UVBarot_xy', nk, xyz_U, nk)
    call col_shape('calc_UVBarot_xy', nk, xyz_V, nk)
    call col_shape('calc_UVBarot_xy', nk, xyz_H, nk)
    call col_vec('calc_UVBarot_xy', 'z_CDK', z_CDK, nk)
    call chk(c_col_uvbarot(plan, xyz_U, xyz_V, xyz_H, int(size(xyz_U, 3), c_int), z_CDK, xy_UBarot, xy_VBarot), &
             'calc_UVBarot_xy')
  end subroutine calc_UVBarot_xy

  ! calc_ADVFlxZ_QUICK (HBEBaroc_hspm_vfvm_mod.f90:233-317)
  subroutine calc_ADVFlxZ_QUICK_xyr(xyr_ADVFlxZ, xyr_OMG, xyz_TRC, z_CDK, z_RFDK, LinFreeSurf)
    real(DP), intent(in) :: xyr_OMG(:,:,:), xyz_TRC(:,:,:), z_CDK(:), z_RFDK(:)
    logical, intent(in) :: LinFreeSurf
    real(DP), intent(out) :: xyr_ADVFlxZ(0:im-1, jm, size(xyz_TRC, 3) + 1)
    integer(c_int) :: lin
    integer :: nk
    nk = size(xyz_TRC, 3)
    call col_shape('calc_ADVFlxZ_QUICK_xyr', nk, xyz_TRC, nk)
    call col_shape('calc_ADVFlxZ_QUICK_xyr', nk, xyr_OMG, nk + 1)
    call col_vec('calc_ADVFlxZ_QUICK_xyr', 'z_CDK', z_CDK, nk)
    call col_vec('calc_ADVFlxZ_QUICK_xyr', 'z_RFDK', z_RFDK, nk + 1)
    lin = 0
    if (LinFreeSurf) lin = 1
    call chk(c_col_quick(plan, xyr_OMG, xyz_TRC, int(size(xyz_TRC, 3), c_int), z_CDK, z_RFDK, lin, xyr_ADVFlxZ), &
             'calc_ADVFlxZ_QUICK_xyr')
  end subroutine calc_ADVFlxZ_QUICK_xyr

  ! calc_VDiffTRCEq (DOGCM_Phys_hspm_vfvm_mod.f90:752-861), or calc_VDiffMOMEq (:865-975) when
  ! xy_BtmFrictCoef is present
  subroutine calc_VDiff_xyz(xyz_qA, xyz_q0, xyz_RHS, xyz_Kv, xyz_H, z_CDK, z_FDK, dt, alpha, &
                            SeaSfcBCType, SeaBtmBCType, xy_BtmFrictCoef)
    real(DP), intent(in) :: xyz_q0(:,:,:), xyz_RHS(:,:,:), xyz_Kv(:,:,:), xyz_H(:,:,:), z_CDK(:), z_FDK(:)
    real(DP), intent(in) :: dt, alpha
    character, intent(in) :: SeaSfcBCType, SeaBtmBCType
    real(DP), intent(in), target, optional :: xy_BtmFrictCoef(0:im-1, jm)
    real(DP), intent(out) :: xyz_qA(0:im-1, jm, size(xyz_q0, 3))
    type(c_ptr) :: bf
    integer :: nk
    nk = size(xyz_q0, 3)
    call col_shape('calc_VDiff_xyz', nk, xyz_q0, nk)
    call col_shape('calc_VDiff_xyz', nk, xyz_RHS, nk)
    call col_shape('calc_VDiff_xyz', nk, xyz_Kv, nk)
    call col_shape('calc_VDiff_xyz', nk, xyz_H, nk)
    call col_vec('calc_VDiff_xyz', 'z_CDK', z_CDK, nk)
    call col_vec('calc_VDiff_xyz', 'z_FDK', z_FDK, nk + 1)
    bf = c_null_ptr
    if (present(xy_BtmFrictCoef)) bf = c_loc(xy_BtmFrictCoef)
    call chk(c_col_vdiff(plan, xyz_q0, xyz_RHS, xyz_Kv, xyz_H, int(size(xyz_q0, 3), c_int), z_CDK, z_FDK, dt, &
                         alpha, int(ichar(SeaSfcBCType), c_int), int(ichar(SeaBtmBCType), c_int), bf, xyz_qA), &
             'calc_VDiff_xyz')
  end subroutine calc_VDiff_xyz

  ! Energy and enstrophy spectra of streamfunctions (cascaded from SPMODEL,
  ! SpmlUtil_mod.f90:131-132), on the unit sphere: E(n,+-m) = 1/2 n(n+1) c_m w(n,+-m)**2 and
  ! the enstrophy with (n(n+1))**2, c_0 = 1, c_m = 2; the sums over (n, m) are the global means
  ! of |grad psi|**2/2 and (lapla psi)**2/2.  nma(0:nm, -nm:nm, :) is zero where |m| > n.
  function nma_EnergyFromStreamfunc_wa(wa_StrFunc) result(nma)
    real(DP), intent(in) :: wa_StrFunc(:,:)
    real(DP) :: nma(0:nm, -nm:nm, size(wa_StrFunc, 2))
    call spec_nma(0, wa_StrFunc, nma, 'nma_EnergyFromStreamfunc_wa')
  end function nma_EnergyFromStreamfunc_wa

  function nma_EnstrophyFromStreamfunc_wa(wa_StrFunc) result(nma)
    real(DP), intent(in) :: wa_StrFunc(:,:)
    real(DP) :: nma(0:nm, -nm:nm, size(wa_StrFunc, 2))
    call spec_nma(1, wa_StrFunc, nma, 'nma_EnstrophyFromStreamfunc_wa')
  end function nma_EnstrophyFromStreamfunc_wa

  function na_EnergyFromStreamfunc_wa(wa_StrFunc) result(na)
    real(DP), intent(in) :: wa_StrFunc(:,:)
    real(DP) :: na(0:nm, size(wa_StrFunc, 2))
    call spec_na(0, wa_StrFunc, na, 'na_EnergyFromStreamfunc_wa')
  end function na_EnergyFromStreamfunc_wa

  function na_EnstrophyFromStreamfunc_wa(wa_StrFunc) result(na)
    real(DP), intent(in) :: wa_StrFunc(:,:)
    real(DP) :: na(0:nm, size(wa_StrFunc, 2))
    call spec_na(1, wa_StrFunc, na, 'na_EnstrophyFromStreamfunc_wa')
  end function na_EnstrophyFromStreamfunc_wa

  subroutine spec_nma(kind, wa, out, where)
    integer, intent(in) :: kind
    real(DP), intent(in) :: wa(:,:)
    real(DP), intent(out) :: out(:,:,:)
    character(*), intent(in) :: where
    real(DP), allocatable, target :: buf(:,:,:)
    if (size(wa, 1) /= lm) call arg_error(where, 'spectral arrays must have lm rows')
    allocate(buf(0:nm, -nm:nm, size(wa, 2)))
    call chk(c_spec_energy(plan, int(kind, c_int), wa, int(size(wa, 2), c_int), c_loc(buf), c_null_ptr), where)
    out = buf
  end subroutine spec_nma

  subroutine spec_na(kind, wa, out, where)
    integer, intent(in) :: kind
    real(DP), intent(in) :: wa(:,:)
    real(DP), intent(out) :: out(:,:)
    character(*), intent(in) :: where
    real(DP), allocatable, target :: buf(:,:)
    if (size(wa, 1) /= lm) call arg_error(where, 'spectral arrays must have lm rows')
    allocate(buf(0:nm, size(wa, 2)))
    call chk(c_spec_energy(plan, int(kind, c_int), wa, int(size(wa, 2), c_int), c_null_ptr, c_loc(buf)), where)
    out = buf
  end subroutine spec_na

  ! LPhys_RediGM_hspm_vfvm_AddMixingTerm (LPhys_RediGM_hspm_vfvm_mod.f90:218-372): adds the
  ! Redi and GM tendencies of potential temperature and salinity to xyz_PTemp_RHS and
  ! xyz_Salt_RHS.  Layers KS:KE of the fields, xy_Topo, z_CDK(KS:KE) and z_RFDK(KS-1:KE).
  ! Parameters default to RediGM_nml's defaults of APEOGCircExp.nml (DiffCoefRedi = DiffCoefGM
  ! = 800, InteriorTaperingName = 'DM95', PBLTaperingName = 'LDD95', SlopeMax = 4e-3,
  ! Sd = 1e-3) and Earth's Omega, RefDens 1027, Grav 9.79764.
  subroutine calc_RediGM_xyz(xyz_PTemp_RHS, xyz_Salt_RHS, xyz_PTemp, xyz_Salt, xyz_H, xyz_Z, xy_Topo, &
                             z_CDK, z_RFDK, DiffCoefRedi, DiffCoefGM, SlopeMax, Sd, InteriorTaperingName, &
                             Omega, RefDens, Grav)
    real(DP), intent(inout) :: xyz_PTemp_RHS(:,:,:), xyz_Salt_RHS(:,:,:)
    real(DP), intent(in) :: xyz_PTemp(:,:,:), xyz_Salt(:,:,:), xyz_H(:,:,:), xyz_Z(:,:,:)
    real(DP), intent(in) :: xy_Topo(0:im-1, jm), z_CDK(:), z_RFDK(:)
    real(DP), intent(in), optional :: DiffCoefRedi, DiffCoefGM, SlopeMax, Sd, Omega, RefDens, Grav
    character(*), intent(in), optional :: InteriorTaperingName
    type(redigm_config_t) :: cfg
    integer :: nk
    nk = size(xyz_PTemp, 3)
    call col_shape('calc_RediGM_xyz', nk, xyz_PTemp, nk)
    call col_shape('calc_RediGM_xyz', nk, xyz_Salt, nk)
    call col_shape('calc_RediGM_xyz', nk, xyz_H, nk)
    call col_shape('calc_RediGM_xyz', nk, xyz_Z, nk)
    call col_shape('calc_RediGM_xyz', nk, xyz_PTemp_RHS, nk)
    call col_shape('calc_RediGM_xyz', nk, xyz_Salt_RHS, nk)
    call col_vec('calc_RediGM_xyz', 'z_CDK', z_CDK, nk)
    call col_vec('calc_RediGM_xyz', 'z_RFDK', z_RFDK, nk + 1)
    call chk(c_redigm_default_config(cfg), 'calc_RediGM_xyz')
    if (present(DiffCoefRedi)) cfg%kappa_redi = DiffCoefRedi
    if (present(DiffCoefGM)) cfg%kappa_gm = DiffCoefGM
    if (present(SlopeMax)) cfg%slope_max = SlopeMax
    if (present(Sd)) cfg%sd = Sd
    if (present(Omega)) cfg%omega = Omega
    if (present(RefDens)) cfg%ref_dens = RefDens
    if (present(Grav)) cfg%grav = Grav
    if (present(InteriorTaperingName)) then
      select case (trim(InteriorTaperingName))
      case ('DM95')
        cfg%interior_taper = 2
      case ('GKW91')
        cfg%interior_taper = 1
      case default
        call arg_error('calc_RediGM_xyz', 'InteriorTaperingName must be DM95 or GKW91')
      end select
    end if
    call chk(c_redigm(plan, cfg, xyz_PTemp, xyz_Salt, xyz_H, xyz_Z, int(nk, c_int), xy_Topo, z_CDK, z_RFDK, &
                      xyz_PTemp_RHS, xyz_Salt_RHS), 'calc_RediGM_xyz')
  end subroutine calc_RediGM_xyz

end module SpmlUtil_mod
