!-------------------------------------------------------------
! Drop-in test: the SpmlUtil_mod API (the reference's names and shapes) driven from
! Fortran through ISO_C_BINDING onto the GPU library.  Checks closed forms and
! invariants at T21 (im = 64, jm = 32) with an Earth radius:
!   round trips, Laplacian eigenvalues, solid-body rotation, velocity <-> vorticity and
!   divergence, global mean = s_0^0, multi-layer forms, and concurrent OpenMP-style
!   per-layer calls (as in SpmlUtil_mod.f90:360-366).
! Prints PASS and exits 0, or stops with status 1.
!-------------------------------------------------------------
program spml_dropin_test
  use SpmlUtil_mod
  implicit none
  integer, parameter :: DP = kind(1d0)
  integer, parameter :: IM = 64, JM = 32, KM = 4, NM = 21
  real(DP), parameter :: R = 6.371d6, U0 = 10d0
  real(DP) :: w(( NM + 1)**2), w2((NM + 1)**2), vor((NM + 1)**2), div((NM + 1)**2)
  real(DP) :: xy(0:IM-1, JM), u(0:IM-1, JM), v(0:IM-1, JM)
  real(DP) :: wa((NM + 1)**2, KM), wa2((NM + 1)**2, KM), xya(0:IM-1, JM, KM)
  real(DP) :: xya_lon(0:IM-1, JM, KM), xya_lat(0:IM-1, JM, KM)
  real(DP) :: err, ref
  integer :: l, n, m, k, nfail
  integer :: nmv(2)

  nfail = 0
  call SpmlUtil_Init(IM, JM, KM, NM, 0, R)
  if (.not. isInitialzed()) call fail('init')

  ! red-spectrum coefficients (deterministic)
  do l = 1, (NM + 1)**2
    nmv = nm_l(l)
    w(l) = sin(0.37d0 * l) / (nmv(1) + 1d0)**1.5d0
  end do
  if (l_nm(3, -2) /= 3*4 - 2 + 1) call fail('l_nm')

  xy = xy_w(w)
  w2 = w_xy(xy)
  err = maxval(abs(w2 - w))
  call check('round trip w -> xy -> w', err, 1d-13)

  w2 = w_Lapla_w(w)
  err = 0d0
  do l = 1, (NM + 1)**2
    nmv = nm_l(l)
    err = max(err, abs(w2(l) + nmv(1) * (nmv(1) + 1) * w(l)))
  end do
  call check('w_Lapla_w = -n(n+1) w', err, 1d-12)
  w2 = w_LaplaInv_w(w_Lapla_w(w))
  err = maxval(abs(w2(2:) - w(2:)))
  call check('LaplaInv(Lapla(w)) = w (n > 0)', err, 1d-13)

  ! solid-body rotation: u = U0 cos(phi), v = 0 -> vor = 2 U0 sin(phi)/R
  u = U0 * xy_CosLat * xy_CosLat
  v = 0d0
  call calc_UVCosLat2VorDiv(u, v, vor, div)
  ref = 2d0 * U0 / (sqrt(3d0) * R)
  vor(l_nm(1, 0)) = vor(l_nm(1, 0)) - ref
  call check('solid-body vorticity', maxval(abs(vor)) / ref, 1d-12)
  call check('solid-body divergence', maxval(abs(div)) / ref, 1d-12)

  ! vorticity/divergence -> (u, v) -> back
  vor = w * 1d-5
  div = cshift(w, 3) * 1d-6
  vor(1) = 0d0; div(1) = 0d0
  call calc_VorDiv2UV(vor, div, u, v)
  call calc_UVCosLat2VorDiv(u * xy_CosLat, v * xy_CosLat, w2, wa(:, 1))
  call check('vor round trip', maxval(abs(w2 - vor)) / maxval(abs(vor)), 1d-12)
  call check('div round trip', maxval(abs(wa(:, 1) - div)) / maxval(abs(div)), 1d-12)

  ! global mean of the grid field is s_0^0
  call check('AvrLonLat_xy = w(1)', abs(AvrLonLat_xy(xy_w(w)) - w(1)), 1d-14)
  call check('IntLonLat_xy = 4 pi w(1)', abs(IntLonLat_xy(xy_w(w)) - 4d0 * acos(-1d0) * w(1)), 1d-13)
  call check('Interpolate_w at a node', abs(Interpolate_w(w, x_Lon(5), y_Lat(7)) - xy(5, 7)), 1d-13)

  ! multi-layer forms
  do k = 1, KM
    wa(:, k) = w * k
  end do
  xya = xya_wa(wa)
  wa2 = wa_xya(xya)
  call check('xya_wa / wa_xya', maxval(abs(wa2 - wa)), 1d-12)
  xya = xya_GradLat_wa(wa)
  err = 0d0
  do k = 1, KM
    err = max(err, maxval(abs(xya(:, :, k) - xy_GradLat_w(wa(:, k)))))
  end do
  call check('xya_GradLat_wa per layer', err, 1d-12)
  ! fused gradients against the separate calls
  xya = xya_GradLat_wa(wa)
  call calc_GradLonLat_xya_wa(wa, xya_lon, xya_lat)
  call check('calc_GradLonLat_xya_wa lat', maxval(abs(xya_lat - xya)), 1d-12)
  xya = xya_GradLon_wa(wa)
  call check('calc_GradLonLat_xya_wa lon', maxval(abs(xya_lon - xya)), 1d-12)

  ! per-layer calls from concurrent threads (each OpenMP thread gets its own stream)
  !$omp parallel do private(k)
  do k = 1, KM
    xya(:, :, k) = xy_w(wa(:, k))
  end do
  !$omp end parallel do
  err = 0d0
  do k = 1, KM
    err = max(err, maxval(abs(w_xy(xya(:, :, k)) - wa(:, k))))
  end do
  call check('per-layer calls', err, 1d-12)

  call SpmlUtil_Final()
  if (nfail > 0) then
    write(*, '(a,i0,a)') 'FAIL (', nfail, ' checks)'
    error stop 1
  end if
  write(*, '(a)') 'PASS'

contains

  subroutine check(what, val, tol)
    character(*), intent(in) :: what
    real(DP), intent(in) :: val, tol
    if (val <= tol) then
      write(*, '(a,a,es10.3)') what, ': ok ', val
    else
      write(*, '(a,a,es10.3,a,es10.3)') what, ': FAILED ', val, ' > ', tol
      nfail = nfail + 1
    end if
  end subroutine check

  subroutine fail(what)
    character(*), intent(in) :: what
    write(*, '(a,a)') 'FAILED: ', what
    nfail = nfail + 1
  end subroutine fail

end program spml_dropin_test
