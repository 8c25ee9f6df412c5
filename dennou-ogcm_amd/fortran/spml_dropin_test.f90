!-------------------------------------------------------------
! Drop-in test: the SpmlUtil_mod API (the reference's names and shapes) driven from
! Fortran through ISO_C_BINDING onto the GPU library.  Checks closed forms and
! invariants at T21 (im = 64, jm = 32) with an Earth radius:
!   round trips, Laplacian eigenvalues, solid-body rotation, velocity <-> vorticity and
!   divergence, global mean = s_0^0, multi-layer forms, the vertical grid of
!   get_SpmlGridInfo (Clenshaw-Curtis exactness on [-1, 0]), and per-layer calls from 4
!   concurrent OpenMP threads (as in SpmlUtil_mod.f90:360-366), bit for bit equal to the
!   serial calls.
! Prints PASS and exits 0, or stops with status 1.
! With the argument "badsize" it passes a halo-sized (KA) z_CDK to a column wrapper after
! the checks, which must stop with the shim's shape error (status 1).
!-------------------------------------------------------------
program spml_dropin_test
  use SpmlUtil_mod
  !$ use omp_lib
  implicit none
  integer, parameter :: DP = kind(1d0)
  integer, parameter :: IM = 64, JM = 32, KM = 4, NM = 21
  real(DP), parameter :: R = 6.371d6, U0 = 10d0
  real(DP) :: w(( NM + 1)**2), w2((NM + 1)**2), vor((NM + 1)**2), div((NM + 1)**2)
  real(DP) :: xy(0:IM-1, JM), u(0:IM-1, JM), v(0:IM-1, JM)
  real(DP) :: wa((NM + 1)**2, KM), wa2((NM + 1)**2, KM), xya(0:IM-1, JM, KM)
  real(DP) :: xya_lon(0:IM-1, JM, KM), xya_lat(0:IM-1, JM, KM)
  real(DP) :: err, ref
  real(DP) :: z_sig(0:KM), z_sigw(0:KM), par_xy(0:IM-1, JM, 2*KM), par_w((NM + 1)**2, 2*KM)
  real(DP) :: ser_xy(0:IM-1, JM, 2*KM), ser_w((NM + 1)**2, 2*KM), wl((NM + 1)**2, 2*KM)
  integer :: l, n, m, k, nfail, tid(2*KM), nthr
  integer :: nmv(2)
  character(len=16) :: arg

  nfail = 0
  call SpmlUtil_Init(IM, JM, KM, NM, KM, R)
  if (.not. isInitialzed()) call fail('init')

  ! vertical grid: Chebyshev-Lobatto levels of [-1, 0] from the top, Clenshaw-Curtis weights
  ! exact for polynomials of degree <= tMax = KM
  call get_SpmlGridInfo(z_Sig_=z_sig, z_Sig_Weight_=z_sigw)
  err = 0d0
  do k = 0, KM
    err = max(err, abs(z_sig(k) - (-0.5d0 + 0.5d0 * cos(acos(-1d0) * k / KM))))
  end do
  call check('g_Sig (Chebyshev-Lobatto levels)', err, 1d-15)
  err = 0d0
  do n = 0, KM
    err = max(err, abs(sum(z_sigw * z_sig**n) - (0d0 - (-1d0)**(n + 1)) / (n + 1)))
  end do
  call check('g_Sig_WEIGHT integrates sigma**n exactly', err, 1d-15)
  call check('g_Sig_WEIGHT public array', maxval(abs(g_Sig_WEIGHT - z_sigw)), 0d0)

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

  ! per-layer calls from 4 concurrent threads (each OpenMP thread gets its own stream), as
  ! the reference's !$omp parallel do over layers makes them; every result must equal the
  ! serial call bit for bit
  do k = 1, 2*KM
    wl(:, k) = cshift(w, k) * (1d0 + 0.1d0 * k)
    ser_xy(:, :, k) = xy_GradLat_w(wl(:, k))
    ser_w(:, k) = w_DivMu_xy(ser_xy(:, :, k))
  end do
  tid = -1
  nthr = 1
  !$omp parallel do private(k) num_threads(4) schedule(static, 1)
  do k = 1, 2*KM
    !$ tid(k) = omp_get_thread_num()
    par_xy(:, :, k) = xy_GradLat_w(wl(:, k))
    par_w(:, k) = w_DivMu_xy(par_xy(:, :, k))
  end do
  !$omp end parallel do
  !$ nthr = count([(any(tid == k), k = 0, 3)])
  write(*, '(a,i0)') 'threads that made per-layer calls: ', nthr
  if (nthr < 4) call fail('per-layer calls did not run on 4 threads')
  call check('per-layer calls from 4 threads = serial (grid)', maxval(abs(par_xy - ser_xy)), 0d0)
  call check('per-layer calls from 4 threads = serial (spectral)', maxval(abs(par_w - ser_w)), 0d0)
  !$omp parallel do private(k) num_threads(4)
  do k = 1, KM
    xya(:, :, k) = xy_w(wa(:, k))
  end do
  !$omp end parallel do
  err = 0d0
  do k = 1, KM
    err = max(err, maxval(abs(w_xy(xya(:, :, k)) - wa(:, k))))
  end do
  call check('per-layer calls', err, 1d-12)

  call column_checks()
  call redigm_checks()
  call spectrum_checks()

  if (command_argument_count() > 0) then
    call get_command_argument(1, arg)
    if (trim(arg) == 'badsize') call bad_size()
  end if

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

  ! the column kernels through the shim's extensions, on closed forms (uniform layers):
  ! omega of a constant divergence falls linearly to 0 at the bottom; the pressure of a
  ! constant density sits at the layer centres; the barotropic mean of a constant is that
  ! constant; the QUICK flux of a constant tracer is omega*T; without diffusivity the solve
  ! adds dt*rhs/H (tracer form) or dt*rhs/(1 + alpha*dt/(H*cdk)*bfric) at a Dirichlet
  ! bottom (momentum form)
  subroutine column_checks()
    integer, parameter :: NK = 5
    real(DP), parameter :: D0 = 2d-6, H0 = 1.5d0, RHO = 1.2d0, G = 9.8d0, T0 = 4d0, DT = 600d0
    real(DP) :: div(0:IM-1, JM, NK), h(0:IM-1, JM, NK), omg(0:IM-1, JM, NK+1), flx(0:IM-1, JM, NK+1)
    real(DP) :: p(0:IM-1, JM, NK), q0(0:IM-1, JM, NK), rhs(0:IM-1, JM, NK), kv(0:IM-1, JM, NK), qa(0:IM-1, JM, NK)
    real(DP) :: ub(0:IM-1, JM), vb(0:IM-1, JM), bf(0:IM-1, JM)
    real(DP) :: cdk(NK), rfdk(NK+1), fdk(NK+1), e1, e2, want
    integer :: kk
    cdk = 1d0 / NK
    rfdk = dble(NK)
    fdk = 1d0 / NK
    div = D0
    h = H0
    call calc_OMG_xyr(omg, div, h, cdk)
    e1 = 0d0
    do kk = 1, NK + 1
      e1 = max(e1, maxval(abs(omg(:, :, kk) + D0 * H0 * dble(NK + 1 - kk) / NK)))
    end do
    call check('calc_OMG_xyr (linear omega)', e1 / (D0 * H0), 1d-14)
    p = RHO
    call calc_HydPres_xyz(p, p * 0d0 + RHO, h, cdk, G)
    e1 = 0d0
    do kk = 1, NK
      e1 = max(e1, maxval(abs(p(:, :, kk) - G * H0 * RHO * (kk - 0.5d0) / NK)))
    end do
    call check('calc_HydPres_xyz (layer centres)', e1 / (G * H0 * RHO), 1d-14)
    q0 = 0.3d0
    rhs = -0.7d0
    call calc_UVBarot_xy(ub, vb, q0, rhs, h, cdk)
    call check('calc_UVBarot_xy (constant)', max(maxval(abs(ub - 0.3d0)), maxval(abs(vb + 0.7d0))), 1d-15)
    q0 = T0
    call calc_ADVFlxZ_QUICK_xyr(flx, omg, q0, cdk, rfdk, .false.)
    e1 = max(maxval(abs(flx(:, :, 1))), maxval(abs(flx(:, :, NK + 1))))
    e1 = max(e1, maxval(abs(flx(:, :, 2:NK) - T0 * omg(:, :, 2:NK))))
    call check('calc_ADVFlxZ_QUICK_xyr (constant tracer)', e1 / (D0 * H0 * T0), 1d-14)
    kv = 0d0
    rhs = 1d-4
    call calc_VDiff_xyz(qa, q0, rhs, kv, h, cdk, fdk, DT, 1d0, 'D', 'N')
    call check('calc_VDiff_xyz (no diffusivity)', maxval(abs(qa - (q0 + DT * rhs / h))), 1d-15)
    bf = 3d-3
    call calc_VDiff_xyz(qa, q0, rhs, kv, h, cdk, fdk, DT, 1d0, 'D', 'D', bf)
    want = T0 + DT * 1d-4 / (1d0 + 1d0 * (DT / (H0 * cdk(NK))) * 3d-3)
    e2 = maxval(abs(qa(:, :, NK) - want))
    e1 = maxval(abs(qa(:, :, 1:NK-1) - (T0 + DT * 1d-4)))
    call check('calc_VDiff_xyz momentum form (bottom friction)', max(e1, e2), 1d-14)
  end subroutine column_checks

  ! the Redi/GM mixing through calc_RediGM_xyz: a flat-bottomed, stratified ocean with a
  ! tilted temperature field gets a nonzero tendency whose tracer content is zero (the
  ! horizontal divergence has no global mean and the vertical flux divergence telescopes);
  ! a horizontally uniform ocean has no isopycnal slope and gets no tendency
  subroutine redigm_checks()
    integer, parameter :: NK = 6
    real(DP), parameter :: DEPTH = 4000d0
    real(DP) :: pt(0:IM-1, JM, NK), sa(0:IM-1, JM, NK), h(0:IM-1, JM, NK), z(0:IM-1, JM, NK)
    real(DP) :: prhs(0:IM-1, JM, NK), srhs(0:IM-1, JM, NK), topo(0:IM-1, JM)
    real(DP) :: fk(0:NK), ck(0:NK+1), cdk(NK), rfdk(NK+1), content, scale
    integer :: kk, i, j
    do kk = 0, NK
      fk(kk) = -0.5d0 + 0.5d0 * cos(acos(-1d0) * kk / NK)
    end do
    do kk = 1, NK
      cdk(kk) = fk(kk - 1) - fk(kk)
      ck(kk) = 0.5d0 * (fk(kk - 1) + fk(kk))
    end do
    ck(0) = fk(0) + 0.5d0 * cdk(1)
    ck(NK + 1) = fk(NK) - 0.5d0 * cdk(NK)
    do kk = 1, NK + 1
      rfdk(kk) = 1d0 / (ck(kk - 1) - ck(kk))
    end do
    topo = DEPTH
    do kk = 1, NK
      h(:, :, kk) = DEPTH
      z(:, :, kk) = DEPTH * ck(kk)
      do j = 1, JM
        do i = 0, IM - 1
          pt(i, j, kk) = 275d0 + 20d0 * exp(z(i, j, kk) / 800d0) &
               & + 3d0 * cos(y_Lat(j))**2 * cos(x_Lon(i)) * exp(z(i, j, kk) / 1500d0)
          sa(i, j, kk) = 34.5d0 + 0.5d0 * exp(z(i, j, kk) / 1000d0) + 0.2d0 * sin(y_Lat(j)) * exp(z(i, j, kk) / 2000d0)
        end do
      end do
    end do
    prhs = 0d0
    srhs = 0d0
    call calc_RediGM_xyz(prhs, srhs, pt, sa, h, z, topo, cdk, rfdk)
    content = 0d0
    scale = 0d0
    do kk = 1, NK
      content = content + cdk(kk) * AvrLonLat_xy(prhs(:, :, kk))
      scale = scale + cdk(kk) * AvrLonLat_xy(abs(prhs(:, :, kk)))
    end do
    call check('calc_RediGM_xyz (nonzero tendency)', merge(0d0, 1d0, scale > 0d0), 0d0)
    call check('calc_RediGM_xyz (tracer content conserved)', abs(content) / scale, 1d-11)
    do kk = 1, NK
      pt(:, :, kk) = 275d0 + 20d0 * exp(z(:, :, kk) / 800d0)
      sa(:, :, kk) = 35d0
    end do
    prhs = 0d0
    srhs = 0d0
    call calc_RediGM_xyz(prhs, srhs, pt, sa, h, z, topo, cdk, rfdk, InteriorTaperingName='GKW91')
    call check('calc_RediGM_xyz (no slope, no tendency)', max(maxval(abs(prhs)), maxval(abs(srhs))), 1d-20)
  end subroutine redigm_checks

  ! energy and enstrophy spectra: for psi = a P_1^0 + b P_3^2 cos(2 lambda) the spectra hold
  ! a**2 at (1, 0) and 1/2 * 12 * 2 * (b/2)**2 at (3, 2) (s_3^2 = b/2), and their sums equal the
  ! global means of |grad psi|**2/2 and (lapla psi)**2/2 on the grid
  subroutine spectrum_checks()
    real(DP) :: ws((NM + 1)**2, 1), en(0:NM, -NM:NM, 1), ena(0:NM, 1), zn(0:NM, -NM:NM, 1), zna(0:NM, 1)
    real(DP) :: glon(0:IM-1, JM), glat(0:IM-1, JM), zeta(0:IM-1, JM), e_grid, z_grid
    real(DP), parameter :: a = 0.7d0, b = -0.4d0
    ws = 0d0
    ws(l_nm(1, 0), 1) = a
    ws(l_nm(3, 2), 1) = b / 2d0
    en = nma_EnergyFromStreamfunc_wa(ws)
    ena = na_EnergyFromStreamfunc_wa(ws)
    zn = nma_EnstrophyFromStreamfunc_wa(ws)
    zna = na_EnstrophyFromStreamfunc_wa(ws)
    call check('nma_EnergyFromStreamfunc_wa (1,0) and (3,2)', &
         max(abs(en(1, 0, 1) - a**2), abs(en(3, 2, 1) - 12d0 * (b / 2d0)**2), &
             abs(sum(en) - en(1, 0, 1) - en(3, 2, 1))), 1d-15)
    call check('na_ = sum over m of nma_', maxval(abs(ena(:, 1) - sum(en(:, :, 1), 2))), 1d-15)
    glon = xy_GradLon_w(ws(:, 1))
    glat = xy_GradLat_w(ws(:, 1))
    e_grid = AvrLonLat_xy(0.5d0 * (glon**2 + glat**2))
    call check('energy spectrum sums to the mean of |grad psi|**2/2', abs(sum(ena) - e_grid) / e_grid, 1d-13)
    zeta = xy_w(w_Lapla_w(ws(:, 1)))
    z_grid = AvrLonLat_xy(0.5d0 * zeta**2)
    call check('enstrophy spectrum sums to the mean of (lapla psi)**2/2', abs(sum(zna) - z_grid) / z_grid, 1d-13)
  end subroutine spectrum_checks

  ! a halo-sized (KA = nk + 2) layer-thickness vector: the shim must refuse it
  subroutine bad_size()
    integer, parameter :: NK = 3
    real(DP) :: f(0:IM-1, JM, NK), qa(0:IM-1, JM, NK), cdk(NK + 2), fdk(NK + 1)
    f = 1d0
    cdk = 1d0
    fdk = 1d0
    write(*, '(a)') 'passing a KA-sized z_CDK'
    call calc_VDiff_xyz(qa, f, f, f, f, cdk, fdk, 1d0, 1d0, 'D', 'N')
    write(*, '(a)') 'BADSIZE NOT REFUSED'
  end subroutine bad_size

  subroutine fail(what)
    character(*), intent(in) :: what
    write(*, '(a,a)') 'FAILED: ', what
    nfail = nfail + 1
  end subroutine fail

end program spml_dropin_test
