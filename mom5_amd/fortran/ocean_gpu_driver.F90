!-----------------------------------------------------------------------
! ocean_gpu_driver -- a Fortran host that stays in charge of the model
! state and advances it through the ISO_C_BINDING boundary (mom_gpu_mod).
!
! It builds a small cyclic-x / closed-y box with an island, fills the
! fields the way MOM's types hold them (Fortran arrays with halos), copies
! them to the device and runs baroclinic steps in update_ocean_model order
! (ocean_core/ocean_model.F90:1441-2108):
!   compute_tmask_limit -> per tracer horz/vert advection, explicit update
!   -> vert_diffuse_implicit -> update_ocean_barotropic.
! It prints the masked tracer content before and after and stops with an
! error if the content drifts by more than 1e-12 (relative).
!-----------------------------------------------------------------------
program ocean_gpu_driver
  use, intrinsic :: iso_c_binding
  use mom_gpu_mod
  implicit none

  integer, parameter :: ni = 48, nj = 32, nk = 10, ntr = 2, nsteps = 4
  integer, parameter :: isc = 1, iec = ni, jsc = 1, jec = nj
  integer, parameter :: isd = 0, ied = ni + 1, jsd = 0, jed = nj + 1
  real(c_double), parameter :: rho0 = 1035.0d0, grav = 9.8d0
  real(c_double), parameter :: dx = 2.0d4, dy = 2.0d4, pi = 3.14159265358979323846d0

  type(mom_gpu_domain_t) :: dom
  type(mom_gpu_grid_t) :: gd
  type(mom_gpu_bt_params_t) :: bp
  type(c_ptr) :: ctx

  real(c_double), target :: dxt(isd:ied,jsd:jed), dyt(isd:ied,jsd:jed), dat(isd:ied,jsd:jed)
  real(c_double), target :: datr(isd:ied,jsd:jed), f2(isd:ied,jsd:jed), dxr(isd:ied,jsd:jed)
  real(c_double), target :: dyr(isd:ied,jsd:jed), fcor(isd:ied,jsd:jed)
  integer(c_int), target :: kmt(isd:ied,jsd:jed), kmu(isd:ied,jsd:jed)
  real(c_double), target :: tmask(isd:ied,jsd:jed,nk), umask(isd:ied,jsd:jed,nk), dzt(nk)
  real(c_double), target :: uh(isd:ied,jsd:jed,nk), vh(isd:ied,jsd:jed,nk)
  real(c_double), target :: wb(isd:ied,jsd:jed,0:nk), rdz(isd:ied,jsd:jed,nk)
  real(c_double), target :: rdzr(isd:ied,jsd:jed,nk), dzw(isd:ied,jsd:jed,0:nk)
  real(c_double), target :: dcbt(isd:ied,jsd:jed,nk,2), tr(isd:ied,jsd:jed,nk,ntr)
  real(c_double), target :: zero2(isd:ied,jsd:jed), eta(isd:ied,jsd:jed)
  real(c_double), target :: ud(isd:ied,jsd:jed,2), thk(isd:ied,jsd:jed), rhog(isd:ied,jsd:jed)
  real(c_double), target :: psi(isd:ied,jsd:jed), mins(ntr), maxs(ntr)

  type(c_ptr) :: d_dxt, d_dyt, d_dat, d_datr, d_f2, d_dxr, d_dyr, d_f, d_kmt, d_kmu
  type(c_ptr) :: d_tmask, d_umask, d_dzt
  type(c_ptr) :: d_uh, d_vh, d_wb, d_rdz, d_rdzr, d_dzw, d_dcbt, d_zero2, d_rhog, d_thk
  type(c_ptr) :: d_eta(2), d_ud(2), d_conv, d_forc
  type(c_ptr), target :: d_T(ntr,2), d_th(ntr), d_lim(ntr), d_stf(ntr), d_btf(ntr)
  type(c_ptr), target :: tab_now(ntr), tab_nxt(ntr), tab_lim(ntr)
  type(c_ptr), target :: d_B(ntr,2), tab_in(ntr), tab_out(ntr)
  integer(c_int), target :: dtypes(ntr)
  type(mom_gpu_tracer_step_t) :: st
  real(c_double) :: errv(7)
  integer :: bnow, bnxt
  real(c_double) :: c0(ntr), c1(ntr), dt, dtbt, rel, gk(nk), gmean
  integer :: i, j, k, n, step, now, nxt
  integer(c_int) :: rc
  integer(c_size_t) :: b2, b3, b3w, bi2

  ! ---- grid: Cartesian box, cyclic in x, walls at j=jsd and j=jed ----
  do k = 1, nk
     dzt(k) = 10.0d0 * 1.5d0**(k-1)
  enddo
  kmt = 0
  do j = jsc, jec
     do i = isc, iec
        kmt(i,j) = nk
        if (i >= 20 .and. i <= 26 .and. j >= 12 .and. j <= 18) kmt(i,j) = 0   ! island
        if (j == jsc .or. j == jec) kmt(i,j) = nk/2
     enddo
  enddo
  kmt(isd,:) = kmt(iec,:)
  kmt(ied,:) = kmt(isc,:)
  kmu = 0
  do j = jsd, jed-1
     do i = isd, ied-1
        kmu(i,j) = min(kmt(i,j), kmt(i+1,j), kmt(i,j+1), kmt(i+1,j+1))
     enddo
  enddo
  kmu(ied,:) = kmu(isc,:)
  do k = 1, nk
     where (kmt >= k)
        tmask(:,:,k) = 1.0d0
     elsewhere
        tmask(:,:,k) = 0.0d0
     end where
     where (kmu >= k)
        umask(:,:,k) = 1.0d0
     elsewhere
        umask(:,:,k) = 0.0d0
     end where
  enddo
  dxt = dx; dyt = dy; dat = dx*dy; datr = 1.0d0/(dx*dy); dxr = 1.0d0/dx; dyr = 1.0d0/dy
  fcor = 1.0d-4; f2 = dx

  ! ---- transports from a streamfunction at U points, w by continuity ----
  ! overturning profile with zero thickness-weighted mean, so that the
  ! column-integrated divergence vanishes and wrho_bt(nk) = 0
  gmean = 0.0d0
  do k = 1, nk
     gk(k) = real(nk/2 - k, 8)
     gmean = gmean + gk(k)*dzt(k)
  enddo
  gk = gk - gmean/sum(dzt)
  psi = 0.0d0
  do j = jsd, jed
     do i = isd, ied
        psi(i,j) = 1.0d4 * sin(2.0d0*pi*real(i,8)/real(ni,8)) * sin(pi*real(j,8)/real(nj,8))
     enddo
  enddo
  uh = 0.0d0; vh = 0.0d0; wb = 0.0d0
  do k = 1, nk
     do j = jsc, jec
        do i = isd, iec
           uh(i,j,k) = rho0*dzt(k)*(psi(i,j)*umask(i,j,k) - psi(i,j-1)*umask(i,j-1,k))/dy &
                + rho0*dzt(k)*1.0d-3*gk(k)*tmask(i,j,k)*tmask(i+1,j,k)*tmask(i,j,nk)*tmask(i+1,j,nk)
        enddo
     enddo
     do j = jsd, jec
        do i = isc, iec
           vh(i,j,k) = -rho0*dzt(k)*(psi(i,j)*umask(i,j,k) - psi(i-1,j)*umask(i-1,j,k))/dx
        enddo
     enddo
     ! halo columns hold the cyclic wrap, as mpp_update_domains leaves them
     uh(isd,:,k) = uh(iec,:,k); uh(ied,:,k) = uh(isc,:,k)
     vh(isd,:,k) = vh(iec,:,k); vh(ied,:,k) = vh(isc,:,k)
  enddo
  do k = 1, nk
     do j = jsc, jec
        do i = isc, iec
           wb(i,j,k) = (datr(i,j)*(dy*uh(i,j,k) - dy*uh(i-1,j,k) + dx*vh(i,j,k) - dx*vh(i,j-1,k)) &
                * tmask(i,j,k) + wb(i,j,k-1))*tmask(i,j,k)
        enddo
     enddo
  enddo
  wb(isd,:,:) = wb(iec,:,:); wb(ied,:,:) = wb(isc,:,:)
  do k = 1, nk
     rdz(:,:,k) = rho0*dzt(k)
     rdzr(:,:,k) = 1.0d0/(rho0*dzt(k) + 1.0d-40)
  enddo
  dzw(:,:,0) = 0.5d0*dzt(1)
  do k = 1, nk-1
     dzw(:,:,k) = 0.5d0*(dzt(k) + dzt(k+1))
  enddo
  dzw(:,:,nk) = 0.5d0*dzt(nk)
  dcbt(:,:,:,1) = 1.0d-4; dcbt(:,:,:,2) = 1.0d-5
  do j = jsd, jed
     do i = isd, ied
        do k = 1, nk
           tr(i,j,k,1) = (2.0d0 + 25.0d0*exp(-sum(dzt(1:k))/800.0d0) + &
                0.5d0*sin(0.3d0*real(i*j+k,8)))*tmask(i,j,k)
           tr(i,j,k,2) = (35.0d0 + 0.2d0*cos(0.7d0*real(i+3*j+k,8)))*tmask(i,j,k)
        enddo
     enddo
  enddo
  zero2 = 0.0d0
  eta = 0.05d0*tmask(:,:,1)*sin(2.0d0*pi*real(spread([(i, i=isd,ied)],2,jed-jsd+1),8)/real(ni,8))
  eta(isd,:) = eta(iec,:); eta(ied,:) = eta(isc,:)
  ud = 0.0d0
  thk = 0.0d0
  do k = 1, nk
     thk = thk + dzt(k)*umask(:,:,k)
  enddo
  rhog = grav*rho0
  mins = [-2.0d0, 1.0d0]; maxs = [32.0d0, 42.0d0]
  dtbt = 20.0d0
  dt = 0.5d0*dtbt*40.0d0

  ! ---- device copies (what MOM's init would upload once) ----
  b2 = int(8*(ied-isd+1)*(jed-jsd+1), c_size_t)
  bi2 = int(4*(ied-isd+1)*(jed-jsd+1), c_size_t)
  b3 = b2*nk; b3w = b2*(nk+1)
  d_dxt = up(c_loc(dxt), b2); d_dyt = up(c_loc(dyt), b2); d_dat = up(c_loc(dat), b2)
  d_datr = up(c_loc(datr), b2); d_f2 = up(c_loc(f2), b2); d_dxr = up(c_loc(dxr), b2)
  d_dyr = up(c_loc(dyr), b2); d_f = up(c_loc(fcor), b2)
  d_kmt = up(c_loc(kmt), bi2); d_kmu = up(c_loc(kmu), bi2)
  d_tmask = up(c_loc(tmask), b3); d_umask = up(c_loc(umask), b3)
  d_dzt = up(c_loc(dzt), int(8*nk, c_size_t))
  dom = mom_gpu_domain_t(isd, ied, jsd, jed, isc, iec, jsc, jec, nk, 1, 0)
  gd = mom_gpu_grid_t(d_dxt, d_dyt, d_dat, d_datr, d_f2, d_f2, d_f2, d_f2, d_f2, d_f2, &
       d_dxr, d_dyr, d_f, d_kmt, d_kmu, d_tmask, d_umask, d_dzt)
  gd%dyte = d_dyt; gd%dytn = d_dyt; gd%dyu = d_dyt
  call mom_gpu_check(mom_gpu_create(dom, gd, ctx), c_null_ptr, 'mom_gpu_create')

  d_uh = up(c_loc(uh), b3); d_vh = up(c_loc(vh), b3); d_wb = up(c_loc(wb), b3w)
  d_rdz = up(c_loc(rdz), b3); d_rdzr = up(c_loc(rdzr), b3); d_dzw = up(c_loc(dzw), b3w)
  d_dcbt = up(c_loc(dcbt), 2*b3); d_zero2 = up(c_loc(zero2), b2); d_rhog = up(c_loc(rhog), b2)
  d_thk = up(c_loc(thk), b2)
  d_eta(1) = up(c_loc(eta), b2); d_eta(2) = up(c_loc(eta), b2)
  d_ud(1) = up(c_loc(ud), 2*b2); d_ud(2) = up(c_loc(ud), 2*b2)
  d_forc = up(c_loc(ud), 2*b2); d_conv = up(c_loc(zero2), b2)
  do n = 1, ntr
     d_T(n,1) = up(c_loc(tr(isd,jsd,1,n)), b3)
     d_T(n,2) = up(c_loc(tr(isd,jsd,1,n)), b3)
     d_th(n) = up(c_loc(rdz), b3)
     d_lim(n) = up(c_loc(rdz), b3)
     d_stf(n) = d_zero2; d_btf(n) = d_zero2
     call mom_gpu_check(mom_gpu_fill(ctx, d_lim(n), 0.0d0, int(b3/8, c_size_t)), ctx, 'fill')
     call mom_gpu_check(mom_gpu_tracer_content(ctx, d_T(n,1), d_rdz, c0(n)), ctx, 'content')
  enddo
  bp = mom_gpu_bt_params_t(40, dtbt, 0.2d0, rho0, 0, 1)

  ! ---- time loop in update_ocean_model order ----
  now = 1
  do step = 1, nsteps
     nxt = 3 - now
     do n = 1, ntr
        tab_now(n) = d_T(n,now); tab_nxt(n) = d_T(n,nxt); tab_lim(n) = d_lim(n)
     enddo
     rc = mom_gpu_compute_tmask_limit(ctx, ntr, tab_now, mins, maxs, 0, 1, 1, tab_lim)
     call mom_gpu_check(rc, ctx, 'compute_tmask_limit')
     do n = 1, ntr
        call mom_gpu_check(mom_gpu_fill(ctx, d_th(n), 0.0d0, int(b3/8, c_size_t)), ctx, 'fill')
        rc = mom_gpu_horz_advect_tracer(ctx, MOM_ADVECT_MDFL_SWEBY, d_T(n,now), d_T(n,now), &
             d_uh, d_vh, d_wb, d_rdz, d_lim(n), 0, dt, d_th(n), c_null_ptr, c_null_ptr, &
             c_null_ptr, c_null_ptr)
        call mom_gpu_check(rc, ctx, 'horz_advect_tracer')
        rc = mom_gpu_vert_advect_tracer(ctx, MOM_ADVECT_MDFL_SWEBY, d_T(n,now), d_T(n,now), &
             d_wb, d_lim(n), d_th(n), c_null_ptr, c_null_ptr)
        call mom_gpu_check(rc, ctx, 'vert_advect_tracer')
        rc = mom_gpu_update_tracer_explicit(ctx, d_T(n,now), d_th(n), d_rdz, d_rdzr, dt, d_T(n,nxt))
        call mom_gpu_check(rc, ctx, 'update_tracer_explicit')
     enddo
     rc = mom_gpu_vert_diffuse_implicit(ctx, ntr, tab_nxt, d_stf, d_btf, d_dcbt, c_null_ptr, 1, &
          rho0, d_rdz, d_dzw, dt, 1.0d0)
     call mom_gpu_check(rc, ctx, 'vert_diffuse_implicit')
     rc = mom_gpu_update_ocean_barotropic(ctx, bp, c_null_ptr, d_eta(now), d_ud(now), d_forc, &
          d_thk, d_zero2, d_zero2, d_zero2, d_zero2, d_zero2, d_rhog, d_eta(nxt), d_ud(nxt), d_conv)
     call mom_gpu_check(rc, ctx, 'update_ocean_barotropic')
     now = nxt
  enddo
  call mom_gpu_check(mom_gpu_sync(ctx), ctx, 'sync')

  do n = 1, ntr
     call mom_gpu_check(mom_gpu_tracer_content(ctx, d_T(n,now), d_rdz, c1(n)), ctx, 'content')
     rel = abs(c1(n) - c0(n))/abs(c0(n))
     write(*,'(a,i2,a,es24.16,a,es24.16,a,es9.2)') 'tracer', n, ' content ', c0(n), ' -> ', c1(n), &
          '  rel drift ', rel
     if (rel > 1.0d-12) then
        write(*,'(a)') 'FATAL: tracer content not conserved'
        error stop 2
     endif
  enddo
  ! ---- the same tracer steps through the batched entry point ----
  do n = 1, ntr
     d_B(n,1) = up(c_loc(tr(isd,jsd,1,n)), b3)
     d_B(n,2) = up(c_loc(tr(isd,jsd,1,n)), b3)
  enddo
  dtypes = 8
  bnow = 1
  do step = 1, nsteps
     bnxt = 3 - bnow
     do n = 1, ntr
        tab_in(n) = d_B(n,bnow); tab_out(n) = d_B(n,bnxt)
     enddo
     st = mom_gpu_tracer_step_t(ntr, c_loc(tab_in), c_loc(tab_out), c_loc(dtypes), c_null_ptr, &
          c_loc(d_stf), c_loc(d_btf), 1, d_uh, d_vh, d_wb, d_rdz, c_null_ptr, d_rdzr, c_null_ptr, &
          d_dzw, d_dcbt, rho0, dt, 1.0d0)
     call mom_gpu_check(mom_gpu_step_tracers(ctx, st), ctx, 'step_tracers')
     bnow = bnxt
  enddo
  do n = 1, ntr
     call mom_gpu_check(mom_gpu_tracer_error(ctx, d_B(n,bnow), 8, d_T(n,now), 0.0d0, errv), ctx, &
          'tracer_error')
     write(*,'(a,i2,a,i0,a,i0,a)') 'tracer', n, ' batched vs per-tracer: ', nint(errv(6)), &
          ' of ', nint(errv(5)), ' wet cells differ'
     if (errv(6) /= 0.0d0) then
        write(*,'(a)') 'FATAL: batched step differs from the per-tracer calls'
        error stop 3
     endif
  enddo

  call mom_gpu_check(mom_gpu_destroy(ctx), c_null_ptr, 'destroy')
  write(*,'(a,i0,a)') 'ocean_gpu_driver: ', nsteps, ' steps done on the GPU, content conserved'

contains

  function up(src, bytes) result(dst)
    type(c_ptr), intent(in) :: src
    integer(c_size_t), intent(in) :: bytes
    type(c_ptr) :: dst
    call mom_gpu_check(mom_gpu_malloc(dst, bytes), c_null_ptr, 'mom_gpu_malloc')
    call mom_gpu_check(mom_gpu_memcpy_htod(dst, src, bytes), c_null_ptr, 'mom_gpu_memcpy_htod')
  end function up

end program ocean_gpu_driver
