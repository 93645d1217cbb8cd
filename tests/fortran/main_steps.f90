! MAIN's time loop (1d.f:196-295) on the reference's own arrays through the
! lica_mi355x module (include/lica_mi355x.f90).  Arguments: lica.in,
! grid.in, N3 (the deck's NS node count), steps, output file.  Writes the
! columns (1,1,0:N3) of W, P, PSI_ZN, PSI_CN and T, then DVMAX, CFLM, DT,
! VOL_TOT_ORI, TIME, with 17 significant digits.
program main_steps
  use iso_c_binding
  use lica_mi355x
  implicit none
  integer, parameter :: m1 = 2, m2 = 2
  integer(c_int64_t) :: n3, nsteps, msub, itracking, ini_psi, m
  integer(c_int) :: ierr
  type(c_ptr) :: ctx
  real(c_double), allocatable, dimension(:,:,:) :: u, v, w, p, t, psi_xn, psi_yn, psi_zn, psi_cn
  real(c_double), allocatable, dimension(:,:,:) :: rk3xo, rk3yo, rk3zo, anpso
  real(c_double) :: cflm, dvmax, dt, vol_tot_ori, time, qvol_ori, vol1
  character(len=512) :: lin, gin, arg, outf
  integer :: k

  call get_command_argument(1, lin)
  call get_command_argument(2, gin)
  call get_command_argument(3, arg)
  read (arg, *) n3
  call get_command_argument(4, arg)
  read (arg, *) nsteps
  call get_command_argument(5, outf)

  allocate(u(0:m1,0:m2,0:n3), v(0:m1,0:m2,0:n3), w(0:m1,0:m2,0:n3), p(0:m1,0:m2,0:n3), t(0:m1,0:m2,0:n3))
  allocate(psi_xn(0:m1,0:m2,0:n3), psi_yn(0:m1,0:m2,0:n3), psi_zn(0:m1,0:m2,0:n3), psi_cn(0:m1,0:m2,0:n3))
  allocate(rk3xo(0:m1,0:m2,0:n3), rk3yo(0:m1,0:m2,0:n3), rk3zo(0:m1,0:m2,0:n3), anpso(0:m1,0:m2,0:n3))
  u = 0.d0; v = 0.d0; w = 0.d0; p = 0.d0; t = 0.d0
  psi_xn = 0.d0; psi_yn = 0.d0; psi_zn = 0.d0; psi_cn = 0.d0
  rk3xo = 0.d0; rk3yo = 0.d0; rk3zo = 0.d0; anpso = 0.d0
  qvol_ori = 0.d0; vol1 = 0.d0; itracking = 2; ini_psi = 0

  ierr = lica_init(ctx, trim(lin)//c_null_char, trim(gin)//c_null_char)
  if (ierr /= lica_ok) stop 2
  ierr = lica_f_state(ctx, w, p, t, psi_zn, psi_cn, vol_tot_ori, time, dt)
  if (ierr /= lica_ok) stop 3
  do m = 1, nsteps
    ierr = lica_f_determine_dt(ctx, cflm, u, v, w, psi_cn, dt)
    if (ierr /= lica_ok) stop 4
    rk3zo = 0.d0
    anpso = 0.d0
    do msub = 1, 3
      ierr = lica_f_rhsnlhs(ctx, msub, itracking, u, v, w, p, psi_xn, psi_yn, psi_zn, psi_cn, &
                            rk3xo, rk3yo, rk3zo, anpso, qvol_ori, vol1)
      if (ierr /= lica_ok) stop 5
      ierr = lica_f_convrge1(ctx, dvmax, u, v, w)
      if (ierr /= lica_ok) stop 6
      ierr = lica_f_transport(ctx, u, v, w, psi_cn, vol_tot_ori)
      if (ierr /= lica_ok) stop 7
      ierr = lica_f_grid_coupling(ctx, ini_psi, psi_xn, psi_yn, psi_zn, psi_cn)
      if (ierr /= lica_ok) stop 8
    end do
    ierr = lica_end_step(ctx)
    if (ierr /= lica_ok) stop 9
  end do
  ierr = lica_f_state(ctx, rk3xo, rk3yo, t, psi_xn, psi_yn, vol1, time, qvol_ori)
  if (ierr /= lica_ok) stop 10

  open (unit=11, file=trim(outf), status='replace')
  do k = 0, int(n3)
    write (11, '(5ES26.17E3)') w(1,1,k), p(1,1,k), psi_zn(1,1,k), psi_cn(1,1,k), t(1,1,k)
  end do
  write (11, '(5ES26.17E3)') dvmax, cflm, dt, vol_tot_ori, time
  close (11)
  call lica_batch_destroy(ctx)
end program main_steps
