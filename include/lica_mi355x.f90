module lica_mi355x
  use iso_c_binding
  implicit none
  integer(c_int), parameter :: lica_ok = 0
  interface
    ! MAIN init 1d.f:70-190 (ALLOCINIT, GEOM, MAKEFLD, LVSINIT, GRID_COUPLING(1))
    integer(c_int) function lica_init(ctx, lica_in, grid_in) bind(c, name='lica_init')
      import :: c_int, c_ptr, c_char
      type(c_ptr) :: ctx
      character(kind=c_char) :: lica_in(*), grid_in(*)
    end function
    ! the module state lica_init leaves, into the driver's arrays
    integer(c_int) function lica_f_state(ctx, w, p, t, psi_zn, psi_cn, vol_tot_ori, time, dt) &
        bind(c, name='lica_f_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: w(*), p(*), t(*), psi_zn(*), psi_cn(*), vol_tot_ori, time, dt
    end function
    ! DETERMINE_DT(CFLM,U,V,W,PSI_CN) 1d.f:683; DT (module variable) in and out
    integer(c_int) function lica_f_determine_dt(ctx, cflm, u, v, w, psi_cn, dt) bind(c, name='lica_f_determine_dt')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: cflm, u(*), v(*), w(*), psi_cn(*), dt
    end function
    ! RHSNLHS(ITRACKING,U,V,W,P,PSI_XN,...,VOL1) 1d.f:1444; MSUB (MAIN's loop variable) first
    integer(c_int) function lica_f_rhsnlhs(ctx, msub, itracking, u, v, w, p, psi_xn, psi_yn, psi_zn, psi_cn, &
        rk3xo, rk3yo, rk3zo, anpso, qvol_ori, vol1) bind(c, name='lica_f_rhsnlhs')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t) :: msub, itracking
      real(c_double) :: u(*), v(*), w(*), p(*), psi_xn(*), psi_yn(*), psi_zn(*), psi_cn(*)
      real(c_double) :: rk3xo(*), rk3yo(*), rk3zo(*), anpso(*), qvol_ori, vol1
    end function
    ! CONVRGE1(DVMAX,U,V,W) 1d.f:3306
    integer(c_int) function lica_f_convrge1(ctx, dvmax, u, v, w) bind(c, name='lica_f_convrge1')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: dvmax, u(*), v(*), w(*)
    end function
    ! TRANSPORT(U,V,W,PSI_CN,VOL_TOT_ORI) two_phase.f:701
    integer(c_int) function lica_f_transport(ctx, u, v, w, psi_cn, vol_tot_ori) bind(c, name='lica_f_transport')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: u(*), v(*), w(*), psi_cn(*), vol_tot_ori
    end function
    ! GRID_COUPLING(INI_PSI,PSI_XN,PSI_YN,PSI_ZN,PSI_CN) two_phase.f:4863, INI_PSI = 0
    integer(c_int) function lica_f_grid_coupling(ctx, ini_psi, psi_xn, psi_yn, psi_zn, psi_cn) &
        bind(c, name='lica_f_grid_coupling')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t) :: ini_psi
      real(c_double) :: psi_xn(*), psi_yn(*), psi_zn(*), psi_cn(*)
    end function
    ! POISSON(P,DIVGSUM,DENF_XI,DENF_YI,DENF_ZI,DENF_Y,DENF_C) 1d.f:3492
    integer(c_int) function lica_f_poisson(ctx, p, divgsum, denf_xi, denf_yi, denf_zi, denf_y, denf_c) &
        bind(c, name='lica_f_poisson')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: p(*), divgsum(*), denf_xi(*), denf_yi(*), denf_zi(*), denf_y(*), denf_c(*)
    end function
    ! TEMP_GRAD(PSI_CN) chphase.f:55
    integer(c_int) function lica_f_temp_grad(ctx, psi_cn) bind(c, name='lica_f_temp_grad')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: psi_cn(*)
    end function
    ! TIME = TIME + DT (1d.f:260) is done by the last TRANSPORT; this waits and reports
    integer(c_int) function lica_end_step(ctx) bind(c, name='lica_end_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    ! the whole 1d.f:196-295 loop NSTEPS times, state resident on the GPU
    integer(c_int) function lica_step(ctx, nsteps) bind(c, name='lica_step')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nsteps
    end function
    ! WRITEFIELD (1d.f:1121-1232) of the context's case
    integer(c_int) function lica_batch_write_fld(ctx, inst, path, ihist) bind(c, name='lica_batch_write_fld')
      import :: c_int, c_ptr, c_int64_t, c_char
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: inst, ihist
      character(kind=c_char), dimension(*) :: path
    end function
    ! IREAD=1, ITRACKING=1 (1d.f:117-143): PREFLD, LVSINIT from the band read,
    ! GRID_COUPLING(1); flags = IRESET + 2*IPZERO + 4*IDTOLD of lica.in line 2
    integer(c_int) function lica_batch_read_fld(ctx, inst, path, flags) bind(c, name='lica_batch_read_fld')
      import :: c_int, c_ptr, c_int64_t, c_char
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: inst, flags
      character(kind=c_char), dimension(*) :: path
    end function
    subroutine lica_batch_destroy(ctx) bind(c, name='lica_batch_destroy')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
  end interface
end module lica_mi355x
