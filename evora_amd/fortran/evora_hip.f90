!-----------------------------------------------------------------------------
! evora_hip.f90 -- iso_c_binding interface to libevora_hip (include/evora_hip.h)
!
! Fortran hosts bind the C-ABI through this module.  evh_config mirrors the
! C struct field by field (bind(c) gives it the C layout); every entry point
! returns an integer(c_int) status, EVH_OK = 0.
!-----------------------------------------------------------------------------
module evora_hip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: EVH_OK = 0
  integer(c_int), parameter :: EVH_ERR_NEG_PRESSURE = 1
  integer(c_int), parameter :: EVH_ERR_RIEMANN_PRESSURE = 2
  integer(c_int), parameter :: EVH_ERR_HLL = 3
  integer(c_int), parameter :: EVH_ERR_HLLC = 4
  integer(c_int), parameter :: EVH_ERR_DT = 5
  integer(c_int), parameter :: EVH_ERR_SOLVER = 6
  integer(c_int), parameter :: EVH_ERR_LIMITER = 7
  integer(c_int), parameter :: EVH_ERR_CONFIG = 8
  integer(c_int), parameter :: EVH_ERR_HIP = 9
  integer(c_int), parameter :: EVH_ERR_STATE = 10
  integer(c_int), parameter :: EVH_FP64 = 0, EVH_FP32_MIXED = 1
  integer(c_int), parameter :: EVH_MODE_FUSED = 0, EVH_MODE_STAGED = 1, EVH_MODE_FUSED_EXACT = 2

  type, bind(c) :: evh_config
     integer(c_int) :: nx, ny, nz
     integer(c_int) :: boundtype, solver, limiter, secondorder, dualenergy
     integer(c_int) :: onedimension, twodimension, precision
     real(c_double) :: g, g1, g1inv, gg1, cfl, de_thresh1
     real(c_double) :: dx, dy, dz, dxinv, dyinv, dzinv
     real(c_double) :: volume, mach_fp64
     integer(c_int) :: conduction, simplekappa
     real(c_double) :: c_cond, temp0, kappa0
     integer(c_int) :: gravity, supercomoving
     real(c_double) :: grav, fb, h2, kfx2, kfy2, kfz2
     integer(c_int) :: nyqx, nyqy, nyqz
     real(c_double) :: dx2inv, dy2inv, dz2inv
     integer(c_int) :: hydro_off
  end type evh_config

  interface
     integer(c_int) function evh_create(cfg, h) bind(c, name='evh_create')
       import :: c_int, c_ptr, evh_config
       type(evh_config), intent(in) :: cfg
       type(c_ptr), intent(out) :: h
     end function evh_create

     integer(c_int) function evh_thermalstep(h, dt) bind(c, name='evh_thermalstep')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(inout) :: dt
     end function evh_thermalstep

     integer(c_int) function evh_thermalconduction(h, dt) bind(c, name='evh_thermalconduction')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), value :: dt
     end function evh_thermalconduction

     integer(c_int) function evh_gravpotential(h, cosmo) bind(c, name='evh_gravpotential')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), value :: cosmo
     end function evh_gravpotential

     integer(c_int) function evh_gravintegration(h, dt) bind(c, name='evh_gravintegration')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), value :: dt
     end function evh_gravintegration

     ! one rank's z-slab (parallel.f90:26-66): halo buffers for the ghost
     ! planes communicate (parallel.f90:317-381) receives, and the boundary
     ! planes it sends, as device addresses
     integer(c_int) function evh_slab(h, lo, hi, lo_halo, hi_halo) bind(c, name='evh_slab')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: lo, hi
       type(c_ptr) :: lo_halo, hi_halo
     end function evh_slab

     integer(c_int) function evh_slab_faces(h, lo_dst, hi_dst) bind(c, name='evh_slab_faces')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, lo_dst, hi_dst
     end function evh_slab_faces

     ! the same for T before thermalconduction (main.f90:170-175)
     integer(c_int) function evh_slab_temp(h, lo_halo, hi_halo) bind(c, name='evh_slab_temp')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       type(c_ptr) :: lo_halo, hi_halo
     end function evh_slab_temp

     integer(c_int) function evh_slab_temp_faces(h, lo_dst, hi_dst) bind(c, name='evh_slab_temp_faces')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, lo_dst, hi_dst
     end function evh_slab_temp_faces

     integer(c_int) function evh_destroy(h) bind(c, name='evh_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function evh_destroy

     integer(c_int) function evh_set_mode(h, mode) bind(c, name='evh_set_mode')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: mode
     end function evh_set_mode

     integer(c_int) function evh_upload_q(h, q) bind(c, name='evh_upload_q')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: q(*)
     end function evh_upload_q

     integer(c_int) function evh_download_q(h, q) bind(c, name='evh_download_q')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: q(*)
     end function evh_download_q

     integer(c_int) function evh_primitive(h) bind(c, name='evh_primitive')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function evh_primitive

     integer(c_int) function evh_bound(h) bind(c, name='evh_bound')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function evh_bound

     integer(c_int) function evh_hydrostep(h, dt) bind(c, name='evh_hydrostep')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(inout) :: dt
     end function evh_hydrostep

     integer(c_int) function evh_compevo(h, dt) bind(c, name='evh_compevo')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), value :: dt
     end function evh_compevo

     integer(c_int) function evh_sweep(h, dt) bind(c, name='evh_sweep')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), value :: dt
     end function evh_sweep

     integer(c_int) function evh_sums(h, out) bind(c, name='evh_sums')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: out(3)
     end function evh_sums

     integer(c_int) function evh_last_error(h, code, cell, msg, msglen) bind(c, name='evh_last_error')
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: h
       integer(c_int), intent(out) :: code
       integer(c_int), intent(out) :: cell(3)
       character(kind=c_char), intent(out) :: msg(*)
       integer(c_int), value :: msglen
     end function evh_last_error
  end interface

end module evora_hip
