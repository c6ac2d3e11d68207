!-----------------------------------------------------------------------------
! hydro_dispatch.f90 -- runtime-selectable hydro backend for the unchanged
! driver (source/main.f90:109-203).
!
! Defines primitive, bound, hydrostep, compevo and sweep with the reference's
! implicit interfaces (muscl.f90:26,45,171; boundary.f90:26; timestep.f90:26)
! and arguments sized from module global, exactly as the reference routines
! are.  The reference objects are linked with these five symbols renamed to
! evcpu_* (objcopy --redefine-sym), so that
!
!   EVORA_BACKEND=cpu (default)  calls the reference routines unchanged;
!   EVORA_BACKEND=hip            runs the step on the GPU through libevora_hip.
!
! It also defines evolution (snapshot.f90:159-193), main.f90's periodic
! report of total mass, energy and entropy: once the state lives on the
! device the three sums are taken there (evh_sums, FP64 compensated), so an
! evolution step moves 24 bytes instead of the whole of q.
!
! Further environment switches:
!   EVORA_PRECISION=fp64|mixed   FP64 (default) or the FP32-mixed form
!   EVORA_EXACT=1                FP64 with the reference's own arithmetic in
!                                the single-pass kernel (bit-identical to the
!                                CPU path except where a power enters)
!   EVORA_MAXSTEPS=N             after N steps, set tend = t in hydrostep so
!                                that updatetime (timestep.f90:163-166) ends
!                                the run with a dt = 0 step (both backends)
!
! On the GPU the conserved state stays on the device.  It is uploaded at the
! first primitive call and downloaded after sweep only when the driver is
! about to read q: when snapshot will run (the tsnap test of
! main.f90:192-195, mirrored with a shadow counter) and on the last step
! (tend - t <= dt, the test of timestep.f90:163).  w, evo and s are never
! filled on the host; pure hydro never reads them there.  Thermal
! conduction (thermalstep, thermalconduction with simplekappa, without dual
! energy; on 1-, 2- and 3-D grids, also with hydro = .false., when main.f90
! calls neither hydrostep nor sweep) and self-gravity on 1-, 2- and 3-D
! grids (gravpotential, gravintegration; the potential stays on the device) run
! on the device too.  Physics that
! update q on the host between steps (chemistry, floors), supercomoving
! runs (whose host-side time and snapshot rules this dispatch does not
! mirror), conduction with the full kappa or with dual energy, and
! multi-rank runs are refused on the GPU backend.
!-----------------------------------------------------------------------------
module hydro_dispatch_state
  use iso_c_binding
  use evora_hip
  implicit none

  integer :: backend = -1            ! -1 undecided, 0 cpu, 1 hip
  type(c_ptr) :: ctx = c_null_ptr
  logical :: on_device = .false.
  logical :: last_step = .false.
  integer :: maxsteps = -1, nsteps_seen = 0
  integer :: k3s = 1                 ! shadow of main's snapshot counter
  integer(c_int) :: precision = EVH_FP64
  logical :: exact = .false.

contains

  subroutine decide_backend()
    character(len=64) :: v
    integer :: l, st
    if (backend .ge. 0) return
    backend = 0
    call get_environment_variable('EVORA_BACKEND', v, l, st)
    if (st .eq. 0 .and. l .gt. 0) then
       if (v(1:l) .eq. 'hip' .or. v(1:l) .eq. 'HIP') backend = 1
    endif
    call get_environment_variable('EVORA_PRECISION', v, l, st)
    if (st .eq. 0 .and. l .gt. 0) then
       if (v(1:l) .eq. 'mixed') precision = EVH_FP32_MIXED
    endif
    call get_environment_variable('EVORA_EXACT', v, l, st)
    if (st .eq. 0 .and. l .gt. 0) exact = v(1:l) .eq. '1'
    call get_environment_variable('EVORA_MAXSTEPS', v, l, st)
    if (st .eq. 0 .and. l .gt. 0) read (v(1:l), *) maxsteps
  end subroutine decide_backend

  ! reference messages and stop kinds for each status (comp.f90:74-83,
  ! riemann.f90:37-41,61-66,87-90,111-116,148-151, timestep.f90:51-55,
  ! muscl.f90:164-167)
  subroutine hip_check(code)
    use global
    integer(c_int), intent(in) :: code
    integer(c_int) :: c, cell(3)
    character(kind=c_char) :: msg(256)
    character(len=256) :: text
    integer :: i
    if (code .eq. EVH_OK) return
    c = evh_last_error(ctx, c, cell, msg, 256_c_int)
    text = ' '
    do i = 1, 256
       if (msg(i) .eq. c_null_char) exit
       text(i:i) = msg(i)
    enddo
    select case (code)
    case (EVH_ERR_NEG_PRESSURE)
       print *, 'ERROR negative pressure occured while computing'
       print *, 'primitive variables!'
       print *, 'time:', t
       print *, 'cell:', cell
       stop
    case (EVH_ERR_RIEMANN_PRESSURE)
       if (solver .eq. 1) then
          print *, 'ERROR: hll solver unphysical pressure'
       else
          print *, 'ERROR: hllc solver unphysical pressure'
       endif
       print *, 'cell:', cell
       stop
    case (EVH_ERR_HLL)
       print *, 'ERROR: hll solver'
       stop
    case (EVH_ERR_HLLC)
       print *, 'ERROR: hllc solver', cell
       stop
    case (EVH_ERR_DT)
       print *, 'ERROR: hydrostep is negative'
       print *, 'infinite pressure ?'
       call cleanstop
    case (EVH_ERR_SOLVER)
       print *, 'ERROR choose proper solver!'
       call cleanstop
    case (EVH_ERR_LIMITER)
       print *, 'ERROR: wrong slope limiter!'
       stop
    case default
       print *, 'ERROR: evora_hip: ', trim(text)
       call cleanstop
    end select
  end subroutine hip_check

  subroutine hip_create()
    use global
    type(evh_config) :: cfg
    if (chemistry .or. supercomoving .or. tempfloor .or. jeansfloor &
         .or. cooldt .or. noncie .or. parallel .or. (.not. hydro .and. .not. conduction)) then
       print *, 'ERROR: EVORA_BACKEND=hip runs hydro (with thermal conduction and gravity) on one rank only'
       call cleanstop
    endif
    if (conduction .and. (.not. simplekappa .or. dualenergy)) then
       print *, 'ERROR: EVORA_BACKEND=hip runs thermal conduction with simplekappa and without dualenergy only'
       call cleanstop
    endif
    cfg%nx = nx ; cfg%ny = ny ; cfg%nz = nz
    cfg%boundtype = boundtype ; cfg%solver = solver ; cfg%limiter = limiter
    cfg%secondorder = merge(1, 0, secondorder)
    cfg%dualenergy = merge(1, 0, dualenergy)
    cfg%onedimension = merge(1, 0, onedimension)
    cfg%twodimension = merge(1, 0, twodimension)
    cfg%precision = precision
    cfg%g = g ; cfg%g1 = g1 ; cfg%g1inv = g1inv ; cfg%gg1 = gg1 ; cfg%cfl = cfl
    cfg%de_thresh1 = de_thresh1
    cfg%dx = dx ; cfg%dy = dy ; cfg%dz = dz
    cfg%dxinv = dxinv ; cfg%dyinv = dyinv ; cfg%dzinv = dzinv
    cfg%volume = volume
    cfg%mach_fp64 = 5.0d0
    cfg%conduction = merge(1, 0, conduction)
    cfg%simplekappa = merge(1, 0, simplekappa)
    cfg%c_cond = c_cond ; cfg%temp0 = temp0 ; cfg%kappa0 = kappa0
    cfg%gravity = merge(1, 0, gravity)
    cfg%supercomoving = merge(1, 0, supercomoving)
    cfg%grav = grav ; cfg%fb = fb ; cfg%h2 = h2
    cfg%kfx2 = kfx2 ; cfg%kfy2 = kfy2 ; cfg%kfz2 = kfz2
    cfg%nyqx = nyqx ; cfg%nyqy = nyqy ; cfg%nyqz = nyqz
    cfg%dx2inv = dx2inv ; cfg%dy2inv = dy2inv ; cfg%dz2inv = dz2inv
    cfg%hydro_off = merge(0, 1, hydro)
    call hip_check(evh_create(cfg, ctx))
    ! 2-D and 3-D grids start in the fused form; the exact form on request (FP64)
    if (exact .and. precision .eq. EVH_FP64 .and. .not. onedimension) &
         call hip_check(evh_set_mode(ctx, EVH_MODE_FUSED_EXACT))
  end subroutine hip_create

  ! one more time step begins (the step's first time limit was taken):
  ! EVORA_MAXSTEPS and the last-step test of timestep.f90:163
  subroutine count_step()
    use global
    nsteps_seen = nsteps_seen + 1
    if (maxsteps .ge. 0 .and. nsteps_seen .gt. maxsteps) tend = t
    last_step = (tend - t .le. dt)
  end subroutine count_step

  ! main.f90:181-195 after this step (t already advanced by updatetime):
  ! download q when the driver is about to read it; evolution takes its
  ! sums on the device and needs no q
  subroutine download_if_read(q)
    use global
    real(8), dimension(nq,lx:mx,ly:my,lz:mz) :: q
    real(8) :: tsnap_next
    logical :: need_q
    need_q = last_step
    tsnap_next = (t - tstart) / (tend - tstart) * nsnap
    if ((cut .or. slice .or. bov) .and. tsnap_next .ge. dble(k3s)) then
       need_q = .true.
       k3s = k3s + 1
    endif
    if (need_q) call hip_check(evh_download_q(ctx, q))
  end subroutine download_if_read

end module hydro_dispatch_state

subroutine primitive(w, q)
  use global
  use hydro_dispatch_state
  implicit none
  real(8), dimension(nq,lx:mx,ly:my,lz:mz) :: q
  real(8), dimension(nw,lx2:mx2,ly2:my2,lz2:mz2) :: w

  call decide_backend()
  if (backend .eq. 0) then
     call evcpu_primitive(w, q)
     return
  endif
  if (.not. c_associated(ctx)) call hip_create()
  if (.not. on_device) then
     call hip_check(evh_upload_q(ctx, q))
     on_device = .true.
  endif
  call hip_check(evh_primitive(ctx))
end subroutine primitive

subroutine bound(w)
  use global
  use hydro_dispatch_state
  implicit none
  real(8), dimension(nw,lx2:mx2,ly2:my2,lz2:mz2) :: w

  call decide_backend()
  if (backend .eq. 0) then
     call evcpu_bound(w)
     return
  endif
  call hip_check(evh_bound(ctx))
end subroutine bound

subroutine hydrostep(w)
  use global
  use hydro_dispatch_state
  implicit none
  real(8), dimension(nw,lx2:mx2,ly2:my2,lz2:mz2) :: w

  call decide_backend()
  if (backend .eq. 0) then
     call evcpu_hydrostep(w)
  else
     call hip_check(evh_hydrostep(ctx, dt))
  endif
  call count_step()
end subroutine hydrostep

subroutine compevo(w, evo, s)
  use global
  use hydro_dispatch_state
  implicit none
  real(8), dimension(nw,lx2:mx2,ly2:my2,lz2:mz2)   :: w
  real(8), dimension(nw,lx1:mx1,ly1:my1,lz1:mz1)   :: evo
  real(8), dimension(nw,3,lx1:mx1,ly1:my1,lz1:mz1) :: s

  call decide_backend()
  if (backend .eq. 0) then
     call evcpu_compevo(w, evo, s)
     return
  endif
  call hip_check(evh_compevo(ctx, dt))
end subroutine compevo

subroutine sweep(w, q, evo, s)
  use global
  use hydro_dispatch_state
  implicit none
  real(8), dimension(nq,lx:mx,ly:my,lz:mz)         :: q
  real(8), dimension(nw,lx2:mx2,ly2:my2,lz2:mz2)   :: w
  real(8), dimension(nw,lx1:mx1,ly1:my1,lz1:mz1)   :: evo
  real(8), dimension(nw,3,lx1:mx1,ly1:my1,lz1:mz1) :: s

  call decide_backend()
  if (backend .eq. 0) then
     call evcpu_sweep(w, q, evo, s)
     return
  endif
  call hip_check(evh_sweep(ctx, dt))
  ! with gravity or conduction the step ends in gravintegration or
  ! thermalconduction, which download
  if (.not. (conduction .or. gravity)) call download_if_read(q)
end subroutine sweep

! gravitation.f90:26-52 (main.f90:159): on the device the potential stays
! there for gravintegration; the host phi is not filled (main.f90 does not
! read it)
subroutine gravpotential(q, phi)
  use global
  use hydro_dispatch_state
  implicit none
  real(8), intent(in), dimension(nq,lx:mx,ly:my,lz:mz) :: q
  complex(8), intent(out), dimension(lx:mx,ly:my,lz:mz) :: phi

  call decide_backend()
  if (backend .eq. 0) then
     call evcpu_gravpotential(q, phi)
     return
  endif
  call hip_check(evh_gravpotential(ctx, cosmo))
end subroutine gravpotential

! gravitation.f90:172-254 (main.f90:163)
subroutine gravintegration(q, phi, phibuf2, phibuf4)
  use global
  use hydro_dispatch_state
  implicit none
  real(8), intent(inout), dimension(nq,lx:mx,ly:my,lz:mz) :: q
  complex(8), intent(in), dimension(lx:mx,ly:my,lz:mz) :: phi
  real(8), intent(in), dimension(lx:mx,ly:my) :: phibuf2, phibuf4

  call decide_backend()
  if (backend .eq. 0) then
     call evcpu_gravintegration(q, phi, phibuf2, phibuf4)
     return
  endif
  call hip_check(evh_gravintegration(ctx, dt))
  if (.not. conduction) call download_if_read(q)
end subroutine gravintegration

! snapshot.f90:159-193.  Before the state is on the device (the call of
! main.f90:100) and on the CPU backend this is the reference routine; on the
! device the sums come from evh_sums and the line written is the
! reference's (same file, header and format).  The GPU backend runs on one
! rank, so the parallel reductions and the barrier of the reference are not
! needed.
subroutine evolution(k2, q)
  use global
  use hydro_dispatch_state
  implicit none
  integer, intent(inout) :: k2
  real(8), intent(in), dimension(nq,lx:mx,ly:my,lz:mz) :: q
  real(c_double) :: sums(3)
  real(8) :: totalmass, totalenergy, totalentropy
  integer :: u

  call decide_backend()
  if (backend .eq. 0 .or. .not. on_device) then
     call evcpu_evolution(k2, q)
     return
  endif
  call hip_check(evh_sums(ctx, sums))
  totalmass = sums(1) * volume
  totalenergy = sums(2) * volume
  totalentropy = sums(3) * volume
  if (master) then
     open (newunit=u, file="data/evo", access="sequential", status="unknown", position="append")
     if (k2 .eq. 0) write (u,'(A,A15,23A16)') "#","t","dt","a","z","totalmass","totalenergy","totalentropy"
     write (u,'(24E16.6)') t, dt, cosmo, redshift, totalmass, totalenergy, totalentropy
     close (u)
  endif
  k2 = k2 + 1
end subroutine evolution

! timestep.f90:72-104 (main.f90:121)
subroutine thermalstep(q, w)
  use global
  use hydro_dispatch_state
  implicit none
  real(8), dimension(nq,lx:mx,ly:my,lz:mz) :: q
  real(8), dimension(nw,lx2:mx2,ly2:my2,lz2:mz2) :: w

  call decide_backend()
  if (backend .eq. 0) then
     call evcpu_thermalstep(q, w)
  else
     call hip_check(evh_thermalstep(ctx, dt))
  endif
  ! without hydro (main.f90:119) this is the step's first limit
  if (.not. hydro) call count_step()
  last_step = (tend - t .le. dt)  ! dt may have shrunk: the test of timestep.f90:163 again
end subroutine thermalstep

! conduction.f90:26-123 (main.f90:174).  main.f90:171-172 call primitive and
! bound before it; on the device those are folded into the conduction kernel
! (the dispatch's primitive and bound find the state already current).
subroutine thermalconduction(q, w)
  use global
  use hydro_dispatch_state
  implicit none
  real(8), dimension(nq,lx:mx,ly:my,lz:mz) :: q
  real(8), dimension(nw,lx2:mx2,ly2:my2,lz2:mz2) :: w

  call decide_backend()
  if (backend .eq. 0) then
     call evcpu_thermalconduction(q, w)
     return
  endif
  call hip_check(evh_thermalconduction(ctx, dt))
  call download_if_read(q)
end subroutine thermalconduction
