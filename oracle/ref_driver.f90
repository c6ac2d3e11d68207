!-----------------------------------------------------------------------------
! ref_driver.f90 -- test harness around the reference's own hot-path objects
! (TEST INFRASTRUCTURE: oracle/Makefile links it with global, comp, muscl,
! riemann, boundary, timestep and parameter.f90 compiled in place from the
! reference tree into oracle/_ref/).
!
! It replaces only main.f90's role for pure hydro on one rank:
!   parameters  readparameter + compparameter   (parameter.f90:26-96,164-265)
!   ranges      lx=1..nx etc. as decomposition does for mpisize=1
!               (parallel.f90:33-64)
!   arrays      q, w, evo, s allocated and zeroed as main.f90:72-77
!   loop        primitive, bound, dt=1d99, (hydrostep), [thermalstep],
!               updatetime, compevo, (sweep), {gravpotential,
!               gravintegration}, [primitive, bound, thermalconduction]
!               (main.f90:109-175; () with hydro, [] with conduction, {} with gravity: the
!               reference's potential needs FFTW-2, absent here, so runs
!               with gravity are for the GPU drop-in, EVORA_BACKEND=hip)
!
! usage: evref <parafile> <mode> <in.bin> <out.bin> [nsteps]
!   mode run    : in = q(nq,nx,ny,nz) f64; nsteps passes; out = q
!   mode stage  : in = q; one pass; out = w, evo, s, q_after_sweep, dt
!   mode riemann: in = n records of (wl(5), wr(5), dir); out = f(6) per record
!   mode slope  : in = n records of (wl(5), w(5), wr(5)); out = s(5) per record
!   mode green  : in = phi(nx,ny,nz) complex(8); multigreen
!                 (gravitation.f90:55-89); out = phi
!   mode gravint: in = q(nq,nx,ny,nz), phi(nx,ny,nz) real(8), dt;
!                 gravintegration (gravitation.f90:172-198) with phi as the
!                 real part of the potential; out = q
! Each step prints "STEP k t dt" to stdout; the step loop's wall time is
! printed as "LOOP_SECONDS s" (system_clock, I/O excluded).  Reference errors print their own
! "ERROR ..." lines and `stop` (exit status 0), exactly as in the reference.
!-----------------------------------------------------------------------------
program evref
  use global
  implicit none

  character(len=256) :: mode, fin, fout, arg
  real(8), allocatable :: q(:,:,:,:), w(:,:,:,:), evo(:,:,:,:), s(:,:,:,:,:)
  real(8), allocatable :: rec(:), phir(:,:,:), phibuf2(:,:), phibuf4(:,:)
  complex(8), allocatable :: phic(:,:,:)
  real(8) :: wl(nw), wr(nw), wc(nw), f(nf), sl(nw)
  integer :: nsteps, k, n, i, u, fsize, dir, k4, k2, l, st
  integer(8) :: c0, c1, crate
  logical :: go, dump, evolve

  call get_command_argument(2, mode)
  call get_command_argument(3, fin)
  call get_command_argument(4, fout)
  nsteps = 1
  if (command_argument_count() .ge. 5) then
     call get_command_argument(5, arg)
     read (arg, *) nsteps
  endif

  mpirank = 0 ; mpisize = 1
  master = .true. ; notmaster = .false. ; last = .true. ; notlast = .false.
  even = .true. ; noteven = .false. ; parallel = .false.

  call readparameter()
  call compparameter()
  call compunits()     ! main.f90:62

  lx  = 1      ; mx  = nx
  lx1 = lx - 1 ; lx2 = lx - 2 ; lx0 = lx + 1
  mx1 = mx + 1 ; mx2 = mx + 2 ; mx0 = mx - 1
  ly  = 1      ; my  = ny
  ly1 = ly - 1 ; ly2 = ly - 2 ; ly0 = ly + 1
  my1 = my + 1 ; my2 = my + 2 ; my0 = my - 1
  lz  = 1      ; mz  = nz
  lz1 = lz - 1 ; lz2 = lz - 2 ; lz0 = lz + 1
  mz1 = mz + 1 ; mz2 = mz + 2 ; mz0 = mz - 1

  if (trim(mode) .eq. 'riemann' .or. trim(mode) .eq. 'slope') then
     inquire (file=trim(fin), size=fsize)
     n = fsize / 8
     allocate(rec(n))
     open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
     read (u) rec
     close (u)
     open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
     if (trim(mode) .eq. 'riemann') then
        do i = 0, n / 11 - 1
           wl = rec(11*i+1:11*i+5)
           wr = rec(11*i+6:11*i+10)
           dir = nint(rec(11*i+11))
           call riemannsolver(wl, wr, f, dir)
           write (u) f
        enddo
     else
        do i = 0, n / 15 - 1
           wl = rec(15*i+1:15*i+5)
           wc = rec(15*i+6:15*i+10)
           wr = rec(15*i+11:15*i+15)
           call comp_slope(wl, wc, wr, sl)
           write (u) sl
        enddo
     endif
     close (u)
     stop
  endif

  if (trim(mode) .eq. 'green' .or. trim(mode) .eq. 'gravint') then
     allocate(phic(lx:mx,ly:my,lz:mz), phibuf2(lx:mx,ly:my), phibuf4(lx:mx,ly:my))
     phibuf2 = 0.0
     phibuf4 = 0.0
     open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
     if (trim(mode) .eq. 'green') then
        read (u) phic
        call multigreen(phic)
     else
        allocate(q(nq,lx:mx,ly:my,lz:mz), phir(lx:mx,ly:my,lz:mz))
        read (u) q
        read (u) phir
        read (u) dt
        phic = cmplx(phir, 0.0d0, kind=8)
        call gravintegration(q, phic, phibuf2, phibuf4)
     endif
     close (u)
     open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
     if (trim(mode) .eq. 'green') then
        write (u) phic
     else
        write (u) q
     endif
     close (u)
     stop
  endif

  allocate(q(nq,lx:mx,ly:my,lz:mz))
  allocate(w(nw,lx2:mx2,ly2:my2,lz2:mz2))
  allocate(evo(nw,lx1:mx1,ly1:my1,lz1:mz1))
  allocate(s(nw,3,lx1:mx1,ly1:my1,lz1:mz1))
  w = 0.0
  evo = 0.0
  s = 0.0

  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read (u) q
  close (u)

  if (gravity) then
     call get_environment_variable('EVORA_BACKEND', arg, l, st)
     if (st .ne. 0 .or. arg(1:max(l, 1)) .ne. 'hip') then
        print *, 'ERROR: the reference potential needs FFTW-2 (not in this image); run gravity with EVORA_BACKEND=hip'
        stop
     endif
     allocate(phic(lx:mx,ly:my,lz:mz), phibuf2(lx:mx,ly:my), phibuf4(lx:mx,ly:my))
     phibuf2 = 0.0
     phibuf4 = 0.0
  endif

  go = .true.
  dump = .false.
  k4 = 1
  ! EVREF_EVOLUTION=1: after each step, the periodic totals of
  ! main.f90:184-191 (not comoving).  The counter starts at 1, as after the
  ! call of main.f90:100 before the loop, which this harness does not make:
  ! the reference evolution ends in the MPI barrier of parallel.f90, which is
  ! not linked here, so the option is for the GPU drop-in (whose evolution
  ! takes the sums on the device) only.
  call get_environment_variable('EVREF_EVOLUTION', arg, l, st)
  evolve = st .eq. 0 .and. l .gt. 0
  k2 = 1
  if (trim(mode) .eq. 'stage') nsteps = 1
  call system_clock(c0, crate)
  do k = 1, nsteps
     if (.not. go) exit
     call primitive(w, q)
     call bound(w)
     dt = 1d99
     if (hydro) call hydrostep(w)              ! main.f90:119
     if (conduction) call thermalstep(q, w)    ! main.f90:121
     call updatetime(go, dump, k4)
     if (secondorder) call compevo(w, evo, s)
     if (trim(mode) .eq. 'stage') then
        open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
        write (u) w
        write (u) evo
        write (u) s
     endif
     if (hydro) call sweep(w, q, evo, s)       ! main.f90:155
     if (gravity) then                          ! main.f90:157-164
        call gravpotential(q, phic)
        call gravintegration(q, phic, phibuf2, phibuf4)
     endif
     if (conduction) then                       ! main.f90:170-175
        call primitive(w, q)
        call bound(w)
        call thermalconduction(q, w)
     endif
     write (*, '(A,I6,2ES25.16)') 'STEP ', k, t, dt
     if (evolve) then
        tsnap = (t - tstart) / (tend - tstart) * nsnap
        if (tsnap .ge. dble(k2)) call evolution(k2, q)
     endif
  enddo
  call system_clock(c1)
  write (*, '(A,ES16.8)') 'LOOP_SECONDS ', dble(c1 - c0) / dble(crate)

  if (trim(mode) .eq. 'stage') then
     write (u) q
     write (u) dt
     close (u)
  else
     open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
     write (u) q
     close (u)
  endif

end program evref
