!-----------------------------------------------------------------------------
! ic_readdump.f90 -- initial-condition plugin ic(q) that reads q from a dump
!
! Same interface as the compile-time IC plugins of the reference
! (ic/*.f90, e.g. ic/shocktube.f90:26-30; selected by Makefile:30 IC=), so a
! build with IC=readdump starts from any state written in the reference's
! dump format (write.f90:47-65: record 1 = nn,nx,ny,nz,lz,mz as int32,
! record 2 = q(nn,1:nx,1:ny,lz:mz) as real(8)).  The file name is taken from
! the environment variable EVORA_IC_DUMP (default data/ic.000).  It lets the
! CPU reference and the GPU drop-in start from bit-identical states, and
! doubles as a restart reader.
!-----------------------------------------------------------------------------
subroutine ic(q)
  use global
  implicit none

  real(8), dimension(nq,lx:mx,ly:my,lz:mz) :: q

  character(len=256) :: fname
  integer :: l, st, u, ierror
  integer :: nn, hx, hy, hz, hlz, hmz

  call get_environment_variable('EVORA_IC_DUMP', fname, l, st)
  if (st .ne. 0 .or. l .eq. 0) fname = 'data/ic.000'

  open (newunit=u, file=trim(fname), form='unformatted', status='old', iostat=ierror)
  if (ierror .ne. 0) then
     print *, 'ERROR file not found: ', trim(fname)
     call cleanstop
  endif
  read (u) nn, hx, hy, hz, hlz, hmz
  if (nn .ne. nq .or. hx .ne. nx .or. hy .ne. ny .or. hz .ne. nz .or. hlz .ne. lz .or. hmz .ne. mz) then
     print *, 'ERROR dump header', nn, hx, hy, hz, hlz, hmz, ' does not match the grid', nq, nx, ny, nz, lz, mz
     call cleanstop
  endif
  read (u) q
  close (u)

end subroutine ic
