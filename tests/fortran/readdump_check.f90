! Test driver: set the grid of module global, call the ic(q) plugin of
! evora_amd/fortran/ic_readdump.f90 and write q raw to the file in argv(1).
program readdump_check
  use global
  implicit none
  real(8), allocatable :: q(:,:,:,:)
  character(len=256) :: fout, arg
  integer :: u
  call get_command_argument(1, fout)
  call get_command_argument(2, arg); read (arg, *) nx
  call get_command_argument(3, arg); read (arg, *) ny
  call get_command_argument(4, arg); read (arg, *) nz
  lx = 1; mx = nx; ly = 1; my = ny; lz = 1; mz = nz
  allocate(q(nq, lx:mx, ly:my, lz:mz))
  call ic(q)
  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u) q
  close (u)
end program readdump_check
