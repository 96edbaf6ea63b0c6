! Exercises the drop-in json_module, mpi, mpiio and wuming_utils the way the
! reference's Weibel app does (proj/weibel/app.f90:162-258, 535, 550-627):
! load a configuration, read typed members, serialize, deserialize, update a
! member, print; an MPI_Reduce over one rank; stream I/O at a displacement.
! Prints what it read, one item per line, for tests/test_cpu_boundary.py.
program json_check
  use iso_fortran_env, only: int64
  use mpi
  use json_module
  use mpiio
  use wuming_utils
  implicit none
  type(json_core) :: json
  type(json_file) :: file
  type(json_value), pointer :: root, p
  character(len=:), allocatable :: datadir, missing, text
  integer :: max_it, nproc, fh, ierr, i4
  real(8) :: mass_ratio, e_l(3), e_g(3), r8
  logical :: found
  integer(int64) :: disp
  character(len=256) :: cfg

  call get_command_argument(1, cfg)
  call json%initialize()
  call file%initialize()
  call file%load(trim(cfg))
  if (file%failed()) stop 'load failed'
  call file%get(root)
  call json%get(root, 'config', p)
  call json%get(p, 'max_it', max_it)
  call json%get(p, 'datadir', datadir)
  call json%get(p, 'restart_file', missing, found)
  print '(a,i0)', 'max_it ', max_it
  print '(a,a)', 'datadir ', datadir
  print '(a,l1)', 'restart_found ', found
  call json%get(root, 'parameter', p)
  call json%get(p, 'mass_ratio', mass_ratio)
  call json%get(p, 'num_process', nproc)
  print '(a,es24.17)', 'mass_ratio ', mass_ratio
  print '(a,i0)', 'num_process ', nproc
  call json%get(p, 'no_such_member', nproc)
  print '(a,l1)', 'failed_after_missing ', json%failed()
  call json%serialize(root, text)
  call json%destroy()
  call file%destroy()
  ! deserialize, update, print (save_restart, app.f90:563-575)
  call json%initialize()
  call file%initialize()
  call file%deserialize(text)
  call file%get(root)
  call json%get(root, 'config', p)
  call json%update(p, 'restart_file', '0000011_restart', found)
  print '(a,l1)', 'updated_existing ', found
  call json%print(root, trim(cfg) // '.out')
  call json%destroy()
  call file%destroy()
  ! one-rank reduction (energy_history, app.f90:535)
  e_l = [1.5d0, 2.5d0, 3.5d0]
  e_g = 0
  call MPI_Reduce(e_l, e_g, 3, MPI_DOUBLE_PRECISION, MPI_SUM, 0, MPI_COMM_WORLD, ierr)
  print '(a,3f6.2)', 'reduced ', e_g
  ! stream I/O at displacements
  call mpiio_open_file(trim(cfg) // '.raw', fh, disp, 'w')
  call mpiio_write_atomic(fh, disp, 7)
  call mpiio_write_atomic(fh, disp, 0.25d0)
  call mpiio_close_file(fh)
  call mpiio_open_file(trim(cfg) // '.raw', fh, disp, 'a')
  print '(a,i0)', 'append_disp ', disp
  call mpiio_close_file(fh)
  call mpiio_open_file(trim(cfg) // '.raw', fh, disp, 'r')
  call mpiio_read_atomic(fh, disp, i4)
  call mpiio_read_atomic(fh, disp, r8)
  call mpiio_close_file(fh)
  print '(a,i0,1x,f5.2)', 'read_back ', i4, r8
  print '(a,i0)', 'endian ', mpiio_get_endian_flag()
  r8 = uniform_rand()
  print '(a,l1)', 'uniform_in_range ', r8 >= 0 .and. r8 < 1
end program json_check
