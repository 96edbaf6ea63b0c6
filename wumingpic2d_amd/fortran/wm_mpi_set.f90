! Drop-in module `mpi_set` (the public interface of common/mpi_set.f90:7-19)
! for one process that holds the whole y range on one device.  The MPI
! handles it publishes are those of the single-rank module mpi, which the
! apps pass to their own MPI calls; the device library does not use them.
module mpi_set
  use iso_c_binding
  use wmpic_capi
  use mpi, only: MPI_INTEGER, MPI_DOUBLE_PRECISION, MPI_CHARACTER, MPI_SUM, MPI_WTIME
  implicit none
  private
  public :: mpi_set__init, MPI_WTIME

  integer, public, parameter   :: mnpi  = MPI_INTEGER
  integer, public, parameter   :: mnpr  = MPI_DOUBLE_PRECISION
  integer, public, parameter   :: mnpc  = MPI_CHARACTER
  integer, public, parameter   :: opsum = MPI_SUM
  integer, public              :: nerr = 0, ncomw = 0, nsize = 1, nrank = 0
  integer, public              :: nstat(1) = 0
  integer, public              :: nys, nye, nup, ndown

contains

  subroutine mpi_set__init(nygs,nyge,nproc)
    integer, intent(in) :: nygs, nyge, nproc
    integer(c_int) :: a, b, c, d, e, f
    if (wm_mpi_set_init(nygs, nyge, nproc, a, b, c, d, e, f) /= 0) then
       stop 'error in proc no.'
    endif
    nys = a; nye = b; nup = c; ndown = d; nrank = e; nsize = f
  end subroutine mpi_set__init

end module mpi_set
