! Drop-in module `mpi_set` (the public interface of common/mpi_set.f90:7-19)
! for one process that holds the whole y range on one device.  The MPI
! handles it publishes are plain integers that the device library ignores.
module mpi_set
  use iso_c_binding
  use wmpic_capi
  implicit none
  private
  public :: mpi_set__init

  integer, public, parameter   :: mnpi  = 0
  integer, public, parameter   :: mnpr  = 0
  integer, public, parameter   :: mnpc  = 0
  integer, public, parameter   :: opsum = 0
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
