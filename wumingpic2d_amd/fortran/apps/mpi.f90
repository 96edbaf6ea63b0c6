! Module `mpi` for one process: the constants the reference's apps and
! utilities name, with the MPI routines as external procedures (implicit
! interfaces, as with mpif.h) implemented for MPI_COMM_WORLD of size 1 in
! wm_mpi_stub.c.  The apps link against it unchanged; the hot path's
! decomposition over ranks is the library's single device context.
module mpi
  implicit none
  public

  integer, parameter :: MPI_SUCCESS = 0
  integer, parameter :: MPI_COMM_WORLD = 0
  integer, parameter :: MPI_PROC_NULL = -1
  integer, parameter :: MPI_STATUS_SIZE = 6
  integer, parameter :: MPI_OFFSET_KIND = 8
  integer, parameter :: MPI_ADDRESS_KIND = 8
  integer, parameter :: MPI_INFO_NULL = 0
  ! datatypes: codes of wm_mpi_stub.c
  integer, parameter :: MPI_INTEGER = 1, MPI_INTEGER4 = 1
  integer, parameter :: MPI_INTEGER8 = 2
  integer, parameter :: MPI_DOUBLE_PRECISION = 3, MPI_REAL8 = 3
  integer, parameter :: MPI_REAL = 4, MPI_REAL4 = 4
  integer, parameter :: MPI_CHARACTER = 5
  integer, parameter :: MPI_LOGICAL = 6
  integer, parameter :: MPI_BYTE = 7
  ! reductions (one rank: the result is the rank's own data)
  integer, parameter :: MPI_SUM = 1, MPI_MAX = 2, MPI_MIN = 3, MPI_PROD = 4

  double precision, external :: MPI_Wtime
end module mpi
