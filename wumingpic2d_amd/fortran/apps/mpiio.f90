! Module `mpiio` (utils/iocore/mpiio.f90's interface) for one process:
! files opened for stream access, scalars and arrays written and read at a
! byte displacement that advances past them.
module mpiio
  use iso_fortran_env, only: int32, int64, real32, real64
  implicit none
  private

  public :: mpiio_open_file, mpiio_close_file, mpiio_get_endian_flag
  public :: mpiio_write_atomic, mpiio_read_atomic

  interface mpiio_write_atomic
     module procedure w_i4, w_i8, w_r4, w_r8, w_i4s, w_i8s, w_r4s, w_r8s, w_char
  end interface mpiio_write_atomic
  interface mpiio_read_atomic
     module procedure r_i4, r_i8, r_r4, r_r8, r_i4s, r_i8s, r_r4s, r_r8s
  end interface mpiio_read_atomic

contains

  ! 1 on a little-endian machine, 16777216 on a big-endian one
  integer function mpiio_get_endian_flag() result(flag)
    flag = transfer([1_1, 0_1, 0_1, 0_1], flag)
  end function mpiio_get_endian_flag

  ! mode 'r' (read), 'w' (new or truncated), 'a' (append: disp = its size)
  subroutine mpiio_open_file(filename, fh, disp, mode)
    character(len=*), intent(in) :: filename, mode
    integer, intent(out) :: fh
    integer(int64), intent(out) :: disp
    logical :: exists
    integer(int64) :: sz
    disp = 0
    select case (mode)
    case ('r')
       open(newunit=fh, file=filename, access='stream', form='unformatted', status='old', &
            action='read')
    case ('w')
       open(newunit=fh, file=filename, access='stream', form='unformatted', status='replace', &
            action='readwrite')
    case ('a')
       inquire(file=filename, exist=exists)
       if (exists) then
          open(newunit=fh, file=filename, access='stream', form='unformatted', status='old', &
               action='readwrite')
          inquire(unit=fh, size=sz)
          disp = max(sz, 0_int64)
       else
          open(newunit=fh, file=filename, access='stream', form='unformatted', status='new', &
               action='readwrite')
       end if
    case default
       write(0,*) 'Error: invalid file open mode ', mode
       stop
    end select
  end subroutine mpiio_open_file

  subroutine mpiio_close_file(fh)
    integer, intent(in) :: fh
    close(fh)
  end subroutine mpiio_close_file

  subroutine w_i4(fh, disp, x)
    integer, intent(in) :: fh
    integer(int64), intent(inout) :: disp
    integer(int32), intent(in) :: x
    write(fh, pos=disp + 1) x
    disp = disp + 4
  end subroutine w_i4
  subroutine w_i8(fh, disp, x)
    integer, intent(in) :: fh
    integer(int64), intent(inout) :: disp
    integer(int64), intent(in) :: x
    write(fh, pos=disp + 1) x
    disp = disp + 8
  end subroutine w_i8
  subroutine w_r4(fh, disp, x)
    integer, intent(in) :: fh
    integer(int64), intent(inout) :: disp
    real(real32), intent(in) :: x
    write(fh, pos=disp + 1) x
    disp = disp + 4
  end subroutine w_r4
  subroutine w_r8(fh, disp, x)
    integer, intent(in) :: fh
    integer(int64), intent(inout) :: disp
    real(real64), intent(in) :: x
    write(fh, pos=disp + 1) x
    disp = disp + 8
  end subroutine w_r8
  subroutine w_i4s(fh, disp, x)
    integer, intent(in) :: fh
    integer(int64), intent(inout) :: disp
    integer(int32), intent(in) :: x(:)
    write(fh, pos=disp + 1) x
    disp = disp + 4 * size(x, kind=int64)
  end subroutine w_i4s
  subroutine w_i8s(fh, disp, x)
    integer, intent(in) :: fh
    integer(int64), intent(inout) :: disp
    integer(int64), intent(in) :: x(:)
    write(fh, pos=disp + 1) x
    disp = disp + 8 * size(x, kind=int64)
  end subroutine w_i8s
  subroutine w_r4s(fh, disp, x)
    integer, intent(in) :: fh
    integer(int64), intent(inout) :: disp
    real(real32), intent(in) :: x(:)
    write(fh, pos=disp + 1) x
    disp = disp + 4 * size(x, kind=int64)
  end subroutine w_r4s
  subroutine w_r8s(fh, disp, x)
    integer, intent(in) :: fh
    integer(int64), intent(inout) :: disp
    real(real64), intent(in) :: x(:)
    write(fh, pos=disp + 1) x
    disp = disp + 8 * size(x, kind=int64)
  end subroutine w_r8s
  subroutine w_char(fh, disp, x)
    integer, intent(in) :: fh
    integer(int64), intent(inout) :: disp
    character(len=*), intent(in) :: x
    write(fh, pos=disp + 1) x
    disp = disp + len(x, kind=int64)
  end subroutine w_char

  subroutine r_i4(fh, disp, x)
    integer, intent(in) :: fh
    integer(int64), intent(inout) :: disp
    integer(int32), intent(out) :: x
    read(fh, pos=disp + 1) x
    disp = disp + 4
  end subroutine r_i4
  subroutine r_i8(fh, disp, x)
    integer, intent(in) :: fh
    integer(int64), intent(inout) :: disp
    integer(int64), intent(out) :: x
    read(fh, pos=disp + 1) x
    disp = disp + 8
  end subroutine r_i8
  subroutine r_r4(fh, disp, x)
    integer, intent(in) :: fh
    integer(int64), intent(inout) :: disp
    real(real32), intent(out) :: x
    read(fh, pos=disp + 1) x
    disp = disp + 4
  end subroutine r_r4
  subroutine r_r8(fh, disp, x)
    integer, intent(in) :: fh
    integer(int64), intent(inout) :: disp
    real(real64), intent(out) :: x
    read(fh, pos=disp + 1) x
    disp = disp + 8
  end subroutine r_r8
  subroutine r_i4s(fh, disp, x)
    integer, intent(in) :: fh
    integer(int64), intent(inout) :: disp
    integer(int32), intent(out) :: x(:)
    read(fh, pos=disp + 1) x
    disp = disp + 4 * size(x, kind=int64)
  end subroutine r_i4s
  subroutine r_i8s(fh, disp, x)
    integer, intent(in) :: fh
    integer(int64), intent(inout) :: disp
    integer(int64), intent(out) :: x(:)
    read(fh, pos=disp + 1) x
    disp = disp + 8 * size(x, kind=int64)
  end subroutine r_i8s
  subroutine r_r4s(fh, disp, x)
    integer, intent(in) :: fh
    integer(int64), intent(inout) :: disp
    real(real32), intent(out) :: x(:)
    read(fh, pos=disp + 1) x
    disp = disp + 4 * size(x, kind=int64)
  end subroutine r_r4s
  subroutine r_r8s(fh, disp, x)
    integer, intent(in) :: fh
    integer(int64), intent(inout) :: disp
    real(real64), intent(out) :: x(:)
    read(fh, pos=disp + 1) x
    disp = disp + 8 * size(x, kind=int64)
  end subroutine r_r8s

end module mpiio
