! Module `wuming_utils` (the public interface of utils/wuming_utils.f90):
! elapsed time, the random seed, uniform and normal deviates, and a random
! permutation, for one process.
module wuming_utils
  implicit none
  private

  public :: get_etime
  public :: init_random_seed
  public :: uniform_rand
  public :: normal_rand
  public :: shuffle

  real(8), parameter :: twopi = 8*atan(1.0d0)

contains

  ! wall-clock seconds
  function get_etime() result(y)
    real(8) :: y
    integer(8) :: count, rate
    call system_clock(count, rate)
    y = dble(count) / dble(rate)
  end function get_etime

  ! the processor's default seed (one process: the reference's rank-0 seed)
  subroutine init_random_seed()
    call random_seed()
  end subroutine init_random_seed

  function uniform_rand() result(y)
    real(8) :: y
    call random_number(y)
  end function uniform_rand

  ! normal deviates in pairs (Box-Muller); the second of a pair is kept for
  ! the next call
  function normal_rand() result(y)
    real(8) :: y
    real(8) :: u(2), rad
    real(8), save :: kept
    logical, save :: have = .false.
    if (have) then
       y = kept
       have = .false.
       return
    end if
    call random_number(u)
    rad = sqrt(-2*log(1 - u(1)) + 1.0d-30)
    kept = rad * cos(twopi * u(2))
    y = rad * sin(twopi * u(2))
    have = .true.
  end function normal_rand

  ! a uniformly random permutation of x (Fisher-Yates)
  subroutine shuffle(x)
    integer, intent(inout) :: x(:)
    integer :: i, k, t
    real(8) :: u
    do i = size(x), 2, -1
       call random_number(u)
       k = 1 + min(i - 1, int(u * i))
       t = x(i)
       x(i) = x(k)
       x(k) = t
    end do
  end subroutine shuffle

end module wuming_utils
