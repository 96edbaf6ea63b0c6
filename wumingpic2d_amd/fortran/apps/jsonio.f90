! Module `jsonio` (utils/iocore/jsonio.f90's interface): the metadata of an
! attribute or dataset of a raw file (datatype, byte offset, size, shape,
! description) as a JSON object, and for attributes their value under
! "data".
module jsonio
  use iso_fortran_env, only: int32, int64, real32, real64
  use json_module
  implicit none
  private

  public :: json_IK, json_RK, json_core, json_value, json_file
  public :: jsonio_put_metadata, jsonio_put_attribute, jsonio_get_metadata
  public :: jsonio_get_attribute

  interface jsonio_put_metadata
     module procedure put_metadata_4, put_metadata_8
  end interface jsonio_put_metadata
  interface jsonio_put_attribute
     module procedure put_i4, put_i4s, put_i8, put_r8, put_r8s
  end interface jsonio_put_attribute
  interface jsonio_get_metadata
     module procedure get_metadata_4, get_metadata_8
  end interface jsonio_get_metadata
  interface jsonio_get_attribute
     module procedure get_i4, get_r8
  end interface jsonio_get_attribute

contains

  subroutine put_metadata_8(json, dst, name, datatype, disp, dsize, ndim, dshape, desc)
    type(json_core), intent(inout) :: json
    type(json_value), pointer :: dst
    character(len=*), intent(in) :: name, datatype, desc
    integer(int64), intent(in) :: disp, dsize
    integer(int64), intent(in) :: ndim
    integer(int64), intent(in) :: dshape(ndim)
    type(json_value), pointer :: obj
    call json%create_object(obj, name)
    call json%add(obj, 'datatype', datatype)
    call json%add(obj, 'offset', disp)
    call json%add(obj, 'size', dsize)
    call json%add(obj, 'ndim', ndim)
    call json%add(obj, 'shape', dshape)
    call json%add(obj, 'description', trim(desc))
    call json%add(dst, obj)
  end subroutine put_metadata_8

  subroutine put_metadata_4(json, dst, name, datatype, disp, dsize, ndim, dshape, desc)
    type(json_core), intent(inout) :: json
    type(json_value), pointer :: dst
    character(len=*), intent(in) :: name, datatype, desc
    integer(int64), intent(in) :: disp, dsize
    integer(int32), intent(in) :: ndim
    integer(int32), intent(in) :: dshape(ndim)
    call put_metadata_8(json, dst, name, datatype, disp, dsize, int(ndim, int64), &
                        int(dshape, int64), desc)
  end subroutine put_metadata_4

  subroutine put_i4(json, dst, data, name, disp, desc)
    type(json_core), intent(inout) :: json
    type(json_value), pointer :: dst
    integer(int32), intent(in) :: data
    character(len=*), intent(in) :: name, desc
    integer(int64), intent(in) :: disp
    call put_metadata_8(json, dst, name, 'i4', disp, 4_int64, 1_int64, [1_int64], desc)
    call json%add_by_path(dst, name // '.data', data)
  end subroutine put_i4

  subroutine put_i4s(json, dst, data, name, disp, desc)
    type(json_core), intent(inout) :: json
    type(json_value), pointer :: dst
    integer(int32), intent(in) :: data(:)
    character(len=*), intent(in) :: name, desc
    integer(int64), intent(in) :: disp
    call put_metadata_8(json, dst, name, 'i4', disp, 4 * size(data, kind=int64), 1_int64, &
                        [size(data, kind=int64)], desc)
    call json%add_by_path(dst, name // '.data', data)
  end subroutine put_i4s

  subroutine put_i8(json, dst, data, name, disp, desc)
    type(json_core), intent(inout) :: json
    type(json_value), pointer :: dst
    integer(int64), intent(in) :: data
    character(len=*), intent(in) :: name, desc
    integer(int64), intent(in) :: disp
    call put_metadata_8(json, dst, name, 'i8', disp, 8_int64, 1_int64, [1_int64], desc)
    call json%add_by_path(dst, name // '.data', data)
  end subroutine put_i8

  subroutine put_r8(json, dst, data, name, disp, desc)
    type(json_core), intent(inout) :: json
    type(json_value), pointer :: dst
    real(real64), intent(in) :: data
    character(len=*), intent(in) :: name, desc
    integer(int64), intent(in) :: disp
    call put_metadata_8(json, dst, name, 'f8', disp, 8_int64, 1_int64, [1_int64], desc)
    call json%add_by_path(dst, name // '.data', data)
  end subroutine put_r8

  subroutine put_r8s(json, dst, data, name, disp, desc)
    type(json_core), intent(inout) :: json
    type(json_value), pointer :: dst
    real(real64), intent(in) :: data(:)
    character(len=*), intent(in) :: name, desc
    integer(int64), intent(in) :: disp
    call put_metadata_8(json, dst, name, 'f8', disp, 8 * size(data, kind=int64), 1_int64, &
                        [size(data, kind=int64)], desc)
    call json%add_by_path(dst, name // '.data', data)
  end subroutine put_r8s

  subroutine get_metadata_8(json, src, name, disp, dsize, ndim, dshape)
    type(json_core), intent(inout) :: json
    type(json_value), pointer :: src
    character(len=*), intent(in) :: name
    integer(int64), intent(out) :: disp, dsize
    integer(int64), intent(out) :: ndim
    integer(int64), intent(out) :: dshape(:)
    integer(int64), allocatable :: s(:)
    integer(int32), allocatable :: s4(:)
    disp = 0
    dsize = 0
    ndim = 0
    dshape = 0
    call json%get(src, name // '.offset', disp)
    call json%get(src, name // '.size', dsize)
    call json%get(src, name // '.ndim', ndim)
    call json%get(src, name // '.shape', s4)
    if (allocated(s4)) then
       s = int(s4, int64)
       dshape(1:min(size(s), size(dshape))) = s(1:min(size(s), size(dshape)))
    end if
  end subroutine get_metadata_8

  subroutine get_metadata_4(json, src, name, disp, dsize, ndim, dshape)
    type(json_core), intent(inout) :: json
    type(json_value), pointer :: src
    character(len=*), intent(in) :: name
    integer(int64), intent(out) :: disp, dsize
    integer(int32), intent(out) :: ndim
    integer(int32), intent(out) :: dshape(:)
    integer(int64) :: n8, s8(size(dshape))
    call get_metadata_8(json, src, name, disp, dsize, n8, s8)
    ndim = int(n8, int32)
    dshape = int(s8, int32)
  end subroutine get_metadata_4

  subroutine get_i4(json, src, name, disp, data)
    type(json_core), intent(inout) :: json
    type(json_value), pointer :: src
    character(len=*), intent(in) :: name
    integer(int64), intent(out) :: disp
    integer(int32), intent(out) :: data
    data = 0
    disp = 0
    call json%get(src, name // '.offset', disp)
    call json%get(src, name // '.data', data)
  end subroutine get_i4

  subroutine get_r8(json, src, name, disp, data)
    type(json_core), intent(inout) :: json
    type(json_value), pointer :: src
    character(len=*), intent(in) :: name
    integer(int64), intent(out) :: disp
    real(real64), intent(out) :: data
    data = 0
    disp = 0
    call json%get(src, name // '.offset', disp)
    call json%get(src, name // '.data', data)
  end subroutine get_r8

end module jsonio
