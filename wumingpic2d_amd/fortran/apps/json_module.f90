! Module `json_module`: the part of the json-fortran 8.2.1 API that the
! reference's apps and I/O modules call (json_core, json_file, json_value;
! proj/weibel/app.f90:162-258, 550-627), over libwmpic's JSON document model
! (include/wmpic_json.h).  A json_value is a handle on a node of a document;
! documents are freed through json_file%destroy or json_core%destroy(root).
module json_module
  use iso_c_binding
  use iso_fortran_env, only: int32, int64, real64
  implicit none
  private

  integer, parameter, public :: json_IK = int32
  integer, parameter, public :: json_RK = real64
  integer, parameter, public :: json_CK = kind('a')
  integer, parameter, public :: json_LK = kind(.true.)

  type, public :: json_value
     type(c_ptr) :: h = c_null_ptr
     character(len=:), allocatable :: name   ! the key it takes when added
  end type json_value

  type, public :: json_core
     logical :: err = .false.
   contains
     procedure :: initialize => core_initialize
     procedure :: destroy => core_destroy
     procedure :: failed => core_failed
     procedure :: print => core_print
     procedure :: serialize => core_serialize
     procedure :: create_object => core_create_object
     procedure :: create_array => core_create_array
     procedure, private :: add_child, add_int, add_int64, add_real, add_str, add_logical
     procedure, private :: add_ints, add_int64s, add_reals
     generic :: add => add_child, add_int, add_int64, add_real, add_str, add_logical, add_ints, &
                       add_int64s, add_reals
     procedure, private :: get_child, get_int, get_int64, get_real, get_str, get_logical
     procedure, private :: get_ints, get_reals
     generic :: get => get_child, get_int, get_int64, get_real, get_str, get_logical, get_ints, &
                       get_reals
     procedure, private :: update_str, update_int, update_real
     generic :: update => update_str, update_int, update_real
     procedure, private :: abp_int, abp_int64, abp_real, abp_ints, abp_int64s, abp_reals, abp_str
     generic :: add_by_path => abp_int, abp_int64, abp_real, abp_ints, abp_int64s, abp_reals, &
                               abp_str
  end type json_core

  type, public :: json_file
     type(c_ptr) :: root = c_null_ptr
     logical :: err = .false.
   contains
     procedure :: initialize => file_initialize
     procedure :: load => file_load
     procedure :: deserialize => file_deserialize
     procedure :: destroy => file_destroy
     procedure :: failed => file_failed
     procedure, private :: file_get_root, file_get_int, file_get_real, file_get_str
     generic :: get => file_get_root, file_get_int, file_get_real, file_get_str
  end type json_file

  interface
     type(c_ptr) function c_new_object() bind(C, name="wm_json_new_object")
       import :: c_ptr
     end function c_new_object
     type(c_ptr) function c_load(path) bind(C, name="wm_json_load")
       import :: c_ptr, c_char
       character(kind=c_char), intent(in) :: path(*)
     end function c_load
     type(c_ptr) function c_parse(text, n) bind(C, name="wm_json_parse")
       import :: c_ptr, c_char, c_int
       character(kind=c_char), intent(in) :: text(*)
       integer(c_int), value :: n
     end function c_parse
     subroutine c_free(v) bind(C, name="wm_json_free")
       import :: c_ptr
       type(c_ptr), value :: v
     end subroutine c_free
     type(c_ptr) function c_child(v, path) bind(C, name="wm_json_child")
       import :: c_ptr, c_char
       type(c_ptr), value :: v
       character(kind=c_char), intent(in) :: path(*)
     end function c_child
     integer(c_int) function c_get_int(v, path, out) bind(C, name="wm_json_get_int")
       import :: c_ptr, c_char, c_int, c_long_long
       type(c_ptr), value :: v
       character(kind=c_char), intent(in) :: path(*)
       integer(c_long_long), intent(out) :: out
     end function c_get_int
     integer(c_int) function c_get_real(v, path, out) bind(C, name="wm_json_get_real")
       import :: c_ptr, c_char, c_int, c_double
       type(c_ptr), value :: v
       character(kind=c_char), intent(in) :: path(*)
       real(c_double), intent(out) :: out
     end function c_get_real
     integer(c_int) function c_get_bool(v, path, out) bind(C, name="wm_json_get_bool")
       import :: c_ptr, c_char, c_int
       type(c_ptr), value :: v
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int), intent(out) :: out
     end function c_get_bool
     integer(c_int) function c_get_str(v, path, buf, cap) bind(C, name="wm_json_get_str")
       import :: c_ptr, c_char, c_int
       type(c_ptr), value :: v
       character(kind=c_char), intent(in) :: path(*)
       character(kind=c_char) :: buf(*)
       integer(c_int), value :: cap
     end function c_get_str
     integer(c_int) function c_array_size(v, path) bind(C, name="wm_json_array_size")
       import :: c_ptr, c_char, c_int
       type(c_ptr), value :: v
       character(kind=c_char), intent(in) :: path(*)
     end function c_array_size
     integer(c_int) function c_add_child(v, name, child) bind(C, name="wm_json_add_child")
       import :: c_ptr, c_char, c_int
       type(c_ptr), value :: v, child
       character(kind=c_char), intent(in) :: name(*)
     end function c_add_child
     integer(c_int) function c_add_int(v, name, x) bind(C, name="wm_json_add_int")
       import :: c_ptr, c_char, c_int, c_long_long
       type(c_ptr), value :: v
       character(kind=c_char), intent(in) :: name(*)
       integer(c_long_long), value :: x
     end function c_add_int
     integer(c_int) function c_add_real(v, name, x) bind(C, name="wm_json_add_real")
       import :: c_ptr, c_char, c_int, c_double
       type(c_ptr), value :: v
       character(kind=c_char), intent(in) :: name(*)
       real(c_double), value :: x
     end function c_add_real
     integer(c_int) function c_add_bool(v, name, x) bind(C, name="wm_json_add_bool")
       import :: c_ptr, c_char, c_int
       type(c_ptr), value :: v
       character(kind=c_char), intent(in) :: name(*)
       integer(c_int), value :: x
     end function c_add_bool
     integer(c_int) function c_add_str(v, name, x) bind(C, name="wm_json_add_str")
       import :: c_ptr, c_char, c_int
       type(c_ptr), value :: v
       character(kind=c_char), intent(in) :: name(*), x(*)
     end function c_add_str
     integer(c_int) function c_add_ints(v, name, x, n) bind(C, name="wm_json_add_ints")
       import :: c_ptr, c_char, c_int, c_long_long
       type(c_ptr), value :: v
       character(kind=c_char), intent(in) :: name(*)
       integer(c_long_long), intent(in) :: x(*)
       integer(c_int), value :: n
     end function c_add_ints
     integer(c_int) function c_add_reals(v, name, x, n) bind(C, name="wm_json_add_reals")
       import :: c_ptr, c_char, c_int, c_double
       type(c_ptr), value :: v
       character(kind=c_char), intent(in) :: name(*)
       real(c_double), intent(in) :: x(*)
       integer(c_int), value :: n
     end function c_add_reals
     integer(c_int) function c_update_str(v, path, x) bind(C, name="wm_json_update_str")
       import :: c_ptr, c_char, c_int
       type(c_ptr), value :: v
       character(kind=c_char), intent(in) :: path(*), x(*)
     end function c_update_str
     integer(c_int) function c_update_int(v, path, x) bind(C, name="wm_json_update_int")
       import :: c_ptr, c_char, c_int, c_long_long
       type(c_ptr), value :: v
       character(kind=c_char), intent(in) :: path(*)
       integer(c_long_long), value :: x
     end function c_update_int
     integer(c_int) function c_update_real(v, path, x) bind(C, name="wm_json_update_real")
       import :: c_ptr, c_char, c_int, c_double
       type(c_ptr), value :: v
       character(kind=c_char), intent(in) :: path(*)
       real(c_double), value :: x
     end function c_update_real
     integer(c_int) function c_print(v, path) bind(C, name="wm_json_print")
       import :: c_ptr, c_char, c_int
       type(c_ptr), value :: v
       character(kind=c_char), intent(in) :: path(*)
     end function c_print
     integer(c_int) function c_serialize(v, buf, cap) bind(C, name="wm_json_serialize")
       import :: c_ptr, c_char, c_int
       type(c_ptr), value :: v
       character(kind=c_char) :: buf(*)
       integer(c_int), value :: cap
     end function c_serialize
  end interface

contains

  function cs(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len(s) + 1) :: c
    c = s // c_null_char
  end function cs

  ! a handle of the node at path below p (p itself for an empty path)
  function node(p, path) result(h)
    type(json_value), pointer, intent(in) :: p
    character(len=*), intent(in) :: path
    type(c_ptr) :: h
    h = c_null_ptr
    if (associated(p)) h = c_child(p%h, cs(path))
  end function node

  subroutine mark(self, ok, found)
    class(json_core), intent(inout) :: self
    logical, intent(in) :: ok
    logical, intent(out), optional :: found
    if (present(found)) then
       found = ok
    else if (.not. ok) then
       self%err = .true.
    end if
  end subroutine mark

  ! ------------------------------------------------------------ json_core
  subroutine core_initialize(self)
    class(json_core), intent(inout) :: self
    self%err = .false.
  end subroutine core_initialize

  subroutine core_destroy(self, p)
    class(json_core), intent(inout) :: self
    type(json_value), pointer, optional :: p
    self%err = .false.
    if (present(p)) then
       if (associated(p)) then
          call c_free(p%h)
          deallocate(p)
       end if
       nullify(p)
    end if
  end subroutine core_destroy

  logical function core_failed(self)
    class(json_core), intent(in) :: self
    core_failed = self%err
  end function core_failed

  subroutine core_print(self, p, filename)
    class(json_core), intent(inout) :: self
    type(json_value), pointer, intent(in) :: p
    character(len=*), intent(in) :: filename
    if (.not. associated(p)) then
       self%err = .true.
       return
    end if
    if (c_print(p%h, cs(filename)) == 0) self%err = .true.
  end subroutine core_print

  subroutine core_serialize(self, p, str)
    class(json_core), intent(inout) :: self
    type(json_value), pointer, intent(in) :: p
    character(len=:), allocatable, intent(out) :: str
    integer :: n
    character(kind=c_char), allocatable :: buf(:)
    str = ''
    if (.not. associated(p)) then
       self%err = .true.
       return
    end if
    n = c_serialize(p%h, [c_null_char], 0)
    allocate(buf(max(n, 1)))
    n = c_serialize(p%h, buf, n)
    allocate(character(len=n) :: str)
    str = transfer(buf(1:n), str)
  end subroutine core_serialize

  subroutine core_create_object(self, p, name)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: name
    allocate(p)
    p%h = c_new_object()
    p%name = name
    if (.false.) self%err = .true.
  end subroutine core_create_object

  subroutine core_create_array(self, p, name)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: name
    allocate(p)
    p%h = c_parse(cs('[]'), 2)
    p%name = name
    if (.false.) self%err = .true.
  end subroutine core_create_array

  subroutine add_child(self, p, child)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p, child
    if (.not. (associated(p) .and. associated(child))) then
       self%err = .true.
       return
    end if
    if (c_add_child(p%h, cs(child%name), child%h) == 0) self%err = .true.
  end subroutine add_child

  subroutine add_int(self, p, name, val)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: name
    integer(int32), intent(in) :: val
    call mark(self, c_add_int(p%h, cs(name), int(val, c_long_long)) /= 0)
  end subroutine add_int

  subroutine add_int64(self, p, name, val)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: name
    integer(int64), intent(in) :: val
    call mark(self, c_add_int(p%h, cs(name), int(val, c_long_long)) /= 0)
  end subroutine add_int64

  subroutine add_real(self, p, name, val)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: name
    real(real64), intent(in) :: val
    call mark(self, c_add_real(p%h, cs(name), val) /= 0)
  end subroutine add_real

  subroutine add_str(self, p, name, val)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: name, val
    call mark(self, c_add_str(p%h, cs(name), cs(val)) /= 0)
  end subroutine add_str

  subroutine add_logical(self, p, name, val)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: name
    logical, intent(in) :: val
    call mark(self, c_add_bool(p%h, cs(name), merge(1_c_int, 0_c_int, val)) /= 0)
  end subroutine add_logical

  subroutine add_ints(self, p, name, val)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: name
    integer(int32), intent(in) :: val(:)
    call mark(self, c_add_ints(p%h, cs(name), int(val, c_long_long), size(val)) /= 0)
  end subroutine add_ints

  subroutine add_int64s(self, p, name, val)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: name
    integer(int64), intent(in) :: val(:)
    call mark(self, c_add_ints(p%h, cs(name), int(val, c_long_long), size(val)) /= 0)
  end subroutine add_int64s

  subroutine add_reals(self, p, name, val)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: name
    real(real64), intent(in) :: val(:)
    call mark(self, c_add_reals(p%h, cs(name), val, size(val)) /= 0)
  end subroutine add_reals

  ! add_by_path: a member (path 'a.b' adds member b of a)
  subroutine abp_int(self, p, path, val)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: path
    integer(int32), intent(in) :: val
    call mark(self, c_update_int(p%h, cs(path), int(val, c_long_long)) >= 0)
  end subroutine abp_int

  subroutine abp_int64(self, p, path, val)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: path
    integer(int64), intent(in) :: val
    call mark(self, c_update_int(p%h, cs(path), int(val, c_long_long)) >= 0)
  end subroutine abp_int64

  subroutine abp_real(self, p, path, val)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: path
    real(real64), intent(in) :: val
    call mark(self, c_update_real(p%h, cs(path), val) >= 0)
  end subroutine abp_real

  subroutine abp_str(self, p, path, val)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: path, val
    call mark(self, c_update_str(p%h, cs(path), cs(val)) >= 0)
  end subroutine abp_str

  subroutine abp_array(self, p, path, ints, reals)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: path
    integer(int64), intent(in), optional :: ints(:)
    real(real64), intent(in), optional :: reals(:)
    type(c_ptr) :: parent
    integer :: dot
    dot = index(path, '.', back=.true.)
    parent = p%h
    if (dot > 0) parent = c_child(p%h, cs(path(1:dot-1)))
    if (present(ints)) then
       call mark(self, c_add_ints(parent, cs(path(dot+1:)), int(ints, c_long_long), size(ints)) /= 0)
    else
       call mark(self, c_add_reals(parent, cs(path(dot+1:)), reals, size(reals)) /= 0)
    end if
  end subroutine abp_array

  subroutine abp_ints(self, p, path, val)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: path
    integer(int32), intent(in) :: val(:)
    call abp_array(self, p, path, ints=int(val, int64))
  end subroutine abp_ints

  subroutine abp_int64s(self, p, path, val)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: path
    integer(int64), intent(in) :: val(:)
    call abp_array(self, p, path, ints=val)
  end subroutine abp_int64s

  subroutine abp_reals(self, p, path, val)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: path
    real(real64), intent(in) :: val(:)
    call abp_array(self, p, path, reals=val)
  end subroutine abp_reals

  subroutine get_child(self, p, path, child, found)
    class(json_core), intent(inout) :: self
    type(json_value), pointer, intent(in) :: p
    character(len=*), intent(in) :: path
    type(json_value), pointer :: child
    logical, intent(out), optional :: found
    type(c_ptr) :: h
    h = node(p, path)
    nullify(child)
    if (c_associated(h)) then
       allocate(child)
       child%h = h
       child%name = path
    end if
    call mark(self, c_associated(h), found)
  end subroutine get_child

  subroutine get_int(self, p, path, val, found)
    class(json_core), intent(inout) :: self
    type(json_value), pointer, intent(in) :: p
    character(len=*), intent(in) :: path
    integer(int32), intent(inout) :: val
    logical, intent(out), optional :: found
    integer(c_long_long) :: v
    logical :: ok
    ok = .false.
    if (associated(p)) ok = c_get_int(p%h, cs(path), v) /= 0
    if (ok) val = int(v, int32)
    call mark(self, ok, found)
  end subroutine get_int

  subroutine get_int64(self, p, path, val, found)
    class(json_core), intent(inout) :: self
    type(json_value), pointer, intent(in) :: p
    character(len=*), intent(in) :: path
    integer(int64), intent(inout) :: val
    logical, intent(out), optional :: found
    integer(c_long_long) :: v
    logical :: ok
    ok = .false.
    if (associated(p)) ok = c_get_int(p%h, cs(path), v) /= 0
    if (ok) val = int(v, int64)
    call mark(self, ok, found)
  end subroutine get_int64

  subroutine get_real(self, p, path, val, found)
    class(json_core), intent(inout) :: self
    type(json_value), pointer, intent(in) :: p
    character(len=*), intent(in) :: path
    real(real64), intent(inout) :: val
    logical, intent(out), optional :: found
    real(c_double) :: v
    logical :: ok
    ok = .false.
    if (associated(p)) ok = c_get_real(p%h, cs(path), v) /= 0
    if (ok) val = v
    call mark(self, ok, found)
  end subroutine get_real

  subroutine get_logical(self, p, path, val, found)
    class(json_core), intent(inout) :: self
    type(json_value), pointer, intent(in) :: p
    character(len=*), intent(in) :: path
    logical, intent(inout) :: val
    logical, intent(out), optional :: found
    integer(c_int) :: v
    logical :: ok
    ok = .false.
    if (associated(p)) ok = c_get_bool(p%h, cs(path), v) /= 0
    if (ok) val = v /= 0
    call mark(self, ok, found)
  end subroutine get_logical

  subroutine get_str(self, p, path, val, found)
    class(json_core), intent(inout) :: self
    type(json_value), pointer, intent(in) :: p
    character(len=*), intent(in) :: path
    character(len=:), allocatable, intent(inout) :: val
    logical, intent(out), optional :: found
    integer :: n
    character(kind=c_char), allocatable :: buf(:)
    n = -1
    if (associated(p)) n = c_get_str(p%h, cs(path), [c_null_char], 0)
    if (n >= 0) then
       allocate(buf(max(n, 1)))
       n = c_get_str(p%h, cs(path), buf, n)
       if (allocated(val)) deallocate(val)
       allocate(character(len=n) :: val)
       if (n > 0) val = transfer(buf(1:n), val)
    end if
    call mark(self, n >= 0, found)
  end subroutine get_str

  subroutine get_ints(self, p, path, val, found)
    class(json_core), intent(inout) :: self
    type(json_value), pointer, intent(in) :: p
    character(len=*), intent(in) :: path
    integer(int32), allocatable, intent(inout) :: val(:)
    logical, intent(out), optional :: found
    integer :: n, k
    integer(c_long_long) :: v
    character(len=16) :: idx
    n = -1
    if (associated(p)) n = c_array_size(p%h, cs(path))
    if (n >= 0) then
       if (allocated(val)) deallocate(val)
       allocate(val(n))
       do k = 1, n
          write(idx, '("(", i0, ")")') k
          if (c_get_int(p%h, cs(path // trim(idx)), v) /= 0) val(k) = int(v, int32)
       end do
    end if
    call mark(self, n >= 0, found)
  end subroutine get_ints

  subroutine get_reals(self, p, path, val, found)
    class(json_core), intent(inout) :: self
    type(json_value), pointer, intent(in) :: p
    character(len=*), intent(in) :: path
    real(real64), allocatable, intent(inout) :: val(:)
    logical, intent(out), optional :: found
    integer :: n, k
    real(c_double) :: v
    character(len=16) :: idx
    n = -1
    if (associated(p)) n = c_array_size(p%h, cs(path))
    if (n >= 0) then
       if (allocated(val)) deallocate(val)
       allocate(val(n))
       do k = 1, n
          write(idx, '("(", i0, ")")') k
          if (c_get_real(p%h, cs(path // trim(idx)), v) /= 0) val(k) = v
       end do
    end if
    call mark(self, n >= 0, found)
  end subroutine get_reals

  subroutine update_str(self, p, path, val, found)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: path, val
    logical, intent(out) :: found
    found = c_update_str(p%h, cs(path), cs(val)) == 1
    if (.false.) self%err = .true.
  end subroutine update_str

  subroutine update_int(self, p, path, val, found)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: path
    integer(int32), intent(in) :: val
    logical, intent(out) :: found
    found = c_update_int(p%h, cs(path), int(val, c_long_long)) == 1
    if (.false.) self%err = .true.
  end subroutine update_int

  subroutine update_real(self, p, path, val, found)
    class(json_core), intent(inout) :: self
    type(json_value), pointer :: p
    character(len=*), intent(in) :: path
    real(real64), intent(in) :: val
    logical, intent(out) :: found
    found = c_update_real(p%h, cs(path), val) == 1
    if (.false.) self%err = .true.
  end subroutine update_real

  ! ------------------------------------------------------------ json_file
  subroutine file_initialize(self)
    class(json_file), intent(inout) :: self
    self%err = .false.
  end subroutine file_initialize

  subroutine file_load(self, filename)
    class(json_file), intent(inout) :: self
    character(len=*), intent(in) :: filename
    if (c_associated(self%root)) call c_free(self%root)
    self%root = c_load(cs(filename))
    self%err = .not. c_associated(self%root)
  end subroutine file_load

  subroutine file_deserialize(self, str)
    class(json_file), intent(inout) :: self
    character(len=*), intent(in) :: str
    if (c_associated(self%root)) call c_free(self%root)
    self%root = c_parse(str, len(str, c_int))
    self%err = .not. c_associated(self%root)
  end subroutine file_deserialize

  subroutine file_destroy(self)
    class(json_file), intent(inout) :: self
    if (c_associated(self%root)) call c_free(self%root)
    self%root = c_null_ptr
    self%err = .false.
  end subroutine file_destroy

  logical function file_failed(self)
    class(json_file), intent(in) :: self
    file_failed = self%err
  end function file_failed

  ! the document's root (owned by the file)
  subroutine file_get_root(self, p)
    class(json_file), intent(inout) :: self
    type(json_value), pointer :: p
    allocate(p)
    p%h = self%root
    p%name = ''
  end subroutine file_get_root

  subroutine file_get_int(self, path, val, found)
    class(json_file), intent(inout) :: self
    character(len=*), intent(in) :: path
    integer(int32), intent(inout) :: val
    logical, intent(out), optional :: found
    integer(c_long_long) :: v
    logical :: ok
    ok = c_get_int(self%root, cs(path), v) /= 0
    if (ok) val = int(v, int32)
    if (present(found)) then
       found = ok
    else if (.not. ok) then
       self%err = .true.
    end if
  end subroutine file_get_int

  subroutine file_get_real(self, path, val, found)
    class(json_file), intent(inout) :: self
    character(len=*), intent(in) :: path
    real(real64), intent(inout) :: val
    logical, intent(out), optional :: found
    real(c_double) :: v
    logical :: ok
    ok = c_get_real(self%root, cs(path), v) /= 0
    if (ok) val = v
    if (present(found)) then
       found = ok
    else if (.not. ok) then
       self%err = .true.
    end if
  end subroutine file_get_real

  subroutine file_get_str(self, path, val, found)
    class(json_file), intent(inout) :: self
    character(len=*), intent(in) :: path
    character(len=:), allocatable, intent(inout) :: val
    logical, intent(out), optional :: found
    type(json_core) :: core
    type(json_value), pointer :: r
    allocate(r)
    r%h = self%root
    call core%get(r, path, val, found)
    if (core%err) self%err = .true.
    deallocate(r)
  end subroutine file_get_str

end module json_module
