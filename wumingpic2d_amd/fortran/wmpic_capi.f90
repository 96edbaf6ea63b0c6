! wmpic_capi -- iso_c_binding interfaces of libwmpic (include/wmpic.h) and the
! device context shared by the drop-in modules particle, field, sort,
! boundary_periodic and mpi_set of this directory.
!
! Non-zero statuses end the run the way the reference's fatal paths do:
! the library's message (which for those paths is the reference's own text)
! is written to unit 6 and the program stops.
module wmpic_capi
  use iso_c_binding
  implicit none
  private

  public :: wm_ctx, wmpic_context, wmpic_check
  public :: wm_particle_init, wm_particle_solv, wm_field_init, wm_field_fdtd_i
  public :: wm_sort_init, wm_sort_bucket, wm_bcp_init, wm_bcp_particle_x
  public :: wm_bcp_particle_y, wm_bcp_dfield, wm_bcp_curre, wm_bcp_phi
  public :: wm_mpi_set_init, wm_set_capacity, WM_BC_PERIODIC, WM_BC_INITIALISED
  public :: wm_bcs_init, wm_bcs_particle_x, wm_bcs_injection, wm_bcs_particle_y, wm_bcs_dfield
  public :: wm_bcs_curre, wm_bcs_phi, wm_bcs_mom, wm_bcr_init, wm_bcr_particle_x
  public :: wm_bcr_particle_y, wm_bcr_dfield, wm_bcr_curre, wm_bcr_phi, wm_bcr_mom
  public :: wm_mom_calc_init, wm_mom_calc_accl, wm_mom_calc_nvt, wm_bcp_mom
  public :: wm_paraio_init, wm_paraio_output, wm_paraio_input, wm_paraio_param, wm_paraio_mom
  public :: wm_paraio_ptcl, wm_paraio_orb

  integer(c_int), parameter :: WM_BC_PERIODIC = 0
  ! field__fdtd_i: the family whose boundary module was initialised
  integer(c_int), parameter :: WM_BC_INITIALISED = -1
  type(c_ptr), save :: wm_ctx = c_null_ptr

  interface
     integer(c_int) function wm_create(ctx, device) bind(C, name="wm_create")
       import :: c_ptr, c_int
       type(c_ptr), intent(out) :: ctx
       integer(c_int), value :: device
     end function wm_create
     type(c_ptr) function wm_last_error(ctx) bind(C, name="wm_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function wm_last_error
     integer(c_int) function wm_set_capacity(ctx, cap, mcap) bind(C, name="wm_set_capacity")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int), value :: cap, mcap
     end function wm_set_capacity
     integer(c_int) function wm_mpi_set_init(nygs, nyge, nproc, nys, nye, nup, ndown, nrank, &
          nsize) bind(C, name="wm_mpi_set_init")
       import :: c_int
       integer(c_int), value :: nygs, nyge, nproc
       integer(c_int), intent(out) :: nys, nye, nup, ndown, nrank, nsize
     end function wm_mpi_set_init
     integer(c_int) function wm_particle_init(ctx, ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, &
          nye, delx, delt, c, q, r) bind(C, name="wm_particle_init")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye
       real(c_double), value :: delx, delt, c
       real(c_double), intent(in) :: q(*), r(*)
     end function wm_particle_init
     integer(c_int) function wm_particle_solv(ctx, gp, up, uf, cumcnt, nxs, nxe) &
          bind(C, name="wm_particle_solv")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: gp(*)
       real(c_double), intent(in) :: up(*), uf(*)
       integer(c_int), intent(in) :: cumcnt(*)
       integer(c_int), value :: nxs, nxe
     end function wm_particle_solv
     integer(c_int) function wm_field_init(ctx, ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, &
          mnpr, ncomw, opsum, nerr, delx, delt, c, q, r, gfac) bind(C, name="wm_field_init")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye
       integer(c_int), value :: mnpr, ncomw, opsum, nerr
       real(c_double), value :: delx, delt, c, gfac
       real(c_double), intent(in) :: q(*), r(*)
     end function wm_field_init
     integer(c_int) function wm_field_fdtd_i(ctx, uf, up, gp, cumcnt, nxs, nxe, family) &
          bind(C, name="wm_field_fdtd_i")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: uf(*)
       real(c_double), intent(in) :: up(*), gp(*)
       integer(c_int), intent(in) :: cumcnt(*)
       integer(c_int), value :: nxs, nxe, family
     end function wm_field_fdtd_i
     integer(c_int) function wm_sort_init(ctx, ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye) &
          bind(C, name="wm_sort_init")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int), value :: ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye
     end function wm_sort_init
     integer(c_int) function wm_sort_bucket(ctx, gp, up, cumcnt, np2, nxs, nxe) &
          bind(C, name="wm_sort_bucket")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: gp(*)
       real(c_double), intent(in) :: up(*)
       integer(c_int) :: cumcnt(*)
       integer(c_int), intent(in) :: np2(*)
       integer(c_int), value :: nxs, nxe
     end function wm_sort_bucket
     integer(c_int) function wm_bcp_init(ctx, ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, &
          nup, ndown, mnpi, mnpr, ncomw, nerr, nstat, nstat_len, delx, delt, c) &
          bind(C, name="wm_bcp_init")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye
       integer(c_int), value :: nup, ndown, mnpi, mnpr, ncomw, nerr, nstat_len
       integer(c_int), intent(in) :: nstat(*)
       real(c_double), value :: delx, delt, c
     end function wm_bcp_init
     integer(c_int) function wm_bcp_particle_x(ctx, up, np2) bind(C, name="wm_bcp_particle_x")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: up(*)
       integer(c_int), intent(in) :: np2(*)
     end function wm_bcp_particle_x
     integer(c_int) function wm_bcp_particle_y(ctx, up, np2) bind(C, name="wm_bcp_particle_y")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: up(*)
       integer(c_int) :: np2(*)
     end function wm_bcp_particle_y
     integer(c_int) function wm_bcp_dfield(ctx, df, nxs, nxe, nys, nye, nxgs, nxge) &
          bind(C, name="wm_bcp_dfield")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: df(*)
       integer(c_int), value :: nxs, nxe, nys, nye, nxgs, nxge
     end function wm_bcp_dfield
     integer(c_int) function wm_bcp_curre(ctx, uj, nxs, nxe, nys, nye, nxgs, nxge) &
          bind(C, name="wm_bcp_curre")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: uj(*)
       integer(c_int), value :: nxs, nxe, nys, nye, nxgs, nxge
     end function wm_bcp_curre
     integer(c_int) function wm_bcp_phi(ctx, phi, nxs, nxe, nys, nye, l) bind(C, name="wm_bcp_phi")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: phi(*)
       integer(c_int), value :: nxs, nxe, nys, nye, l
     end function wm_bcp_phi
     integer(c_int) function wm_bcs_init(ctx, ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, &
          nup, ndown, mnpi, mnpr, ncomw, nerr, nstat, nstat_len, delx, delt, c) &
          bind(C, name="wm_bcs_init")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye
       integer(c_int), value :: nup, ndown, mnpi, mnpr, ncomw, nerr, nstat_len
       integer(c_int), intent(in) :: nstat(*)
       real(c_double), value :: delx, delt, c
     end function wm_bcs_init
     integer(c_int) function wm_bcs_particle_x(ctx, up, np2, nxs, nxe) &
          bind(C, name="wm_bcs_particle_x")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: up(*)
       integer(c_int), intent(in) :: np2(*)
       integer(c_int), value :: nxs, nxe
     end function wm_bcs_particle_x
     integer(c_int) function wm_bcs_particle_y(ctx, up, np2) bind(C, name="wm_bcs_particle_y")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: up(*)
       integer(c_int) :: np2(*)
     end function wm_bcs_particle_y
     integer(c_int) function wm_bcs_dfield(ctx, df, nxs, nxe, nys, nye, nxgs, nxge) &
          bind(C, name="wm_bcs_dfield")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: df(*)
       integer(c_int), value :: nxs, nxe, nys, nye, nxgs, nxge
     end function wm_bcs_dfield
     integer(c_int) function wm_bcs_curre(ctx, uj, nxs, nxe, nys, nye, nxgs, nxge) &
          bind(C, name="wm_bcs_curre")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: uj(*)
       integer(c_int), value :: nxs, nxe, nys, nye, nxgs, nxge
     end function wm_bcs_curre
     integer(c_int) function wm_bcs_phi(ctx, phi, nxs, nxe, nys, nye, l) bind(C, name="wm_bcs_phi")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: phi(*)
       integer(c_int), value :: nxs, nxe, nys, nye, l
     end function wm_bcs_phi
     integer(c_int) function wm_bcs_mom(ctx, mom) bind(C, name="wm_bcs_mom")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: mom(*)
     end function wm_bcs_mom
     integer(c_int) function wm_bcr_init(ctx, ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, &
          nup, ndown, mnpi, mnpr, ncomw, nerr, nstat, nstat_len, delx, delt, c) &
          bind(C, name="wm_bcr_init")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye
       integer(c_int), value :: nup, ndown, mnpi, mnpr, ncomw, nerr, nstat_len
       integer(c_int), intent(in) :: nstat(*)
       real(c_double), value :: delx, delt, c
     end function wm_bcr_init
     integer(c_int) function wm_bcr_particle_x(ctx, up, np2, nxs, nxe) &
          bind(C, name="wm_bcr_particle_x")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: up(*)
       integer(c_int), intent(in) :: np2(*)
       integer(c_int), value :: nxs, nxe
     end function wm_bcr_particle_x
     integer(c_int) function wm_bcr_particle_y(ctx, up, np2) bind(C, name="wm_bcr_particle_y")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: up(*)
       integer(c_int) :: np2(*)
     end function wm_bcr_particle_y
     integer(c_int) function wm_bcr_dfield(ctx, df, nxs, nxe, nys, nye, nxgs, nxge) &
          bind(C, name="wm_bcr_dfield")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: df(*)
       integer(c_int), value :: nxs, nxe, nys, nye, nxgs, nxge
     end function wm_bcr_dfield
     integer(c_int) function wm_bcr_curre(ctx, uj, nxs, nxe, nys, nye, nxgs, nxge) &
          bind(C, name="wm_bcr_curre")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: uj(*)
       integer(c_int), value :: nxs, nxe, nys, nye, nxgs, nxge
     end function wm_bcr_curre
     integer(c_int) function wm_bcr_phi(ctx, phi, nxs, nxe, nys, nye, l) bind(C, name="wm_bcr_phi")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: phi(*)
       integer(c_int), value :: nxs, nxe, nys, nye, l
     end function wm_bcr_phi
     integer(c_int) function wm_bcr_mom(ctx, mom) bind(C, name="wm_bcr_mom")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: mom(*)
     end function wm_bcr_mom
     integer(c_int) function wm_bcs_injection(ctx, up, np2, nxs, nxe, u0) &
          bind(C, name="wm_bcs_injection")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: up(*)
       integer(c_int), intent(in) :: np2(*)
       integer(c_int), value :: nxs, nxe
       real(c_double), value :: u0
     end function wm_bcs_injection
     integer(c_int) function wm_mom_calc_init(ctx, ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, &
          nye, delx, delt, c, q, r) bind(C, name="wm_mom_calc_init")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye
       real(c_double), value :: delx, delt, c
       real(c_double), intent(in) :: q(*), r(*)
     end function wm_mom_calc_init
     integer(c_int) function wm_mom_calc_accl(ctx, gp, up, uf, cumcnt, nxs, nxe) &
          bind(C, name="wm_mom_calc_accl")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: gp(*)
       real(c_double), intent(in) :: up(*), uf(*)
       integer(c_int), intent(in) :: cumcnt(*)
       integer(c_int), value :: nxs, nxe
     end function wm_mom_calc_accl
     integer(c_int) function wm_mom_calc_nvt(ctx, mom, up, np2) bind(C, name="wm_mom_calc_nvt")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: mom(*)
       real(c_double), intent(in) :: up(*)
       integer(c_int), intent(in) :: np2(*)
     end function wm_mom_calc_nvt
     integer(c_int) function wm_paraio_init(ctx, ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, &
          nproc, nrank, delx, delt, c, q, r, dir) bind(C, name="wm_paraio_init")
       import :: c_ptr, c_int, c_double, c_char
       type(c_ptr), value :: ctx
       integer(c_int), value :: ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, nproc, nrank
       real(c_double), value :: delx, delt, c
       real(c_double), intent(in) :: q(*), r(*)
       character(kind=c_char), intent(in) :: dir(*)
     end function wm_paraio_init
     integer(c_int) function wm_paraio_output(ctx, up, uf, np2, nxs, nxe, it, filename) &
          bind(C, name="wm_paraio_output")
       import :: c_ptr, c_int, c_double, c_char
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: up(*), uf(*)
       integer(c_int), intent(in) :: np2(*)
       integer(c_int), value :: nxs, nxe, it
       character(kind=c_char), intent(in) :: filename(*)
     end function wm_paraio_output
     integer(c_int) function wm_paraio_input(ctx, up, uf, np2, nxs, nxe, it, filename) &
          bind(C, name="wm_paraio_input")
       import :: c_ptr, c_int, c_double, c_char
       type(c_ptr), value :: ctx
       real(c_double) :: up(*), uf(*)
       integer(c_int) :: np2(*)
       integer(c_int), intent(out) :: nxs, nxe, it
       character(kind=c_char), intent(in) :: filename(*)
     end function wm_paraio_input
     integer(c_int) function wm_paraio_param(ctx, n0, wpe, wpi, wge, wgi, vti, vte, filename) &
          bind(C, name="wm_paraio_param")
       import :: c_ptr, c_int, c_double, c_char
       type(c_ptr), value :: ctx
       integer(c_int), value :: n0
       real(c_double), value :: wpe, wpi, wge, wgi, vti, vte
       character(kind=c_char), intent(in) :: filename(*)
     end function wm_paraio_param
     integer(c_int) function wm_paraio_mom(ctx, mom, uf, it) bind(C, name="wm_paraio_mom")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: mom(*)
       real(c_double), intent(in) :: uf(*)
       integer(c_int), value :: it
     end function wm_paraio_mom
     integer(c_int) function wm_paraio_ptcl(ctx, up, uf, np2, it) bind(C, name="wm_paraio_ptcl")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: up(*), uf(*)
       integer(c_int), intent(in) :: np2(*)
       integer(c_int), value :: it
     end function wm_paraio_ptcl
     integer(c_int) function wm_paraio_orb(ctx, up, uf, np2, it) bind(C, name="wm_paraio_orb")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: up(*), uf(*)
       integer(c_int), intent(in) :: np2(*)
       integer(c_int), value :: it
     end function wm_paraio_orb
     integer(c_int) function wm_bcp_mom(ctx, mom) bind(C, name="wm_bcp_mom")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: mom(*)
     end function wm_bcp_mom
  end interface

contains

  ! The context of this process's device (created on first use).
  subroutine wmpic_context()
    integer(c_int) :: st
    integer :: dev, ios
    character(len=16) :: env
    if (c_associated(wm_ctx)) return
    dev = 0
    call get_environment_variable('WMPIC_DEVICE', env, status=ios)
    if (ios == 0) read(env, *) dev
    st = wm_create(wm_ctx, int(dev, c_int))
    if (st /= 0) then
       write(6,*) 'wmpic: no usable GPU'
       stop
    endif
    ! device slots per row: WMPIC_CAPACITY (default: the reference's np)
    call get_environment_variable('WMPIC_CAPACITY', env, status=ios)
    if (ios == 0) then
       read(env, *) dev
       st = wm_set_capacity(wm_ctx, int(dev, c_int), 0_c_int)
    endif
  end subroutine wmpic_context

  subroutine wmpic_check(st)
    integer(c_int), intent(in) :: st
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    character(len=512) :: text
    if (st == 0) return
    call c_f_pointer(wm_last_error(wm_ctx), msg, [512])
    text = ' '
    do n = 1, 512
       if (msg(n) == c_null_char) exit
       text(n:n) = msg(n)
    enddo
    write(6,*) trim(text)
    stop
  end subroutine wmpic_check

end module wmpic_capi
