! Drop-in module `boundary_reconnection` (the public interface of proj/reconnection/boundary_reconnection.f90:
! init :26-58, particle_x :61-99, particle_y :102-251, dfield :254-361,
! curre :364-502, phi :505-579, mom :582-647) for one rank holding the
! domain: the walls, ghosts and folds are device kernels.  Initialising it
! selects the family the field module's set_boundary_* arguments stand for.
module boundary_reconnection
  use iso_c_binding
  use wmpic_capi
  implicit none
  private
  public :: boundary_reconnection__init
  public :: boundary_reconnection__dfield
  public :: boundary_reconnection__particle_x
  public :: boundary_reconnection__particle_y
  public :: boundary_reconnection__curre
  public :: boundary_reconnection__phi
  public :: boundary_reconnection__mom

  integer, save :: ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye

contains

  subroutine boundary_reconnection__init(ndim_in,np_in,nsp_in,nxgs_in,nxge_in,nygs_in,nyge_in,nys_in,nye_in, &
                            nup_in,ndown_in,mnpi_in,mnpr_in,ncomw_in,nerr_in,nstat_in,          &
                            delx_in,delt_in,c_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in
    integer, intent(in) :: nup_in, ndown_in, mnpi_in, mnpr_in, ncomw_in, nerr_in, nstat_in(:)
    real(8), intent(in) :: delx_in, delt_in, c_in
    integer(c_int) :: ns(max(1, size(nstat_in)))
    ns = 0
    if (size(nstat_in) > 0) ns(1:size(nstat_in)) = nstat_in
    call wmpic_context()
    call wmpic_check(wm_bcr_init(wm_ctx, ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, &
         nyge_in, nys_in, nye_in, nup_in, ndown_in, mnpi_in, mnpr_in, ncomw_in, nerr_in, ns, &
         int(size(nstat_in), c_int), delx_in, delt_in, c_in))
    ndim = ndim_in; np = np_in; nsp = nsp_in
    nxgs = nxgs_in; nxge = nxge_in; nygs = nygs_in; nyge = nyge_in; nys = nys_in; nye = nye_in
  end subroutine boundary_reconnection__init

  subroutine boundary_reconnection__particle_x(up,np2,nxs,nxe)
    integer, intent(in)    :: nxs, nxe
    integer, intent(in)    :: np2(nys:nye,nsp)
    real(8), intent(inout) :: up(ndim,np,nys:nye,nsp)
    call wmpic_context()
    call wmpic_check(wm_bcr_particle_x(wm_ctx, up, np2, nxs, nxe))
  end subroutine boundary_reconnection__particle_x

  subroutine boundary_reconnection__particle_y(up,np2)
    integer, intent(inout) :: np2(nys:nye,nsp)
    real(8), intent(inout) :: up(ndim,np,nys:nye,nsp)
    call wmpic_context()
    call wmpic_check(wm_bcr_particle_y(wm_ctx, up, np2))
  end subroutine boundary_reconnection__particle_y

  subroutine boundary_reconnection__dfield(df,nxs,nxe,nys,nye,nxgs,nxge)
    integer, intent(in)    :: nxs, nxe, nys, nye, nxgs, nxge
    real(8), intent(inout) :: df(6,nxgs-2:nxge+2,nys-2:nye+2)
    call wmpic_context()
    call wmpic_check(wm_bcr_dfield(wm_ctx, df, nxs, nxe, nys, nye, nxgs, nxge))
  end subroutine boundary_reconnection__dfield

  subroutine boundary_reconnection__curre(uj,nxs,nxe,nys,nye,nxgs,nxge)
    integer, intent(in)    :: nxs, nxe, nys, nye, nxgs, nxge
    real(8), intent(inout) :: uj(3,nxgs-2:nxge+2,nys-2:nye+2)
    call wmpic_context()
    call wmpic_check(wm_bcr_curre(wm_ctx, uj, nxs, nxe, nys, nye, nxgs, nxge))
  end subroutine boundary_reconnection__curre

  subroutine boundary_reconnection__phi(phi,nxs,nxe,nys,nye,l)
    integer, intent(in)    :: nxs, nxe, nys, nye, l
    real(8), intent(inout) :: phi(nxs-1:nxe+1,nys-1:nye+1)
    call wmpic_context()
    call wmpic_check(wm_bcr_phi(wm_ctx, phi, nxs, nxe, nys, nye, l))
  end subroutine boundary_reconnection__phi

  subroutine boundary_reconnection__mom(mom)
    real(8), intent(inout) :: mom(7,nxgs-1:nxge+1,nys-1:nye+1,nsp)
    call wmpic_context()
    call wmpic_check(wm_bcr_mom(wm_ctx, mom))
  end subroutine boundary_reconnection__mom

end module boundary_reconnection
