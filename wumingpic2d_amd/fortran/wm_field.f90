! Drop-in module `field` (the public interface of common/field.f90:22-95).
! The three boundary procedures keep their interface blocks so that callers
! compile unchanged; the device runs the boundary family registered by the
! boundary module's initialiser (periodic: boundary_periodic__init), since
! host procedures cannot be called on device data once per CG iteration.
module field
  use iso_c_binding
  use wmpic_capi
  implicit none
  private
  public :: field__init, field__fdtd_i

  integer, save :: ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye

contains

  subroutine field__init(ndim_in,np_in,nsp_in,nxgs_in,nxge_in,nygs_in,nyge_in,nys_in,nye_in, &
                         mnpr_in,ncomw_in,opsum_in,nerr_in,                                  &
                         delx_in,delt_in,c_in,q_in,r_in,gfac_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in
    integer, intent(in) :: mnpr_in, ncomw_in, opsum_in, nerr_in
    real(8), intent(in) :: delx_in, delt_in, c_in, q_in(nsp_in), r_in(nsp_in), gfac_in
    call wmpic_context()
    call wmpic_check(wm_field_init(wm_ctx, ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, &
         nyge_in, nys_in, nye_in, mnpr_in, ncomw_in, opsum_in, nerr_in, delx_in, delt_in, c_in, &
         q_in, r_in, gfac_in))
    ndim = ndim_in; np = np_in; nsp = nsp_in
    nxgs = nxgs_in; nxge = nxge_in; nygs = nygs_in; nyge = nyge_in; nys = nys_in; nye = nye_in
  end subroutine field__init

  subroutine field__fdtd_i(uf,up,gp,cumcnt,nxs,nxe, &
       & set_boundary_dfield, &
       & set_boundary_curre, &
       & set_boundary_phi)

    interface
       subroutine set_boundary_dfield(df,nxs,nxe,nys,nye,nxgs,nxge)
         real(8), intent(inout) :: df(6,nxgs-2:nxge+2,nys-2:nye+2)
         integer, intent(in)    :: nxs, nxe, nys, nye, nxgs, nxge
       end subroutine set_boundary_dfield
       subroutine set_boundary_curre(uj,nxs,nxe,nys,nye,nxgs,nxge)
         real(8), intent(inout) :: uj(3,nxgs-2:nxge+2,nys-2:nye+2)
         integer, intent(in)    :: nxs, nxe, nys, nye, nxgs, nxge
       end subroutine set_boundary_curre
       subroutine set_boundary_phi(phi,nxs,nxe,nys,nye,l)
         real(8), intent(inout) :: phi(nxs-1:nxe+1,nys-1:nye+1)
         integer, intent(in)    :: nxs, nxe, nys, nye, l
       end subroutine set_boundary_phi
    end interface

    integer, intent(in)    :: nxs, nxe
    integer, intent(in)    :: cumcnt(nxgs:nxge+1,nys:nye,nsp)
    real(8), intent(in)    :: gp(ndim,np,nys:nye,nsp)
    real(8), intent(in)    :: up(ndim,np,nys:nye,nsp)
    real(8), intent(inout) :: uf(6,nxgs-2:nxge+2,nys-2:nye+2)
    call wmpic_context()
    ! the set_boundary_* arguments are the initialised boundary module's
    ! (boundary_periodic, boundary_shock or boundary_reconnection), which
    ! the device applies itself
    call wmpic_check(wm_field_fdtd_i(wm_ctx, uf, up, gp, cumcnt, nxs, nxe, WM_BC_INITIALISED))
  end subroutine field__fdtd_i

end module field
