! Drop-in module `mom_calc` (the public interface of common/mom_calc.f90:7-18,
! 48-54, 167-171) running the output-step moment diagnostics on the device
! through libwmpic.
module mom_calc
  use iso_c_binding
  use wmpic_capi
  implicit none
  private
  public :: mom_calc__init, mom_calc__accl, mom_calc__nvt

  integer, save :: ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye

contains

  subroutine mom_calc__init(ndim_in,np_in,nsp_in,nxgs_in,nxge_in,nygs_in,nyge_in,nys_in,nye_in, &
                            delx_in,delt_in,c_in,q_in,r_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in
    real(8), intent(in) :: delx_in, delt_in, c_in, q_in(nsp_in), r_in(nsp_in)
    call wmpic_context()
    call wmpic_check(wm_mom_calc_init(wm_ctx, ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, &
         nyge_in, nys_in, nye_in, delx_in, delt_in, c_in, q_in, r_in))
    ndim = ndim_in; np = np_in; nsp = nsp_in
    nxgs = nxgs_in; nxge = nxge_in; nygs = nygs_in; nyge = nyge_in; nys = nys_in; nye = nye_in
  end subroutine mom_calc__init

  subroutine mom_calc__accl(gp,up,uf,cumcnt,nxs,nxe)
    integer, intent(in)  :: nxs, nxe
    integer, intent(in)  :: cumcnt(nxgs:nxge+1,nys:nye,nsp)
    real(8), intent(in)  :: up(ndim,np,nys:nye,nsp)
    real(8), intent(in)  :: uf(6,nxgs-2:nxge+2,nys-2:nye+2)
    real(8), intent(out) :: gp(ndim,np,nys:nye,nsp)
    call wmpic_context()
    call wmpic_check(wm_mom_calc_accl(wm_ctx, gp, up, uf, cumcnt, nxs, nxe))
  end subroutine mom_calc__accl

  subroutine mom_calc__nvt(mom,up,np2)
    integer, intent(in)  :: np2(nys:nye,nsp)
    real(8), intent(in)  :: up(ndim,np,nys:nye,nsp)
    real(8), intent(out) :: mom(7,nxgs-1:nxge+1,nys-1:nye+1,1:nsp)
    call wmpic_context()
    call wmpic_check(wm_mom_calc_nvt(wm_ctx, mom, up, np2))
  end subroutine mom_calc__nvt

end module mom_calc
