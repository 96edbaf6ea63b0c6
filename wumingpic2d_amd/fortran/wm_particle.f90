! Drop-in module `particle` (the public interface of common/particle.f90:18-54)
! running the push on the device through libwmpic.
module particle
  use iso_c_binding
  use wmpic_capi
  implicit none
  private
  public :: particle__init, particle__solv

  integer, save :: ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye

contains

  subroutine particle__init(ndim_in,np_in,nsp_in,nxgs_in,nxge_in,nygs_in,nyge_in,nys_in,nye_in, &
                            delx_in,delt_in,c_in,q_in,r_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in
    real(8), intent(in) :: delx_in, delt_in, c_in, q_in(nsp_in), r_in(nsp_in)
    call wmpic_context()
    call wmpic_check(wm_particle_init(wm_ctx, ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, &
         nyge_in, nys_in, nye_in, delx_in, delt_in, c_in, q_in, r_in))
    ndim = ndim_in; np = np_in; nsp = nsp_in
    nxgs = nxgs_in; nxge = nxge_in; nygs = nygs_in; nyge = nyge_in; nys = nys_in; nye = nye_in
  end subroutine particle__init

  subroutine particle__solv(gp,up,uf,cumcnt,nxs,nxe)
    integer, intent(in)  :: nxs, nxe
    integer, intent(in)  :: cumcnt(nxgs:nxge+1,nys:nye,nsp)
    real(8), intent(in)  :: up(ndim,np,nys:nye,nsp)
    real(8), intent(in)  :: uf(6,nxgs-2:nxge+2,nys-2:nye+2)
    real(8), intent(out) :: gp(ndim,np,nys:nye,nsp)
    call wmpic_context()
    call wmpic_check(wm_particle_solv(wm_ctx, gp, up, uf, cumcnt, nxs, nxe))
  end subroutine particle__solv

end module particle
