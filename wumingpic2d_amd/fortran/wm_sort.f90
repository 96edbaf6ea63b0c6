! Drop-in module `sort` (the public interface of common/sort.f90:16-45).
module sort
  use iso_c_binding
  use wmpic_capi
  implicit none
  private
  public :: sort__init, sort__bucket

  integer, save :: ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye

contains

  subroutine sort__init(ndim_in,np_in,nsp_in,nxgs_in,nxge_in,nygs_in,nyge_in,nys_in,nye_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in
    call wmpic_context()
    call wmpic_check(wm_sort_init(wm_ctx, ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, &
         nyge_in, nys_in, nye_in))
    ndim = ndim_in; np = np_in; nsp = nsp_in
    nxgs = nxgs_in; nxge = nxge_in; nygs = nygs_in; nyge = nyge_in; nys = nys_in; nye = nye_in
  end subroutine sort__init

  subroutine sort__bucket(gp,up,cumcnt,np2,nxs,nxe)
    integer, intent(in)    :: nxs, nxe
    integer, intent(in)    :: np2(nys:nye,nsp)
    integer, intent(out)   :: cumcnt(nxgs:nxge+1,nys:nye,nsp)
    real(8), intent(in)    :: up(ndim,np,nys:nye,nsp)
    real(8), intent(out)   :: gp(ndim,np,nys:nye,nsp)
    call wmpic_context()
    call wmpic_check(wm_sort_bucket(wm_ctx, gp, up, cumcnt, np2, nxs, nxe))
  end subroutine sort__bucket

end module sort
