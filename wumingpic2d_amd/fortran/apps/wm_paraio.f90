! Drop-in module `paraio` (the public interface of common/paraio.f90:9-16):
! the restart files and the run's outputs in the reference's JSON + raw
! format, written and read by libwmpic for one process.
module paraio
  use iso_c_binding
  use wmpic_capi
  implicit none
  private

  public :: paraio__init, paraio__finalize, paraio__output, paraio__input
  public :: paraio__param, paraio__mom, paraio__ptcl, paraio__orb

  integer, save :: ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye

contains

  subroutine paraio__init(ndim_in,np_in,nsp_in,nxgs_in,nxge_in,nygs_in,nyge_in,nys_in,nye_in, &
       & nproc_in,nrank_in,delx_in,delt_in,c_in,q_in,r_in,dir_in)
    integer, intent(in) :: ndim_in, np_in, nsp_in
    integer, intent(in) :: nxgs_in, nxge_in, nygs_in, nyge_in, nys_in, nye_in
    integer, intent(in) :: nproc_in, nrank_in
    real(8), intent(in) :: delx_in, delt_in, c_in, q_in(nsp_in), r_in(nsp_in)
    character(len=*), intent(in) :: dir_in
    call wmpic_context()
    call wmpic_check(wm_paraio_init(wm_ctx, ndim_in, np_in, nsp_in, nxgs_in, nxge_in, nygs_in, &
         nyge_in, nys_in, nye_in, nproc_in, nrank_in, delx_in, delt_in, c_in, q_in, r_in, &
         trim(dir_in) // c_null_char))
    ndim = ndim_in; np = np_in; nsp = nsp_in
    nxgs = nxgs_in; nxge = nxge_in; nygs = nygs_in; nyge = nyge_in; nys = nys_in; nye = nye_in
  end subroutine paraio__init

  subroutine paraio__finalize()
  end subroutine paraio__finalize

  subroutine paraio__output(up,uf,np2,nxs,nxe,it,filename)
    integer, intent(in) :: np2(nys:nye,nsp), nxs, nxe
    integer, intent(in) :: it
    real(8), intent(in) :: up(ndim,np,nys:nye,nsp)
    real(8), intent(in) :: uf(6,nxgs-2:nxge+2,nys-2:nye+2)
    character(len=*), intent(in) :: filename
    call wmpic_check(wm_paraio_output(wm_ctx, up, uf, np2, nxs, nxe, it, &
         trim(filename) // c_null_char))
  end subroutine paraio__output

  subroutine paraio__input(up,uf,np2,nxs,nxe,it,filename)
    real(8), intent(out) :: up(ndim,np,nys:nye,nsp)
    real(8), intent(out) :: uf(6,nxgs-2:nxge+2,nys-2:nye+2)
    integer, intent(out) :: np2(nys:nye,nsp)
    integer, intent(out) :: nxs, nxe, it
    character(len=*), intent(in) :: filename
    call wmpic_check(wm_paraio_input(wm_ctx, up, uf, np2, nxs, nxe, it, &
         trim(filename) // c_null_char))
  end subroutine paraio__input

  subroutine paraio__param(n0, wpe, wpi, wge, wgi, vti, vte, filename)
    integer, intent(in) :: n0
    real(8), intent(in) :: wpe, wpi, wge, wgi, vti, vte
    character(len=*), intent(in) :: filename
    call wmpic_check(wm_paraio_param(wm_ctx, n0, wpe, wpi, wge, wgi, vti, vte, &
         trim(filename) // c_null_char))
  end subroutine paraio__param

  subroutine paraio__mom(mom,uf,it)
    integer, intent(in) :: it
    real(8), intent(in) :: uf(6,nxgs-2:nxge+2,nys-2:nye+2)
    real(8), intent(inout) :: mom(7,nxgs-1:nxge+1,nys-1:nye+1,nsp)
    call wmpic_check(wm_paraio_mom(wm_ctx, mom, uf, it))
  end subroutine paraio__mom

  subroutine paraio__ptcl(up,uf,np2,it)
    integer, intent(in) :: it
    integer, intent(in) :: np2(nys:nye,nsp)
    real(8), intent(in) :: up(ndim,np,nys:nye,nsp)
    real(8), intent(in) :: uf(6,nxgs-2:nxge+2,nys-2:nye+2)
    call wmpic_check(wm_paraio_ptcl(wm_ctx, up, uf, np2, it))
  end subroutine paraio__ptcl

  subroutine paraio__orb(up,uf,np2,it)
    integer, intent(in) :: it
    integer, intent(in) :: np2(nys:nye,nsp)
    real(8), intent(in) :: up(ndim,np,nys:nye,nsp)
    real(8), intent(in) :: uf(6,nxgs-2:nxge+2,nys-2:nye+2)
    call wmpic_check(wm_paraio_orb(wm_ctx, up, uf, np2, it))
  end subroutine paraio__orb

end module paraio
