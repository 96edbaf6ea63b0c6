! Test driver: the time loop of proj/weibel/app.f90:100-107 written against
! the drop-in modules particle, field, sort, boundary_periodic, mom_calc and
! mpi_set (the module names, procedure names and argument lists of the
! reference), so the same source would compile against the reference's own
! modules.  After the steps it computes the moments of an output step
! (proj/weibel/app.f90:119-121).
!
! Usage: weibel_loop.x <dir>   reads <dir>/params.txt and <dir>/state.bin,
! runs the steps, writes <dir>/out.bin (the layout of state.bin, then mom).
program weibel_loop
  use mpi_set
  use particle
  use field
  use sort
  use mom_calc
  use boundary_periodic, &
       & bc__init        => boundary_periodic__init,       &
       & bc__dfield      => boundary_periodic__dfield,     &
       & bc__particle_x  => boundary_periodic__particle_x, &
       & bc__particle_y  => boundary_periodic__particle_y, &
       & bc__curre       => boundary_periodic__curre,      &
       & bc__phi         => boundary_periodic__phi,        &
       & bc__mom         => boundary_periodic__mom
  implicit none
  integer :: ndim, np, nsp, nxgs, nxge, nygs, nyge, nsteps, it, u
  integer :: nxs, nxe
  real(8) :: delx, delt, c, gfac
  real(8), allocatable :: q(:), r(:)
  integer, allocatable :: np2(:,:), cumcnt(:,:,:)
  real(8), allocatable :: uf(:,:,:), up(:,:,:,:), gp(:,:,:,:), mom(:,:,:,:)
  character(len=512) :: dir

  call get_command_argument(1, dir)
  open(newunit=u, file=trim(dir)//'/params.txt', status='old')
  read(u,*) ndim, np, nsp, nxgs, nxge, nygs, nyge
  read(u,*) delx, delt, c, gfac
  allocate(q(nsp), r(nsp))
  read(u,*) q
  read(u,*) r
  read(u,*) nsteps
  close(u)

  call mpi_set__init(nygs, nyge, 1)
  nxs = nxgs
  nxe = nxge
  allocate(np2(nys:nye,nsp))
  allocate(cumcnt(nxgs:nxge+1,nys:nye,nsp))
  allocate(uf(6,nxgs-2:nxge+2,nys-2:nye+2))
  allocate(up(ndim,np,nys:nye,nsp))
  allocate(gp(ndim,np,nys:nye,nsp))
  allocate(mom(7,nxgs-1:nxge+1,nys-1:nye+1,nsp))

  open(newunit=u, file=trim(dir)//'/state.bin', access='stream', form='unformatted', status='old')
  read(u) up, np2, cumcnt, uf
  close(u)

  call bc__init(ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, &
       & nup, ndown, mnpi, mnpr, ncomw, nerr, nstat, delx, delt, c)
  call particle__init(ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, delx, delt, c, q, r)
  call field__init(ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, &
       & mnpr, ncomw, opsum, nerr, delx, delt, c, q, r, gfac)
  call sort__init(ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye)
  call mom_calc__init(ndim, np, nsp, nxgs, nxge, nygs, nyge, nys, nye, delx, delt, c, q, r)
  gp = up

  do it = 1, nsteps
     call particle__solv(gp, up, uf, cumcnt, nxs, nxe)
     call field__fdtd_i(uf, up, gp, cumcnt, nxs, nxe, &
          & bc__dfield, bc__curre, bc__phi)
     call bc__particle_x(gp, np2)
     call bc__particle_y(gp, np2)
     call sort__bucket(up, gp, cumcnt, np2, nxs, nxe)
  enddo

  call mom_calc__accl(gp, up, uf, cumcnt, nxs, nxe)
  call mom_calc__nvt(mom, gp, np2)
  call bc__mom(mom)

  open(newunit=u, file=trim(dir)//'/out.bin', access='stream', form='unformatted', status='replace')
  write(u) up, np2, cumcnt, uf, mom
  close(u)
end program weibel_loop
