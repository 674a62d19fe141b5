! Minimal Fortran driver through les_gpu_mod: the call shape a maintainer's
! flosol replacement uses (create, upload, step, download, destroy).
program drive_step
use iso_c_binding
use les_gpu_mod
implicit none
type(les_params) :: prm
type(c_ptr) :: ctx
integer(c_int) :: st
integer, parameter :: nxt = 16, nyt = 16, nzt = 9, ldx = nxt + 2
real(c_double), allocatable :: u(:,:,:)
prm = les_params(nxt, nyt, nzt, 5, 2, 2, 1, 0, 1, 0, LES_WALL_ROUGH, LES_BC_WALL, LES_BC_STRESS_FREE, &
                 0, 1, 0, 0, 0, 1, 0, 1, 0, 1, 0, &
                 6.283185307179586d0, 6.283185307179586d0, 1d0, 2d-4, 100d0, 1d0, 300d0, 1.14d-5, &
                 1d-4, 0d0, 0.4d0, 0d0, 0.03d0, 0d0, 0d0, 0d0, 0d0, 1d0, &
                 LES_STOKES_MONO, 0, 0d0, 0d0, 2d-4, 0d0, 0d0, 0d0, 1d0, 0d0, 0d0, 0.96d0, 0.16d0)
st = les_create(prm, ctx)
call les_check(st, "les_create")
allocate(u(ldx, nyt, 0:nzt))
u = 1d0
call les_check(les_set_field(ctx, LES_F_U, u, ldx, 0, nzt + 1), "set u")
call les_check(les_step(ctx, 1), "les_step")
call les_check(les_get_field(ctx, LES_F_U, u, ldx, 0, nzt + 1), "get u")
write(*,*) "fortran driver ok", u(1, 1, 1)
st = les_destroy(ctx)
end program drive_step
