! Stand-ins for the reference's modules param, sim_param (003_sim_param.f90)
! and a les_bridge holding the context: the names, kinds and array bounds the
! drop-ins in les_dropin.f90 use (102_initial.f90:498-513 allocates the
! sim_param arrays with these bounds).
module param
use iso_c_binding
implicit none
integer :: nxt, nynpy, nz, ldx, jt
real(c_double) :: dz
end module param

module sim_param
use iso_c_binding
implicit none
real(c_double), allocatable, target :: u(:, :, :), v(:, :, :), w(:, :, :)
real(c_double), allocatable, target :: dudx(:, :, :), dudy(:, :, :), dudz(:, :, :), &
    dvdx(:, :, :), dvdy(:, :, :), dvdz(:, :, :), dwdx(:, :, :), dwdy(:, :, :), dwdz(:, :, :)
real(c_double), allocatable, target :: RHSx(:, :, :), RHSy(:, :, :), RHSz(:, :, :), &
    RHSx_f(:, :, :), RHSy_f(:, :, :), RHSz_f(:, :, :)
real(c_double), allocatable, target :: txx(:, :, :), txy(:, :, :), txz(:, :, :), tyy(:, :, :), &
    tyz(:, :, :), tzz(:, :, :), divtx(:, :, :), divty(:, :, :), divtz(:, :, :), p(:, :, :)
real(c_double), allocatable, target :: dpdx(:, :, :), dpdy(:, :, :), dpdz(:, :, :)
contains
    subroutine sim_param_alloc(ldx, ny, nz)
        integer, intent(in) :: ldx, ny, nz
        allocate (u(ldx, ny, 0:nz), v(ldx, ny, 0:nz), w(ldx, ny, 0:nz))
        allocate (dudx(ldx, ny, 0:nz), dudy(ldx, ny, 0:nz), dudz(ldx, ny, 0:nz), dvdx(ldx, ny, 0:nz), &
                  dvdy(ldx, ny, 0:nz), dvdz(ldx, ny, 0:nz), dwdx(ldx, ny, 0:nz), dwdy(ldx, ny, 0:nz), &
                  dwdz(ldx, ny, 0:nz))
        allocate (RHSx(ldx, ny, 0:nz), RHSy(ldx, ny, 0:nz), RHSz(ldx, ny, 0:nz), RHSx_f(ldx, ny, 0:nz), &
                  RHSy_f(ldx, ny, 0:nz), RHSz_f(ldx, ny, 0:nz))
        allocate (txx(ldx, ny, 0:nz), txy(ldx, ny, 0:nz), txz(ldx, ny, 0:nz), tyy(ldx, ny, 0:nz), &
                  tyz(ldx, ny, 0:nz), tzz(ldx, ny, 0:nz), divtx(ldx, ny, 0:nz), divty(ldx, ny, 0:nz), &
                  divtz(ldx, ny, 0:nz), p(ldx, ny, 0:nz))
        allocate (dpdx(ldx, ny, nz), dpdy(ldx, ny, nz), dpdz(ldx, ny, nz))
        u = 0; v = 0; w = 0; dudx = 0; dudy = 0; dudz = 0; dvdx = 0; dvdy = 0; dvdz = 0
        dwdx = 0; dwdy = 0; dwdz = 0; RHSx = 0; RHSy = 0; RHSz = 0; RHSx_f = 0; RHSy_f = 0; RHSz_f = 0
        txx = 0; txy = 0; txz = 0; tyy = 0; tyz = 0; tzz = 0; divtx = 0; divty = 0; divtz = 0; p = 0
        dpdx = 0; dpdy = 0; dpdz = 0
    end subroutine sim_param_alloc
end module sim_param

module les_bridge
use iso_c_binding
implicit none
type(c_ptr), save :: les_ctx
end module les_bridge
