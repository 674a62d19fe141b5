! flosol (103_flosol.f90:23-229, theta off) routine by routine through the
! drop-ins of les_dropin.f90 with the reference's interfaces and arrays
! (ldx = nxt + 2 columns, the reference's bounds): the caller's Fortran
! does what flosol itself does (newstep's RHS_f = RHS, the RHS assembly,
! dpdz, RHS -= grad p), each routine is a drop-in call.  Inputs and outputs
! are raw float64 files in C order (levels, nyt, nxt).
program drive_routines
use iso_c_binding
use les_gpu_mod
use les_derivatives
use les_bridge
use param
use sim_param
implicit none
type(les_params) :: prm
integer :: nsteps, step, i, u_in, u_out, nzt
real(c_double), allocatable :: tmp(:, :, :)
character(len=512) :: dir
interface
    subroutine wallstress_bottom()
    end subroutine
    subroutine wallstress_top()
    end subroutine
    subroutine sgs_stag()
    end subroutine
    subroutine divstress_uv(divt, tx, ty, tz)
    import :: c_double
    real(c_double) :: divt(*), tx(*), ty(*), tz(*)
    end subroutine
    subroutine divstress_w(divt, tx, ty, tz)
    import :: c_double
    real(c_double) :: divt(*), tx(*), ty(*), tz(*)
    end subroutine
    subroutine convec(cx, cy, cz)
    import :: c_double
    real(c_double) :: cx(*), cy(*), cz(*)
    end subroutine
    subroutine Admas_Bashforth()
    end subroutine
    subroutine press_stag_array(func, dfdx, dfdy)
    import :: c_double
    real(c_double) :: func(*), dfdx(*), dfdy(*)
    end subroutine
    subroutine project()
    end subroutine
    subroutine filter_data(f)
    import :: c_double
    real(c_double) :: f(*)
    end subroutine
end interface

call get_command_argument(1, dir)
open(newunit=u_in, file=trim(dir)//"/params.bin", access="stream", form="unformatted", status="old")
read(u_in) prm
read(u_in) nsteps
close(u_in)
nxt = prm%nxt; nynpy = prm%nyt; nzt = prm%nzt; nz = nzt; ldx = nxt + 2
dz = prm%lz_tot / (nzt - 1)
call les_check(les_create(prm, les_ctx), "les_create")
call sim_param_alloc(ldx, nynpy, nz)
allocate(tmp(nxt, nynpy, 0:nz))
open(newunit=u_in, file=trim(dir)//"/fields.bin", access="stream", form="unformatted", status="old")
read(u_in) tmp; u(1:nxt, :, :) = tmp
read(u_in) tmp; v(1:nxt, :, :) = tmp
read(u_in) tmp; w(1:nxt, :, :) = tmp
close(u_in)
do i = 1, 3   ! init's filter_data(u, v, w), 102_initial.f90:55-57
    if (i == 1) call filter_data(u)
    if (i == 2) call filter_data(v)
    if (i == 3) call filter_data(w)
end do
do step = 1, nsteps
    jt = step
    RHSx_f = RHSx; RHSy_f = RHSy; RHSz_f = RHSz                  ! newstep, 103_flosol.f90:311-313
    call les_check(les_set_field(les_ctx, LES_F_U, u, ldx, 0, nz + 1), "u")
    call les_check(les_set_field(les_ctx, LES_F_V, v, ldx, 0, nz + 1), "v")
    call les_check(les_obukhov(les_ctx, jt), "obukhov")        ! scalars_module procedure (:26)
    call ddx(dudx, u); call ddy(dudy, u)                          ! :30-38
    call ddx(dvdx, v); call ddy(dvdy, v)
    call ddx(dwdx, w); call ddy(dwdy, w)
    call ddz_uv(dudz, u); call ddz_uv(dvdz, v); call ddz_w(dwdz, w)
    call wallstress_bottom(); call wallstress_top()               ! :48-52
    call sgs_stag()                                               ! :59
    call divstress_uv(divtx, txx, txy, txz)                      ! :82-84
    call divstress_uv(divty, txy, tyy, tyz)
    call divstress_w(divtz, txz, tyz, tzz)
    call convec(RHSx, RHSy, RHSz)                                 ! :87
    RHSx(:, :, 1:nz - 1) = -RHSx(:, :, 1:nz - 1) - divtx(:, :, 1:nz - 1)   ! :115-117
    RHSy(:, :, 1:nz - 1) = -RHSy(:, :, 1:nz - 1) - divty(:, :, 1:nz - 1)
    RHSz(:, :, 1:nz - 1) = -RHSz(:, :, 1:nz - 1) - divtz(:, :, 1:nz - 1)
    call Admas_Bashforth()                                        ! :162
    call press_stag_array(p, dpdx, dpdy)                          ! :168
    dpdz(1:nxt, :, 1:nz - 1) = (p(1:nxt, :, 1:nz - 1) - p(1:nxt, :, 0:nz - 2)) / dz   ! :174-175
    RHSx(:, :, 1:nz - 1) = RHSx(:, :, 1:nz - 1) - dpdx(:, :, 1:nz - 1)    ! :179-181
    RHSy(:, :, 1:nz - 1) = RHSy(:, :, 1:nz - 1) - dpdy(:, :, 1:nz - 1)
    RHSz(:, :, 1:nz - 1) = RHSz(:, :, 1:nz - 1) - dpdz(:, :, 1:nz - 1)
    call project()                                                ! :188
    call filter_data(u); call filter_data(v); call filter_data(w)          ! :227-229
end do
open(newunit=u_out, file=trim(dir)//"/out.bin", access="stream", form="unformatted", status="replace")
write(u_out) u(1:nxt, :, :)
write(u_out) v(1:nxt, :, :)
write(u_out) w(1:nxt, :, :)
write(u_out) RHSx(1:nxt, :, :)
write(u_out) p(1:nxt, :, :)
close(u_out)
call les_check(les_destroy(les_ctx), "les_destroy")
write(*,*) "fortran routines ok"
end program drive_routines
