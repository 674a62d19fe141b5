!-----------------------------------------------------------------------
! Per-routine drop-ins with the reference's own interfaces.
!
! The reference calls these routines from flosol (103_flosol.f90:23-229)
! through implicit interfaces (external subroutines) or, for the
! derivatives, as procedures of module `derivatives`.  Linking these in
! place of the reference's objects routes each call to the GPU: a wrapper
! copies the module arrays the routine reads (sim_param, the reference's
! (ldx, nynpy, 0:nz) arrays) to the device, calls the C-ABI entry point and
! copies back what the routine writes, over the level ranges the reference
! routine writes.  State only the library touches (Cs_opt2, F_*, u_lag,
! the surface planes) stays on the device between calls.  The context is
! les_bridge's les_ctx, created once after init (see INTEGRATION.md);
! jt and the sizes come from module param.
!-----------------------------------------------------------------------
module les_derivatives
! replaces module derivatives (004_derivatives.f90) for ddx, ddy, ddz_uv, ddz_w
use iso_c_binding
use les_gpu_mod
use les_bridge, only: les_ctx
use param, only: ldx, nz
implicit none
private
public :: ddx, ddy, ddz_uv, ddz_w
contains
    ! d = op(f) on levels 0..nz through the device work fields
    subroutine deriv(which, d, f)
        integer, intent(in) :: which
        real(c_double), intent(out) :: d(ldx, *)
        real(c_double), intent(in) :: f(ldx, *)
        integer(c_int) :: st
        call les_check(les_set_field(les_ctx, LES_F_WORK0, f, ldx, 0, nz + 1), "deriv: f")
        select case (which)
        case (1)
            st = les_ddx(les_ctx, LES_F_WORK1, LES_F_WORK0)
        case (2)
            st = les_ddy(les_ctx, LES_F_WORK1, LES_F_WORK0)
        case (3)
            st = les_ddz_uv(les_ctx, LES_F_WORK1, LES_F_WORK0)
        case default
            st = les_ddz_w(les_ctx, LES_F_WORK1, LES_F_WORK0)
        end select
        call les_check(st, "deriv")
        call les_check(les_get_field(les_ctx, LES_F_WORK1, d, ldx, 0, nz + 1), "deriv: d")
    end subroutine deriv
    ! ddx, 004_derivatives.f90:18-41
    subroutine ddx(dfdx, f)
        real(c_double), intent(out) :: dfdx(ldx, *)
        real(c_double), intent(in) :: f(ldx, *)
        call deriv(1, dfdx, f)
    end subroutine ddx
    ! ddy, 004_derivatives.f90:45-81
    subroutine ddy(dfdy, f)
        real(c_double), intent(out) :: dfdy(ldx, *)
        real(c_double), intent(in) :: f(ldx, *)
        call deriv(2, dfdy, f)
    end subroutine ddy
    ! ddz_uv, 004_derivatives.f90:85-116 (levels the wall routine sets are
    ! left to it, as in the reference: the device holds them from before)
    subroutine ddz_uv(dfdz, f)
        real(c_double), intent(inout) :: dfdz(ldx, *)
        real(c_double), intent(in) :: f(ldx, *)
        call les_check(les_set_field(les_ctx, LES_F_WORK1, dfdz, ldx, 0, nz + 1), "ddz_uv: d")
        call deriv(3, dfdz, f)
    end subroutine ddz_uv
    ! ddz_w, 004_derivatives.f90:155-183
    subroutine ddz_w(dfdz, f)
        real(c_double), intent(inout) :: dfdz(ldx, *)
        real(c_double), intent(in) :: f(ldx, *)
        call les_check(les_set_field(les_ctx, LES_F_WORK1, dfdz, ldx, 0, nz + 1), "ddz_w: d")
        call deriv(4, dfdz, f)
    end subroutine ddz_w
end module les_derivatives

! wallstress_bottom, 104_wallstress.f90:98-190: txz, tyz, dudz, dvdz at level 1
subroutine wallstress_bottom()
use iso_c_binding
use les_gpu_mod
use les_bridge, only: les_ctx
use param, only: ldx, nz
use sim_param, only: u, v, dudz, dvdz, txz, tyz
implicit none
call les_check(les_set_field(les_ctx, LES_F_U, u, ldx, 0, nz + 1), "wallstress_bottom: u")
call les_check(les_set_field(les_ctx, LES_F_V, v, ldx, 0, nz + 1), "wallstress_bottom: v")
call les_check(les_set_field(les_ctx, LES_F_DUDZ, dudz, ldx, 0, nz + 1), "wallstress_bottom: dudz")
call les_check(les_set_field(les_ctx, LES_F_DVDZ, dvdz, ldx, 0, nz + 1), "wallstress_bottom: dvdz")
call les_check(les_wallstress_bottom(les_ctx), "wallstress_bottom")
call les_check(les_get_field(les_ctx, LES_F_TXZ, txz(1, 1, 1), ldx, 1, 1), "wallstress_bottom: txz")
call les_check(les_get_field(les_ctx, LES_F_TYZ, tyz(1, 1, 1), ldx, 1, 1), "wallstress_bottom: tyz")
call les_check(les_get_field(les_ctx, LES_F_DUDZ, dudz(1, 1, 1), ldx, 1, 1), "wallstress_bottom: dudz")
call les_check(les_get_field(les_ctx, LES_F_DVDZ, dvdz(1, 1, 1), ldx, 1, 1), "wallstress_bottom: dvdz")
end subroutine wallstress_bottom

! wallstress_top, 104_wallstress.f90:194-260: txz, tyz, dudz, dvdz at level nz
subroutine wallstress_top()
use iso_c_binding
use les_gpu_mod
use les_bridge, only: les_ctx
use param, only: ldx, nz
use sim_param, only: u, v, dudz, dvdz, txz, tyz
implicit none
call les_check(les_set_field(les_ctx, LES_F_U, u, ldx, 0, nz + 1), "wallstress_top: u")
call les_check(les_set_field(les_ctx, LES_F_V, v, ldx, 0, nz + 1), "wallstress_top: v")
call les_check(les_wallstress_top(les_ctx), "wallstress_top")
call les_check(les_get_field(les_ctx, LES_F_TXZ, txz(1, 1, nz), ldx, nz, 1), "wallstress_top: txz")
call les_check(les_get_field(les_ctx, LES_F_TYZ, tyz(1, 1, nz), ldx, nz, 1), "wallstress_top: tyz")
call les_check(les_get_field(les_ctx, LES_F_DUDZ, dudz(1, 1, nz), ldx, nz, 1), "wallstress_top: dudz")
call les_check(les_get_field(les_ctx, LES_F_DVDZ, dvdz(1, 1, nz), ldx, nz, 1), "wallstress_top: dvdz")
end subroutine wallstress_top

! sgs_stag, 106_sgs_stag.f90:5-318: txx, txy, tyy, tzz at 1..nz-1, txz, tyz
! at 1..nz (with lagrange_Sdep on update steps; Cs_opt2, F_* and u_lag stay
! on the device)
subroutine sgs_stag()
use iso_c_binding
use les_gpu_mod
use les_bridge, only: les_ctx
use param, only: ldx, nz, jt
use sim_param, only: u, v, w, dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz, &
                     txx, txy, txz, tyy, tyz, tzz
implicit none
call les_check(les_set_field(les_ctx, LES_F_U, u, ldx, 0, nz + 1), "sgs_stag: u")
call les_check(les_set_field(les_ctx, LES_F_V, v, ldx, 0, nz + 1), "sgs_stag: v")
call les_check(les_set_field(les_ctx, LES_F_W, w, ldx, 0, nz + 1), "sgs_stag: w")
call les_check(les_set_field(les_ctx, LES_F_DUDX, dudx, ldx, 0, nz + 1), "sgs_stag: dudx")
call les_check(les_set_field(les_ctx, LES_F_DUDY, dudy, ldx, 0, nz + 1), "sgs_stag: dudy")
call les_check(les_set_field(les_ctx, LES_F_DUDZ, dudz, ldx, 0, nz + 1), "sgs_stag: dudz")
call les_check(les_set_field(les_ctx, LES_F_DVDX, dvdx, ldx, 0, nz + 1), "sgs_stag: dvdx")
call les_check(les_set_field(les_ctx, LES_F_DVDY, dvdy, ldx, 0, nz + 1), "sgs_stag: dvdy")
call les_check(les_set_field(les_ctx, LES_F_DVDZ, dvdz, ldx, 0, nz + 1), "sgs_stag: dvdz")
call les_check(les_set_field(les_ctx, LES_F_DWDX, dwdx, ldx, 0, nz + 1), "sgs_stag: dwdx")
call les_check(les_set_field(les_ctx, LES_F_DWDY, dwdy, ldx, 0, nz + 1), "sgs_stag: dwdy")
call les_check(les_set_field(les_ctx, LES_F_DWDZ, dwdz, ldx, 0, nz + 1), "sgs_stag: dwdz")
call les_check(les_set_field(les_ctx, LES_F_TXZ, txz, ldx, 0, nz + 1), "sgs_stag: txz")
call les_check(les_set_field(les_ctx, LES_F_TYZ, tyz, ldx, 0, nz + 1), "sgs_stag: tyz")
call les_check(les_sgs_stag(les_ctx, jt), "sgs_stag")
call les_check(les_get_field(les_ctx, LES_F_TXX, txx(1, 1, 1), ldx, 1, nz - 1), "sgs_stag: txx")
call les_check(les_get_field(les_ctx, LES_F_TXY, txy(1, 1, 1), ldx, 1, nz - 1), "sgs_stag: txy")
call les_check(les_get_field(les_ctx, LES_F_TYY, tyy(1, 1, 1), ldx, 1, nz - 1), "sgs_stag: tyy")
call les_check(les_get_field(les_ctx, LES_F_TZZ, tzz(1, 1, 1), ldx, 1, nz - 1), "sgs_stag: tzz")
call les_check(les_get_field(les_ctx, LES_F_TXZ, txz(1, 1, 1), ldx, 1, nz), "sgs_stag: txz")
call les_check(les_get_field(les_ctx, LES_F_TYZ, tyz(1, 1, 1), ldx, 1, nz), "sgs_stag: tyz")
end subroutine sgs_stag

! divstress_uv / divstress_w (107_divstress.f90:4-104) through the device
! work fields: divt at 1..nz-1 (divstress_w: 1..nz)
subroutine les_divstress_call(wver, divt, tx, ty, tz)
use iso_c_binding
use les_gpu_mod
use les_bridge, only: les_ctx
use param, only: ldx, nynpy, nz
implicit none
integer, intent(in) :: wver
real(c_double), intent(inout) :: divt(ldx, nynpy, 0:nz)
real(c_double), intent(in) :: tx(ldx, nynpy, 0:nz), ty(ldx, nynpy, 0:nz), tz(ldx, nynpy, 0:nz)
integer(c_int) :: st
call les_check(les_set_field(les_ctx, LES_F_WORK0, tx, ldx, 0, nz + 1), "divstress: tx")
call les_check(les_set_field(les_ctx, LES_F_WORK1, ty, ldx, 0, nz + 1), "divstress: ty")
call les_check(les_set_field(les_ctx, LES_F_WORK2, tz, ldx, 0, nz + 1), "divstress: tz")
if (wver == 1) then
    st = les_divstress_w(les_ctx, LES_F_DIVTZ, LES_F_WORK0, LES_F_WORK1, LES_F_WORK2)
    call les_check(st, "divstress_w")
    call les_check(les_get_field(les_ctx, LES_F_DIVTZ, divt(1, 1, 1), ldx, 1, nz), "divstress_w: divt")
else
    st = les_divstress_uv(les_ctx, LES_F_DIVTX, LES_F_WORK0, LES_F_WORK1, LES_F_WORK2)
    call les_check(st, "divstress_uv")
    call les_check(les_get_field(les_ctx, LES_F_DIVTX, divt(1, 1, 1), ldx, 1, nz - 1), "divstress_uv: divt")
end if
end subroutine les_divstress_call

subroutine divstress_uv(divt, tx, ty, tz)
use iso_c_binding
use param, only: ldx, nynpy, nz
implicit none
real(c_double) :: divt(ldx, nynpy, 0:nz), tx(ldx, nynpy, 0:nz), ty(ldx, nynpy, 0:nz), tz(ldx, nynpy, 0:nz)
call les_divstress_call(0, divt, tx, ty, tz)
end subroutine divstress_uv

subroutine divstress_w(divt, tx, ty, tz)
use iso_c_binding
use param, only: ldx, nynpy, nz
implicit none
real(c_double) :: divt(ldx, nynpy, 0:nz), tx(ldx, nynpy, 0:nz), ty(ldx, nynpy, 0:nz), tz(ldx, nynpy, 0:nz)
call les_divstress_call(1, divt, tx, ty, tz)
end subroutine divstress_w

! convec, 108_convec.f90:11-202: cx, cy, cz at 1..nz-1
subroutine convec(cx, cy, cz)
use iso_c_binding
use les_gpu_mod
use les_bridge, only: les_ctx
use param, only: ldx, nynpy, nz
use sim_param, only: u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy
implicit none
real(c_double) :: cx(ldx, nynpy, 0:nz), cy(ldx, nynpy, 0:nz), cz(ldx, nynpy, 0:nz)
call les_check(les_set_field(les_ctx, LES_F_U, u, ldx, 0, nz + 1), "convec: u")
call les_check(les_set_field(les_ctx, LES_F_V, v, ldx, 0, nz + 1), "convec: v")
call les_check(les_set_field(les_ctx, LES_F_W, w, ldx, 0, nz + 1), "convec: w")
call les_check(les_set_field(les_ctx, LES_F_DUDY, dudy, ldx, 0, nz + 1), "convec: dudy")
call les_check(les_set_field(les_ctx, LES_F_DUDZ, dudz, ldx, 0, nz + 1), "convec: dudz")
call les_check(les_set_field(les_ctx, LES_F_DVDX, dvdx, ldx, 0, nz + 1), "convec: dvdx")
call les_check(les_set_field(les_ctx, LES_F_DVDZ, dvdz, ldx, 0, nz + 1), "convec: dvdz")
call les_check(les_set_field(les_ctx, LES_F_DWDX, dwdx, ldx, 0, nz + 1), "convec: dwdx")
call les_check(les_set_field(les_ctx, LES_F_DWDY, dwdy, ldx, 0, nz + 1), "convec: dwdy")
call les_check(les_convec(les_ctx), "convec")
call les_check(les_get_field(les_ctx, LES_F_CX, cx(1, 1, 1), ldx, 1, nz - 1), "convec: cx")
call les_check(les_get_field(les_ctx, LES_F_CY, cy(1, 1, 1), ldx, 1, nz - 1), "convec: cy")
call les_check(les_get_field(les_ctx, LES_F_CZ, cz(1, 1, 1), ldx, 1, nz - 1), "convec: cz")
end subroutine convec

! Admas_Bashforth, 103_flosol.f90:319-389: u, v, w at 1..nz-1 (0 <- BOGUS)
subroutine Admas_Bashforth()
use iso_c_binding
use les_gpu_mod
use les_bridge, only: les_ctx
use param, only: ldx, nz, jt
use sim_param, only: u, v, w, RHSx, RHSy, RHSz, RHSx_f, RHSy_f, RHSz_f
implicit none
call les_check(les_set_field(les_ctx, LES_F_U, u, ldx, 0, nz + 1), "ab2: u")
call les_check(les_set_field(les_ctx, LES_F_V, v, ldx, 0, nz + 1), "ab2: v")
call les_check(les_set_field(les_ctx, LES_F_W, w, ldx, 0, nz + 1), "ab2: w")
call les_check(les_set_field(les_ctx, LES_F_RHSX, RHSx, ldx, 0, nz + 1), "ab2: RHSx")
call les_check(les_set_field(les_ctx, LES_F_RHSY, RHSy, ldx, 0, nz + 1), "ab2: RHSy")
call les_check(les_set_field(les_ctx, LES_F_RHSZ, RHSz, ldx, 0, nz + 1), "ab2: RHSz")
call les_check(les_set_field(les_ctx, LES_F_RHSX_F, RHSx_f, ldx, 0, nz + 1), "ab2: RHSx_f")
call les_check(les_set_field(les_ctx, LES_F_RHSY_F, RHSy_f, ldx, 0, nz + 1), "ab2: RHSy_f")
call les_check(les_set_field(les_ctx, LES_F_RHSZ_F, RHSz_f, ldx, 0, nz + 1), "ab2: RHSz_f")
call les_check(les_ab2(les_ctx, jt), "Admas_Bashforth")
call les_check(les_get_field(les_ctx, LES_F_U, u, ldx, 0, nz), "ab2: u")
call les_check(les_get_field(les_ctx, LES_F_V, v, ldx, 0, nz), "ab2: v")
call les_check(les_get_field(les_ctx, LES_F_W, w, ldx, 0, nz), "ab2: w")
call les_check(les_get_field(les_ctx, LES_F_RHSX_F, RHSx_f, ldx, 0, nz + 1), "ab2: RHSx_f")
call les_check(les_get_field(les_ctx, LES_F_RHSY_F, RHSy_f, ldx, 0, nz + 1), "ab2: RHSy_f")
call les_check(les_get_field(les_ctx, LES_F_RHSZ_F, RHSz_f, ldx, 0, nz + 1), "ab2: RHSz_f")
end subroutine Admas_Bashforth

! press_stag_array, 109_press_stag_array.f90:107-240: p at 0..nz-1, dpdx,
! dpdy at 1..nz-1
subroutine press_stag_array(func, dfdx, dfdy)
use iso_c_binding
use les_gpu_mod
use les_bridge, only: les_ctx
use param, only: ldx, nynpy, nz, jt
use sim_param, only: u, v, w, divtz
implicit none
real(c_double) :: func(ldx, nynpy, 0:nz), dfdx(ldx, nynpy, nz), dfdy(ldx, nynpy, nz)
call les_check(les_set_field(les_ctx, LES_F_U, u, ldx, 0, nz + 1), "press: u")
call les_check(les_set_field(les_ctx, LES_F_V, v, ldx, 0, nz + 1), "press: v")
call les_check(les_set_field(les_ctx, LES_F_W, w, ldx, 0, nz + 1), "press: w")
call les_check(les_set_field(les_ctx, LES_F_DIVTZ, divtz, ldx, 0, nz + 1), "press: divtz")
call les_check(les_press_stag_array(les_ctx, jt), "press_stag_array")
call les_check(les_get_field(les_ctx, LES_F_P, func, ldx, 0, nz), "press: p")
call les_check(les_get_field(les_ctx, LES_F_DPDX, dfdx, ldx, 1, nz - 1), "press: dpdx")
call les_check(les_get_field(les_ctx, LES_F_DPDY, dfdy, ldx, 1, nz - 1), "press: dpdy")
end subroutine press_stag_array

! project, 110_project.f90:4-62, with data_exchange's ghost rows
! (202_parallel.f90:722-733): u, v, w at 0..nz
subroutine project()
use iso_c_binding
use les_gpu_mod
use les_bridge, only: les_ctx
use param, only: ldx, nz
use sim_param, only: u, v, w, dpdx, dpdy, dpdz
implicit none
call les_check(les_set_field(les_ctx, LES_F_U, u, ldx, 0, nz + 1), "project: u")
call les_check(les_set_field(les_ctx, LES_F_V, v, ldx, 0, nz + 1), "project: v")
call les_check(les_set_field(les_ctx, LES_F_W, w, ldx, 0, nz + 1), "project: w")
call les_check(les_set_field(les_ctx, LES_F_DPDX, dpdx(1, 1, 1), ldx, 1, nz - 1), "project: dpdx")
call les_check(les_set_field(les_ctx, LES_F_DPDY, dpdy(1, 1, 1), ldx, 1, nz - 1), "project: dpdy")
call les_check(les_set_field(les_ctx, LES_F_DPDZ, dpdz(1, 1, 1), ldx, 1, nz - 1), "project: dpdz")
call les_check(les_project(les_ctx), "project")
call les_check(les_get_field(les_ctx, LES_F_U, u, ldx, 0, nz + 1), "project: u")
call les_check(les_get_field(les_ctx, LES_F_V, v, ldx, 0, nz + 1), "project: v")
call les_check(les_get_field(les_ctx, LES_F_W, w, ldx, 0, nz + 1), "project: w")
end subroutine project

! filter_data, 006_test_filter.f90:82-105, in place on levels 0..nz
subroutine filter_data(f)
use iso_c_binding
use les_gpu_mod
use les_bridge, only: les_ctx
use param, only: ldx, nynpy, nz
implicit none
real(c_double) :: f(ldx, nynpy, 0:nz)
call les_check(les_set_field(les_ctx, LES_F_WORK0, f, ldx, 0, nz + 1), "filter_data: f")
call les_check(les_filter_data(les_ctx, LES_F_WORK0), "filter_data")
call les_check(les_get_field(les_ctx, LES_F_WORK0, f, ldx, 0, nz + 1), "filter_data: f")
end subroutine filter_data
