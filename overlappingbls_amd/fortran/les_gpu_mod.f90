!-----------------------------------------------------------------------
! les_gpu_mod: ISO_C_BINDING interface to the MI355X LES timestep library
! (include/les_gpu.h).  The reference's Fortran driver keeps its namelists
! and call sequence; the per-step routines reach the GPU through these
! interfaces.  Array arguments are the reference's own contiguous arrays,
! passed by address with their leading dimension (ldx = nxt + 2).
!-----------------------------------------------------------------------
module les_gpu_mod
use iso_c_binding
implicit none
private

integer(c_int), parameter, public :: LES_OK = 0, LES_E_BADARG = -1, LES_E_HIP = -2, &
                                     LES_E_CFL = -3, LES_E_NOMEM = -4, LES_E_NODEV = -5, &
                                     LES_E_IO = -6
integer(c_int), parameter, public :: LES_FP64 = 0, LES_MIXED = 1
integer(c_int), parameter, public :: LES_WALL_ROUGH = 0, LES_WALL_SMOOTH = 1
integer(c_int), parameter, public :: LES_BC_WALL = 0, LES_BC_STRESS_FREE = 1
integer(c_int), parameter, public :: LES_STOKES_DP = 0, LES_STOKES_MONO = 1, LES_STOKES_DYNAMIC = 2
integer(c_int), parameter, public :: LES_OBJ_USTAR2 = 0, LES_OBJ_T_FLUX = 1

! field identifiers (enum les_field)
integer(c_int), parameter, public :: LES_F_U = 0, LES_F_V = 1, LES_F_W = 2, LES_F_THETA = 3, &
    LES_F_RHSX = 4, LES_F_RHSY = 5, LES_F_RHSZ = 6, LES_F_RHST = 7, &
    LES_F_RHSX_F = 8, LES_F_RHSY_F = 9, LES_F_RHSZ_F = 10, LES_F_RHST_F = 11, &
    LES_F_DUDX = 12, LES_F_DUDY = 13, LES_F_DUDZ = 14, LES_F_DVDX = 15, LES_F_DVDY = 16, &
    LES_F_DVDZ = 17, LES_F_DWDX = 18, LES_F_DWDY = 19, LES_F_DWDZ = 20, &
    LES_F_TXX = 21, LES_F_TXY = 22, LES_F_TXZ = 23, LES_F_TYY = 24, LES_F_TYZ = 25, LES_F_TZZ = 26, &
    LES_F_DIVTX = 27, LES_F_DIVTY = 28, LES_F_DIVTZ = 29, &
    LES_F_P = 30, LES_F_DPDX = 31, LES_F_DPDY = 32, LES_F_DPDZ = 33, &
    LES_F_CS_OPT2 = 34, LES_F_F_LM = 35, LES_F_F_MM = 36, LES_F_F_QN = 37, LES_F_F_NN = 38, &
    LES_F_NU_T = 39, LES_F_U_LAG = 40, LES_F_V_LAG = 41, LES_F_W_LAG = 42, &
    LES_F_S11 = 43, LES_F_S12 = 44, LES_F_S13 = 45, LES_F_S22 = 46, LES_F_S23 = 47, LES_F_S33 = 48, &
    LES_F_CX = 49, LES_F_CY = 50, LES_F_CZ = 51, &
    LES_F_DTDX = 52, LES_F_DTDY = 53, LES_F_DTDZ = 54, LES_F_BETA_SCAL = 55, LES_F_SGS_T3 = 56, &
    LES_F_WORK0 = 57, LES_F_WORK1 = 58, LES_F_WORK2 = 59, &
    LES_F_USTAR_SFC = 100, LES_F_PSI_M_SFC = 101, LES_F_PHI_M_SFC = 102, LES_F_PSI_H_SFC = 103, &
    LES_F_PHI_H_SFC = 104, LES_F_USTAR_TOP_SFC = 105

! time-averaged slice identifiers (enum les_slice): flow_slice then theta_slice
integer(c_int), parameter, public :: LES_S_U = 0, LES_S_V = 1, LES_S_W = 2, LES_S_P = 3, &
    LES_S_DUDT = 4, LES_S_DVDT = 5, LES_S_DWDT = 6, LES_S_DUDZ = 7, LES_S_DVDZ = 8, &
    LES_S_U2 = 9, LES_S_V2 = 10, LES_S_W2 = 11, LES_S_P2 = 12, LES_S_UV = 13, LES_S_UW = 14, &
    LES_S_VW = 15, LES_S_WP = 16, LES_S_TXX = 17, LES_S_TXY = 18, LES_S_TXZ = 19, LES_S_TYY = 20, &
    LES_S_TYZ = 21, LES_S_TZZ = 22, LES_S_U3 = 23, LES_S_V3 = 24, LES_S_W3 = 25, &
    LES_S_U2_FLUC = 26, LES_S_V2_FLUC = 27, LES_S_UV_FLUC = 28, LES_S_U2W_FLUC = 29, &
    LES_S_V2W_FLUC = 30, LES_S_DISSIP = 31, LES_S_UTXZ_FLUC = 32, LES_S_VTYZ_FLUC = 33, &
    LES_S_WTZZ = 34, LES_S_CS = 35, LES_S_THETA = 36, LES_S_THETA2 = 37, LES_S_T2_FLUC = 38, &
    LES_S_DTDZ = 39, LES_S_SGS_T3 = 40, LES_S_WT = 41, LES_S_NU_T = 42, LES_N_SLICES = 43

type, bind(C), public :: les_params
    integer(c_int32_t) :: nxt, nyt, nzt, model, cs_count, dyn_init, inilag, restart_vel, sgs, molec, &
                          walltype, lbc_mom, ubc_mom, ubc, damping_method, coriolis_forcing, &
                          use_mean_p_force, ocean_flag, theta_flag, restart_theta, theta_init_time, &
                          passive_scalar, lbc, stokes_flag
    real(c_double) :: lx_tot, ly_tot, lz_tot, dt, z_i, u_scale, T_scale, nu_molec, zo, d0, Pr, wt_s, &
                      dTdz_top, coriol, ug, vg, mean_p_force, lz_npz
    integer(c_int32_t) :: stokes_action, flag_crossvel
    real(c_double) :: u_star, rad_stress, alpha_w, u_cross, v_cross, amp_w, lambda_w, agl_ust, &
                      inv_strength, T_s, Co
end type les_params

public :: les_create, les_destroy, les_set_field, les_get_field, les_step, les_run, &
          les_ddx, les_ddy, les_ddz_uv, les_ddz_w, les_filter_data, les_press_stag_array, &
          les_sgs_stag, les_convec, les_theta_all_in_one, les_project, les_set_precision, &
          les_fftn, les_ifftn, les_zero_padding, les_zero_unpadding, les_test_filter, &
          les_obukhov, les_wallstress_bottom, les_wallstress_top, les_divstress_uv, &
          les_divstress_w, les_set_stokes, les_get_cfl, les_set_avgslice, les_get_avgslice, &
          les_ab2, les_tdma_solve, les_bind_field, les_upload, les_download, &
          les_lin_set, les_lin_get, les_lin_zero, les_lin_set_objective, les_set_sgs_frozen, les_tlm_window, &
          les_adjoint_window, les_step_tlm, les_step_adj, &
          les_write_restart, les_read_restart, les_write_avgslice, &
          les_lagrange_sdep, les_interpolag_sdep, les_set_sgs_init_flags, les_get_sgs_init_flags, &
          les_set_sponge, les_set_wind_stress, les_get_profiles, les_sync, &
          les_nfamilies, les_profile_step, les_profile_run, &
          les_check

interface
    integer(c_int) function les_create(prm, ctx) bind(C, name="les_create")
        import :: c_int, c_ptr, les_params
        type(les_params), intent(in) :: prm
        type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function les_destroy(ctx) bind(C, name="les_destroy")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
    end function
    integer(c_int) function les_set_precision(ctx, p) bind(C, name="les_set_precision")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: p
    end function
    integer(c_int) function les_set_stokes(ctx, ust, vst) bind(C, name="les_set_stokes")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: ctx
        real(c_double), intent(in) :: ust(*), vst(*)
    end function
    integer(c_int) function les_set_sponge(ctx, sponge) bind(C, name="les_set_sponge")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: ctx
        real(c_double), intent(in) :: sponge(*)
    end function
    integer(c_int) function les_get_profiles(ctx, ust, vst, sponge) bind(C, name="les_get_profiles")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: ctx
        real(c_double), intent(out) :: ust(*), vst(*), sponge(*)
    end function
    integer(c_int) function les_set_wind_stress(ctx, ustar, agl_deg) bind(C, name="les_set_wind_stress")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: ctx
        real(c_double), value :: ustar, agl_deg
    end function
    integer(c_int) function les_set_sgs_init_flags(ctx, f_lm_mm_init, f_qn_nn_init) &
            bind(C, name="les_set_sgs_init_flags")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: f_lm_mm_init, f_qn_nn_init
    end function
    integer(c_int) function les_get_sgs_init_flags(ctx, f_lm_mm_init, f_qn_nn_init) &
            bind(C, name="les_get_sgs_init_flags")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), intent(out) :: f_lm_mm_init, f_qn_nn_init
    end function
    integer(c_int) function les_lagrange_sdep(ctx, jt) bind(C, name="les_lagrange_sdep")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: jt
    end function
    integer(c_int) function les_interpolag_sdep(ctx) bind(C, name="les_interpolag_sdep")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
    end function
    integer(c_int) function les_sync(ctx) bind(C, name="les_sync")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
    end function
    ! per-family device time (les_nfamilies() entries per array)
    integer(c_int) function les_nfamilies() bind(C, name="les_nfamilies")
        import :: c_int
    end function
    integer(c_int) function les_profile_step(ctx, jt, ms, bytes, launches) bind(C, name="les_profile_step")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: ctx
        integer(c_int), value :: jt
        real(c_double), intent(out) :: ms(*), bytes(*)
        integer(c_int), intent(out) :: launches(*)
    end function
    integer(c_int) function les_profile_run(ctx, jt0, nsteps, family, ms, bytes, launches) &
            bind(C, name="les_profile_run")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: ctx
        integer(c_int), value :: jt0, nsteps, family
        real(c_double), intent(out) :: ms(*), bytes(*)
        integer(c_int), intent(out) :: launches(*)
    end function
    integer(c_int) function les_set_field(ctx, f, h, ld, lev0, nlev) bind(C, name="les_set_field")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: ctx
        integer(c_int), value :: f, ld, lev0, nlev
        real(c_double), intent(in) :: h(*)
    end function
    integer(c_int) function les_get_field(ctx, f, h, ld, lev0, nlev) bind(C, name="les_get_field")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: ctx
        integer(c_int), value :: f, ld, lev0, nlev
        real(c_double), intent(out) :: h(*)
    end function
    integer(c_int) function les_ddx(ctx, d, f) bind(C, name="les_ddx")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: d, f
    end function
    integer(c_int) function les_ddy(ctx, d, f) bind(C, name="les_ddy")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: d, f
    end function
    integer(c_int) function les_ddz_uv(ctx, d, f) bind(C, name="les_ddz_uv")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: d, f
    end function
    integer(c_int) function les_ddz_w(ctx, d, f) bind(C, name="les_ddz_w")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: d, f
    end function
    integer(c_int) function les_filter_data(ctx, f) bind(C, name="les_filter_data")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: f
    end function
    integer(c_int) function les_test_filter(ctx, f, level, which) bind(C, name="les_test_filter")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: f, level, which
    end function
    integer(c_int) function les_fftn(ctx, fin, ld, fout) bind(C, name="les_fftn")
        import :: c_int, c_ptr, c_double, c_double_complex
        type(c_ptr), value :: ctx
        real(c_double), intent(in) :: fin(*)
        integer(c_int), value :: ld
        complex(c_double_complex), intent(out) :: fout(*)
    end function
    integer(c_int) function les_ifftn(ctx, fin, fout, ld) bind(C, name="les_ifftn")
        import :: c_int, c_ptr, c_double, c_double_complex
        type(c_ptr), value :: ctx
        complex(c_double_complex), intent(in) :: fin(*)
        real(c_double), intent(out) :: fout(*)
        integer(c_int), value :: ld
    end function
    integer(c_int) function les_zero_padding(ctx, var, ld, big) bind(C, name="les_zero_padding")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: ctx
        real(c_double), intent(in) :: var(*)
        integer(c_int), value :: ld
        real(c_double), intent(out) :: big(*)
    end function
    integer(c_int) function les_zero_unpadding(ctx, big, var, ld) bind(C, name="les_zero_unpadding")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: ctx
        real(c_double), intent(in) :: big(*)
        real(c_double), intent(out) :: var(*)
        integer(c_int), value :: ld
    end function
    integer(c_int) function les_obukhov(ctx, jt) bind(C, name="les_obukhov")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: jt
    end function
    integer(c_int) function les_wallstress_bottom(ctx) bind(C, name="les_wallstress_bottom")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
    end function
    integer(c_int) function les_wallstress_top(ctx) bind(C, name="les_wallstress_top")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
    end function
    integer(c_int) function les_sgs_stag(ctx, jt) bind(C, name="les_sgs_stag")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: jt
    end function
    integer(c_int) function les_divstress_uv(ctx, divt, tx, ty, tz) bind(C, name="les_divstress_uv")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: divt, tx, ty, tz
    end function
    integer(c_int) function les_divstress_w(ctx, divt, tx, ty, tz) bind(C, name="les_divstress_w")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: divt, tx, ty, tz
    end function
    integer(c_int) function les_convec(ctx) bind(C, name="les_convec")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
    end function
    integer(c_int) function les_theta_all_in_one(ctx, jt) bind(C, name="les_theta_all_in_one")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: jt
    end function
    integer(c_int) function les_press_stag_array(ctx, jt) bind(C, name="les_press_stag_array")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: jt
    end function
    integer(c_int) function les_project(ctx) bind(C, name="les_project")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
    end function
    integer(c_int) function les_step(ctx, jt) bind(C, name="les_step")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: jt
    end function
    integer(c_int) function les_run(ctx, jt0, nsteps, failed) bind(C, name="les_run")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: jt0, nsteps
        integer(c_int), intent(out) :: failed
    end function
    integer(c_int) function les_get_cfl(ctx, cfl) bind(C, name="les_get_cfl")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: ctx
        real(c_double), intent(out) :: cfl
    end function
    integer(c_int) function les_set_avgslice(ctx, use_avgslice, c_count, p_count) &
            bind(C, name="les_set_avgslice")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: use_avgslice, c_count, p_count
    end function
    integer(c_int) function les_get_avgslice(ctx, slice, window, host) bind(C, name="les_get_avgslice")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: ctx
        integer(c_int), value :: slice, window
        real(c_double), intent(out) :: host(*)
    end function
    integer(c_int) function les_ab2(ctx, jt) bind(C, name="les_ab2")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: jt
    end function
    integer(c_int) function les_tdma_solve(ctx, hx, hy, hz, ph) bind(C, name="les_tdma_solve")
        import :: c_ptr, c_int, c_double
        type(c_ptr), value :: ctx
        real(c_double), intent(in) :: hx(*), hy(*), hz(*)
        real(c_double), intent(out) :: ph(*)
    end function
    ! path: path_restart//c_null_char (checkpoint_final / read_flow_field)
    integer(c_int) function les_write_restart(ctx, path, nums) bind(C, name="les_write_restart")
        import :: c_int, c_ptr, c_char
        type(c_ptr), value :: ctx
        character(kind=c_char), intent(in) :: path(*)
        integer(c_int), value :: nums
    end function
    integer(c_int) function les_read_restart(ctx, path, nums, read_theta) bind(C, name="les_read_restart")
        import :: c_int, c_ptr, c_char
        type(c_ptr), value :: ctx
        character(kind=c_char), intent(in) :: path(*)
        integer(c_int), value :: nums, read_theta
    end function
    ! path: path//c_null_char (write_avgslice appends to path//'result/aver_*.out')
    integer(c_int) function les_write_avgslice(ctx, path, nums, average_dim_num) bind(C, name="les_write_avgslice")
        import :: c_int, c_ptr, c_char
        type(c_ptr), value :: ctx
        character(kind=c_char), intent(in) :: path(*)
        integer(c_int), value :: nums, average_dim_num
    end function
    integer(c_int) function les_bind_field(ctx, f, h, ld, lev0, nlev) bind(C, name="les_bind_field")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: ctx
        integer(c_int), value :: f, ld, lev0, nlev
        real(c_double), intent(inout), target :: h(*)
    end function
    integer(c_int) function les_upload(ctx) bind(C, name="les_upload")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
    end function
    integer(c_int) function les_download(ctx) bind(C, name="les_download")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
    end function
    ! ---- tangent-linear model and adjoint (no reference counterpart) ----
    integer(c_int) function les_lin_set(ctx, f, h, ld, lev0, nlev) bind(C, name="les_lin_set")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: ctx
        integer(c_int), value :: f, ld, lev0, nlev
        real(c_double), intent(in) :: h(*)
    end function
    integer(c_int) function les_lin_get(ctx, f, h, ld, lev0, nlev) bind(C, name="les_lin_get")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: ctx
        integer(c_int), value :: f, ld, lev0, nlev
        real(c_double), intent(out) :: h(*)
    end function
    integer(c_int) function les_lin_zero(ctx) bind(C, name="les_lin_zero")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
    end function
    integer(c_int) function les_lin_set_objective(ctx, kind) bind(C, name="les_lin_set_objective")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: kind
    end function
    integer(c_int) function les_set_sgs_frozen(ctx, frozen) bind(C, name="les_set_sgs_frozen")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: frozen
    end function
    integer(c_int) function les_tlm_window(ctx, jt0, nsteps, obj_steps, dJ) bind(C, name="les_tlm_window")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: ctx
        integer(c_int), value :: jt0, nsteps, obj_steps
        real(c_double), intent(out) :: dJ
    end function
    integer(c_int) function les_adjoint_window(ctx, jt0, nsteps, k_ckpt, obj_steps, J) &
            bind(C, name="les_adjoint_window")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: ctx
        integer(c_int), value :: jt0, nsteps, k_ckpt, obj_steps
        real(c_double), intent(out) :: J
    end function
    integer(c_int) function les_step_tlm(ctx, jt) bind(C, name="les_step_tlm")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: jt
    end function
    integer(c_int) function les_step_adj(ctx, jt) bind(C, name="les_step_adj")
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int), value :: jt
    end function
end interface

contains

    ! The reference prints and stops where the library returns a status.
    subroutine les_check(st, where)
        integer(c_int), intent(in) :: st
        character(*), intent(in) :: where
        if (st /= LES_OK) then
            if (st == LES_E_CFL) then
                write(*,*) "CFL IS TOO LARGE, PLEASE REDUCE THE TIME STEP (", where, ")"
            else
                write(*,*) "les_gpu error ", st, " in ", where
            end if
            stop
        end if
    end subroutine les_check

end module les_gpu_mod
