! A maintainer's flosol replacement in miniature, through les_gpu_mod: the
! module arrays (ldx = nxt+2 columns, levels 0:nzt) are registered once,
! uploaded, advanced by les_run, downloaded; then a frozen-SGS adjoint window
! gives dJ/dtheta and dJ/du at its first step.  Inputs and outputs are raw
! float64 files in C order (levels, nyt, nxt) written / read by the test.
program drive_window
use iso_c_binding
use les_gpu_mod
implicit none
type(les_params) :: prm
type(c_ptr) :: ctx
integer, parameter :: nf = 17
integer :: nxt, nyt, nzt, ldx, nsteps, i, u_in, u_out
real(c_double) :: J
character(len=512) :: dir
integer(c_int), parameter :: ids(nf) = [LES_F_U, LES_F_V, LES_F_W, LES_F_THETA, LES_F_RHSX, LES_F_RHSY, &
    LES_F_RHSZ, LES_F_RHST, LES_F_CS_OPT2, LES_F_F_LM, LES_F_F_MM, LES_F_F_QN, LES_F_F_NN, LES_F_U_LAG, &
    LES_F_V_LAG, LES_F_W_LAG, LES_F_SGS_T3]
real(c_double), allocatable, target :: fld(:,:,:,:), tmp(:,:,:), g(:,:,:)

call get_command_argument(1, dir)
open(newunit=u_in, file=trim(dir)//"/params.bin", access="stream", form="unformatted", status="old")
read(u_in) prm
read(u_in) nsteps
close(u_in)
nxt = prm%nxt; nyt = prm%nyt; nzt = prm%nzt; ldx = nxt + 2
call les_check(les_create(prm, ctx), "les_create")
allocate(fld(ldx, nyt, 0:nzt, nf), tmp(nxt, nyt, 0:nzt), g(ldx, nyt, 1:nzt-1))
fld = 0d0
open(newunit=u_in, file=trim(dir)//"/fields.bin", access="stream", form="unformatted", status="old")
do i = 1, nf
    read(u_in) tmp
    fld(1:nxt, :, :, i) = tmp
    ! the element's address (sequence association): no temporary copy
    call les_check(les_bind_field(ctx, ids(i), fld(1, 1, 0, i), ldx, 0, nzt + 1), "bind")
end do
close(u_in)
call les_check(les_upload(ctx), "upload")
do i = 1, 3   ! init's filter_data(u, v, w), 102_initial.f90:55-57
    call les_check(les_filter_data(ctx, ids(i)), "filter_data")
end do
call les_check(les_run(ctx, 1, nsteps, i), "les_run")
call les_check(les_download(ctx), "download")
open(newunit=u_out, file=trim(dir)//"/out.bin", access="stream", form="unformatted", status="replace")
write(u_out) fld(1:nxt, :, :, 1)
write(u_out) fld(1:nxt, :, :, 4)
! adjoint window of 3 steps from nsteps+1, checkpoint every 2, J over the last 2
call les_check(les_set_sgs_frozen(ctx, 1), "frozen")
call les_check(les_lin_zero(ctx), "lin_zero")
call les_check(les_adjoint_window(ctx, nsteps + 1, 3, 2, 2, J), "adjoint_window")
write(u_out) J
call les_check(les_lin_get(ctx, LES_F_THETA, g, ldx, 1, nzt - 1), "dJ/dtheta")
write(u_out) g(1:nxt, :, :)
call les_check(les_lin_get(ctx, LES_F_U, g, ldx, 1, nzt - 1), "dJ/du")
write(u_out) g(1:nxt, :, :)
close(u_out)
call les_check(les_destroy(ctx), "les_destroy")
write(*,*) "fortran window ok", J
end program drive_window
