/*
 * les_gpu.h -- C-ABI of the MI355X boundary-layer LES timestep.
 *
 * Drop-in boundary for the per-step routines of the GAbelois/OverlappingBLs
 * solver (the SBL/code Fortran files).  Every entry point mirrors one reference routine;
 * the citation on each declaration names the routine it replaces.  The
 * Fortran side binds these with ISO_C_BINDING (see the les_gpu_mod module
 * shipped with the package).
 *
 * Conventions
 *  - Grid: nxt x nyt horizontal points, z nodes 0..nzt.  Physical uvp levels
 *    are 1..nzt-1 (z = (k-1/2)dz), w levels are 1..nzt (z = (k-1)dz); level 0
 *    is the ghost below the wall.  One context holds the whole domain, so the
 *    reference's z-slab ghosts are ordinary neighbouring levels.
 *  - A 3-D field is (nzt+1) planes of nyt rows of nxt reals, x fastest.
 *    Host arrays passed to les_set_field/les_get_field may carry a row
 *    stride `ld` >= nxt (ld = ldx = nxt+2 for the reference's arrays); the
 *    padding columns nxt+1..ldx are not represented on the device.
 *  - All reals are FP64 at the boundary.  les_set_precision selects the
 *    internal arithmetic of the step (LES_FP64 or LES_MIXED).
 *  - Every call returns LES_OK (0) or a negative status.  The reference
 *    prints and stops on the same conditions; its Fortran wrapper turns a
 *    nonzero status into write + stop.
 *  - One host thread per context; calls are synchronous unless stated.
 */
#ifndef LES_GPU_H
#define LES_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define LES_OK          0
#define LES_E_BADARG   -1   /* invalid argument / unsupported option       */
#define LES_E_HIP      -2   /* HIP runtime error                           */
#define LES_E_CFL      -3   /* CFL > 0.6 (901_debug.f90:123-128)            */
#define LES_E_NOMEM    -4   /* device allocation failed                    */
#define LES_E_NODEV    -5   /* no GPU visible                              */
#define LES_E_IO       -6   /* restart file missing, short or not writable */
#define LES_E_SIZE     -7   /* request beyond the allocation-free entry point's fixed scratch */

/* ---- precision --------------------------------------------------------- */
#define LES_FP64   0
#define LES_MIXED  1

/* ---- option codes (namelist strings of 001_param.f90) -------------------- */
#define LES_WALL_ROUGH   0   /* walltype = 'rough'  */
#define LES_WALL_SMOOTH  1   /* walltype = 'smooth' */
#define LES_BC_WALL        0 /* lbc_mom / ubc_mom = 'wall'        */
#define LES_BC_STRESS_FREE 1 /* lbc_mom / ubc_mom = 'stress free' */
/* ocean_param stokes_action (014_stokes_drift.f90:45-56) */
#define LES_STOKES_DP       0 /* 'donelan.pierson1987' (not supported)        */
#define LES_STOKES_MONO     1 /* 'mono': amp_w, lambda_w, agl_Ust             */
#define LES_STOKES_DYNAMIC  2 /* 'dynamic': the same with gwave.in's record 1 */

/*
 * Run parameters, already nondimensionalised as init_nondimensional leaves
 * them (102_initial.f90:176-256): lengths / z_i, dt * u_scale / z_i,
 * coriol = f z_i / u_scale, ug = ug_dim / u_scale.  Dimensional inputs that
 * the reference rescales inside the step (wt_s, dTdz_top, nu_molec) stay
 * dimensional, as in the namelist.
 */
typedef struct les_params {
    int32_t nxt, nyt, nzt;     /* domain_param; nzt counts the top ghost    */
    int32_t model;             /* flow_param model (1 Smagorinsky, 5 Lagrangian
                                  scale-dependent)                          */
    int32_t cs_count;          /* time_param cs_count                       */
    int32_t dyn_init;          /* flow_param DYN_init                       */
    int32_t inilag;            /* flow_param inilag                         */
    int32_t restart_vel;       /* flow_param restart_vel                    */
    int32_t sgs;               /* flow_param sgs                            */
    int32_t molec;             /* flow_param molec                          */
    int32_t walltype;          /* LES_WALL_*                                */
    int32_t lbc_mom, ubc_mom;  /* LES_BC_*                                  */
    int32_t ubc;               /* flow_param ubc (1: sponge on)             */
    int32_t damping_method;    /* flow_param damping_method (1, 2, 3)       */
    int32_t coriolis_forcing;  /* flow_param coriolis_forcing               */
    int32_t use_mean_p_force;  /* flow_param use_mean_p_force               */
    int32_t ocean_flag;        /* flow_param ocean_flag                     */
    int32_t theta_flag;        /* temperature_param theta_flag              */
    int32_t restart_theta;     /* temperature_param restart_theta           */
    int32_t theta_init_time;   /* temperature_param theta_init_time         */
    int32_t passive_scalar;    /* 001_param.f90:512 compile-time switch     */
    int32_t lbc;               /* temperature_param lbc (1 flux, 0 temperature) */
    int32_t stokes_flag;       /* ocean_param stokes_flag                   */
    double lx_tot, ly_tot, lz_tot;   /* / z_i                               */
    double dt;                       /* nondimensional                      */
    double z_i, u_scale, T_scale;    /* scales                              */
    double nu_molec;                 /* m^2/s                               */
    double zo, d0;                   /* roughness, displacement / z_i       */
    double Pr;                       /* turbulent Prandtl number            */
    double wt_s;                     /* surface kinematic heat flux, K m/s  */
    double dTdz_top;                 /* K/m                                 */
    double coriol, ug, vg;           /* nondimensional Coriolis forcing     */
    double mean_p_force;             /* nondimensional                      */
    double lz_npz;                   /* lz*npz of step_scalar (= lz_tot)     */
    /* ocean branch, Stokes drift, damping and lbc = 0 (001_param.f90:304-333,
     * 102_initial.f90:226-250, 008_bottombc.f90:116-127) */
    int32_t stokes_action;           /* LES_STOKES_*                        */
    int32_t flag_crossvel;           /* ocean_param flag_crossVel           */
    double u_star;                   /* wind-stress friction velocity / u_scale */
    double rad_stress;               /* wind-stress angle (radians)         */
    double alpha_w;                  /* thermal expansion alpha_w * T_scale */
    double u_cross, v_cross;         /* cross-flow velocity / u_scale       */
    double amp_w, lambda_w;          /* wave amplitude and length, metres   */
    double agl_ust;                  /* Stokes drift direction, degrees     */
    double inv_strength;             /* K/m, damping method 3               */
    double T_s;                      /* surface temperature / T_scale       */
    double Co;                       /* flow_param Co (model 1)             */
} les_params;

/* ---- field identifiers ----------------------------------------------------
 * 3-D fields hold nzt+1 planes.  The plane fields (LES_F_*_SFC) hold one
 * plane.  The names follow the reference's module variables.            */
enum les_field {
    LES_F_U = 0, LES_F_V, LES_F_W, LES_F_THETA,            /* sim_param     */
    LES_F_RHSX, LES_F_RHSY, LES_F_RHSZ, LES_F_RHST,         /* RHS (this step)*/
    LES_F_RHSX_F, LES_F_RHSY_F, LES_F_RHSZ_F, LES_F_RHST_F, /* RHS history   */
    LES_F_DUDX, LES_F_DUDY, LES_F_DUDZ,
    LES_F_DVDX, LES_F_DVDY, LES_F_DVDZ,
    LES_F_DWDX, LES_F_DWDY, LES_F_DWDZ,
    LES_F_TXX, LES_F_TXY, LES_F_TXZ, LES_F_TYY, LES_F_TYZ, LES_F_TZZ,
    LES_F_DIVTX, LES_F_DIVTY, LES_F_DIVTZ,
    LES_F_P, LES_F_DPDX, LES_F_DPDY, LES_F_DPDZ,
    LES_F_CS_OPT2, LES_F_F_LM, LES_F_F_MM, LES_F_F_QN, LES_F_F_NN,   /* sgsmodule */
    LES_F_NU_T, LES_F_U_LAG, LES_F_V_LAG, LES_F_W_LAG,
    LES_F_S11, LES_F_S12, LES_F_S13, LES_F_S22, LES_F_S23, LES_F_S33, /* w-node strain */
    LES_F_CX, LES_F_CY, LES_F_CZ,                       /* convec output       */
    LES_F_DTDX, LES_F_DTDY, LES_F_DTDZ, LES_F_BETA_SCAL, LES_F_SGS_T3,
    LES_F_WORK0, LES_F_WORK1, LES_F_WORK2,              /* caller scratch      */
    LES_N_FIELDS3D,
    /* one-plane surface fields (bottombc module) */
    LES_F_USTAR_SFC = 100, LES_F_PSI_M_SFC, LES_F_PHI_M_SFC,
    LES_F_PSI_H_SFC, LES_F_PHI_H_SFC,
    LES_F_USTAR_TOP_SFC,   /* ocean branch: ustar_avg of the top wall (obukhov) */
    LES_F_END
};

typedef struct les_ctx les_ctx;

/* ---- lifecycle (102_initial.f90:4-61 init) -------------------------------- */
/* Allocates the device state for one domain and precomputes wavenumbers,
 * test-filter kernels and the pressure TDMA factorisation
 * (002_fft.f90:121-189, 006_test_filter.f90:41-77, 109_press_stag_array.f90:4-103). */
int les_create(const les_params* prm, les_ctx** out);
int les_destroy(les_ctx* ctx);
/* LES_MIXED: the work fields (derivatives, strain, stresses, divergences,
 * convective terms, theta gradients, beta, p and grad p), the momentum RHS
 * history (RHSx/y/z and their previous-step copies) and the Lagrangian
 * velocity sums (u_lag, v_lag, w_lag) are FP32; u, v, w, theta, RHS_T,
 * Cs_opt2, F_LM .. F_NN, the TDMA, the mean mode and the statistics stay
 * FP64.  The FP32 state is converted when the mode changes. */
int les_set_precision(les_ctx* ctx, int precision);
/* Stokes-drift profiles ust(0:nzt), vst(0:nzt) (014_stokes_drift.f90:29-115). */
int les_set_stokes(les_ctx* ctx, const double* ust, const double* vst);
/* Sponge profile sponge(0:nzt) (007_topbc.f90:15-85); computed by les_create
 * for damping_method 1, 2 and 3, overridable here. */
int les_set_sponge(les_ctx* ctx, const double* sponge);
/* The Stokes drift and sponge profiles the context holds (nzt+1 values
 * each, levels 0..nzt; any pointer may be null). */
int les_get_profiles(les_ctx* ctx, double* ust, double* vst, double* sponge);
/* Ocean wind stress used by wallstress_bottom from now on
 * (104_wallstress.f90:118-136): ustar = friction velocity / u_scale, agl in
 * degrees.  les_create takes u_star and rad_stress from les_params; with
 * flag_dynStress the caller reads ustar.in's record nums+ttshift_dynStress+1
 * (104_wallstress.f90:120-126) and passes it here before each step. */
int les_set_wind_stress(les_ctx* ctx, double ustar, double agl_deg);

/* ---- registration (the reference's module arrays stay the caller's) ------- */
/* Register a caller array for field `field` (levels lev0..lev0+nlev-1, row
 * stride ld); les_upload copies every registered array to the device,
 * les_download copies the device fields back (the sync points around a
 * device-resident les_run).  The arrays must stay valid while registered. */
int les_bind_field(les_ctx* ctx, int field, double* host, int ld, int lev0, int nlev);
int les_upload(les_ctx* ctx);
int les_download(les_ctx* ctx);

/* ---- field transfer -------------------------------------------------------- */
/* Copy levels lev0..lev0+nlev-1 of a field between host and device.  The
 * host block is nlev planes of nyt rows with row stride ld (doubles). */
int les_set_field(les_ctx* ctx, int field, const double* host, int ld, int lev0, int nlev);
int les_get_field(les_ctx* ctx, int field, double* host, int ld, int lev0, int nlev);
/* ---- restart files ------------------------------------------------------------
 * The reference's restart files, read and written on the host:
 *   <path>vel_tt<nums>.out  u, v, w, RHSx, RHSy, RHSz, Cs_opt2, F_LM, F_MM, F_QN, F_NN
 *   <path>temp_tt<nums>.out theta, RHS_T, sgs_t3 (wall level), psi_m (theta_flag only)
 * Each 3-D field is levels 1..nzt-1 as nzt-1 planes of nyt rows of nxt
 * native-endian doubles; each plane field one such plane (the MPI-IO layout of
 * 202_parallel.f90:1084-1117, 1252-1284).  nums is the %08d step counter of
 * the names; path is path_restart, ending in '/' (the directory must
 * exist).  Returns LES_E_IO when a file cannot be opened or has the wrong
 * size.
 * les_write_restart: checkpoint_final, 012_io.f90:496-559.
 * les_read_restart: read_flow_field, 102_01_ic.f90:682-734, and with
 * read_theta init_temperature_field, 102_initial.f90:740-797; then the
 * boundary ghost rows of data_exchange (202_parallel.f90:722-733) and
 * w(nzt) = 0 (110_project.f90:54).  Other fields and levels are left as
 * they are; the caller sets the restart flags (les_params.restart_vel,
 * restart_theta) and the SGS init flags as the reference's run would. */
int les_write_restart(les_ctx* ctx, const char* path, int nums);
int les_read_restart(les_ctx* ctx, const char* path, int nums, int read_theta);

/* Scalar state carried across steps (lagrange_Sdep's saved init flags). */
int les_set_sgs_init_flags(les_ctx* ctx, int f_lm_mm_init, int f_qn_nn_init);
int les_get_sgs_init_flags(les_ctx* ctx, int* f_lm_mm_init, int* f_qn_nn_init);

/* ---- derivative and transform routines (fields on the device) ------------ */
/* derivatives::ddx, 004_derivatives.f90:18-41.  All levels 0..nzt. */
int les_ddx(les_ctx* ctx, int dfdx, int f);
/* derivatives::ddy, 004_derivatives.f90:45-81. */
int les_ddy(les_ctx* ctx, int dfdy, int f);
/* derivatives::ddz_uv, 004_derivatives.f90:85-116 (global-domain form). */
int les_ddz_uv(les_ctx* ctx, int dfdz, int f);
/* derivatives::ddz_w, 004_derivatives.f90:155-183. */
int les_ddz_w(les_ctx* ctx, int dfdz, int f);
/* filter_data, 006_test_filter.f90:82-105: removes the x and y Nyquist modes. */
int les_filter_data(les_ctx* ctx, int f);
/* test_filter on one plane, 006_test_filter.f90:19-37; which = 2 (G_test,
 * width 2 delta) or 4 (G_test_test, width 4 delta). */
int les_test_filter(les_ctx* ctx, int f, int level, int which);

/* ---- one-plane host-array transforms (drop-ins for the reference's calls) -- */
/* fftn_mpi, 203_fftn_mpi.f90:4-35: plane (ld x nyt) -> half spectrum
 * (nxt/2+1) x nyt complex (interleaved re,im), Nyquist row/column zeroed,
 * unnormalised. */
int les_fftn(les_ctx* ctx, const double* in, int ld, double* out_c);
/* ifftn_mpi, 203_fftn_mpi.f90:39-67 (unnormalised). */
int les_ifftn(les_ctx* ctx, const double* in_c, double* out, int ld);
/* zero_padding, 108_01_padd.f90:73-132: (ld x nyt) -> (3nxt/2 x 3nyt/2). */
int les_zero_padding(les_ctx* ctx, const double* var, int ld, double* var_big);
/* No counterpart in the reference.  zero_padding (as les_zero_padding) of
 * nplanes planes already in device memory, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc): nothing is allocated; the intermediate of the two
 * passes lives in a static device array of the library, 48 MiB per GPU held
 * until the process ends.  d_var: nplanes planes of nyt rows of row stride
 * ld >= nxt (plane stride ld*nyt); d_var_big: nplanes planes of
 * (3nxt/2) x (3nyt/2).  Requires nplanes * nyt * 3nxt/2 <= 6*2^20 (48 MiB of
 * doubles), else LES_E_SIZE.  Synchronous on the context's stream.  All calls
 * on one GPU share the array: calls that overlap on one GPU (other threads,
 * other contexts) race. */
int les_zero_padding_noalloc(les_ctx* ctx, const double* d_var, int ld, double* d_var_big, int nplanes);
/* zero_unpadding, 108_01_padd.f90:136-197: (3nxt/2 x 3nyt/2) -> (ld x nyt). */
int les_zero_unpadding(les_ctx* ctx, const double* var_big, double* var, int ld);

/* ---- physics routines (fields on the device) ------------------------------- */
/* obukhov, 010_scalars.f90:675-856 (land branch). */
int les_obukhov(les_ctx* ctx, int jt);
/* wallstress_bottom / wallstress_top, 104_wallstress.f90:98-260. */
int les_wallstress_bottom(les_ctx* ctx);
int les_wallstress_top(les_ctx* ctx);
/* sgs_stag, 106_sgs_stag.f90:5-318, including lagrange_Sdep
 * (106_03_lagrange_Sdep.f90:6-362) and interpolag_Sdep
 * (106_06_interpolag_Sdep.f90:4-336) on update steps. */
int les_sgs_stag(les_ctx* ctx, int jt);
/* lagrange_Sdep alone, 106_03_lagrange_Sdep.f90:6-362 (with its
 * interpolag_Sdep); reads the w-node strain LES_F_S11..LES_F_S33. */
int les_lagrange_sdep(les_ctx* ctx, int jt);
/* interpolag_Sdep alone, 106_06_interpolag_Sdep.f90:4-336: moves F_LM, F_MM,
 * F_QN, F_NN to their upstream positions and zeroes u_lag, v_lag, w_lag. */
int les_interpolag_sdep(les_ctx* ctx);
/* divstress_uv, 107_divstress.f90:4-44 / divstress_w, 107_divstress.f90:48-104. */
int les_divstress_uv(les_ctx* ctx, int divt, int tx, int ty, int tz);
int les_divstress_w(les_ctx* ctx, int divt, int tx, int ty, int tz);
/* convec, 108_convec.f90:11-202 -> LES_F_CX, LES_F_CY, LES_F_CZ. */
int les_convec(les_ctx* ctx);
/* theta_all_in_one, 010_scalars.f90:111-200 (with scalar_RHS_calc,
 * calc_beta, step_scalar). */
int les_theta_all_in_one(les_ctx* ctx, int jt);
/* press_stag_array, 109_press_stag_array.f90:107-319, then dpdz
 * (103_flosol.f90:174-175). */
int les_press_stag_array(les_ctx* ctx, int jt);
/* project, 110_project.f90:4-62, followed by data_exchange ghost rules
 * (202_parallel.f90:722-733). */
int les_project(les_ctx* ctx);

/* Admas_Bashforth, 103_flosol.f90:319-389: u, v, w += dt (1.5 RHS - 0.5 RHS_f
 * + force), Euler (RHS_f <- RHS) at jt = 1 unless restart_vel, level 0 <- BOGUS. */
int les_ab2(les_ctx* ctx, int jt);
/* tdmaSolver_pipelined with press_stag_array's right-hand side
 * (109_press_stag_array.f90:201-319) on host half spectra: Hx, Hy, Hz hold
 * levels 0..nzt of (nxt/2+1) x nyt complex values (re, im interleaved, kx
 * fastest); p_hat receives levels 0..nzt-1 in the same layout. */
int les_tdma_solve(les_ctx* ctx, const double* hx, const double* hy, const double* hz, double* p_hat);

/* ---- whole step --------------------------------------------------------- */
/* flosol, 103_flosol.f90:4-290, minus file I/O (the ke/dudt diagnostics are
 * computed only on the steps that sample the time-averaged slices).
 * Returns LES_E_CFL if check_cfl would abort (the state is then undefined,
 * as the reference stops the program). */
int les_step(les_ctx* ctx, int jt);
/* nsteps consecutive flosol steps jt0..jt0+nsteps-1, enqueued without host
 * synchronisation; CFL is checked once at the end for every step.
 * *failed_jt receives the first step whose CFL exceeded 0.6 (or 0). */
int les_run(les_ctx* ctx, int jt0, int nsteps, int* failed_jt);
/* Last CFL value computed by newstep/check_cfl (901_debug.f90:66-143). */
int les_get_cfl(les_ctx* ctx, double* cfl);

/* ---- time-averaged x-z slices ----------------------------------------------- */
/* y-averaged slices of flow_slice (011_scalars2.f90:28-242) and theta_slice
 * (:246-318), accumulated with weight fr = c_count / p_count
 * (102_initial.f90:170) by the step when use_avgslice and mod(jt, c_count)
 * == 0 (103_flosol.f90:231-236); the theta slices only when theta_flag and
 * jt >= theta_init_time.  On a step with mod(jt, p_count) == 0 the running
 * sums close a window: they are copied to the window set (what the reference
 * writes with write_avgslice) and restart from zero.  The d/dt slices need
 * the flosol diagnostics dudt, dvdt, dwdt (103_flosol.f90:119-213, energy
 * :393-429), which the step then computes on those steps.  Every slice is
 * nxt x (nzt-1) values, x fastest, levels 1..nzt-1, FP64. */
enum les_slice {
    LES_S_U = 0, LES_S_V, LES_S_W, LES_S_P, LES_S_DUDT, LES_S_DVDT, LES_S_DWDT, LES_S_DUDZ, LES_S_DVDZ,
    LES_S_U2, LES_S_V2, LES_S_W2, LES_S_P2, LES_S_UV, LES_S_UW, LES_S_VW, LES_S_WP,
    LES_S_TXX, LES_S_TXY, LES_S_TXZ, LES_S_TYY, LES_S_TYZ, LES_S_TZZ,
    LES_S_U3, LES_S_V3, LES_S_W3,
    LES_S_U2_FLUC, LES_S_V2_FLUC, LES_S_UV_FLUC, LES_S_U2W_FLUC, LES_S_V2W_FLUC, LES_S_DISSIP,
    LES_S_UTXZ_FLUC, LES_S_VTYZ_FLUC, LES_S_WTZZ, LES_S_CS,
    LES_N_FLOW_SLICES,
    LES_S_THETA = LES_N_FLOW_SLICES, LES_S_THETA2, LES_S_T2_FLUC, LES_S_DTDZ, LES_S_SGS_T3, LES_S_WT, LES_S_NU_T,
    LES_N_SLICES
};
/* use_avgslice, c_count, p_count of the namelist (001_param.f90:73,140-141);
 * (re)starts both sums and windows from zero. */
int les_set_avgslice(les_ctx* ctx, int use_avgslice, int c_count, int p_count);
/* window 0: running sums; 1: the last closed window. */
int les_get_avgslice(les_ctx* ctx, int slice, int window, double* host);
/* write_avgslice (011_scalars2.f90:393-442) for every slice of the last
 * closed window (the theta slices when theta_flag): appends to
 * <path>result/aver_<name>.out, names as flow_slice / theta_slice pass them
 * (:203-244, 305-311).  average_dim_num 1: one line per level 1..nzt-1 of
 * nums*dt and the nxt values; 2: one line of nums*dt and the x-mean of each
 * level.  Values in Fortran E14.5 (fmt_5168, :21).  Call it where the
 * reference does, on the steps that close a window (mod(jt, p_count) == 0).
 * The directory <path>result/ must exist (LES_E_IO otherwise); LES_E_BADARG
 * before les_set_avgslice. */
int les_write_avgslice(les_ctx* ctx, const char* path, int nums, int average_dim_num);

/* ---- tangent-linear model and adjoint -------------------------------------- */
/* No reference counterpart: the reference has no derivative code.  These
 * differentiate the FP64 step of les_step (flosol, 103_flosol.f90:4-290) with
 * the Lagrangian SGS coefficients Cs_opt2 and F_LM..F_NN held at their
 * forward values ("frozen SGS"; Nu_t = Cs_opt2 Delta^2 |S| still varies with
 * the strain).  The wall model, obukhov's stability functions, advection,
 * the pressure projection and the scalar with calc_beta are all linearised.
 * Supported: LES_FP64, theta_flag = 1, the land branch (ocean_flag = 0)
 * with lbc = 0 or 1, lbc_mom = wall, ubc_mom = stress free, damping method 1,
 * steps jt >= 2 with jt > theta_init_time; else LES_E_BADARG.
 *
 * The linear state ("lin") is one tangent or adjoint of the step's state:
 * LES_F_U, LES_F_V, LES_F_W, LES_F_THETA, LES_F_RHSX, LES_F_RHSY,
 * LES_F_RHSZ, LES_F_RHST (levels 1..nzt-1; other levels are ignored and
 * read back as 0), LES_F_PSI_M_SFC (one plane) and the surface scalar flux
 * sgs_t3(:,:,1), addressed as LES_F_SGS_T3 with lev0 = 1, nlev = 1 (the next
 * step's obukhov reads its plane mean; with lbc = 0 it is the bulk flux
 * (T_s - theta(1)) kappa ustar / (log(dz/(2 z_os)) - psi_h), 010_scalars.f90
 * :359-394, a function of the state).  The objective J of the windows is the
 * mean over the plane and over the last obj_steps steps of ustar^2 (the
 * surface momentum flux |tau_w|, LES_OBJ_USTAR2, the default) or of the
 * surface scalar flux sgs_t3(:,:,1) (LES_OBJ_T_FLUX). */
#define LES_OBJ_USTAR2 0
#define LES_OBJ_T_FLUX 1
int les_lin_set_objective(les_ctx* ctx, int kind);
int les_lin_set(les_ctx* ctx, int field, const double* host, int ld, int lev0, int nlev);
int les_lin_get(les_ctx* ctx, int field, double* host, int ld, int lev0, int nlev);
int les_lin_zero(les_ctx* ctx);
/* Hold Cs_opt2 and F_* (skip lagrange_Sdep/interpolag_Sdep and the
 * Lagrangian accumulation) in every forward step, so that finite
 * differences of the forward step see the model the derivative is of. */
int les_set_sgs_frozen(les_ctx* ctx, int frozen);
/* Tangent-linear window: the state advances nsteps from jt0 and lin (the
 * tangent of the state at jt0) becomes the tangent at jt0 + nsteps; *dJ
 * (may be null) receives the tangent of J. */
int les_tlm_window(les_ctx* ctx, int jt0, int nsteps, int obj_steps, double* dJ);
int les_step_tlm(les_ctx* ctx, int jt);
/* Adjoint window: on entry the context holds the state at jt0 and lin the
 * adjoint seed y of the state at jt0 + nsteps (zero for J alone).  The
 * forward runs nsteps steps saving a checkpoint every k_ckpt steps; then,
 * segment by segment from the last, it recomputes the segment's steps with
 * tapes and runs their adjoint steps backwards (the last segment is taped
 * during the forward sweep).  k_ckpt = 0 chooses the spacing from free
 * device memory, up to the whole window taped at once.  On return the state
 * is at jt0 + nsteps, lin holds d(<y, x(jt0+nsteps)> + J)/dx(jt0) and *J
 * (may be null) the objective.  Device memory: nsteps/k_ckpt + 1
 * checkpoints (17 fields each) and k_ckpt tapes (12 fields + 9 padded
 * fields each; 550 MB at 128^3, 4.4 GB at 256^3). */
int les_adjoint_window(les_ctx* ctx, int jt0, int nsteps, int k_ckpt, int obj_steps, double* J);
int les_step_adj(les_ctx* ctx, int jt);
/* Dot-product test of stage `stage` of the linearised step (0 ..
 * les_lin_nstages()-1) or of the whole step (-1), about forward step jt from
 * the context's state (which advances one step): lhs = <M dx, y>, rhs =
 * <dx, M^T y> for seeded random dx, y on the stage's input and output
 * levels.  The stage's linear buffers are left zero. */
int les_lin_check(les_ctx* ctx, int stage, int jt, unsigned long long seed, double* lhs, double* rhs);
int les_lin_nstages(void);
const char* les_lin_stage_name(int stage);
/* Device time (ms, HIP events) of each stage's tangent and adjoint about
 * forward step jt from the context's state (which advances one step);
 * arrays of les_lin_nstages() entries. */
int les_lin_stage_times(les_ctx* ctx, int jt, double* tl_ms, double* adj_ms);

/* ---- timing support ------------------------------------------------------ */
/* Kernel families timed by les_time_kernel (no reference counterpart: the
 * reference has no timers; these back the roofline figures). */
#define LES_K_TDMA    0   /* batched Thomas solve of press_stag_array        */
#define LES_K_XDERIV  1   /* ddx of one field (x-line FFT pass)              */
#define LES_K_YDERIV  2   /* ddy of one field (y-line FFT pass)              */
#define LES_K_PAD     3   /* zero_padding of one field (x then y pass)       */
#define LES_K_COPY    4   /* plain device copy of one field (calibration)    */
/* Average device time (ms) of `reps` launches of one kernel family on the
 * context's stream, bracketed by HIP events, and the algorithmic HBM bytes
 * one launch must move.  Operates on scratch fields only. */
int les_time_kernel(les_ctx* ctx, int kernel, int reps, double* ms, double* bytes);

/* Per-family device time of one step (no reference counterpart).  Runs
 * flosol step jt as les_run does (the state advances one step) with every
 * launch of an instrumented kernel family bracketed by HIP events on the
 * stream it is enqueued on; ms[f] receives family f's summed device time,
 * bytes[f] its algorithmic HBM bytes (each launch's inputs read once and its
 * outputs written once) and launches[f] its launch count; arrays of
 * les_nfamilies() entries, names from les_family_name.  One profiled step
 * at a time per process. */
int les_nfamilies(void);
const char* les_family_name(int fam);
int les_profile_step(les_ctx* ctx, int jt, double* ms, double* bytes, int* launches);
/* The same over nsteps steps from jt0 (as les_run(ctx, jt0, nsteps, ...)),
 * with only family `family` bracketed (-1: every family), so that one
 * family's launches are timed in steps that otherwise run undisturbed;
 * ms/bytes/launches are summed over the steps.  No reference counterpart. */
int les_profile_run(les_ctx* ctx, int jt0, int nsteps, int family, double* ms, double* bytes, int* launches);

/* HIP stream the context enqueues on (hipStream_t as void*). */
void* les_stream(les_ctx* ctx);
/* Synchronise the context's stream. */
int les_sync(les_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* LES_GPU_H */
