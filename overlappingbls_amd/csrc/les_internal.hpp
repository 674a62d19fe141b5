// Internal declarations of the MI355X LES timestep library.
#pragma once
#include <initializer_list>
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <vector>
#include "../../include/les_gpu.h"
#include "fft_lds.hpp"

namespace les {

constexpr double BOGUS = -1234567890.0;  // 001_param.f90:40
constexpr double VONK = 0.4;             // 001_param.f90:32
constexpr double GRAV = 9.81;            // 001_param.f90:33
constexpr double PI = 3.1415926535897932384626433;
constexpr double B_SMOOTH = 5.2;         // 008_bottombc.f90:25
constexpr double OPFTIME = 1.5;          // 005_sgs.f90:12

// Spectral multipliers applied between the forward and inverse transform of
// a line kernel.
enum LineOp : int {
    OP_DERIV = 0,   // i k / n     (ddx / ddy)
    OP_NYQ = 1,     // 1/n, Nyquist zeroed (filter_data)
    OP_PAD = 2,     // n -> 3n/2, |m| < n/2 kept, 1/n (zero_padding)
    OP_UNPAD = 3,   // 3n/2 -> n, |m| < n/2 kept, 1/(3n/2) (zero_unpadding)
};

enum SpecOp : int {
    SP_FWD = 0,     // forward c2c along y, ky Nyquist row zeroed (fftn_mpi)
    SP_INV = 1,     // backward c2c along y (ifftn_mpi)
    SP_FILT = 2,    // forward, Nyquist row zeroed, * G, backward (test_filter)
    SP_MULINV = 3,  // * G, backward
};

// Twiddle tables for one transform length, both precisions.
struct Twiddle {
    int n = 0;
    FftPlan plan{};
    cx<double>* d = nullptr;
    cx<float>* f = nullptr;
};

struct Scalars;  // device-side per-step scalars (obukhov)

// Forward values of one step that its tangent-linear and adjoint read (FP64).
struct Tape {
    double* obk;                 // obukhov plane sums [4]
    double *u1, *v1;             // u, v at level 1 as the step found them
    double* uv1f;                // test-filtered u, v at level 1 (2 planes)
    double *thf, *tmean;         // filtered theta (before step_scalar), its plane means
    double* S[6];                // S11 S12 S13 S22 S23 S33
    double *nu, *cs;             // Nu_t, Cs_opt2
    double* dT[3];               // dTdx, dTdy, dTdz
    double *ub, *vb, *wb;        // padded u, v, w
    double* ob[3];               // padded vorticity
    double* tb[3];               // padded theta gradients
};
enum TapePoint : int { TAPE_OBUKHOV = 0, TAPE_WALL = 1, TAPE_THETA = 2, TAPE_END = 3 };
struct LinCtx;
// a caller array registered with les_bind_field
struct Binding {
    int id;
    double* h;
    int ld, lev0, nlev;
};

}  // namespace les

struct les_ctx {
    les_params prm;
    int nx, ny, nz, nx2, ny2, lhx, lhy;
    size_t plane, plane2, nlev, ns;  // ns = lhx*ny half-spectrum plane
    double dx, dy, dz;
    int precision;
    int dev;
    hipStream_t stream;
    hipStream_t aux;                // second stream (default priority): theta's chain, the padded u, v, w
    hipStream_t sums;               // third stream (default priority): obukhov's ordered plane sums
    hipEvent_t ev_fork, ev_join;
    hipEvent_t ev_obk, ev_th;       // obukhov sums done (sums) / plane means of theta done (aux)
    hipEvent_t ev_uvw;              // convec's padded u, v, w done on aux
    // device fields (FP64), nlev planes each
    double* f3[LES_N_FIELDS3D];
    // work fields in the precision of the current mode (double: f3[id];
    // mixed: FP32 copies in f32buf); null for the FP64 state fields
    void* wk[LES_N_FIELDS3D];
    float* f32buf[LES_N_FIELDS3D];  // FP32 work fields, allocated on the first switch to LES_MIXED
    les::cx<float>* phf;            // FP32 pressure spectrum of the mixed mode
    double* hzm;                    // FP64 plane sums of the z pressure RHS (nlev)
    double* lmean;                  // plane means of u and v (2 nlev), mixed-mode lagrange_Sdep
    // time-averaged slices (flow_slice / theta_slice) and their diagnostics
    int st_use, st_c, st_p;
    double st_fr;
    double* st_diag[3];             // dudt, dvdt, dwdt (FP64 fields)
    double *st_ke, *st_tmp, *st_dke;  // ke, filtered ke, one derivative of it
    double* st_mean;                // plane means of u, v, theta (3 nlev)
    double *st_run, *st_win;        // LES_N_SLICES slices of nxt x (nzt-1)
    double *ustar, *psi_m, *phi_m, *psi_h, *phi_h;  // one plane each
    double* ustar_top;              // ocean obukhov: ustar_avg of the top wall (one plane)
    double wind_ustar, wind_rad;    // ocean wind stress (u_star / u_scale, radians)
    double* dmean;                  // plane means of u, v, w for damping methods 2, 3 (3 nlev)
    double* noalloc;                // this device's copy of the allocation-free entry point's scratch
    double *u_big, *v_big, *w_big;                  // nlev planes of nx2*ny2
    // wavenumber tables (half spectrum, 002_fft.f90:143-167)
    double *kx2d, *ky2d, *k2, *Gtest, *Gtest2;
    // support of the two test filters (Nyquist column and row excluded, which
    // the forward transforms zero): columns kx < lag_kc (lag_kc4 for the 4
    // delta filter), rows with min(y, ny - y) < lag_mc
    int lag_kc, lag_kc4, lag_mc;
    unsigned char* notnyq;
    double *coef_b, *coef_c;  // factorised TDMA, nz planes of ns
    double *ust, *vst, *sponge;
    les::Twiddle tw_x, tw_y, tw_x2, tw_y2;
    // work memory
    double* work3[8];          // 8 extra nlev real fields
    double* big3[6];           // 6 padded fields (nlev planes of plane2)
    double* tbig[3];           // padded theta gradients
    double* th_mean;           // plane means of theta for calc_beta (nlev)
    double* th1_copy;          // theta(1) for the obukhov sums (one plane)
    les::cx<double>* spec3[4]; // half-spectrum fields (nlev planes of ns): Hx, Hy, Hz, p_hat
    double* lag_buf;           // lagrange_Sdep: 21 fields x 3 (raw, 2d, 4d)
    les::cx<double>* lag_spec; // 21 spectra
    double* fold[4];           // interpolag copies of F_*
    double* scal;              // device scalars
    double* red;               // reduction scratch (nlev * 8)
    double* cfl_dev;           // CFL per step of les_run
    double* cfl_host;          // pinned
    int cfl_cap;
    int f_lm_mm_init, f_qn_nn_init;
    int sgs_frozen;            // skip the Lagrangian SGS update (Cs_opt2, F_* held)
    int lin_obj;               // objective of the derivative windows (LES_OBJ_*)
    int fuse;                  // fused kernels (LES_FUSE=0 at creation: the separate ones)
    int overlap;               // theta's RHS and update on aux (LES_OVERLAP=0: in order)
    int lag_done;              // this step's u_lag accumulation already done (k_max3_lag)
    int obk_early;             // this step's obukhov sums already launched on aux (les_run)
    double* tscr[5];           // theta's private scratch while it runs on aux:
                               // work3[0..2], work3[6], big3[3] stand-ins
    les::LinCtx* lin;          // linearised-step buffers (les_lin.hip), or null
    double cfl;
    hipError_t err;            // first asynchronous HIP error since the last finish()
    std::vector<void*> allocs;
    std::vector<les::Binding> binds;
};

// Record the first failing HIP runtime call of a context.
inline void les_note(les_ctx* c, hipError_t e) {
    if (e != hipSuccess && c->err == hipSuccess) c->err = e;
}

// Work field id in the precision W of the current mode.
template <typename W>
inline W* WF(les_ctx* c, int id) {
    return static_cast<W*>(c->wk[id]);
}

namespace les {

// ---- per-family device time (les_prof.hip) ----------------------------------
// Kernel families of the step.  While les_profile_step runs a step, every
// launch of an instrumented family is bracketed by two HIP events on the
// stream it is enqueued on, and its algorithmic bytes (inputs read once,
// outputs written once) are recorded.
enum Fam : int {
    FAM_XDERIV = 0, FAM_YDERIV, FAM_XFILTER, FAM_YFILTER, FAM_XPAD, FAM_YPAD, FAM_XUNPAD, FAM_YUNPAD,
    FAM_XR2C, FAM_XC2R, FAM_YSPEC, FAM_TDMA, FAM_VORT_PAD_X, FAM_CROSS_UNPAD_X, FAM_SCALAR_UNPAD_X,
    FAM_LAG_CONTRACT_X, FAM_RHS_AB2, FAM_GRADP_PROJECT, FAM_SGS_STRESS, FAM_SCALAR_FLUX, FAM_DDZ,
    FAM_DIVSTRESS, FAM_CFL_MAX, FAM_PLANE_SUMS, FAM_PRESS_PREP, FAM_LAGRANGE_POINT, FAM_THETA_POINT,
    FAM_N
};
const char* fam_name(int f);
struct FamProf;
extern FamProf* g_prof;   // non-null only inside les_profile_step
// RAII bracket of one launch (or a dispatcher's launches) of family fam
struct FamScope {
    FamScope(hipStream_t s, int fam, double bytes);
    ~FamScope();
    int slot = -1;
    hipStream_t s = nullptr;
};

// ---- line transforms (k_lines.hip) ---------------------------------------
// Real x-lines: nrows rows, row r at in + r*ld_in (length nin) -> out + r*ld_out.
template <typename T, typename TI = T>
void xline_real(hipStream_t s, const TI* in, T* out, long nrows, int ld_in, int ld_out,
                const Twiddle& tin, const Twiddle& tout, LineOp op, double kfac);
// Real y-lines on nplanes planes: (nin rows x ncols) -> (nout rows x ncols).
template <typename T, typename TI = T>
void yline_real(hipStream_t s, const TI* in, T* out, int nplanes, int ncols, int ld_in, int ld_out,
                long ps_in, long ps_out, const Twiddle& tin, const Twiddle& tout, LineOp op, double kfac);
// Pruned half spectra (the test-filtered fields of lagrange_Sdep): only the
// columns kx < kst are stored, and only the rows y with min(y, n - y) < mc.
// x r2c (fftn_mpi first half): nrows real rows (length n) -> half-spectrum
// rows of stride ldc; columns kx < kst are stored (kst < 0: all n/2 + 1).
template <typename T, typename TI = T>
void x_r2c(hipStream_t s, const TI* in, cx<T>* out, long nrows, int ld_in, int ldc,
           const Twiddle& tw, double scale, int kst = -1);
// x c2r (ifftn_mpi second half); deriv != 0 applies i*kx (kx Nyquist = 0)
// first; modes kx >= kst are taken as zero and not read (kst < 0: none).
template <typename T>
void x_c2r(hipStream_t s, const cx<T>* in, T* out, long nrows, int ldc, int ld_out,
           const Twiddle& tw, int deriv, double kfac, int kst = -1);
// complex y-lines over nplanes half-spectrum planes: ncols columns of row
// stride ld (< 0: ncols), G of row stride ldG (< 0: ld).  With mc >= 0,
// SP_FWD stores and SP_MULINV reads only the rows with min(y, n - y) < mc
// (the rows outside a filter's support); the other ops take every row.
template <typename T>
void yline_spec(hipStream_t s, const cx<T>* in, cx<T>* out, int nplanes, int ncols, long ps,
                const Twiddle& tw, SpecOp op, const double* G, int ld = -1, int ldG = -1, int mc = -1);

// register-radix versions for the production line lengths (k_lines_rg.hip);
// return false when the
// line length has no specialisation
template <typename T, typename TI>
bool xline_real_ct(hipStream_t s, const TI* in, T* out, long nrows, int ld_in, int ld_out, const Twiddle& tin,
                   const Twiddle& tout, LineOp op, double kfac);
template <typename T, typename TI>
bool yline_real_ct(hipStream_t s, const TI* in, T* out, int nplanes, int ncols, int ld_in, int ld_out, long ps_in,
                   long ps_out, const Twiddle& tin, const Twiddle& tout, LineOp op, double kfac);
template <typename T, typename TI>
bool x_r2c_ct(hipStream_t s, const TI* in, cx<T>* out, long nrows, int ld_in, int ldc, const Twiddle& tw,
              double scale, int kst);
template <typename T>
bool x_c2r_ct(hipStream_t s, const cx<T>* in, T* out, long nrows, int ldc, int ld_out, const Twiddle& tw, int deriv,
              double kfac, int kst);
template <typename T>
bool yline_spec_ct(hipStream_t s, const cx<T>* in, cx<T>* out, int nplanes, int ncols, long ps, const Twiddle& tw,
                   SpecOp op, const double* G, int ld, int ldG, int mc);
// First x passes of a pad / unpad fused with the pointwise function that
// produces their input (k_lines_rg.hip); false when the line lengths have no
// register-radix specialisation (the caller then runs the separate kernels).
// ConvIn: padded u, v, w and b1..b3 = padded vorticity (cross product) or
// padded theta gradients (scalar advection); ust, vst the Galilean shift.
template <typename T>
struct ConvIn {
    const T *ub, *vb, *wb, *b1, *b2, *b3;
    const double *ust, *vst;
};
// VortIn: the velocity gradients of the vorticity
template <typename T>
struct VortIn {
    const T *uy, *uz, *vx, *vz, *wx, *wy;
};
// cross product (scalar = 0: components to o0, o1, o2) or scalar advection
// (scalar = 1: to o0) on levels 1..nz-1 of the 3/2 grid, then the x unpad:
// o* hold (nz - 1) * ny2 rows of length tout.n
template <typename T>
bool cross_unpad_x(hipStream_t s, const ConvIn<T>& a, int scalar, T* o0, T* o1, T* o2, int nz, int ny2,
                   const Twiddle& tin, const Twiddle& tout);
// vorticity on levels 1..nz, then the x pad: o0..o2 hold nz * ny rows of
// length tout.n
template <typename T>
bool vort_pad_x(hipStream_t s, const VortIn<T>& g, T* o0, T* o1, T* o2, int nz, int ny, const Twiddle& tin,
                const Twiddle& tout);
// process-wide switch: 0 forces the generic runtime-length kernels
extern int g_use_ct;

// ---- TDMA (k_tdma.hip) --------------------------------------------------
// Pressure TDMA; dsc is an FP64 store for d' (may equal ph when T = double),
// hzm optional FP64 plane sums for the mean mode.
template <typename T>
void tdma_press(les_ctx* c, const cx<T>* Hx, const cx<T>* Hy, const cx<T>* Hz, cx<T>* ph, double* dsc,
                const double* hzm);

// ---- pointwise / reductions (k_point.hip) -------------------------------
// Kernels templated on W read and write the work fields in precision W
// (double, or float in the mixed mode); the prognostic state, the SGS
// Lagrangian state, the surface planes and the reductions are FP64 always.
template <typename T>
void fill_levels(hipStream_t s, T* f, size_t plane, int lev0, int nlev, double v);
// v into n values at each of up to FILL_MAX addresses, in one launch
constexpr int FILL_MAX = 8;
template <typename T>
void fill_runs(hipStream_t s, long n, double v, std::initializer_list<T*> ps);
template <typename TO, typename TF>
void convert(hipStream_t s, TO* out, const TF* in, long n);
void level_means(hipStream_t s, const double* f, size_t plane, int lev0, int nlev, double inv_n, double* out);
void level_means_seq(hipStream_t s, const double* f, size_t plane, int lev0, int nlev, double nn, double* out);
void obukhov_dev(les_ctx* c, int jt);
void obukhov_theta_copy(les_ctx* c);   // theta(1) -> th1_copy on the context's stream
void obukhov_sums_dev(les_ctx* c, int jt, hipStream_t s);
void obukhov_finish_dev(les_ctx* c, int jt);
// run work on the auxiliary stream after everything enqueued so far on the
// context's stream; join() makes the context's stream wait for it
inline void fork(les_ctx* c) {
    les_note(c, hipEventRecord(c->ev_fork, c->stream));
    les_note(c, hipStreamWaitEvent(c->aux, c->ev_fork, 0));
}
inline void join(les_ctx* c) {
    les_note(c, hipEventRecord(c->ev_join, c->aux));
    les_note(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
}
template <typename W> void wallstress_dev(les_ctx* c, int top, const double* u1, const double* v1);
template <typename W> void strain_dev(les_ctx* c);
template <typename W> void lag_accumulate_dev(les_ctx* c);
template <typename W> void nu_t_stress_dev(les_ctx* c);
template <typename W> void strain_stress_dev(les_ctx* c);
template <typename W> void ddz_w_accum_dev(les_ctx* c, const W* f);
void cs_init_dev(les_ctx* c);
template <typename W> void interpolag_dev(les_ctx* c);
template <typename W> void lagrange_prep_dev(les_ctx* c, W* buf21);
template <typename W> void lagrange_contract_dev(les_ctx* c, const W* b2, const W* b4, int init_lm, int init_qn);
template <typename W, typename TI> void ddz_uv_dev(les_ctx* c, W* d, const TI* f);
template <typename W, typename TI> void ddz_w_dev(les_ctx* c, W* d, const TI* f);
template <typename W> void divstress_combine_dev(les_ctx* c, W* out, const W* a, const W* b, const W* d, int w_version);
template <typename W> void vorticity_dev(les_ctx* c, W* v1, W* v2, W* v3);
template <typename W>
void cross_product_dev(les_ctx* c, const W* b1, const W* b2, const W* b3, W* cx, W* cy, W* cz);
template <typename W> void theta_prep_dev(les_ctx* c);
template <typename W> void scalar_adv_dev(les_ctx* c, const W* dsx, const W* dsy, const W* dsz, W* rm);
template <typename W> void scalar_flux_dev(les_ctx* c, W* temp, int comp);
template <typename W> void scalar_accum_dev(les_ctx* c, const W* dtemp, const W* adv);
template <typename W> void calc_beta_dev(les_ctx* c, const double* means);
void step_scalar_dev(les_ctx* c);
template <typename W> void press_rhs_prep_dev(les_ctx* c, W* rx, W* ry, W* rz);
template <typename W>
void press_rhs_prep_on(les_ctx* c, W* rx, W* ry, W* rz, const double* u, const double* v, const double* w,
                       const W* divtz, double modi);
template <typename W> void dpdz_on(les_ctx* c, W* dpdz, const W* p);
template <typename W> void press_hz_sums_dev(les_ctx* c, double* out, double modi);
template <typename W> void dpdz_dev(les_ctx* c);
// RHS assembly + Admas_Bashforth, and RHS -= grad p + project, each one pass
template <typename W> void rhs_ab2_dev(les_ctx* c, int jt);
template <typename W> void ab2_dev(les_ctx* c);
template <typename W> void gradp_project_dev(les_ctx* c, int jt, bool theta_rows = true);
void theta_rows_dev(les_ctx* c);
template <typename W> void project_dev(les_ctx* c);
// check_cfl's maxima to out; lag: with this step's u_lag accumulation
// (lag_accumulate_dev) in the same pass over u, v, w
constexpr int MAX3_BLOCKS = 2048;  // partial maxima per component (c->red + 64)
void max3_dev(les_ctx* c, double* out, bool lag = false);

// ---- forward step and the linearised step (les_api.hip, les_lin.hip) ---------
// modi: lbc = 0's surface-temperature term of the bottom row
// (109_press_stag_array.f90:186-191), 0 otherwise
template <typename W>
void press_solve(les_ctx* c, const double* u, const double* v, const double* w, const W* divtz, W* pf, W* dpdx,
                 W* dpdy, W* dpdz, double modi = 0.0);
// one FP64 flosol step; tape != null records the values the linearised step
// reads; stats = false skips the time-averaged slices
void step_fp64(les_ctx* c, int jt, Tape* tape, bool stats);
void tape_save(les_ctx* c, Tape* t, int point);
void lin_destroy(les_ctx* c);

// ---- time-averaged slices (k_stats.hip) -----------------------------------
// a step samples the slices (and computes the dudt diagnostics) when
// use_avgslice and mod(jt, c_count) == 0 (103_flosol.f90:231)
inline bool stats_step(const les_ctx* c, int jt) { return c->st_use && c->st_run && (jt % c->st_c) == 0; }
template <typename W> void flow_slices_dev(les_ctx* c);
template <typename W> void theta_slices_dev(les_ctx* c);
void ke_dev(les_ctx* c, double* ke);
void add_levels_dev(les_ctx* c, double* d, const double* a, int skip_level1);

}  // namespace les

#define LES_HIP_CHECK(expr)                                  \
    do {                                                     \
        hipError_t e_ = (expr);                              \
        if (e_ != hipSuccess) return LES_E_HIP;              \
    } while (0)
