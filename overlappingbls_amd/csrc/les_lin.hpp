// Internal declarations of the linearised step: tangent-linear and adjoint
// kernels (k_lin.hip) and their orchestration (les_lin.hip).
#pragma once
#include "les_ops.hpp"

namespace les {

// Linear variables: the tangent (or adjoint) of each forward field the step
// reads or writes.  L_U .. L_PSI are the step's state; the rest are the
// step's intermediates.  Plain fields: nlev planes; L_PSI, L_Q, L_PHI, L_PSH,
// L_DUS: one plane; L_UB .. L_RM: nlev planes of the 3/2 grid.  L_Q is the
// surface scalar flux sgs_t3(:,:,1), which the next step's obukhov reads
// (its plane mean); L_PSH is psi_h, L_DUS the per-point ustar of the wall
// model.
enum LinVar : int {
    L_U, L_V, L_W, L_TH, L_RX, L_RY, L_RZ, L_RT, L_PSI, L_Q,
    L_RXF, L_RYF, L_RZF, L_RTF, L_PHI, L_PSH, L_DUS,
    L_UX, L_UY, L_UZ, L_VX, L_VY, L_VZ, L_WX, L_WY, L_WZ,
    L_TX, L_TY, L_TZ,
    L_S11, L_S12, L_S13, L_S22, L_S23, L_S33, L_NU,
    L_TXX, L_TXY, L_TXZ, L_TYY, L_TYZ, L_TZZ,
    L_DVX, L_DVY, L_DVZ,
    L_O1, L_O2, L_O3, L_CX, L_CY, L_CZ,
    L_ADV, L_FX, L_FY, L_FZ, L_BETA,
    L_P, L_DPX, L_DPY, L_DPZ,
    L_UB, L_VB, L_WB, L_OB1, L_OB2, L_OB3, L_CB1, L_CB2, L_CB3, L_TB1, L_TB2, L_TB3, L_RM,
    L_COUNT
};
constexpr int L_NSTATE = L_Q + 1;
inline bool lin_plane_only(int v) { return v == L_PSI || v == L_Q || v == L_PHI || v == L_PSH || v == L_DUS; }
inline bool lin_padded(int v) { return v >= L_UB; }

constexpr int LIN_NPART = 256;  // partial sums per reduction
constexpr int LIN_MAXCLR = 32;  // clears per batched launch

struct ClearBatch {
    double* p[LIN_MAXCLR];
    long n[LIN_MAXCLR];
    int count;
};

struct LinBufs {
    ClearBatch clr;  // queued clears (host side)
    double* v[L_COUNT];
    double* uv1f;    // tangent / adjoint of the test-filtered u, v at level 1 (2 planes)
    double* us2;     // 2 ustar d(ustar) (one plane)
    double* red;     // small reductions (64)
    double* small;   // per-level scratch (nlev)
    double* dsc;     // TDMA sweep store (ns * nlev complex)
    double* part;    // LIN_NPART partial sums
    double* acc;     // accumulators (8)
};

struct LP {
    long plane, plane2;
    int nz;
    double dz, dt, l2, nu;
};
LP lin_params(const les_ctx* c);

struct ObkArgs {
    double nn, dz, zo_, g_;
};
// the bulk surface flux of lbc = 0: (T_s - theta(1)) kappa ustar / (log_zos - psi_h)
struct SfcArgs {
    int bulk;
    double T_s, log_zos;
};
struct WallArgs {
    int smooth;
    double dsmooth, dz, d0, zo;
};

struct LinGrad {
    double *ux, *uy, *uz, *vx, *vy, *vz, *wx, *wy, *wz;
};
struct LinStrain {
    double *s11, *s12, *s13, *s22, *s23, *s33;
};
struct LinTau {
    double *xx, *xy, *xz, *yy, *yz, *zz;
};
struct LinBig {
    double *ub, *vb, *wb, *o1, *o2, *o3, *c1, *c2, *c3, *t1, *t2, *t3, *rm;
};
struct TapeS {
    const double* S[6];
    const double *nu, *cs;
};
struct TapeBig {
    const double *ub, *vb, *wb, *o1, *o2, *o3, *t1, *t2, *t3;
};
struct LinRhs {
    double *rx, *ry, *rz, *cx, *cy, *cz, *dx, *dy, *dz, *beta, *u, *v, *w;
};
struct RhsFlags {
    int buoy, cori, damp;
    double coriol;
};

inline LinGrad lin_grad(LinBufs& b) {
    return {b.v[L_UX], b.v[L_UY], b.v[L_UZ], b.v[L_VX], b.v[L_VY], b.v[L_VZ], b.v[L_WX], b.v[L_WY], b.v[L_WZ]};
}
inline LinStrain lin_strain(LinBufs& b) {
    return {b.v[L_S11], b.v[L_S12], b.v[L_S13], b.v[L_S22], b.v[L_S23], b.v[L_S33]};
}
inline LinTau lin_tau(LinBufs& b) {
    return {b.v[L_TXX], b.v[L_TXY], b.v[L_TXZ], b.v[L_TYY], b.v[L_TYZ], b.v[L_TZZ]};
}
inline LinBig lin_big(LinBufs& b) {
    return {b.v[L_UB], b.v[L_VB], b.v[L_WB], b.v[L_OB1], b.v[L_OB2], b.v[L_OB3], b.v[L_CB1],
            b.v[L_CB2], b.v[L_CB3], b.v[L_TB1], b.v[L_TB2], b.v[L_TB3], b.v[L_RM]};
}
inline TapeS tape_s(const Tape& t) { return {{t.S[0], t.S[1], t.S[2], t.S[3], t.S[4], t.S[5]}, t.nu, t.cs}; }
inline TapeBig tape_big(const Tape& t) {
    return {t.ub, t.vb, t.wb, t.ob[0], t.ob[1], t.ob[2], t.tb[0], t.tb[1], t.tb[2]};
}

// ---- host wrappers of the kernels (k_lin.hip) ----------------------------------
void lin_axpy(les_ctx* c, double* y, const double* x, double a, long n);
void lin_rand(les_ctx* c, double* f, long n, unsigned long long seed, unsigned long long salt, long off);
void lin_dot_part(les_ctx* c, const double* a, const double* b, long n, unsigned long long seed,
                  unsigned long long salt, long off, int sq, double* part);
void lin_acc(les_ctx* c, const double* part, double w, double* acc);
void lin_clear(les_ctx* c, LinBufs& b, double* p, long n);
void lin_flush(les_ctx* c, LinBufs& b);
void clear_levels(les_ctx* c, LinBufs& b, double* f, int l0, int l1);
void clear_levels2(les_ctx* c, LinBufs& b, double* f, int l0, int l1);

void obk_tl(les_ctx* c, const Tape& t, LinBufs& b, const ObkArgs& a);
void obk_adj(les_ctx* c, const Tape& t, LinBufs& b, const ObkArgs& a);
void wall_tl(les_ctx* c, const Tape& t, LinBufs& b, const ObkArgs& oa, const WallArgs& wa, double* dus2);
void wall_adj(les_ctx* c, const Tape& t, LinBufs& b, const ObkArgs& oa, const WallArgs& wa, double wobj);
void ddz_uv_tl(les_ctx* c, double* d, const double* f);
void ddz_w_tl(les_ctx* c, double* d, const double* f);
void grad_adj(les_ctx* c, double* lf, const double* tx, const double* ty, double* lz, int wform);
void strain_tl(les_ctx* c, LinBufs& b);
void strain_adj(les_ctx* c, LinBufs& b);
void stress_tl(les_ctx* c, const Tape& t, LinBufs& b);
void stress_adj(les_ctx* c, const Tape& t, LinBufs& b);
void div_tl(les_ctx* c, double* out, const double* a, const double* b, const double* f, int wver);
void div_adj(les_ctx* c, LinBufs& b, double* const* t);
void vort_tl(les_ctx* c, LinBufs& b);
void vort_adj(les_ctx* c, LinBufs& b);
void cross_tl(les_ctx* c, const Tape& t, LinBufs& b);
void cross_adj(les_ctx* c, const Tape& t, LinBufs& b);
void sadv_tl(les_ctx* c, const Tape& t, LinBufs& b);
void sadv_adj(les_ctx* c, const Tape& t, LinBufs& b);
void sflux_tl(les_ctx* c, const Tape& t, LinBufs& b);
void sflux_adj(les_ctx* c, const Tape& t, LinBufs& b);
void sfc_tl(les_ctx* c, const Tape& t, LinBufs& b, const ObkArgs& oa, const WallArgs& wa, const SfcArgs& sa);
void sfc_adj(les_ctx* c, const Tape& t, LinBufs& b, const ObkArgs& oa, const WallArgs& wa, const SfcArgs& sa);
void lin_addc(les_ctx* c, double* y, double a, long n);
void sdiv_tl(les_ctx* c, LinBufs& b, const double* t1, const double* t2);
void sdiv_adj(les_ctx* c, LinBufs& b, const double* t1, const double* t2);
void beta_tl(les_ctx* c, const Tape& t, LinBufs& b, double g);
void beta_adj(les_ctx* c, const Tape& t, LinBufs& b, double g);
void ab2_tl(les_ctx* c, double* x, const double* r, const double* rf);
void ab2_adj(les_ctx* c, const double* lx, double* lr, double* lrf);
void rhs_tl(les_ctx* c, const LinRhs& a, const RhsFlags& f);
void rhs_adj(les_ctx* c, const LinRhs& a, const RhsFlags& f);
void sub_tl(les_ctx* c, double* r, const double* d);
void sub_adj(les_ctx* c, const double* lr, double* ld);
void project_tl(les_ctx* c, LinBufs& b, int sf);
void project_adj(les_ctx* c, LinBufs& b, int sf);
void step_scalar_tl(les_ctx* c, LinBufs& b);
void step_scalar_adj(les_ctx* c, LinBufs& b);
void press_adj(les_ctx* c, LinBufs& b);

}  // namespace les
