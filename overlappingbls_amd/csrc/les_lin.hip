// Discrete tangent-linear model and adjoint of the FP64 step, with
// checkpointed adjoint windows.  The reference has neither; this is the
// derivative of the step that les_step runs (103_flosol.f90:4-290 as
// restated in les_api.hip), with the Lagrangian SGS coefficient Cs_opt2 and
// the Lagrangian averages F_* held at their forward values (frozen SGS): the
// eddy viscosity Nu_t = Cs_opt2 Delta^2 |S| stays a function of the strain.
//
// The step is a chain of stages, each a map between linear variables (the
// tangents of the forward's fields).  tlm runs the stages in order, adj runs
// their transposes in reverse.  A stage's forward values come from a tape
// recorded by a forward step (step_fp64 with a Tape).  An adjoint window
// checkpoints the state every k steps, then for each segment, last first,
// reruns the segment's forward steps into k tapes and sweeps its adjoint
// steps backwards.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#include "les_lin.hpp"

namespace les {

// the forward fields an adjoint window checkpoints (everything a step reads
// from the previous one), plus psi_m and the init flags of lagrange_Sdep
static const int CKPT_IDS[] = {LES_F_U,     LES_F_V,    LES_F_W,     LES_F_THETA, LES_F_RHSX,  LES_F_RHSY,
                               LES_F_RHSZ,  LES_F_RHST, LES_F_CS_OPT2, LES_F_F_LM, LES_F_F_MM, LES_F_F_QN,
                               LES_F_F_NN, LES_F_U_LAG, LES_F_V_LAG, LES_F_W_LAG, LES_F_SGS_T3};
constexpr int N_CKPT = sizeof(CKPT_IDS) / sizeof(CKPT_IDS[0]);

struct Ckpt {
    double* f[N_CKPT];
    double* psi;
    int flm, fqn;
};

struct LinCtx {
    LinBufs b;
    ObkArgs oa;
    std::vector<Tape> tapes;
    std::vector<Ckpt> ckpts;
    double* seed[L_NSTATE];
};

namespace {

template <typename T>
hipError_t lalloc(les_ctx* c, T** p, size_t count) {
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, sizeof(T) * (count ? count : 1));
    if (e != hipSuccess) return e;
    c->allocs.push_back(v);
    *p = static_cast<T*>(v);
    return hipSuccess;
}

size_t lv_size(const les_ctx* c, int v) {
    if (lin_plane_only(v)) return c->plane;
    if (lin_padded(v)) return c->plane2 * c->nlev;
    return c->plane * c->nlev;
}
size_t lv_plane(const les_ctx* c, int v) { return lin_padded(v) ? c->plane2 : c->plane; }

ObkArgs obk_args(const les_ctx* c);

hipError_t lin_alloc(les_ctx* c) {
    if (c->lin) return hipSuccess;
    LinCtx* L = new LinCtx();
    LinBufs& b = L->b;
    hipError_t e = hipSuccess;
    for (int v = 0; v < L_COUNT && e == hipSuccess; ++v) e = lalloc(c, &b.v[v], lv_size(c, v));
    if (e == hipSuccess) e = lalloc(c, &b.uv1f, 2 * c->plane);
    if (e == hipSuccess) e = lalloc(c, &b.us2, c->plane);
    if (e == hipSuccess) e = lalloc(c, &b.red, 64);
    if (e == hipSuccess) e = lalloc(c, &b.small, c->nlev);
    if (e == hipSuccess) e = lalloc(c, &b.dsc, 2 * c->ns * c->nlev);
    if (e == hipSuccess) e = lalloc(c, &b.part, LIN_NPART);
    if (e == hipSuccess) e = lalloc(c, &b.acc, 8);
    for (int v = 0; v < L_NSTATE && e == hipSuccess; ++v) e = lalloc(c, &L->seed[v], lv_size(c, v));
    if (e != hipSuccess) {
        delete L;
        return e;
    }
    c->lin = L;
    L->oa = obk_args(c);
    for (int v = 0; v < L_COUNT; ++v) fill_levels(c->stream, b.v[v], lv_size(c, v), 0, 1, 0.0);
    fill_levels(c->stream, b.acc, 8, 0, 1, 0.0);
    return hipSuccess;
}

hipError_t tape_alloc(les_ctx* c, Tape& t) {
    const size_t F3 = c->plane * c->nlev, B3 = c->plane2 * c->nlev, np = c->plane;
    hipError_t e = lalloc(c, &t.obk, 8);
    if (e == hipSuccess) e = lalloc(c, &t.u1, np);
    if (e == hipSuccess) e = lalloc(c, &t.v1, np);
    if (e == hipSuccess) e = lalloc(c, &t.uv1f, 2 * np);
    if (e == hipSuccess) e = lalloc(c, &t.thf, F3);
    if (e == hipSuccess) e = lalloc(c, &t.tmean, c->nlev);
    for (int m = 0; m < 6 && e == hipSuccess; ++m) e = lalloc(c, &t.S[m], F3);
    if (e == hipSuccess) e = lalloc(c, &t.nu, F3);
    if (e == hipSuccess) e = lalloc(c, &t.cs, F3);
    for (int m = 0; m < 3 && e == hipSuccess; ++m) e = lalloc(c, &t.dT[m], F3);
    if (e == hipSuccess) e = lalloc(c, &t.ub, B3);
    if (e == hipSuccess) e = lalloc(c, &t.vb, B3);
    if (e == hipSuccess) e = lalloc(c, &t.wb, B3);
    for (int m = 0; m < 3 && e == hipSuccess; ++m) e = lalloc(c, &t.ob[m], B3);
    for (int m = 0; m < 3 && e == hipSuccess; ++m) e = lalloc(c, &t.tb[m], B3);
    return e;
}

hipError_t ensure_tapes(les_ctx* c, int n) {
    LinCtx* L = c->lin;
    while ((int)L->tapes.size() < n) {
        Tape t;
        std::memset(&t, 0, sizeof(t));
        const hipError_t e = tape_alloc(c, t);
        if (e != hipSuccess) return e;
        L->tapes.push_back(t);
    }
    return hipSuccess;
}

// device bytes of one tape (tape_alloc)
size_t tape_bytes(const les_ctx* c) {
    const size_t F3 = c->plane * c->nlev, B3 = c->plane2 * c->nlev, np = c->plane;
    return 8 * (8 + 4 * np + 12 * F3 + c->nlev + 9 * B3);
}

// Checkpoint spacing for k_ckpt = 0: tapes already held plus those that fit
// in the free device memory less a sixteenth of the device, within 1..nsteps.
int auto_ckpt(les_ctx* c, int nsteps) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 1;
    const size_t have = c->lin ? c->lin->tapes.size() : 0;
    const size_t margin = tot / 16;
    const size_t fit = have + (fr > margin ? (fr - margin) / tape_bytes(c) : 0);
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)nsteps, fit));
}

hipError_t ensure_ckpts(les_ctx* c, int n) {
    LinCtx* L = c->lin;
    const size_t F3 = c->plane * c->nlev;
    while ((int)L->ckpts.size() < n) {
        Ckpt k;
        std::memset(&k, 0, sizeof(k));
        hipError_t e = hipSuccess;
        for (int i = 0; i < N_CKPT && e == hipSuccess; ++i) e = lalloc(c, &k.f[i], F3);
        if (e == hipSuccess) e = lalloc(c, &k.psi, c->plane);
        if (e != hipSuccess) return e;
        L->ckpts.push_back(k);
    }
    return hipSuccess;
}

void dcopy(les_ctx* c, double* d, const double* s, size_t n) {
    les_note(c, hipMemcpyAsync(d, s, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
}

void ckpt_save(les_ctx* c, Ckpt& k) {
    const size_t F3 = c->plane * c->nlev;
    for (int i = 0; i < N_CKPT; ++i) dcopy(c, k.f[i], c->f3[CKPT_IDS[i]], F3);
    dcopy(c, k.psi, c->psi_m, c->plane);
    k.flm = c->f_lm_mm_init;
    k.fqn = c->f_qn_nn_init;
}
void ckpt_load(les_ctx* c, const Ckpt& k) {
    const size_t F3 = c->plane * c->nlev;
    for (int i = 0; i < N_CKPT; ++i) dcopy(c, c->f3[CKPT_IDS[i]], k.f[i], F3);
    dcopy(c, c->psi_m, k.psi, c->plane);
    c->f_lm_mm_init = k.flm;
    c->f_qn_nn_init = k.fqn;
}

// ---- per-step constants ------------------------------------------------------------
struct Run {
    les_ctx* c;
    LinBufs& b;
    const Tape& t;
    double wobj;      // weight of the objective at this step
    double* dus2;     // tlm: 2 us d(us) plane (ustar^2 objective steps), or null
    ObkArgs oa;
    WallArgs wa;
    RhsFlags rf;
    double gbeta;
    int sf;
    SfcArgs sa;
};

ObkArgs obk_args(const les_ctx* c) {
    const les_params& P = c->prm;
    ObkArgs a;
    a.nn = (double)(c->nx * c->ny);
    a.dz = c->dz;
    double sl = 0.0;  // zo_ as the forward forms it (ordered sum of a uniform plane)
    const double lz = log(P.zo);
    for (size_t i = 0; i < c->plane; ++i) sl += lz;
    a.zo_ = exp(sl / a.nn);
    a.g_ = GRAV / (P.u_scale * P.u_scale / P.z_i);
    return a;
}
WallArgs wall_args(const les_ctx* c) {
    const les_params& P = c->prm;
    WallArgs w;
    w.smooth = P.walltype == LES_WALL_SMOOTH;
    w.dsmooth = log(0.5 * c->dz / (P.nu_molec / P.u_scale / P.z_i)) / VONK + B_SMOOTH;
    w.dz = c->dz;
    w.d0 = P.d0;
    w.zo = P.zo;
    return w;
}
SfcArgs sfc_args(const les_ctx* c) {
    const les_params& P = c->prm;
    SfcArgs a;
    a.bulk = P.lbc == 0;
    a.T_s = P.T_s;
    a.log_zos = log(c->dz / (2.0 * ((1.0 / 10.0) * P.zo)));   // z_os = zo/10, 008_bottombc.f90:126
    return a;
}
RhsFlags rhs_flags(const les_ctx* c) {
    const les_params& P = c->prm;
    RhsFlags f;
    f.buoy = P.theta_flag && !P.passive_scalar;
    f.cori = P.coriolis_forcing;
    f.coriol = P.coriol;
    f.damp = (P.ubc_mom == LES_BC_STRESS_FREE && P.ubc == 1 && P.damping_method == 1);
    return f;
}

double* V(Run& r, int v) { return r.b.v[v]; }
int NZ(Run& r) { return r.c->nz; }
int NL(Run& r) { return (int)r.c->nlev; }
size_t NP(Run& r) { return r.c->plane; }
size_t NP2(Run& r) { return r.c->plane2; }
double* W3(Run& r, int i) { return r.c->work3[i]; }

// lambda_dst(levels l0..l1) += a * unpad(lambda_src(levels l0..l1)) (the
// transpose of a pad, a = 9/4) or += a * pad(...) (transpose of an unpad, a = 4/9)
void add_unpad(Run& r, double* dst, double* src, int l0, int l1, double a) {
    if (l1 < l0) return;
    unpad_acc_f(r.c, src + (size_t)l0 * NP2(r), dst + (size_t)l0 * NP(r), l1 - l0 + 1, a);
}
void add_pad(Run& r, double* dst, double* src, int l0, int l1, double a) {
    if (l1 < l0) return;
    pad_acc_f(r.c, src + (size_t)l0 * NP(r), dst + (size_t)l0 * NP2(r), l1 - l0 + 1, a);
}
void clear_all(Run& r, int v) { lin_clear(r.c, r.b, V(r, v), (long)lv_size(r.c, v)); }

// ---- stages ---------------------------------------------------------------------------
// 0: obukhov (010_scalars.f90:758-843): u(1), v(1), psi_m, theta(1) -> psi_m, phi_m
void s_obk_tl(Run& r) { obk_tl(r.c, r.t, r.b, r.oa); }
void s_obk_adj(Run& r) { obk_adj(r.c, r.t, r.b, r.oa); }

// 1: velocity gradients (004_derivatives.f90; 103_flosol.f90:58-80)
void s_deriv_tl(Run& r) {
    les_ctx* c = r.c;
    const int nl = NL(r);
    ddx_f<double, double>(c, V(r, L_UX), V(r, L_U), 0, nl);
    ddy_f<double, double>(c, V(r, L_UY), V(r, L_U), 0, nl);
    ddx_f<double, double>(c, V(r, L_VX), V(r, L_V), 0, nl);
    ddy_f<double, double>(c, V(r, L_VY), V(r, L_V), 0, nl);
    ddx_f<double, double>(c, V(r, L_WX), V(r, L_W), 0, nl);
    ddy_f<double, double>(c, V(r, L_WY), V(r, L_W), 0, nl);
    ddz_uv_tl(c, V(r, L_UZ), V(r, L_U));
    ddz_uv_tl(c, V(r, L_VZ), V(r, L_V));
    ddz_w_tl(c, V(r, L_WZ), V(r, L_W));
}
// lambda_f += the adjoint of its three gradients; the gradient adjoints are
// not cleared: their first source in the adjoint step writes them (the
// vorticity and strain for u, v, w; the scalar SGS flux for theta)
void grad_adj_field(Run& r, int f, int fx, int fy, int fz, int wform) {
    les_ctx* c = r.c;
    const int nl = NL(r);
    ddx_f<double, double>(c, W3(r, 0), V(r, fx), 0, nl);
    ddy_f<double, double>(c, W3(r, 1), V(r, fy), 0, nl);
    grad_adj(c, V(r, f), W3(r, 0), W3(r, 1), V(r, fz), wform);
}
void s_deriv_adj(Run& r) {
    grad_adj_field(r, L_U, L_UX, L_UY, L_UZ, 0);
    grad_adj_field(r, L_V, L_VX, L_VY, L_VZ, 0);
    grad_adj_field(r, L_W, L_WX, L_WY, L_WZ, 1);
}

// 2: filter_data(theta) in place (010_scalars.f90:111-118)
void s_thfilt(Run& r) { filter_f<double>(r.c, V(r, L_TH), 0, NL(r)); }

// 3: theta gradients
void s_thgrad_tl(Run& r) {
    les_ctx* c = r.c;
    ddx_f<double, double>(c, V(r, L_TX), V(r, L_TH), 0, NL(r));
    ddy_f<double, double>(c, V(r, L_TY), V(r, L_TH), 0, NL(r));
    ddz_uv_tl(c, V(r, L_TZ), V(r, L_TH));
}
void s_thgrad_adj(Run& r) { grad_adj_field(r, L_TH, L_TX, L_TY, L_TZ, 0); }

// 4: RHS_T history (theta_all_in_one :130-140)
void s_thhist_tl(Run& r) { dcopy(r.c, V(r, L_RTF), V(r, L_RT), r.c->plane * r.c->nlev); }
void s_thhist_adj(Run& r) {
    lin_axpy(r.c, V(r, L_RT), V(r, L_RTF), 1.0, (long)(r.c->plane * r.c->nlev));
    clear_all(r, L_RTF);
}

// 5: padded theta gradients (:466-470)
void s_thpad_tl(Run& r) {
    les_ctx* c = r.c;
    const int nz = NZ(r);
    const size_t np = NP(r), np2 = NP2(r);
    pad_f<double, double>(c, V(r, L_TX) + np, V(r, L_TB1) + np2, nz - 1);
    pad_f<double, double>(c, V(r, L_TY) + np, V(r, L_TB2) + np2, nz - 1);
    pad_f<double, double>(c, V(r, L_TZ) + 2 * np, V(r, L_TB3) + 2 * np2, nz - 1);
}
void s_thpad_adj(Run& r) {
    const int nz = NZ(r);
    add_unpad(r, V(r, L_TX), V(r, L_TB1), 1, nz - 1, 9.0 / 4.0);
    add_unpad(r, V(r, L_TY), V(r, L_TB2), 1, nz - 1, 9.0 / 4.0);
    add_unpad(r, V(r, L_TZ), V(r, L_TB3), 2, nz - 1, 9.0 / 4.0);
    // no clearing: the scalar advection's adjoint writes levels 1..nz-1 of all three
}

// 6: wall stress at the bottom (104_wallstress.f90:98-182)
void s_wall_tl(Run& r) {
    les_ctx* c = r.c;
    const size_t np = NP(r);
    dcopy(c, r.b.uv1f, V(r, L_U) + np, np);
    dcopy(c, r.b.uv1f + np, V(r, L_V) + np, np);
    test_filter_f<double>(c, r.b.uv1f, 2, c->Gtest);
    wall_tl(c, r.t, r.b, r.oa, r.wa, r.dus2);
}
void s_wall_adj(Run& r) {
    les_ctx* c = r.c;
    const size_t np = NP(r);
    wall_adj(c, r.t, r.b, r.oa, r.wa, r.c->lin_obj == LES_OBJ_USTAR2 ? r.wobj : 0.0);
    test_filter_f<double>(c, r.b.uv1f, 2, c->Gtest);
    lin_axpy(c, V(r, L_U) + np, r.b.uv1f, 1.0, (long)np);
    lin_axpy(c, V(r, L_V) + np, r.b.uv1f + np, 1.0, (long)np);
}

// 7, 8: strain, Nu_t and tau_ij (106_sgs_stag.f90)
void s_strain_tl(Run& r) { strain_tl(r.c, r.b); }
void s_strain_adj(Run& r) { strain_adj(r.c, r.b); }
void s_stress_tl(Run& r) { stress_tl(r.c, r.t, r.b); }
void s_stress_adj(Run& r) { stress_adj(r.c, r.t, r.b); }

// 9: divstress (107_divstress.f90)
void s_div_tl(Run& r) {
    les_ctx* c = r.c;
    const int nl = NL(r);
    double *a = W3(r, 0), *b = W3(r, 1);
    ddx_f<double, double>(c, a, V(r, L_TXX), 0, nl);
    ddy_f<double, double>(c, b, V(r, L_TXY), 0, nl);
    div_tl(c, V(r, L_DVX), a, b, V(r, L_TXZ), 0);
    ddx_f<double, double>(c, a, V(r, L_TXY), 0, nl);
    ddy_f<double, double>(c, b, V(r, L_TYY), 0, nl);
    div_tl(c, V(r, L_DVY), a, b, V(r, L_TYZ), 0);
    ddx_f<double, double>(c, a, V(r, L_TXZ), 0, nl);
    ddy_f<double, double>(c, b, V(r, L_TYZ), 0, nl);
    div_tl(c, V(r, L_DVZ), a, b, V(r, L_TZZ), 1);
}
void s_div_adj(Run& r) {
    les_ctx* c = r.c;
    const int nl = NL(r);
    double* t[6];
    for (int i = 0; i < 6; ++i) t[i] = W3(r, i);
    ddx_f<double, double>(c, t[0], V(r, L_DVX), 0, nl);
    ddy_f<double, double>(c, t[1], V(r, L_DVX), 0, nl);
    ddx_f<double, double>(c, t[2], V(r, L_DVY), 0, nl);
    ddy_f<double, double>(c, t[3], V(r, L_DVY), 0, nl);
    ddx_f<double, double>(c, t[4], V(r, L_DVZ), 0, nl);
    ddy_f<double, double>(c, t[5], V(r, L_DVZ), 0, nl);
    div_adj(c, r.b, t);
}

// 10-14: convec (108_convec.f90)
void s_cpad_tl(Run& r) {
    const int nl = NL(r);
    pad_f<double, double>(r.c, V(r, L_U), V(r, L_UB), nl);
    pad_f<double, double>(r.c, V(r, L_V), V(r, L_VB), nl);
    pad_f<double, double>(r.c, V(r, L_W), V(r, L_WB), nl);
}
void s_cpad_adj(Run& r) {
    const int nz = NZ(r);
    add_unpad(r, V(r, L_U), V(r, L_UB), 1, nz - 1, 9.0 / 4.0);
    add_unpad(r, V(r, L_V), V(r, L_VB), 1, nz - 1, 9.0 / 4.0);
    add_unpad(r, V(r, L_W), V(r, L_WB), 1, nz - 1, 9.0 / 4.0);
    // no clearing: the scalar advection's adjoint writes levels 1..nz-1 first
}
void s_vort_tl(Run& r) { vort_tl(r.c, r.b); }
void s_vort_adj(Run& r) { vort_adj(r.c, r.b); }
void s_vpad_tl(Run& r) {
    const int nz = NZ(r);
    const size_t np = NP(r), np2 = NP2(r);
    for (int m = 0; m < 3; ++m) pad_f<double, double>(r.c, V(r, L_O1 + m) + np, V(r, L_OB1 + m) + np2, nz);
}
void s_vpad_adj(Run& r) {
    const int nz = NZ(r);
    for (int m = 0; m < 3; ++m) {
        add_unpad(r, V(r, L_O1 + m), V(r, L_OB1 + m), 1, nz, 9.0 / 4.0);  // the cross product writes these
    }
}
void s_cross_tl(Run& r) { cross_tl(r.c, r.t, r.b); }
void s_cross_adj(Run& r) { cross_adj(r.c, r.t, r.b); }
void s_cunpad_tl(Run& r) {
    const int nz = NZ(r);
    const size_t np = NP(r), np2 = NP2(r);
    for (int m = 0; m < 3; ++m) unpad_f<double>(r.c, V(r, L_CB1 + m) + np2, V(r, L_CX + m) + np, nz - 1);
}
void s_cunpad_adj(Run& r) {
    const int nz = NZ(r);
    for (int m = 0; m < 3; ++m) {
        add_pad(r, V(r, L_CB1 + m), V(r, L_CX + m), 1, nz - 1, 4.0 / 9.0);  // rhs writes lambda_C
    }
}

// 15-20: scalar RHS, calc_beta, step_scalar (010_scalars.f90:466-671)
void s_sadv_tl(Run& r) { sadv_tl(r.c, r.t, r.b); }
void s_sadv_adj(Run& r) { sadv_adj(r.c, r.t, r.b); }
void s_sunpad_tl(Run& r) { unpad_f<double>(r.c, V(r, L_RM) + NP2(r), V(r, L_ADV) + NP(r), NZ(r) - 1); }
void s_sunpad_adj(Run& r) {
    add_pad(r, V(r, L_RM), V(r, L_ADV), 1, NZ(r) - 1, 4.0 / 9.0);  // k_sdiv_adj writes lambda_ADV
}
void s_sflux_tl(Run& r) { sflux_tl(r.c, r.t, r.b); }
void s_sflux_adj(Run& r) { sflux_adj(r.c, r.t, r.b); }
void s_sfc_tl(Run& r) { sfc_tl(r.c, r.t, r.b, r.oa, r.wa, r.sa); }
void s_sfc_adj(Run& r) { sfc_adj(r.c, r.t, r.b, r.oa, r.wa, r.sa); }
void s_sdiv_tl(Run& r) {
    ddx_f<double, double>(r.c, W3(r, 0), V(r, L_FX), 0, NL(r));
    ddy_f<double, double>(r.c, W3(r, 1), V(r, L_FY), 0, NL(r));
    sdiv_tl(r.c, r.b, W3(r, 0), W3(r, 1));
}
void s_sdiv_adj(Run& r) {
    ddx_f<double, double>(r.c, W3(r, 0), V(r, L_RT), 0, NL(r));
    ddy_f<double, double>(r.c, W3(r, 1), V(r, L_RT), 0, NL(r));
    sdiv_adj(r.c, r.b, W3(r, 0), W3(r, 1));
}
void s_beta_tl(Run& r) { beta_tl(r.c, r.t, r.b, r.gbeta); }
void s_beta_adj(Run& r) { beta_adj(r.c, r.t, r.b, r.gbeta); }
void s_sstep_tl(Run& r) { step_scalar_tl(r.c, r.b); }
void s_sstep_adj(Run& r) { step_scalar_adj(r.c, r.b); }

// 21-26: RHS, AB2, pressure, projection, filter (103_flosol.f90:115-290)
LinRhs rhs_vars(Run& r) {
    return {V(r, L_RX),  V(r, L_RY),  V(r, L_RZ),  V(r, L_CX),   V(r, L_CY), V(r, L_CZ), V(r, L_DVX),
            V(r, L_DVY), V(r, L_DVZ), V(r, L_BETA), V(r, L_U), V(r, L_V),  V(r, L_W)};
}
void s_rhs_tl(Run& r) { rhs_tl(r.c, rhs_vars(r), r.rf); }
void s_rhs_adj(Run& r) { rhs_adj(r.c, rhs_vars(r), r.rf); }
void s_ab2_tl(Run& r) {
    for (int m = 0; m < 3; ++m) ab2_tl(r.c, V(r, L_U + m), V(r, L_RX + m), V(r, L_RXF + m));
}
void s_ab2_adj(Run& r) {
    for (int m = 0; m < 3; ++m) ab2_adj(r.c, V(r, L_U + m), V(r, L_RX + m), V(r, L_RXF + m));
}
void s_press_tl(Run& r) {
    press_solve<double>(r.c, V(r, L_U), V(r, L_V), V(r, L_W), V(r, L_DVZ), V(r, L_P), V(r, L_DPX), V(r, L_DPY),
                        V(r, L_DPZ));
}
void s_press_adj(Run& r) { press_adj(r.c, r.b); }
void s_gradp_tl(Run& r) {
    for (int m = 0; m < 3; ++m) sub_tl(r.c, V(r, L_RX + m), V(r, L_DPX + m));
}
void s_gradp_adj(Run& r) {
    for (int m = 0; m < 3; ++m) sub_adj(r.c, V(r, L_RX + m), V(r, L_DPX + m));
}
void s_proj_tl(Run& r) { project_tl(r.c, r.b, r.sf); }
void s_proj_adj(Run& r) { project_adj(r.c, r.b, r.sf); }
void s_filt(Run& r) {
    for (int m = 0; m < 3; ++m) filter_f<double>(r.c, V(r, L_U + m), 0, NL(r));
}

struct Reg {
    int v, l0, l1;  // levels l0..l1 of linear variable v (plane variables: 0..0)
};
struct Stage {
    const char* name;
    void (*tl)(Run&);
    void (*adj)(Run&);
    std::vector<Reg> in, out;
};

std::vector<Stage> stages(int nz) {
    const int a = 1, b = nz - 1;
    auto R = [](int v, int l0, int l1) { return Reg{v, l0, l1}; };
    std::vector<Stage> s;
    s.push_back({"obukhov", s_obk_tl, s_obk_adj,
                 {R(L_U, 1, 1), R(L_V, 1, 1), R(L_PSI, 0, 0), R(L_TH, 1, 1), R(L_Q, 0, 0)},
                 {R(L_PSI, 0, 0), R(L_PHI, 0, 0), R(L_PSH, 0, 0)}});
    s.push_back({"velocity_gradients", s_deriv_tl, s_deriv_adj,
                 {R(L_U, a, b), R(L_V, a, b), R(L_W, a, b)},
                 {R(L_UX, a, b), R(L_UY, a, b), R(L_VX, a, b), R(L_VY, a, b), R(L_WX, a, b), R(L_WY, a, b),
                  R(L_WZ, a, b), R(L_UZ, 2, b), R(L_VZ, 2, b)}});
    s.push_back({"theta_filter", s_thfilt, s_thfilt, {R(L_TH, a, b)}, {R(L_TH, a, b)}});
    s.push_back({"theta_gradients", s_thgrad_tl, s_thgrad_adj, {R(L_TH, a, b)},
                 {R(L_TX, a, b), R(L_TY, a, b), R(L_TZ, 2, b)}});
    s.push_back({"theta_history", s_thhist_tl, s_thhist_adj, {R(L_RT, a, b)}, {R(L_RTF, a, b)}});
    s.push_back({"theta_pad", s_thpad_tl, s_thpad_adj, {R(L_TX, a, b), R(L_TY, a, b), R(L_TZ, 2, b)},
                 {R(L_TB1, a, b), R(L_TB2, a, b), R(L_TB3, 2, b)}});
    s.push_back({"wallstress", s_wall_tl, s_wall_adj,
                 {R(L_U, 1, 1), R(L_V, 1, 1), R(L_PSI, 0, 0), R(L_PHI, 0, 0)},
                 {R(L_TXZ, 1, 1), R(L_TYZ, 1, 1), R(L_UZ, 1, 1), R(L_VZ, 1, 1), R(L_DUS, 0, 0)}});
    s.push_back({"strain", s_strain_tl, s_strain_adj,
                 {R(L_UX, a, b), R(L_UY, a, b), R(L_UZ, a, b), R(L_VX, a, b), R(L_VY, a, b), R(L_VZ, a, b),
                  R(L_WX, a, b), R(L_WY, a, b), R(L_WZ, a, b)},
                 {R(L_S11, 1, nz), R(L_S12, 1, nz), R(L_S13, 1, nz), R(L_S22, 1, nz), R(L_S23, 1, nz),
                  R(L_S33, 1, nz)}});
    s.push_back({"nu_t_stress", s_stress_tl, s_stress_adj,
                 {R(L_S11, 1, nz), R(L_S12, 1, nz), R(L_S13, 1, nz), R(L_S22, 1, nz), R(L_S23, 1, nz),
                  R(L_S33, 1, nz)},
                 {R(L_NU, 1, nz), R(L_TXX, a, b), R(L_TXY, a, b), R(L_TYY, a, b), R(L_TZZ, a, b), R(L_TXZ, 2, b),
                  R(L_TYZ, 2, b)}});
    s.push_back({"divstress", s_div_tl, s_div_adj,
                 {R(L_TXX, a, b), R(L_TXY, a, b), R(L_TYY, a, b), R(L_TZZ, a, b), R(L_TXZ, a, b), R(L_TYZ, a, b)},
                 {R(L_DVX, a, b), R(L_DVY, a, b), R(L_DVZ, 1, nz)}});
    s.push_back({"convec_pad", s_cpad_tl, s_cpad_adj, {R(L_U, a, b), R(L_V, a, b), R(L_W, a, b)},
                 {R(L_UB, a, b), R(L_VB, a, b), R(L_WB, a, b)}});
    s.push_back({"vorticity", s_vort_tl, s_vort_adj,
                 {R(L_UY, a, b), R(L_UZ, a, b), R(L_VX, a, b), R(L_VZ, a, b), R(L_WX, a, b), R(L_WY, a, b)},
                 {R(L_O1, 1, nz), R(L_O2, 1, nz), R(L_O3, a, b)}});
    s.push_back({"vorticity_pad", s_vpad_tl, s_vpad_adj, {R(L_O1, 1, nz), R(L_O2, 1, nz), R(L_O3, a, b)},
                 {R(L_OB1, 1, nz), R(L_OB2, 1, nz), R(L_OB3, a, b)}});
    s.push_back({"cross_product", s_cross_tl, s_cross_adj,
                 {R(L_UB, a, b), R(L_VB, a, b), R(L_WB, 2, b), R(L_OB1, 1, nz), R(L_OB2, 1, nz), R(L_OB3, a, b)},
                 {R(L_CB1, a, b), R(L_CB2, a, b), R(L_CB3, 2, b)}});
    s.push_back({"convec_unpad", s_cunpad_tl, s_cunpad_adj, {R(L_CB1, a, b), R(L_CB2, a, b), R(L_CB3, 2, b)},
                 {R(L_CX, a, b), R(L_CY, a, b), R(L_CZ, 2, b)}});
    s.push_back({"scalar_advection", s_sadv_tl, s_sadv_adj,
                 {R(L_UB, a, b), R(L_VB, a, b), R(L_WB, 2, b), R(L_TB1, a, b), R(L_TB2, a, b), R(L_TB3, 2, b)},
                 {R(L_RM, a, b)}});
    s.push_back({"scalar_advection_unpad", s_sunpad_tl, s_sunpad_adj, {R(L_RM, a, b)}, {R(L_ADV, a, b)}});
    s.push_back({"scalar_sgs_flux", s_sflux_tl, s_sflux_adj,
                 {R(L_NU, 1, nz), R(L_TX, a, b), R(L_TY, a, b), R(L_TZ, 2, b)},
                 {R(L_FX, a, b), R(L_FY, a, b), R(L_FZ, 2, nz)}});
    s.push_back({"surface_flux", s_sfc_tl, s_sfc_adj, {R(L_TH, 1, 1), R(L_DUS, 0, 0), R(L_PSH, 0, 0)},
                 {R(L_Q, 0, 0), R(L_FZ, 1, 1)}});
    s.push_back({"scalar_rhs", s_sdiv_tl, s_sdiv_adj,
                 {R(L_ADV, a, b), R(L_FX, a, b), R(L_FY, a, b), R(L_FZ, 1, nz)}, {R(L_RT, a, b)}});
    s.push_back({"calc_beta", s_beta_tl, s_beta_adj, {R(L_TH, a, b)}, {R(L_BETA, 2, b)}});
    s.push_back({"step_scalar", s_sstep_tl, s_sstep_adj, {R(L_TH, a, b), R(L_RT, a, b), R(L_RTF, a, b)},
                 {R(L_TH, a, b)}});
    s.push_back({"rhs", s_rhs_tl, s_rhs_adj,
                 {R(L_CX, a, b), R(L_CY, a, b), R(L_CZ, 2, b), R(L_DVX, a, b), R(L_DVY, a, b), R(L_DVZ, a, b),
                  R(L_BETA, 2, b), R(L_U, a, b), R(L_V, a, b), R(L_W, a, b)},
                 {R(L_RX, a, b), R(L_RY, a, b), R(L_RZ, a, b)}});
    s.push_back({"ab2", s_ab2_tl, s_ab2_adj,
                 {R(L_U, a, b), R(L_V, a, b), R(L_W, a, b), R(L_RX, a, b), R(L_RY, a, b), R(L_RZ, a, b),
                  R(L_RXF, a, b), R(L_RYF, a, b), R(L_RZF, a, b)},
                 {R(L_U, a, b), R(L_V, a, b), R(L_W, a, b)}});
    s.push_back({"press_stag_array", s_press_tl, s_press_adj,
                 {R(L_U, a, b), R(L_V, a, b), R(L_W, a, b), R(L_DVZ, 1, nz)},
                 {R(L_P, 0, b), R(L_DPX, a, b), R(L_DPY, a, b), R(L_DPZ, a, b)}});
    s.push_back({"rhs_gradp", s_gradp_tl, s_gradp_adj,
                 {R(L_RX, a, b), R(L_RY, a, b), R(L_RZ, a, b), R(L_DPX, a, b), R(L_DPY, a, b), R(L_DPZ, a, b)},
                 {R(L_RX, a, b), R(L_RY, a, b), R(L_RZ, a, b)}});
    s.push_back({"project", s_proj_tl, s_proj_adj,
                 {R(L_U, a, b), R(L_V, a, b), R(L_W, a, b), R(L_DPX, a, b), R(L_DPY, a, b), R(L_DPZ, a, b)},
                 {R(L_U, a, b), R(L_V, a, b), R(L_W, a, b)}});
    s.push_back({"filter", s_filt, s_filt, {R(L_U, a, b), R(L_V, a, b), R(L_W, a, b)},
                 {R(L_U, a, b), R(L_V, a, b), R(L_W, a, b)}});
    return s;
}

std::vector<Reg> state_regs(int nz) {
    std::vector<Reg> r;
    for (int v = L_U; v <= L_RT; ++v) r.push_back({v, 1, nz - 1});
    r.push_back({L_PSI, 0, 0});
    r.push_back({L_Q, 0, 0});
    return r;
}

Run make_run(les_ctx* c, const Tape& t, double wobj, double* dus2) {
    Run r{c, c->lin->b, t, wobj, dus2, c->lin->oa, wall_args(c), rhs_flags(c), 0.0, 0, sfc_args(c)};
    const les_params& P = c->prm;
    r.gbeta = GRAV * (P.z_i / (P.u_scale * P.u_scale));
    r.sf = P.ubc_mom == LES_BC_STRESS_FREE;
    return r;
}

// tangent step: lin <- M lin (state variables), intermediates overwritten
void tlm_step(les_ctx* c, const Tape& t, double wobj) {
    LinBufs& b = c->lin->b;
    for (int m = 0; m < 3; ++m) std::swap(b.v[L_RX + m], b.v[L_RXF + m]);
    const bool us2 = c->lin_obj == LES_OBJ_USTAR2;
    Run r = make_run(c, t, wobj, (wobj != 0.0 && us2) ? b.us2 : nullptr);
    const std::vector<Stage> st = stages(c->nz);
    for (const Stage& s : st) {
        s.tl(r);
        lin_flush(c, b);
    }
    if (wobj != 0.0) {
        lin_dot_part(c, us2 ? b.us2 : b.v[L_Q], nullptr, (long)c->plane, 0, 0, 0, 2, b.part);
        lin_acc(c, b.part, wobj, b.acc);
    }
}

// adjoint step: lin <- M^T lin (+ the objective's direct term at this step)
void adj_step(les_ctx* c, const Tape& t, double wobj) {
    LinBufs& b = c->lin->b;
    Run r = make_run(c, t, wobj, nullptr);
    const std::vector<Stage> st = stages(c->nz);
    // the surface flux objective reads this step's output sgs_t3(:,:,1)
    if (wobj != 0.0 && c->lin_obj == LES_OBJ_T_FLUX) lin_addc(c, b.v[L_Q], wobj, (long)c->plane);
    for (int i = (int)st.size() - 1; i >= 0; --i) {
        st[i].adj(r);
        lin_flush(c, b);
    }
    for (int m = 0; m < 3; ++m) std::swap(b.v[L_RX + m], b.v[L_RXF + m]);
}

// zero the levels of the state variables outside 1..nz-1
void clear_state_ghosts(les_ctx* c) {
    LinBufs& b = c->lin->b;
    for (int v = L_U; v <= L_RT; ++v) {
        clear_levels(c, b, b.v[v], 0, 0);
        clear_levels(c, b, b.v[v], c->nz, c->nz);
    }
    lin_flush(c, b);
}
void clear_intermediates(les_ctx* c) {
    LinBufs& b = c->lin->b;
    for (int v = L_NSTATE; v < L_COUNT; ++v) fill_levels(c->stream, b.v[v], lv_size(c, v), 0, 1, 0.0);
}

// the objective: J += wobj * sum over the plane of ustar^2, or of the
// surface flux sgs_t3(:,:,1)
void obj_accumulate(les_ctx* c, double wobj) {
    LinBufs& b = c->lin->b;
    if (c->lin_obj == LES_OBJ_T_FLUX)
        lin_dot_part(c, c->f3[LES_F_SGS_T3] + c->plane, nullptr, (long)c->plane, 0, 0, 0, 2, b.part);
    else
        lin_dot_part(c, c->ustar, nullptr, (long)c->plane, 0, 0, 0, 1, b.part);
    lin_acc(c, b.part, wobj, b.acc + 1);
}

int lin_check_supported(const les_ctx* c, int jt0, int nsteps) {
    const les_params& P = c->prm;
    if (c->precision != LES_FP64) return LES_E_BADARG;
    if (!P.theta_flag || jt0 <= P.theta_init_time || jt0 < 2 || nsteps < 0) return LES_E_BADARG;
    if (P.lbc_mom != LES_BC_WALL || P.ubc_mom != LES_BC_STRESS_FREE) return LES_E_BADARG;
    // linearised: the land step (prescribed surface flux or temperature) with damping method 1
    if (P.ocean_flag || (P.ubc == 1 && P.damping_method != 1)) return LES_E_BADARG;
    return LES_OK;
}

int lin_var(int id) {
    switch (id) {
        case LES_F_U: return L_U;
        case LES_F_V: return L_V;
        case LES_F_W: return L_W;
        case LES_F_THETA: return L_TH;
        case LES_F_RHSX: return L_RX;
        case LES_F_RHSY: return L_RY;
        case LES_F_RHSZ: return L_RZ;
        case LES_F_RHST: return L_RT;
        case LES_F_PSI_M_SFC: return L_PSI;
        case LES_F_SGS_T3: return L_Q;
        default: return -1;
    }
}
// the level a one-plane linear variable is addressed at (sgs_t3: level 1)
int lin_plane_lev(int id) { return id == LES_F_SGS_T3 ? 1 : 0; }

// random values on a region of variable v (zeros elsewhere in the variable
// when fresh), or the inner product of the region with the same values
void region_bounds(const les_ctx* c, const Reg& g, long* off, long* n) {
    const long pl = (long)lv_plane(c, g.v);
    *off = (long)g.l0 * pl;
    *n = (long)(g.l1 - g.l0 + 1) * pl;
}
void fill_region(les_ctx* c, const Reg& g, unsigned long long seed) {
    long off, n;
    region_bounds(c, g, &off, &n);
    lin_rand(c, c->lin->b.v[g.v] + off, n, seed, (unsigned long long)(g.v + 1), off);
}
double dot_region(les_ctx* c, const Reg& g, unsigned long long seed) {
    long off, n;
    region_bounds(c, g, &off, &n);
    LinBufs& b = c->lin->b;
    fill_levels(c->stream, b.acc + 2, 1, 0, 1, 0.0);
    lin_dot_part(c, b.v[g.v] + off, nullptr, n, seed, (unsigned long long)(g.v + 1), off, 0, b.part);
    lin_acc(c, b.part, 1.0, b.acc + 2);
    double h = 0.0;
    les_note(c, hipMemcpyAsync(&h, b.acc + 2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    les_note(c, hipStreamSynchronize(c->stream));
    return h;
}

void clear_lin(les_ctx* c) {
    LinBufs& b = c->lin->b;
    for (int v = 0; v < L_COUNT; ++v) fill_levels(c->stream, b.v[v], lv_size(c, v), 0, 1, 0.0);
}

}  // namespace

void lin_destroy(les_ctx* c) {
    delete c->lin;  // device memory is released with the context's allocations
    c->lin = nullptr;
}

void tape_save(les_ctx* c, Tape* t, int point) {
    const size_t F3 = c->plane * c->nlev, np = c->plane;
    switch (point) {
        case TAPE_OBUKHOV:
            dcopy(c, t->obk, c->red, 4);
            dcopy(c, t->u1, c->f3[LES_F_U] + np, np);
            dcopy(c, t->v1, c->f3[LES_F_V] + np, np);
            break;
        case TAPE_WALL:
            dcopy(c, t->uv1f, c->work3[7], 2 * np);
            break;
        case TAPE_THETA:
            dcopy(c, t->thf, c->f3[LES_F_THETA], F3);
            dcopy(c, t->tmean, c->th_mean, c->nlev);
            break;
        default: {
            // Work fields and padded scratch the next step rewrites before
            // reading: the tape takes the buffers the step wrote and hands
            // its own to the context (FP64 mode: a work field's wk is its f3)
            auto take = [&](int id, double*& tp) {
                std::swap(c->f3[id], tp);
                c->wk[id] = c->f3[id];
            };
            for (int m = 0; m < 6; ++m) take(LES_F_S11 + m, t->S[m]);
            take(LES_F_NU_T, t->nu);
            take(LES_F_DTDX, t->dT[0]);
            take(LES_F_DTDY, t->dT[1]);
            take(LES_F_DTDZ, t->dT[2]);
            std::swap(c->u_big, t->ub);
            std::swap(c->v_big, t->vb);
            std::swap(c->w_big, t->wb);
            for (int m = 0; m < 3; ++m) std::swap(c->big3[m], t->ob[m]);
            for (int m = 0; m < 3; ++m) std::swap(c->tbig[m], t->tb[m]);
            dcopy(c, t->cs, c->f3[LES_F_CS_OPT2], F3);   // state: copied
            break;
        }
    }
}

}  // namespace les

using namespace les;

// ============================================================================
// C-ABI
// ============================================================================
extern "C" {

int les_set_sgs_frozen(les_ctx* c, int frozen) {
    if (!c) return LES_E_BADARG;
    c->sgs_frozen = frozen ? 1 : 0;
    return LES_OK;
}

int les_lin_set(les_ctx* c, int id, const double* h, int ld, int lev0, int nlev) {
    const int v = lin_var(id);
    if (!c || !h || v < 0 || ld < c->nx) return LES_E_BADARG;
    const bool pl = lin_plane_only(v);
    if (pl ? (lev0 != lin_plane_lev(id) || nlev != 1) : (lev0 < 0 || nlev < 0 || lev0 + nlev > (int)c->nlev))
        return LES_E_BADARG;
    if (lin_alloc(c) != hipSuccess) return LES_E_NOMEM;
    double* d = c->lin->b.v[v] + (pl ? 0 : (size_t)lev0 * c->plane);
    les_note(c, hipMemcpy2DAsync(d, (size_t)c->nx * 8, h, (size_t)ld * 8, (size_t)c->nx * 8, (size_t)nlev * c->ny,
                                 hipMemcpyHostToDevice, c->stream));
    return finish(c);
}

int les_lin_get(les_ctx* c, int id, double* h, int ld, int lev0, int nlev) {
    const int v = lin_var(id);
    if (!c || !h || v < 0 || ld < c->nx) return LES_E_BADARG;
    const bool pl = lin_plane_only(v);
    if (pl ? (lev0 != lin_plane_lev(id) || nlev != 1) : (lev0 < 0 || nlev < 0 || lev0 + nlev > (int)c->nlev))
        return LES_E_BADARG;
    if (lin_alloc(c) != hipSuccess) return LES_E_NOMEM;
    const double* d = c->lin->b.v[v] + (pl ? 0 : (size_t)lev0 * c->plane);
    les_note(c, hipMemcpy2DAsync(h, (size_t)ld * 8, d, (size_t)c->nx * 8, (size_t)c->nx * 8, (size_t)nlev * c->ny,
                                 hipMemcpyDeviceToHost, c->stream));
    return finish(c);
}

int les_lin_set_objective(les_ctx* c, int kind) {
    if (!c || (kind != LES_OBJ_USTAR2 && kind != LES_OBJ_T_FLUX)) return LES_E_BADARG;
    c->lin_obj = kind;
    return LES_OK;
}

int les_lin_zero(les_ctx* c) {
    if (!c) return LES_E_BADARG;
    if (lin_alloc(c) != hipSuccess) return LES_E_NOMEM;
    clear_lin(c);
    return finish(c);
}

// tangent-linear window: state advances nsteps from jt0; lin (tangent of
// the state at jt0) becomes the tangent at jt0 + nsteps.  dJ: tangent of
// the objective (mean over the plane and the last obj_steps steps of ustar^2
// or of the surface flux, les_lin_set_objective)
int les_tlm_window(les_ctx* c, int jt0, int nsteps, int obj_steps, double* dJ) {
    if (!c || obj_steps < 0 || obj_steps > nsteps) return LES_E_BADARG;
    int st = lin_check_supported(c, jt0, nsteps);
    if (st != LES_OK) return st;
    if (lin_alloc(c) != hipSuccess || ensure_tapes(c, 1) != hipSuccess) return LES_E_NOMEM;
    LinCtx* L = c->lin;
    clear_intermediates(c);
    fill_levels(c->stream, L->b.acc, 8, 0, 1, 0.0);
    const double w = obj_steps > 0 ? 1.0 / ((double)obj_steps * (double)(c->nx * c->ny)) : 0.0;
    for (int i = 0; i < nsteps; ++i) {
        const double wobj = (i >= nsteps - obj_steps) ? w : 0.0;
        step_fp64(c, jt0 + i, &L->tapes[0], false);
        clear_state_ghosts(c);
        tlm_step(c, L->tapes[0], wobj);
    }
    double acc = 0.0;
    les_note(c, hipMemcpyAsync(&acc, L->b.acc, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    st = finish(c);
    if (dJ) *dJ = acc;
    return st;
}

int les_step_tlm(les_ctx* c, int jt) { return les_tlm_window(c, jt, 1, 0, nullptr); }

// adjoint window: the context holds the state at jt0 and, in lin, the
// adjoint of the state at jt0 + nsteps (zero for a pure objective).  On
// return the state is at jt0 + nsteps, lin holds the gradient with respect
// to the state at jt0 of <lin seed, x(jt0 + nsteps)> + J, and *J the
// objective (mean over the plane and the last obj_steps steps of ustar^2 or
// of the surface flux).  Checkpoints every k_ckpt steps; the last segment is taped during the
// forward sweep, every earlier one recomputed once with tapes.  k_ckpt = 0
// picks the spacing from free device memory: as many tapes as fit with a
// sixteenth of the device left free, up to the whole window (then nothing
// is recomputed).
int les_adjoint_window(les_ctx* c, int jt0, int nsteps, int k_ckpt, int obj_steps, double* J) {
    if (!c || k_ckpt < 0 || nsteps < 1 || obj_steps < 0 || obj_steps > nsteps) return LES_E_BADARG;
    int st = lin_check_supported(c, jt0, nsteps);
    if (st != LES_OK) return st;
    if (k_ckpt == 0) k_ckpt = auto_ckpt(c, nsteps);
    const int nseg = (nsteps + k_ckpt - 1) / k_ckpt;
    const int kt = nsteps < k_ckpt ? nsteps : k_ckpt;
    if (lin_alloc(c) != hipSuccess || ensure_tapes(c, kt > 0 ? kt : 1) != hipSuccess ||
        ensure_ckpts(c, nseg + 1) != hipSuccess)
        return LES_E_NOMEM;
    LinCtx* L = c->lin;
    LinBufs& b = L->b;
    for (int v = 0; v < L_NSTATE; ++v) dcopy(c, L->seed[v], b.v[v], lv_size(c, v));
    fill_levels(c->stream, b.acc, 8, 0, 1, 0.0);
    const double w = obj_steps > 0 ? 1.0 / ((double)obj_steps * (double)(c->nx * c->ny)) : 0.0;
    auto wobj_of = [&](int i) { return (i >= nsteps - obj_steps) ? w : 0.0; };
    // forward sweep with checkpoints; the last segment is taped on the way
    const int ilast = (nseg - 1) * k_ckpt;
    for (int i = 0; i < nsteps; ++i) {
        if (i % k_ckpt == 0) ckpt_save(c, L->ckpts[i / k_ckpt]);
        step_fp64(c, jt0 + i, i >= ilast ? &L->tapes[i - ilast] : nullptr, false);
        if (wobj_of(i) != 0.0) obj_accumulate(c, wobj_of(i));
    }
    ckpt_save(c, L->ckpts[nseg]);
    // reverse sweep
    clear_intermediates(c);
    for (int v = 0; v < L_NSTATE; ++v) dcopy(c, b.v[v], L->seed[v], lv_size(c, v));
    clear_state_ghosts(c);
    for (int sgi = nseg - 1; sgi >= 0; --sgi) {
        const int i0 = sgi * k_ckpt;
        const int m = (nsteps - i0) < k_ckpt ? (nsteps - i0) : k_ckpt;
        if (sgi != nseg - 1) {
            ckpt_load(c, L->ckpts[sgi]);
            for (int j = 0; j < m; ++j) step_fp64(c, jt0 + i0 + j, &L->tapes[j], false);
        }
        for (int j = m - 1; j >= 0; --j) {
            adj_step(c, L->tapes[j], wobj_of(i0 + j));
            clear_state_ghosts(c);
        }
    }
    ckpt_load(c, L->ckpts[nseg]);
    double acc = 0.0;
    les_note(c, hipMemcpyAsync(&acc, b.acc + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    st = finish(c);
    if (J) *J = acc;
    return st;
}

int les_step_adj(les_ctx* c, int jt) { return les_adjoint_window(c, jt, 1, 1, 0, nullptr); }

// Device time of each stage's tangent and adjoint (ms, HIP events on the
// context's stream) about forward step jt from the context's state (which
// advances one step); arrays of les_lin_nstages() entries.
int les_lin_stage_times(les_ctx* c, int jt, double* tl_ms, double* adj_ms) {
    if (!c || !tl_ms || !adj_ms) return LES_E_BADARG;
    int st = lin_check_supported(c, jt, 1);
    if (st != LES_OK) return st;
    if (lin_alloc(c) != hipSuccess || ensure_tapes(c, 1) != hipSuccess) return LES_E_NOMEM;
    const std::vector<Stage> all = stages(c->nz);
    const int n = (int)all.size();
    std::vector<hipEvent_t> ev(2 * (n + 1));
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) return LES_E_HIP;
    LinCtx* L = c->lin;
    step_fp64(c, jt, &L->tapes[0], false);
    clear_lin(c);
    Run r = make_run(c, L->tapes[0], 0.0, nullptr);
    les_note(c, hipEventRecord(ev[0], c->stream));
    for (int i = 0; i < n; ++i) {
        all[i].tl(r);
        lin_flush(c, L->b);
        les_note(c, hipEventRecord(ev[i + 1], c->stream));
    }
    les_note(c, hipEventRecord(ev[n + 1], c->stream));
    for (int i = n - 1; i >= 0; --i) {
        all[i].adj(r);
        lin_flush(c, L->b);
        les_note(c, hipEventRecord(ev[n + 1 + (n - i)], c->stream));
    }
    st = finish(c);
    for (int i = 0; i < n && st == LES_OK; ++i) {
        float a = 0.f, b = 0.f;
        if (hipEventElapsedTime(&a, ev[i], ev[i + 1]) != hipSuccess) st = LES_E_HIP;
        if (hipEventElapsedTime(&b, ev[n + 1 + (n - 1 - i)], ev[n + 1 + (n - i)]) != hipSuccess) st = LES_E_HIP;
        tl_ms[i] = a;
        adj_ms[i] = b;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    clear_lin(c);
    const int st2 = finish(c);
    return st != LES_OK ? st : st2;
}

int les_lin_nstages(void) { return (int)stages(8).size(); }

const char* les_lin_stage_name(int stage) {
    static const std::vector<Stage> s = stages(8);
    if (stage < 0 || stage >= (int)s.size()) return nullptr;
    return s[stage].name;
}

// Dot-product test of one stage (stage >= 0) or of the whole step (-1)
// about the forward step jt from the context's state (which advances one
// step): lhs = <M dx, y>, rhs = <dx, M^T y> for counter-based random dx, y.
int les_lin_check(les_ctx* c, int stage, int jt, unsigned long long seed, double* lhs, double* rhs) {
    if (!c || !lhs || !rhs) return LES_E_BADARG;
    int st = lin_check_supported(c, jt, 1);
    if (st != LES_OK) return st;
    const std::vector<Stage> all = stages(c->nz);
    if (stage < -1 || stage >= (int)all.size()) return LES_E_BADARG;
    if (lin_alloc(c) != hipSuccess || ensure_tapes(c, 1) != hipSuccess) return LES_E_NOMEM;
    LinCtx* L = c->lin;
    const Tape& t = L->tapes[0];
    step_fp64(c, jt, &L->tapes[0], false);
    const std::vector<Reg> in = stage < 0 ? state_regs(c->nz) : all[stage].in;
    const std::vector<Reg> out = stage < 0 ? state_regs(c->nz) : all[stage].out;
    const unsigned long long sx = seed * 2 + 1, sy = seed * 2 + 2;
    // tangent
    clear_lin(c);
    for (const Reg& g : in) fill_region(c, g, sx);
    if (stage < 0) {
        tlm_step(c, t, 0.0);
    } else {
        Run r = make_run(c, t, 0.0, nullptr);
        all[stage].tl(r);
        lin_flush(c, L->b);
    }
    double l = 0.0;
    for (const Reg& g : out) l += dot_region(c, g, sy);
    // adjoint
    clear_lin(c);
    for (const Reg& g : out) fill_region(c, g, sy);
    if (stage < 0) {
        adj_step(c, t, 0.0);
    } else {
        Run r = make_run(c, t, 0.0, nullptr);
        all[stage].adj(r);
        lin_flush(c, L->b);
    }
    double rr = 0.0;
    for (const Reg& g : in) rr += dot_region(c, g, sx);
    clear_lin(c);
    st = finish(c);
    *lhs = l;
    *rhs = rr;
    return st;
}

}  // extern "C"
