// C-ABI of the MI355X boundary-layer LES step (include/les_gpu.h): context
// lifecycle, field transfer, per-routine entry points and the flosol
// orchestration (103_flosol.f90:4-290).  All work is enqueued on one HIP
// stream per context; entry points synchronise before returning.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <utility>
#include <vector>
#include "les_lag.hpp"
#include "les_ops.hpp"

using namespace les;

namespace les {

// press_stag_array (109_press_stag_array.f90:104-238) on explicit fields:
// (u, v, w, divtz(1), divtz(nz)) -> p (levels 0..nz-1), dp/dx, dp/dy, dp/dz
// (levels 1..nz-1).  Linear in its inputs; the linearised step calls it on
// tangent fields.
template <typename W>
void press_solve(les_ctx* c, const double* u, const double* v, const double* w, const W* divtz, W* pf, W* dpdx,
                 W* dpdy, W* dpdz, double modi) {
    const int nz = c->nz;
    const size_t np = c->plane, ns = c->ns;
    W *rx = scr<W>(c->work3[0]), *ry = scr<W>(c->work3[1]), *rz = scr<W>(c->work3[2]);
    press_rhs_prep_on<W>(c, rx, ry, rz, u, v, w, divtz, modi);
    cx<W>*Hx = scr<W>(c->spec3[0]), *Hy = scr<W>(c->spec3[1]), *Hz = scr<W>(c->spec3[2]);
    const long rows = (long)(nz + 1) * c->ny;
    x_r2c<W>(c->stream, rx, Hx, rows, c->nx, c->lhx, c->tw_x, 1.0);
    x_r2c<W>(c->stream, ry, Hy, rows, c->nx, c->lhx, c->tw_x, 1.0);
    x_r2c<W>(c->stream, rz, Hz, rows, c->nx, c->lhx, c->tw_x, 1.0);
    yline_spec<W>(c->stream, Hx, Hx, nz + 1, c->lhx, (long)ns, c->tw_y, SP_FWD, nullptr);
    yline_spec<W>(c->stream, Hy, Hy, nz + 1, c->lhx, (long)ns, c->tw_y, SP_FWD, nullptr);
    yline_spec<W>(c->stream, Hz, Hz, nz + 1, c->lhx, (long)ns, c->tw_y, SP_FWD, nullptr);
    cx<W>* ph;
    const double* hzm = nullptr;
    if (sizeof(W) == sizeof(double)) {
        ph = scr<W>(c->spec3[3]);
    } else {
        ph = reinterpret_cast<cx<W>*>(c->phf);
        press_hz_sums_dev<W>(c, c->hzm, modi);
        hzm = c->hzm;
    }
    tdma_press<W>(c, Hx, Hy, Hz, ph, reinterpret_cast<double*>(c->spec3[3]), hzm);
    yline_spec<W>(c->stream, ph, ph, nz, c->lhx, (long)ns, c->tw_y, SP_INV, nullptr);
    x_c2r<W>(c->stream, ph, pf, (long)nz * c->ny, c->lhx, c->nx, c->tw_x, 0, 0.0);
    x_c2r<W>(c->stream, ph + ns, dpdx + np, (long)(nz - 1) * c->ny, c->lhx, c->nx, c->tw_x, 1, kfx(c));
    ddy_f<W, W>(c, dpdy, pf, 1, nz - 1);
    dpdz_on<W>(c, dpdz, pf);
}
template void press_solve<double>(les_ctx*, const double*, const double*, const double*, const double*, double*,
                                  double*, double*, double*, double);
template void press_solve<float>(les_ctx*, const double*, const double*, const double*, const float*, float*, float*,
                                 float*, float*, double);

// lbc = 0's surface-temperature term of the pressure's bottom row on step jt
// (109_press_stag_array.f90:186-191): g z_i/u_scale^2 (T_s - <T_s>) for the
// uniform surface temperature of the patches, <.> a sum in the reference's order
double press_modi(const les_ctx* c, int jt) {
    const les_params& P = c->prm;
    if (!(P.theta_flag && jt > P.theta_init_time && P.lbc == 0)) return 0.0;
    double st = 0.0;
    for (size_t q = 0; q < c->plane; ++q) st += P.T_s;
    return GRAV * (P.z_i / (P.u_scale * P.u_scale)) * (P.T_s - st / (double)(c->nx * c->ny));
}

}  // namespace les

namespace {

int good_size(int n) {
    if (n < 8 || n % 4) return 0;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    return n == 1;
}

template <typename T>
hipError_t dalloc(les_ctx* c, T** p, size_t count) {
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, sizeof(T) * (count ? count : 1));
    if (e != hipSuccess) return e;
    c->allocs.push_back(v);
    *p = static_cast<T*>(v);
    return hipSuccess;
}

hipError_t make_twiddle(les_ctx* c, Twiddle& t, int n) {
    t.n = n;
    t.plan = make_plan(n);
    std::vector<cx<double>> hd(n);
    std::vector<cx<float>> hf(n);
    for (int m = 0; m < n; ++m) {
        long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
        hd[m] = {(double)cosl(a), -(double)sinl(a)};
        hf[m] = {(float)cosl(a), -(float)sinl(a)};
    }
    hipError_t e = dalloc(c, &t.d, n);
    if (e != hipSuccess) return e;
    e = dalloc(c, &t.f, n);
    if (e != hipSuccess) return e;
    e = hipMemcpy(t.d, hd.data(), sizeof(cx<double>) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    return hipMemcpy(t.f, hf.data(), sizeof(cx<float>) * n, hipMemcpyHostToDevice);
}

// Wavenumbers, Nyquist mask and cutoff filters: init_wavenumber
// (002_fft.f90:121-189) and init_test_filter (006_test_filter.f90:41-77).
hipError_t init_spectral(les_ctx* c) {
    const int lhx = c->lhx, ny = c->ny;
    const size_t ns = c->ns;
    const double lx = c->prm.lx_tot, ly = c->prm.ly_tot;
    std::vector<double> kx2d(ns), ky2d(ns), k2(ns), G2(ns), G4(ns);
    std::vector<unsigned char> nn(ns, 1);
    for (int j = 0; j < ny; ++j) {
        const int m = ((j + ny / 2) % ny) - ny / 2;
        for (int i = 0; i < lhx; ++i) {
            const size_t q = (size_t)j * lhx + i;
            kx2d[q] = (i == lhx - 1) ? 0.0 : (double)i * (2.0 * PI / lx);
            ky2d[q] = (double)m * (2.0 * PI / ly);
        }
    }
    for (int j = 0; j < ny; ++j) {
        const size_t q = (size_t)j * lhx + lhx - 1;
        ky2d[q] = 0.0;
        nn[q] = 0;
    }
    for (int i = 0; i < lhx; ++i) {
        const size_t q = (size_t)(ny / 2) * lhx + i;
        kx2d[q] = 0.0;
        ky2d[q] = 0.0;
        nn[q] = 0;
    }
    for (size_t q = 0; q < ns; ++q) k2[q] = kx2d[q] * kx2d[q] + ky2d[q] * ky2d[q];
    const double alphas[2] = {2.0, 4.0};
    std::vector<double>* Gs[2] = {&G2, &G4};
    for (int a = 0; a < 2; ++a) {
        const double delta = alphas[a] * sqrt(c->dx * c->dy);
        const double kc2 = (PI / delta) * (PI / delta);
        for (size_t q = 0; q < ns; ++q) (*Gs[a])[q] = (k2[q] >= kc2) ? 0.0 : 1.0 / (double)(c->nx * c->ny);
    }
    // filter support for the pruned spectra of lagrange_Sdep
    int kc[2] = {1, 1}, mc = 1;
    for (int j = 0; j < ny; ++j) {
        if (2 * j == ny) continue;
        for (int i = 0; i + 1 < lhx; ++i) {
            const size_t q = (size_t)j * lhx + i;
            for (int a = 0; a < 2; ++a) {
                if ((*Gs[a])[q] == 0.0) continue;
                kc[a] = std::max(kc[a], i + 1);
                mc = std::max(mc, std::min(j, ny - j) + 1);
            }
        }
    }
    c->lag_kc = std::max(kc[0], kc[1]);
    c->lag_kc4 = kc[1];
    c->lag_mc = mc;
    // TDMA coefficients with the wall row at level 1 and the top row at
    // nzt-1, factorised (109_press_stag_array.f90:4-103)
    const int nz = c->nz;
    const double dz = c->dz;
    std::vector<double> cb(ns * nz, 0.0), cc(ns * nz, 0.0);
    for (size_t q = 0; q < ns; ++q) {
        if (!nn[q]) continue;
        double cprev = 0.0;
        for (int k = 1; k <= nz - 1; ++k) {
            double a = 0.0, b = 0.0, cv = 0.0;
            if (k2[q] > 0) {
                if (k == 1) { a = 0.0; b = -(k2[q] + 1.0 / (dz * dz)); cv = 1.0 / (dz * dz); }
                else if (k == nz - 1) { a = 1.0 / (dz * dz); b = -(k2[q] + 1.0 / (dz * dz)); cv = 0.0; }
                else { a = 1.0 / (dz * dz); b = -(k2[q] + 2.0 / (dz * dz)); cv = 1.0 / (dz * dz); }
            } else {
                if (k == 1) { a = 0.0; b = 1.0 / dz; }
                else { a = -1.0 / dz; b = 1.0 / dz; }
            }
            const double bp = b - a * cprev;
            const double cp = cv / bp;
            cb[(size_t)k * ns + q] = bp;
            cc[(size_t)k * ns + q] = cp;
            cprev = cp;
        }
    }
    hipError_t e;
    if ((e = dalloc(c, &c->kx2d, ns)) != hipSuccess) return e;
    if ((e = dalloc(c, &c->ky2d, ns)) != hipSuccess) return e;
    if ((e = dalloc(c, &c->k2, ns)) != hipSuccess) return e;
    if ((e = dalloc(c, &c->Gtest, ns)) != hipSuccess) return e;
    if ((e = dalloc(c, &c->Gtest2, ns)) != hipSuccess) return e;
    if ((e = dalloc(c, &c->notnyq, ns)) != hipSuccess) return e;
    if ((e = dalloc(c, &c->coef_b, ns * nz)) != hipSuccess) return e;
    if ((e = dalloc(c, &c->coef_c, ns * nz)) != hipSuccess) return e;
    const std::pair<void*, const void*> up[] = {{c->kx2d, kx2d.data()}, {c->ky2d, ky2d.data()}, {c->k2, k2.data()},
                                                {c->Gtest, G2.data()}, {c->Gtest2, G4.data()}};
    for (const auto& pr : up)
        if ((e = hipMemcpy(pr.first, pr.second, ns * 8, hipMemcpyHostToDevice)) != hipSuccess) return e;
    if ((e = hipMemcpy(c->notnyq, nn.data(), ns, hipMemcpyHostToDevice)) != hipSuccess) return e;
    if ((e = hipMemcpy(c->coef_b, cb.data(), ns * nz * 8, hipMemcpyHostToDevice)) != hipSuccess) return e;
    return hipMemcpy(c->coef_c, cc.data(), ns * nz * 8, hipMemcpyHostToDevice);
}

// setsponge (007_topbc.f90:15-85) at the global level index k, with the uvp
// height zuvp(k) = (k - 1/2) dz (init_meshgrid, 102_initial.f90:275-278).
// Method 2 zeroes sponge(nz) of every z-slab, a per-slab value; in one domain
// only the top slab's remains (level nzt).
std::vector<double> sponge_profile(const les_ctx* c) {
    const les_params& P = c->prm;
    std::vector<double> sp(c->nlev, 0.0);
    if (P.ubc != 1) return sp;
    const int nzt = c->nz;
    if (P.damping_method == 1) {                                     // :26-36
        const double factor = 9.0 / (double)(nzt - 3 * nzt / 4 + 1);
        const double top = P.z_i / (50.0 * P.u_scale);
        for (int k = 1; k <= nzt - 1; ++k)
            if (k > 3 * nzt / 4 + 1) sp[k] = top * pow(5.0, (double)(k - nzt) * factor);
    } else if (P.damping_method == 2) {                              // :37-50
        const double z_d = 0.75 * P.lz_tot, cfrdmp = 3.9;
        for (int k = 1; k <= nzt - 1; ++k) {
            const double z = ((double)k - 0.5) * c->dz;
            if (z >= z_d && z <= P.lz_tot) sp[k] = cfrdmp * 0.5 * (1.0 - cos(PI * (z - z_d) / (P.lz_tot - z_d)));
        }
    } else {                                                         // :55-82
        const double g_hat = GRAV * (P.z_i / (P.u_scale * P.u_scale));
        // fac and 8.0*atan(1.0) are default-real (single precision) in the reference
        const double fac = P.ocean_flag ? (double)0.22f : 1.5;
        const double bv = P.ocean_flag ? sqrt(g_hat * P.alpha_w * P.inv_strength / P.T_scale * P.z_i)
                                       : sqrt(g_hat * P.inv_strength / P.T_scale * P.z_i);
        const double sigma_z = 0.8 * P.lz_tot;
        const double sigma_max = fac * bv / (double)(8.0f * atanf(1.0f));
        for (int k = 1; k <= nzt; ++k) {
            const double z = ((double)k - 0.5) * c->dz;
            if (z >= sigma_z && z <= P.lz_tot) {
                const double ell = (z - sigma_z) / (P.lz_tot - sigma_z);
                sp[k] = ell * ell * sigma_max;
            }
        }
    }
    return sp;
}

// init_stokes_drift (014_stokes_drift.f90:29-115): the monochromatic wave's
// drift, exp(-2 k z) under a stress-free top and the finite-depth form under
// a wall top (mono_shallow_water); 'dynamic' is the monochromatic wave of
// gwave.in's first record, whose values the caller passes as amp_w, lambda_w
// and agl_ust.  False for the Donelan-Pierson spectrum (not provided).
bool stokes_profile(const les_ctx* c, std::vector<double>& ust, std::vector<double>& vst) {
    const les_params& P = c->prm;
    ust.assign(c->nlev, 0.0);
    vst.assign(c->nlev, 0.0);
    if (!P.stokes_flag) return true;
    if (P.stokes_action != LES_STOKES_MONO && P.stokes_action != LES_STOKES_DYNAMIC) return false;
    const double rad = P.agl_ust / 180.0 * PI;
    const double amp = P.amp_w / P.z_i, lam = P.lambda_w / P.z_i;
    const double wn = 2.0 * PI / lam;
    const double gnd = GRAV * P.z_i / (P.u_scale * P.u_scale);
    if (P.ubc_mom == LES_BC_STRESS_FREE) {                           // monochromatic, :146-160
        const double omega = sqrt(gnd * wn);
        const double us = omega * wn * amp * amp;
        for (int k = 0; k <= c->nz; ++k) {                           // :59-63
            const double z = ((double)k - 0.5) * c->dz;
            ust[k] = us * exp(-2.0 * wn * z) * cos(rad);
            vst[k] = us * exp(-2.0 * wn * z) * sin(rad);
        }
    } else {                                                         // mono_shallow_water, :164-178
        const double H = P.lz_tot;
        const double omega = sqrt(gnd * wn * tanh(wn * H));
        const double us = omega * wn * amp * amp / tanh(wn * H);
        for (int k = 0; k <= c->nz; ++k) {                           // :76-81
            const double z = ((double)k - 0.5) * c->dz;
            const double sh = sinh(wn * H);
            ust[k] = us * cosh(2.0 * wn * (-z + H)) / 2.0 / (sh * sh) * cos(rad);
            vst[k] = us * cosh(2.0 * wn * (-z + H)) / 2.0 / (sh * sh) * sin(rad);
        }
    }
    return true;
}

// The momentum RHS history and the Lagrangian velocity sums u_lag, v_lag,
// w_lag: state stored in the precision of the mode (FP32 in LES_MIXED,
// converted when the mode changes).
bool is_work_state(int id) {
    return id == LES_F_RHSX || id == LES_F_RHSY || id == LES_F_RHSZ || id == LES_F_RHSX_F || id == LES_F_RHSY_F ||
           id == LES_F_RHSZ_F || id == LES_F_U_LAG || id == LES_F_V_LAG || id == LES_F_W_LAG;
}
// Work fields: recomputed inside every step, stored in the precision of the
// mode (FP32 in LES_MIXED), and the work-precision state above.  Every
// other 3-D field is FP64 state.
bool is_work_field(int id) {
    return (id >= LES_F_DUDX && id <= LES_F_DPDZ) || id == LES_F_NU_T || (id >= LES_F_S11 && id <= LES_F_BETA_SCAL) ||
           (id >= LES_F_WORK0 && id <= LES_F_WORK2) || is_work_state(id);
}
// exchange two fields' storage in both precisions
void swap_fields(les_ctx* c, int a, int b) {
    std::swap(c->f3[a], c->f3[b]);
    std::swap(c->f32buf[a], c->f32buf[b]);
    std::swap(c->wk[a], c->wk[b]);
}

// storage of a 3-D field: double or float pointer
struct FieldRef {
    double* d;
    float* f;
};
FieldRef field_ref(les_ctx* c, int id) {
    if (c->precision == LES_MIXED && is_work_field(id)) return {nullptr, c->f32buf[id]};
    return {c->f3[id], nullptr};
}

double* surf(les_ctx* c, int id) {
    switch (id) {
        case LES_F_USTAR_SFC: return c->ustar;
        case LES_F_PSI_M_SFC: return c->psi_m;
        case LES_F_PHI_M_SFC: return c->phi_m;
        case LES_F_PSI_H_SFC: return c->psi_h;
        case LES_F_PHI_H_SFC: return c->phi_h;
        case LES_F_USTAR_TOP_SFC: return c->ustar_top;
        default: return nullptr;
    }
}


bool valid_field(int f) { return f >= 0 && f < LES_N_FIELDS3D; }

// ---- routines ---------------------------------------------------------------
// The wall model's filtered first-level velocity is FP64 in both modes.
template <typename W>
void wallstress(les_ctx* c, int top) {
    const les_params& P = c->prm;
    const int sf = top ? (P.ubc_mom == LES_BC_STRESS_FREE) : (P.lbc_mom == LES_BC_STRESS_FREE);
    if (sf) {
        wallstress_dev<W>(c, top, nullptr, nullptr);
        return;
    }
    const int lev = top ? c->nz - 1 : 1;
    double* w = c->work3[7];  // two planes: u1, v1
    les_note(c, hipMemcpyAsync(w, F(c, LES_F_U) + (size_t)lev * c->plane, c->plane * 8, hipMemcpyDeviceToDevice,
                               c->stream));
    les_note(c, hipMemcpyAsync(w + c->plane, F(c, LES_F_V) + (size_t)lev * c->plane, c->plane * 8,
                               hipMemcpyDeviceToDevice, c->stream));
    test_filter_f<double>(c, w, 2, c->Gtest);
    wallstress_dev<W>(c, top, w, w + c->plane);
}

template <typename W>
void lagrange(les_ctx* c, int jt) {
    const les_params& P = c->prm;
    const int nz = c->nz;
    const long np21 = 21L * nz;  // planes
    interpolag_dev<W>(c);
    W* raw = scr<W>(c->lag_buf);
    W* b2 = raw + (size_t)np21 * c->plane;
    W* b4 = b2 + (size_t)np21 * c->plane;
    // Spectra pruned to the filters' support: kc columns per row, rows
    // min(y, ny - y) < mc.  Modes outside it are zero after either filter,
    // so they are neither stored nor transformed back.
    const int kc = c->lag_kc, kc4 = c->lag_kc4, mc = c->lag_mc;
    const long psp = (long)kc * c->ny;
    cx<W>* spf = scr<W>(c->lag_spec);
    cx<W>* spt = spf + (size_t)np21 * psp;
    lagrange_prep_dev<W>(c, raw);
    x_r2c<W>(c->stream, raw, spf, np21 * c->ny, c->nx, kc, c->tw_x, 1.0, kc);
    yline_spec<W>(c->stream, spf, spf, (int)np21, kc, psp, c->tw_y, SP_FWD, nullptr, kc, -1, mc);
    yline_spec<W>(c->stream, spf, spt, (int)np21, kc, psp, c->tw_y, SP_MULINV, c->Gtest, kc, c->lhx, mc);
    const int init_lm = P.inilag && !c->f_lm_mm_init && (jt == P.cs_count || jt == P.dyn_init);
    const int init_qn = P.inilag && !c->f_qn_nn_init && (jt == P.cs_count || jt == P.dyn_init);
    // each filter's fields contracted as they are transformed back (L M, M M
    // of the 2 delta filter kept in the spent product buffer), else the
    // filtered fields are stored and contracted by k_lag_contract
    const double delta = 1.0 * pow(c->dx * c->dy * c->dz, 1.0 / 3.0);
    const LagUpdate u{c->f3[LES_F_F_LM], c->f3[LES_F_F_MM], c->f3[LES_F_F_QN], c->f3[LES_F_F_NN],
                      c->f3[LES_F_CS_OPT2], init_lm, init_qn, nz, OPFTIME * delta,
                      P.dt * (double)P.cs_count, log(2.0) / (log(4.0) - log(2.0))};
    double* lmb = c->lag_buf;
    double* mmb = c->lag_buf + c->plane * c->nlev;
    const double cst = 2.0 * (delta * delta);
    if (c->fuse && g_use_ct &&
        lag_contract_x<W>(c->stream, 2, spt, kc, kc, nz, c->ny, lmb, mmb, u, cst, c->tw_x)) {
        yline_spec<W>(c->stream, spf, spt, (int)np21, kc4, psp, c->tw_y, SP_MULINV, c->Gtest2, kc, c->lhx, mc);
        lag_contract_x<W>(c->stream, 4, spt, kc, kc4, nz, c->ny, lmb, mmb, u, cst, c->tw_x);
    } else {
        x_c2r<W>(c->stream, spt, b2, np21 * c->ny, kc, c->nx, c->tw_x, 0, 0.0, kc);
        yline_spec<W>(c->stream, spf, spt, (int)np21, kc4, psp, c->tw_y, SP_MULINV, c->Gtest2, kc, c->lhx, mc);
        x_c2r<W>(c->stream, spt, b4, np21 * c->ny, kc, c->nx, c->tw_x, 0, 0.0, kc4);
        lagrange_contract_dev<W>(c, b2, b4, init_lm, init_qn);
    }
    if (init_lm) c->f_lm_mm_init = 1;
    if (init_qn) c->f_qn_nn_init = 1;
}

// keep_s: the strain fields S_ij must hold this step's values afterwards
// (the routine entry, taped steps, statistics steps).  Otherwise, on the
// steps without the dynamic update, the strain is formed inside the Nu_t /
// tau kernel and not stored.
template <typename W>
void sgs_stag(les_ctx* c, int jt, bool keep_s = true) {
    const les_params& P = c->prm;
    const bool init = (jt == 1 && P.inilag);
    const bool update = !c->sgs_frozen && (jt >= P.dyn_init || P.restart_vel) && (jt % P.cs_count) == 0;
    if (!c->sgs_frozen && !c->lag_done) lag_accumulate_dev<W>(c);
    if (c->fuse && !keep_s && !init && !update) {
        strain_stress_dev<W>(c);
        return;
    }
    strain_dev<W>(c);
    if (init) cs_init_dev(c);
    else if (update) lagrange<W>(c, jt);
    nu_t_stress_dev<W>(c);
}

template <typename W>
void divstress(les_ctx* c, W* out, const W* tx, const W* ty, const W* tz, int wver) {
    W* a = scr<W>(c->work3[0]);
    W* b = scr<W>(c->work3[1]);
    ddx_f<W, W>(c, a, tx, 0, (int)c->nlev);
    ddy_f<W, W>(c, b, ty, 0, (int)c->nlev);
    divstress_combine_dev<W>(c, out, a, b, tz, wver);
}

template <typename W>
void convec(les_ctx* c, bool uvw_on_aux = false) {
    const int nz = c->nz;
    const size_t np = c->plane, np2 = c->plane2;
    W* ub = scr<W>(c->u_big);
    W* vb = scr<W>(c->v_big);
    W* wb = scr<W>(c->w_big);
    W* big[6];
    for (int i = 0; i < 6; ++i) big[i] = scr<W>(c->big3[i]);
    if (!uvw_on_aux) {
        pad_f<W, double>(c, F(c, LES_F_U), ub, (int)c->nlev);
        pad_f<W, double>(c, F(c, LES_F_V), vb, (int)c->nlev);
        pad_f<W, double>(c, F(c, LES_F_W), wb, (int)c->nlev);
    }
    // vorticity and the x pass of its pads in one kernel (x-padded rows in
    // big[3..5]), then the y passes into big[0..2]
    const VortIn<W> vg{WF<W>(c, LES_F_DUDY), WF<W>(c, LES_F_DUDZ), WF<W>(c, LES_F_DVDX),
                       WF<W>(c, LES_F_DVDZ), WF<W>(c, LES_F_DWDX), WF<W>(c, LES_F_DWDY)};
    const bool fz = c->fuse && g_use_ct;
    if (fz && vort_pad_x<W>(c->stream, vg, big[3], big[4], big[5], nz, c->ny, c->tw_x, c->tw_x2)) {
        for (int m = 0; m < 3; ++m)
            yline_real<W>(c->stream, big[3 + m], big[m] + np2, nz, c->nx2, c->nx2, c->nx2, (long)c->ny * c->nx2,
                          (long)np2, c->tw_y, c->tw_y2, OP_PAD, 0.0);
    } else {
        W *v1 = scr<W>(c->work3[2]), *v2 = scr<W>(c->work3[3]), *v3 = scr<W>(c->work3[4]);
        vorticity_dev<W>(c, v1, v2, v3);
        pad_f<W, W>(c, v1 + np, big[0] + np2, nz);
        pad_f<W, W>(c, v2 + np, big[1] + np2, nz);
        pad_f<W, W>(c, v3 + np, big[2] + np2, nz);
    }
    W* cxf = WF<W>(c, LES_F_CX);
    W* cyf = WF<W>(c, LES_F_CY);
    W* czf = WF<W>(c, LES_F_CZ);
    // cross product and the x pass of its unpads in one kernel (x-unpadded
    // rows in big[3..5]), then the y passes
    const ConvIn<W> ci{ub, vb, wb, big[0], big[1], big[2], c->ust, c->vst};
    if (uvw_on_aux) les_note(c, hipStreamWaitEvent(c->stream, c->ev_uvw, 0));
    if (fz && cross_unpad_x<W>(c->stream, ci, 0, big[3], big[4], big[5], nz, c->ny2, c->tw_x2, c->tw_x)) {
        W* dst[3] = {cxf, cyf, czf};
        for (int m = 0; m < 3; ++m)
            yline_real<W>(c->stream, big[3 + m], dst[m] + np, nz - 1, c->nx, c->nx, c->nx, (long)c->ny2 * c->nx,
                          (long)np, c->tw_y2, c->tw_y, OP_UNPAD, 0.0);
    } else {
        cross_product_dev<W>(c, big[0], big[1], big[2], big[3], big[4], big[5]);
        unpad_f<W>(c, big[3] + np2, cxf + np, nz - 1);
        unpad_f<W>(c, big[4] + np2, cyf + np, nz - 1);
        unpad_f<W>(c, big[5] + np2, czf + np, nz - 1);
    }
    fill_runs<W>(c->stream, (long)np, BOGUS, {cxf, cyf, czf, cxf + (size_t)nz * np, cyf + (size_t)nz * np});
}

// theta_all_in_one (010_scalars.f90:111-671) in two parts.  The first needs
// only theta, so the step runs it before the wall model to overlap the
// ordered plane sums; the second needs Nu_t (sgs_stag) and the padded
// velocity (convec).
bool theta_active(const les_ctx* c, int jt) {
    return c->prm.theta_flag && jt >= c->prm.theta_init_time;
}

// theta's gradients on the 3/2 grid for the advection (:466-470)
template <typename W>
void theta_pads(les_ctx* c, W* dtdx, W* dtdy, W* dtdz) {
    const int nz = c->nz;
    const size_t np = c->plane, np2 = c->plane2;
    pad_f<W, W>(c, dtdx + np, scr<W>(c->tbig[0]) + np2, nz - 1);
    pad_f<W, W>(c, dtdy + np, scr<W>(c->tbig[1]) + np2, nz - 1);
    pad_f<W, W>(c, dtdz + 2 * np, scr<W>(c->tbig[2]) + 2 * np2, nz - 1);
}

// filter_data(theta), dTdx, dTdy, dTdz, the theta RHS history (:111-140) and
// the padded gradients (:466-470); calc_beta's plane means of theta start on
// the auxiliary stream
template <typename W>
void theta_pre(les_ctx* c, int jt) {
    if (!theta_active(c, jt)) return;
    const int nz = c->nz;
    const size_t np = c->plane, np2 = c->plane2;
    double* th = F(c, LES_F_THETA);
    W* dtdx = WF<W>(c, LES_F_DTDX);
    W* dtdy = WF<W>(c, LES_F_DTDY);
    W* dtdz = WF<W>(c, LES_F_DTDZ);
    filter_f<double>(c, th, 0, (int)c->nlev);
    fork(c);
    level_means_seq(c->aux, th, np, 0, nz, (double)(c->nx * c->ny), c->th_mean);   // calc_hor_avg order
    les_note(c, hipEventRecord(c->ev_th, c->aux));
    ddx_f<W, double>(c, dtdx, th, 0, (int)c->nlev);
    ddy_f<W, double>(c, dtdy, th, 0, (int)c->nlev);
    ddz_uv_dev<W, double>(c, dtdz, th);
    theta_prep_dev<W>(c);
    theta_pads<W>(c, dtdx, dtdy, dtdz);
}

// convec's zero_padding of u, v, w on the auxiliary stream, first in the
// step there: they read only u, v, w, which nothing changes before the
// step's update, so they overlap the CFL pass, the derivatives, the wall
// model and the SGS stresses; convec waits for ev_uvw before the cross
// product.
// The x passes take the auxiliary scratch (tscr[3]).
template <typename W>
void uvw_pads_overlap(les_ctx* c) {
    std::swap(c->stream, c->aux);
    std::swap(c->work3[6], c->tscr[3]);
    pad_f<W, double>(c, F(c, LES_F_U), scr<W>(c->u_big), (int)c->nlev);
    pad_f<W, double>(c, F(c, LES_F_V), scr<W>(c->v_big), (int)c->nlev);
    pad_f<W, double>(c, F(c, LES_F_W), scr<W>(c->w_big), (int)c->nlev);
    les_note(c, hipEventRecord(c->ev_uvw, c->stream));
    std::swap(c->work3[6], c->tscr[3]);
    std::swap(c->stream, c->aux);
}

// theta_pre on the auxiliary stream, after the padded u, v, w there.  The
// momentum path reads none of theta's fields until calc_beta, which waits for
// ev_th (recorded after the plane means); the x passes of the pads take the
// auxiliary scratch, so the momentum pads on the context's stream run
// concurrently.  Same kernels in the same order as theta_pre.
template <typename W>
void theta_pre_overlap(les_ctx* c, int jt) {
    if (!theta_active(c, jt)) return;
    const int nz = c->nz;
    const size_t np = c->plane;
    std::swap(c->stream, c->aux);
    std::swap(c->work3[6], c->tscr[3]);
    double* th = F(c, LES_F_THETA);
    W* dtdx = WF<W>(c, LES_F_DTDX);
    W* dtdy = WF<W>(c, LES_F_DTDY);
    W* dtdz = WF<W>(c, LES_F_DTDZ);
    filter_f<double>(c, th, 0, (int)c->nlev);
    ddx_f<W, double>(c, dtdx, th, 0, (int)c->nlev);
    ddy_f<W, double>(c, dtdy, th, 0, (int)c->nlev);
    ddz_uv_dev<W, double>(c, dtdz, th);
    theta_prep_dev<W>(c);
    theta_pads<W>(c, dtdx, dtdy, dtdz);
    level_means_seq(c->stream, th, np, 0, nz, (double)(c->nx * c->ny), c->th_mean);   // calc_hor_avg order
    les_note(c, hipEventRecord(c->ev_th, c->stream));
    std::swap(c->work3[6], c->tscr[3]);
    std::swap(c->stream, c->aux);
}

// advection, SGS fluxes, calc_beta and step_scalar (:466-671)
template <typename W>
void theta_rhs(les_ctx* c);

template <typename W>
void theta_post(les_ctx* c, int jt, Tape* tape = nullptr) {
    const les_params& P = c->prm;
    const size_t np = c->plane;
    if (!theta_active(c, jt)) {
        fill_levels(c->stream, WF<W>(c, LES_F_BETA_SCAL), np, 0, (int)c->nlev, 0.0);
        return;
    }
    theta_rhs<W>(c);
    les_note(c, hipStreamWaitEvent(c->stream, c->ev_th, 0));
    calc_beta_dev<W>(c, c->th_mean);
    if (tape) tape_save(c, tape, TAPE_THETA);
    if (jt == P.theta_init_time && !P.restart_theta)
        les_note(c, hipMemcpyAsync(F(c, LES_F_RHST_F), F(c, LES_F_RHST), np * c->nlev * 8, hipMemcpyDeviceToDevice,
                                   c->stream));
    step_scalar_dev(c);
}

// The same with theta's RHS and update on the auxiliary stream, concurrent
// with the momentum RHS and the pressure solve: calc_beta (which the momentum
// RHS reads, and which must read theta before step_scalar updates it) runs on
// the context's stream first; the rest runs on aux with its own scratch
// (theta reads nothing the momentum path writes after convec, and writes
// nothing it reads or writes: project's boundary rows of theta, levels 0 and
// nz, are written on aux after the update, not by k_project_bc, so that they
// come last as in the serial order).  The caller joins the streams before the
// step ends.
template <typename W>
void theta_post_overlap(les_ctx* c, int jt, Tape* tape = nullptr) {
    const les_params& P = c->prm;
    const size_t np = c->plane;
    if (!theta_active(c, jt)) {
        fill_levels(c->stream, WF<W>(c, LES_F_BETA_SCAL), np, 0, (int)c->nlev, 0.0);
        return;
    }
    les_note(c, hipStreamWaitEvent(c->stream, c->ev_th, 0));
    calc_beta_dev<W>(c, c->th_mean);
    // the filtered theta and its plane means, before the update on aux
    if (tape) tape_save(c, tape, TAPE_THETA);
    fork(c);
    double** scratch[5] = {&c->work3[0], &c->work3[1], &c->work3[2], &c->work3[6], &c->big3[3]};
    std::swap(c->stream, c->aux);
    for (int i = 0; i < 5; ++i) std::swap(*scratch[i], c->tscr[i]);
    theta_rhs<W>(c);
    if (jt == P.theta_init_time && !P.restart_theta)
        les_note(c, hipMemcpyAsync(F(c, LES_F_RHST_F), F(c, LES_F_RHST), np * c->nlev * 8, hipMemcpyDeviceToDevice,
                                   c->stream));
    step_scalar_dev(c);
    theta_rows_dev(c);  // project's theta rows (left out of k_project_bc on the context's stream)
    for (int i = 0; i < 5; ++i) std::swap(*scratch[i], c->tscr[i]);
    std::swap(c->stream, c->aux);
}

template <typename W>
void theta_fluxes(les_ctx* c, const W* adv);

// theta's RHS: advection and the SGS flux divergence (:466-594)
template <typename W>
void theta_rhs(les_ctx* c) {
    const int nz = c->nz;
    const size_t np = c->plane, np2 = c->plane2;
    W* rm = scr<W>(c->big3[3]);
    W* adv = scr<W>(c->work3[2]);
    // advection and the x pass of its unpad in one kernel (x-unpadded rows in
    // big3[3]), then the y pass
    const ConvIn<W> si{scr<W>(c->u_big), scr<W>(c->v_big), scr<W>(c->w_big), scr<W>(c->tbig[0]),
                       scr<W>(c->tbig[1]), scr<W>(c->tbig[2]), c->ust, c->vst};
    if (c->fuse && g_use_ct && cross_unpad_x<W>(c->stream, si, 1, rm, rm, rm, nz, c->ny2, c->tw_x2, c->tw_x)) {
        yline_real<W>(c->stream, rm, adv + np, nz - 1, c->nx, c->nx, c->nx, (long)c->ny2 * c->nx, (long)np,
                      c->tw_y2, c->tw_y, OP_UNPAD, 0.0);
    } else {
        scalar_adv_dev<W>(c, scr<W>(c->tbig[0]), scr<W>(c->tbig[1]), scr<W>(c->tbig[2]), rm);
        unpad_f<W>(c, rm + np2, adv + np, nz - 1);
    }
    theta_fluxes<W>(c, adv);
}

// the SGS flux divergence added to the advection adv (:513-594)
template <typename W>
void theta_fluxes(les_ctx* c, const W* adv) {
    W* temp = scr<W>(c->work3[0]);
    W* dtemp = scr<W>(c->work3[1]);
    scalar_flux_dev<W>(c, temp, 0);
    ddx_f<W, W>(c, dtemp, temp, 0, (int)c->nlev);
    scalar_accum_dev<W>(c, dtemp, adv);
    scalar_flux_dev<W>(c, temp, 1);
    ddy_f<W, W>(c, dtemp, temp, 0, (int)c->nlev);
    scalar_accum_dev<W>(c, dtemp, (const W*)nullptr);
    scalar_flux_dev<W>(c, temp, 2);
    if (c->fuse) {
        ddz_w_accum_dev<W>(c, temp);  // the vertical term added as it is differenced
    } else {
        ddz_w_dev<W, W>(c, dtemp, temp);
        scalar_accum_dev<W>(c, dtemp, (const W*)nullptr);
    }
}

template <typename W>
void theta_all_in_one(les_ctx* c, int jt) {
    theta_pre<W>(c, jt);
    theta_post<W>(c, jt);
}

// FP64: H spectra in spec3[0..2], p_hat (and d') in spec3[3].  Mixed: FP32
// H spectra in spec3[0..2], d' in spec3[3] (FP64), FP32 p_hat in phf, and the
// mean mode from FP64 plane sums.
template <typename W>
void press_stag_array(les_ctx* c, int jt) {
    press_solve<W>(c, F(c, LES_F_U), F(c, LES_F_V), F(c, LES_F_W), WF<W>(c, LES_F_DIVTZ), WF<W>(c, LES_F_P),
                   WF<W>(c, LES_F_DPDX), WF<W>(c, LES_F_DPDY), WF<W>(c, LES_F_DPDZ), press_modi(c, jt));
}

// energy and d(ke)/dx_i added to the dudt diagnostics (103_flosol.f90:192-213,
// :393-429); dkedz at the wall level is not set by the reference (0 here)
void diagnostics_ke(les_ctx* c) {
    const int nl = (int)c->nlev;
    ke_dev(c, c->st_ke);
    les_note(c, hipMemcpyAsync(c->st_tmp, c->st_ke, c->plane * c->nlev * 8, hipMemcpyDeviceToDevice, c->stream));
    filter_f<double>(c, c->st_tmp, 0, nl);
    ddx_f<double, double>(c, c->st_dke, c->st_tmp, 0, nl);
    add_levels_dev(c, c->st_diag[0], c->st_dke, 0);
    ddy_f<double, double>(c, c->st_dke, c->st_tmp, 0, nl);
    add_levels_dev(c, c->st_diag[1], c->st_dke, 0);
    ddz_uv_dev<double, double>(c, c->st_dke, c->st_ke);
    add_levels_dev(c, c->st_diag[2], c->st_dke, 1);
}

// copy the running sums of slices [s0, s1) to the window set, zero them
void close_window(les_ctx* c, int s0, int s1) {
    const size_t L = (size_t)c->nx * (c->nz - 1);
    les_note(c, hipMemcpyAsync(c->st_win + s0 * L, c->st_run + s0 * L, sizeof(double) * L * (s1 - s0),
                               hipMemcpyDeviceToDevice, c->stream));
    fill_levels(c->stream, c->st_run + s0 * L, L, 0, s1 - s0, 0.0);
}

// flow_slice and theta_slice (011_scalars2.f90:28-318) at the end of flosol
template <typename W>
void avg_slices(les_ctx* c, int jt) {
    const les_params& P = c->prm;
    flow_slices_dev<W>(c);
    if (jt % c->st_p == 0) close_window(c, 0, LES_N_FLOW_SLICES);
    if (P.theta_flag && jt >= P.theta_init_time) {
        theta_slices_dev<W>(c);
        if (jt % c->st_p == 0) close_window(c, LES_N_FLOW_SLICES, LES_N_SLICES);
    }
}

// obukhov's ordered plane sums on their own stream, after everything
// enqueued so far on the context's stream (which the auxiliary stream also
// waits for); ev_obk marks their end.  A sum is one wavefront's dependent
// chain, so on a stream of its own it does not hold up the auxiliary work.
void obukhov_sums_fork(les_ctx* c, int jt) {
    obukhov_theta_copy(c);
    fork(c);
    les_note(c, hipStreamWaitEvent(c->sums, c->ev_fork, 0));
    obukhov_sums_dev(c, jt, c->sums);
    les_note(c, hipEventRecord(c->ev_obk, c->sums));
}

template <typename W>
void step_body(les_ctx* c, int jt, Tape* tape = nullptr, bool stats = true) {
    // newstep: RHS_f = RHS (103_flosol.f90:311-313) as a pointer exchange
    for (int m = 0; m < 3; ++m) swap_fields(c, LES_F_RHSX + m, LES_F_RHSX_F + m);
    // obukhov's ordered plane sums overlap the CFL pass (when les_run
    // launched them before it), the horizontal derivatives and the theta
    // gradients
    if (!c->obk_early) obukhov_sums_fork(c, jt);
    c->obk_early = 0;
    const bool st = stats && stats_step(c, jt);
    const bool ovl = c->overlap && !st;  // statistics steps stay in order
    if (ovl) uvw_pads_overlap<W>(c);             // convec's padded u, v, w on aux
    const int nl = (int)c->nlev;
    const double *u = F(c, LES_F_U), *v = F(c, LES_F_V), *w = F(c, LES_F_W);
    ddx_f<W, double>(c, WF<W>(c, LES_F_DUDX), u, 0, nl);
    ddy_f<W, double>(c, WF<W>(c, LES_F_DUDY), u, 0, nl);
    ddx_f<W, double>(c, WF<W>(c, LES_F_DVDX), v, 0, nl);
    ddy_f<W, double>(c, WF<W>(c, LES_F_DVDY), v, 0, nl);
    ddx_f<W, double>(c, WF<W>(c, LES_F_DWDX), w, 0, nl);
    ddy_f<W, double>(c, WF<W>(c, LES_F_DWDY), w, 0, nl);
    ddz_uv_dev<W, double>(c, WF<W>(c, LES_F_DUDZ), u);
    ddz_uv_dev<W, double>(c, WF<W>(c, LES_F_DVDZ), v);
    ddz_w_dev<W, double>(c, WF<W>(c, LES_F_DWDZ), w);
    // theta's filter, gradients and pads: on aux (submitted after the
    // derivatives, so that the host, which has a bounded number of launches
    // outstanding, hands the momentum stream its work first), or in order
    if (ovl) theta_pre_overlap<W>(c, jt);
    else theta_pre<W>(c, jt);
    les_note(c, hipStreamWaitEvent(c->stream, c->ev_obk, 0));
    obukhov_finish_dev(c, jt);
    if (tape) tape_save(c, tape, TAPE_OBUKHOV);
    wallstress<W>(c, 0);
    if (tape) tape_save(c, tape, TAPE_WALL);
    wallstress<W>(c, 1);
    sgs_stag<W>(c, jt, tape != nullptr || (stats && stats_step(c, jt)));
    W *txx = WF<W>(c, LES_F_TXX), *txy = WF<W>(c, LES_F_TXY), *txz = WF<W>(c, LES_F_TXZ);
    W *tyy = WF<W>(c, LES_F_TYY), *tyz = WF<W>(c, LES_F_TYZ), *tzz = WF<W>(c, LES_F_TZZ);
    divstress<W>(c, WF<W>(c, LES_F_DIVTX), txx, txy, txz, 0);
    divstress<W>(c, WF<W>(c, LES_F_DIVTY), txy, tyy, tyz, 0);
    divstress<W>(c, WF<W>(c, LES_F_DIVTZ), txz, tyz, tzz, 1);
    convec<W>(c, ovl);
    if (ovl) theta_post_overlap<W>(c, jt, tape);
    else theta_post<W>(c, jt, tape);
    rhs_ab2_dev<W>(c, jt);
    press_stag_array<W>(c, jt);
    gradp_project_dev<W>(c, jt, !(ovl && theta_active(c, jt)));
    if (st) diagnostics_ke(c);
    filter_f<double>(c, F(c, LES_F_U), 0, nl);
    filter_f<double>(c, F(c, LES_F_V), 0, nl);
    filter_f<double>(c, F(c, LES_F_W), 0, nl);
    if (ovl) join(c);
    if (st) avg_slices<W>(c, jt);
    if (tape) tape_save(c, tape, TAPE_END);
}

bool mixed(const les_ctx* c) { return c->precision == LES_MIXED; }

// Run a routine templated on the work precision in the current mode.
#define LES_DISPATCH(fn, ...)                          \
    do {                                               \
        if (mixed(c)) fn<float>(__VA_ARGS__);          \
        else fn<double>(__VA_ARGS__);                  \
    } while (0)

// A derivative writes an FP64 field from an FP64 source, or a work field in
// the mode's precision from either; an FP64 destination with an FP32 source
// (mixed mode) is rejected.
enum DerivKind { D_X, D_Y, D_ZUV, D_ZW };
int deriv_entry(les_ctx* c, int d, int f, DerivKind k) {
    if (!c || !valid_field(d) || !valid_field(f)) return LES_E_BADARG;
    if ((k == D_ZUV || k == D_ZW) && d == f) return LES_E_BADARG;
    const FieldRef rd = field_ref(c, d), rf = field_ref(c, f);
    const int nl = (int)c->nlev;
    if (rd.d && rf.d) {
        switch (k) {
            case D_X: ddx_f<double, double>(c, rd.d, rf.d, 0, nl); break;
            case D_Y: ddy_f<double, double>(c, rd.d, rf.d, 0, nl); break;
            case D_ZUV: ddz_uv_dev<double, double>(c, rd.d, rf.d); break;
            default: ddz_w_dev<double, double>(c, rd.d, rf.d); break;
        }
    } else if (rd.f && rf.d) {
        switch (k) {
            case D_X: ddx_f<float, double>(c, rd.f, rf.d, 0, nl); break;
            case D_Y: ddy_f<float, double>(c, rd.f, rf.d, 0, nl); break;
            case D_ZUV: ddz_uv_dev<float, double>(c, rd.f, rf.d); break;
            default: ddz_w_dev<float, double>(c, rd.f, rf.d); break;
        }
    } else if (rd.f && rf.f) {
        switch (k) {
            case D_X: ddx_f<float, float>(c, rd.f, rf.f, 0, nl); break;
            case D_Y: ddy_f<float, float>(c, rd.f, rf.f, 0, nl); break;
            case D_ZUV: ddz_uv_dev<float, float>(c, rd.f, rf.f); break;
            default: ddz_w_dev<float, float>(c, rd.f, rf.f); break;
        }
    } else {
        return LES_E_BADARG;
    }
    return finish(c);
}
// divstress on four fields that share one storage precision
int divstress_entry(les_ctx* c, int divt, int tx, int ty, int tz, int wver) {
    if (!c || !valid_field(divt) || !valid_field(tx) || !valid_field(ty) || !valid_field(tz)) return LES_E_BADARG;
    const FieldRef r[4] = {field_ref(c, divt), field_ref(c, tx), field_ref(c, ty), field_ref(c, tz)};
    bool all_d = true, all_f = true;
    for (const FieldRef& x : r) {
        all_d = all_d && x.d;
        all_f = all_f && x.f;
    }
    if (all_d) divstress<double>(c, r[0].d, r[1].d, r[2].d, r[3].d, wver);
    else if (all_f) divstress<float>(c, r[0].f, r[1].f, r[2].f, r[3].f, wver);
    else return LES_E_BADARG;
    return finish(c);
}

}  // namespace

void les::step_fp64(les_ctx* c, int jt, Tape* tape, bool stats) { step_body<double>(c, jt, tape, stats); }

int les::finish(les_ctx* c) {
    les_note(c, hipStreamSynchronize(c->sums));
    les_note(c, hipStreamSynchronize(c->aux));
    hipError_t e = hipStreamSynchronize(c->stream);
    les_note(c, e);
    les_note(c, hipGetLastError());
    const hipError_t first = c->err;
    c->err = hipSuccess;
    return first == hipSuccess ? LES_OK : LES_E_HIP;
}

// ---- scratch of the allocation-free entry point -----------------------------
// les_zero_padding_noalloc keeps the x-padded intermediate of zero_padding in
// this array instead of context memory.  Every GPU the library runs on holds
// its own copy until the process ends.  Host code never uses the array's name
// as an address: les_create looks up each device's copy once
// (hipGetSymbolAddress) and the entry point passes that pointer.
constexpr size_t NOALLOC_DOUBLES = (size_t)6 << 20;   // 48 MiB
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];
constexpr int NOALLOC_MAX_DEV = 64;
static double* g_scratch_dev[NOALLOC_MAX_DEV];        // per device, set once

static std::mutex g_scratch_mu;

// the device copy of scratch for the current device (looked up once)
static hipError_t noalloc_scratch(int dev, double** out) {
    if (dev < 0 || dev >= NOALLOC_MAX_DEV) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_dev[dev] = static_cast<double*>(ptr);
    }
    *out = g_scratch_dev[dev];
    return hipSuccess;
}

// ============================================================================
// C-ABI
// ============================================================================
extern "C" {

int les_create(const les_params* prm, les_ctx** out) {
    if (!prm || !out) return LES_E_BADARG;
    *out = nullptr;
    if (!good_size(prm->nxt) || !good_size(prm->nyt) || prm->nzt < 4) return LES_E_BADARG;
    if (prm->model != 5 || !prm->sgs || prm->cs_count < 1) return LES_E_BADARG;
    if (prm->lbc != 0 && prm->lbc != 1) return LES_E_BADARG;
    // lbc = 0 reads ustar_avg and psi_h, which the ocean branch never sets on
    // the surface ranks (010_scalars.f90:380-390, :696-716)
    if (prm->lbc == 0 && prm->ocean_flag) return LES_E_BADARG;
    if (prm->damping_method < 1 || prm->damping_method > 3) return LES_E_BADARG;
    if (3 * prm->nxt / 2 > 3072 || 3 * prm->nyt / 2 > 3072) return LES_E_BADARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return LES_E_NODEV;
    // LES_LINE_KERNELS=generic forces the runtime-length line kernels (process-wide)
    if (const char* lk = getenv("LES_LINE_KERNELS")) g_use_ct = strcmp(lk, "generic") != 0;
    les_ctx* c = new les_ctx();
    c->prm = *prm;
    // LES_FUSE=0: the separate kernels instead of the fused ones (bitwise the
    // same results; for checking that)
    c->fuse = 1;
    if (const char* fz = getenv("LES_FUSE")) c->fuse = strcmp(fz, "0") != 0;
    // LES_OVERLAP=0: theta's RHS and update in order on the context's stream
    c->overlap = 1;
    if (const char* ov = getenv("LES_OVERLAP")) c->overlap = strcmp(ov, "0") != 0;
    c->nx = prm->nxt; c->ny = prm->nyt; c->nz = prm->nzt;
    c->nx2 = 3 * c->nx / 2; c->ny2 = 3 * c->ny / 2;
    c->lhx = c->nx / 2 + 1; c->lhy = c->ny / 2 + 1;
    c->plane = (size_t)c->nx * c->ny;
    c->plane2 = (size_t)c->nx2 * c->ny2;
    c->nlev = (size_t)c->nz + 1;
    c->ns = (size_t)c->lhx * c->ny;
    c->dx = prm->lx_tot / c->nx;
    c->dy = prm->ly_tot / c->ny;
    c->dz = prm->lz_tot / (c->nz - 1);
    c->precision = LES_FP64;
    if (hipGetDevice(&c->dev) != hipSuccess) { delete c; return LES_E_NODEV; }
    if (noalloc_scratch(c->dev, &c->noalloc) != hipSuccess) { delete c; return LES_E_HIP; }
    c->err = hipSuccess;
    if (hipStreamCreate(&c->stream) != hipSuccess) { delete c; return LES_E_HIP; }
    if (hipStreamCreate(&c->aux) != hipSuccess || hipStreamCreate(&c->sums) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_obk, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_th, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_uvw, hipEventDisableTiming) != hipSuccess) {
        les_destroy(c);
        return LES_E_HIP;
    }
    const size_t F3 = c->plane * c->nlev;
    hipError_t e = hipSuccess;
    for (int f = 0; f < LES_N_FIELDS3D && e == hipSuccess; ++f) e = dalloc(c, &c->f3[f], F3);
    double** surfs[6] = {&c->ustar, &c->psi_m, &c->phi_m, &c->psi_h, &c->phi_h, &c->ustar_top};
    for (int s = 0; s < 6 && e == hipSuccess; ++s) e = dalloc(c, surfs[s], c->plane);
    if (e == hipSuccess) e = dalloc(c, &c->dmean, 3 * c->nlev);
    const size_t B3 = c->plane2 * c->nlev;
    if (e == hipSuccess) e = dalloc(c, &c->u_big, B3);
    if (e == hipSuccess) e = dalloc(c, &c->v_big, B3);
    if (e == hipSuccess) e = dalloc(c, &c->w_big, B3);
    for (int i = 0; i < 6 && e == hipSuccess; ++i) e = dalloc(c, &c->big3[i], B3);
    for (int i = 0; i < 3 && e == hipSuccess; ++i) e = dalloc(c, &c->tbig[i], B3);
    if (e == hipSuccess) e = dalloc(c, &c->th_mean, c->nlev);
    if (e == hipSuccess) e = dalloc(c, &c->th1_copy, c->plane);
    for (int i = 0; i < 8 && e == hipSuccess; ++i) e = dalloc(c, &c->work3[i], i == 6 ? (F3 * 3) / 2 + c->plane2 : F3);
    for (int i = 0; i < 3 && e == hipSuccess; ++i) e = dalloc(c, &c->tscr[i], F3);
    if (e == hipSuccess) e = dalloc(c, &c->tscr[3], (F3 * 3) / 2 + c->plane2);
    if (e == hipSuccess) e = dalloc(c, &c->tscr[4], B3);
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = dalloc(c, &c->spec3[i], c->ns * c->nlev);
    const size_t np21 = (size_t)21 * c->nz;
    if (e == hipSuccess) e = dalloc(c, &c->lag_buf, 3 * np21 * c->plane);
    if (e == hipSuccess) e = dalloc(c, &c->lag_spec, 2 * np21 * c->ns);
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = dalloc(c, &c->fold[i], F3);
    if (e == hipSuccess) e = dalloc(c, &c->scal, 64);
    if (e == hipSuccess) e = dalloc(c, &c->red, 64 + 3 * MAX3_BLOCKS + c->nlev);
    c->cfl_cap = 4096;
    if (e == hipSuccess) e = dalloc(c, &c->cfl_dev, (size_t)c->cfl_cap);
    if (e == hipSuccess) e = dalloc(c, &c->ust, c->nlev);
    if (e == hipSuccess) e = dalloc(c, &c->vst, c->nlev);
    if (e == hipSuccess) e = dalloc(c, &c->sponge, c->nlev);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->cfl_host, sizeof(double) * c->cfl_cap, hipHostMallocDefault);
    if (e != hipSuccess) {
        les_destroy(c);
        return LES_E_NOMEM;
    }
    // twiddles, wavenumbers, filters, TDMA factorisation
    if (make_twiddle(c, c->tw_x, c->nx) != hipSuccess || make_twiddle(c, c->tw_y, c->ny) != hipSuccess ||
        make_twiddle(c, c->tw_x2, c->nx2) != hipSuccess || make_twiddle(c, c->tw_y2, c->ny2) != hipSuccess ||
        init_spectral(c) != hipSuccess) {
        les_destroy(c);
        return LES_E_HIP;
    }
    for (int f = 0; f < LES_N_FIELDS3D; ++f) les_note(c, hipMemsetAsync(c->f3[f], 0, F3 * 8, c->stream));
    for (int f = 0; f < LES_N_FIELDS3D; ++f) c->wk[f] = is_work_field(f) ? (void*)c->f3[f] : nullptr;
    les_note(c, hipMemsetAsync(c->ust, 0, c->nlev * 8, c->stream));
    les_note(c, hipMemsetAsync(c->vst, 0, c->nlev * 8, c->stream));
    les_note(c, hipMemsetAsync(c->psi_m, 0, c->plane * 8, c->stream));
    les_note(c, hipMemsetAsync(c->psi_h, 0, c->plane * 8, c->stream));
    les_note(c, hipMemsetAsync(c->ustar, 0, c->plane * 8, c->stream));
    fill_levels(c->stream, c->phi_m, c->plane, 0, 1, 1.0);
    fill_levels(c->stream, c->phi_h, c->plane, 0, 1, 1.0);
    les_note(c, hipMemsetAsync(c->ustar_top, 0, c->plane * 8, c->stream));
    les_note(c, hipMemsetAsync(c->dmean, 0, 3 * c->nlev * 8, c->stream));
    std::vector<double> sp = sponge_profile(c);
    les_note(c, hipMemcpyAsync(c->sponge, sp.data(), c->nlev * 8, hipMemcpyHostToDevice, c->stream));
    std::vector<double> ust, vst;
    if (!stokes_profile(c, ust, vst)) {
        les_destroy(c);
        return LES_E_BADARG;
    }
    les_note(c, hipMemcpyAsync(c->ust, ust.data(), c->nlev * 8, hipMemcpyHostToDevice, c->stream));
    les_note(c, hipMemcpyAsync(c->vst, vst.data(), c->nlev * 8, hipMemcpyHostToDevice, c->stream));
    c->wind_ustar = prm->u_star;      // 102_initial.f90:201, 246
    c->wind_rad = prm->rad_stress;
    if (finish(c) != LES_OK) {
        les_destroy(c);
        return LES_E_HIP;
    }
    *out = c;
    return LES_OK;
}

int les_destroy(les_ctx* c) {
    if (!c) return LES_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->aux) (void)hipStreamSynchronize(c->aux);
    if (c->sums) (void)hipStreamSynchronize(c->sums);
    lin_destroy(c);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->cfl_host) (void)hipHostFree(c->cfl_host);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->ev_obk) (void)hipEventDestroy(c->ev_obk);
    if (c->ev_th) (void)hipEventDestroy(c->ev_th);
    if (c->ev_uvw) (void)hipEventDestroy(c->ev_uvw);
    if (c->aux) (void)hipStreamDestroy(c->aux);
    if (c->sums) (void)hipStreamDestroy(c->sums);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return LES_OK;
}

// LES_MIXED: FFTs, nonlinear terms, SGS and scalar RHS on FP32 work fields;
// the prognostic and SGS state, the TDMA, the mean mode, the wall model and
// the surface statistics stay FP64.  The FP32 buffers are allocated on the
// first switch and kept until les_destroy.
int les_set_precision(les_ctx* c, int precision) {
    if (!c || (precision != LES_FP64 && precision != LES_MIXED)) return LES_E_BADARG;
    if (precision == LES_MIXED && !c->phf) {
        const size_t F3 = c->plane * c->nlev;
        hipError_t e = hipSuccess;
        for (int f = 0; f < LES_N_FIELDS3D && e == hipSuccess; ++f)
            if (is_work_field(f)) e = dalloc(c, &c->f32buf[f], F3);
        if (e == hipSuccess) e = dalloc(c, &c->hzm, c->nlev);
        if (e == hipSuccess) e = dalloc(c, &c->lmean, 2 * c->nlev);
        cx<float>* ph = nullptr;
        if (e == hipSuccess) e = dalloc(c, &ph, c->ns * c->nlev);
        if (e != hipSuccess) return LES_E_NOMEM;
        for (int f = 0; f < LES_N_FIELDS3D; ++f)
            if (c->f32buf[f]) les_note(c, hipMemsetAsync(c->f32buf[f], 0, F3 * 4, c->stream));
        c->phf = ph;
    }
    // the work-precision state follows the mode
    if (precision != c->precision) {
        const long n = (long)(c->plane * c->nlev);
        for (int f = 0; f < LES_N_FIELDS3D; ++f) {
            if (!is_work_state(f)) continue;
            if (precision == LES_MIXED) convert<float, double>(c->stream, c->f32buf[f], c->f3[f], n);
            else convert<double, float>(c->stream, c->f3[f], c->f32buf[f], n);
        }
    }
    c->precision = precision;
    for (int f = 0; f < LES_N_FIELDS3D; ++f) {
        if (!is_work_field(f)) c->wk[f] = nullptr;
        else c->wk[f] = (precision == LES_MIXED) ? (void*)c->f32buf[f] : (void*)c->f3[f];
    }
    return finish(c);
}

int les_set_stokes(les_ctx* c, const double* ust, const double* vst) {
    if (!c || !ust || !vst) return LES_E_BADARG;
    les_note(c, hipMemcpyAsync(c->ust, ust, c->nlev * 8, hipMemcpyHostToDevice, c->stream));
    les_note(c, hipMemcpyAsync(c->vst, vst, c->nlev * 8, hipMemcpyHostToDevice, c->stream));
    return finish(c);
}

int les_set_sponge(les_ctx* c, const double* sp) {
    if (!c || !sp) return LES_E_BADARG;
    les_note(c, hipMemcpyAsync(c->sponge, sp, c->nlev * 8, hipMemcpyHostToDevice, c->stream));
    return finish(c);
}

int les_get_profiles(les_ctx* c, double* ust, double* vst, double* sponge) {
    if (!c) return LES_E_BADARG;
    if (ust) les_note(c, hipMemcpyAsync(ust, c->ust, c->nlev * 8, hipMemcpyDeviceToHost, c->stream));
    if (vst) les_note(c, hipMemcpyAsync(vst, c->vst, c->nlev * 8, hipMemcpyDeviceToHost, c->stream));
    if (sponge) les_note(c, hipMemcpyAsync(sponge, c->sponge, c->nlev * 8, hipMemcpyDeviceToHost, c->stream));
    return finish(c);
}

// wallstress_bottom's dynamic wind stress (104_wallstress.f90:119-127)
int les_set_wind_stress(les_ctx* c, double ustar, double agl_deg) {
    if (!c) return LES_E_BADARG;
    c->wind_ustar = ustar;
    c->wind_rad = agl_deg / 180.0 * PI;   // deg2rad, 014_stokes_drift.f90:300-307
    return LES_OK;
}

// FP32 work fields of the mixed mode are converted through an FP64 staging
// field (work3[7], one full 3-D field).
int les_set_field(les_ctx* c, int id, const double* h, int ld, int lev0, int nlev) {
    if (!c || !h || ld < c->nx) return LES_E_BADARG;
    double* d;
    float* df = nullptr;
    if (valid_field(id)) {
        if (lev0 < 0 || nlev < 0 || lev0 + nlev > (int)c->nlev) return LES_E_BADARG;
        const FieldRef r = field_ref(c, id);
        if (r.f) {
            df = r.f + (size_t)lev0 * c->plane;
            d = c->work3[7];
        } else {
            d = r.d + (size_t)lev0 * c->plane;
        }
    } else {
        d = surf(c, id);
        if (!d || lev0 != 0 || nlev != 1) return LES_E_BADARG;
    }
    les_note(c, hipMemcpy2DAsync(d, (size_t)c->nx * 8, h, (size_t)ld * 8, (size_t)c->nx * 8, (size_t)nlev * c->ny,
                     hipMemcpyHostToDevice, c->stream));
    if (df) convert(c->stream, df, (const double*)d, (long)nlev * (long)c->plane);
    return finish(c);
}

int les_get_field(les_ctx* c, int id, double* h, int ld, int lev0, int nlev) {
    if (!c || !h || ld < c->nx) return LES_E_BADARG;
    const double* d;
    if (valid_field(id)) {
        if (lev0 < 0 || nlev < 0 || lev0 + nlev > (int)c->nlev) return LES_E_BADARG;
        const FieldRef r = field_ref(c, id);
        if (r.f) {
            convert(c->stream, c->work3[7], (const float*)(r.f + (size_t)lev0 * c->plane), (long)nlev * (long)c->plane);
            d = c->work3[7];
        } else {
            d = r.d + (size_t)lev0 * c->plane;
        }
    } else {
        d = surf(c, id);
        if (!d || lev0 != 0 || nlev != 1) return LES_E_BADARG;
    }
    les_note(c, hipMemcpy2DAsync(h, (size_t)ld * 8, d, (size_t)c->nx * 8, (size_t)c->nx * 8, (size_t)nlev * c->ny,
                     hipMemcpyDeviceToHost, c->stream));
    return finish(c);
}

// Registration of the caller's arrays (the reference's module arrays) for
// bulk synchronisation; the arrays must outlive the context or the next
// les_bind_field of the same field.
int les_bind_field(les_ctx* c, int id, double* h, int ld, int lev0, int nlev) {
    if (!c || !h || ld < c->nx) return LES_E_BADARG;
    if (valid_field(id)) {
        if (lev0 < 0 || nlev < 0 || lev0 + nlev > (int)c->nlev) return LES_E_BADARG;
    } else if (!surf(c, id) || lev0 != 0 || nlev != 1) {
        return LES_E_BADARG;
    }
    c->binds.push_back({id, h, ld, lev0, nlev});
    return LES_OK;
}
int les_upload(les_ctx* c) {
    if (!c) return LES_E_BADARG;
    for (const Binding& b : c->binds) {
        const int st = les_set_field(c, b.id, b.h, b.ld, b.lev0, b.nlev);
        if (st != LES_OK) return st;
    }
    return LES_OK;
}
int les_download(les_ctx* c) {
    if (!c) return LES_E_BADARG;
    for (const Binding& b : c->binds) {
        const int st = les_get_field(c, b.id, b.h, b.ld, b.lev0, b.nlev);
        if (st != LES_OK) return st;
    }
    return LES_OK;
}

int les_set_sgs_init_flags(les_ctx* c, int a, int b) {
    if (!c) return LES_E_BADARG;
    c->f_lm_mm_init = a;
    c->f_qn_nn_init = b;
    return LES_OK;
}

int les_get_sgs_init_flags(les_ctx* c, int* a, int* b) {
    if (!c || !a || !b) return LES_E_BADARG;
    *a = c->f_lm_mm_init;
    *b = c->f_qn_nn_init;
    return LES_OK;
}

int les_ddx(les_ctx* c, int d, int f) { return deriv_entry(c, d, f, D_X); }
int les_ddy(les_ctx* c, int d, int f) { return deriv_entry(c, d, f, D_Y); }
int les_ddz_uv(les_ctx* c, int d, int f) { return deriv_entry(c, d, f, D_ZUV); }
int les_ddz_w(les_ctx* c, int d, int f) { return deriv_entry(c, d, f, D_ZW); }
int les_filter_data(les_ctx* c, int f) {
    if (!c || !valid_field(f)) return LES_E_BADARG;
    const FieldRef r = field_ref(c, f);
    if (r.f) filter_f<float>(c, r.f, 0, (int)c->nlev);
    else filter_f<double>(c, r.d, 0, (int)c->nlev);
    return finish(c);
}
int les_test_filter(les_ctx* c, int f, int level, int which) {
    if (!c || !valid_field(f) || level < 0 || level > c->nz || (which != 2 && which != 4)) return LES_E_BADARG;
    const FieldRef r = field_ref(c, f);
    const double* G = which == 2 ? c->Gtest : c->Gtest2;
    if (r.f) test_filter_f<float>(c, r.f + (size_t)level * c->plane, 1, G);
    else test_filter_f<double>(c, r.d + (size_t)level * c->plane, 1, G);
    return finish(c);
}

int les_fftn(les_ctx* c, const double* in, int ld, double* out_c) {
    if (!c || !in || !out_c || ld < c->nx) return LES_E_BADARG;
    double* w = c->work3[7];
    les_note(c, hipMemcpy2DAsync(w, (size_t)c->nx * 8, in, (size_t)ld * 8, (size_t)c->nx * 8, c->ny, hipMemcpyHostToDevice,
                     c->stream));
    cx<double>* sp = c->spec3[2];
    x_r2c<double>(c->stream, w, sp, c->ny, c->nx, c->lhx, c->tw_x, 1.0);
    yline_spec<double>(c->stream, sp, sp, 1, c->lhx, (long)c->ns, c->tw_y, SP_FWD, nullptr);
    les_note(c, hipMemcpyAsync(out_c, sp, c->ns * 16, hipMemcpyDeviceToHost, c->stream));
    return finish(c);
}

int les_ifftn(les_ctx* c, const double* in_c, double* out, int ld) {
    if (!c || !in_c || !out || ld < c->nx) return LES_E_BADARG;
    cx<double>* sp = c->spec3[2];
    les_note(c, hipMemcpyAsync(sp, in_c, c->ns * 16, hipMemcpyHostToDevice, c->stream));
    yline_spec<double>(c->stream, sp, sp, 1, c->lhx, (long)c->ns, c->tw_y, SP_INV, nullptr);
    double* w = c->work3[7];
    x_c2r<double>(c->stream, sp, w, c->ny, c->lhx, c->nx, c->tw_x, 0, 0.0);
    les_note(c, hipMemcpy2DAsync(out, (size_t)ld * 8, w, (size_t)c->nx * 8, (size_t)c->nx * 8, c->ny, hipMemcpyDeviceToHost,
                     c->stream));
    return finish(c);
}

int les_zero_padding(les_ctx* c, const double* var, int ld, double* big) {
    if (!c || !var || !big || ld < c->nx) return LES_E_BADARG;
    double* w = c->work3[7];
    les_note(c, hipMemcpy2DAsync(w, (size_t)c->nx * 8, var, (size_t)ld * 8, (size_t)c->nx * 8, c->ny, hipMemcpyHostToDevice,
                     c->stream));
    pad_f<double, double>(c, w, c->big3[0], 1);
    les_note(c, hipMemcpyAsync(big, c->big3[0], c->plane2 * 8, hipMemcpyDeviceToHost, c->stream));
    return finish(c);
}

// zero_padding of nplanes device planes with the intermediate in the
// library's static scratch: no device memory is allocated
int les_zero_padding_noalloc(les_ctx* c, const double* d_var, int ld, double* d_big, int nplanes) {
    if (!c || !d_var || !d_big || ld < c->nx || nplanes < 1) return LES_E_BADARG;
    if ((size_t)nplanes * c->ny * c->nx2 > NOALLOC_DOUBLES) return LES_E_SIZE;
    double* tmp = c->noalloc;
    xline_real<double, double>(c->stream, d_var, tmp, (long)nplanes * c->ny, ld, c->nx2, c->tw_x, c->tw_x2, OP_PAD,
                               0.0);
    yline_real<double>(c->stream, tmp, d_big, nplanes, c->nx2, c->nx2, c->nx2, (long)c->ny * c->nx2,
                       (long)c->plane2, c->tw_y, c->tw_y2, OP_PAD, 0.0);
    return finish(c);
}

int les_zero_unpadding(les_ctx* c, const double* big, double* var, int ld) {
    if (!c || !var || !big || ld < c->nx) return LES_E_BADARG;
    les_note(c, hipMemcpyAsync(c->big3[0], big, c->plane2 * 8, hipMemcpyHostToDevice, c->stream));
    double* w = c->work3[7];
    unpad_f<double>(c, c->big3[0], w, 1);
    les_note(c, hipMemcpy2DAsync(var, (size_t)ld * 8, w, (size_t)c->nx * 8, (size_t)c->nx * 8, c->ny, hipMemcpyDeviceToHost,
                     c->stream));
    return finish(c);
}

int les_obukhov(les_ctx* c, int jt) {
    if (!c) return LES_E_BADARG;
    obukhov_dev(c, jt);
    return finish(c);
}
int les_wallstress_bottom(les_ctx* c) {
    if (!c) return LES_E_BADARG;
    LES_DISPATCH(wallstress, c, 0);
    return finish(c);
}
int les_wallstress_top(les_ctx* c) {
    if (!c) return LES_E_BADARG;
    LES_DISPATCH(wallstress, c, 1);
    return finish(c);
}
int les_sgs_stag(les_ctx* c, int jt) {
    if (!c) return LES_E_BADARG;
    LES_DISPATCH(sgs_stag, c, jt);
    return finish(c);
}
int les_lagrange_sdep(les_ctx* c, int jt) {
    if (!c) return LES_E_BADARG;
    LES_DISPATCH(lagrange, c, jt);
    return finish(c);
}
int les_interpolag_sdep(les_ctx* c) {
    if (!c) return LES_E_BADARG;
    LES_DISPATCH(interpolag_dev, c);
    return finish(c);
}
int les_divstress_uv(les_ctx* c, int divt, int tx, int ty, int tz) { return divstress_entry(c, divt, tx, ty, tz, 0); }
int les_divstress_w(les_ctx* c, int divt, int tx, int ty, int tz) { return divstress_entry(c, divt, tx, ty, tz, 1); }
int les_convec(les_ctx* c) {
    if (!c) return LES_E_BADARG;
    LES_DISPATCH(convec, c);
    return finish(c);
}
int les_theta_all_in_one(les_ctx* c, int jt) {
    if (!c) return LES_E_BADARG;
    LES_DISPATCH(theta_all_in_one, c, jt);
    return finish(c);
}
int les_press_stag_array(les_ctx* c, int jt) {
    if (!c) return LES_E_BADARG;
    LES_DISPATCH(press_stag_array, c, jt);
    return finish(c);
}
// Admas_Bashforth (103_flosol.f90:319-389) on the device fields: Euler at
// jt = 1 unless restart_vel (RHS_f <- RHS, :331-338), level 0 <- BOGUS
int les_ab2(les_ctx* c, int jt) {
    if (!c) return LES_E_BADARG;
    if (jt == 1 && !c->prm.restart_vel) {
        const size_t n = c->plane * c->nlev * (mixed(c) ? 4 : 8);
        for (int m = 0; m < 3; ++m)
            les_note(c, hipMemcpyAsync(c->wk[LES_F_RHSX_F + m], c->wk[LES_F_RHSX + m], n, hipMemcpyDeviceToDevice,
                                       c->stream));
    }
    LES_DISPATCH(ab2_dev, c);
    return finish(c);
}
// The pressure TDMA alone (109_press_stag_array.f90:201-319): host half
// spectra Hx, Hy, Hz (levels 0..nzt, complex interleaved, kx fastest, lhx x
// nyt per level) -> p_hat (levels 0..nzt-1).  FP64.
int les_tdma_solve(les_ctx* c, const double* hx, const double* hy, const double* hz, double* ph) {
    if (!c || !hx || !hy || !hz || !ph) return LES_E_BADARG;
    const size_t bytes = c->ns * 16;
    les_note(c, hipMemcpyAsync(c->spec3[0], hx, bytes * c->nlev, hipMemcpyHostToDevice, c->stream));
    les_note(c, hipMemcpyAsync(c->spec3[1], hy, bytes * c->nlev, hipMemcpyHostToDevice, c->stream));
    les_note(c, hipMemcpyAsync(c->spec3[2], hz, bytes * c->nlev, hipMemcpyHostToDevice, c->stream));
    tdma_press<double>(c, c->spec3[0], c->spec3[1], c->spec3[2], c->spec3[3],
                       reinterpret_cast<double*>(c->spec3[3]), nullptr);
    les_note(c, hipMemcpyAsync(ph, c->spec3[3], bytes * c->nz, hipMemcpyDeviceToHost, c->stream));
    return finish(c);
}
int les_project(les_ctx* c) {
    if (!c) return LES_E_BADARG;
    LES_DISPATCH(project_dev, c);
    return finish(c);
}

int les_step(les_ctx* c, int jt) {
    int failed = 0;
    return les_run(c, jt, 1, &failed);
}

int les_run(les_ctx* c, int jt0, int nsteps, int* failed_jt) {
    if (!c || nsteps < 0 || !failed_jt) return LES_E_BADARG;
    *failed_jt = 0;
    int done = 0;
    while (done < nsteps) {
        const int chunk = (nsteps - done) < c->cfl_cap ? (nsteps - done) : c->cfl_cap;
        for (int s = 0; s < chunk; ++s) {
            // check_cfl of newstep (901_debug.f90:85-99), with sgs_stag's u_lag
            // accumulation in the same pass (u, v, w do not change in between)
            const bool lag = c->fuse && !c->sgs_frozen;
            // the sums read u, v, psi_m, sgs_t3 and theta at the wall level,
            // which the CFL pass leaves alone
            obukhov_sums_fork(c, jt0 + done + s);
            c->obk_early = 1;
            max3_dev(c, c->cfl_dev + s, lag);
            c->lag_done = lag;
            LES_DISPATCH(step_body, c, jt0 + done + s);
            c->lag_done = 0;
        }
        les_note(c, hipMemcpyAsync(c->cfl_host, c->cfl_dev, sizeof(double) * chunk, hipMemcpyDeviceToHost, c->stream));
        const int st = finish(c);
        if (st != LES_OK) return st;
        for (int s = 0; s < chunk; ++s) {
            c->cfl = c->cfl_host[s];
            if (c->cfl > 0.6) {   // CFL_max = 0.6 (901_debug.f90:80, :123)
                *failed_jt = jt0 + done + s;
                return LES_E_CFL;
            }
        }
        done += chunk;
    }
    return LES_OK;
}

int les_set_avgslice(les_ctx* c, int use_avgslice, int c_count, int p_count) {
    if (!c || c_count < 1 || p_count < 1) return LES_E_BADARG;
    const size_t F3 = c->plane * c->nlev;
    const size_t S = (size_t)LES_N_SLICES * (size_t)c->nx * (c->nz - 1);
    if (!c->st_run) {
        hipError_t e = hipSuccess;
        for (int i = 0; i < 3 && e == hipSuccess; ++i) e = dalloc(c, &c->st_diag[i], F3);
        if (e == hipSuccess) e = dalloc(c, &c->st_ke, F3);
        if (e == hipSuccess) e = dalloc(c, &c->st_tmp, F3);
        if (e == hipSuccess) e = dalloc(c, &c->st_dke, F3);
        if (e == hipSuccess) e = dalloc(c, &c->st_mean, 3 * c->nlev);
        double *run = nullptr, *win = nullptr;
        if (e == hipSuccess) e = dalloc(c, &run, S);
        if (e == hipSuccess) e = dalloc(c, &win, S);
        if (e != hipSuccess) return LES_E_NOMEM;
        c->st_run = run;
        c->st_win = win;
        for (int i = 0; i < 3; ++i) fill_levels(c->stream, c->st_diag[i], F3, 0, 1, 0.0);
        fill_levels(c->stream, c->st_dke, F3, 0, 1, 0.0);
    }
    fill_levels(c->stream, c->st_run, S, 0, 1, 0.0);
    fill_levels(c->stream, c->st_win, S, 0, 1, 0.0);
    c->st_use = use_avgslice;
    c->st_c = c_count;
    c->st_p = p_count;
    c->st_fr = (1.0 / (double)p_count) * (double)c_count;   // 102_initial.f90:170
    return finish(c);
}

int les_get_avgslice(les_ctx* c, int slice, int window, double* host) {
    if (!c || !host || !c->st_run || slice < 0 || slice >= LES_N_SLICES || (window != 0 && window != 1))
        return LES_E_BADARG;
    const size_t L = (size_t)c->nx * (c->nz - 1);
    les_note(c, hipMemcpyAsync(host, (window ? c->st_win : c->st_run) + slice * L, sizeof(double) * L,
                               hipMemcpyDeviceToHost, c->stream));
    return finish(c);
}

int les_get_cfl(les_ctx* c, double* cfl) {
    if (!c || !cfl) return LES_E_BADARG;
    *cfl = c->cfl;
    return LES_OK;
}

int les_time_kernel(les_ctx* c, int kernel, int reps, double* ms, double* bytes) {
    if (!c || !ms || !bytes || reps < 1 || kernel < 0 || kernel > 4) return LES_E_BADARG;
    const size_t F3 = c->plane * c->nlev;
    const int nz = c->nz;
    const bool m = mixed(c);
    const double ws = m ? 4.0 : 8.0;  // work-field word size
    double* a = c->work3[0];
    double* b = c->work3[1];
    auto launch = [&]() {
        switch (kernel) {
            case LES_K_TDMA:
                if (m)
                    tdma_press<float>(c, scr<float>(c->spec3[0]), scr<float>(c->spec3[1]), scr<float>(c->spec3[2]),
                                      c->phf, reinterpret_cast<double*>(c->spec3[3]), c->hzm);
                else
                    tdma_press<double>(c, c->spec3[0], c->spec3[1], c->spec3[2], c->spec3[3],
                                       reinterpret_cast<double*>(c->spec3[3]), nullptr);
                break;
            case LES_K_XDERIV:
                if (m) ddx_f<float, float>(c, scr<float>(b), scr<float>(a), 0, (int)c->nlev);
                else ddx_f<double, double>(c, b, a, 0, (int)c->nlev);
                break;
            case LES_K_YDERIV:
                if (m) ddy_f<float, float>(c, scr<float>(b), scr<float>(a), 0, (int)c->nlev);
                else ddy_f<double, double>(c, b, a, 0, (int)c->nlev);
                break;
            case LES_K_PAD:
                if (m) pad_f<float, float>(c, scr<float>(a), scr<float>(c->big3[0]), (int)c->nlev);
                else pad_f<double, double>(c, a, c->big3[0], (int)c->nlev);
                break;
            default: les_note(c, hipMemcpyAsync(b, a, F3 * (size_t)ws, hipMemcpyDeviceToDevice, c->stream)); break;
        }
    };
    // algorithmic bytes: TDMA reads Hx, Hy (levels 1..nz-1) and Hz (0..nz),
    // writes p (0..nz-1); the mixed TDMA also reads the FP64 plane sums
    switch (kernel) {
        case LES_K_TDMA:
            *bytes = (double)c->ns * 2.0 * ws * (double)(2 * (nz - 1) + (nz + 1) + nz) + (m ? 8.0 * (nz + 1) : 0.0);
            break;
        case LES_K_PAD: *bytes = (double)c->nlev * ws * ((double)c->plane + (double)c->plane2); break;
        default: *bytes = 2.0 * (double)F3 * ws; break;
    }
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return LES_E_HIP;
    if (hipEventCreate(&e1) != hipSuccess) {
        les_note(c, hipEventDestroy(e0));
        return LES_E_HIP;
    }
    launch();  // warm
    les_note(c, hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; ++r) launch();
    les_note(c, hipEventRecord(e1, c->stream));
    const int st = finish(c);
    float t = 0.f;
    const hipError_t et = hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (et != hipSuccess) return LES_E_HIP;
    *ms = (double)t / reps;
    return st;
}

void* les_stream(les_ctx* c) { return c ? (void*)c->stream : nullptr; }

int les_sync(les_ctx* c) {
    if (!c) return LES_E_BADARG;
    return finish(c);
}

}  // extern "C"
