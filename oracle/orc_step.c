/*
 * ORACLE -- test infrastructure only (see orc_internal.h).
 *
 * FP64 CPU restatement of the boundary-layer LES timestep
 * (SBL/code/103_flosol.f90:4-290) in global-domain form: one context holds
 * levels 0..nzt, the bottom z-slab's rules apply at level 0/1 and the top
 * z-slab's at nzt-1/nzt, and every MPI ghost exchange becomes a read of the
 * neighbouring level.  This is the npz >= 2 behaviour of the reference;
 * with npz = 1 its top boundary rows differ (the coordz branches at
 * 103_flosol.f90:48-52, 109_press_stag_array.f90:20-50 skip the top).
 *
 * Each routine cites the reference lines it follows.  Quirks reproduced:
 *   - interpolag_Sdep top level uses the un-averaged u_lag(nz-1)
 *     (106_06_interpolag_Sdep.f90:195-199);
 *   - interpolag_Sdep overwrites zp(:,:,1) from zp(:,:,2) (:252-255);
 *   - the 1e-24 clip is the single-precision literal real(1.e-24, rprec)
 *     (106_03_lagrange_Sdep.f90:261-331).
 * Not reproduced (never read by the step): dTdz at the wall level, which the
 * reference divides by an uninitialised patch count (010_scalars.f90:349,
 * 008_bottombc.f90:37); padding columns nxt+1..ldx; the ke/dudt/dissip
 * diagnostics.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc_internal.h"

#define CLIP24 ((double)1.e-24f)   /* real(1.e-24, kind=rprec) */

static double* F(les_ctx* c, int id) { return c->f3[id]; }
int orc_destroy(les_ctx* c);

/* --------------------------------------------------------------------------
 * lifecycle
 * ------------------------------------------------------------------------ */
static int good_size(int n) {
    if (n < 4 || n % 2) return 0;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    return n == 1;
}

static void init_wavenumbers(les_ctx* c) {
    /* init_wavenumber, 002_fft.f90:121-189 */
    const int lhx = c->lhx, ny = c->ny;
    const double lx = c->prm.lx_tot, ly = c->prm.ly_tot;
    for (int i = 0; i < lhx; ++i) c->kx[i] = (double)i * (2.0 * ORC_PI / lx);
    for (int j = 0; j < c->lhy; ++j) c->ky[j] = (double)j * (2.0 * ORC_PI / ly);
    c->kx[lhx - 1] = 0.0;           /* :135-137 */
    c->ky[c->lhy - 1] = 0.0;        /* :138-140 */
    for (int j = 0; j < ny; ++j) {
        int m = ((j + ny / 2) % ny) - ny / 2;   /* :147-148 */
        for (int i = 0; i < lhx; ++i) {
            size_t q = (size_t)j * lhx + i;
            c->kx2d[q] = c->kx[i];
            c->ky2d[q] = (double)m * (2.0 * ORC_PI / ly);
            c->notnyq[q] = 1;
        }
    }
    for (int j = 0; j < ny; ++j) {         /* :157-160 */
        size_t q = (size_t)j * lhx + lhx - 1;
        c->ky2d[q] = 0.0;
        c->notnyq[q] = 0;
    }
    for (int i = 0; i < lhx; ++i) {        /* :161-165 */
        size_t q = (size_t)(ny / 2) * lhx + i;
        c->kx2d[q] = 0.0;
        c->ky2d[q] = 0.0;
        c->notnyq[q] = 0;
    }
    for (size_t q = 0; q < (size_t)lhx * ny; ++q)
        c->k2[q] = c->kx2d[q] * c->kx2d[q] + c->ky2d[q] * c->ky2d[q];
}

static void init_test_filter(les_ctx* c, double alpha, double* G) {
    /* init_test_filter, 006_test_filter.f90:41-77, ifilter = 1 */
    const double delta = alpha * sqrt(c->dx * c->dy);
    const double kc2 = (ORC_PI / delta) * (ORC_PI / delta);
    for (size_t q = 0; q < (size_t)c->lhx * c->ny; ++q) {
        G[q] = 1.0 / (double)(c->nx * c->ny);
        if (c->k2[q] >= kc2) G[q] = 0.0;
    }
}

static void init_coef_press(les_ctx* c) {
    /* init_coef_press + tdmaMcoef_pipelined, 109_press_stag_array.f90:4-103,
     * global form: bottom row at level 1, top row at level nzt-1.  Arrays
     * are [level 0..nz-1][ky][kx]. */
    const size_t ns = (size_t)c->lhx * c->ny;
    const int nz = c->nz;
    const double dz = c->dz, idz2 = 1.0 / (dz * dz);
    double *a = c->coef_a, *b = c->coef_b, *cc = c->coef_c;
    memset(a, 0, sizeof(double) * ns * nz);
    memset(b, 0, sizeof(double) * ns * nz);
    memset(cc, 0, sizeof(double) * ns * nz);
    for (size_t q = 0; q < ns; ++q) {
        const double k2 = c->k2[q];
        for (int k = 1; k <= nz - 1; ++k) {
            size_t o = (size_t)k * ns + q;
            if (k == 1) {                       /* :20-30 */
                if (k2 > 0) { a[o] = 0.0; b[o] = -(k2 + 1.0 / (dz * dz)); cc[o] = 1.0 / (dz * dz); }
                if (k2 == 0 && c->notnyq[q]) { a[o] = 0.0; b[o] = 1.0 / dz; }
            } else if (k == nz - 1) {           /* :31-43 */
                if (k2 > 0) { a[o] = 1.0 / (dz * dz); b[o] = -(k2 + 1.0 / (dz * dz)); cc[o] = 0.0; }
                if (k2 == 0 && c->notnyq[q]) { a[o] = -1.0 / dz; b[o] = 1.0 / dz; }
            } else {                            /* :51-62 */
                if (k2 > 0) { a[o] = idz2; b[o] = -(k2 + 2.0 / (dz * dz)); cc[o] = 1.0 / (dz * dz); }
                if (k2 == 0 && c->notnyq[q]) { a[o] = -1.0 / dz; b[o] = 1.0 / dz; }
            }
        }
        /* tdmaMcoef_pipelined :92-97 (coef_cp(0) = 0 at the bottom) */
        if (c->notnyq[q]) {
            for (int k = 1; k <= nz - 1; ++k) {
                size_t o = (size_t)k * ns + q, om = (size_t)(k - 1) * ns + q;
                b[o] = b[o] - a[o] * cc[om];
                cc[o] = cc[o] / b[o];
            }
        }
    }
    (void)idz2;
}

static void init_sponge(les_ctx* c) {
    /* setsponge, 007_topbc.f90:15-85, global level k; zuvp(k) = (k - 1/2) dz
     * (init_meshgrid, 102_initial.f90:275-278) */
    const les_params* P = &c->prm;
    const int nzt = c->nz;
    memset(c->sponge, 0, sizeof(double) * c->nlev);
    if (P->ubc != 1) return;
    if (P->damping_method == 1) {                                /* :26-36 */
        double factor = 9.0 / (double)(nzt - 3 * nzt / 4 + 1);
        double sponge_top = P->z_i / (50.0 * P->u_scale);
        for (int k = 1; k <= nzt - 1; ++k)
            if (k > 3 * nzt / 4 + 1) c->sponge[k] = sponge_top * pow(5.0, (double)(k - nzt) * factor);
    } else if (P->damping_method == 2) {                         /* :37-50 */
        const double z_d = 0.75 * P->lz_tot, cfrdmp = 3.9;
        for (int k = 1; k <= nzt - 1; ++k) {
            const double z = ((double)k - 0.5) * c->dz;
            if (z >= z_d && z <= P->lz_tot)
                c->sponge[k] = cfrdmp * 0.5 * (1.0 - cos(ORC_PI * (z - z_d) / (P->lz_tot - z_d)));
        }
    } else {                                                     /* :55-82 */
        const double g_hat = ORC_G * (P->z_i / (P->u_scale * P->u_scale));
        double fac, bv;
        if (P->ocean_flag) {
            fac = (double)0.22f;                                 /* default-real literal */
            bv = sqrt(g_hat * P->alpha_w * P->inv_strength / P->T_scale * P->z_i);
        } else {
            fac = 1.5;
            bv = sqrt(g_hat * P->inv_strength / P->T_scale * P->z_i);
        }
        const double sigma_z = 0.8 * P->lz_tot;
        const double sigma_max = fac * bv / (double)(8.0f * atanf(1.0f));   /* 8.0*atan(1.0), default real */
        for (int k = 1; k <= nzt; ++k) {
            const double z = ((double)k - 0.5) * c->dz;
            if (z >= sigma_z && z <= P->lz_tot) {
                const double ell = (z - sigma_z) / (P->lz_tot - sigma_z);
                c->sponge[k] = ell * ell * sigma_max;
            }
        }
    }
}

/* init_stokes_drift, 014_stokes_drift.f90:29-115 (monochromatic, its
 * shallow-water form under a wall top, and 'dynamic' with the values of
 * gwave.in's first record); zuvp(k) = (k - 1/2) dz */
static int init_stokes(les_ctx* c) {
    const les_params* P = &c->prm;
    memset(c->ust, 0, sizeof(double) * c->nlev);
    memset(c->vst, 0, sizeof(double) * c->nlev);
    if (!P->stokes_flag) return LES_OK;
    if (P->stokes_action != LES_STOKES_MONO && P->stokes_action != LES_STOKES_DYNAMIC) return LES_E_BADARG;
    const double rad = P->agl_ust / 180.0 * ORC_PI;              /* deg2rad, :300-307 */
    const double amp = P->amp_w / P->z_i, lam = P->lambda_w / P->z_i;
    const double wn = 2.0 * ORC_PI / lam;
    const double gnd = ORC_G * P->z_i / (P->u_scale * P->u_scale);
    if (P->ubc_mom == LES_BC_STRESS_FREE) {                      /* monochromatic, :146-160 */
        const double omega = sqrt(gnd * wn);
        const double us = omega * wn * amp * amp;
        for (int k = 0; k <= c->nz; ++k) {                       /* :59-63 */
            const double z = ((double)k - 0.5) * c->dz;
            c->ust[k] = us * exp(-2.0 * wn * z) * cos(rad);
            c->vst[k] = us * exp(-2.0 * wn * z) * sin(rad);
        }
    } else {                                                     /* mono_shallow_water, :164-178 */
        const double H = P->lz_tot;
        const double omega = sqrt(gnd * wn * tanh(wn * H));
        const double us = omega * wn * amp * amp / tanh(wn * H);
        for (int k = 0; k <= c->nz; ++k) {                       /* :76-81 */
            const double z = ((double)k - 0.5) * c->dz;
            const double sh = sinh(wn * H);
            c->ust[k] = us * cosh(2.0 * wn * (-z + H)) / 2.0 / (sh * sh) * cos(rad);
            c->vst[k] = us * cosh(2.0 * wn * (-z + H)) / 2.0 / (sh * sh) * sin(rad);
        }
    }
    return LES_OK;
}

int orc_create(const les_params* prm, les_ctx** out) {
    *out = NULL;
    if (!good_size(prm->nxt) || !good_size(prm->nyt) || prm->nzt < 4) return LES_E_BADARG;
    if (3 * prm->nxt / 2 > ORC_NMAX || 3 * prm->nyt / 2 > ORC_NMAX) return LES_E_BADARG;
    if (prm->model != 5 || !prm->sgs || prm->cs_count < 1) return LES_E_BADARG;
    if (prm->lbc != 0 && prm->lbc != 1) return LES_E_BADARG;
    /* lbc = 0 reads ustar_avg and psi_h, which the ocean branch never sets on
     * the surface ranks (010_scalars.f90:380-390, :696-716) */
    if (prm->lbc == 0 && prm->ocean_flag) return LES_E_BADARG;
    if (prm->damping_method < 1 || prm->damping_method > 3) return LES_E_BADARG;
    les_ctx* c = (les_ctx*)calloc(1, sizeof(les_ctx));
    c->prm = *prm;
    c->nx = prm->nxt; c->ny = prm->nyt; c->nz = prm->nzt;
    c->nx2 = 3 * c->nx / 2; c->ny2 = 3 * c->ny / 2;
    c->lhx = c->nx / 2 + 1; c->lhy = c->ny / 2 + 1;
    c->plane = (size_t)c->nx * c->ny;
    c->plane2 = (size_t)c->nx2 * c->ny2;
    c->nlev = (size_t)c->nz + 1;
    c->dx = prm->lx_tot / c->nx;            /* 102_initial.f90:137-139 */
    c->dy = prm->ly_tot / c->ny;
    c->dz = prm->lz_tot / (c->nz - 1);
    orc_plan_init(&c->px, c->nx);
    orc_plan_init(&c->py, c->ny);
    orc_plan_init(&c->px2, c->nx2);
    orc_plan_init(&c->py2, c->ny2);
    const size_t ns = (size_t)c->lhx * c->ny;
    c->kx = (double*)calloc(c->lhx, sizeof(double));
    c->ky = (double*)calloc(c->lhy, sizeof(double));
    c->kx2d = (double*)calloc(ns, sizeof(double));
    c->ky2d = (double*)calloc(ns, sizeof(double));
    c->k2 = (double*)calloc(ns, sizeof(double));
    c->notnyq = (unsigned char*)calloc(ns, 1);
    c->Gtest = (double*)calloc(ns, sizeof(double));
    c->Gtest2 = (double*)calloc(ns, sizeof(double));
    c->coef_a = (double*)calloc(ns * c->nz, sizeof(double));
    c->coef_b = (double*)calloc(ns * c->nz, sizeof(double));
    c->coef_c = (double*)calloc(ns * c->nz, sizeof(double));
    c->ust = (double*)calloc(c->nlev, sizeof(double));
    c->vst = (double*)calloc(c->nlev, sizeof(double));
    c->sponge = (double*)calloc(c->nlev, sizeof(double));
    for (int f = 0; f < LES_N_FIELDS3D; ++f) c->f3[f] = (double*)calloc(c->plane * c->nlev, sizeof(double));
    c->ustar_avg = (double*)calloc(c->plane, sizeof(double));
    c->psi_m = (double*)calloc(c->plane, sizeof(double));
    c->phi_m = (double*)calloc(c->plane, sizeof(double));
    c->psi_h = (double*)calloc(c->plane, sizeof(double));
    c->phi_h = (double*)calloc(c->plane, sizeof(double));
    for (size_t q = 0; q < c->plane; ++q) c->phi_m[q] = c->phi_h[q] = 1.0;
    c->u_big = (double*)calloc(c->plane2 * c->nlev, sizeof(double));
    c->v_big = (double*)calloc(c->plane2 * c->nlev, sizeof(double));
    c->w_big = (double*)calloc(c->plane2 * c->nlev, sizeof(double));
    init_wavenumbers(c);
    init_test_filter(c, 2.0, c->Gtest);      /* 102_initial.f90:39 */
    init_test_filter(c, 4.0, c->Gtest2);     /* 102_initial.f90:42 */
    init_coef_press(c);
    init_sponge(c);
    c->ustar_top = (double*)calloc(c->plane, sizeof(double));
    c->wind_ustar = prm->u_star;                  /* 102_initial.f90:201, 246 */
    c->wind_rad = prm->rad_stress;
    if (init_stokes(c) != LES_OK) {
        orc_destroy(c);
        return LES_E_BADARG;
    }
    *out = c;
    return LES_OK;
}

int orc_destroy(les_ctx* c) {
    if (!c) return LES_OK;
    orc_plan_free(&c->px); orc_plan_free(&c->py);
    orc_plan_free(&c->px2); orc_plan_free(&c->py2);
    free(c->kx); free(c->ky); free(c->kx2d); free(c->ky2d); free(c->k2); free(c->notnyq);
    free(c->Gtest); free(c->Gtest2);
    free(c->coef_a); free(c->coef_b); free(c->coef_c);
    free(c->ust); free(c->vst); free(c->sponge);
    for (int f = 0; f < LES_N_FIELDS3D; ++f) free(c->f3[f]);
    free(c->ustar_avg); free(c->psi_m); free(c->phi_m); free(c->psi_h); free(c->phi_h); free(c->ustar_top);
    free(c->u_big); free(c->v_big); free(c->w_big);
    free(c->dudt); free(c->dvdt); free(c->dwdt); free(c->ke); free(c->ke_temp); free(c->dke);
    free(c->sl_run); free(c->sl_win);
    free(c);
    return LES_OK;
}

int orc_set_stokes(les_ctx* c, const double* ust, const double* vst) {
    memcpy(c->ust, ust, sizeof(double) * c->nlev);
    memcpy(c->vst, vst, sizeof(double) * c->nlev);
    return LES_OK;
}

int orc_set_sponge(les_ctx* c, const double* sp) {
    memcpy(c->sponge, sp, sizeof(double) * c->nlev);
    return LES_OK;
}

int orc_get_profiles(les_ctx* c, double* ust, double* vst, double* sponge) {
    if (ust) memcpy(ust, c->ust, sizeof(double) * c->nlev);
    if (vst) memcpy(vst, c->vst, sizeof(double) * c->nlev);
    if (sponge) memcpy(sponge, c->sponge, sizeof(double) * c->nlev);
    return LES_OK;
}

int orc_set_wind_stress(les_ctx* c, double ustar, double agl_deg) {
    c->wind_ustar = ustar;
    c->wind_rad = agl_deg / 180.0 * ORC_PI;       /* deg2rad, 014_stokes_drift.f90:300-307 */
    return LES_OK;
}

static double* surf_field(les_ctx* c, int id) {
    switch (id) {
        case LES_F_USTAR_SFC: return c->ustar_avg;
        case LES_F_PSI_M_SFC: return c->psi_m;
        case LES_F_PHI_M_SFC: return c->phi_m;
        case LES_F_PSI_H_SFC: return c->psi_h;
        case LES_F_PHI_H_SFC: return c->phi_h;
        case LES_F_USTAR_TOP_SFC: return c->ustar_top;
        default: return NULL;
    }
}

int orc_set_field(les_ctx* c, int id, const double* h, int ld, int lev0, int nlev) {
    double* d;
    if (id >= 0 && id < LES_N_FIELDS3D) {
        if (lev0 < 0 || lev0 + nlev > (int)c->nlev) return LES_E_BADARG;
        d = c->f3[id];
    } else {
        d = surf_field(c, id);
        if (!d || lev0 != 0 || nlev != 1) return LES_E_BADARG;
    }
    if (ld < c->nx) return LES_E_BADARG;
    for (int k = 0; k < nlev; ++k)
        for (int j = 0; j < c->ny; ++j)
            memcpy(d + IDX(c, lev0 + k, j, 0), h + ((size_t)k * c->ny + j) * ld, sizeof(double) * c->nx);
    return LES_OK;
}

int orc_get_field(les_ctx* c, int id, double* h, int ld, int lev0, int nlev) {
    double* d;
    if (id >= 0 && id < LES_N_FIELDS3D) {
        if (lev0 < 0 || lev0 + nlev > (int)c->nlev) return LES_E_BADARG;
        d = c->f3[id];
    } else {
        d = surf_field(c, id);
        if (!d || lev0 != 0 || nlev != 1) return LES_E_BADARG;
    }
    if (ld < c->nx) return LES_E_BADARG;
    for (int k = 0; k < nlev; ++k)
        for (int j = 0; j < c->ny; ++j)
            memcpy(h + ((size_t)k * c->ny + j) * ld, d + IDX(c, lev0 + k, j, 0), sizeof(double) * c->nx);
    return LES_OK;
}

/* les_bind_field / les_upload / les_download: the caller's arrays are
 * registered once and synchronised in bulk (no reference counterpart; the
 * reference's module arrays are these host arrays) */
int orc_bind_field(les_ctx* c, int id, double* h, int ld, int lev0, int nlev) {
    if (!h || c->nbind >= 128) return LES_E_BADARG;
    double* probe = (id >= 0 && id < LES_N_FIELDS3D) ? c->f3[id] : surf_field(c, id);
    if (!probe || ld < c->nx) return LES_E_BADARG;
    if (id >= 0 && id < LES_N_FIELDS3D ? (lev0 < 0 || nlev < 0 || lev0 + nlev > (int)c->nlev) : (lev0 != 0 || nlev != 1))
        return LES_E_BADARG;
    c->bind[c->nbind].id = id; c->bind[c->nbind].h = h; c->bind[c->nbind].ld = ld;
    c->bind[c->nbind].lev0 = lev0; c->bind[c->nbind].nlev = nlev;
    ++c->nbind;
    return LES_OK;
}
int orc_upload(les_ctx* c) {
    for (int i = 0; i < c->nbind; ++i) {
        int st = orc_set_field(c, c->bind[i].id, c->bind[i].h, c->bind[i].ld, c->bind[i].lev0, c->bind[i].nlev);
        if (st != LES_OK) return st;
    }
    return LES_OK;
}
int orc_download(les_ctx* c) {
    for (int i = 0; i < c->nbind; ++i) {
        int st = orc_get_field(c, c->bind[i].id, c->bind[i].h, c->bind[i].ld, c->bind[i].lev0, c->bind[i].nlev);
        if (st != LES_OK) return st;
    }
    return LES_OK;
}

int orc_set_sgs_frozen(les_ctx* c, int frozen) { c->sgs_frozen = frozen ? 1 : 0; return LES_OK; }
int orc_set_sgs_init_flags(les_ctx* c, int a, int b) { c->f_lm_mm_init = a; c->f_qn_nn_init = b; return LES_OK; }
int orc_get_sgs_init_flags(les_ctx* c, int* a, int* b) { *a = c->f_lm_mm_init; *b = c->f_qn_nn_init; return LES_OK; }
int orc_get_cfl(les_ctx* c, double* cfl) { *cfl = c->cfl; return LES_OK; }

/* --------------------------------------------------------------------------
 * derivatives (004_derivatives.f90) and filters (006_test_filter.f90)
 * ------------------------------------------------------------------------ */
static void ddx_arr(les_ctx* c, double* d, const double* f) {
    #pragma omp parallel for schedule(static)
    for (int k = 0; k <= c->nz; ++k) orc_ddx_plane(c, LEV(c, d, k), LEV(c, f, k));
}
static void ddy_arr(les_ctx* c, double* d, const double* f) {
    #pragma omp parallel for schedule(static)
    for (int k = 0; k <= c->nz; ++k) orc_ddy_plane(c, LEV(c, d, k), LEV(c, f, k));
}
/* ddz_uv 004_derivatives.f90:85-116: level 0 BOGUS, 2..nzt-1 FD; level 1
 * (bottom) and nzt (top) are left to the wall routines. */
static void ddz_uv_arr(les_ctx* c, double* d, const double* f) {
    const double cst = 1.0 / c->dz;
    for (size_t q = 0; q < c->plane; ++q) d[q] = ORC_BOGUS;
    for (int k = 2; k <= c->nz - 1; ++k) {
        double* dk = LEV(c, d, k);
        const double *fk = LEV(c, f, k), *fm = LEV(c, f, k - 1);
        for (size_t q = 0; q < c->plane; ++q) dk[q] = cst * (fk[q] - fm[q]);
    }
}
/* ddz_w 004_derivatives.f90:155-183: 0..nzt-1 FD, 0 BOGUS, nzt zero. */
static void ddz_w_arr(les_ctx* c, double* d, const double* f) {
    const double cst = 1.0 / c->dz;
    for (int k = 0; k <= c->nz - 1; ++k) {
        double* dk = LEV(c, d, k);
        const double *fk = LEV(c, f, k + 1), *fm = LEV(c, f, k);
        for (size_t q = 0; q < c->plane; ++q) dk[q] = cst * (fk[q] - fm[q]);
    }
    for (size_t q = 0; q < c->plane; ++q) { d[q] = ORC_BOGUS; LEV(c, d, c->nz)[q] = 0.0; }
}
/* filter_data 006_test_filter.f90:82-105 */
static void filter_arr(les_ctx* c, double* f) {
    const double cst = 1.0 / (double)(c->nx * c->ny);
    const size_t ns = (size_t)c->lhx * c->ny;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k <= c->nz; ++k) {
        double* fk = LEV(c, f, k);
        cplx* tmp = (cplx*)malloc(sizeof(cplx) * ns);
        for (size_t q = 0; q < c->plane; ++q) fk[q] = cst * fk[q];
        orc_fftn_plane(c, fk, tmp);
        orc_ifftn_plane(c, tmp, fk);
        free(tmp);
    }
}

int orc_ddx(les_ctx* c, int d, int f) { ddx_arr(c, F(c, d), F(c, f)); return LES_OK; }
int orc_ddy(les_ctx* c, int d, int f) { ddy_arr(c, F(c, d), F(c, f)); return LES_OK; }
int orc_ddz_uv(les_ctx* c, int d, int f) { ddz_uv_arr(c, F(c, d), F(c, f)); return LES_OK; }
int orc_ddz_w(les_ctx* c, int d, int f) { ddz_w_arr(c, F(c, d), F(c, f)); return LES_OK; }
int orc_filter_data(les_ctx* c, int f) { filter_arr(c, F(c, f)); return LES_OK; }
int orc_test_filter(les_ctx* c, int f, int level, int which) {
    if (level < 0 || level > c->nz || (which != 2 && which != 4)) return LES_E_BADARG;
    orc_test_filter_plane(c, LEV(c, F(c, f), level), which == 2 ? c->Gtest : c->Gtest2);
    return LES_OK;
}

/* one-plane host-array transforms */
static void plane_in(les_ctx* c, double* d, const double* h, int ld) {
    for (int j = 0; j < c->ny; ++j) memcpy(d + (size_t)j * c->nx, h + (size_t)j * ld, sizeof(double) * c->nx);
}
static void plane_out(les_ctx* c, double* h, const double* d, int ld) {
    for (int j = 0; j < c->ny; ++j) memcpy(h + (size_t)j * ld, d + (size_t)j * c->nx, sizeof(double) * c->nx);
}
int orc_fftn(les_ctx* c, const double* in, int ld, double* out_c) {
    double* t = (double*)malloc(sizeof(double) * c->plane);
    plane_in(c, t, in, ld);
    orc_fftn_plane(c, t, (cplx*)out_c);
    free(t);
    return LES_OK;
}
int orc_ifftn(les_ctx* c, const double* in_c, double* out, int ld) {
    double* t = (double*)malloc(sizeof(double) * c->plane);
    orc_ifftn_plane(c, (const cplx*)in_c, t);
    plane_out(c, out, t, ld);
    free(t);
    return LES_OK;
}
int orc_zero_padding(les_ctx* c, const double* var, int ld, double* big) {
    double* t = (double*)malloc(sizeof(double) * c->plane);
    plane_in(c, t, var, ld);
    orc_pad_plane(c, t, big);
    free(t);
    return LES_OK;
}
int orc_zero_unpadding(les_ctx* c, const double* big, double* var, int ld) {
    double* t = (double*)malloc(sizeof(double) * c->plane);
    orc_unpad_plane(c, big, t);
    plane_out(c, var, t, ld);
    free(t);
    return LES_OK;
}

/* --------------------------------------------------------------------------
 * obukhov, 010_scalars.f90:675-856 (land branch, patch_flag = 1,
 * num_patch = 1, so the homogeneous branch at :768-775)
 * ------------------------------------------------------------------------ */
int orc_obukhov(les_ctx* c, int jt) {
    const les_params* P = &c->prm;
    const size_t np = c->plane;
    const double* u1 = LEV(c, F(c, LES_F_U), 1);
    const double* v1 = LEV(c, F(c, LES_F_V), 1);
    const double nn = (double)(c->nx * c->ny);
    const double dz = c->dz;
    if (P->ocean_flag) {                                       /* :696-716 */
        /* the top wall's ranks set their own ustar_avg (to ustar_top here)
         * and psi/phi; the surface ranks' planes are left as they are */
        if (P->ubc_mom != LES_BC_WALL) return LES_OK;
        const double* ut = LEV(c, F(c, LES_F_U), c->nz - 1);
        const double* vt = LEV(c, F(c, LES_F_V), c->nz - 1);
        double su = 0.0;
        for (size_t q = 0; q < np; ++q) su += sqrt(ut[q] * ut[q] + vt[q] * vt[q]);
        const double nf = (double)(float)(c->nx * c->ny);      /* float(nxt*nynpy) */
        double us;
        if (P->walltype == LES_WALL_SMOOTH)
            us = su / nf / (log(0.5 * dz / (P->nu_molec / P->u_scale / P->z_i)) / ORC_VONK + ORC_B_SMOOTH);
        else
            us = su / nf * ORC_VONK / log(0.5 * dz / P->zo);
        for (size_t q = 0; q < np; ++q) c->ustar_top[q] = us;
        return LES_OK;
    }
    if (!P->theta_flag || jt < P->theta_init_time) {          /* :718-738 */
        if (!P->restart_theta) for (size_t q = 0; q < np; ++q) c->psi_m[q] = 0.0;
        double su = 0.0, sp = 0.0;
        for (size_t q = 0; q < np; ++q) su += sqrt(u1[q] * u1[q] + v1[q] * v1[q]);
        for (size_t q = 0; q < np; ++q) sp += c->psi_m[q];
        for (size_t q = 0; q < np; ++q) {
            c->phi_m[q] = 1.0; c->psi_h[q] = 0.0; c->phi_h[q] = 1.0;
            if (P->walltype == LES_WALL_SMOOTH)
                c->ustar_avg[q] = su / nn / (log(0.5 * dz / (P->nu_molec / P->u_scale / P->z_i)) / ORC_VONK + ORC_B_SMOOTH);
            else
                c->ustar_avg[q] = su / nn * ORC_VONK / (log(0.5 * dz / P->zo) - sp / nn);
        }
        return LES_OK;
    }
    double* sgs_t3_1 = LEV(c, F(c, LES_F_SGS_T3), 1);
    if (jt == P->theta_init_time && !P->restart_theta)         /* :745-749 */
        for (size_t q = 0; q < np; ++q) sgs_t3_1[q] = P->wt_s / P->u_scale / P->T_scale;
    const double g_ = ORC_G / (P->u_scale * P->u_scale / P->z_i);   /* :758 */
    double* tmp = (double*)malloc(sizeof(double) * np);
    for (size_t q = 0; q < np; ++q) tmp[q] = log(P->zo);
    double zo_ = exp(orc_hor_avg(c, tmp));                   /* :762-763 */
    double wt_ = orc_hor_avg(c, sgs_t3_1);                   /* :765 */
    for (size_t q = 0; q < np; ++q) tmp[q] = sqrt(u1[q] * u1[q] + v1[q] * v1[q]);
    double U_tmp = orc_hor_avg(c, tmp);
    double psi_m_tmp = orc_hor_avg(c, c->psi_m);
    double ustar_ = U_tmp * ORC_VONK / (log(0.5 * dz / zo_) - psi_m_tmp);   /* :770 */
    double theta_ = orc_hor_avg(c, LEV(c, F(c, LES_F_THETA), 1));
    free(tmp);
    for (size_t q = 0; q < np; ++q) {                          /* :793-843 */
        double wt_avg = wt_, us = ustar_, th = theta_;
        c->ustar_avg[q] = us;
        double L = -(us * us * us) / (ORC_VONK * g_ / th * wt_avg);
        if (L < 0.0 && wt_avg != 0.0) {
            double x = pow(1.0 - 16.0 * dz / 2.0 / L, 0.25);
            c->psi_m[q] = 2.0 * log((1.0 + x) / 2.0) + log((1.0 + x * x) / 2.0) - 2.0 * atan(x) + ORC_PI / 2.0;
            c->psi_h[q] = 2.0 * log((1.0 + x * x) / 2.0);
            c->phi_m[q] = 1.0 / x;
            c->phi_h[q] = 1.0 / (x * x);
        } else if (L > 0.0 && wt_avg != 0.0) {
            double zeta = 0.5 * dz / L;
            c->phi_m[q] = 1.0 + 4.8 * zeta;
            c->phi_h[q] = 1.0 + 7.8 * zeta;
            c->psi_m[q] = -1.0 * 4.8 * zeta;
            c->psi_h[q] = -1.0 * 7.8 * zeta;
        } else {
            c->psi_m[q] = 0.0; c->psi_h[q] = 0.0; c->phi_m[q] = 1.0; c->phi_h[q] = 1.0;
        }
    }
    return LES_OK;
}

/* --------------------------------------------------------------------------
 * wallstress_bottom / wallstress_top, 104_wallstress.f90:98-260
 * ------------------------------------------------------------------------ */
int orc_wallstress_bottom(les_ctx* c) {
    const les_params* P = &c->prm;
    const size_t np = c->plane;
    double* txz = LEV(c, F(c, LES_F_TXZ), 1);
    double* tyz = LEV(c, F(c, LES_F_TYZ), 1);
    double* dudz = LEV(c, F(c, LES_F_DUDZ), 1);
    double* dvdz = LEV(c, F(c, LES_F_DVDZ), 1);
    if (P->ocean_flag) {                                        /* :118-136 */
        const double us = c->wind_ustar, rad = c->wind_rad;
        const double* dudz2 = LEV(c, F(c, LES_F_DUDZ), 2);
        const double* dvdz2 = LEV(c, F(c, LES_F_DVDZ), 2);
        for (size_t q = 0; q < np; ++q) {
            txz[q] = us * us * cos(rad);
            tyz[q] = us * us * sin(rad);
            const double a = dudz2[q] * cos(rad) + dvdz2[q] * sin(rad);
            dudz[q] = a * cos(rad);
            dvdz[q] = a * sin(rad);
        }
        return LES_OK;
    }
    if (P->lbc_mom == LES_BC_STRESS_FREE) {                     /* :178-182 */
        for (size_t q = 0; q < np; ++q) txz[q] = tyz[q] = dudz[q] = dvdz[q] = 0.0;
        return LES_OK;
    }
    const double* u = LEV(c, F(c, LES_F_U), 1);
    const double* v = LEV(c, F(c, LES_F_V), 1);
    double* u1 = (double*)malloc(sizeof(double) * np);
    double* v1 = (double*)malloc(sizeof(double) * np);
    memcpy(u1, u, sizeof(double) * np);
    memcpy(v1, v, sizeof(double) * np);
    orc_test_filter_plane(c, u1, c->Gtest);                     /* :143-144 */
    orc_test_filter_plane(c, v1, c->Gtest);
    const double dz = c->dz, d0 = P->d0;
    for (size_t q = 0; q < np; ++q) {
        double denom;
        if (P->walltype == LES_WALL_SMOOTH)
            denom = log(0.5 * dz / (P->nu_molec / P->u_scale / P->z_i)) / ORC_VONK + ORC_B_SMOOTH;
        else
            denom = (log((0.5 * dz - d0) / P->zo) - c->psi_m[q]) / ORC_VONK;   /* :150 */
        double u_avg = sqrt(u1[q] * u1[q] + v1[q] * v1[q]);     /* :155 */
        double us = u_avg / denom;
        c->ustar_avg[q] = us;
        double cst = -(us * us) / u_avg;                         /* :160 */
        txz[q] = cst * u1[q];
        tyz[q] = cst * v1[q];
        double du = us / ((0.5 * dz - d0) * ORC_VONK) * u[q] / u_avg * c->phi_m[q];
        double dv = us / ((0.5 * dz - d0) * ORC_VONK) * v[q] / u_avg * c->phi_m[q];
        dudz[q] = (u[q] == 0.0) ? 0.0 : du;                     /* :174-175 */
        dvdz[q] = (v[q] == 0.0) ? 0.0 : dv;
    }
    free(u1); free(v1);
    return LES_OK;
}

int orc_wallstress_top(les_ctx* c) {
    const les_params* P = &c->prm;
    const size_t np = c->plane;
    const int nz = c->nz;
    double* txz = LEV(c, F(c, LES_F_TXZ), nz);
    double* tyz = LEV(c, F(c, LES_F_TYZ), nz);
    double* dudz = LEV(c, F(c, LES_F_DUDZ), nz);
    double* dvdz = LEV(c, F(c, LES_F_DVDZ), nz);
    if (P->ubc_mom == LES_BC_STRESS_FREE) {                     /* :250-254 */
        for (size_t q = 0; q < np; ++q) txz[q] = tyz[q] = dudz[q] = dvdz[q] = 0.0;
        return LES_OK;
    }
    const double* u = LEV(c, F(c, LES_F_U), nz - 1);
    const double* v = LEV(c, F(c, LES_F_V), nz - 1);
    double* u1 = (double*)malloc(sizeof(double) * np);
    double* v1 = (double*)malloc(sizeof(double) * np);
    memcpy(u1, u, sizeof(double) * np);
    memcpy(v1, v, sizeof(double) * np);
    orc_test_filter_plane(c, u1, c->Gtest);                     /* :219-220 */
    orc_test_filter_plane(c, v1, c->Gtest);
    const double dz = c->dz, d0 = P->d0;
    for (size_t q = 0; q < np; ++q) {
        double denom;
        if (P->walltype == LES_WALL_SMOOTH)
            denom = log(0.5 * dz / (P->nu_molec / P->u_scale / P->z_i)) / ORC_VONK + ORC_B_SMOOTH;
        else
            denom = log((0.5 * dz - d0) / P->zo) / ORC_VONK;     /* :226 */
        double u_avg = sqrt(u1[q] * u1[q] + v1[q] * v1[q]);
        double ta = u_avg / denom;
        double cst = -(ta * ta) / u_avg;
        txz[q] = -cst * u1[q];                                   /* :237-238 */
        tyz[q] = -cst * v1[q];
        double du = -ta / (0.5 * dz * ORC_VONK) * u[q] / u_avg;
        double dv = -ta / (0.5 * dz * ORC_VONK) * v[q] / u_avg;
        dudz[q] = (u[q] == 0.0) ? 0.0 : du;                     /* :246-247 */
        dvdz[q] = (v[q] == 0.0) ? 0.0 : dv;
    }
    free(u1); free(v1);
    return LES_OK;
}

/* --------------------------------------------------------------------------
 * interpolag_Sdep, 106_06_interpolag_Sdep.f90:4-336 (global form: periodic
 * x/y wrap, F(level 0) := F(1) at the bottom (:56-58), F(nzt+1) := F(nzt)
 * at the top (:69-72)).
 * ------------------------------------------------------------------------ */
static void interpolag_sdep(les_ctx* c) {
    const les_params* P = &c->prm;
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const size_t np = c->plane, nl = c->nlev;
    const double csc = (double)P->cs_count;
    double* u_lag = F(c, LES_F_U_LAG);
    double* v_lag = F(c, LES_F_V_LAG);
    double* w_lag = F(c, LES_F_W_LAG);
    double* ut = (double*)malloc(sizeof(double) * np * nl);
    double* vt = (double*)malloc(sizeof(double) * np * nl);
    for (size_t q = 0; q < np * nl; ++q) { ut[q] = u_lag[q] / csc; vt[q] = v_lag[q] / csc; }   /* :174-175 */
    /* top level first, reading the raw sum at nz-1 (:193-199) */
    for (size_t q = 0; q < np; ++q) {
        LEV(c, u_lag, nz)[q] = LEV(c, u_lag, nz - 1)[q] + c->ust[nz - 1];
        LEV(c, v_lag, nz)[q] = LEV(c, v_lag, nz - 1)[q] + c->vst[nz - 1];
    }
    for (size_t q = 0; q < np; ++q) {                            /* :181-187 */
        double a = LEV(c, ut, 1)[q];
        LEV(c, u_lag, 1)[q] = a + c->ust[1];
        double b = LEV(c, vt, 1)[q];
        LEV(c, v_lag, 1)[q] = b + c->vst[1];
    }
    for (int k = 2; k <= nz - 1; ++k) {                          /* :207-213 */
        for (size_t q = 0; q < np; ++q) {
            double a = 0.5 * (LEV(c, ut, k)[q] + LEV(c, ut, k - 1)[q]);
            LEV(c, u_lag, k)[q] = a + 0.5 * (c->ust[k] + c->ust[k - 1]);
            double b = 0.5 * (LEV(c, vt, k)[q] + LEV(c, vt, k - 1)[q]);
            LEV(c, v_lag, k)[q] = b + 0.5 * (c->vst[k] + c->vst[k - 1]);
        }
    }
    for (size_t q = 0; q < np * nl; ++q) w_lag[q] = w_lag[q] / csc;   /* :215 */
    for (size_t q = 0; q < np; ++q) {
        LEV(c, w_lag, 1)[q] = 0.25 * LEV(c, w_lag, 2)[q];        /* :221 */
        LEV(c, w_lag, nz)[q] = 0.25 * LEV(c, w_lag, nz - 1)[q];  /* :225 */
    }
    /* displacements (:230-232) and the wall-level zp fix (:239-257) */
    double* xp = ut;   /* reuse */
    double* yp = vt;
    double* zp = (double*)malloc(sizeof(double) * np * nl);
    for (int k = 1; k <= nz; ++k)
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q;
            xp[o] = -u_lag[o] * P->dt * csc / c->dx;
            yp[o] = -v_lag[o] * P->dt * csc / c->dy;
            zp[o] = -w_lag[o] * P->dt * csc / c->dz;
        }
    for (size_t q = 0; q < np; ++q) {
        double z2 = zp[2 * np + q], z1 = zp[np + q];
        z2 = z2 + fmin(z2, 0.0);
        z1 = z1 + fmax(z1, 0.0);
        (void)z1;
        z2 = fmin(1.0, z2);
        double z1n = fmin(1.0, z2);
        z2 = fmax(-1.0, z2);
        z1n = fmax(-1.0, z2);
        zp[2 * np + q] = z2;
        zp[np + q] = z1n;
    }
    /* trilinear interpolation from the upwind cube (:264-328) */
    double* FF[4] = {F(c, LES_F_F_LM), F(c, LES_F_F_MM), F(c, LES_F_F_QN), F(c, LES_F_F_NN)};
    double* old[4];
    for (int m = 0; m < 4; ++m) {
        old[m] = (double*)malloc(sizeof(double) * np * nl);
        memcpy(old[m], FF[m], sizeof(double) * np * nl);
    }
    #pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k) {
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t o = IDX(c, k, j, i);
                int ax = (copysign(1.0, xp[o]) > 0) ? 1 : -1;
                int ay = (copysign(1.0, yp[o]) > 0) ? 1 : -1;
                int az = (copysign(1.0, zp[o]) > 0) ? 1 : -1;
                int i2 = (i + ax + nx) % nx, j2 = (j + ay + ny) % ny;
                int k2 = k + az;
                if (k2 < 1) k2 = 1;
                if (k2 > nz) k2 = nz;
                double cx = fabs(xp[o]), cy = fabs(yp[o]), cz = fabs(zp[o]);
                double fx = 1.0 - cx, fy = 1.0 - cy, fz = 1.0 - cz;
                for (int m = 0; m < 4; ++m) {
                    const double* A = old[m];
                    #define AT(kk, jj, ii) A[IDX(c, kk, jj, ii)]
                    FF[m][o] = fx * fy * (AT(k, j, i) * fz + AT(k2, j, i) * cz)
                             + fx * cy * (AT(k, j2, i) * fz + AT(k2, j2, i) * cz)
                             + cx * fy * (AT(k, j, i2) * fz + AT(k2, j, i2) * cz)
                             + cx * cy * (AT(k, j2, i2) * fz + AT(k2, j2, i2) * cz);
                    #undef AT
                }
            }
    }
    for (int m = 0; m < 4; ++m) free(old[m]);
    memset(u_lag, 0, sizeof(double) * np * nl);                   /* :330-332 */
    memset(v_lag, 0, sizeof(double) * np * nl);
    memset(w_lag, 0, sizeof(double) * np * nl);
    free(ut); free(vt); free(zp);
}

/* --------------------------------------------------------------------------
 * lagrange_Sdep, 106_03_lagrange_Sdep.f90:6-362
 * ------------------------------------------------------------------------ */
static double smag(double s11, double s12, double s13, double s22, double s23, double s33) {
    return sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33 + 2.0 * (s12 * s12 + s13 * s13 + s23 * s23)));
}

static void lagrange_sdep(les_ctx* c, int jt) {
    const les_params* P = &c->prm;
    const int nz = c->nz;
    const size_t np = c->plane;
    const double delta = 1.0 * pow(c->dx * c->dy * c->dz, 1.0 / 3.0);   /* :47 */
    const double tf1 = 2.0, tf2 = 4.0, tf1_2 = tf1 * tf1, tf2_2 = tf2 * tf2;
    const double opftdelta = ORC_OPFTIME * delta;
    const double powcoeff = -1.0 / 8.0;
    const double lagran_dt = P->dt * (double)P->cs_count;
    const double cst = 2.0 * (delta * delta);
    const double bexp = log(tf1) / (log(tf2) - log(tf1));
    /* u_temp = u (:60-61) before interpolag zeroes nothing of u */
    const double* u = F(c, LES_F_U);
    const double* v = F(c, LES_F_V);
    const double* w = F(c, LES_F_W);
    interpolag_sdep(c);                                          /* :63 */
    const int init_lm = P->inilag && !c->f_lm_mm_init && (jt == P->cs_count || jt == P->dyn_init);
    const int init_qn = P->inilag && !c->f_qn_nn_init && (jt == P->cs_count || jt == P->dyn_init);
    #pragma omp parallel for schedule(dynamic)
    for (int jz = 1; jz <= nz; ++jz) {
        double* buf = (double*)malloc(sizeof(double) * np * 46);
        double* ub = buf + 0 * np; double* vb = buf + 1 * np; double* wb = buf + 2 * np;
        double* uh = buf + 3 * np; double* vh = buf + 4 * np; double* wh = buf + 5 * np;
        double* L[6]; double* Q[6]; double* Sb[6]; double* Sh[6]; double* SSb[6]; double* SSh[6];
        for (int m = 0; m < 6; ++m) {
            L[m] = buf + (6 + m) * np; Q[m] = buf + (12 + m) * np;
            Sb[m] = buf + (18 + m) * np; Sh[m] = buf + (24 + m) * np;
            SSb[m] = buf + (30 + m) * np; SSh[m] = buf + (36 + m) * np;
        }
        double* Sm = buf + 42 * np;
        const double* Sij[6] = {LEV(c, F(c, LES_F_S11), jz), LEV(c, F(c, LES_F_S12), jz), LEV(c, F(c, LES_F_S13), jz),
                                LEV(c, F(c, LES_F_S22), jz), LEV(c, F(c, LES_F_S23), jz), LEV(c, F(c, LES_F_S33), jz)};
        for (size_t q = 0; q < np; ++q) {                       /* :72-86 */
            if (jz == 1) {
                ub[q] = LEV(c, u, 1)[q]; vb[q] = LEV(c, v, 1)[q]; wb[q] = 0.25 * LEV(c, w, 2)[q];
            } else if (jz == nz) {
                ub[q] = LEV(c, u, nz - 1)[q]; vb[q] = LEV(c, v, nz - 1)[q]; wb[q] = 0.25 * LEV(c, w, nz - 1)[q];
            } else {
                ub[q] = 0.5 * (LEV(c, u, jz)[q] + LEV(c, u, jz - 1)[q]);
                vb[q] = 0.5 * (LEV(c, v, jz)[q] + LEV(c, v, jz - 1)[q]);
                wb[q] = LEV(c, w, jz)[q];
            }
            uh[q] = ub[q]; vh[q] = vb[q]; wh[q] = wb[q];
            double p11 = ub[q] * ub[q], p12 = ub[q] * vb[q], p13 = ub[q] * wb[q];
            double p22 = vb[q] * vb[q], p23 = vb[q] * wb[q], p33 = wb[q] * wb[q];
            L[0][q] = Q[0][q] = p11; L[1][q] = Q[1][q] = p12; L[2][q] = Q[2][q] = p13;
            L[3][q] = Q[3][q] = p22; L[4][q] = Q[4][q] = p23; L[5][q] = Q[5][q] = p33;
            double s = smag(Sij[0][q], Sij[1][q], Sij[2][q], Sij[3][q], Sij[4][q], Sij[5][q]);   /* :138-139 */
            Sm[q] = s;
            for (int m = 0; m < 6; ++m) {
                Sb[m][q] = Sh[m][q] = Sij[m][q];
                SSb[m][q] = SSh[m][q] = s * Sij[m][q];
            }
        }
        /* 2 delta filters (:103-118, :156-161, :190-195) */
        orc_test_filter_plane(c, ub, c->Gtest); orc_test_filter_plane(c, vb, c->Gtest); orc_test_filter_plane(c, wb, c->Gtest);
        for (int m = 0; m < 6; ++m) {
            orc_test_filter_plane(c, L[m], c->Gtest);
            orc_test_filter_plane(c, Sb[m], c->Gtest);
            orc_test_filter_plane(c, SSb[m], c->Gtest);
        }
        /* 4 delta filters (:120-135, :163-168, :197-202) */
        orc_test_filter_plane(c, uh, c->Gtest2); orc_test_filter_plane(c, vh, c->Gtest2); orc_test_filter_plane(c, wh, c->Gtest2);
        for (int m = 0; m < 6; ++m) {
            orc_test_filter_plane(c, Q[m], c->Gtest2);
            orc_test_filter_plane(c, Sh[m], c->Gtest2);
            orc_test_filter_plane(c, SSh[m], c->Gtest2);
        }
        double* FLM = LEV(c, F(c, LES_F_F_LM), jz); double* FMM = LEV(c, F(c, LES_F_F_MM), jz);
        double* FQN = LEV(c, F(c, LES_F_F_QN), jz); double* FNN = LEV(c, F(c, LES_F_F_NN), jz);
        double* Cs = LEV(c, F(c, LES_F_CS_OPT2), jz);
        for (size_t q = 0; q < np; ++q) {
            double l11 = L[0][q] - ub[q] * ub[q], l12 = L[1][q] - ub[q] * vb[q], l13 = L[2][q] - ub[q] * wb[q];
            double l22 = L[3][q] - vb[q] * vb[q], l23 = L[4][q] - vb[q] * wb[q], l33 = L[5][q] - wb[q] * wb[q];
            double q11 = Q[0][q] - uh[q] * uh[q], q12 = Q[1][q] - uh[q] * vh[q], q13 = Q[2][q] - uh[q] * wh[q];
            double q22 = Q[3][q] - vh[q] * vh[q], q23 = Q[4][q] - vh[q] * wh[q], q33 = Q[5][q] - wh[q] * wh[q];
            double S_bar = smag(Sb[0][q], Sb[1][q], Sb[2][q], Sb[3][q], Sb[4][q], Sb[5][q]);   /* :170-174 */
            double S_hat = smag(Sh[0][q], Sh[1][q], Sh[2][q], Sh[3][q], Sh[4][q], Sh[5][q]);
            double M[6], N[6];
            for (int m = 0; m < 6; ++m) {                        /* :204-216 */
                M[m] = cst * (SSb[m][q] - tf1_2 * S_bar * Sb[m][q]);
                N[m] = cst * (SSh[m][q] - tf2_2 * S_hat * Sh[m][q]);
            }
            double LM = l11 * M[0] + l22 * M[3] + l33 * M[5] + 2.0 * (l12 * M[1] + l13 * M[2] + l23 * M[4]);
            double MM = M[0] * M[0] + M[3] * M[3] + M[5] * M[5] + 2.0 * (M[1] * M[1] + M[2] * M[2] + M[4] * M[4]);
            double QN = q11 * N[0] + q22 * N[3] + q33 * N[5] + 2.0 * (q12 * N[1] + q13 * N[2] + q23 * N[4]);
            double NN = N[0] * N[0] + N[3] * N[3] + N[5] * N[5] + 2.0 * (N[1] * N[1] + N[2] * N[2] + N[4] * N[4]);
            if (init_lm) { FMM[q] = MM; FLM[q] = 0.03 * MM; }    /* :223-233 */
            double Tn = fmax(FLM[q] * FMM[q], CLIP24);           /* :261-267 */
            Tn = opftdelta * pow(Tn, powcoeff);
            Tn = fmax(CLIP24, Tn);
            double dumfac = lagran_dt / Tn;
            double epsi = dumfac / (1.0 + dumfac);
            FLM[q] = epsi * LM + (1.0 - epsi) * FLM[q];
            FMM[q] = epsi * MM + (1.0 - epsi) * FMM[q];
            FLM[q] = fmax(CLIP24, FLM[q]);
            double cs2d = fmax(CLIP24, FLM[q] / FMM[q]);         /* :271-276 */
            if (init_qn) { FNN[q] = NN; FQN[q] = 0.03 * NN; }    /* :278-288 */
            Tn = fmax(FQN[q] * FNN[q], CLIP24);                  /* :292-300 */
            Tn = opftdelta * pow(Tn, powcoeff);
            Tn = fmax(CLIP24, Tn);
            dumfac = lagran_dt / Tn;
            epsi = dumfac / (1.0 + dumfac);
            FQN[q] = epsi * QN + (1.0 - epsi) * FQN[q];
            FNN[q] = epsi * NN + (1.0 - epsi) * FNN[q];
            FQN[q] = fmax(CLIP24, FQN[q]);
            double cs4d = fmax(CLIP24, FQN[q] / FNN[q]);         /* :302-306 */
            double beta = pow(cs4d / cs2d, bexp);                /* :308-309 */
            if (jz == nz) beta = 1.0;                            /* :323-325 */
            double betaclip = fmax(beta, 1.0 / (tf1 * tf2));     /* :327 */
            Cs[q] = fmax(CLIP24, cs2d / betaclip);               /* :328-331 */
        }
        free(buf);
    }
    if (init_lm) c->f_lm_mm_init = 1;
    if (init_qn) c->f_qn_nn_init = 1;
}

/* --------------------------------------------------------------------------
 * sgs_stag, 106_sgs_stag.f90:5-318 (model 5, sgs = .true.)
 * ------------------------------------------------------------------------ */
int orc_sgs_stag(les_ctx* c, int jt) {
    const les_params* P = &c->prm;
    const int nz = c->nz;
    const size_t np = c->plane, nl = c->nlev;
    const double delta = 1.0 * pow(c->dx * c->dy * c->dz, 1.0 / 3.0);   /* :33 */
    const double nu = P->molec ? (P->nu_molec / (P->u_scale * P->z_i)) : 0.0;
    double *dudx = F(c, LES_F_DUDX), *dudy = F(c, LES_F_DUDY), *dudz = F(c, LES_F_DUDZ);
    double *dvdx = F(c, LES_F_DVDX), *dvdy = F(c, LES_F_DVDY), *dvdz = F(c, LES_F_DVDZ);
    double *dwdx = F(c, LES_F_DWDX), *dwdy = F(c, LES_F_DWDY), *dwdz = F(c, LES_F_DWDZ);
    double *S11 = F(c, LES_F_S11), *S12 = F(c, LES_F_S12), *S13 = F(c, LES_F_S13);
    double *S22 = F(c, LES_F_S22), *S23 = F(c, LES_F_S23), *S33 = F(c, LES_F_S33);
    /* wall level on uvp nodes (:47-68) */
    for (size_t q = 0; q < np; ++q) {
        size_t o1 = np + q, o2 = 2 * np + q;
        double ux = dudx[o1], uy = dudy[o1], uz = dudz[o1];
        double vx = dvdx[o1], vy = dvdy[o1], vz = dvdz[o1];
        double wx = 0.5 * (dwdx[o1] + dwdx[o2]), wy = 0.5 * (dwdy[o1] + dwdy[o2]), wz = dwdz[o1];
        S11[o1] = ux; S12[o1] = 0.5 * (uy + vx); S13[o1] = 0.5 * (uz + wx);
        S22[o1] = vy; S23[o1] = 0.5 * (vz + wy); S33[o1] = wz;
    }
    /* top level on uvp nodes (:83-104) */
    for (size_t q = 0; q < np; ++q) {
        size_t ot = (size_t)nz * np + q, om = (size_t)(nz - 1) * np + q;
        double ux = dudx[om], uy = dudy[om], uz = dudz[ot];
        double vx = dvdx[om], vy = dvdy[om], vz = dvdz[ot];
        double wx = 0.5 * (dwdx[om] + dwdx[ot]), wy = 0.5 * (dwdy[om] + dwdy[ot]), wz = dwdz[om];
        S11[ot] = ux; S12[ot] = 0.5 * (uy + vx); S13[ot] = 0.5 * (uz + wx);
        S22[ot] = vy; S23[ot] = 0.5 * (vz + wy); S33[ot] = wz;
    }
    /* w-nodes 2..nzt-1 (:126-146) */
    for (int k = 2; k <= nz - 1; ++k)
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q, om = o - np;
            double ux = 0.5 * (dudx[o] + dudx[om]), uy = 0.5 * (dudy[o] + dudy[om]), uz = dudz[o];
            double vx = 0.5 * (dvdx[o] + dvdx[om]), vy = 0.5 * (dvdy[o] + dvdy[om]), vz = dvdz[o];
            double wx = dwdx[o], wy = dwdy[o], wz = 0.5 * (dwdz[o] + dwdz[om]);
            S11[o] = ux; S12[o] = 0.5 * (uy + vx); S13[o] = 0.5 * (uz + wx);
            S22[o] = vy; S23[o] = 0.5 * (vz + wy); S33[o] = wz;
        }
    /* u_lag += u (:151-155); skipped with the SGS frozen (no reference
     * counterpart: holds Cs_opt2 and F_* for derivative checks) */
    if (!c->sgs_frozen) {
        double *ul = F(c, LES_F_U_LAG), *vl = F(c, LES_F_V_LAG), *wl = F(c, LES_F_W_LAG);
        const double *u = F(c, LES_F_U), *v = F(c, LES_F_V), *w = F(c, LES_F_W);
        for (size_t q = 0; q < np * nl; ++q) { ul[q] = ul[q] + u[q]; vl[q] = vl[q] + v[q]; wl[q] = wl[q] + w[q]; }
    }
    /* dynamic coefficient (:180-199) */
    double* Cs = F(c, LES_F_CS_OPT2);
    if (jt == 1 && P->inilag) {
        for (size_t q = np; q < np * nl; ++q) Cs[q] = 0.03;
    } else if (!c->sgs_frozen && (jt >= P->dyn_init || P->restart_vel) && (jt % P->cs_count) == 0) {
        lagrange_sdep(c, jt);
    }
    /* |S| and Nu_t (:206-216), l = delta */
    double* Nu = F(c, LES_F_NU_T);
    for (int k = 1; k <= nz; ++k)
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q;
            double s = smag(S11[o], S12[o], S13[o], S22[o], S23[o], S33[o]);
            Nu[o] = s * Cs[o] * (delta * delta);
        }
    double *txx = F(c, LES_F_TXX), *txy = F(c, LES_F_TXY), *txz = F(c, LES_F_TXZ);
    double *tyy = F(c, LES_F_TYY), *tyz = F(c, LES_F_TYZ), *tzz = F(c, LES_F_TZZ);
    for (size_t q = 0; q < np; ++q) {                            /* :219-233 */
        size_t o = np + q;
        double k2 = Nu[o];
        txx[o] = -2.0 * (k2 + nu) * S11[o];
        txy[o] = -2.0 * (k2 + nu) * S12[o];
        tyy[o] = -2.0 * (k2 + nu) * S22[o];
        tzz[o] = -2.0 * (k2 + nu) * S33[o];
    }
    for (size_t q = 0; q < np; ++q) {                            /* :240-254 */
        size_t o = (size_t)(nz - 1) * np + q, ot = (size_t)nz * np + q;
        double k2 = Nu[ot];
        txx[o] = -2.0 * (k2 + nu) * S11[ot];
        txy[o] = -2.0 * (k2 + nu) * S12[ot];
        tyy[o] = -2.0 * (k2 + nu) * S22[ot];
        tzz[o] = -2.0 * (k2 + nu) * S33[ot];
    }
    for (int k = 2; k <= nz - 2; ++k)                            /* :263-280 */
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q, op = o + np;
            double k2 = 0.5 * (Nu[o] + Nu[op]);
            txx[o] = -2.0 * (k2 + nu) * 0.5 * (S11[o] + S11[op]);
            txy[o] = -2.0 * (k2 + nu) * 0.5 * (S12[o] + S12[op]);
            tyy[o] = -2.0 * (k2 + nu) * 0.5 * (S22[o] + S22[op]);
            tzz[o] = -2.0 * (k2 + nu) * 0.5 * (S33[o] + S33[op]);
        }
    for (int k = 2; k <= nz - 1; ++k)                            /* :283-292 */
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q;
            double k2 = Nu[o];
            txz[o] = -2.0 * (k2 + nu) * S13[o];
            tyz[o] = -2.0 * (k2 + nu) * S23[o];
        }
    for (size_t q = 0; q < np; ++q) {                            /* :305-315 */
        txx[q] = txy[q] = txz[q] = tyy[q] = tyz[q] = tzz[q] = ORC_BOGUS;
        size_t ot = (size_t)nz * np + q;
        txx[ot] = txy[ot] = tyy[ot] = tzz[ot] = ORC_BOGUS;
    }
    return LES_OK;
}

/* --------------------------------------------------------------------------
 * divstress_uv / divstress_w, 107_divstress.f90:4-104
 * ------------------------------------------------------------------------ */
int orc_divstress_uv(les_ctx* c, int divt, int tx, int ty, int tz) {
    const size_t np = c->plane, nl = c->nlev;
    double* a = (double*)malloc(sizeof(double) * np * nl);
    double* b = (double*)malloc(sizeof(double) * np * nl);
    double* d = (double*)malloc(sizeof(double) * np * nl);
    ddx_arr(c, a, F(c, tx));
    ddy_arr(c, b, F(c, ty));
    ddz_w_arr(c, d, F(c, tz));
    double* o = F(c, divt);
    for (int k = 1; k <= c->nz - 1; ++k)
        for (size_t q = 0; q < np; ++q) {
            size_t r = (size_t)k * np + q;
            o[r] = a[r] + b[r] + d[r];
        }
    for (size_t q = 0; q < np; ++q) { o[q] = ORC_BOGUS; LEV(c, o, c->nz)[q] = ORC_BOGUS; }
    free(a); free(b); free(d);
    return LES_OK;
}

int orc_divstress_w(les_ctx* c, int divt, int tx, int ty, int tz) {
    const size_t np = c->plane, nl = c->nlev;
    const int nz = c->nz;
    double* a = (double*)malloc(sizeof(double) * np * nl);
    double* b = (double*)malloc(sizeof(double) * np * nl);
    double* d = (double*)calloc(np * nl, sizeof(double));
    ddx_arr(c, a, F(c, tx));
    ddy_arr(c, b, F(c, ty));
    ddz_uv_arr(c, d, F(c, tz));
    double* o = F(c, divt);
    for (size_t q = 0; q < np; ++q) {
        o[q] = ORC_BOGUS;
        o[np + q] = a[np + q] + b[np + q];                       /* :85 */
        size_t t = (size_t)nz * np + q;
        o[t] = a[t] + b[t];                                      /* :98 */
    }
    for (int k = 2; k <= nz - 1; ++k)
        for (size_t q = 0; q < np; ++q) {
            size_t r = (size_t)k * np + q;
            o[r] = a[r] + b[r] + d[r];
        }
    free(a); free(b); free(d);
    return LES_OK;
}

/* --------------------------------------------------------------------------
 * convec + cross_product, 108_convec.f90:11-202
 * ------------------------------------------------------------------------ */
int orc_convec(les_ctx* c) {
    const int nz = c->nz;
    const size_t np = c->plane, np2 = c->plane2, nl = c->nlev;
    const double *u = F(c, LES_F_U), *v = F(c, LES_F_V), *w = F(c, LES_F_W);
    const double *dudy = F(c, LES_F_DUDY), *dudz = F(c, LES_F_DUDZ);
    const double *dvdx = F(c, LES_F_DVDX), *dvdz = F(c, LES_F_DVDZ);
    const double *dwdx = F(c, LES_F_DWDX), *dwdy = F(c, LES_F_DWDY);
    double* vort = (double*)calloc(3 * np * nl, sizeof(double));
    double* vb = (double*)calloc(3 * np2 * nl, sizeof(double));
    double* cb = (double*)calloc(3 * np2 * nl, sizeof(double));
    double *v1 = vort, *v2 = vort + np * nl, *v3 = vort + 2 * np * nl;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k <= nz; ++k) {                              /* :38-42 */
        orc_pad_plane(c, LEV(c, u, k), c->u_big + (size_t)k * np2);
        orc_pad_plane(c, LEV(c, v, k), c->v_big + (size_t)k * np2);
        orc_pad_plane(c, LEV(c, w, k), c->w_big + (size_t)k * np2);
    }
    for (int k = 1; k <= nz; ++k)                                /* :44-69 */
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q;
            if (k == 1) {
                v1[o] = 0.5 * (dwdy[o] + dwdy[o + np]) - dvdz[o];
                v2[o] = dudz[o] - 0.5 * (dwdx[o] + dwdx[o + np]);
            } else if (k == nz) {
                v1[o] = 0.5 * (dwdy[o] + dwdy[o - np]) - dvdz[o];
                v2[o] = dudz[o] - 0.5 * (dwdx[o] + dwdx[o - np]);
            } else {
                v1[o] = dwdy[o] - dvdz[o];
                v2[o] = dudz[o] - dwdx[o];
            }
            v3[o] = dvdx[o] - dudy[o];
        }
    double *b1 = vb, *b2 = vb + np2 * nl, *b3 = vb + 2 * np2 * nl;
    #pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k) {                              /* :71-75 */
        orc_pad_plane(c, LEV(c, v1, k), b1 + (size_t)k * np2);
        orc_pad_plane(c, LEV(c, v2, k), b2 + (size_t)k * np2);
        orc_pad_plane(c, LEV(c, v3, k), b3 + (size_t)k * np2);
    }
    const double *ub = c->u_big, *vbg = c->v_big, *wb = c->w_big;
    double *cxb = cb, *cyb = cb + np2 * nl, *czb = cb + 2 * np2 * nl;
    const double *ust = c->ust, *vst = c->vst;
    for (size_t q = 0; q < np2; ++q) {                           /* :116-126, :148-159, :181-190 */
        size_t o1 = np2 + q, o2 = 2 * np2 + q;
        cxb[o1] = (vbg[o1] + vst[1]) * (-b3[o1]) + 0.5 * wb[o2] * b2[o2];
        cyb[o1] = (ub[o1] + ust[1]) * (b3[o1]) + 0.5 * wb[o2] * (-b1[o2]);
        czb[o1] = 0.0;
        czb[(size_t)nz * np2 + q] = 0.0;
    }
    for (int k = 2; k <= nz - 1; ++k)                            /* :142-146, :175-179, :196-200 */
        for (size_t q = 0; q < np2; ++q) {
            size_t o = (size_t)k * np2 + q, op = o + np2, om = o - np2;
            cxb[o] = (vbg[o] + vst[k]) * (-b3[o]) + 0.5 * (wb[op] * (b2[op]) + wb[o] * (b2[o]));
            cyb[o] = (ub[o] + ust[k]) * (b3[o]) + 0.5 * (wb[op] * (-b1[op]) + wb[o] * (-b1[o]));
            czb[o] = 0.5 * ((ub[o] + ub[om] + ust[k] + ust[k - 1]) * (-b2[o]) +
                            (vbg[o] + vbg[om] + vst[k] + vst[k - 1]) * (b1[o]));
        }
    double *cx = F(c, LES_F_CX), *cy = F(c, LES_F_CY), *cz = F(c, LES_F_CZ);
    #pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz - 1; ++k) {                          /* :84-88 */
        orc_unpad_plane(c, cxb + (size_t)k * np2, LEV(c, cx, k));
        orc_unpad_plane(c, cyb + (size_t)k * np2, LEV(c, cy, k));
        orc_unpad_plane(c, czb + (size_t)k * np2, LEV(c, cz, k));
    }
    for (size_t q = 0; q < np; ++q) {                            /* :90-96 */
        cx[q] = cy[q] = cz[q] = ORC_BOGUS;
        LEV(c, cx, nz)[q] = ORC_BOGUS;
        LEV(c, cy, nz)[q] = ORC_BOGUS;
    }
    free(vort); free(vb); free(cb);
    return LES_OK;
}

/* --------------------------------------------------------------------------
 * theta_all_in_one, 010_scalars.f90:111-200, with scalar_RHS_calc (:265-620,
 * lbc = 1, patch_flag = 1), calc_beta (:204-261), step_scalar (:624-671)
 * ------------------------------------------------------------------------ */
int orc_theta_all_in_one(les_ctx* c, int jt) {
    const les_params* P = &c->prm;
    const int nz = c->nz;
    const size_t np = c->plane, np2 = c->plane2, nl = c->nlev;
    double* beta = F(c, LES_F_BETA_SCAL);
    if (!(P->theta_flag && jt >= P->theta_init_time)) {          /* 103_flosol.f90:96-100 */
        memset(beta, 0, sizeof(double) * np * nl);
        return LES_OK;
    }
    double* th = F(c, LES_F_THETA);
    double *dTdx = F(c, LES_F_DTDX), *dTdy = F(c, LES_F_DTDY), *dTdz = F(c, LES_F_DTDZ);
    double *RHS = F(c, LES_F_RHST), *RHSf = F(c, LES_F_RHST_F);
    double* sgs_t3 = F(c, LES_F_SGS_T3);
    const double* Nu = F(c, LES_F_NU_T);
    filter_arr(c, th);                                           /* :135 */
    ddx_arr(c, dTdx, th);
    ddy_arr(c, dTdy, th);
    ddz_uv_arr(c, dTdz, th);
    for (size_t q = 0; q < np; ++q) LEV(c, dTdz, nz)[q] = P->dTdz_top / P->T_scale * P->z_i;   /* :140-142 */
    memcpy(RHSf, RHS, sizeof(double) * np * nl);                 /* :163 */
    /* ---- scalar_RHS_calc ---- */
    /* surface flux (:359-394): lbc = 1 the prescribed wt_s; lbc = 0 the bulk
     * Monin-Obukhov flux from the surface temperature T_s, the scalar
     * roughness z_os = zo/10 (008_bottombc.f90:126) and the filtered theta
     * of level 1, with ustar_avg as wallstress_bottom (or obukhov) left it */
    double* surf_flux = (double*)malloc(sizeof(double) * np);
    {
        const double* th1 = LEV(c, th, 1);
        const double z_os = (1.0 / 10.0) * P->zo;
        for (size_t q = 0; q < np; ++q) {
            if (P->lbc == 1) surf_flux[q] = P->wt_s / P->T_scale / P->u_scale;
            else surf_flux[q] = (P->T_s - th1[q]) * ORC_VONK * c->ustar_avg[q] / (log(c->dz / (2.0 * z_os)) - c->psi_h[q]);
        }
    }
    double* dsx = (double*)calloc(np2 * nl, sizeof(double));
    double* dsy = (double*)calloc(np2 * nl, sizeof(double));
    double* dsz = (double*)calloc(np2 * nl, sizeof(double));
    double* rm = (double*)calloc(np2 * nl, sizeof(double));
    #pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; ++k) {                              /* :428-432 (levels read below) */
        orc_pad_plane(c, LEV(c, dTdx, k), dsx + (size_t)k * np2);
        orc_pad_plane(c, LEV(c, dTdy, k), dsy + (size_t)k * np2);
        if (k >= 2) orc_pad_plane(c, LEV(c, dTdz, k), dsz + (size_t)k * np2);
    }
    const double *ub = c->u_big, *vb = c->v_big, *wb = c->w_big;
    for (int k = 2; k <= nz - 1; ++k)                            /* :466-474 */
        for (size_t q = 0; q < np2; ++q) {
            size_t o = (size_t)k * np2 + q, op = o + np2;
            rm[o] = (ub[o] + c->ust[k]) * dsx[o] + (vb[o] + c->vst[k]) * dsy[o] +
                    (wb[o] * dsz[o] + wb[op] * dsz[op]) / 2.0;
        }
    for (size_t q = 0; q < np2; ++q) {                           /* :476-484 */
        size_t o = np2 + q, o2 = 2 * np2 + q;
        rm[o] = (ub[o] + c->ust[1]) * dsx[o] + (vb[o] + c->vst[1]) * dsy[o] + (0.5 * wb[o2]) * dsz[o2];
    }
    #pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz - 1; ++k) orc_unpad_plane(c, rm + (size_t)k * np2, LEV(c, RHS, k));   /* :497-499 */
    free(dsx); free(dsy); free(dsz); free(rm);
    double* temp = (double*)calloc(np * nl, sizeof(double));
    double* dtemp = (double*)calloc(np * nl, sizeof(double));
    const double iPr = 1.0 / P->Pr;
    /* x flux (:513-541) */
    for (size_t q = 0; q < np; ++q) {
        temp[np + q] = iPr * Nu[np + q] * dTdx[np + q];
        size_t o = (size_t)(nz - 1) * np + q;
        temp[o] = iPr * Nu[(size_t)nz * np + q] * dTdx[o];
    }
    for (int k = 2; k <= nz - 2; ++k)
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q;
            temp[o] = iPr * 0.5 * (Nu[o] + Nu[o + np]) * dTdx[o];
        }
    ddx_arr(c, dtemp, temp);
    for (int k = 1; k <= nz - 1; ++k)                            /* :545-569 */
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q;
            RHS[o] = (-1.0 * RHS[o] + dtemp[o]);
        }
    /* y flux */
    memset(temp, 0, sizeof(double) * np * nl);
    for (size_t q = 0; q < np; ++q) {
        temp[np + q] = iPr * Nu[np + q] * dTdy[np + q];
        size_t o = (size_t)(nz - 1) * np + q;
        temp[o] = iPr * Nu[(size_t)nz * np + q] * dTdy[o];
    }
    for (int k = 2; k <= nz - 2; ++k)
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q;
            temp[o] = iPr * 0.5 * (Nu[o] + Nu[o + np]) * dTdy[o];
        }
    ddy_arr(c, dtemp, temp);
    for (int k = 1; k <= nz - 1; ++k)                            /* :576-603 */
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q;
            RHS[o] = RHS[o] + dtemp[o];
        }
    /* z flux (:576-618) */
    memset(temp, 0, sizeof(double) * np * nl);
    for (size_t q = 0; q < np; ++q) {
        temp[np + q] = -1.0 * surf_flux[q];
        sgs_t3[np + q] = surf_flux[q];
    }
    free(surf_flux);
    for (int k = 2; k <= nz; ++k)
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q;
            temp[o] = iPr * Nu[o] * dTdz[o];
            sgs_t3[o] = -1.0 * temp[o];
        }
    ddz_w_arr(c, dtemp, temp);
    for (int k = 1; k <= nz - 1; ++k)
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q;
            RHS[o] = RHS[o] + dtemp[o];
        }
    free(temp); free(dtemp);
    /* ---- calc_beta (:204-261) ---- */
    const double g_hat = ORC_G * (P->z_i / (P->u_scale * P->u_scale));
    double* bar = (double*)calloc(nl, sizeof(double));
    for (int k = 0; k <= nz - 1; ++k) bar[k] = orc_hor_avg(c, LEV(c, th, k));
    memset(beta, 0, sizeof(double) * np * nl);
    for (int k = 2; k <= nz - 1; ++k)
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q;
            double above, below;
            if (P->ocean_flag) {                                 /* :239-248 */
                above = P->alpha_w * (th[o] - bar[k]);
                below = P->alpha_w * (th[o - np] - bar[k - 1]);
            } else {
                above = (th[o] - bar[k]) / bar[k];
                below = (th[o - np] - bar[k - 1]) / bar[k - 1];
            }
            beta[o] = g_hat * (above + below) / 2.0;
        }
    free(bar);
    if (jt == P->theta_init_time && !P->restart_theta) memcpy(RHSf, RHS, sizeof(double) * np * nl);   /* :183-185 */
    /* ---- step_scalar (:624-671) ---- */
    for (int k = 1; k <= nz - 1; ++k)
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q;
            th[o] = th[o] + P->dt * (1.5 * RHS[o] - 0.5 * RHSf[o]);
        }
    for (size_t q = 0; q < np; ++q) {
        size_t o = (size_t)nz * np + q, om = o - np;
        if (P->lz_npz > 1.0) th[o] = th[om] + P->dTdz_top / P->T_scale * P->z_i * c->dz;
        else th[o] = th[om];
    }
    return LES_OK;
}

/* --------------------------------------------------------------------------
 * press_stag_array's right-hand side (109_press_stag_array.f90:201-225) and
 * tdmaSolver_pipelined (:244-319): half spectra Hx, Hy, Hz (levels 0..nz)
 * -> p_hat (levels 0..nz-1)
 * ------------------------------------------------------------------------ */
static void tdma_solve(les_ctx* c, const cplx* Hx, const cplx* Hy, const cplx* Hz, cplx* ph) {
    const int nz = c->nz;
    const size_t ns = (size_t)c->lhx * c->ny;
    const double dz = c->dz;
    cplx* rhs = (cplx*)calloc(ns * nz, sizeof(cplx));           /* RHS_col(0:nz-1) */
    for (size_t q = 0; q < ns; ++q) {
        const double k2 = c->k2[q], kx = c->kx2d[q], ky = c->ky2d[q];
        if (k2 > 0)                                              /* :203-209 */
            rhs[ns + q] = I * (kx * Hx[ns + q] + ky * Hy[ns + q]) + (Hz[2 * ns + q] - Hz[ns + q] + Hz[q]) / dz;
        rhs[q] = Hz[q];
        for (int k = 2; k <= nz - 1; ++k) {                      /* :213-223 */
            size_t o = (size_t)k * ns + q;
            if (k2 > 0) rhs[o] = I * (kx * Hx[o] + ky * Hy[o]) + (Hz[o + ns] - Hz[o]) / dz;
            if (k2 == 0 && c->notnyq[q]) rhs[o] = Hz[o];
        }
    }
    for (size_t q = 0; q < ns; ++q) {
        if (!c->notnyq[q]) continue;                             /* Nyquist stays 0 */
        for (int k = 1; k <= nz - 1; ++k) {                      /* forward :256-266 */
            size_t o = (size_t)k * ns + q;
            rhs[o] = (rhs[o] - c->coef_a[o] * rhs[o - ns]) / c->coef_b[o];
        }
        ph[(size_t)(nz - 1) * ns + q] = rhs[(size_t)(nz - 1) * ns + q];   /* c'(nz-1) = 0 */
        for (int k = nz - 2; k >= 1; --k) {                      /* backward :290-297 */
            size_t o = (size_t)k * ns + q;
            ph[o] = rhs[o] - c->coef_c[o] * ph[o + ns];
        }
        ph[q] = ph[ns + q] - dz * rhs[q];                        /* :313-315 */
    }
    free(rhs);
}

int orc_tdma_solve(les_ctx* c, const double* hx, const double* hy, const double* hz, double* ph) {
    tdma_solve(c, (const cplx*)hx, (const cplx*)hy, (const cplx*)hz, (cplx*)ph);
    return LES_OK;
}

/* --------------------------------------------------------------------------
 * press_stag_array + tdmaSolver_pipelined, 109_press_stag_array.f90:107-319,
 * then dpdz (103_flosol.f90:174-175)
 * ------------------------------------------------------------------------ */
int orc_press_stag_array(les_ctx* c, int jt) {
    const les_params* P = &c->prm;
    const int nz = c->nz, lhx = c->lhx, ny = c->ny;
    const size_t np = c->plane, ns = (size_t)lhx * ny;
    const double cst = 1.0 / (double)(c->nx * c->ny);
    const double tadv1 = 1.5, dz = c->dz;
    cplx* Hx = (cplx*)calloc(ns * (nz + 1), sizeof(cplx));
    cplx* Hy = (cplx*)calloc(ns * (nz + 1), sizeof(cplx));
    cplx* Hz = (cplx*)calloc(ns * (nz + 1), sizeof(cplx));
    const double *u = F(c, LES_F_U), *v = F(c, LES_F_V), *w = F(c, LES_F_W);
    #pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz - 1; ++k) {                          /* :158-166 */
        double* r = (double*)malloc(sizeof(double) * np);
        for (size_t q = 0; q < np; ++q) r[q] = cst / tadv1 * (LEV(c, u, k)[q] / P->dt);
        orc_fftn_plane(c, r, Hx + (size_t)k * ns);
        for (size_t q = 0; q < np; ++q) r[q] = cst / tadv1 * (LEV(c, v, k)[q] / P->dt);
        orc_fftn_plane(c, r, Hy + (size_t)k * ns);
        for (size_t q = 0; q < np; ++q) r[q] = cst / tadv1 * (LEV(c, w, k)[q] / P->dt);
        orc_fftn_plane(c, r, Hz + (size_t)k * ns);
        free(r);
    }
    {
        double* r = (double*)malloc(sizeof(double) * np);
        const double* divtz = F(c, LES_F_DIVTZ);
        /* surface-temperature term of lbc = 0 (:186-191): g z_i/u_scale^2
         * (T_s - <T_s>) of the uniform T_s plane, <.> summed in order */
        double modi = 0.0;
        if (P->theta_flag && jt > P->theta_init_time && P->lbc == 0) {
            double st = 0.0;
            for (size_t q = 0; q < np; ++q) st += P->T_s;
            modi = ORC_G * (P->z_i / (P->u_scale * P->u_scale)) * (P->T_s - st / (double)(c->nx * c->ny));
        }
        for (size_t q = 0; q < np; ++q) r[q] = cst * (LEV(c, divtz, 1)[q] - modi);   /* :192-196 */
        orc_fftn_plane(c, r, Hz);
        for (size_t q = 0; q < np; ++q) r[q] = cst * LEV(c, divtz, nz)[q];   /* :198-199 */
        orc_fftn_plane(c, r, Hz + (size_t)nz * ns);
        free(r);
    }
    cplx* ph = (cplx*)calloc(ns * nz, sizeof(cplx));            /* p_hat(0:nz-1) */
    tdma_solve(c, Hx, Hy, Hz, ph);
    double* p = F(c, LES_F_P);
    double *dpdx = F(c, LES_F_DPDX), *dpdy = F(c, LES_F_DPDY), *dpdz = F(c, LES_F_DPDZ);
    orc_ifftn_plane(c, ph, p);                                    /* :231 */
    #pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz - 1; ++k) {                          /* :232-238 */
        cplx* t = (cplx*)malloc(sizeof(cplx) * ns);
        const cplx* pk = ph + (size_t)k * ns;
        for (size_t q = 0; q < ns; ++q) t[q] = (I * c->kx2d[q]) * pk[q];
        orc_ifftn_plane(c, t, LEV(c, dpdx, k));
        for (size_t q = 0; q < ns; ++q) t[q] = (I * c->ky2d[q]) * pk[q];
        orc_ifftn_plane(c, t, LEV(c, dpdy, k));
        orc_ifftn_plane(c, pk, LEV(c, p, k));
        free(t);
    }
    for (int k = 1; k <= nz - 1; ++k)                            /* 103_flosol.f90:174-175 */
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q;
            dpdz[o] = (p[o] - p[o - np]) / dz;
        }
    free(Hx); free(Hy); free(Hz); free(ph);
    return LES_OK;
}

/* --------------------------------------------------------------------------
 * project, 110_project.f90:4-62, then data_exchange's boundary-ghost rules
 * (202_parallel.f90:722-733)
 * ------------------------------------------------------------------------ */
int orc_project(les_ctx* c) {
    const les_params* P = &c->prm;
    const int nz = c->nz;
    const size_t np = c->plane;
    double *u = F(c, LES_F_U), *v = F(c, LES_F_V), *w = F(c, LES_F_W), *th = F(c, LES_F_THETA);
    const double *dpdx = F(c, LES_F_DPDX), *dpdy = F(c, LES_F_DPDY), *dpdz = F(c, LES_F_DPDZ);
    const double tadv1 = 1.5;
    for (int k = 1; k <= nz - 1; ++k)
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q;
            double r = -tadv1 * dpdx[o];
            u[o] = u[o] + P->dt * r;
            r = -tadv1 * dpdy[o];
            v[o] = v[o] + P->dt * r;
        }
    for (int k = 2; k <= nz - 1; ++k)
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q;
            double r = -tadv1 * dpdz[o];
            w[o] = w[o] + P->dt * r;
        }
    for (size_t q = 0; q < np; ++q) {
        if (P->ubc_mom == LES_BC_STRESS_FREE) {                   /* :50-54 */
            LEV(c, u, nz - 1)[q] = LEV(c, u, nz - 2)[q];
            LEV(c, v, nz - 1)[q] = LEV(c, v, nz - 2)[q];
        }
        LEV(c, w, nz)[q] = 0.0;                                   /* :55 */
        LEV(c, w, 1)[q] = 0.0;                                    /* :58-60 */
        u[q] = v[q] = w[q] = ORC_BOGUS;                           /* data_exchange :722-733 */
        LEV(c, u, nz)[q] = ORC_BOGUS;
        LEV(c, v, nz)[q] = ORC_BOGUS;
        if (P->theta_flag) { th[q] = ORC_BOGUS; LEV(c, th, nz)[q] = ORC_BOGUS; }
    }
    return LES_OK;
}

/* --------------------------------------------------------------------------
 * flosol, 103_flosol.f90:4-290
 * ------------------------------------------------------------------------ */
/* ---- time-averaged slices ------------------------------------------------ */
static size_t slice_len(const les_ctx* c) { return (size_t)c->nx * (size_t)(c->nz - 1); }

int orc_set_avgslice(les_ctx* c, int use_avgslice, int c_count, int p_count) {
    if (c_count < 1 || p_count < 1) return LES_E_BADARG;
    c->use_avgslice = use_avgslice;
    c->c_count = c_count;
    c->p_count = p_count;
    c->fr = (1.0 / (double)p_count) * (double)c_count;             /* 102_initial.f90:170 */
    const size_t F3 = c->plane * c->nlev, S = (size_t)LES_N_SLICES * slice_len(c);
    double** bufs[6] = {&c->dudt, &c->dvdt, &c->dwdt, &c->ke, &c->ke_temp, &c->dke};
    for (int b = 0; b < 6; ++b) {
        free(*bufs[b]);
        *bufs[b] = (double*)calloc(F3, sizeof(double));
    }
    free(c->sl_run); free(c->sl_win);
    c->sl_run = (double*)calloc(S, sizeof(double));
    c->sl_win = (double*)calloc(S, sizeof(double));
    return LES_OK;
}

int orc_get_avgslice(les_ctx* c, int slice, int window, double* host) {
    if (!c->sl_run || slice < 0 || slice >= LES_N_SLICES || (window != 0 && window != 1)) return LES_E_BADARG;
    const size_t L = slice_len(c);
    memcpy(host, (window ? c->sl_win : c->sl_run) + (size_t)slice * L, sizeof(double) * L);
    return LES_OK;
}

static int stats_step(const les_ctx* c, int jt) {
    return c->use_avgslice && c->sl_run && (jt % c->c_count) == 0;    /* 103_flosol.f90:231 */
}

/* avg_y3d (202_parallel.f90:966-989): out(jx, jz) = sum over y / nyt, for the
 * expression e(k, j, i) at levels 1..nz-1; accumulated into slice s with
 * weight fr. */
#define SLICE_ACC(c, s, EXPR)                                                        \
    do {                                                                             \
        double* acc_ = (c)->sl_run + (size_t)(s) * slice_len(c);                     \
        for (int k = 1; k <= (c)->nz - 1; ++k)                                       \
            for (int i = 0; i < (c)->nx; ++i) {                                      \
                double t_ = 0.0;                                                     \
                for (int j = 0; j < (c)->ny; ++j) {                                  \
                    const size_t o = IDX(c, k, j, i);                                \
                    (void)o;                                                         \
                    t_ += (EXPR);                                                    \
                }                                                                    \
                acc_[(size_t)(k - 1) * (c)->nx + i] += (c)->fr * (t_ / (c)->ny);     \
            }                                                                        \
    } while (0)

static void close_window(les_ctx* c, int s0, int s1) {
    const size_t L = slice_len(c);
    memcpy(c->sl_win + (size_t)s0 * L, c->sl_run + (size_t)s0 * L, sizeof(double) * L * (size_t)(s1 - s0));
    memset(c->sl_run + (size_t)s0 * L, 0, sizeof(double) * L * (size_t)(s1 - s0));
}

/* flow_slice, 011_scalars2.f90:28-242 (coordz == 0 branches at level 1) */
static void flow_slice(les_ctx* c, int jt) {
    const int nz = c->nz;
    const size_t np = c->plane, nl = c->nlev;
    const double *u = F(c, LES_F_U), *v = F(c, LES_F_V), *w = F(c, LES_F_W), *p = F(c, LES_F_P);
    const double *ust = c->ust, *vst = c->vst;
    double* po = (double*)malloc(sizeof(double) * np * nl);   /* p_org, :63-70 */
    for (int k = 0; k <= nz - 1; ++k)
        for (size_t q = 0; q < np; ++q) {
            const size_t o = (size_t)k * np + q;
            po[o] = p[o] - 0.5 * (u[o] * u[o] + v[o] * v[o] + w[o] * w[o]) - (u[o] * ust[k] + v[o] * vst[k]) -
                    0.5 * (ust[k] * ust[k] + vst[k] * vst[k]);
        }
    double* ub = (double*)malloc(sizeof(double) * nl);
    double* vb = (double*)malloc(sizeof(double) * nl);
    for (int k = 0; k <= nz - 1; ++k) {                           /* calc_hor_avg, :134-135 */
        ub[k] = orc_hor_avg(c, LEV(c, u, k));
        vb[k] = orc_hor_avg(c, LEV(c, v, k));
    }
    const double* dudz = F(c, LES_F_DUDZ);
    const double* dvdz = F(c, LES_F_DVDZ);
    const double *txx = F(c, LES_F_TXX), *txy = F(c, LES_F_TXY), *txz = F(c, LES_F_TXZ);
    const double *tyy = F(c, LES_F_TYY), *tyz = F(c, LES_F_TYZ), *tzz = F(c, LES_F_TZZ);
    const double* cs = F(c, LES_F_CS_OPT2);
    const double *s11 = F(c, LES_F_S11), *s12 = F(c, LES_F_S12), *s13 = F(c, LES_F_S13);
    const double *s22 = F(c, LES_F_S22), *s23 = F(c, LES_F_S23), *s33 = F(c, LES_F_S33);
    const double delta = 1.0 * pow(c->dx * c->dy * c->dz, 1.0 / 3.0);   /* l = delta, 106_sgs_stag.f90:33,181 */
    const double l2 = delta * delta;
    const double *du = c->dudt, *dv = c->dvdt, *dw = c->dwdt;
    const size_t P = np;
    /* arg1, arg2 (:77-86): staggered u, v at w-nodes, 0 at the wall level */
#define ARG1(o, k) ((k) == 1 ? 0.0 : (u[o] + u[(o) - P]) / 2.0)
#define ARG2(o, k) ((k) == 1 ? 0.0 : (v[o] + v[(o) - P]) / 2.0)
#define ARG3(o) ((po[o] + po[(o) - P]) / 2.0)
    /* arg7, arg8 (:140-157): staggered fluctuations, 0 at the wall level */
#define ARG7(o, k) ((k) == 1 ? 0.0 : (u[o] - ub[k] + u[(o) - P] - ub[(k) - 1]) / 2.0)
#define ARG8(o, k) ((k) == 1 ? 0.0 : (v[o] - vb[k] + v[(o) - P] - vb[(k) - 1]) / 2.0)
#define MAGS(o) sqrt(2.0 * (s11[o] * s11[o] + s22[o] * s22[o] + s33[o] * s33[o] + \
                            2.0 * (s12[o] * s12[o] + s13[o] * s13[o] + s23[o] * s23[o])))
    SLICE_ACC(c, LES_S_U, u[o]);
    SLICE_ACC(c, LES_S_V, v[o]);
    SLICE_ACC(c, LES_S_W, w[o]);
    SLICE_ACC(c, LES_S_P, po[o]);
    SLICE_ACC(c, LES_S_DUDT, du[o]);
    SLICE_ACC(c, LES_S_DVDT, dv[o]);
    SLICE_ACC(c, LES_S_DWDT, dw[o]);
    SLICE_ACC(c, LES_S_DUDZ, dudz[o]);
    SLICE_ACC(c, LES_S_DVDZ, dvdz[o]);
    SLICE_ACC(c, LES_S_U2, u[o] * u[o]);
    SLICE_ACC(c, LES_S_V2, v[o] * v[o]);
    SLICE_ACC(c, LES_S_W2, w[o] * w[o]);
    SLICE_ACC(c, LES_S_P2, po[o] * po[o]);
    SLICE_ACC(c, LES_S_UV, u[o] * v[o]);
    SLICE_ACC(c, LES_S_UW, ARG1(o, k) * w[o]);
    SLICE_ACC(c, LES_S_VW, ARG2(o, k) * w[o]);
    SLICE_ACC(c, LES_S_WP, ARG3(o) * w[o]);
    SLICE_ACC(c, LES_S_TXX, txx[o]);
    SLICE_ACC(c, LES_S_TXY, txy[o]);
    SLICE_ACC(c, LES_S_TXZ, txz[o]);
    SLICE_ACC(c, LES_S_TYY, tyy[o]);
    SLICE_ACC(c, LES_S_TYZ, tyz[o]);
    SLICE_ACC(c, LES_S_TZZ, tzz[o]);
    SLICE_ACC(c, LES_S_U3, u[o] * u[o] * u[o]);
    SLICE_ACC(c, LES_S_V3, v[o] * v[o] * v[o]);
    SLICE_ACC(c, LES_S_W3, w[o] * w[o] * w[o]);
    SLICE_ACC(c, LES_S_U2_FLUC, (u[o] - ub[k]) * (u[o] - ub[k]));
    SLICE_ACC(c, LES_S_V2_FLUC, (v[o] - vb[k]) * (v[o] - vb[k]));
    SLICE_ACC(c, LES_S_UV_FLUC, (u[o] - ub[k]) * (v[o] - vb[k]));
    SLICE_ACC(c, LES_S_U2W_FLUC, ARG7(o, k) * ARG7(o, k) * w[o]);
    SLICE_ACC(c, LES_S_V2W_FLUC, ARG8(o, k) * ARG8(o, k) * w[o]);
    /* dissip = -|S|^3 Cs_opt2 l^2 on w-nodes (106_sgs_stag.f90:212) */
    SLICE_ACC(c, LES_S_DISSIP, -(MAGS(o) * MAGS(o) * MAGS(o)) * cs[o] * l2);
    SLICE_ACC(c, LES_S_UTXZ_FLUC, ARG7(o, k) * txz[o]);
    SLICE_ACC(c, LES_S_VTYZ_FLUC, ARG8(o, k) * tyz[o]);
    SLICE_ACC(c, LES_S_WTZZ, w[o] * tzz[o]);
    SLICE_ACC(c, LES_S_CS, sqrt(cs[o]));
#undef ARG1
#undef ARG2
#undef ARG3
#undef ARG7
#undef ARG8
#undef MAGS
    free(po); free(ub); free(vb);
    if (jt % c->p_count == 0) close_window(c, 0, LES_N_FLOW_SLICES);   /* :186-241 */
}

/* theta_slice, 011_scalars2.f90:246-318 */
static void theta_slice(les_ctx* c, int jt) {
    const int nz = c->nz;
    const double *th = F(c, LES_F_THETA), *w = F(c, LES_F_W), *dtdz = F(c, LES_F_DTDZ);
    const double *t3 = F(c, LES_F_SGS_T3), *nut = F(c, LES_F_NU_T);
    double* tb = (double*)malloc(sizeof(double) * c->nlev);
    for (int k = 0; k <= nz - 1; ++k) tb[k] = orc_hor_avg(c, LEV(c, th, k));
    const size_t P = c->plane;
    SLICE_ACC(c, LES_S_THETA, th[o]);
    SLICE_ACC(c, LES_S_THETA2, th[o] * th[o]);
    SLICE_ACC(c, LES_S_T2_FLUC, (th[o] - tb[k]) * (th[o] - tb[k]));
    SLICE_ACC(c, LES_S_DTDZ, dtdz[o]);
    SLICE_ACC(c, LES_S_SGS_T3, t3[o]);
    SLICE_ACC(c, LES_S_WT, w[o] * ((th[o] + th[o - P]) / 2.0));
    SLICE_ACC(c, LES_S_NU_T, nut[o]);
    free(tb);
    if (jt % c->p_count == 0) close_window(c, LES_N_FLOW_SLICES, LES_N_SLICES);   /* :303-315 */
}

/* energy (103_flosol.f90:393-429) and d(ke)/dx_i added to dudt (:196-213):
 * ke = (u^2 + v^2 + w_uvp^2)/2 at 1..nz-1, 0 at the wall level; the top
 * ghost ke(nz) and dkedz(1) are not set by the reference (0 here). */
static void diag_ke(les_ctx* c) {
    const int nz = c->nz;
    const size_t np = c->plane, nl = c->nlev;
    const double *u = F(c, LES_F_U), *v = F(c, LES_F_V), *w = F(c, LES_F_W);
    memset(c->ke, 0, sizeof(double) * np * nl);
    for (int k = 1; k <= nz - 1; ++k)
        for (size_t q = 0; q < np; ++q) {
            const size_t o = (size_t)k * np + q;
            const double tw = .5 * (w[o] + w[o + np]);
            c->ke[o] = .5 * (u[o] * u[o] + v[o] * v[o] + tw * tw);
        }
    memcpy(c->ke_temp, c->ke, sizeof(double) * np * nl);
    filter_arr(c, c->ke_temp);
    ddx_arr(c, c->dke, c->ke_temp);
    for (int k = 1; k <= nz - 1; ++k)
        for (size_t q = 0; q < np; ++q) c->dudt[(size_t)k * np + q] += c->dke[(size_t)k * np + q];
    ddy_arr(c, c->dke, c->ke_temp);
    for (int k = 1; k <= nz - 1; ++k)
        for (size_t q = 0; q < np; ++q) c->dvdt[(size_t)k * np + q] += c->dke[(size_t)k * np + q];
    memset(c->dke, 0, sizeof(double) * np * nl);
    ddz_uv_arr(c, c->dke, c->ke);
    for (int k = 1; k <= nz - 1; ++k)
        for (size_t q = 0; q < np; ++q) {
            const size_t o = (size_t)k * np + q;
            c->dwdt[o] += (k == 1) ? 0.0 : c->dke[o];
        }
}

static double max_lev(les_ctx* c, const double* f) {
    double m = -INFINITY;
    for (int k = 1; k <= c->nz - 1; ++k)
        for (size_t q = 0; q < c->plane; ++q) { double x = LEV(c, f, k)[q]; if (x > m) m = x; }
    return m;
}

/* Admas_Bashforth, 103_flosol.f90:319-389: Euler at jt = 1 unless
 * restart_vel (:331-338), level 0 <- BOGUS (:385-387) */
int orc_ab2(les_ctx* c, int jt) {
    const les_params* P = &c->prm;
    const int nz = c->nz;
    const size_t np = c->plane, nl = c->nlev;
    double *RX = F(c, LES_F_RHSX), *RY = F(c, LES_F_RHSY), *RZ = F(c, LES_F_RHSZ);
    double *RXf = F(c, LES_F_RHSX_F), *RYf = F(c, LES_F_RHSY_F), *RZf = F(c, LES_F_RHSZ_F);
    double *u = F(c, LES_F_U), *v = F(c, LES_F_V), *w = F(c, LES_F_W);
    if (jt == 1 && !P->restart_vel) {
        memcpy(RXf, RX, sizeof(double) * np * nl);
        memcpy(RYf, RY, sizeof(double) * np * nl);
        memcpy(RZf, RZ, sizeof(double) * np * nl);
    }
    const double force = P->use_mean_p_force ? P->mean_p_force : 0.0;
    const double tadv1 = 1.5, tadv2 = 1.0 - tadv1;
    for (int k = 1; k <= nz - 1; ++k)                            /* :369-377 */
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q;
            u[o] = u[o] + P->dt * (tadv1 * RX[o] + tadv2 * RXf[o] + 0.0 + force);
            v[o] = v[o] + P->dt * (tadv1 * RY[o] + tadv2 * RYf[o] + 0.0);
            w[o] = w[o] + P->dt * (tadv1 * RZ[o] + tadv2 * RZf[o] + 0.0);
        }
    for (size_t q = 0; q < np; ++q) u[q] = v[q] = w[q] = ORC_BOGUS;
    return LES_OK;
}

int orc_step(les_ctx* c, int jt) {
    const les_params* P = &c->prm;
    const int nz = c->nz;
    const size_t np = c->plane, nl = c->nlev;
    /* newstep: check_cfl (901_debug.f90:85-99) then RHS_f = RHS (103:311-313) */
    double umax = max_lev(c, F(c, LES_F_U)), vmax = max_lev(c, F(c, LES_F_V)), wmax = max_lev(c, F(c, LES_F_W));
    c->cfl = umax * P->dt / c->dx + vmax * P->dt / c->dy + wmax * P->dt / c->dz;
    if (c->cfl > 0.6) return LES_E_CFL;
    memcpy(F(c, LES_F_RHSX_F), F(c, LES_F_RHSX), sizeof(double) * np * nl);
    memcpy(F(c, LES_F_RHSY_F), F(c, LES_F_RHSY), sizeof(double) * np * nl);
    memcpy(F(c, LES_F_RHSZ_F), F(c, LES_F_RHSZ), sizeof(double) * np * nl);
    orc_obukhov(c, jt);                                          /* :26 */
    orc_ddx(c, LES_F_DUDX, LES_F_U); orc_ddy(c, LES_F_DUDY, LES_F_U);   /* :30-38 */
    orc_ddx(c, LES_F_DVDX, LES_F_V); orc_ddy(c, LES_F_DVDY, LES_F_V);
    orc_ddx(c, LES_F_DWDX, LES_F_W); orc_ddy(c, LES_F_DWDY, LES_F_W);
    orc_ddz_uv(c, LES_F_DUDZ, LES_F_U); orc_ddz_uv(c, LES_F_DVDZ, LES_F_V);
    orc_ddz_w(c, LES_F_DWDZ, LES_F_W);
    orc_wallstress_bottom(c);                                    /* :48-52 */
    orc_wallstress_top(c);
    orc_sgs_stag(c, jt);                                         /* :59 */
    orc_divstress_uv(c, LES_F_DIVTX, LES_F_TXX, LES_F_TXY, LES_F_TXZ);   /* :82-84 */
    orc_divstress_uv(c, LES_F_DIVTY, LES_F_TXY, LES_F_TYY, LES_F_TYZ);
    orc_divstress_w(c, LES_F_DIVTZ, LES_F_TXZ, LES_F_TYZ, LES_F_TZZ);
    orc_convec(c);                                               /* :87 */
    orc_theta_all_in_one(c, jt);                                 /* :96-100 */
    /* RHS assembly (:115-157), damping (007_topbc.f90:113-118), AB2 (:319-389) */
    double *RX = F(c, LES_F_RHSX), *RY = F(c, LES_F_RHSY), *RZ = F(c, LES_F_RHSZ);
    const double *cx = F(c, LES_F_CX), *cy = F(c, LES_F_CY), *cz = F(c, LES_F_CZ);
    const double *dx = F(c, LES_F_DIVTX), *dy = F(c, LES_F_DIVTY), *dzt = F(c, LES_F_DIVTZ);
    const double* beta = F(c, LES_F_BETA_SCAL);
    double *u = F(c, LES_F_U), *v = F(c, LES_F_V), *w = F(c, LES_F_W);
    /* add_damping_layer (007_topbc.f90:92-124): method 1 damps w; methods 2
     * and 3 damp u, v, w toward their plane means */
    const int damp_sf = (P->ubc_mom == LES_BC_STRESS_FREE && P->ubc == 1);
    const int damp1 = damp_sf && P->damping_method == 1;
    const int damp23 = damp_sf && P->damping_method != 1;
    double* mn = NULL;
    if (damp23) {
        mn = (double*)calloc(3 * c->nlev, sizeof(double));
        for (int k = 1; k <= nz - 1; ++k) {
            mn[3 * k] = orc_hor_avg(c, LEV(c, u, k));
            mn[3 * k + 1] = orc_hor_avg(c, LEV(c, v, k));
            mn[3 * k + 2] = orc_hor_avg(c, LEV(c, w, k));
        }
    }
    for (int k = 1; k <= nz - 1; ++k)
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q;
            RX[o] = -cx[o] - dx[o];
            RY[o] = -cy[o] - dy[o];
            RZ[o] = -cz[o] - dzt[o];
            if (P->theta_flag && !P->passive_scalar) RZ[o] = RZ[o] + beta[o];
            if (P->ocean_flag) {                                 /* :134-149 */
                RX[o] = RX[o] + P->coriol * (v[o] + c->vst[k]) - P->coriol * P->vg;
                RY[o] = RY[o] - P->coriol * (u[o] + c->ust[k]) + P->coriol * P->ug;
                if (P->flag_crossvel) {
                    RX[o] = RX[o] - P->coriol * P->v_cross;
                    RY[o] = RY[o] + P->coriol * P->u_cross;
                }
            } else if (P->coriolis_forcing) {                    /* :150-156 */
                RX[o] = RX[o] + P->coriol * (v[o] - P->vg);
                RY[o] = RY[o] + P->coriol * (P->ug - u[o]);
            }
            if (damp1) RZ[o] = RZ[o] - c->sponge[k] * w[o];
            if (damp23) {
                const double sk = 0.5 * (c->sponge[k] + c->sponge[k + 1]);
                RX[o] = RX[o] - sk * (u[o] - mn[3 * k]);
                RY[o] = RY[o] - sk * (v[o] - mn[3 * k + 1]);
                RZ[o] = RZ[o] - c->sponge[k] * (w[o] - mn[3 * k + 2]);
            }
        }
    free(mn);
    const int st = stats_step(c, jt);
    if (st)                                                      /* dudt, :119-156 */
        for (int k = 1; k <= nz - 1; ++k)
            for (size_t q = 0; q < np; ++q) {
                size_t o = (size_t)k * np + q;
                c->dudt[o] = -dx[o];
                c->dvdt[o] = -dy[o];
                c->dwdt[o] = -dzt[o];
                if (P->ocean_flag) {
                    c->dudt[o] = c->dudt[o] + P->coriol * (v[o] + c->vst[k]) - P->coriol * P->vg;
                    c->dvdt[o] = c->dvdt[o] - P->coriol * (u[o] + c->ust[k]) + P->coriol * P->ug;
                    if (P->flag_crossvel) {
                        c->dudt[o] = c->dudt[o] - P->coriol * P->v_cross;
                        c->dvdt[o] = c->dvdt[o] + P->coriol * P->u_cross;
                    }
                } else if (P->coriolis_forcing) {
                    c->dudt[o] = c->dudt[o] + P->coriol * (v[o] - P->vg);
                    c->dvdt[o] = c->dvdt[o] + P->coriol * (P->ug - u[o]);
                }
            }
    orc_ab2(c, jt);                                              /* :319-389 */
    orc_press_stag_array(c, jt);                                 /* :168-175 */
    const double *dpdx = F(c, LES_F_DPDX), *dpdy = F(c, LES_F_DPDY), *dpdz = F(c, LES_F_DPDZ);
    for (int k = 1; k <= nz - 1; ++k)                            /* :179-181 */
        for (size_t q = 0; q < np; ++q) {
            size_t o = (size_t)k * np + q;
            RX[o] = RX[o] - dpdx[o];
            RY[o] = RY[o] - dpdy[o];
            RZ[o] = RZ[o] - dpdz[o];
            if (st) {                                            /* :183-185 */
                c->dudt[o] = c->dudt[o] - dpdx[o];
                c->dvdt[o] = c->dvdt[o] - dpdy[o];
                c->dwdt[o] = c->dwdt[o] - dpdz[o];
            }
        }
    orc_project(c);                                              /* :188-189 */
    if (st) diag_ke(c);                                          /* :192-213 */
    orc_filter_data(c, LES_F_U);                                 /* :227-229 */
    orc_filter_data(c, LES_F_V);
    orc_filter_data(c, LES_F_W);
    if (st) {                                                    /* :231-236 */
        flow_slice(c, jt);
        if (P->theta_flag && jt >= P->theta_init_time) theta_slice(c, jt);
    }
    return LES_OK;
}

int orc_run(les_ctx* c, int jt0, int nsteps, int* failed_jt) {
    *failed_jt = 0;
    for (int s = 0; s < nsteps; ++s) {
        int st = orc_step(c, jt0 + s);
        if (st != LES_OK) { *failed_jt = jt0 + s; return st; }
    }
    return LES_OK;
}

int orc_set_precision(les_ctx* c, int precision) {
    (void)c;
    return precision == LES_FP64 ? LES_OK : LES_E_BADARG;
}

int orc_lagrange_sdep(les_ctx* c, int jt) { lagrange_sdep(c, jt); return LES_OK; }
int orc_interpolag_sdep(les_ctx* c) { interpolag_sdep(c); return LES_OK; }
