/*
 * ORACLE -- test infrastructure only.  CPU FP64 restatement of the
 * GAbelois/OverlappingBLs timestep (the SBL/code Fortran files).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker.  The product (overlappingbls_amd) never calls it.
 *
 * Internal declarations shared by the oracle's translation units.
 */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H

#include <complex.h>
#include <stddef.h>
#include "../include/les_gpu.h"

typedef double complex cplx;

/* longest transform: the 3/2 grid of nxt, nyt <= 2048 (les_create's limit) */
#define ORC_NMAX 3072

#define ORC_BOGUS (-1234567890.0)   /* 001_param.f90:40 */
#define ORC_VONK 0.4                /* 001_param.f90:32 */
#define ORC_G 9.81                  /* 001_param.f90:33 */
#define ORC_PI 3.1415926535897932384626433
#define ORC_B_SMOOTH 5.2            /* 008_bottombc.f90:25 */
#define ORC_OPFTIME 1.5             /* 005_sgs.f90:12 */

/* complex FFT plan, n = 2^a 3^b, Stockham autosort */
typedef struct {
    int n;
    int nstage;
    int radix[32];
    cplx* tw_fwd;   /* exp(-2 pi i m / n) */
    cplx* tw_bwd;   /* exp(+2 pi i m / n) */
} orc_plan;

int orc_plan_init(orc_plan* p, int n);
void orc_plan_free(orc_plan* p);
/* sign -1 forward (FFTW_FORWARD), +1 backward; unnormalised; x, work length n */
void orc_fft(const orc_plan* p, cplx* x, cplx* work, int sign);
/* FFTW r2c / c2r semantics (unnormalised) */
void orc_r2c(const orc_plan* p, const double* in, ptrdiff_t istride, cplx* out);
void orc_c2r(const orc_plan* p, const cplx* in, double* out, ptrdiff_t ostride);

struct les_ctx {
    les_params prm;
    int nx, ny, nz;          /* nxt, nyt, nzt (levels 0..nz)                 */
    int nx2, ny2;            /* 3/2 grid                                     */
    int lhx, lhy;
    size_t plane, plane2;    /* nx*ny, nx2*ny2                               */
    size_t nlev;             /* nz+1                                         */
    double dx, dy, dz;
    orc_plan px, py, px2, py2;
    double *kx, *ky;         /* 1-D wavenumbers with Nyquist zeroed (002_fft.f90:131-140) */
    double *kx2d, *ky2d, *k2;/* (lhx, nyt), Nyquist rules 002_fft.f90:143-167 */
    unsigned char* notnyq;   /* (lhx, nyt)                                   */
    double *Gtest, *Gtest2;  /* G_test (2 delta), G_test_test (4 delta)      */
    double *coef_a, *coef_b, *coef_c;  /* TDMA (lhx*nyt) x nz, factorised    */
    double *ust, *vst;       /* Stokes drift (0:nz)                          */
    double *sponge;          /* (0:nz)                                        */
    double* f3[LES_N_FIELDS3D];
    double *ustar_avg, *psi_m, *phi_m, *psi_h, *phi_h;  /* surface planes */
    double* ustar_top;       /* ocean obukhov: ustar_avg of the top wall     */
    double wind_ustar, wind_rad;  /* ocean wind stress (u_star, rad_stress) */
    double *u_big, *v_big, *w_big;     /* padded velocity kept for theta (convec) */
    int f_lm_mm_init, f_qn_nn_init;
    int sgs_frozen;
    /* les_bind_field registrations (host arrays the caller owns) */
    int nbind;
    struct { int id; double* h; int ld, lev0, nlev; } bind[128];
    double cfl;
    int nthreads;
    /* time-averaged slices (011_scalars2.f90) and the flosol diagnostics
     * they read (103_flosol.f90:119-213) */
    int use_avgslice, c_count, p_count;
    double fr;
    double *dudt, *dvdt, *dwdt, *ke, *ke_temp, *dke;   /* nlev planes each */
    double *sl_run, *sl_win;   /* LES_N_SLICES x (nz-1) x nx */
};

#define IDX(c, k, j, i) (((size_t)(k) * (c)->ny + (size_t)(j)) * (c)->nx + (size_t)(i))
#define LEV(c, f, k) ((f) + (size_t)(k) * (c)->plane)

/* transforms (orc_transforms.c) */
void orc_fftn_plane(les_ctx* c, const double* in, cplx* out);       /* 203_fftn_mpi.f90:4-35  */
void orc_ifftn_plane(les_ctx* c, const cplx* in, double* out);      /* 203_fftn_mpi.f90:39-67 */
void orc_pad_plane(les_ctx* c, const double* var, double* big);   /* 108_01_padd.f90:73-132 */
void orc_unpad_plane(les_ctx* c, const double* big, double* var); /* 108_01_padd.f90:136-197 */
void orc_test_filter_plane(les_ctx* c, double* f, const double* G);
void orc_ddx_plane(les_ctx* c, double* d, const double* f);
void orc_ddy_plane(les_ctx* c, double* d, const double* f);
double orc_hor_avg(les_ctx* c, const double* plane);          /* calc_hor_avg3 */

#endif
