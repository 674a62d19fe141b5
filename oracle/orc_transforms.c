/*
 * ORACLE -- test infrastructure only (see orc_internal.h).
 *
 * Plane transforms of the reference, restated line by line on one whole
 * x-y plane (npy = 1, so every transpose of 203_fftn_mpi.f90:71-371 is the
 * identity permutation of a global plane).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc_internal.h"

/* fftn_mpi, 203_fftn_mpi.f90:4-35.  out is (lhx, nyt) complex, kx fastest. */
void orc_fftn_plane(les_ctx* c, const double* in, cplx* out) {
    const int nx = c->nx, ny = c->ny, lhx = c->lhx;
    cplx col[ORC_NMAX], work[ORC_NMAX];
    /* x r2c per row (:21-23) */
    for (int j = 0; j < ny; ++j) orc_r2c(&c->px, in + (size_t)j * nx, 1, out + (size_t)j * lhx);
    /* in_hat_x(lhx,:) = 0 (:25) */
    for (int j = 0; j < ny; ++j) out[(size_t)j * lhx + lhx - 1] = 0;
    /* y c2c forward per kx column (:29-31) */
    for (int i = 0; i < lhx; ++i) {
        for (int j = 0; j < ny; ++j) col[j] = out[(size_t)j * lhx + i];
        orc_fft(&c->py, col, work, -1);
        for (int j = 0; j < ny; ++j) out[(size_t)j * lhx + i] = col[j];
    }
    /* out(:, nyt/2+1) = 0 (:33) */
    for (int i = 0; i < lhx; ++i) out[(size_t)(ny / 2) * lhx + i] = 0;
}

/* ifftn_mpi, 203_fftn_mpi.f90:39-67. */
void orc_ifftn_plane(les_ctx* c, const cplx* in, double* out) {
    const int nx = c->nx, ny = c->ny, lhx = c->lhx;
    cplx* tmp = (cplx*)malloc(sizeof(cplx) * (size_t)lhx * ny);
    cplx col[ORC_NMAX], work[ORC_NMAX];
    for (int i = 0; i < lhx; ++i) {
        for (int j = 0; j < ny; ++j) col[j] = in[(size_t)j * lhx + i];
        orc_fft(&c->py, col, work, +1);
        for (int j = 0; j < ny; ++j) tmp[(size_t)j * lhx + i] = col[j];
    }
    for (int j = 0; j < ny; ++j) orc_c2r(&c->px, tmp + (size_t)j * lhx, out + (size_t)j * nx, 1);
    free(tmp);
}

/* test_filter, 006_test_filter.f90:19-37 (G carries the 1/(nxt nyt)). */
void orc_test_filter_plane(les_ctx* c, double* f, const double* G) {
    const size_t ns = (size_t)c->lhx * c->ny;
    cplx* tmp = (cplx*)malloc(sizeof(cplx) * ns);
    orc_fftn_plane(c, f, tmp);
    for (size_t q = 0; q < ns; ++q) tmp[q] = G[q] * tmp[q];
    orc_ifftn_plane(c, tmp, f);
    free(tmp);
}

/* ddx on one plane, 004_derivatives.f90:31-39. */
void orc_ddx_plane(les_ctx* c, double* d, const double* f) {
    const int nx = c->nx, ny = c->ny, lhx = c->lhx;
    const double cst = 1.0 / nx;
    double tmp[ORC_NMAX];
    cplx arr[ORC_NMAX];
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) tmp[i] = cst * f[(size_t)j * nx + i];
        orc_r2c(&c->px, tmp, 1, arr);
        for (int i = 0; i < lhx; ++i) arr[i] = (I * c->kx[i]) * arr[i];
        orc_c2r(&c->px, arr, d + (size_t)j * nx, 1);
    }
}

/* ddy on one plane, 004_derivatives.f90:60-79. */
void orc_ddy_plane(les_ctx* c, double* d, const double* f) {
    const int nx = c->nx, ny = c->ny, lhy = c->lhy;
    const double cst = 1.0 / ny;
    double tmp[ORC_NMAX];
    cplx arr[ORC_NMAX];
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) tmp[j] = cst * f[(size_t)j * nx + i];
        orc_r2c(&c->py, tmp, 1, arr);
        for (int j = 0; j < lhy; ++j) arr[j] = (I * c->ky[j]) * arr[j];
        orc_c2r(&c->py, arr, d + i, nx);
    }
}

/* zero_padding, 108_01_padd.f90:73-132: (nxt, nyt) -> (nxt2, nyt2). */
void orc_pad_plane(les_ctx* c, const double* var, double* big) {
    const int nx = c->nx, ny = c->ny, nx2 = c->nx2, ny2 = c->ny2;
    double* xbig = (double*)malloc(sizeof(double) * (size_t)nx2 * ny);
    double tmp[ORC_NMAX];
    cplx hat[ORC_NMAX], hatbig[ORC_NMAX];
    /* x direction: tmp = var/nxt, r2c(nxt), keep 1..nxt/2, c2r(nxt2) (:101-117) */
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) tmp[i] = var[(size_t)j * nx + i] / nx;
        orc_r2c(&c->px, tmp, 1, hat);
        for (int k = 0; k < nx2 / 2 + 1; ++k) hatbig[k] = 0;
        for (int k = 0; k < nx / 2; ++k) hatbig[k] = hat[k];
        orc_c2r(&c->px2, hatbig, xbig + (size_t)j * nx2, 1);
    }
    /* y direction: var_y/nyt, r2c(nyt), keep 1..nyt/2, c2r(nyt2) (:119-131) */
    for (int i = 0; i < nx2; ++i) {
        for (int j = 0; j < ny; ++j) tmp[j] = xbig[(size_t)j * nx2 + i] / ny;
        orc_r2c(&c->py, tmp, 1, hat);
        for (int k = 0; k < ny2 / 2 + 1; ++k) hatbig[k] = 0;
        for (int k = 0; k < ny / 2; ++k) hatbig[k] = hat[k];
        orc_c2r(&c->py2, hatbig, big + i, nx2);
    }
    free(xbig);
}

/* zero_unpadding, 108_01_padd.f90:136-197: (nxt2, nyt2) -> (nxt, nyt). */
void orc_unpad_plane(les_ctx* c, const double* big, double* var) {
    const int nx = c->nx, ny = c->ny, nx2 = c->nx2, ny2 = c->ny2;
    double* xs = (double*)malloc(sizeof(double) * (size_t)nx * ny2);
    double tmp[ORC_NMAX];
    cplx hatbig[ORC_NMAX], hat[ORC_NMAX];
    /* tmp = var_big/nxt2; r2c(nxt2); keep 1..nxt/2; Nyquist 0; c2r(nxt) (:157-172) */
    for (int j = 0; j < ny2; ++j) {
        for (int i = 0; i < nx2; ++i) tmp[i] = big[(size_t)j * nx2 + i] / nx2;
        orc_r2c(&c->px2, tmp, 1, hatbig);
        for (int k = 0; k < nx / 2; ++k) hat[k] = hatbig[k];
        hat[nx / 2] = 0;
        orc_c2r(&c->px, hat, xs + (size_t)j * nx, 1);
    }
    /* var_y_big/nyt2; r2c(nyt2); keep 1..nyt/2; Nyquist 0; c2r(nyt) (:176-190) */
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny2; ++j) tmp[j] = xs[(size_t)j * nx + i] / ny2;
        orc_r2c(&c->py2, tmp, 1, hatbig);
        for (int k = 0; k < ny / 2; ++k) hat[k] = hatbig[k];
        hat[ny / 2] = 0;
        orc_c2r(&c->py, hat, var + i, nx);
    }
    free(xs);
}

/* calc_hor_avg3, 202_parallel.f90:1039-1059: sum / (nxt nyt). */
double orc_hor_avg(les_ctx* c, const double* plane) {
    double s = 0.0;
    for (size_t q = 0; q < c->plane; ++q) s += plane[q];
    return s / (double)(c->nx * c->ny);
}
