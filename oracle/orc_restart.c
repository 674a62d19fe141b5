/* Oracle restatement of the reference's restart files (test infrastructure
 * only): checkpoint_final (SBL/code/012_io.f90:496-559) and the readers
 * read_flow_field (102_01_ic.f90:682-734) and init_temperature_field
 * (102_initial.f90:740-797).  Each field goes through
 * write_file_mpi_3d / write_file_mpi_xy (202_parallel.f90:1084-1117,
 * 1252-1284): an MPI-IO subarray of the global (nxt, nyt, nzt-1) array in
 * Fortran order with a 'native' view, so the file is the global array's
 * doubles, x fastest, levels 1..nzt-1, fields one after another at
 * displacements disp3d = nxt*nyt*(nzt-1)*8 (012_io.f90:520-521).  Written
 * here element by element from the oracle's level-major arrays. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "orc_internal.h"

static const int vel_ids[11] = {LES_F_U, LES_F_V, LES_F_W, LES_F_RHSX, LES_F_RHSY, LES_F_RHSZ,
                                LES_F_CS_OPT2, LES_F_F_LM, LES_F_F_MM, LES_F_F_QN, LES_F_F_NN};

static FILE* open_named(const char* path, const char* stem, int nums, const char* mode) {
    char fn[4096];
    if (!path || nums < 0 || nums > 99999999) return NULL;
    if (snprintf(fn, sizeof fn, "%s%s%08d.out", path, stem, nums) >= (int)sizeof fn) return NULL;
    return fopen(fn, mode);
}

static const double* field(les_ctx* c, int id) {
    return id == LES_F_PSI_M_SFC ? c->psi_m : c->f3[id];
}

/* levels k0..k1 of a field, element by element in Fortran order (i, j, k) */
static int put(les_ctx* c, FILE* f, int id, int k0, int k1) {
    const double* a = field(c, id);
    for (int k = k0; k <= k1; ++k)
        for (int j = 0; j < c->ny; ++j)
            for (int i = 0; i < c->nx; ++i) {
                const double v = a[(size_t)k * c->plane + (size_t)j * c->nx + i];
                if (fwrite(&v, sizeof v, 1, f) != 1) return LES_E_IO;
            }
    return LES_OK;
}

static int take(les_ctx* c, FILE* f, int id, int k0, int k1) {
    double* a = (double*)field(c, id);
    for (int k = k0; k <= k1; ++k)
        for (int j = 0; j < c->ny; ++j)
            for (int i = 0; i < c->nx; ++i)
                if (fread(&a[(size_t)k * c->plane + (size_t)j * c->nx + i], sizeof(double), 1, f) != 1)
                    return LES_E_IO;
    return LES_OK;
}

int orc_write_restart(les_ctx* c, const char* path, int nums) {
    FILE* f = open_named(path, "vel_tt", nums, "wb");
    int st = f ? LES_OK : LES_E_IO;
    for (int m = 0; m < 11 && st == LES_OK; ++m) st = put(c, f, vel_ids[m], 1, c->nz - 1);
    if (f && fclose(f) != 0) st = LES_E_IO;
    if (st != LES_OK || !c->prm.theta_flag) return st;
    f = open_named(path, "temp_tt", nums, "wb");
    if (!f) return LES_E_IO;
    st = put(c, f, LES_F_THETA, 1, c->nz - 1);
    if (st == LES_OK) st = put(c, f, LES_F_RHST, 1, c->nz - 1);
    if (st == LES_OK) st = put(c, f, LES_F_SGS_T3, 1, 1);
    if (st == LES_OK) st = put(c, f, LES_F_PSI_M_SFC, 0, 0);
    if (fclose(f) != 0) st = LES_E_IO;
    return st;
}

/* after reading: data_exchange's ghost rows at the physical boundaries
 * (202_parallel.f90:722-733) and project's w(nz) = 0 (110_project.f90:54) */
int orc_read_restart(les_ctx* c, const char* path, int nums, int read_theta) {
    if (read_theta && !c->prm.theta_flag) return LES_E_BADARG;
    FILE* f = open_named(path, "vel_tt", nums, "rb");
    int st = f ? LES_OK : LES_E_IO;
    for (int m = 0; m < 11 && st == LES_OK; ++m) st = take(c, f, vel_ids[m], 1, c->nz - 1);
    if (st == LES_OK && fgetc(f) != EOF) st = LES_E_IO; /* longer than the grid's file */
    if (f) fclose(f);
    if (st != LES_OK) return st;
    const size_t top = (size_t)c->nz * c->plane;
    for (size_t q = 0; q < c->plane; ++q) {
        c->f3[LES_F_U][q] = ORC_BOGUS;
        c->f3[LES_F_V][q] = ORC_BOGUS;
        c->f3[LES_F_W][q] = ORC_BOGUS;
        c->f3[LES_F_U][top + q] = ORC_BOGUS;
        c->f3[LES_F_V][top + q] = ORC_BOGUS;
        c->f3[LES_F_W][top + q] = 0.0;
    }
    if (!read_theta) return LES_OK;
    f = open_named(path, "temp_tt", nums, "rb");
    if (!f) return LES_E_IO;
    st = take(c, f, LES_F_THETA, 1, c->nz - 1);
    if (st == LES_OK) st = take(c, f, LES_F_RHST, 1, c->nz - 1);
    if (st == LES_OK) st = take(c, f, LES_F_SGS_T3, 1, 1);
    if (st == LES_OK) st = take(c, f, LES_F_PSI_M_SFC, 0, 0);
    if (st == LES_OK && fgetc(f) != EOF) st = LES_E_IO;
    fclose(f);
    if (st != LES_OK) return st;
    for (size_t q = 0; q < c->plane; ++q) {
        c->f3[LES_F_THETA][q] = ORC_BOGUS;
        c->f3[LES_F_THETA][top + q] = ORC_BOGUS;
    }
    return LES_OK;
}

/* write_avgslice (011_scalars2.f90:393-442): one file per slice,
 * path//'result/aver_'//name//'.out' opened for appending, records written
 * with fmt_5168 = "(1400(E14.5))" (:21).  Names as flow_slice and
 * theta_slice pass them (:203-244, 305-311), indexed by les_slice. */
int orc_get_avgslice(les_ctx* c, int slice, int window, double* host); /* orc_step.c */

static const char* const orc_slice_names[LES_N_SLICES] = {
    "u", "v", "w", "p", "dudt", "dvdt", "dwdt", "dudz", "dvdz", "u2", "v2", "w2", "p2", "uv", "uw", "vw",
    "wp", "txx", "txy", "txz", "tyy", "tyz", "tzz", "u3", "v3", "w3", "u2fluc", "v2fluc", "uvfluc",
    "u2wfluc", "v2wfluc", "dissip", "utxz_fluc", "vtyz_fluc", "wtzz", "Cs", "theta", "theta2", "Tfluc",
    "dTdz", "sgs_t3", "wt", "Nu_t"};

/* E14.5 edit descriptor: the value as 0.d1d2d3d4d5 x 10^e, d1 != 0 unless
 * the value is zero, in a 14-character field padded on the left; the
 * exponent as E+ee, or +eee when |e| > 99 */
static void orc_e14_5(FILE* f, double x) {
    char field[32];
    if (x != x) {
        snprintf(field, sizeof field, "NaN");
    } else if (isinf(x)) {
        snprintf(field, sizeof field, "%s", x < 0 ? "-Infinity" : "Infinity");
    } else if (x == 0.0) {
        snprintf(field, sizeof field, "%s0.00000E+00", signbit(x) ? "-" : "");
    } else {
        char sci[32];
        snprintf(sci, sizeof sci, "%.4E", fabs(x)); /* d.ddddE+xx, rounded to 5 digits */
        const int e = atoi(strchr(sci, 'E') + 1) + 1;
        char mant[6] = {sci[0], sci[2], sci[3], sci[4], sci[5], 0};
        if (e >= -99 && e <= 99)
            snprintf(field, sizeof field, "%s0.%sE%c%02d", x < 0 ? "-" : "", mant, e < 0 ? '-' : '+', e < 0 ? -e : e);
        else
            snprintf(field, sizeof field, "%s0.%s%c%03d", x < 0 ? "-" : "", mant, e < 0 ? '-' : '+', e < 0 ? -e : e);
    }
    fprintf(f, "%14s", field);
}

/* One record of fmt_5168: item i written, a new line after every 1400 items
 * (format reversion) and at the end of the record */
static void orc_record_item(FILE* f, double x, int i, int n) {
    orc_e14_5(f, x);
    if ((i + 1) % 1400 == 0 || i + 1 == n) fputc('\n', f);
}

int orc_write_avgslice(les_ctx* c, const char* path, int nums, int average_dim_num) {
    if (!path || !c->sl_run || (average_dim_num != 1 && average_dim_num != 2)) return LES_E_BADARG;
    const int nx = c->nx, nl = c->nz - 1, ns = c->prm.theta_flag ? LES_N_SLICES : LES_N_FLOW_SLICES;
    const double t = nums * c->prm.dt;
    double* a = malloc(sizeof(double) * (size_t)nx * nl);
    if (!a) return LES_E_NOMEM;
    int st = LES_OK;
    for (int s = 0; s < ns && st == LES_OK; ++s) {
        char fn[4096];
        st = orc_get_avgslice(c, s, 1, a);
        if (st != LES_OK) break;
        snprintf(fn, sizeof fn, "%sresult/aver_%s.out", path, orc_slice_names[s]);
        FILE* f = fopen(fn, "a");
        if (!f) { st = LES_E_IO; break; }
        if (average_dim_num == 1) {
            for (int k = 0; k < nl; ++k) {
                orc_record_item(f, t, 0, nx + 1);
                for (int i = 0; i < nx; ++i) orc_record_item(f, a[(size_t)k * nx + i], i + 1, nx + 1);
            }
        } else {
            orc_record_item(f, t, 0, nl + 1);
            for (int k = 0; k < nl; ++k) {
                double sum = 0.0; /* sum(arr2d(:, k), dim=1) in x order */
                for (int i = 0; i < nx; ++i) sum += a[(size_t)k * nx + i];
                orc_record_item(f, sum / nx, k + 1, nl + 1);
            }
        }
        if (fclose(f) != 0) st = LES_E_IO;
    }
    free(a);
    return st;
}
