// Output files of the reference, written and read on the host: the restart
// files and the time-averaged slices.
//
// Restart files: checkpoint_final (012_io.f90:496-546,
// :548-559) writes them; read_flow_field (102_01_ic.f90:682-734) and
// init_temperature_field (102_initial.f90:740-797) read them.
//
// Format: MPI-IO with a 'native' view of the global array (nxt, nyt, nzt-1)
// in Fortran order (mpi_type_create_subarray, 202_parallel.f90:1084-1117;
// sizes and starts 102_initial.f90:154-162).  The file is therefore plain
// native-endian doubles: x fastest, then y, then z over the global levels
// 1..nzt-1, one whole field after another.  A plane field
// (write_file_mpi_xy, 202_parallel.f90:1252-1284) is the (nxt, nyt) array.
//   vel_tt<nums>.out:  u, v, w, RHSx, RHSy, RHSz, Cs_opt2, F_LM, F_MM, F_QN, F_NN
//   temp_tt<nums>.out: theta, RHS_T, sgs_t3(:,:,1), psi_m
// The host side moves fields through les_get_field / les_set_field, so both
// precisions read and write the FP64 values the caller would see.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "les_internal.hpp"

namespace {

constexpr int VEL_FIELDS[] = {LES_F_U,       LES_F_V,    LES_F_W,    LES_F_RHSX, LES_F_RHSY, LES_F_RHSZ,
                              LES_F_CS_OPT2, LES_F_F_LM, LES_F_F_MM, LES_F_F_QN, LES_F_F_NN};
constexpr int N_VEL = sizeof(VEL_FIELDS) / sizeof(VEL_FIELDS[0]);

// path_restart//'vel_tt'//i8.8//'.out' (012_io.f90:524)
bool file_name(std::string& out, const char* path, const char* stem, int nums) {
    if (!path || nums < 0 || nums > 99999999) return false;
    char b[32];
    std::snprintf(b, sizeof b, "%s%08d.out", stem, nums);
    out = std::string(path) + b;
    return true;
}

struct File {
    std::FILE* f = nullptr;
    bool close() {  // false when buffered data could not be written
        const int rc = std::fclose(f);
        f = nullptr;
        return rc == 0;
    }
    ~File() {
        if (f) std::fclose(f);
    }
};

// levels 1..nzt-1 of a field <-> the next block of the file
int put_field(les_ctx* c, std::FILE* f, int id, std::vector<double>& buf, int lev0, int nlev) {
    const size_t n = (size_t)nlev * c->plane;
    buf.resize(n);
    const int st = les_get_field(c, id, buf.data(), c->nx, lev0, nlev);
    if (st != LES_OK) return st;
    return std::fwrite(buf.data(), sizeof(double), n, f) == n ? LES_OK : LES_E_IO;
}
int take_field(les_ctx* c, std::FILE* f, int id, std::vector<double>& buf, int lev0, int nlev) {
    const size_t n = (size_t)nlev * c->plane;
    buf.resize(n);
    if (std::fread(buf.data(), sizeof(double), n, f) != n) return LES_E_IO;
    return les_set_field(c, id, buf.data(), c->nx, lev0, nlev);
}

// the file holds exactly `bytes` bytes
bool has_size(std::FILE* f, long long bytes) {
    if (std::fseek(f, 0, SEEK_END) != 0) return false;
    const long long n = std::ftell(f);
    std::rewind(f);
    return n == bytes;
}

// one plane of v at level lev of field id
int fill_plane(les_ctx* c, int id, int lev, double v) {
    std::vector<double> p((size_t)c->plane, v);
    return les_set_field(c, id, p.data(), c->nx, lev, 1);
}

}  // namespace

int les_write_restart(les_ctx* c, const char* path, int nums) {
    if (!c) return LES_E_BADARG;
    std::string fn;
    if (!file_name(fn, path, "vel_tt", nums)) return LES_E_BADARG;
    const int nl = c->nz - 1;
    std::vector<double> buf;
    {
        File f;
        f.f = std::fopen(fn.c_str(), "wb");
        if (!f.f) return LES_E_IO;
        for (int m = 0; m < N_VEL; ++m) {
            const int st = put_field(c, f.f, VEL_FIELDS[m], buf, 1, nl);
            if (st != LES_OK) return st;
        }
        if (!f.close()) return LES_E_IO;
    }
    if (!c->prm.theta_flag) return LES_OK;
    if (!file_name(fn, path, "temp_tt", nums)) return LES_E_BADARG;
    File f;
    f.f = std::fopen(fn.c_str(), "wb");
    if (!f.f) return LES_E_IO;
    int st = put_field(c, f.f, LES_F_THETA, buf, 1, nl);
    if (st == LES_OK) st = put_field(c, f.f, LES_F_RHST, buf, 1, nl);
    if (st == LES_OK) st = put_field(c, f.f, LES_F_SGS_T3, buf, 1, 1);
    if (st == LES_OK) st = put_field(c, f.f, LES_F_PSI_M_SFC, buf, 0, 1);
    if (st == LES_OK && !f.close()) st = LES_E_IO;
    return st;
}

int les_read_restart(les_ctx* c, const char* path, int nums, int read_theta) {
    if (!c) return LES_E_BADARG;
    if (read_theta && !c->prm.theta_flag) return LES_E_BADARG;  // 102_initial.f90:798-803
    std::string fn;
    if (!file_name(fn, path, "vel_tt", nums)) return LES_E_BADARG;
    const int nl = c->nz - 1;
    const long long F = (long long)c->plane * nl * (long long)sizeof(double);
    const long long P = (long long)c->plane * (long long)sizeof(double);
    std::vector<double> buf;
    {
        File f;
        f.f = std::fopen(fn.c_str(), "rb");
        if (!f.f || !has_size(f.f, N_VEL * F)) return LES_E_IO;
        for (int m = 0; m < N_VEL; ++m) {
            const int st = take_field(c, f.f, VEL_FIELDS[m], buf, 1, nl);
            if (st != LES_OK) return st;
        }
    }
    // the ghost rows data_exchange leaves at the physical boundaries
    // (202_parallel.f90:722-733): levels 0 of u, v, w and nzt of u, v are
    // BOGUS; w at nzt is the top boundary's 0 (110_project.f90:48-55)
    int st = LES_OK;
    for (int id : {LES_F_U, LES_F_V, LES_F_W})
        if (st == LES_OK) st = fill_plane(c, id, 0, les::BOGUS);
    for (int id : {LES_F_U, LES_F_V})
        if (st == LES_OK) st = fill_plane(c, id, c->nz, les::BOGUS);
    if (st == LES_OK) st = fill_plane(c, LES_F_W, c->nz, 0.0);
    if (st != LES_OK || !read_theta) return st;
    if (!file_name(fn, path, "temp_tt", nums)) return LES_E_BADARG;
    File f;
    f.f = std::fopen(fn.c_str(), "rb");
    if (!f.f || !has_size(f.f, 2 * F + 2 * P)) return LES_E_IO;
    st = take_field(c, f.f, LES_F_THETA, buf, 1, nl);
    if (st == LES_OK) st = take_field(c, f.f, LES_F_RHST, buf, 1, nl);
    if (st == LES_OK) st = take_field(c, f.f, LES_F_SGS_T3, buf, 1, 1);
    if (st == LES_OK) st = take_field(c, f.f, LES_F_PSI_M_SFC, buf, 0, 1);
    if (st == LES_OK) st = fill_plane(c, LES_F_THETA, 0, les::BOGUS);
    if (st == LES_OK) st = fill_plane(c, LES_F_THETA, c->nz, les::BOGUS);
    return st;
}

// ---- time-averaged slices: write_avgslice (011_scalars2.f90:393-442) -----------
// Appends to <path>result/aver_<name>.out, names as flow_slice / theta_slice
// pass them (011_scalars2.f90:203-244, 305-311).  average_dim_num 1: one line
// per level 1..nzt-1 holding nums*dt and the slice's nxt values; 2: one line
// holding nums*dt and the x-mean of each level (sum(arr2d, dim=1)/nxt, summed
// in x order).  Every value is written with Fortran's E14.5 edit descriptor
// (fmt_5168 = "(1400(E14.5))", :21): right-justified in 14 characters as
// [-]0.ddddd followed by E+ee, or +eee without the E when the exponent has
// three digits.  The 5 significant digits come from the C library's
// correctly rounded conversion.
namespace {

const char* const SLICE_NAMES[LES_N_SLICES] = {
    "u", "v", "w", "p", "dudt", "dvdt", "dwdt", "dudz", "dvdz",
    "u2", "v2", "w2", "p2", "uv", "uw", "vw", "wp",
    "txx", "txy", "txz", "tyy", "tyz", "tzz",
    "u3", "v3", "w3",
    "u2fluc", "v2fluc", "uvfluc", "u2wfluc", "v2wfluc", "dissip",
    "utxz_fluc", "vtyz_fluc", "wtzz", "Cs",
    "theta", "theta2", "Tfluc", "dTdz", "sgs_t3", "wt", "Nu_t"};

// Fortran Ew.d output of x into out (w characters, no terminator)
void fortran_e(char* out, double x, int w, int d) {
    char t[64];
    int n;
    if (std::isnan(x)) {
        n = std::snprintf(t, sizeof t, "NaN");
    } else if (std::isinf(x)) {
        n = std::snprintf(t, sizeof t, "%s", x < 0 ? "-Infinity" : "Infinity");
    } else {
        char m[64];
        int e10 = 0;
        std::string digits(d, '0');
        if (x != 0.0) {
            std::snprintf(m, sizeof m, "%.*e", d - 1, std::fabs(x));  // D.DDDDe+XX
            digits = std::string(1, m[0]) + std::string(m + 2, m + 1 + d);
            e10 = std::atoi(std::strchr(m, 'e') + 1) + 1;  // 0.DDDDD x 10^e10
        }
        const char* sg = std::signbit(x) ? "-" : "";
        if (std::abs(e10) <= 99)
            n = std::snprintf(t, sizeof t, "%s0.%sE%c%02d", sg, digits.c_str(), e10 < 0 ? '-' : '+', std::abs(e10));
        else
            n = std::snprintf(t, sizeof t, "%s0.%s%c%03d", sg, digits.c_str(), e10 < 0 ? '-' : '+', std::abs(e10));
    }
    if (n > w) {
        for (int i = 0; i < w; ++i) out[i] = '*';
        return;
    }
    for (int i = 0; i < w - n; ++i) out[i] = ' ';
    for (int i = 0; i < n; ++i) out[w - n + i] = t[i];
}

void put_line(std::string& line, double v) {
    char b[14];
    fortran_e(b, v, 14, 5);
    line.append(b, 14);
}

// item i of an n-item record of fmt_5168 = "(1400(E14.5))": format reversion
// starts a new line after every 1400 items; the record ends with a newline
void put_item(std::string& text, double v, int i, int n) {
    put_line(text, v);
    if ((i + 1) % 1400 == 0 || i + 1 == n) text += "\n";
}

}  // namespace

int les_write_avgslice(les_ctx* c, const char* path, int nums, int average_dim_num) {
    if (!c || !path || !c->st_run || (average_dim_num != 1 && average_dim_num != 2)) return LES_E_BADARG;
    const int nx = c->nx, nl = c->nz - 1;
    const int ns = c->prm.theta_flag ? (int)LES_N_SLICES : (int)LES_N_FLOW_SLICES;
    const double time = nums * c->prm.dt;
    std::vector<double> a((size_t)nx * nl);
    for (int s = 0; s < ns; ++s) {
        const int st = les_get_avgslice(c, s, 1, a.data());
        if (st != LES_OK) return st;
        std::string text;
        if (average_dim_num == 1) {
            for (int k = 0; k < nl; ++k) {
                put_item(text, time, 0, nx + 1);
                for (int i = 0; i < nx; ++i) put_item(text, a[(size_t)k * nx + i], i + 1, nx + 1);
            }
        } else {
            put_item(text, time, 0, nl + 1);
            for (int k = 0; k < nl; ++k) {
                double sum = 0.0;
                for (int i = 0; i < nx; ++i) sum += a[(size_t)k * nx + i];
                put_item(text, sum / nx, k + 1, nl + 1);
            }
        }
        const std::string fn = std::string(path) + "result/aver_" + SLICE_NAMES[s] + ".out";
        File f;
        f.f = std::fopen(fn.c_str(), "ab");
        if (!f.f) return LES_E_IO;
        if (std::fwrite(text.data(), 1, text.size(), f.f) != text.size() || !f.close()) return LES_E_IO;
    }
    return LES_OK;
}
