// Time-averaged y-averaged x-z slices (flow_slice, theta_slice:
// 011_scalars2.f90:28-318) and the flosol diagnostics they read (dudt,
// dvdt, dwdt: 103_flosol.f90:119-213; energy :393-429).
//
// One thread per (x, level) walks its column in y in order, as avg_y3d
// (202_parallel.f90:966-989) sums it, and adds fr * (sum / nyt) to the
// running slice.  Accumulators and diagnostics are FP64 in both precisions.
#include "les_internal.hpp"

namespace les {

namespace {

struct SP {
    long plane;
    int nx, ny, nz;
    double fr, l2;
};

SP sp(const les_ctx* c) {
    SP s;
    s.plane = (long)c->plane;
    s.nx = c->nx;
    s.ny = c->ny;
    s.nz = c->nz;
    s.fr = c->st_fr;
    const double delta = 1.0 * pow(c->dx * c->dy * c->dz, 1.0 / 3.0);  // l = delta, 106_sgs_stag.f90:33,181
    s.l2 = delta * delta;
    return s;
}

#define COL_LOOP(x, k, s)                                                                       \
    const long t_ = (long)blockIdx.x * blockDim.x + threadIdx.x;                                \
    if (t_ >= (long)(s).nx * ((s).nz - 1)) return;                                              \
    const int k = 1 + (int)(t_ / (s).nx);                                                       \
    const int x = (int)(t_ - (long)(k - 1) * (s).nx);

template <typename W>
struct FlowIn {
    const double *u, *v, *w, *cs, *ust, *vst, *ub, *vb, *dudt, *dvdt, *dwdt;
    const W *p, *dudz, *dvdz, *txx, *txy, *txz, *tyy, *tyz, *tzz, *s11, *s12, *s13, *s22, *s23, *s33;
};

template <typename W>
__device__ __forceinline__ double p_org(const FlowIn<W>& a, long o, int k) {  // :63-70
    const double u = a.u[o], v = a.v[o], w = a.w[o];
    return (double)a.p[o] - 0.5 * (u * u + v * v + w * w) - (u * a.ust[k] + v * a.vst[k]) -
           0.5 * (a.ust[k] * a.ust[k] + a.vst[k] * a.vst[k]);
}

template <typename W>
__global__ __launch_bounds__(256) void k_flow_slices(FlowIn<W> a, double* __restrict__ run, SP s) {
    COL_LOOP(x, k, s)
    constexpr int NS = LES_N_FLOW_SLICES;
    double t[NS];
#pragma unroll
    for (int m = 0; m < NS; ++m) t[m] = 0.0;
    const double ubk = a.ub[k], vbk = a.vb[k], ubm = a.ub[k - 1], vbm = a.vb[k - 1];
    for (int j = 0; j < s.ny; ++j) {
        const long o = (long)k * s.plane + (long)j * s.nx + x;
        const long om = o - s.plane;
        const double u = a.u[o], v = a.v[o], w = a.w[o];
        const double po = p_org(a, o, k);
        const double pom = p_org(a, om, k - 1);
        const double arg1 = (k == 1) ? 0.0 : (u + a.u[om]) / 2.0;          // :77-86
        const double arg2 = (k == 1) ? 0.0 : (v + a.v[om]) / 2.0;
        const double arg3 = (po + pom) / 2.0;                                // :87
        const double arg7 = (k == 1) ? 0.0 : (u - ubk + a.u[om] - ubm) / 2.0;  // :140-157
        const double arg8 = (k == 1) ? 0.0 : (v - vbk + a.v[om] - vbm) / 2.0;
        const double txz = a.txz[o], tyz = a.tyz[o], tzz = a.tzz[o];
        const double s11 = a.s11[o], s12 = a.s12[o], s13 = a.s13[o];
        const double s22 = a.s22[o], s23 = a.s23[o], s33 = a.s33[o];
        const double mags = sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33 + 2.0 * (s12 * s12 + s13 * s13 + s23 * s23)));
        const double cs = a.cs[o];
        t[LES_S_U] += u;
        t[LES_S_V] += v;
        t[LES_S_W] += w;
        t[LES_S_P] += po;
        t[LES_S_DUDT] += a.dudt[o];
        t[LES_S_DVDT] += a.dvdt[o];
        t[LES_S_DWDT] += a.dwdt[o];
        t[LES_S_DUDZ] += (double)a.dudz[o];
        t[LES_S_DVDZ] += (double)a.dvdz[o];
        t[LES_S_U2] += u * u;
        t[LES_S_V2] += v * v;
        t[LES_S_W2] += w * w;
        t[LES_S_P2] += po * po;
        t[LES_S_UV] += u * v;
        t[LES_S_UW] += arg1 * w;
        t[LES_S_VW] += arg2 * w;
        t[LES_S_WP] += arg3 * w;
        t[LES_S_TXX] += (double)a.txx[o];
        t[LES_S_TXY] += (double)a.txy[o];
        t[LES_S_TXZ] += txz;
        t[LES_S_TYY] += (double)a.tyy[o];
        t[LES_S_TYZ] += tyz;
        t[LES_S_TZZ] += tzz;
        t[LES_S_U3] += u * u * u;
        t[LES_S_V3] += v * v * v;
        t[LES_S_W3] += w * w * w;
        t[LES_S_U2_FLUC] += (u - ubk) * (u - ubk);
        t[LES_S_V2_FLUC] += (v - vbk) * (v - vbk);
        t[LES_S_UV_FLUC] += (u - ubk) * (v - vbk);
        t[LES_S_U2W_FLUC] += arg7 * arg7 * w;
        t[LES_S_V2W_FLUC] += arg8 * arg8 * w;
        t[LES_S_DISSIP] += -(mags * mags * mags) * cs * s.l2;             // 106_sgs_stag.f90:212
        t[LES_S_UTXZ_FLUC] += arg7 * txz;
        t[LES_S_VTYZ_FLUC] += arg8 * tyz;
        t[LES_S_WTZZ] += w * tzz;
        t[LES_S_CS] += sqrt(cs);
    }
    const long L = (long)s.nx * (s.nz - 1);
    const long e = (long)(k - 1) * s.nx + x;
#pragma unroll
    for (int m = 0; m < NS; ++m) run[m * L + e] += s.fr * (t[m] / s.ny);
}

template <typename W>
struct ThetaIn {
    const double *th, *w, *t3, *tb;
    const W *dtdz, *nut;
};

template <typename W>
__global__ __launch_bounds__(256) void k_theta_slices(ThetaIn<W> a, double* __restrict__ run, SP s) {
    COL_LOOP(x, k, s)
    double t[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double tbk = a.tb[k];
    for (int j = 0; j < s.ny; ++j) {
        const long o = (long)k * s.plane + (long)j * s.nx + x;
        const double th = a.th[o];
        t[0] += th;
        t[1] += th * th;
        t[2] += (th - tbk) * (th - tbk);
        t[3] += (double)a.dtdz[o];
        t[4] += a.t3[o];
        t[5] += a.w[o] * ((th + a.th[o - s.plane]) / 2.0);                 // :270
        t[6] += (double)a.nut[o];
    }
    const long L = (long)s.nx * (s.nz - 1);
    const long e = (long)(k - 1) * s.nx + x;
    double* r = run + (long)LES_N_FLOW_SLICES * L;
#pragma unroll
    for (int m = 0; m < 7; ++m) r[m * L + e] += s.fr * (t[m] / s.ny);
}

// energy (103_flosol.f90:405-419): ke = (u^2 + v^2 + w_uvp^2)/2 at 1..nz-1,
// 0 at the wall level and at the top ghost (not set by the reference)
__global__ void k_ke(double* __restrict__ ke, const double* __restrict__ u, const double* __restrict__ v,
                     const double* __restrict__ w, long plane, int nz) {
    const long n = plane * (nz + 1);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int k = (int)(i / plane);
        if (k == 0 || k == nz) {
            ke[i] = 0.0;
            continue;
        }
        const double tw = .5 * (w[i] + w[i + plane]);
        ke[i] = .5 * (u[i] * u[i] + v[i] * v[i] + tw * tw);
    }
}

// d += a at levels 1..nz-1 (lev1: 0 adds nothing at level 1, for dkedz(1))
__global__ void k_add_levels(double* __restrict__ d, const double* __restrict__ a, long plane, int nz, int skip1) {
    const long n = plane * (nz - 1);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const long i = plane + t;
        if (skip1 && t < plane) continue;
        d[i] += a[i];
    }
}

dim3 grid_n(long n) {
    long b = (n + 255) / 256;
    if (b > 65535L * 16) b = 65535L * 16;
    return dim3((unsigned)b);
}

}  // namespace

template <typename W>
void flow_slices_dev(les_ctx* c) {
    SP s = sp(c);
    const double nn = (double)(c->nx * c->ny);
    level_means_seq(c->stream, c->f3[LES_F_U], c->plane, 0, c->nz, nn, c->st_mean);
    level_means_seq(c->stream, c->f3[LES_F_V], c->plane, 0, c->nz, nn, c->st_mean + c->nlev);
    FlowIn<W> a;
    a.u = c->f3[LES_F_U]; a.v = c->f3[LES_F_V]; a.w = c->f3[LES_F_W]; a.cs = c->f3[LES_F_CS_OPT2];
    a.ust = c->ust; a.vst = c->vst; a.ub = c->st_mean; a.vb = c->st_mean + c->nlev;
    a.dudt = c->st_diag[0]; a.dvdt = c->st_diag[1]; a.dwdt = c->st_diag[2];
    a.p = WF<W>(c, LES_F_P); a.dudz = WF<W>(c, LES_F_DUDZ); a.dvdz = WF<W>(c, LES_F_DVDZ);
    a.txx = WF<W>(c, LES_F_TXX); a.txy = WF<W>(c, LES_F_TXY); a.txz = WF<W>(c, LES_F_TXZ);
    a.tyy = WF<W>(c, LES_F_TYY); a.tyz = WF<W>(c, LES_F_TYZ); a.tzz = WF<W>(c, LES_F_TZZ);
    a.s11 = WF<W>(c, LES_F_S11); a.s12 = WF<W>(c, LES_F_S12); a.s13 = WF<W>(c, LES_F_S13);
    a.s22 = WF<W>(c, LES_F_S22); a.s23 = WF<W>(c, LES_F_S23); a.s33 = WF<W>(c, LES_F_S33);
    const long n = (long)c->nx * (c->nz - 1);
    hipLaunchKernelGGL(k_flow_slices<W>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a, c->st_run, s);
}

template <typename W>
void theta_slices_dev(les_ctx* c) {
    SP s = sp(c);
    const double nn = (double)(c->nx * c->ny);
    level_means_seq(c->stream, c->f3[LES_F_THETA], c->plane, 0, c->nz, nn, c->st_mean + 2 * c->nlev);
    ThetaIn<W> a;
    a.th = c->f3[LES_F_THETA]; a.w = c->f3[LES_F_W]; a.t3 = c->f3[LES_F_SGS_T3]; a.tb = c->st_mean + 2 * c->nlev;
    a.dtdz = WF<W>(c, LES_F_DTDZ); a.nut = WF<W>(c, LES_F_NU_T);
    const long n = (long)c->nx * (c->nz - 1);
    hipLaunchKernelGGL(k_theta_slices<W>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a, c->st_run,
                       s);
}

void ke_dev(les_ctx* c, double* ke) {
    hipLaunchKernelGGL(k_ke, grid_n((long)c->plane * c->nlev), dim3(256), 0, c->stream, ke,
                       (const double*)c->f3[LES_F_U], (const double*)c->f3[LES_F_V], (const double*)c->f3[LES_F_W],
                       (long)c->plane, c->nz);
}

void add_levels_dev(les_ctx* c, double* d, const double* a, int skip_level1) {
    hipLaunchKernelGGL(k_add_levels, grid_n((long)c->plane * (c->nz - 1)), dim3(256), 0, c->stream, d, a,
                       (long)c->plane, c->nz, skip_level1);
}

template void flow_slices_dev<double>(les_ctx*);
template void flow_slices_dev<float>(les_ctx*);
template void theta_slices_dev<double>(les_ctx*);
template void theta_slices_dev<float>(les_ctx*);

}  // namespace les
