// Pointwise, stencil and reduction kernels of the timestep.  Each kernel
// restates one loop nest of the reference in global-domain form (levels
// 0..nzt; the bottom z-slab's rules at levels 0/1, the top z-slab's at
// nzt-1/nzt).  Streaming kernels are HBM-bound: one read of each input and
// one write of each output per point, grid-stride over (level, plane).
#include <utility>
#include "fft_rg.hpp"
#include "les_lag.hpp"

namespace les {

// Kernel-side view of the run constants.
struct KP {
    long plane, plane2;
    int nx, ny, nz;
    double dx, dy, dz, dt;
    double delta;       // filter_size*(dx dy dz)^(1/3), 106_sgs_stag.f90:33
    double nu;          // molecular viscosity (nondim) if molec
};

static KP kp(const les_ctx* c) {
    KP k;
    k.plane = (long)c->plane;
    k.plane2 = (long)c->plane2;
    k.nx = c->nx; k.ny = c->ny; k.nz = c->nz;
    k.dx = c->dx; k.dy = c->dy; k.dz = c->dz; k.dt = c->prm.dt;
    k.delta = 1.0 * pow(c->dx * c->dy * c->dz, 1.0 / 3.0);
    k.nu = c->prm.molec ? (c->prm.nu_molec / (c->prm.u_scale * c->prm.z_i)) : 0.0;
    return k;
}

static inline dim3 grid_for(long n) {
    long b = (n + 255) / 256;
    if (b > 65535L * 16) b = 65535L * 16;
    return dim3((unsigned)b);
}

#define GRID_LOOP(i, n) for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long)gridDim.x * blockDim.x)

__device__ __forceinline__ double smag(double s11, double s12, double s13, double s22, double s23, double s33) {
    return sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33 + 2.0 * (s12 * s12 + s13 * s13 + s23 * s23)));
}

// ---- fills and reductions ------------------------------------------------
template <typename T>
__global__ void k_fill(T* f, long n, double v) {
    GRID_LOOP(i, n) f[i] = (T)v;
}

template <typename T>
void fill_levels(hipStream_t s, T* f, size_t plane, int lev0, int nlev, double v) {
    const long n = (long)plane * nlev;
    hipLaunchKernelGGL(k_fill<T>, grid_for(n), dim3(256), 0, s, f + (size_t)lev0 * plane, n, v);
}
template void fill_levels<double>(hipStream_t, double*, size_t, int, int, double);
template void fill_levels<float>(hipStream_t, float*, size_t, int, int, double);

// up to FILL_MAX runs of n values in one launch (run blockIdx.y)
template <typename T>
struct FillRuns {
    T* p[FILL_MAX];
};
template <typename T>
__global__ void k_fill_runs(FillRuns<T> r, long n, double v) {
    T* f = r.p[blockIdx.y];
    GRID_LOOP(i, n) f[i] = (T)v;
}

template <typename T>
void fill_runs(hipStream_t s, long n, double v, std::initializer_list<T*> ps) {
    FillRuns<T> r{};
    int m = 0;
    dim3 g = grid_for(n);
    for (T* q : ps) {
        r.p[m++] = q;
        if (m == FILL_MAX) {  // launch each full set of FILL_MAX runs
            g.y = (unsigned)m;
            hipLaunchKernelGGL(k_fill_runs<T>, g, dim3(256), 0, s, r, n, v);
            m = 0;
        }
    }
    if (m > 0) {
        g.y = (unsigned)m;
        hipLaunchKernelGGL(k_fill_runs<T>, g, dim3(256), 0, s, r, n, v);
    }
}
template void fill_runs<double>(hipStream_t, long, double, std::initializer_list<double*>);
template void fill_runs<float>(hipStream_t, long, double, std::initializer_list<float*>);

// element-wise conversion between storage precisions
template <typename TO, typename TF>
__global__ void k_convert(TO* __restrict__ out, const TF* __restrict__ in, long n) {
    GRID_LOOP(i, n) out[i] = (TO)in[i];
}

template <typename TO, typename TF>
void convert(hipStream_t s, TO* out, const TF* in, long n) {
    hipLaunchKernelGGL((k_convert<TO, TF>), grid_for(n), dim3(256), 0, s, out, in, n);
}
template void convert<float, double>(hipStream_t, float*, const double*, long);
template void convert<double, float>(hipStream_t, double*, const float*, long);

// one block per level: out[l] = sum(plane l) / nn, in a fixed order (four
// running sums per thread over its strided values, combined in order, then
// a tree over the block), so every run gives the same bits.  Four
// independent sums keep four loads of a thread in flight.
constexpr int LM_THREADS = 1024;
__global__ __launch_bounds__(LM_THREADS) void k_level_means(const double* __restrict__ f, long plane, double nn,
                                                             double* __restrict__ out) {
    __shared__ double sh[LM_THREADS];
    const double* p = f + (long)blockIdx.x * plane;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    long i = threadIdx.x;
    for (; i + 3 * LM_THREADS < plane; i += 4 * LM_THREADS) {
        s0 += p[i];
        s1 += p[i + LM_THREADS];
        s2 += p[i + 2 * LM_THREADS];
        s3 += p[i + 3 * LM_THREADS];
    }
    for (; i < plane; i += LM_THREADS) s0 += p[i];
    sh[threadIdx.x] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    for (int w = LM_THREADS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0] / nn;
}

void level_means(hipStream_t s, const double* f, size_t plane, int lev0, int nlev, double nn, double* out) {
    hipLaunchKernelGGL(k_level_means, dim3(nlev), dim3(LM_THREADS), 0, s, f + (size_t)lev0 * plane, (long)plane, nn,
                       out);
}

// ---- sums in the reference's order ---------------------------------------
// calc_hor_avg and the obukhov plane averages (202_parallel.f90:993-1059,
// 010_scalars.f90:758-775) are Fortran sums over a plane, x fastest: one
// running sum, each element added in index order.  A sequential sum rounds
// differently from a tree sum (the sum of a uniform plane, as psi_m, is not
// n times the value), and the difference is amplified where the mean feeds
// a well-resolved quantity (ustar, psi_m, the wall stress) or a cancellation
// (theta - mean(theta) in calc_beta and the fluctuation statistics), so these
// sums keep the reference's order.  One wavefront per sum walks the values in
// index order, one FP64 add per value (the add's latency bounds the walk at
// about 4 ns per value on this GPU).  The values reach it through LDS: the
// lanes load stages of SEQ_STAGE values (coalesced) two stages ahead of the
// walk, into a ring of three LDS stages, so a load has two stages' walking
// time to arrive even when other kernels load the memory system.
constexpr int SEQ_STAGE = 1024;  // values per stage, 16 per lane

template <typename F>
__device__ __forceinline__ double wave_chain(long n, const F& val, double* lds) {
    const int lane = threadIdx.x & 63;
    constexpr int PER = SEQ_STAGE / 64;
    double r0[PER], r1[PER];
    auto gload = [&](double* r, long b) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const long idx = b + j * 64 + lane;
            r[j] = val(idx < n ? idx : n - 1);   // in bounds; values past n are never added
        }
    };
    auto put = [&](const double* r, int slot) {
        double* L = lds + slot * SEQ_STAGE;
#pragma unroll
        for (int j = 0; j < PER; ++j) L[j * 64 + lane] = r[j];
    };
    const long nst = (n + SEQ_STAGE - 1) / SEQ_STAGE;
    double s = 0.0;
    // the walk of stage st from LDS slot st % 3
    auto walk = [&](long st) {
        const double* L = lds + (int)(st % 3) * SEQ_STAGE;
        const long base = st * SEQ_STAGE;
        const int lim = (n - base < SEQ_STAGE) ? (int)(n - base) : SEQ_STAGE;
        const double2* L2 = reinterpret_cast<const double2*>(L);
        int j = 0;
        for (; j + 64 <= lim; j += 64) {
            double2 t[32];
#pragma unroll
            for (int q = 0; q < 32; ++q) t[q] = L2[(j >> 1) + q];
            // all 32 broadcast reads are issued before the first add consumes one
            asm volatile("" ::: "memory");
#pragma unroll
            for (int q = 0; q < 32; ++q) {
                s += t[q].x;
                s += t[q].y;
            }
        }
        for (; j < lim; ++j) s += L[j];
    };
    gload(r0, 0);
    if (nst > 1) gload(r1, SEQ_STAGE);
    put(r0, 0);
    if (nst > 2) gload(r0, 2 * SEQ_STAGE);
    __syncthreads();
    // stage st + 1 goes to LDS while stage st is walked; the register set
    // just written takes stage st + 3
    for (long st = 0; st < nst; st += 2) {
        if (st + 1 < nst) put(r1, (int)((st + 1) % 3));
        if (st + 3 < nst) gload(r1, (st + 3) * SEQ_STAGE);
        __syncthreads();
        walk(st);
        if (st + 1 >= nst) break;
        if (st + 2 < nst) put(r0, (int)((st + 2) % 3));
        if (st + 4 < nst) gload(r0, (st + 4) * SEQ_STAGE);
        __syncthreads();
        walk(st + 1);
    }
    return s;
}

// The walks run beside other kernels (on the auxiliary stream): at the
// highest wave priority their dependent adds issue ahead of the other waves on
// the SIMD instead of waiting their turn.

// plane means, one wavefront (block) per plane
__global__ __launch_bounds__(64) void k_level_means_seq(const double* __restrict__ f, long plane, double nn,
                                                         double* __restrict__ out) {
    __builtin_amdgcn_s_setprio(3);
    __shared__ __align__(16) double lds[3 * SEQ_STAGE];
    const double* p = f + (long)blockIdx.x * plane;
    const double s = wave_chain(plane, [&](long i) { return p[i]; }, lds);
    if (threadIdx.x == 0) out[blockIdx.x] = s / nn;
}

void level_means_seq(hipStream_t s, const double* f, size_t plane, int lev0, int nlev, double nn, double* out) {
    FamScope fs(s, FAM_PLANE_SUMS, 8.0 * (double)plane * nlev);
    hipLaunchKernelGGL(k_level_means_seq, dim3(nlev), dim3(64), 0, s, f + (size_t)lev0 * plane, (long)plane, nn,
                       out);
}

// obukhov sums at the wall level, in the reference's order, one wavefront
// per sum: [0] sum sqrt(u^2+v^2), [1] sum psi_m, [2] sum sgs_t3(1), [3] sum theta(1)
__global__ __launch_bounds__(64) void k_obukhov_sums(const double* __restrict__ u1, const double* __restrict__ v1,
                                                      const double* __restrict__ psi, const double* __restrict__ t3,
                                                      const double* __restrict__ th1, long plane,
                                                      double* __restrict__ out) {
    __builtin_amdgcn_s_setprio(3);
    __shared__ __align__(16) double lds[3 * SEQ_STAGE];
    const int which = blockIdx.x;
    double s;
    if (which == 0) s = wave_chain(plane, [&](long i) { return sqrt(u1[i] * u1[i] + v1[i] * v1[i]); }, lds);
    else if (which == 1) s = wave_chain(plane, [&](long i) { return psi[i]; }, lds);
    else if (which == 2) s = wave_chain(plane, [&](long i) { return t3[i]; }, lds);
    else s = wave_chain(plane, [&](long i) { return th1[i]; }, lds);
    if (threadIdx.x == 0) out[which] = s;
}

// neutral / theta-off branch, 010_scalars.f90:718-738
__global__ void k_obukhov_neutral(double* ustar, double* phi_m, double* psi_h, double* phi_h,
                                  const double* __restrict__ sums, long plane, double nn, int smooth, double zo,
                                  double dz, double nu_nd) {
    const double su = sums[0], sp = sums[1];
    GRID_LOOP(i, plane) {
        phi_m[i] = 1.0; psi_h[i] = 0.0; phi_h[i] = 1.0;
        if (smooth) ustar[i] = su / nn / (log(0.5 * dz / nu_nd) / VONK + B_SMOOTH);
        else ustar[i] = su / nn * VONK / (log(0.5 * dz / zo) - sp / nn);
    }
}

// theta-active branch with homogeneous surface, 010_scalars.f90:758-843
__global__ void k_obukhov_active(double* ustar, double* psi_m, double* phi_m, double* psi_h, double* phi_h,
                                 const double* __restrict__ sums, long plane, double nn, double zo_, double dz,
                                 double g_) {
    const double wt_ = sums[2] / nn;
    const double U_tmp = sums[0] / nn;
    const double psi_m_tmp = sums[1] / nn;
    const double theta_ = sums[3] / nn;
    const double ustar_ = U_tmp * VONK / (log(0.5 * dz / zo_) - psi_m_tmp);
    const double L = -(ustar_ * ustar_ * ustar_) / (VONK * g_ / theta_ * wt_);
    double pm, ph, fm, fh;
    if (L < 0.0 && wt_ != 0.0) {
        const double x = pow(1.0 - 16.0 * dz / 2.0 / L, 0.25);
        pm = 2.0 * log((1.0 + x) / 2.0) + log((1.0 + x * x) / 2.0) - 2.0 * atan(x) + PI / 2.0;
        ph = 2.0 * log((1.0 + x * x) / 2.0);
        fm = 1.0 / x;
        fh = 1.0 / (x * x);
    } else if (L > 0.0 && wt_ != 0.0) {
        const double zeta = 0.5 * dz / L;
        fm = 1.0 + 4.8 * zeta;
        fh = 1.0 + 7.8 * zeta;
        pm = -1.0 * 4.8 * zeta;
        ph = -1.0 * 7.8 * zeta;
    } else {
        pm = 0.0; ph = 0.0; fm = 1.0; fh = 1.0;
    }
    GRID_LOOP(i, plane) {
        ustar[i] = ustar_;
        psi_m[i] = pm; psi_h[i] = ph; phi_m[i] = fm; phi_h[i] = fh;
    }
}

// ocean branch under a wall top (010_scalars.f90:696-716): the top wall's
// plane-mean ustar_avg from level nz-1 (sum in sums[0]); nf = float(nxt*nynpy)
__global__ void k_obukhov_ocean(double* ustar_top, const double* __restrict__ sums, long plane, double nf,
                                int smooth, double zo, double dz, double nu_nd) {
    const double su = sums[0];
    const double us = smooth ? su / nf / (log(0.5 * dz / nu_nd) / VONK + B_SMOOTH)
                             : su / nf * VONK / log(0.5 * dz / zo);
    GRID_LOOP(i, plane) ustar_top[i] = us;
}

// obukhov's plane sums (and the fills that precede them) on stream s
void obukhov_sums_dev(les_ctx* c, int jt, hipStream_t s) {
    FamScope fs(s, FAM_PLANE_SUMS, 8.0 * 5.0 * (double)c->plane);
    const les_params& P = c->prm;
    const long np = (long)c->plane;
    const double* u1 = c->f3[LES_F_U] + np;
    const double* v1 = c->f3[LES_F_V] + np;
    if (P.ocean_flag) {
        // the surface ranks' planes are left as they are; the top wall's
        // ranks take their ustar_avg from level nz-1
        if (P.ubc_mom == LES_BC_WALL)
            hipLaunchKernelGGL(k_obukhov_sums, dim3(1), dim3(64), 0, s, u1 + (long)(c->nz - 2) * np,
                               v1 + (long)(c->nz - 2) * np, (const double*)nullptr, (const double*)nullptr,
                               (const double*)nullptr, np, c->red);
        return;
    }
    if (!P.theta_flag || jt < P.theta_init_time) {
        if (!P.restart_theta) fill_levels(s, c->psi_m, c->plane, 0, 1, 0.0);
        hipLaunchKernelGGL(k_obukhov_sums, dim3(2), dim3(64), 0, s, u1, v1, c->psi_m, (const double*)nullptr,
                           (const double*)nullptr, np, c->red);
        return;
    }
    double* t3 = c->f3[LES_F_SGS_T3] + np;
    if (jt == P.theta_init_time && !P.restart_theta)
        fill_levels(s, t3, c->plane, 0, 1, P.wt_s / P.u_scale / P.T_scale);
    // theta(1) as the previous step left it: a copy taken on the context's
    // stream, since the step filters theta while these sums may still run
    hipLaunchKernelGGL(k_obukhov_sums, dim3(4), dim3(64), 0, s, u1, v1, c->psi_m, (const double*)t3,
                       (const double*)c->th1_copy, np, c->red);
}

// the rest of obukhov from the sums, on the context's stream
void obukhov_finish_dev(les_ctx* c, int jt) {
    const les_params& P = c->prm;
    const long np = (long)c->plane;
    const double nn = (double)(c->nx * c->ny);
    hipStream_t s = c->stream;
    if (P.ocean_flag) {
        if (P.ubc_mom == LES_BC_WALL)
            hipLaunchKernelGGL(k_obukhov_ocean, grid_for(np), dim3(256), 0, s, c->ustar_top, (const double*)c->red,
                               np, (double)(float)(c->nx * c->ny), (int)(P.walltype == LES_WALL_SMOOTH), P.zo, c->dz,
                               P.nu_molec / P.u_scale / P.z_i);
        return;
    }
    if (!P.theta_flag || jt < P.theta_init_time) {
        hipLaunchKernelGGL(k_obukhov_neutral, grid_for(np), dim3(256), 0, s, c->ustar, c->phi_m, c->psi_h,
                           c->phi_h, (const double*)c->red, np, nn, (int)(P.walltype == LES_WALL_SMOOTH), P.zo,
                           c->dz, P.nu_molec / P.u_scale / P.z_i);
        return;
    }
    // zo_ = exp(mean(log(zo))) of a homogeneous zo, summed in order
    double sl = 0.0;
    const double lz = log(P.zo);
    for (long i = 0; i < np; ++i) sl += lz;
    const double zo_ = exp(sl / nn);
    const double g_ = GRAV / (P.u_scale * P.u_scale / P.z_i);
    hipLaunchKernelGGL(k_obukhov_active, grid_for(np), dim3(256), 0, s, c->ustar, c->psi_m, c->phi_m, c->psi_h,
                       c->phi_h, (const double*)c->red, np, nn, zo_, c->dz, g_);
}

void obukhov_theta_copy(les_ctx* c) {
    les_note(c, hipMemcpyAsync(c->th1_copy, c->f3[LES_F_THETA] + c->plane, c->plane * sizeof(double),
                               hipMemcpyDeviceToDevice, c->stream));
}

void obukhov_dev(les_ctx* c, int jt) {
    obukhov_theta_copy(c);
    obukhov_sums_dev(c, jt, c->stream);
    obukhov_finish_dev(c, jt);
}

// ---- wall stress, 104_wallstress.f90:98-260 ------------------------------
template <typename W>
__global__ void k_wall(W* txz, W* tyz, W* dudz, W* dvdz, double* ustar,
                       const double* __restrict__ u, const double* __restrict__ v,
                       const double* __restrict__ u1, const double* __restrict__ v1,
                       const double* __restrict__ psi_m, const double* __restrict__ phi_m, long plane, int top,
                       int smooth, double dz, double d0, double zo, double nu_nd) {
    GRID_LOOP(i, plane) {
        double denom;
        if (smooth) denom = log(0.5 * dz / nu_nd) / VONK + B_SMOOTH;
        else if (top) denom = log((0.5 * dz - d0) / zo) / VONK;
        else denom = (log((0.5 * dz - d0) / zo) - psi_m[i]) / VONK;
        const double u_avg = sqrt(u1[i] * u1[i] + v1[i] * v1[i]);
        const double us = u_avg / denom;
        const double cst = -(us * us) / u_avg;
        if (!top) {
            ustar[i] = us;
            txz[i] = cst * u1[i];
            tyz[i] = cst * v1[i];
            const double du = us / ((0.5 * dz - d0) * VONK) * u[i] / u_avg * phi_m[i];
            const double dv = us / ((0.5 * dz - d0) * VONK) * v[i] / u_avg * phi_m[i];
            dudz[i] = (u[i] == 0.0) ? 0.0 : du;
            dvdz[i] = (v[i] == 0.0) ? 0.0 : dv;
        } else {
            txz[i] = -cst * u1[i];
            tyz[i] = -cst * v1[i];
            const double du = -us / (0.5 * dz * VONK) * u[i] / u_avg;
            const double dv = -us / (0.5 * dz * VONK) * v[i] / u_avg;
            dudz[i] = (u[i] == 0.0) ? 0.0 : du;
            dvdz[i] = (v[i] == 0.0) ? 0.0 : dv;
        }
    }
}

// ocean surface (104_wallstress.f90:118-136): the wind stress, and dudz,
// dvdz of level 1 as the projection of level 2's on the stress direction
template <typename W>
__global__ void k_wall_ocean(W* txz, W* tyz, W* dudz, W* dvdz, long plane, double us, double rad) {
    GRID_LOOP(i, plane) {
        txz[i] = us * us * cos(rad);
        tyz[i] = us * us * sin(rad);
        const double a = (double)dudz[i + plane] * cos(rad) + (double)dvdz[i + plane] * sin(rad);
        dudz[i] = a * cos(rad);
        dvdz[i] = a * sin(rad);
    }
}

template <typename W>
void wallstress_dev(les_ctx* c, int top, const double* u1, const double* v1) {
    const les_params& P = c->prm;
    const long np = (long)c->plane;
    const int lev = top ? c->nz : 1;
    const int ulev = top ? c->nz - 1 : 1;
    W* txz = WF<W>(c, LES_F_TXZ) + lev * np;
    W* tyz = WF<W>(c, LES_F_TYZ) + lev * np;
    W* dudz = WF<W>(c, LES_F_DUDZ) + lev * np;
    W* dvdz = WF<W>(c, LES_F_DVDZ) + lev * np;
    if (!top && P.ocean_flag) {
        hipLaunchKernelGGL(k_wall_ocean<W>, grid_for(np), dim3(256), 0, c->stream, txz, tyz, dudz, dvdz, np,
                           c->wind_ustar, c->wind_rad);
        return;
    }
    const int sf = top ? (P.ubc_mom == LES_BC_STRESS_FREE) : (P.lbc_mom == LES_BC_STRESS_FREE);
    if (sf) {
        fill_runs<W>(c->stream, np, 0.0, {txz, tyz, dudz, dvdz});
        return;
    }
    hipLaunchKernelGGL(k_wall<W>, grid_for(np), dim3(256), 0, c->stream, txz, tyz, dudz, dvdz, c->ustar,
                       (const double*)(c->f3[LES_F_U] + ulev * np), (const double*)(c->f3[LES_F_V] + ulev * np),
                       u1, v1, (const double*)c->psi_m, (const double*)c->phi_m, np, top,
                       (int)(P.walltype == LES_WALL_SMOOTH), c->dz, P.d0, P.zo, P.nu_molec / P.u_scale / P.z_i);
}

// ---- sgs_stag, 106_sgs_stag.f90 ---------------------------------------------
template <typename W>
struct Grad {
    const W *ux, *uy, *uz, *vx, *vy, *vz, *wx, *wy, *wz;
};
template <typename W>
struct Strain {
    W *s11, *s12, *s13, *s22, *s23, *s33;
};

// S_ij on w-nodes, uvp nodes at the wall (:47-68) and the top (:83-104)
template <typename W>
__global__ void k_strain(Grad<W> g, Strain<W> S, KP p) {
    const long n = p.plane * p.nz;  // levels 1..nz
    GRID_LOOP(t, n) {
        const long i = p.plane + t;
        const int k = (int)(i / p.plane);
        double ux, uy, uz, vx, vy, vz, wx, wy, wz;
        if (k == 1) {
            ux = g.ux[i]; uy = g.uy[i]; uz = g.uz[i]; vx = g.vx[i]; vy = g.vy[i]; vz = g.vz[i];
            wx = 0.5 * (g.wx[i] + g.wx[i + p.plane]);
            wy = 0.5 * (g.wy[i] + g.wy[i + p.plane]);
            wz = g.wz[i];
        } else if (k == p.nz) {
            const long m = i - p.plane;
            ux = g.ux[m]; uy = g.uy[m]; uz = g.uz[i]; vx = g.vx[m]; vy = g.vy[m]; vz = g.vz[i];
            wx = 0.5 * (g.wx[m] + g.wx[i]);
            wy = 0.5 * (g.wy[m] + g.wy[i]);
            wz = g.wz[m];
        } else {
            const long m = i - p.plane;
            ux = 0.5 * (g.ux[i] + g.ux[m]); uy = 0.5 * (g.uy[i] + g.uy[m]); uz = g.uz[i];
            vx = 0.5 * (g.vx[i] + g.vx[m]); vy = 0.5 * (g.vy[i] + g.vy[m]); vz = g.vz[i];
            wx = g.wx[i]; wy = g.wy[i]; wz = 0.5 * (g.wz[i] + g.wz[m]);
        }
        S.s11[i] = ux; S.s12[i] = 0.5 * (uy + vx); S.s13[i] = 0.5 * (uz + wx);
        S.s22[i] = vy; S.s23[i] = 0.5 * (vz + wy); S.s33[i] = wz;
    }
}

template <typename W>
static Grad<W> grads(les_ctx* c) {
    return {WF<W>(c, LES_F_DUDX), WF<W>(c, LES_F_DUDY), WF<W>(c, LES_F_DUDZ), WF<W>(c, LES_F_DVDX),
            WF<W>(c, LES_F_DVDY), WF<W>(c, LES_F_DVDZ), WF<W>(c, LES_F_DWDX), WF<W>(c, LES_F_DWDY),
            WF<W>(c, LES_F_DWDZ)};
}
template <typename W>
static Strain<W> strains(les_ctx* c) {
    return {WF<W>(c, LES_F_S11), WF<W>(c, LES_F_S12), WF<W>(c, LES_F_S13), WF<W>(c, LES_F_S22),
            WF<W>(c, LES_F_S23), WF<W>(c, LES_F_S33)};
}

template <typename W>
void strain_dev(les_ctx* c) {
    FamScope fs(c->stream, FAM_SGS_STRESS, 15.0 * sizeof(W) * (double)c->plane * (double)c->nlev);
    KP p = kp(c);
    hipLaunchKernelGGL(k_strain<W>, grid_for(p.plane * p.nz), dim3(256), 0, c->stream, grads<W>(c), strains<W>(c),
                       p);
}

// u_lag, v_lag, w_lag are kept in the work precision
template <typename W>
__global__ void k_lag_acc(W* ul, W* vl, W* wl, const double* __restrict__ u, const double* __restrict__ v,
                          const double* __restrict__ w, long n) {
    GRID_LOOP(i, n) {
        ul[i] = (W)((double)ul[i] + u[i]);
        vl[i] = (W)((double)vl[i] + v[i]);
        wl[i] = (W)((double)wl[i] + w[i]);
    }
}

template <typename W>
void lag_accumulate_dev(les_ctx* c) {
    FamScope fs(c->stream, FAM_LAGRANGE_POINT, (24.0 + 6.0 * sizeof(W)) * (double)c->plane * (double)c->nlev);
    const long n = (long)(c->plane * c->nlev);
    hipLaunchKernelGGL(k_lag_acc<W>, grid_for(n), dim3(256), 0, c->stream, WF<W>(c, LES_F_U_LAG),
                       WF<W>(c, LES_F_V_LAG), WF<W>(c, LES_F_W_LAG), (const double*)c->f3[LES_F_U],
                       (const double*)c->f3[LES_F_V], (const double*)c->f3[LES_F_W], n);
}

void cs_init_dev(les_ctx* c) { fill_levels(c->stream, c->f3[LES_F_CS_OPT2], c->plane, 1, c->nz, 0.03); }

// Nu_t (:206-216) and tau_ij (:219-315)
template <typename W>
__global__ void k_stress(Strain<W> S, const double* __restrict__ Cs, W* Nu, W* txx, W* txy, W* txz, W* tyy, W* tyz,
                         W* tzz, KP p) {
    // pass 1 is Nu_t; it is computed per point of the levels this point reads
    const long n = p.plane * (p.nz + 1);
    const double l2 = p.delta * p.delta;
    GRID_LOOP(i, n) {
        const int k = (int)(i / p.plane);
        if (k == 0) {
            txx[i] = BOGUS; txy[i] = BOGUS; txz[i] = BOGUS; tyy[i] = BOGUS; tyz[i] = BOGUS; tzz[i] = BOGUS;
            continue;
        }
        const double nu_k = smag(S.s11[i], S.s12[i], S.s13[i], S.s22[i], S.s23[i], S.s33[i]) * Cs[i] * l2;
        Nu[i] = nu_k;
        if (k == p.nz) {
            txx[i] = BOGUS; txy[i] = BOGUS; tyy[i] = BOGUS; tzz[i] = BOGUS;
            continue;  // txz, tyz(nz) from wallstress_top
        }
        if (k == 1) {
            const double f = -2.0 * (nu_k + p.nu);
            txx[i] = f * S.s11[i]; txy[i] = f * S.s12[i]; tyy[i] = f * S.s22[i]; tzz[i] = f * S.s33[i];
            continue;  // txz, tyz(1) from wallstress_bottom
        }
        if (k == p.nz - 1) {
            const long t = i + p.plane;
            const double nu_t = smag(S.s11[t], S.s12[t], S.s13[t], S.s22[t], S.s23[t], S.s33[t]) * Cs[t] * l2;
            const double f = -2.0 * (nu_t + p.nu);
            txx[i] = f * S.s11[t]; txy[i] = f * S.s12[t]; tyy[i] = f * S.s22[t]; tzz[i] = f * S.s33[t];
        } else {
            const long t = i + p.plane;
            const double nu_p = smag(S.s11[t], S.s12[t], S.s13[t], S.s22[t], S.s23[t], S.s33[t]) * Cs[t] * l2;
            const double kk = 0.5 * (nu_k + nu_p);
            const double f = -2.0 * (kk + p.nu);
            txx[i] = f * 0.5 * (S.s11[i] + S.s11[t]);
            txy[i] = f * 0.5 * (S.s12[i] + S.s12[t]);
            tyy[i] = f * 0.5 * (S.s22[i] + S.s22[t]);
            tzz[i] = f * 0.5 * (S.s33[i] + S.s33[t]);
        }
        txz[i] = -2.0 * (nu_k + p.nu) * S.s13[i];
        tyz[i] = -2.0 * (nu_k + p.nu) * S.s23[i];
    }
}

// sgs_stag's strain (:5-158) and Nu_t / tau_ij (:206-315) in one pass, for the
// steps whose strain nothing else reads: the w-node strain of levels k and
// k + 1 is formed from the gradients (each value rounded to W as k_strain
// stores it) and used at once, so the six S_ij fields are neither written
// nor re-read.  Every output is bitwise equal to k_strain + k_stress.
template <typename W>
__device__ __forceinline__ void strain_at(const Grad<W>& g, long i, int k, long P, int nz, double s[6]) {
    double ux, uy, uz, vx, vy, vz, wx, wy, wz;
    if (k == 1) {
        ux = g.ux[i]; uy = g.uy[i]; uz = g.uz[i]; vx = g.vx[i]; vy = g.vy[i]; vz = g.vz[i];
        wx = 0.5 * (g.wx[i] + g.wx[i + P]);
        wy = 0.5 * (g.wy[i] + g.wy[i + P]);
        wz = g.wz[i];
    } else if (k == nz) {
        const long m = i - P;
        ux = g.ux[m]; uy = g.uy[m]; uz = g.uz[i]; vx = g.vx[m]; vy = g.vy[m]; vz = g.vz[i];
        wx = 0.5 * (g.wx[m] + g.wx[i]);
        wy = 0.5 * (g.wy[m] + g.wy[i]);
        wz = g.wz[m];
    } else {
        const long m = i - P;
        ux = 0.5 * (g.ux[i] + g.ux[m]); uy = 0.5 * (g.uy[i] + g.uy[m]); uz = g.uz[i];
        vx = 0.5 * (g.vx[i] + g.vx[m]); vy = 0.5 * (g.vy[i] + g.vy[m]); vz = g.vz[i];
        wx = g.wx[i]; wy = g.wy[i]; wz = 0.5 * (g.wz[i] + g.wz[m]);
    }
    s[0] = (double)(W)ux;
    s[1] = (double)(W)(0.5 * (uy + vx));
    s[2] = (double)(W)(0.5 * (uz + wx));
    s[3] = (double)(W)vy;
    s[4] = (double)(W)(0.5 * (vz + wy));
    s[5] = (double)(W)wz;
}

template <typename W>
__global__ void k_strain_stress(Grad<W> g, const double* __restrict__ Cs, W* Nu, W* txx, W* txy, W* txz, W* tyy,
                                W* tyz, W* tzz, KP p) {
    const long n = p.plane * (p.nz + 1);
    const double l2 = p.delta * p.delta;
    GRID_LOOP(i, n) {
        const int k = (int)(i / p.plane);
        if (k == 0) {
            txx[i] = BOGUS; txy[i] = BOGUS; txz[i] = BOGUS; tyy[i] = BOGUS; tyz[i] = BOGUS; tzz[i] = BOGUS;
            continue;
        }
        double s[6];
        strain_at<W>(g, i, k, p.plane, p.nz, s);
        const double nu_k = smag(s[0], s[1], s[2], s[3], s[4], s[5]) * Cs[i] * l2;
        Nu[i] = nu_k;
        if (k == p.nz) {
            txx[i] = BOGUS; txy[i] = BOGUS; tyy[i] = BOGUS; tzz[i] = BOGUS;
            continue;  // txz, tyz(nz) from wallstress_top
        }
        if (k == 1) {
            const double f = -2.0 * (nu_k + p.nu);
            txx[i] = f * s[0]; txy[i] = f * s[1]; tyy[i] = f * s[3]; tzz[i] = f * s[5];
        } else {
            const long t = i + p.plane;
            double u[6];
            strain_at<W>(g, t, k + 1, p.plane, p.nz, u);
            const double nu_t = smag(u[0], u[1], u[2], u[3], u[4], u[5]) * Cs[t] * l2;
            if (k == p.nz - 1) {
                const double f = -2.0 * (nu_t + p.nu);
                txx[i] = f * u[0]; txy[i] = f * u[1]; tyy[i] = f * u[3]; tzz[i] = f * u[5];
            } else {
                const double kk = 0.5 * (nu_k + nu_t);
                const double f = -2.0 * (kk + p.nu);
                // the sums of two stored strains are taken in W, as k_stress takes them
                txx[i] = f * 0.5 * (double)((W)s[0] + (W)u[0]);
                txy[i] = f * 0.5 * (double)((W)s[1] + (W)u[1]);
                tyy[i] = f * 0.5 * (double)((W)s[3] + (W)u[3]);
                tzz[i] = f * 0.5 * (double)((W)s[5] + (W)u[5]);
            }
            txz[i] = -2.0 * (nu_k + p.nu) * s[2];
            tyz[i] = -2.0 * (nu_k + p.nu) * s[4];
        }
    }
}

template <typename W>
void strain_stress_dev(les_ctx* c) {
    FamScope fs(c->stream, FAM_SGS_STRESS, (16.0 * sizeof(W) + 8.0) * (double)c->plane * (double)c->nlev);
    KP p = kp(c);
    hipLaunchKernelGGL(k_strain_stress<W>, grid_for(p.plane * (p.nz + 1)), dim3(256), 0, c->stream, grads<W>(c),
                       (const double*)c->f3[LES_F_CS_OPT2], WF<W>(c, LES_F_NU_T), WF<W>(c, LES_F_TXX),
                       WF<W>(c, LES_F_TXY), WF<W>(c, LES_F_TXZ), WF<W>(c, LES_F_TYY), WF<W>(c, LES_F_TYZ),
                       WF<W>(c, LES_F_TZZ), p);
}

template <typename W>
void nu_t_stress_dev(les_ctx* c) {
    FamScope fs(c->stream, FAM_SGS_STRESS, (13.0 * sizeof(W) + 8.0) * (double)c->plane * (double)c->nlev);
    KP p = kp(c);
    hipLaunchKernelGGL(k_stress<W>, grid_for(p.plane * (p.nz + 1)), dim3(256), 0, c->stream, strains<W>(c),
                       (const double*)c->f3[LES_F_CS_OPT2], WF<W>(c, LES_F_NU_T), WF<W>(c, LES_F_TXX),
                       WF<W>(c, LES_F_TXY), WF<W>(c, LES_F_TXZ), WF<W>(c, LES_F_TYY), WF<W>(c, LES_F_TYZ),
                       WF<W>(c, LES_F_TZZ), p);
}

// ---- interpolag_Sdep, 106_06_interpolag_Sdep.f90:4-336 -------------------
template <typename W>
__device__ __forceinline__ double ulag_at(const W* __restrict__ ul, const double* __restrict__ st, long i,
                                          int k, long plane, int nz, double csc) {
    if (k == 1) return ul[i] / csc + st[1];                                 // :181-186
    if (k == nz) return ul[i - plane] + st[nz - 1];                         // :195-199 (raw sum)
    const double a = 0.5 * (ul[i] / csc + ul[i - plane] / csc);            // :207-213
    return a + 0.5 * (st[k] + st[k - 1]);
}

template <typename W>
__device__ __forceinline__ double wlag_at(const W* __restrict__ wl, long i, int k, long plane, int nz,
                                          double csc) {
    if (k == 1) return 0.25 * (wl[i + plane] / csc);                        // :221
    if (k == nz) return 0.25 * (wl[i - plane] / csc);                       // :225
    return wl[i] / csc;
}

template <typename W>
__global__ void k_interpolag(const W* __restrict__ ul, const W* __restrict__ vl,
                             const W* __restrict__ wl, const double* __restrict__ ust,
                             const double* __restrict__ vst, const double* __restrict__ o0,
                             const double* __restrict__ o1, const double* __restrict__ o2,
                             const double* __restrict__ o3, double* f0, double* f1, double* f2, double* f3, KP p,
                             double csc) {
    // one level per blockIdx.y: no 64-bit index division
    const int qi = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (qi < (int)p.plane) {
        const int k = 1 + (int)blockIdx.y;
        const long q = qi;
        const long i = (long)k * p.plane + q;
        const int jy = qi / p.nx, jx = qi - jy * p.nx;
        const double xu = ulag_at(ul, ust, i, k, p.plane, p.nz, csc);
        const double yv = ulag_at(vl, vst, i, k, p.plane, p.nz, csc);
        const double xp = -xu * p.dt * csc / p.dx;
        const double yp = -yv * p.dt * csc / p.dy;
        double zp;
        if (k <= 2) {                                                        // :239-257
            const long i2 = q + 2 * p.plane;
            double z2 = -wlag_at(wl, i2, 2, p.plane, p.nz, csc) * p.dt * csc / p.dz;
            z2 = z2 + fmin(z2, 0.0);
            z2 = fmin(1.0, z2);
            z2 = fmax(-1.0, z2);
            zp = z2;
        } else {
            zp = -wlag_at(wl, i, k, p.plane, p.nz, csc) * p.dt * csc / p.dz;
        }
        const int ax = (copysign(1.0, xp) > 0.0) ? 1 : -1;
        const int ay = (copysign(1.0, yp) > 0.0) ? 1 : -1;
        const int az = (copysign(1.0, zp) > 0.0) ? 1 : -1;
        int x2 = jx + ax, y2 = jy + ay;  // periodic neighbours
        x2 = (x2 < 0) ? x2 + p.nx : ((x2 >= p.nx) ? x2 - p.nx : x2);
        y2 = (y2 < 0) ? y2 + p.ny : ((y2 >= p.ny) ? y2 - p.ny : y2);
        int k2 = k + az;
        k2 = k2 < 1 ? 1 : (k2 > p.nz ? p.nz : k2);
        const double cx = fabs(xp), cy = fabs(yp), cz = fabs(zp);
        const double fx = 1.0 - cx, fy = 1.0 - cy, fz = 1.0 - cz;
        const long b00 = (long)k * p.plane, b10 = (long)k2 * p.plane;
        const long r0 = (long)jy * p.nx, r1 = (long)y2 * p.nx;
        const double* src[4] = {o0, o1, o2, o3};
        double* dst[4] = {f0, f1, f2, f3};
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const double* A = src[m];
            dst[m][i] = fx * fy * (A[b00 + r0 + jx] * fz + A[b10 + r0 + jx] * cz) +
                        fx * cy * (A[b00 + r1 + jx] * fz + A[b10 + r1 + jx] * cz) +
                        cx * fy * (A[b00 + r0 + x2] * fz + A[b10 + r0 + x2] * cz) +
                        cx * cy * (A[b00 + r1 + x2] * fz + A[b10 + r1 + x2] * cz);
        }
    }
}

template <typename W>
void interpolag_dev(les_ctx* c) {
    // F_* old read and new written (FP64), u_lag, v_lag, w_lag read and cleared
    FamScope fs(c->stream, FAM_LAGRANGE_POINT, (64.0 + 6.0 * sizeof(W)) * (double)c->plane * (double)c->nlev);
    KP p = kp(c);
    const int ids[4] = {LES_F_F_LM, LES_F_F_MM, LES_F_F_QN, LES_F_F_NN};
    // the old F_* become the interpolation source by a buffer exchange; the
    // kernel writes levels 1..nz of the new ones, level 0 is carried over
    for (int m = 0; m < 4; ++m) {
        std::swap(c->f3[ids[m]], c->fold[m]);
        les_note(c, hipMemcpyAsync(c->f3[ids[m]], c->fold[m], sizeof(double) * c->plane, hipMemcpyDeviceToDevice,
                                   c->stream));
    }
    hipLaunchKernelGGL(k_interpolag<W>, dim3((unsigned)((p.plane + 255) / 256), (unsigned)p.nz), dim3(256), 0,
                       c->stream, (const W*)WF<W>(c, LES_F_U_LAG), (const W*)WF<W>(c, LES_F_V_LAG),
                       (const W*)WF<W>(c, LES_F_W_LAG), (const double*)c->ust, (const double*)c->vst,
                       (const double*)c->fold[0], (const double*)c->fold[1], (const double*)c->fold[2],
                       (const double*)c->fold[3], c->f3[ids[0]], c->f3[ids[1]], c->f3[ids[2]], c->f3[ids[3]], p,
                       (double)c->prm.cs_count);
    const long n = (long)(c->plane * c->nlev);                              // :330-332
    fill_runs<W>(c->stream, n, 0.0, {WF<W>(c, LES_F_U_LAG), WF<W>(c, LES_F_V_LAG), WF<W>(c, LES_F_W_LAG)});
}

// ---- lagrange_Sdep, 106_03_lagrange_Sdep.f90:70-331 -------------------------
// buf21 layout: field f (0..20) of level jz (1..nz) at ((f * nz) + jz-1) * plane.
// f: 0 u_bar, 1 v_bar, 2 w_bar, 3..8 products p11 p12 p13 p22 p23 p33,
//    9..14 S_ij, 15..20 |S| S_ij  (ij order 11 12 13 22 23 33)
// With um/vm (plane means of u, v; FP32 work precision) the w-node velocities
// are taken relative to their plane mean.  L_ij = bar(u_i u_j) - bar(u_i)
// bar(u_j) and Q_ij are unchanged by a per-level constant shift (the test
// filters keep the mean), and the FP32 products then carry the fluctuation
// instead of the mean flow, avoiding the cancellation in L_ij.
template <typename W>
__global__ void k_lag_prep(const double* __restrict__ u, const double* __restrict__ v,
                           const double* __restrict__ w, Strain<W> S, W* __restrict__ buf, KP p,
                           const double* __restrict__ um, const double* __restrict__ vm) {
    const long n = p.plane * p.nz;
    GRID_LOOP(t, n) {
        const long i = p.plane + t;
        const int k = (int)(i / p.plane);
        double ub, vb, wb;
        if (k == 1) { ub = u[i]; vb = v[i]; wb = 0.25 * w[i + p.plane]; }
        else if (k == p.nz) { ub = u[i - p.plane]; vb = v[i - p.plane]; wb = 0.25 * w[i - p.plane]; }
        else { ub = 0.5 * (u[i] + u[i - p.plane]); vb = 0.5 * (v[i] + v[i - p.plane]); wb = w[i]; }
        if (um) {
            if (k == 1) { ub -= um[1]; vb -= vm[1]; }
            else if (k == p.nz) { ub -= um[k - 1]; vb -= vm[k - 1]; }
            else { ub -= 0.5 * (um[k] + um[k - 1]); vb -= 0.5 * (vm[k] + vm[k - 1]); }
        }
        const double s[6] = {S.s11[i], S.s12[i], S.s13[i], S.s22[i], S.s23[i], S.s33[i]};
        const double sm = smag(s[0], s[1], s[2], s[3], s[4], s[5]);
        const long fs = p.plane * p.nz;
        W* o = buf + t;
        o[0 * fs] = ub; o[1 * fs] = vb; o[2 * fs] = wb;
        o[3 * fs] = ub * ub; o[4 * fs] = ub * vb; o[5 * fs] = ub * wb;
        o[6 * fs] = vb * vb; o[7 * fs] = vb * wb; o[8 * fs] = wb * wb;
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            o[(9 + m) * fs] = s[m];
            o[(15 + m) * fs] = sm * s[m];
        }
    }
}

template <typename W>
void lagrange_prep_dev(les_ctx* c, W* buf21) {
    FamScope fs(c->stream, FAM_LAGRANGE_POINT, (24.0 + 27.0 * sizeof(W)) * (double)c->plane * c->nz);
    KP p = kp(c);
    const double *um = nullptr, *vm = nullptr;
    if (sizeof(W) < sizeof(double)) {
        const double nn = (double)(c->nx * c->ny);
        level_means(c->stream, c->f3[LES_F_U], c->plane, 0, (int)c->nlev, nn, c->lmean);
        level_means(c->stream, c->f3[LES_F_V], c->plane, 0, (int)c->nlev, nn, c->lmean + c->nlev);
        um = c->lmean;
        vm = c->lmean + c->nlev;
    }
    hipLaunchKernelGGL(k_lag_prep<W>, grid_for(p.plane * p.nz), dim3(256), 0, c->stream,
                       (const double*)c->f3[LES_F_U], (const double*)c->f3[LES_F_V], (const double*)c->f3[LES_F_W],
                       strains<W>(c), buf21, p, um, vm);
}

// Two waves per SIMD: the FP64 contraction (42 filtered values and three
// pow calls per point) spills to scratch under the default four.
template <typename W>
__global__ __launch_bounds__(256, 2) void k_lag_contract(const W* __restrict__ b2, const W* __restrict__ b4,
                                                         LagUpdate u, KP p) {
    const long n = p.plane * p.nz;
    const long fs = n;
    const double cst = 2.0 * (p.delta * p.delta);
    GRID_LOOP(t, n) {
        const long i = p.plane + t;
        const int k = (int)(i / p.plane);
        const W* a = b2 + t;
        const W* h = b4 + t;
        double LM, MM, QN, NN;
        lag_scale([&](int j) { return (double)a[j * fs]; }, cst, 4.0, LM, MM);
        lag_scale([&](int j) { return (double)h[j * fs]; }, cst, 16.0, QN, NN);
        lag_update(u, i, k, LM, MM, QN, NN);
    }
}

template <typename W>
void lagrange_contract_dev(les_ctx* c, const W* b2, const W* b4, int init_lm, int init_qn) {
    FamScope fs(c->stream, FAM_LAGRANGE_POINT, (42.0 * sizeof(W) + 56.0) * (double)c->plane * c->nz);
    KP p = kp(c);
    const double bexp = log(2.0) / (log(4.0) - log(2.0));
    const LagUpdate u{c->f3[LES_F_F_LM], c->f3[LES_F_F_MM], c->f3[LES_F_F_QN], c->f3[LES_F_F_NN],
                      c->f3[LES_F_CS_OPT2], init_lm, init_qn, p.nz, OPFTIME * p.delta,
                      c->prm.dt * (double)c->prm.cs_count, bexp};
    hipLaunchKernelGGL(k_lag_contract<W>, grid_for(p.plane * p.nz), dim3(256), 0, c->stream, b2, b4, u, p);
}

// ---- vertical differences, 004_derivatives.f90:85-183 --------------------
template <typename W, typename TI>
__global__ void k_ddz_uv(W* d, const TI* __restrict__ f, KP p) {
    const long n = p.plane * (p.nz + 1);
    const double cst = 1.0 / p.dz;
    GRID_LOOP(i, n) {
        const int k = (int)(i / p.plane);
        if (k == 0) d[i] = BOGUS;
        else if (k >= 2 && k <= p.nz - 1) d[i] = cst * ((double)f[i] - (double)f[i - p.plane]);
    }
}

template <typename W, typename TI>
__global__ void k_ddz_w(W* d, const TI* __restrict__ f, KP p) {
    const long n = p.plane * (p.nz + 1);
    const double cst = 1.0 / p.dz;
    GRID_LOOP(i, n) {
        const int k = (int)(i / p.plane);
        if (k == 0) d[i] = BOGUS;
        else if (k == p.nz) d[i] = 0.0;
        else d[i] = cst * ((double)f[i + p.plane] - (double)f[i]);
    }
}

template <typename W, typename TI>
void ddz_uv_dev(les_ctx* c, W* d, const TI* f) {
    FamScope fs(c->stream, FAM_DDZ, (double)(sizeof(TI) + sizeof(W)) * (double)c->plane * (double)c->nlev);
    KP p = kp(c);
    hipLaunchKernelGGL((k_ddz_uv<W, TI>), grid_for(p.plane * (p.nz + 1)), dim3(256), 0, c->stream, d, f, p);
}
template <typename W, typename TI>
void ddz_w_dev(les_ctx* c, W* d, const TI* f) {
    FamScope fs(c->stream, FAM_DDZ, (double)(sizeof(TI) + sizeof(W)) * (double)c->plane * (double)c->nlev);
    KP p = kp(c);
    hipLaunchKernelGGL((k_ddz_w<W, TI>), grid_for(p.plane * (p.nz + 1)), dim3(256), 0, c->stream, d, f, p);
}

// ---- divstress, 107_divstress.f90:36-102 ---------------------------------------
// a = d/dx tx, b = d/dy ty, f = tz (FD taken here: ddz_w for uv, ddz_uv for w)
template <typename W>
__global__ void k_divstress(W* out, const W* __restrict__ a, const W* __restrict__ b, const W* __restrict__ f, KP p,
                            int wver) {
    const long n = p.plane * (p.nz + 1);
    const double cst = 1.0 / p.dz;
    GRID_LOOP(i, n) {
        const int k = (int)(i / p.plane);
        if (!wver) {
            if (k == 0 || k == p.nz) out[i] = BOGUS;
            else out[i] = (double)a[i] + (double)b[i] + cst * ((double)f[i + p.plane] - (double)f[i]);
        } else {
            if (k == 0) out[i] = BOGUS;
            else if (k == 1 || k == p.nz) out[i] = (double)a[i] + (double)b[i];
            else out[i] = (double)a[i] + (double)b[i] + cst * ((double)f[i] - (double)f[i - p.plane]);
        }
    }
}

template <typename W>
void divstress_combine_dev(les_ctx* c, W* out, const W* a, const W* b, const W* d, int wver) {
    FamScope fs(c->stream, FAM_DIVSTRESS, 4.0 * sizeof(W) * (double)c->plane * (double)c->nlev);
    KP p = kp(c);
    hipLaunchKernelGGL(k_divstress<W>, grid_for(p.plane * (p.nz + 1)), dim3(256), 0, c->stream, out, a, b, d, p, wver);
}

// ---- convec, 108_convec.f90:44-69 and cross_product :103-202 -------------------
template <typename W>
__global__ void k_vort(W* v1, W* v2, W* v3, Grad<W> g, KP p) {
    const long n = p.plane * p.nz;
    GRID_LOOP(t, n) {
        const long i = p.plane + t;
        const int k = (int)(i / p.plane);
        if (k == 1) {
            v1[i] = 0.5 * ((double)g.wy[i] + (double)g.wy[i + p.plane]) - (double)g.vz[i];
            v2[i] = (double)g.uz[i] - 0.5 * ((double)g.wx[i] + (double)g.wx[i + p.plane]);
        } else if (k == p.nz) {
            v1[i] = 0.5 * ((double)g.wy[i] + (double)g.wy[i - p.plane]) - (double)g.vz[i];
            v2[i] = (double)g.uz[i] - 0.5 * ((double)g.wx[i] + (double)g.wx[i - p.plane]);
        } else {
            v1[i] = (double)g.wy[i] - (double)g.vz[i];
            v2[i] = (double)g.uz[i] - (double)g.wx[i];
        }
        v3[i] = (double)g.vx[i] - (double)g.uy[i];
    }
}

template <typename W>
void vorticity_dev(les_ctx* c, W* v1, W* v2, W* v3) {
    FamScope fs(c->stream, FAM_VORT_PAD_X, 9.0 * sizeof(W) * (double)c->plane * (double)c->nz);
    KP p = kp(c);
    hipLaunchKernelGGL(k_vort<W>, grid_for(p.plane * p.nz), dim3(256), 0, c->stream, v1, v2, v3, grads<W>(c), p);
}

template <typename W>
__global__ void k_cross(const W* __restrict__ ub, const W* __restrict__ vb, const W* __restrict__ wb,
                        const W* __restrict__ b1, const W* __restrict__ b2, const W* __restrict__ b3,
                        const double* __restrict__ ust, const double* __restrict__ vst, W* cxb, W* cyb, W* czb,
                        KP p) {
    const long P2 = p.plane2;
    const long n = P2 * (p.nz - 1);  // levels 1..nz-1
    GRID_LOOP(t, n) {
        const long i = P2 + t;
        const int k = (int)(i / P2);
        if (k == 1) {
            cxb[i] = (vb[i] + vst[1]) * (-b3[i]) + 0.5 * wb[i + P2] * b2[i + P2];
            cyb[i] = (ub[i] + ust[1]) * (b3[i]) + 0.5 * wb[i + P2] * (-b1[i + P2]);
            czb[i] = 0.0;
        } else {
            cxb[i] = (vb[i] + vst[k]) * (-b3[i]) + 0.5 * (wb[i + P2] * (b2[i + P2]) + wb[i] * (b2[i]));
            cyb[i] = (ub[i] + ust[k]) * (b3[i]) + 0.5 * (wb[i + P2] * (-b1[i + P2]) + wb[i] * (-b1[i]));
            czb[i] = 0.5 * ((ub[i] + ub[i - P2] + ust[k] + ust[k - 1]) * (-b2[i]) +
                            (vb[i] + vb[i - P2] + vst[k] + vst[k - 1]) * (b1[i]));
        }
    }
}

template <typename W>
void cross_product_dev(les_ctx* c, const W* b1, const W* b2, const W* b3, W* cxb, W* cyb, W* czb) {
    KP p = kp(c);
    hipLaunchKernelGGL(k_cross<W>, grid_for(p.plane2 * (p.nz - 1)), dim3(256), 0, c->stream,
                       (const W*)c->u_big, (const W*)c->v_big, (const W*)c->w_big, b1, b2, b3,
                       (const double*)c->ust, (const double*)c->vst, cxb, cyb, czb, p);
}

// ---- scalar transport, 010_scalars.f90:111-671 ----------------------------------
template <typename W>
__global__ void k_theta_prep(W* dTdz, double* rhsf, const double* __restrict__ rhs, KP p, double top) {
    const long n = p.plane * (p.nz + 1);
    GRID_LOOP(i, n) {
        rhsf[i] = rhs[i];
        if (i >= p.plane * p.nz) dTdz[i] = top;
    }
}

template <typename W>
void theta_prep_dev(les_ctx* c) {
    FamScope fs(c->stream, FAM_THETA_POINT, (2.0 * sizeof(W) + 16.0) * (double)c->plane * (double)c->nlev);
    KP p = kp(c);
    const les_params& P = c->prm;
    hipLaunchKernelGGL(k_theta_prep<W>, grid_for(p.plane * (p.nz + 1)), dim3(256), 0, c->stream, WF<W>(c, LES_F_DTDZ),
                       c->f3[LES_F_RHST_F], (const double*)c->f3[LES_F_RHST], p, P.dTdz_top / P.T_scale * P.z_i);
}

// dealiased advection on the 3/2 grid, levels 1..nz-1 (:466-484)
template <typename W>
__global__ void k_scalar_adv(const W* __restrict__ ub, const W* __restrict__ vb, const W* __restrict__ wb,
                             const W* __restrict__ sx, const W* __restrict__ sy, const W* __restrict__ sz,
                             const double* __restrict__ ust, const double* __restrict__ vst, W* rm, KP p) {
    const long P2 = p.plane2;
    const long n = P2 * (p.nz - 1);
    GRID_LOOP(t, n) {
        const long i = P2 + t;
        const int k = (int)(i / P2);
        if (k == 1)
            rm[i] = (ub[i] + ust[1]) * sx[i] + (vb[i] + vst[1]) * sy[i] + (0.5 * wb[i + P2]) * sz[i + P2];
        else
            rm[i] = (ub[i] + ust[k]) * sx[i] + (vb[i] + vst[k]) * sy[i] +
                    ((double)wb[i] * sz[i] + (double)wb[i + P2] * sz[i + P2]) / 2.0;
    }
}

template <typename W>
void scalar_adv_dev(les_ctx* c, const W* dsx, const W* dsy, const W* dsz, W* rm) {
    KP p = kp(c);
    hipLaunchKernelGGL(k_scalar_adv<W>, grid_for(p.plane2 * (p.nz - 1)), dim3(256), 0, c->stream,
                       (const W*)c->u_big, (const W*)c->v_big, (const W*)c->w_big, dsx, dsy, dsz,
                       (const double*)c->ust, (const double*)c->vst, rm, p);
}

// SGS flux components: comp 0 x (:513-541), 1 y (:552-569), 2 z (:578-594)
// Surface flux of scalar_RHS_calc (010_scalars.f90:359-394): lbc = 1 the
// prescribed wt_s (surf_flux); lbc = 0 the bulk Monin-Obukhov flux from the
// surface temperature T_s and the filtered theta of level 1, with ustar_avg
// as the wall model left it and obukhov's psi_h (:380-385)
struct SurfFlux {
    int bulk;
    double surf_flux;                  // lbc = 1
    double T_s, log_zos;               // lbc = 0: log(dz / (2 z_os))
    const double *th1, *ustar, *psi_h;
};
__device__ __forceinline__ double surf_flux_at(const SurfFlux& f, long q) {
    if (!f.bulk) return f.surf_flux;
    return (f.T_s - f.th1[q]) * VONK * f.ustar[q] / (f.log_zos - f.psi_h[q]);
}

// SGS flux components: comp 0 x (:513-541), 1 y (:552-569), 2 z (:578-594)
template <typename W>
__global__ void k_scalar_flux(W* temp, double* sgs_t3, const W* __restrict__ Nu, const W* __restrict__ ds, KP p,
                              int comp, double iPr, SurfFlux sf) {
    const long n = p.plane * (p.nz + 1);
    GRID_LOOP(i, n) {
        const int k = (int)(i / p.plane);
        double t = 0.0;
        if (comp < 2) {
            if (k == 1) t = iPr * Nu[i] * ds[i];
            else if (k == p.nz - 1) t = iPr * Nu[i + p.plane] * ds[i];
            else if (k >= 2 && k <= p.nz - 2) t = iPr * 0.5 * ((double)Nu[i] + (double)Nu[i + p.plane]) * ds[i];
        } else {
            if (k == 1) {
                const double fl = surf_flux_at(sf, i - p.plane);
                t = -1.0 * fl;
                sgs_t3[i] = fl;
            } else if (k >= 2) {
                t = iPr * Nu[i] * ds[i];
                sgs_t3[i] = -1.0 * t;
            }
        }
        temp[i] = t;
    }
}

template <typename W>
void scalar_flux_dev(les_ctx* c, W* temp, int comp) {
    FamScope fs(c->stream, FAM_SCALAR_FLUX, (3.0 * sizeof(W) + (comp == 2 ? 8.0 : 0.0)) * (double)c->plane * (double)c->nlev);
    KP p = kp(c);
    const les_params& P = c->prm;
    const W* ds = WF<W>(c, comp == 0 ? LES_F_DTDX : comp == 1 ? LES_F_DTDY : LES_F_DTDZ);
    SurfFlux sf;
    sf.bulk = (P.lbc == 0);
    sf.surf_flux = P.wt_s / P.T_scale / P.u_scale;
    sf.T_s = P.T_s;
    sf.log_zos = log(c->dz / (2.0 * ((1.0 / 10.0) * P.zo)));   // z_os = zo/10, 008_bottombc.f90:126
    sf.th1 = c->f3[LES_F_THETA] + c->plane;
    sf.ustar = c->ustar;
    sf.psi_h = c->psi_h;
    hipLaunchKernelGGL(k_scalar_flux<W>, grid_for(p.plane * (p.nz + 1)), dim3(256), 0, c->stream, temp,
                       c->f3[LES_F_SGS_T3], (const W*)WF<W>(c, LES_F_NU_T), ds, p, comp, 1.0 / P.Pr, sf);
}

// adv != null: RHS = -adv + d (first flux, adv = the advection term), else
// RHS = RHS + d; levels 1..nz-1
template <typename W>
__global__ void k_scalar_accum(double* rhs, const W* __restrict__ d, const W* __restrict__ adv, KP p) {
    const long n = p.plane * (p.nz - 1);
    GRID_LOOP(t, n) {
        const long i = p.plane + t;
        rhs[i] = adv ? (-1.0 * (double)adv[i] + d[i]) : (rhs[i] + d[i]);
    }
}

// ddz_w of the vertical flux (rounded to W as k_ddz_w stores it) added to the
// RHS at once (k_scalar_accum without adv), levels 1..nz-1
template <typename W>
__global__ void k_ddz_w_accum(double* rhs, const W* __restrict__ f, KP p) {
    const long n = p.plane * (p.nz - 1);
    const double cst = 1.0 / p.dz;
    GRID_LOOP(t, n) {
        const long i = p.plane + t;
        const W d = (W)(cst * ((double)f[i + p.plane] - (double)f[i]));
        rhs[i] = rhs[i] + (double)d;
    }
}

template <typename W>
void ddz_w_accum_dev(les_ctx* c, const W* f) {
    FamScope fs(c->stream, FAM_SCALAR_FLUX, (16.0 + sizeof(W)) * (double)c->plane * (double)c->nlev);
    KP p = kp(c);
    hipLaunchKernelGGL(k_ddz_w_accum<W>, grid_for(p.plane * (p.nz - 1)), dim3(256), 0, c->stream,
                       c->f3[LES_F_RHST], f, p);
}

template <typename W>
void scalar_accum_dev(les_ctx* c, const W* dtemp, const W* adv) {
    FamScope fs(c->stream, FAM_SCALAR_FLUX, (16.0 + (adv ? 2.0 : 1.0) * sizeof(W)) * (double)c->plane * (double)c->nlev);
    KP p = kp(c);
    hipLaunchKernelGGL(k_scalar_accum<W>, grid_for(p.plane * (p.nz - 1)), dim3(256), 0, c->stream,
                       c->f3[LES_F_RHST], dtemp, adv, p);
}

// calc_beta (:214-258); means[k] = plane mean of theta(k); alpha_w > 0
// selects the ocean's thermal-expansion form (:239-248)
template <typename W>
__global__ void k_calc_beta(W* beta, const double* __restrict__ th, const double* __restrict__ bar, KP p,
                            double g_hat, int ocean, double alpha_w) {
    const long n = p.plane * (p.nz + 1);
    GRID_LOOP(i, n) {
        const int k = (int)(i / p.plane);
        double b = 0.0;
        if (k >= 2 && k <= p.nz - 1) {
            double above, below;
            if (ocean) {
                above = alpha_w * (th[i] - bar[k]);
                below = alpha_w * (th[i - p.plane] - bar[k - 1]);
            } else {
                above = (th[i] - bar[k]) / bar[k];
                below = (th[i - p.plane] - bar[k - 1]) / bar[k - 1];
            }
            b = g_hat * (above + below) / 2.0;
        }
        beta[i] = b;
    }
}

template <typename W>
void calc_beta_dev(les_ctx* c, const double* means) {
    FamScope fs(c->stream, FAM_THETA_POINT, (8.0 + sizeof(W)) * (double)c->plane * (double)c->nlev);
    KP p = kp(c);
    const les_params& P = c->prm;
    hipLaunchKernelGGL(k_calc_beta<W>, grid_for(p.plane * (p.nz + 1)), dim3(256), 0, c->stream,
                       WF<W>(c, LES_F_BETA_SCAL), (const double*)c->f3[LES_F_THETA], means, p,
                       GRAV * (P.z_i / (P.u_scale * P.u_scale)), P.ocean_flag, P.alpha_w);
}

// step_scalar (:624-671): AB2 on 1..nz-1, top boundary row
__global__ void k_step_scalar(double* th, const double* __restrict__ rhs, const double* __restrict__ rhsf, KP p,
                              int above1, double topgrad) {
    const long n = p.plane * p.nz;  // levels 1..nz
    GRID_LOOP(t, n) {
        const long i = p.plane + t;
        const int k = (int)(i / p.plane);
        if (k <= p.nz - 1) th[i] = th[i] + p.dt * (1.5 * rhs[i] - 0.5 * rhsf[i]);
    }
}

__global__ void k_scalar_top(double* th, KP p, int above1, double topgrad) {
    GRID_LOOP(q, p.plane) {
        const long i = (long)p.nz * p.plane + q;
        th[i] = above1 ? th[i - p.plane] + topgrad : th[i - p.plane];
    }
}

void step_scalar_dev(les_ctx* c) {
    FamScope fs(c->stream, FAM_THETA_POINT, 32.0 * (double)c->plane * (double)c->nlev);
    KP p = kp(c);
    const les_params& P = c->prm;
    const double topgrad = P.dTdz_top / P.T_scale * P.z_i * c->dz;
    hipLaunchKernelGGL(k_step_scalar, grid_for(p.plane * p.nz), dim3(256), 0, c->stream, c->f3[LES_F_THETA],
                       (const double*)c->f3[LES_F_RHST], (const double*)c->f3[LES_F_RHST_F], p,
                       (int)(P.lz_npz > 1.0), topgrad);
    hipLaunchKernelGGL(k_scalar_top, grid_for(p.plane), dim3(256), 0, c->stream, c->f3[LES_F_THETA], p,
                       (int)(P.lz_npz > 1.0), topgrad);
}

// ---- pressure, 109_press_stag_array.f90:158-199 -------------------------------
// shift != 0 (reduced-precision work fields): u, v, w are taken relative to
// their value at the level's first point before rounding.  This changes only
// the kx = ky = 0 coefficients, which the solve does not read from these
// transforms (Hx, Hy: unused; Hz: replaced by the FP64 plane sums).
// modi: the surface-temperature term of lbc = 0 taken from divtz(1) at the
// bottom (109_press_stag_array.f90:186-196), 0 otherwise
template <typename W>
__global__ void k_press_prep(W* rx, W* ry, W* rz, const double* __restrict__ u, const double* __restrict__ v,
                             const double* __restrict__ w, const W* __restrict__ divtz, KP p, double cst, int shift,
                             double modi) {
    const long n = p.plane * (p.nz + 1);
    const double tadv1 = 1.5;
    GRID_LOOP(i, n) {
        const int k = (int)(i / p.plane);
        if (k == 0) {
            rz[i] = cst * ((double)divtz[i + p.plane] - modi);
            rx[i] = 0.0; ry[i] = 0.0;
        } else if (k == p.nz) {
            rz[i] = cst * divtz[i];
            rx[i] = 0.0; ry[i] = 0.0;
        } else {
            const long i0 = (long)k * p.plane;
            const double su = shift ? u[i0] : 0.0, sv = shift ? v[i0] : 0.0, sw = shift ? w[i0] : 0.0;
            rx[i] = cst / tadv1 * ((u[i] - su) / p.dt);
            ry[i] = cst / tadv1 * ((v[i] - sv) / p.dt);
            rz[i] = cst / tadv1 * ((w[i] - sw) / p.dt);
        }
    }
}

template <typename W>
void press_rhs_prep_on(les_ctx* c, W* rx, W* ry, W* rz, const double* u, const double* v, const double* w,
                       const W* divtz, double modi) {
    FamScope fs(c->stream, FAM_PRESS_PREP, (24.0 + 3.0 * sizeof(W)) * (double)c->plane * (double)c->nlev);
    KP p = kp(c);
    hipLaunchKernelGGL(k_press_prep<W>, grid_for(p.plane * (p.nz + 1)), dim3(256), 0, c->stream, rx, ry, rz, u, v, w,
                       divtz, p, 1.0 / (double)(c->nx * c->ny), (int)(sizeof(W) < sizeof(double)), modi);
}

template <typename W>
void press_rhs_prep_dev(les_ctx* c, W* rx, W* ry, W* rz) {
    press_rhs_prep_on<W>(c, rx, ry, rz, c->f3[LES_F_U], c->f3[LES_F_V], c->f3[LES_F_W],
                         (const W*)WF<W>(c, LES_F_DIVTZ), 0.0);
}

// Plane sums of the z right-hand side of the pressure equation in FP64 (the
// kx = ky = 0 coefficient of its forward transform, formed from the FP64
// velocity rather than from a reduced-precision transform), levels 0..nz.
// HZ_SPLIT blocks per level write partial sums; one thread per level adds
// them in order.
constexpr int HZ_SPLIT = 8;

template <typename W>
__global__ __launch_bounds__(256) void k_hz_part(const double* __restrict__ w, const W* __restrict__ divtz, KP p,
                                                  double cst, double modi, double* __restrict__ part) {
    __shared__ double sh[256];
    const int k = blockIdx.x / HZ_SPLIT, part_id = blockIdx.x % HZ_SPLIT;
    const double tadv1 = 1.5;
    const long chunk = (p.plane + HZ_SPLIT - 1) / HZ_SPLIT;
    const long q0 = (long)part_id * chunk;
    const long q1 = (q0 + chunk < p.plane) ? q0 + chunk : p.plane;
    double s = 0.0;
    for (long q = q0 + threadIdx.x; q < q1; q += 256) {
        const long i = (long)k * p.plane + q;
        if (k == 0) s += cst * ((double)divtz[i + p.plane] - modi);
        else if (k == p.nz) s += cst * divtz[i];
        else s += cst / tadv1 * (w[i] / p.dt);
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int m = 128; m > 0; m >>= 1) {
        if (threadIdx.x < m) sh[threadIdx.x] += sh[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_hz_fold(const double* __restrict__ part, int nlev, double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nlev) return;
    double s = 0.0;
    for (int j = 0; j < HZ_SPLIT; ++j) s += part[k * HZ_SPLIT + j];
    out[k] = s;
}

template <typename W>
void press_hz_sums_dev(les_ctx* c, double* out, double modi) {
    FamScope fs(c->stream, FAM_PRESS_PREP, (8.0 + sizeof(W)) * (double)c->plane * (double)c->nlev);
    KP p = kp(c);
    double* part = c->work3[5];  // (nz+1) * HZ_SPLIT partial sums
    hipLaunchKernelGGL(k_hz_part<W>, dim3((p.nz + 1) * HZ_SPLIT), dim3(256), 0, c->stream,
                       (const double*)c->f3[LES_F_W], (const W*)WF<W>(c, LES_F_DIVTZ), p,
                       1.0 / (double)(c->nx * c->ny), modi, part);
    hipLaunchKernelGGL(k_hz_fold, dim3((p.nz + 1 + 63) / 64), dim3(64), 0, c->stream, (const double*)part, p.nz + 1,
                       out);
}

template <typename W>
__global__ void k_dpdz(W* dpdz, const W* __restrict__ pr, KP p) {
    const long n = p.plane * (p.nz - 1);
    GRID_LOOP(t, n) {
        const long i = p.plane + t;
        dpdz[i] = ((double)pr[i] - (double)pr[i - p.plane]) / p.dz;
    }
}

template <typename W>
void dpdz_on(les_ctx* c, W* dpdz, const W* pr) {
    FamScope fs(c->stream, FAM_PRESS_PREP, 2.0 * sizeof(W) * (double)c->plane * (double)c->nlev);
    KP p = kp(c);
    hipLaunchKernelGGL(k_dpdz<W>, grid_for(p.plane * (p.nz - 1)), dim3(256), 0, c->stream, dpdz, pr, p);
}

template <typename W>
void dpdz_dev(les_ctx* c) {
    dpdz_on<W>(c, WF<W>(c, LES_F_DPDZ), (const W*)WF<W>(c, LES_F_P));
}

// ---- RHS assembly and AB2, 103_flosol.f90:115-162, 319-389 -----------------
template <typename W>
struct RhsArgs {
    W *rx, *ry, *rz, *rxf, *ryf, *rzf;  // the RHS history, in the work precision
    const W *cx, *cy, *cz, *dx, *dy, *dz, *beta;
    const double* sponge;
    double *du, *dv, *dw;  // dudt diagnostics (103_flosol.f90:119-156), or null
};

// Forcing terms of the RHS assembly (103_flosol.f90:121-156) and the
// damping layer (007_topbc.f90:92-124)
struct RhsOpts {
    int buoy;                 // buoyancy (theta_flag, not passive)
    int ocean;                // ocean Coriolis with Stokes drift (:134-149)
    int cori;                 // land Coriolis / geostrophic forcing (:150-156)
    int crossvel;             // ocean cross flow (:143-148)
    int damp1, damp23;        // damping method 1 (w only), 2 or 3 (u, v, w toward plane means)
    int euler;                // Euler start (RHS_f <- RHS)
    double coriol, ug, vg, u_cross, v_cross, force;
    const double *ust, *vst;  // Stokes drift (0:nz)
    const double* mn;         // plane means of u, v, w at levels 1..nz-1 (mn[k-1], mn[nlev+k-1], ...)
    long nlev;
};

// RHS assembly (103_flosol.f90:115-157) fused with Admas_Bashforth
// (:319-389): each point forms RHS from the same-point fields, stores it (and
// RHS_f at an Euler start, the dudt diagnostics on sampling steps), then
// advances u, v, w with it; level 0 of u, v, w <- BOGUS.  The arithmetic and
// its order are those of the loops.
template <typename W>
__global__ void k_rhs_ab2(RhsArgs<W> a, double* u, double* v, double* w, KP p, RhsOpts o) {
    const long n = p.plane * p.nz;  // levels 0..nz-1
    const double tadv1 = 1.5, tadv2 = 1.0 - tadv1;
    GRID_LOOP(i, n) {
        if (i < p.plane) {
            u[i] = BOGUS; v[i] = BOGUS; w[i] = BOGUS;
            continue;
        }
        const int k = (int)(i / p.plane);
        double rx = -(double)a.cx[i] - (double)a.dx[i];
        double ry = -(double)a.cy[i] - (double)a.dy[i];
        double rz = -(double)a.cz[i] - (double)a.dz[i];
        if (o.buoy) rz = rz + a.beta[i];
        if (o.ocean) {
            rx = rx + o.coriol * (v[i] + o.vst[k]) - o.coriol * o.vg;
            ry = ry - o.coriol * (u[i] + o.ust[k]) + o.coriol * o.ug;
            if (o.crossvel) {
                rx = rx - o.coriol * o.v_cross;
                ry = ry + o.coriol * o.u_cross;
            }
        } else if (o.cori) {
            rx = rx + o.coriol * (v[i] - o.vg);
            ry = ry + o.coriol * (o.ug - u[i]);
        }
        if (o.damp1) rz = rz - a.sponge[k] * w[i];
        if (o.damp23) {
            const double sk = 0.5 * (a.sponge[k] + a.sponge[k + 1]);
            rx = rx - sk * (u[i] - o.mn[k - 1]);
            ry = ry - sk * (v[i] - o.mn[o.nlev + k - 1]);
            rz = rz - a.sponge[k] * (w[i] - o.mn[2 * o.nlev + k - 1]);
        }
        a.rx[i] = (W)rx; a.ry[i] = (W)ry; a.rz[i] = (W)rz;
        if (a.du) {
            double du = -(double)a.dx[i], dv = -(double)a.dy[i];
            if (o.ocean) {
                du = du + o.coriol * (v[i] + o.vst[k]) - o.coriol * o.vg;
                dv = dv - o.coriol * (u[i] + o.ust[k]) + o.coriol * o.ug;
                if (o.crossvel) {
                    du = du - o.coriol * o.v_cross;
                    dv = dv + o.coriol * o.u_cross;
                }
            } else if (o.cori) {
                du = du + o.coriol * (v[i] - o.vg);
                dv = dv + o.coriol * (o.ug - u[i]);
            }
            a.du[i] = du; a.dv[i] = dv; a.dw[i] = -(double)a.dz[i];
        }
        double rxf, ryf, rzf;
        if (o.euler) {
            a.rxf[i] = (W)rx; a.ryf[i] = (W)ry; a.rzf[i] = (W)rz;
            rxf = rx; ryf = ry; rzf = rz;
        } else {
            rxf = (double)a.rxf[i]; ryf = (double)a.ryf[i]; rzf = (double)a.rzf[i];
        }
        u[i] = u[i] + p.dt * (tadv1 * rx + tadv2 * rxf + 0.0 + o.force);
        v[i] = v[i] + p.dt * (tadv1 * ry + tadv2 * ryf + 0.0);
        w[i] = w[i] + p.dt * (tadv1 * rz + tadv2 * rzf + 0.0);
    }
}

template <typename W>
void rhs_ab2_dev(les_ctx* c, int jt) {
    // the seven terms and the RHS history (work precision), u, v, w read and written
    FamScope fs(c->stream, FAM_RHS_AB2, (13.0 * sizeof(W) + 48.0) * (double)c->plane * c->nz);
    KP p = kp(c);
    const les_params& P = c->prm;
    RhsArgs<W> a;
    a.rx = WF<W>(c, LES_F_RHSX); a.ry = WF<W>(c, LES_F_RHSY); a.rz = WF<W>(c, LES_F_RHSZ);
    a.rxf = WF<W>(c, LES_F_RHSX_F); a.ryf = WF<W>(c, LES_F_RHSY_F); a.rzf = WF<W>(c, LES_F_RHSZ_F);
    a.cx = WF<W>(c, LES_F_CX); a.cy = WF<W>(c, LES_F_CY); a.cz = WF<W>(c, LES_F_CZ);
    a.dx = WF<W>(c, LES_F_DIVTX); a.dy = WF<W>(c, LES_F_DIVTY); a.dz = WF<W>(c, LES_F_DIVTZ);
    a.beta = WF<W>(c, LES_F_BETA_SCAL);
    const bool st = stats_step(c, jt);
    a.du = st ? c->st_diag[0] : nullptr;
    a.dv = st ? c->st_diag[1] : nullptr;
    a.dw = st ? c->st_diag[2] : nullptr;
    a.sponge = c->sponge;
    RhsOpts o;
    o.buoy = (int)(P.theta_flag && !P.passive_scalar);
    o.ocean = P.ocean_flag;
    o.cori = P.coriolis_forcing;
    o.crossvel = P.ocean_flag && P.flag_crossvel;
    const int damp = (P.ubc_mom == LES_BC_STRESS_FREE && P.ubc == 1);
    o.damp1 = damp && P.damping_method == 1;
    o.damp23 = damp && P.damping_method != 1;
    o.euler = (int)(jt == 1 && !P.restart_vel);
    o.coriol = P.coriol; o.ug = P.ug; o.vg = P.vg;
    o.u_cross = P.u_cross; o.v_cross = P.v_cross;
    o.force = P.use_mean_p_force ? P.mean_p_force : 0.0;
    o.ust = c->ust; o.vst = c->vst;
    o.mn = c->dmean;
    o.nlev = (long)c->nlev;
    if (o.damp23) {
        // calc_hor_avg3 of u, v, w at levels 1..nz-1 (007_topbc.f90:110-112), in order
        const double nn = (double)(c->nx * c->ny);
        level_means_seq(c->stream, c->f3[LES_F_U], c->plane, 1, c->nz - 1, nn, c->dmean);
        level_means_seq(c->stream, c->f3[LES_F_V], c->plane, 1, c->nz - 1, nn, c->dmean + c->nlev);
        level_means_seq(c->stream, c->f3[LES_F_W], c->plane, 1, c->nz - 1, nn, c->dmean + 2 * c->nlev);
    }
    hipLaunchKernelGGL(k_rhs_ab2<W>, grid_for(p.plane * p.nz), dim3(256), 0, c->stream, a, c->f3[LES_F_U],
                       c->f3[LES_F_V], c->f3[LES_F_W], p, o);
}

template <typename W>
__global__ void k_ab2(double* u, double* v, double* w, const W* __restrict__ rx, const W* __restrict__ ry,
                      const W* __restrict__ rz, const W* __restrict__ rxf, const W* __restrict__ ryf,
                      const W* __restrict__ rzf, KP p, double force) {
    const long n = p.plane * p.nz;  // levels 0..nz-1
    const double tadv1 = 1.5, tadv2 = 1.0 - tadv1;
    GRID_LOOP(i, n) {
        if (i < p.plane) {
            u[i] = BOGUS; v[i] = BOGUS; w[i] = BOGUS;
            continue;
        }
        u[i] = u[i] + p.dt * (tadv1 * (double)rx[i] + tadv2 * (double)rxf[i] + 0.0 + force);
        v[i] = v[i] + p.dt * (tadv1 * (double)ry[i] + tadv2 * (double)ryf[i] + 0.0);
        w[i] = w[i] + p.dt * (tadv1 * (double)rz[i] + tadv2 * (double)rzf[i] + 0.0);
    }
}

template <typename W>
void ab2_dev(les_ctx* c) {
    KP p = kp(c);
    const les_params& P = c->prm;
    hipLaunchKernelGGL(k_ab2<W>, grid_for(p.plane * p.nz), dim3(256), 0, c->stream, c->f3[LES_F_U], c->f3[LES_F_V],
                       c->f3[LES_F_W], (const W*)WF<W>(c, LES_F_RHSX), (const W*)WF<W>(c, LES_F_RHSY),
                       (const W*)WF<W>(c, LES_F_RHSZ), (const W*)WF<W>(c, LES_F_RHSX_F),
                       (const W*)WF<W>(c, LES_F_RHSY_F), (const W*)WF<W>(c, LES_F_RHSZ_F), p,
                       P.use_mean_p_force ? P.mean_p_force : 0.0);
}

// RHS -= grad p (103_flosol.f90:179-181; the dudt diagnostics likewise,
// :183-185, when du != null) fused with project (110_project.f90:16-40):
// u, v on levels 1..nz-1, w from level 2
template <typename W>
__global__ void k_gradp_project(W* rx, W* ry, W* rz, double* u, double* v, double* w,
                                const W* __restrict__ px, const W* __restrict__ py, const W* __restrict__ pz, KP p,
                                double* du, double* dv, double* dw) {
    const long n = p.plane * (p.nz - 1);
    const double tadv1 = 1.5;
    GRID_LOOP(t, n) {
        const long i = p.plane + t;
        const int k = (int)(i / p.plane);
        const double gx = px[i], gy = py[i], gz = pz[i];
        rx[i] = (W)((double)rx[i] - gx);
        ry[i] = (W)((double)ry[i] - gy);
        rz[i] = (W)((double)rz[i] - gz);
        if (du) {
            du[i] = du[i] - gx;
            dv[i] = dv[i] - gy;
            dw[i] = dw[i] - gz;
        }
        u[i] = u[i] + p.dt * (-tadv1 * gx);
        v[i] = v[i] + p.dt * (-tadv1 * gy);
        if (k >= 2) w[i] = w[i] + p.dt * (-tadv1 * gz);
    }
}

// project (110_project.f90:16-40) and the boundary rows that follow
template <typename W>
__global__ void k_project(double* u, double* v, double* w, const W* __restrict__ px, const W* __restrict__ py,
                          const W* __restrict__ pz, KP p) {
    const long n = p.plane * (p.nz - 1);
    const double tadv1 = 1.5;
    GRID_LOOP(t, n) {
        const long i = p.plane + t;
        const int k = (int)(i / p.plane);
        double r = -tadv1 * px[i];
        u[i] = u[i] + p.dt * r;
        r = -tadv1 * py[i];
        v[i] = v[i] + p.dt * r;
        if (k >= 2) {
            r = -tadv1 * pz[i];
            w[i] = w[i] + p.dt * r;
        }
    }
}

__global__ void k_project_bc(double* u, double* v, double* w, double* th, KP p, int sf, int theta) {
    GRID_LOOP(q, p.plane) {
        const long top = (long)p.nz * p.plane + q;
        if (sf) {
            u[top - p.plane] = u[top - 2 * p.plane];
            v[top - p.plane] = v[top - 2 * p.plane];
        }
        w[top] = 0.0;
        w[p.plane + q] = 0.0;
        u[q] = BOGUS; v[q] = BOGUS; w[q] = BOGUS;   // data_exchange, 202_parallel.f90:722-733
        u[top] = BOGUS; v[top] = BOGUS;
        if (theta) { th[q] = BOGUS; th[top] = BOGUS; }
    }
}

// theta's rows of k_project_bc, for a theta update that runs on the
// auxiliary stream: enqueued there after that update, so the rows are written
// last, as in the serial order
__global__ void k_theta_rows(double* th, KP p) {
    GRID_LOOP(q, p.plane) {
        th[q] = BOGUS;
        th[(long)p.nz * p.plane + q] = BOGUS;
    }
}

void theta_rows_dev(les_ctx* c) {
    KP p = kp(c);
    hipLaunchKernelGGL(k_theta_rows, grid_for(p.plane), dim3(256), 0, c->stream, c->f3[LES_F_THETA], p);
}

// theta_rows = false: theta's rows are left to theta_rows_dev
template <typename W>
void gradp_project_dev(les_ctx* c, int jt, bool theta_rows) {
    // RHS (work precision) read and written, u, v, w read and written, grad p read
    FamScope fs(c->stream, FAM_GRADP_PROJECT, (48.0 + 9.0 * sizeof(W)) * (double)c->plane * (c->nz - 1));
    KP p = kp(c);
    const les_params& P = c->prm;
    const bool st = stats_step(c, jt);
    hipLaunchKernelGGL(k_gradp_project<W>, grid_for(p.plane * (p.nz - 1)), dim3(256), 0, c->stream,
                       WF<W>(c, LES_F_RHSX), WF<W>(c, LES_F_RHSY), WF<W>(c, LES_F_RHSZ), c->f3[LES_F_U], c->f3[LES_F_V],
                       c->f3[LES_F_W], (const W*)WF<W>(c, LES_F_DPDX), (const W*)WF<W>(c, LES_F_DPDY),
                       (const W*)WF<W>(c, LES_F_DPDZ), p, st ? c->st_diag[0] : nullptr,
                       st ? c->st_diag[1] : nullptr, st ? c->st_diag[2] : nullptr);
    hipLaunchKernelGGL(k_project_bc, grid_for(p.plane), dim3(256), 0, c->stream, c->f3[LES_F_U], c->f3[LES_F_V],
                       c->f3[LES_F_W], c->f3[LES_F_THETA], p, (int)(P.ubc_mom == LES_BC_STRESS_FREE),
                       (int)(P.theta_flag && theta_rows));
}

template <typename W>
void project_dev(les_ctx* c) {
    KP p = kp(c);
    const les_params& P = c->prm;
    hipLaunchKernelGGL(k_project<W>, grid_for(p.plane * (p.nz - 1)), dim3(256), 0, c->stream, c->f3[LES_F_U],
                       c->f3[LES_F_V], c->f3[LES_F_W], (const W*)WF<W>(c, LES_F_DPDX),
                       (const W*)WF<W>(c, LES_F_DPDY), (const W*)WF<W>(c, LES_F_DPDZ), p);
    hipLaunchKernelGGL(k_project_bc, grid_for(p.plane), dim3(256), 0, c->stream, c->f3[LES_F_U], c->f3[LES_F_V],
                       c->f3[LES_F_W], c->f3[LES_F_THETA], p, (int)(P.ubc_mom == LES_BC_STRESS_FREE),
                       P.theta_flag);
}

// ---- check_cfl, 901_debug.f90:85-99 ----------------------------------------------
// Two-stage max over levels 1..nz-1 of u, v, w; the final block forms
// u_max dt/dx + v_max dt/dy + w_max dt/dz.
constexpr int MAXB = MAX3_BLOCKS;

__global__ __launch_bounds__(256) void k_max3_part(const double* __restrict__ u, const double* __restrict__ v,
                                                   const double* __restrict__ w, long lo, long hi,
                                                   double* __restrict__ part) {
    __shared__ double sh[3][256];
    double m[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (long i = lo + (long)blockIdx.x * 256 + threadIdx.x; i < hi; i += (long)gridDim.x * 256) {
        m[0] = fmax(m[0], u[i]);
        m[1] = fmax(m[1], v[i]);
        m[2] = fmax(m[2], w[i]);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) sh[a][threadIdx.x] = m[a];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
#pragma unroll
            for (int a = 0; a < 3; ++a) sh[a][threadIdx.x] = fmax(sh[a][threadIdx.x], sh[a][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int a = 0; a < 3; ++a) part[a * MAXB + blockIdx.x] = sh[a][0];
}

// k_max3_part over levels lo..hi fused with k_lag_acc over all n values:
// u, v, w are read once
template <typename W>
__global__ __launch_bounds__(256) void k_max3_lag(const double* __restrict__ u, const double* __restrict__ v,
                                                  const double* __restrict__ w, W* ul, W* vl, W* wl,
                                                  long n, long lo, long hi, double* __restrict__ part) {
    __shared__ double sh[3][256];
    double m[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const double a = u[i], b = v[i], d = w[i];
        ul[i] = (W)((double)ul[i] + a);
        vl[i] = (W)((double)vl[i] + b);
        wl[i] = (W)((double)wl[i] + d);
        if (i >= lo && i < hi) {
            m[0] = fmax(m[0], a);
            m[1] = fmax(m[1], b);
            m[2] = fmax(m[2], d);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) sh[a][threadIdx.x] = m[a];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
#pragma unroll
            for (int a = 0; a < 3; ++a) sh[a][threadIdx.x] = fmax(sh[a][threadIdx.x], sh[a][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int a = 0; a < 3; ++a) part[a * MAXB + blockIdx.x] = sh[a][0];
}

__global__ __launch_bounds__(256) void k_max3_final(const double* __restrict__ part, int nb, double* out,
                                                    double cu, double cv, double cw) {
    __shared__ double sh[3][256];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double m = -INFINITY;
        for (int b = threadIdx.x; b < nb; b += 256) m = fmax(m, part[a * MAXB + b]);
        sh[a][threadIdx.x] = m;
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
#pragma unroll
            for (int a = 0; a < 3; ++a) sh[a][threadIdx.x] = fmax(sh[a][threadIdx.x], sh[a][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0][0] * cu + sh[1][0] * cv + sh[2][0] * cw;
}

void max3_dev(les_ctx* c, double* out, bool lag) {
    const bool m32 = c->precision == LES_MIXED;  // u_lag in the work precision
    FamScope fs(c->stream, FAM_CFL_MAX,
                (24.0 + (lag ? 6.0 * (m32 ? 4.0 : 8.0) : 0.0)) * (double)c->plane * (double)c->nlev);
    const long lo = (long)c->plane, hi = (long)c->plane * c->nz;
    const long n = lag ? (long)c->plane * (long)c->nlev : hi - lo;
    long nb = (n + 255) / 256;
    if (nb > MAXB) nb = MAXB;
    const double dt = c->prm.dt;
    double* part = c->red + 64;  // red holds >= 64 + 3*MAXB doubles
    if (lag && m32)
        hipLaunchKernelGGL(k_max3_lag<float>, dim3((unsigned)nb), dim3(256), 0, c->stream,
                           (const double*)c->f3[LES_F_U], (const double*)c->f3[LES_F_V],
                           (const double*)c->f3[LES_F_W], WF<float>(c, LES_F_U_LAG), WF<float>(c, LES_F_V_LAG),
                           WF<float>(c, LES_F_W_LAG), n, lo, hi, part);
    else if (lag)
        hipLaunchKernelGGL(k_max3_lag<double>, dim3((unsigned)nb), dim3(256), 0, c->stream,
                           (const double*)c->f3[LES_F_U], (const double*)c->f3[LES_F_V],
                           (const double*)c->f3[LES_F_W], WF<double>(c, LES_F_U_LAG), WF<double>(c, LES_F_V_LAG),
                           WF<double>(c, LES_F_W_LAG), n, lo, hi, part);
    else
        hipLaunchKernelGGL(k_max3_part, dim3((unsigned)nb), dim3(256), 0, c->stream, (const double*)c->f3[LES_F_U],
                           (const double*)c->f3[LES_F_V], (const double*)c->f3[LES_F_W], lo, hi, part);
    hipLaunchKernelGGL(k_max3_final, dim3(1), dim3(256), 0, c->stream, (const double*)part, (int)nb, out,
                       dt / c->dx, dt / c->dy, dt / c->dz);
}

// ---- instantiations for both work precisions --------------------------------
#define LES_INST_POINT(W)                                                                                   \
    template void wallstress_dev<W>(les_ctx*, int, const double*, const double*);                           \
    template void strain_dev<W>(les_ctx*);                                                                  \
    template void nu_t_stress_dev<W>(les_ctx*);                                                             \
    template void ddz_w_accum_dev<W>(les_ctx*, const W*);                                                    \
    template void strain_stress_dev<W>(les_ctx*);                                                           \
    template void lagrange_prep_dev<W>(les_ctx*, W*);                                                       \
    template void lagrange_contract_dev<W>(les_ctx*, const W*, const W*, int, int);                         \
    template void divstress_combine_dev<W>(les_ctx*, W*, const W*, const W*, const W*, int);                \
    template void vorticity_dev<W>(les_ctx*, W*, W*, W*);                                                   \
    template void cross_product_dev<W>(les_ctx*, const W*, const W*, const W*, W*, W*, W*);                 \
    template void theta_prep_dev<W>(les_ctx*);                                                              \
    template void scalar_adv_dev<W>(les_ctx*, const W*, const W*, const W*, W*);                            \
    template void scalar_flux_dev<W>(les_ctx*, W*, int);                                                    \
    template void scalar_accum_dev<W>(les_ctx*, const W*, const W*);                                        \
    template void calc_beta_dev<W>(les_ctx*, const double*);                                                \
    template void press_rhs_prep_dev<W>(les_ctx*, W*, W*, W*);                                              \
    template void press_rhs_prep_on<W>(les_ctx*, W*, W*, W*, const double*, const double*, const double*,   \
                                       const W*, double);                                                   \
    template void dpdz_on<W>(les_ctx*, W*, const W*);                                                       \
    template void press_hz_sums_dev<W>(les_ctx*, double*, double);                                          \
    template void dpdz_dev<W>(les_ctx*);                                                                    \
    template void rhs_ab2_dev<W>(les_ctx*, int);                                                            \
    template void lag_accumulate_dev<W>(les_ctx*);                                                          \
    template void interpolag_dev<W>(les_ctx*);                                                              \
    template void ab2_dev<W>(les_ctx*);                                                                     \
    template void gradp_project_dev<W>(les_ctx*, int, bool);                                                \
    template void project_dev<W>(les_ctx*);
LES_INST_POINT(double)
LES_INST_POINT(float)
#undef LES_INST_POINT
template void ddz_uv_dev<double, double>(les_ctx*, double*, const double*);
template void ddz_w_dev<double, double>(les_ctx*, double*, const double*);
template void ddz_uv_dev<float, double>(les_ctx*, float*, const double*);
template void ddz_w_dev<float, double>(les_ctx*, float*, const double*);
template void ddz_uv_dev<float, float>(les_ctx*, float*, const float*);
template void ddz_w_dev<float, float>(les_ctx*, float*, const float*);

}  // namespace les
