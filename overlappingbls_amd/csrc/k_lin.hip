// Tangent-linear (tl) and adjoint (adj) kernels of the FP64 step.
//
// Every nonlinear pointwise or stencil kernel of the step has a tl kernel
// (the product rule about the taped forward values) and an adj kernel, its
// exact transpose written in gather form: each thread owns one input point
// and sums the contributions of every output that read it, so the adjoint
// needs no atomics and is deterministic.  Linear building blocks (line
// transforms, pads, filters, the pressure solve) are reused from the forward
// step; their transposes are the same transforms (ddx^T = -ddx, filter^T =
// filter, pad^T = (9/4) unpad, ...), applied by the host code.
//
// Level conventions (global form, levels 0..nz): a tangent field is zero
// outside the levels where the forward field is a function of the state; a
// tl kernel writes its whole output range (zeros where the forward writes a
// constant); an adj kernel accumulates (+=) into the adjoints of its inputs
// on their active levels and clears the adjoints of its outputs.
#include "les_lin.hpp"

namespace les {

#define LGRID(i, n) for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (long)gridDim.x * blockDim.x)

static inline dim3 lgrid(long n) {
    long b = (n + 255) / 256;
    if (b > 65535L * 16) b = 65535L * 16;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

LP lin_params(const les_ctx* c) {
    LP p;
    p.plane = (long)c->plane;
    p.plane2 = (long)c->plane2;
    p.nz = c->nz;
    p.dz = c->dz;
    p.dt = c->prm.dt;
    const double delta = pow(c->dx * c->dy * c->dz, 1.0 / 3.0);
    p.l2 = delta * delta;
    p.nu = c->prm.molec ? (c->prm.nu_molec / (c->prm.u_scale * c->prm.z_i)) : 0.0;
    return p;
}

// ---- utilities ---------------------------------------------------------------
__global__ void k_axpy(double* y, const double* __restrict__ x, double a, long n) {
    LGRID(i, n) y[i] = y[i] + a * x[i];
}
void lin_axpy(les_ctx* c, double* y, const double* x, double a, long n) {
    hipLaunchKernelGGL(k_axpy, lgrid(n), dim3(256), 0, c->stream, y, x, a, n);
}
__global__ void k_addc(double* y, double a, long n) {
    LGRID(i, n) y[i] = y[i] + a;
}
void lin_addc(les_ctx* c, double* y, double a, long n) {
    hipLaunchKernelGGL(k_addc, lgrid(n), dim3(256), 0, c->stream, y, a, n);
}

// counter-based uniform(-1, 1): splitmix64 of (seed, salt, index)
__device__ __forceinline__ double hash_unif(unsigned long long seed, unsigned long long salt, unsigned long long i) {
    unsigned long long z = seed * 0x9E3779B97F4A7C15ull + salt * 0xD1B54A32D192ED03ull + i;
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

__global__ void k_rand(double* f, long n, unsigned long long seed, unsigned long long salt, long off) {
    LGRID(i, n) f[i] = hash_unif(seed, salt, (unsigned long long)(off + i));
}
void lin_rand(les_ctx* c, double* f, long n, unsigned long long seed, unsigned long long salt, long off) {
    hipLaunchKernelGGL(k_rand, lgrid(n), dim3(256), 0, c->stream, f, n, seed, salt, off);
}

// partial sums over LIN_NPART blocks, tree order, of: mode 0, a[i] * b[i]
// (b null: the regenerated uniforms of (seed, salt) at index off + i);
// mode 1, a[i]^2; mode 2, a[i]
__global__ __launch_bounds__(256) void k_dot_part(const double* __restrict__ a, const double* __restrict__ b, long n,
                                                   unsigned long long seed, unsigned long long salt, long off, int sq,
                                                   double* __restrict__ part) {
    __shared__ double sh[256];
    double s = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const double x = a[i];
        const double y = sq == 1 ? x : sq == 2 ? 1.0 : (b ? b[i] : hash_unif(seed, salt, (unsigned long long)(off + i)));
        s += x * y;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// acc[0] += w * sum(part[0..np)), in index order
__global__ void k_acc_parts(const double* __restrict__ part, int np, double w, double* acc) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < np; ++i) s += part[i];
        acc[0] = acc[0] + w * s;
    }
}

void lin_dot_part(les_ctx* c, const double* a, const double* b, long n, unsigned long long seed,
                  unsigned long long salt, long off, int sq, double* part) {
    hipLaunchKernelGGL(k_dot_part, dim3(LIN_NPART), dim3(256), 0, c->stream, a, b, n, seed, salt, off, sq, part);
}
void lin_acc(les_ctx* c, const double* part, double w, double* acc) {
    hipLaunchKernelGGL(k_acc_parts, dim3(1), dim3(64), 0, c->stream, part, LIN_NPART, w, acc);
}

// ---- obukhov (theta-active branch, 010_scalars.f90:758-843) ------------------
struct ObkPt {
    double D, us, L, th, wt, arg, x, zeta, pm, fm, ph;
    int br;  // 1: L < 0, 2: L > 0, 0: neutral
};

// the forward's point values from its plane sums (same expressions as the
// forward kernel)
__device__ ObkPt obk_point(const double* s, const ObkArgs& a) {
    ObkPt o;
    const double wt_ = s[2] / a.nn;
    const double U = s[0] / a.nn;
    const double ps = s[1] / a.nn;
    o.wt = wt_;
    o.th = s[3] / a.nn;
    o.D = log(0.5 * a.dz / a.zo_) - ps;
    o.us = U * VONK / o.D;
    o.L = -(o.us * o.us * o.us) / (VONK * a.g_ / o.th * wt_);
    o.arg = 0.0; o.x = 0.0; o.zeta = 0.0;
    if (o.L < 0.0 && wt_ != 0.0) {
        o.br = 1;
        o.arg = 1.0 - 16.0 * a.dz / 2.0 / o.L;
        o.x = pow(o.arg, 0.25);
        o.pm = 2.0 * log((1.0 + o.x) / 2.0) + log((1.0 + o.x * o.x) / 2.0) - 2.0 * atan(o.x) + PI / 2.0;
        o.fm = 1.0 / o.x;
        o.ph = 2.0 * log((1.0 + o.x * o.x) / 2.0);
    } else if (o.L > 0.0 && wt_ != 0.0) {
        o.br = 2;
        o.zeta = 0.5 * a.dz / o.L;
        o.fm = 1.0 + 4.8 * o.zeta;
        o.pm = -1.0 * 4.8 * o.zeta;
        o.ph = -1.0 * 7.8 * o.zeta;
    } else {
        o.br = 0;
        o.pm = 0.0;
        o.fm = 1.0;
        o.ph = 0.0;
    }
    return o;
}

__device__ __forceinline__ double dpm_dx(double x) {
    return 2.0 / (1.0 + x) + 2.0 * x / (1.0 + x * x) - 2.0 / (1.0 + x * x);
}
__device__ __forceinline__ double dph_dx(double x) { return 4.0 * x / (1.0 + x * x); }

// block 0: sum (u du + v dv)/|u|, block 1: sum dpsi, block 2: sum dtheta(1),
// block 3: sum dq (the surface flux obukhov averages into wt_)
__global__ __launch_bounds__(256) void k_obk_tl_sums(const double* __restrict__ u1, const double* __restrict__ v1,
                                                      const double* __restrict__ du, const double* __restrict__ dv,
                                                      const double* __restrict__ dpsi, const double* __restrict__ dth,
                                                      const double* __restrict__ dq, long plane,
                                                      double* __restrict__ out) {
    __shared__ double sh[256];
    const int which = blockIdx.x;
    double s = 0.0;
    for (long i = threadIdx.x; i < plane; i += 256) {
        if (which == 0) s += (u1[i] * du[i] + v1[i] * dv[i]) / sqrt(u1[i] * u1[i] + v1[i] * v1[i]);
        else if (which == 1) s += dpsi[i];
        else if (which == 2) s += dth[i];
        else s += dq[i];
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[which] = sh[0];
}

// L = -us^3 theta / (kappa g wt): dL/L = 3 dus/us + dtheta/theta - dwt/wt
__global__ void k_obk_tl_finish(const double* __restrict__ sums, const double* __restrict__ ds, ObkArgs a,
                                double* dpsi, double* dphi, double* dpsh, long plane) {
    const ObkPt o = obk_point(sums, a);
    const double dU = ds[0] / a.nn, dps = ds[1] / a.nn, dth = ds[2] / a.nn, dwt = ds[3] / a.nn;
    const double dD = -dps;
    const double dus = VONK * dU / o.D - o.us * dD / o.D;
    const double dL = (o.br != 0) ? o.L * (3.0 * dus / o.us + dth / o.th - dwt / o.wt) : 0.0;
    double dpm = 0.0, dfm = 0.0, dph = 0.0;
    if (o.br == 1) {
        const double darg = 8.0 * a.dz * dL / (o.L * o.L);
        const double dx = 0.25 * o.x * darg / o.arg;
        dpm = dpm_dx(o.x) * dx;
        dfm = -dx / (o.x * o.x);
        dph = dph_dx(o.x) * dx;
    } else if (o.br == 2) {
        const double dz = -o.zeta * dL / o.L;
        dfm = 4.8 * dz;
        dpm = -4.8 * dz;
        dph = -7.8 * dz;
    }
    LGRID(i, plane) {
        dpsi[i] = dpm;
        dphi[i] = dfm;
        dpsh[i] = dph;
    }
}

void obk_tl(les_ctx* c, const Tape& t, LinBufs& b, const ObkArgs& a) {
    const long np = (long)c->plane;
    hipLaunchKernelGGL(k_obk_tl_sums, dim3(4), dim3(256), 0, c->stream, (const double*)t.u1, (const double*)t.v1,
                       (const double*)(b.v[L_U] + np), (const double*)(b.v[L_V] + np), (const double*)b.v[L_PSI],
                       (const double*)(b.v[L_TH] + np), (const double*)b.v[L_Q], np, b.red);
    hipLaunchKernelGGL(k_obk_tl_finish, lgrid(np), dim3(256), 0, c->stream, (const double*)t.obk,
                       (const double*)b.red, a, b.v[L_PSI], b.v[L_PHI], b.v[L_PSH], np);
}

// block 0: sum lambda_psi, block 1: sum lambda_phi, block 2: sum lambda_psi_h
__global__ __launch_bounds__(256) void k_obk_adj_sums(const double* __restrict__ lpsi, const double* __restrict__ lphi,
                                                       const double* __restrict__ lpsh, long plane,
                                                       double* __restrict__ out) {
    __shared__ double sh[256];
    const double* p = blockIdx.x == 0 ? lpsi : blockIdx.x == 1 ? lphi : lpsh;
    double s = 0.0;
    for (long i = threadIdx.x; i < plane; i += 256) s += p[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

__global__ void k_obk_adj_finish(const double* __restrict__ sums, const double* __restrict__ ls, ObkArgs a,
                                 const double* __restrict__ u1, const double* __restrict__ v1, double* lu,
                                 double* lv, double* lpsi, double* lphi, double* lpsh, double* lth, double* lq,
                                 long plane) {
    const ObkPt o = obk_point(sums, a);
    const double lpm = ls[0], lfm = ls[1], lph = ls[2];
    double lL = 0.0;
    if (o.br == 1) {
        const double lx = dpm_dx(o.x) * lpm - lfm / (o.x * o.x) + dph_dx(o.x) * lph;
        const double larg = 0.25 * o.x / o.arg * lx;
        lL = larg * 8.0 * a.dz / (o.L * o.L);
    } else if (o.br == 2) {
        const double lz = -4.8 * lpm + 4.8 * lfm - 7.8 * lph;
        lL = lz * (-o.zeta / o.L);
    }
    // neutral branch (wt = 0: L infinite): psi_m, phi_m, psi_h are constants
    const double lus = (o.br != 0) ? lL * 3.0 * o.L / o.us : 0.0;
    const double lthb = (o.br != 0) ? lL * o.L / o.th : 0.0;
    const double lwt = (o.br != 0) ? -lL * o.L / o.wt : 0.0;
    const double lU = lus * VONK / o.D;
    const double lD = -lus * o.us / o.D;
    const double l0 = lU / a.nn, l1 = -lD / a.nn, l3 = lthb / a.nn, l4 = lwt / a.nn;
    LGRID(i, plane) {
        const double m = sqrt(u1[i] * u1[i] + v1[i] * v1[i]);
        lu[i] = lu[i] + l0 * u1[i] / m;
        lv[i] = lv[i] + l0 * v1[i] / m;
        lpsi[i] = l1;
        lth[i] = lth[i] + l3;
        lq[i] = lq[i] + l4;
        lphi[i] = 0.0;
        lpsh[i] = 0.0;
    }
}

void obk_adj(les_ctx* c, const Tape& t, LinBufs& b, const ObkArgs& a) {
    const long np = (long)c->plane;
    hipLaunchKernelGGL(k_obk_adj_sums, dim3(3), dim3(256), 0, c->stream, (const double*)b.v[L_PSI],
                       (const double*)b.v[L_PHI], (const double*)b.v[L_PSH], np, b.red);
    hipLaunchKernelGGL(k_obk_adj_finish, lgrid(np), dim3(256), 0, c->stream, (const double*)t.obk,
                       (const double*)b.red, a, (const double*)t.u1, (const double*)t.v1, b.v[L_U] + np,
                       b.v[L_V] + np, b.v[L_PSI], b.v[L_PHI], b.v[L_PSH], b.v[L_TH] + np, b.v[L_Q], np);
}

// ---- wall stress, bottom (104_wallstress.f90:98-182) ---------------------------
struct WallPt {
    double denom, ua, us, cst, C, phi;
};
__device__ __forceinline__ WallPt wall_point(double u1f, double v1f, const ObkPt& o, const WallArgs& w) {
    WallPt p;
    if (w.smooth) p.denom = w.dsmooth;
    else p.denom = (log((0.5 * w.dz - w.d0) / w.zo) - o.pm) / VONK;
    p.ua = sqrt(u1f * u1f + v1f * v1f);
    p.us = p.ua / p.denom;
    p.cst = -(p.us * p.us) / p.ua;
    p.C = 1.0 / ((0.5 * w.dz - w.d0) * VONK);
    p.phi = o.fm;
    return p;
}

// dus2 (null or a plane) receives 2 us dus (the tangent of ustar^2)
__global__ void k_wall_tl(const double* __restrict__ obk, ObkArgs oa, WallArgs wa, const double* __restrict__ uv1f,
                          const double* __restrict__ u1, const double* __restrict__ v1,
                          const double* __restrict__ duv1f, const double* __restrict__ du,
                          const double* __restrict__ dv, const double* __restrict__ dpsi,
                          const double* __restrict__ dphi, double* dtxz, double* dtyz, double* ddudz, double* ddvdz,
                          double* ddus, double* dus2, long plane) {
    const ObkPt o = obk_point(obk, oa);
    LGRID(i, plane) {
        const double uf = uv1f[i], vf = uv1f[plane + i];
        const double duf = duv1f[i], dvf = duv1f[plane + i];
        const WallPt p = wall_point(uf, vf, o, wa);
        const double dden = wa.smooth ? 0.0 : -dpsi[i] / VONK;
        const double dua = (uf * duf + vf * dvf) / p.ua;
        const double dus = dua / p.denom - p.us * dden / p.denom;
        const double dcst = -2.0 * p.us * dus / p.ua + p.us * p.us * dua / (p.ua * p.ua);
        dtxz[i] = dcst * uf + p.cst * duf;
        dtyz[i] = dcst * vf + p.cst * dvf;
        const double u = u1[i], v = v1[i];
        const double g = dus * p.phi / p.ua - p.us * p.phi * dua / (p.ua * p.ua) + p.us * dphi[i] / p.ua;
        ddudz[i] = (u == 0.0) ? 0.0 : p.C * (g * u + p.us * p.phi / p.ua * du[i]);
        ddvdz[i] = (v == 0.0) ? 0.0 : p.C * (g * v + p.us * p.phi / p.ua * dv[i]);
        ddus[i] = dus;
        if (dus2) dus2[i] = 2.0 * p.us * dus;
    }
}

// wobj: weight of ustar^2 in the objective at this step
__global__ void k_wall_adj(const double* __restrict__ obk, ObkArgs oa, WallArgs wa, const double* __restrict__ uv1f,
                           const double* __restrict__ u1, const double* __restrict__ v1, double* ltxz, double* ltyz,
                           double* ldudz, double* ldvdz, double* ldus, double* luv1f, double* lu, double* lv,
                           double* lpsi, double* lphi, double wobj, long plane) {
    const ObkPt o = obk_point(obk, oa);
    LGRID(i, plane) {
        const double uf = uv1f[i], vf = uv1f[plane + i];
        const WallPt p = wall_point(uf, vf, o, wa);
        const double u = u1[i], v = v1[i];
        const double lt = ltxz[i], ly = ltyz[i];
        const double ldu = (u == 0.0) ? 0.0 : ldudz[i];
        const double ldv = (v == 0.0) ? 0.0 : ldvdz[i];
        const double lcst = lt * uf + ly * vf;
        double luf = p.cst * lt, lvf = p.cst * ly;
        const double q = u * ldu + v * ldv;
        const double lus = -2.0 * p.us / p.ua * lcst + p.C * p.phi * q / p.ua + 2.0 * p.us * wobj + ldus[i];
        double lua = p.us * p.us / (p.ua * p.ua) * lcst - p.C * p.us * p.phi * q / (p.ua * p.ua);
        lu[i] = lu[i] + p.C * p.us * p.phi / p.ua * ldu;
        lv[i] = lv[i] + p.C * p.us * p.phi / p.ua * ldv;
        lphi[i] = lphi[i] + p.C * p.us * q / p.ua;
        lua += lus / p.denom;
        const double lden = -p.us / p.denom * lus;
        if (!wa.smooth) lpsi[i] = lpsi[i] - lden / VONK;
        luf += uf / p.ua * lua;
        lvf += vf / p.ua * lua;
        luv1f[i] = luf;
        luv1f[plane + i] = lvf;
        ltxz[i] = 0.0; ltyz[i] = 0.0; ldudz[i] = 0.0; ldvdz[i] = 0.0; ldus[i] = 0.0;
    }
}

void wall_tl(les_ctx* c, const Tape& t, LinBufs& b, const ObkArgs& oa, const WallArgs& wa, double* dus2) {
    const long np = (long)c->plane;
    hipLaunchKernelGGL(k_wall_tl, lgrid(np), dim3(256), 0, c->stream, (const double*)t.obk, oa, wa,
                       (const double*)t.uv1f, (const double*)t.u1, (const double*)t.v1, (const double*)b.uv1f,
                       (const double*)(b.v[L_U] + np), (const double*)(b.v[L_V] + np), (const double*)b.v[L_PSI],
                       (const double*)b.v[L_PHI], b.v[L_TXZ] + np, b.v[L_TYZ] + np, b.v[L_UZ] + np,
                       b.v[L_VZ] + np, b.v[L_DUS], dus2, np);
}
void wall_adj(les_ctx* c, const Tape& t, LinBufs& b, const ObkArgs& oa, const WallArgs& wa, double wobj) {
    const long np = (long)c->plane;
    hipLaunchKernelGGL(k_wall_adj, lgrid(np), dim3(256), 0, c->stream, (const double*)t.obk, oa, wa,
                       (const double*)t.uv1f, (const double*)t.u1, (const double*)t.v1, b.v[L_TXZ] + np,
                       b.v[L_TYZ] + np, b.v[L_UZ] + np, b.v[L_VZ] + np, b.v[L_DUS], b.uv1f, b.v[L_U] + np,
                       b.v[L_V] + np, b.v[L_PSI], b.v[L_PHI], wobj, np);
}

// ---- vertical differences --------------------------------------------------------
// d(k) = (f(k) - f(k-1))/dz, k = 2..nz-1; 0 on levels 0, 1 and nz
__global__ void k_ddz_uv_tl(double* d, const double* __restrict__ f, LP p) {
    const long n = p.plane * (p.nz + 1);
    LGRID(i, n) {
        const int k = (int)(i / p.plane);
        d[i] = (k >= 2 && k <= p.nz - 1) ? (f[i] - f[i - p.plane]) / p.dz : 0.0;
    }
}
// d(k) = (f(k+1) - f(k))/dz, k = 1..nz-1; 0 on levels 0 and nz
__global__ void k_ddz_w_tl(double* d, const double* __restrict__ f, LP p) {
    const long n = p.plane * (p.nz + 1);
    LGRID(i, n) {
        const int k = (int)(i / p.plane);
        d[i] = (k >= 1 && k <= p.nz - 1) ? (f[i + p.plane] - f[i]) / p.dz : 0.0;
    }
}
void ddz_uv_tl(les_ctx* c, double* d, const double* f) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_ddz_uv_tl, lgrid(p.plane * (p.nz + 1)), dim3(256), 0, c->stream, d, f, p);
}
void ddz_w_tl(les_ctx* c, double* d, const double* f) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_ddz_w_tl, lgrid(p.plane * (p.nz + 1)), dim3(256), 0, c->stream, d, f, p);
}

// adjoint of (fx, fy, fz) = (ddx f, ddy f, ddz f) on levels 1..nz-1 of f:
// lf += -tx - ty + ddz^T(lz); tx = ddx(lfx), ty = ddy(lfy) (either may be
// null).  wform: ddz_w (else ddz_uv).  Clears lz.
__global__ void k_grad_adj(double* lf, const double* __restrict__ tx, const double* __restrict__ ty, double* lz, LP p,
                           int wform) {
    const long n = p.plane * (p.nz - 1);
    LGRID(t, n) {
        const long i = p.plane + t;
        const int k = (int)(i / p.plane);
        double a = 0.0;
        if (tx) a -= tx[i];
        if (ty) a -= ty[i];
        if (lz) {
            if (!wform) {
                if (k >= 2) a += lz[i] / p.dz;                         // d(k) reads f(k)
                if (k + 1 <= p.nz - 1) a -= lz[i + p.plane] / p.dz;    // d(k+1) reads f(k)
            } else {
                if (k - 1 >= 1) a += lz[i - p.plane] / p.dz;           // d(k-1) reads f(k)
                a -= lz[i] / p.dz;                                     // d(k) reads f(k)
            }
        }
        lf[i] = lf[i] + a;
    }
}
// Clears are queued per stage and issued as one launch (segment s of the
// batch in blockIdx.y): an adjoint stage clears the adjoints of its outputs
// after the kernels that read them, and no kernel of the same stage writes
// them again.
__global__ void k_clear_batch(ClearBatch cb) {
    const int s = blockIdx.y;
    double* p = cb.p[s];
    const long n = cb.n[s];
    LGRID(i, n) p[i] = 0.0;
}
void lin_flush(les_ctx* c, LinBufs& b) {
    if (b.clr.count == 0) return;
    long mx = 0;
    for (int i = 0; i < b.clr.count; ++i) mx = b.clr.n[i] > mx ? b.clr.n[i] : mx;
    long gx = (mx + 255) / 256;
    if (gx > 2048) gx = 2048;
    hipLaunchKernelGGL(k_clear_batch, dim3((unsigned)gx, (unsigned)b.clr.count), dim3(256), 0, c->stream, b.clr);
    b.clr.count = 0;
}
void lin_clear(les_ctx* c, LinBufs& b, double* p, long n) {
    if (n <= 0) return;
    if (b.clr.count == LIN_MAXCLR) lin_flush(c, b);
    b.clr.p[b.clr.count] = p;
    b.clr.n[b.clr.count] = n;
    ++b.clr.count;
}
void clear_levels(les_ctx* c, LinBufs& b, double* f, int l0, int l1) {
    if (l1 < l0) return;
    const long lo = (long)l0 * (long)c->plane, hi = (long)(l1 + 1) * (long)c->plane;
    lin_clear(c, b, f + lo, hi - lo);
}
void clear_levels2(les_ctx* c, LinBufs& b, double* f, int l0, int l1) {
    if (l1 < l0) return;
    const long lo = (long)l0 * (long)c->plane2, hi = (long)(l1 + 1) * (long)c->plane2;
    lin_clear(c, b, f + lo, hi - lo);
}
void grad_adj(les_ctx* c, double* lf, const double* tx, const double* ty, double* lz, int wform) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_grad_adj, lgrid(p.plane * (p.nz - 1)), dim3(256), 0, c->stream, lf, tx, ty, lz, p, wform);
}

// ---- strain (106_sgs_stag.f90:47-104) ------------------------------------------
__global__ void k_strain_tl(LinGrad g, LinStrain S, LP p) {
    const long n = p.plane * (p.nz + 1);
    LGRID(i, n) {
        const int k = (int)(i / p.plane);
        if (k == 0) {
            S.s11[i] = 0.0; S.s12[i] = 0.0; S.s13[i] = 0.0; S.s22[i] = 0.0; S.s23[i] = 0.0; S.s33[i] = 0.0;
            continue;
        }
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

// gradient level j (1..nz) gathers from the strain rows that read it
__global__ void k_strain_adj(LinGrad g, LinStrain S, LP p) {
    const long n = p.plane * p.nz;
    const int nz = p.nz;
    LGRID(t, n) {
        const long i = p.plane + t;
        const int j = (int)(i / p.plane);
        const long up = i + p.plane, dn = i - p.plane;
        auto at = [&](double* f, int k, long idx) -> double { return (k >= 1 && k <= nz) ? f[idx] : 0.0; };
        // rows of the "centred" fields (s11, s22, s33, and the halves of s12)
        const bool in_j = (j >= 2 && j <= nz - 1), in_j1 = (j + 1 >= 2 && j + 1 <= nz - 1);
        double a11 = 0.0, a22 = 0.0, a33 = 0.0, a12 = 0.0;
        if (j == 1) { a11 += S.s11[i]; a22 += S.s22[i]; a33 += S.s33[i]; a12 += 0.5 * S.s12[i]; }
        if (in_j) { a11 += 0.5 * S.s11[i]; a22 += 0.5 * S.s22[i]; a33 += 0.5 * S.s33[i]; a12 += 0.25 * S.s12[i]; }
        if (in_j1) {
            a11 += 0.5 * S.s11[up]; a22 += 0.5 * S.s22[up]; a33 += 0.5 * S.s33[up]; a12 += 0.25 * S.s12[up];
        }
        if (j == nz - 1) { a11 += S.s11[up]; a22 += S.s22[up]; a33 += S.s33[up]; a12 += 0.5 * S.s12[up]; }
        // s13, s23: uz/vz on the same row, wx/wy averaged at rows 1 and nz
        const double l13 = S.s13[i], l23 = S.s23[i];
        double wx = 0.0, wy = 0.0;
        if (j == 1 || j == 2) { wx += 0.25 * at(S.s13, 1, j == 1 ? i : dn); wy += 0.25 * at(S.s23, 1, j == 1 ? i : dn); }
        if (in_j) { wx += 0.5 * l13; wy += 0.5 * l23; }
        if (j == nz - 1 || j == nz) {
            const long r = (j == nz) ? i : up;
            wx += 0.25 * S.s13[r];
            wy += 0.25 * S.s23[r];
        }
        g.ux[i] = a11; g.vy[i] = a22; g.wz[i] = a33;  // read by the strain only
        g.uy[i] += a12; g.vx[i] += a12;
        g.uz[i] += 0.5 * l13; g.vz[i] += 0.5 * l23;
        g.wx[i] += wx; g.wy[i] += wy;
    }
}

void strain_tl(les_ctx* c, LinBufs& b) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_strain_tl, lgrid(p.plane * (p.nz + 1)), dim3(256), 0, c->stream, lin_grad(b), lin_strain(b),
                       p);
}
void strain_adj(les_ctx* c, LinBufs& b) {
    const LP p = lin_params(c);
    // lambda_S needs no clearing: k_stress_adj, its only source, writes all of levels 1..nz
    hipLaunchKernelGGL(k_strain_adj, lgrid(p.plane * p.nz), dim3(256), 0, c->stream, lin_grad(b), lin_strain(b), p);
}

// ---- Nu_t and tau_ij (106_sgs_stag.f90:206-315), Cs_opt2 frozen ------------------
__device__ __forceinline__ double smag6(const double* s) {
    return sqrt(2.0 * (s[0] * s[0] + s[3] * s[3] + s[5] * s[5] + 2.0 * (s[1] * s[1] + s[2] * s[2] + s[4] * s[4])));
}
// d|S| = sum w_m s_m ds_m / |S| with w = 2 (diagonal), 4 (off-diagonal)
__device__ __forceinline__ double dsmag6(const double* s, const double* ds, double sm) {
    if (sm == 0.0) return 0.0;
    return 2.0 * (s[0] * ds[0] + s[3] * ds[3] + s[5] * ds[5] + 2.0 * (s[1] * ds[1] + s[2] * ds[2] + s[4] * ds[4])) / sm;
}
__device__ __forceinline__ void load6(const LinStrain& S, long i, double* s) {
    s[0] = S.s11[i]; s[1] = S.s12[i]; s[2] = S.s13[i]; s[3] = S.s22[i]; s[4] = S.s23[i]; s[5] = S.s33[i];
}
__device__ __forceinline__ void load6t(const double* const* T, long i, double* s) {
#pragma unroll
    for (int m = 0; m < 6; ++m) s[m] = T[m][i];
}

// tangent of Nu (levels 1..nz) and of txx txy tyy tzz (1..nz-1), txz tyz
// (2..nz-1; level 1 belongs to the wall model, levels 0 and nz are zero)
__global__ void k_stress_tl(TapeS ts, LinStrain dS, double* dnu, LinTau dt, LP p) {
    const long n = p.plane * (p.nz + 1);
    LGRID(i, n) {
        const int k = (int)(i / p.plane);
        if (k == 0 || k == p.nz) {
            dt.xx[i] = 0.0; dt.xy[i] = 0.0; dt.yy[i] = 0.0; dt.zz[i] = 0.0; dt.xz[i] = 0.0; dt.yz[i] = 0.0;
            if (k == 0) { dnu[i] = 0.0; continue; }
        }
        double s[6], ds[6];
        load6t(ts.S, i, s);
        load6(dS, i, ds);
        const double sm = smag6(s);
        const double dn_k = dsmag6(s, ds, sm) * ts.cs[i] * p.l2;
        dnu[i] = dn_k;
        if (k == p.nz) continue;
        const double nu_k = ts.nu[i];
        const int o[4] = {0, 1, 3, 5};
        double* out[4] = {dt.xx, dt.xy, dt.yy, dt.zz};
        if (k == 1) {
            const double f = -2.0 * (nu_k + p.nu), df = -2.0 * dn_k;
#pragma unroll
            for (int m = 0; m < 4; ++m) out[m][i] = df * s[o[m]] + f * ds[o[m]];
            continue;
        }
        const long t = i + p.plane;
        double s2[6], ds2[6];
        load6t(ts.S, t, s2);
        load6(dS, t, ds2);
        const double sm2 = smag6(s2);
        const double dn_t = dsmag6(s2, ds2, sm2) * ts.cs[t] * p.l2;
        if (k == p.nz - 1) {
            const double f = -2.0 * (ts.nu[t] + p.nu), df = -2.0 * dn_t;
#pragma unroll
            for (int m = 0; m < 4; ++m) out[m][i] = df * s2[o[m]] + f * ds2[o[m]];
        } else {
            const double f = -2.0 * (0.5 * (nu_k + ts.nu[t]) + p.nu), df = -2.0 * 0.5 * (dn_k + dn_t);
#pragma unroll
            for (int m = 0; m < 4; ++m)
                out[m][i] = df * 0.5 * (s[o[m]] + s2[o[m]]) + f * 0.5 * (ds[o[m]] + ds2[o[m]]);
        }
        dt.xz[i] = -2.0 * dn_k * s[2] - 2.0 * (nu_k + p.nu) * ds[2];
        dt.yz[i] = -2.0 * dn_k * s[4] - 2.0 * (nu_k + p.nu) * ds[4];
    }
}

// level k (1..nz): lambda_S(k), written (this is its only source), and the
// total lambda_Nu(k) from the tau rows that read level k (rows k and k-1) and
// from the scalar fluxes (lnu on entry); the host clears lambda_Nu once every
// level has been gathered
__global__ void k_stress_adj(TapeS ts, LinStrain lS, const double* __restrict__ lnu, LinTau lt, LP p) {
    const long n = p.plane * p.nz;
    const int nz = p.nz;
    LGRID(t0, n) {
        const long i = p.plane + t0;
        const int k = (int)(i / p.plane);
        const int o[4] = {0, 1, 3, 5};
        const double* L4[4] = {lt.xx, lt.xy, lt.yy, lt.zz};
        double s[6];
        load6t(ts.S, i, s);
        const double nu_k = ts.nu[i];
        double ls[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        double lnk = lnu[i];
        if (k == 1) {                                        // row 1: f(1) S(1)
            const double f = -2.0 * (nu_k + p.nu);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const double l = L4[m][i];
                ls[o[m]] += f * l;
                lnk += -2.0 * l * s[o[m]];
            }
        }
        if (k >= 2 && k <= nz - 2) {                         // row k: averages of k, k+1
            const long up = i + p.plane;
            const double f = -2.0 * (0.5 * (nu_k + ts.nu[up]) + p.nu);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const double l = L4[m][i];
                const double sa = 0.5 * (s[o[m]] + ts.S[o[m]][up]);
                ls[o[m]] += f * 0.5 * l;
                lnk += -2.0 * 0.5 * l * sa;
            }
        }
        if (k - 1 >= 2 && k - 1 <= nz - 2) {                 // row k-1 reads level k as its upper level
            const long dn = i - p.plane;
            const double f = -2.0 * (0.5 * (ts.nu[dn] + nu_k) + p.nu);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const double l = L4[m][dn];
                const double sa = 0.5 * (ts.S[o[m]][dn] + s[o[m]]);
                ls[o[m]] += f * 0.5 * l;
                lnk += -2.0 * 0.5 * l * sa;
            }
        }
        if (k == nz && nz - 1 >= 2) {                        // row nz-1: f(Nu(nz)) S(nz)
            const long dn = i - p.plane;
            const double f = -2.0 * (nu_k + p.nu);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const double l = L4[m][dn];
                ls[o[m]] += f * l;
                lnk += -2.0 * l * s[o[m]];
            }
        }
        if (k >= 2 && k <= nz - 1) {                         // txz, tyz rows
            const double f = -2.0 * (nu_k + p.nu);
            const double lx = lt.xz[i], ly = lt.yz[i];
            ls[2] += f * lx;
            ls[4] += f * ly;
            lnk += -2.0 * (lx * s[2] + ly * s[4]);
        }
        // Nu(k) = |S(k)| Cs l2
        const double sm = smag6(s);
        if (sm != 0.0) {
            const double a = lnk * ts.cs[i] * p.l2 / sm;
            ls[0] += 2.0 * s[0] * a; ls[3] += 2.0 * s[3] * a; ls[5] += 2.0 * s[5] * a;
            ls[1] += 4.0 * s[1] * a; ls[2] += 4.0 * s[2] * a; ls[4] += 4.0 * s[4] * a;
        }
        lS.s11[i] = ls[0]; lS.s12[i] = ls[1]; lS.s13[i] = ls[2];
        lS.s22[i] = ls[3]; lS.s23[i] = ls[4]; lS.s33[i] = ls[5];
    }
}

void stress_tl(les_ctx* c, const Tape& t, LinBufs& b) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_stress_tl, lgrid(p.plane * (p.nz + 1)), dim3(256), 0, c->stream, tape_s(t), lin_strain(b),
                       b.v[L_NU], lin_tau(b), p);
}
void stress_adj(les_ctx* c, const Tape& t, LinBufs& b) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_stress_adj, lgrid(p.plane * p.nz), dim3(256), 0, c->stream, tape_s(t), lin_strain(b),
                       (const double*)b.v[L_NU], lin_tau(b), p);
    // no clearing: k_div_adj writes levels 1..nz-1 of the six tau adjoints and
    // k_sflux_adj levels 1..nz of lambda_Nu
}

// ---- divstress (107_divstress.f90:36-102) ------------------------------------------
// out(k) = a + b + (f(k+1) - f(k))/dz on 1..nz-1 (uv form), or a + b +
// [2..nz-1](f(k) - f(k-1))/dz on 1..nz (w form); zero elsewhere
__global__ void k_div_tl(double* out, const double* __restrict__ a, const double* __restrict__ b,
                         const double* __restrict__ f, LP p, int wver) {
    const long n = p.plane * (p.nz + 1);
    LGRID(i, n) {
        const int k = (int)(i / p.plane);
        double r = 0.0;
        if (!wver) {
            if (k >= 1 && k <= p.nz - 1) r = a[i] + b[i] + (f[i + p.plane] - f[i]) / p.dz;
        } else {
            if (k == 1 || k == p.nz) r = a[i] + b[i];
            else if (k >= 2) r = a[i] + b[i] + (f[i] - f[i - p.plane]) / p.dz;
        }
        out[i] = r;
    }
}
void div_tl(les_ctx* c, double* out, const double* a, const double* b, const double* f, int wver) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_div_tl, lgrid(p.plane * (p.nz + 1)), dim3(256), 0, c->stream, out, a, b, f, p, wver);
}

// t[0..5] = ddx(ldx), ddy(ldx), ddx(ldy), ddy(ldy), ddx(ldz), ddy(ldz)
__global__ void k_div_adj(LinTau lt, const double* __restrict__ t0, const double* __restrict__ t1,
                          const double* __restrict__ t2, const double* __restrict__ t3, const double* __restrict__ t4,
                          const double* __restrict__ t5, double* ldx, double* ldy, double* ldz, LP p) {
    const long n = p.plane * (p.nz - 1);
    const int nz = p.nz;
    LGRID(t, n) {
        const long i = p.plane + t;
        const int k = (int)(i / p.plane);
        lt.xx[i] = -t0[i];
        lt.xy[i] = -t1[i] - t2[i];
        lt.yy[i] = -t3[i];
        // txz(k): DIVX rows k-1 (+1/dz) and k (-1/dz), DIVZ row k via ddx
        double ax = -t4[i] - ldx[i] / p.dz;
        double ay = -t5[i] - ldy[i] / p.dz;
        if (k - 1 >= 1) {
            ax += ldx[i - p.plane] / p.dz;
            ay += ldy[i - p.plane] / p.dz;
        }
        lt.xz[i] = ax;
        lt.yz[i] = ay;
        // tzz(k): DIVZ rows k (+1/dz, k in 2..nz-1) and k+1 (-1/dz)
        double az = 0.0;
        if (k >= 2) az += ldz[i] / p.dz;
        if (k + 1 <= nz - 1) az -= ldz[i + p.plane] / p.dz;
        lt.zz[i] = az;
    }
}
void div_adj(les_ctx* c, LinBufs& b, double* const* t) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_div_adj, lgrid(p.plane * (p.nz - 1)), dim3(256), 0, c->stream, lin_tau(b),
                       (const double*)t[0], (const double*)t[1], (const double*)t[2], (const double*)t[3],
                       (const double*)t[4], (const double*)t[5], b.v[L_DVX], b.v[L_DVY], b.v[L_DVZ], p);
    // lambda_DIVX, DIVY: k_rhs_adj writes levels 1..nz-1 (no clearing); the
    // pressure adjoint adds into lambda_DIVZ first, so it is cleared
    clear_levels(c, b, b.v[L_DVZ], 1, p.nz);
}

// ---- vorticity (108_convec.f90:44-69) ------------------------------------------------
__global__ void k_vort_tl(double* v1, double* v2, double* v3, LinGrad g, LP p) {
    const long n = p.plane * (p.nz + 1);
    LGRID(i, n) {
        const int k = (int)(i / p.plane);
        if (k == 0) { v1[i] = 0.0; v2[i] = 0.0; v3[i] = 0.0; continue; }
        if (k == 1) {
            v1[i] = 0.5 * (g.wy[i] + g.wy[i + p.plane]) - g.vz[i];
            v2[i] = g.uz[i] - 0.5 * (g.wx[i] + g.wx[i + p.plane]);
        } else if (k == p.nz) {
            v1[i] = 0.5 * (g.wy[i] + g.wy[i - p.plane]) - g.vz[i];
            v2[i] = g.uz[i] - 0.5 * (g.wx[i] + g.wx[i - p.plane]);
        } else {
            v1[i] = g.wy[i] - g.vz[i];
            v2[i] = g.uz[i] - g.wx[i];
        }
        v3[i] = g.vx[i] - g.uy[i];
    }
}
__global__ void k_vort_adj(double* l1, double* l2, double* l3, LinGrad g, LP p) {
    const long n = p.plane * p.nz;
    const int nz = p.nz;
    LGRID(t, n) {
        const long i = p.plane + t;
        const int j = (int)(i / p.plane);
        const long dn = i - p.plane, up = i + p.plane;
        double wy = 0.0, wx = 0.0;
        if (j == 1) { wy += 0.5 * l1[i]; wx += 0.5 * l2[i]; }
        if (j == 2) { wy += 0.5 * l1[dn]; wx += 0.5 * l2[dn]; }
        if (j >= 2 && j <= nz - 1) { wy += l1[i]; wx += l2[i]; }
        if (j == nz) { wy += 0.5 * l1[i]; wx += 0.5 * l2[i]; }
        if (j == nz - 1) { wy += 0.5 * l1[up]; wx += 0.5 * l2[up]; }
        // first (in reverse order) of these six gradient adjoints: written
        g.wy[i] = wy;
        g.wx[i] = -wx;
        g.vz[i] = -l1[i];
        g.uz[i] = l2[i];
        g.vx[i] = l3[i];
        g.uy[i] = -l3[i];
    }
}
void vort_tl(les_ctx* c, LinBufs& b) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_vort_tl, lgrid(p.plane * (p.nz + 1)), dim3(256), 0, c->stream, b.v[L_O1], b.v[L_O2],
                       b.v[L_O3], lin_grad(b), p);
}
void vort_adj(les_ctx* c, LinBufs& b) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_vort_adj, lgrid(p.plane * p.nz), dim3(256), 0, c->stream, b.v[L_O1], b.v[L_O2], b.v[L_O3],
                       lin_grad(b), p);
    for (int m = 0; m < 3; ++m) clear_levels(c, b, b.v[L_O1 + m], 1, p.nz);
}

// ---- cross product on the 3/2 grid (108_convec.f90:103-202) ---------------------------
__global__ void k_cross_tl(TapeBig tb, const double* __restrict__ ust, const double* __restrict__ vst,
                           LinBig d, LP p) {
    const long P2 = p.plane2;
    const long n = P2 * (p.nz + 1);
    LGRID(i, n) {
        const int k = (int)(i / P2);
        if (k == 0 || k == p.nz) { d.c1[i] = 0.0; d.c2[i] = 0.0; d.c3[i] = 0.0; continue; }
        const double *ub = tb.ub, *vb = tb.vb, *wb = tb.wb, *b1 = tb.o1, *b2 = tb.o2, *b3 = tb.o3;
        const long u = i + P2;
        if (k == 1) {
            d.c1[i] = d.vb[i] * (-b3[i]) + (vb[i] + vst[1]) * (-d.o3[i]) + 0.5 * (d.wb[u] * b2[u] + wb[u] * d.o2[u]);
            d.c2[i] = d.ub[i] * b3[i] + (ub[i] + ust[1]) * d.o3[i] - 0.5 * (d.wb[u] * b1[u] + wb[u] * d.o1[u]);
            d.c3[i] = 0.0;
        } else {
            const long m = i - P2;
            d.c1[i] = d.vb[i] * (-b3[i]) + (vb[i] + vst[k]) * (-d.o3[i]) +
                      0.5 * (d.wb[u] * b2[u] + wb[u] * d.o2[u] + d.wb[i] * b2[i] + wb[i] * d.o2[i]);
            d.c2[i] = d.ub[i] * b3[i] + (ub[i] + ust[k]) * d.o3[i] -
                      0.5 * (d.wb[u] * b1[u] + wb[u] * d.o1[u] + d.wb[i] * b1[i] + wb[i] * d.o1[i]);
            d.c3[i] = 0.5 * ((d.ub[i] + d.ub[m]) * (-b2[i]) + (ub[i] + ub[m] + ust[k] + ust[k - 1]) * (-d.o2[i]) +
                             (d.vb[i] + d.vb[m]) * b1[i] + (vb[i] + vb[m] + vst[k] + vst[k - 1]) * d.o1[i]);
        }
    }
}

// padded level j (1..nz) gathers from the cross rows k = j-1, j, j+1
__global__ void k_cross_adj(TapeBig tb, const double* __restrict__ ust, const double* __restrict__ vst, LinBig l,
                            LP p) {
    const long P2 = p.plane2;
    const long n = P2 * p.nz;
    const int nz = p.nz;
    LGRID(t, n) {
        const long i = P2 + t;
        const int j = (int)(i / P2);
        const long dn = i - P2, up = i + P2;
        const double *ub = tb.ub, *vb = tb.vb, *wb = tb.wb, *b1 = tb.o1, *b2 = tb.o2, *b3 = tb.o3;
        const bool row_j = (j <= nz - 1);                   // cx, cy rows exist at 1..nz-1
        const double lx = row_j ? l.c1[i] : 0.0, ly = row_j ? l.c2[i] : 0.0;
        const double lz = (j >= 2 && j <= nz - 1) ? l.c3[i] : 0.0;
        const double lzu = (j + 1 >= 2 && j + 1 <= nz - 1) ? l.c3[up] : 0.0;
        const double lxd = (j >= 2) ? l.c1[dn] : 0.0, lyd = (j >= 2) ? l.c2[dn] : 0.0;
        const double lxo = (j >= 2 && j <= nz - 1) ? lx : 0.0, lyo = (j >= 2 && j <= nz - 1) ? ly : 0.0;
        // the padded vorticity is read by the cross product only: its
        // adjoint on levels 1..nz is written (o3 is zero on level nz)
        if (j <= nz - 1) {
            l.ub[i] += ly * b3[i] - 0.5 * b2[i] * lz - 0.5 * b2[up] * lzu;
            l.vb[i] += -lx * b3[i] + 0.5 * b1[i] * lz + 0.5 * b1[up] * lzu;
        }
        l.o3[i] = (j <= nz - 1) ? -(vb[i] + vst[j]) * lx + (ub[i] + ust[j]) * ly : 0.0;
        if (j >= 2 && j <= nz - 1) l.wb[i] += 0.5 * b2[i] * (lxd + lxo) - 0.5 * b1[i] * (lyd + lyo);
        double a2 = 0.5 * wb[i] * (lxd + lxo);
        double a1 = -0.5 * wb[i] * (lyd + lyo);
        if (j >= 2 && j <= nz - 1) {
            a2 += -0.5 * (ub[i] + ub[dn] + ust[j] + ust[j - 1]) * lz;
            a1 += 0.5 * (vb[i] + vb[dn] + vst[j] + vst[j - 1]) * lz;
        }
        l.o2[i] = a2;
        l.o1[i] = a1;
    }
}
void cross_tl(les_ctx* c, const Tape& t, LinBufs& b) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_cross_tl, lgrid(p.plane2 * (p.nz + 1)), dim3(256), 0, c->stream, tape_big(t),
                       (const double*)c->ust, (const double*)c->vst, lin_big(b), p);
}
void cross_adj(les_ctx* c, const Tape& t, LinBufs& b) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_cross_adj, lgrid(p.plane2 * p.nz), dim3(256), 0, c->stream, tape_big(t),
                       (const double*)c->ust, (const double*)c->vst, lin_big(b), p);
    for (int m = 0; m < 3; ++m) clear_levels2(c, b, b.v[L_CB1 + m], 1, p.nz - 1);
}

// ---- scalar advection on the 3/2 grid (010_scalars.f90:466-484) ------------------------
__global__ void k_sadv_tl(TapeBig tb, const double* __restrict__ ust, const double* __restrict__ vst, LinBig d,
                          LP p) {
    const long P2 = p.plane2;
    const long n = P2 * (p.nz + 1);
    LGRID(i, n) {
        const int k = (int)(i / P2);
        if (k == 0 || k == p.nz) { d.rm[i] = 0.0; continue; }
        const long u = i + P2;
        double r = d.ub[i] * tb.t1[i] + (tb.ub[i] + ust[k]) * d.t1[i] + d.vb[i] * tb.t2[i] +
                   (tb.vb[i] + vst[k]) * d.t2[i];
        if (k == 1) r += 0.5 * (d.wb[u] * tb.t3[u] + tb.wb[u] * d.t3[u]);
        else r += 0.5 * (d.wb[i] * tb.t3[i] + tb.wb[i] * d.t3[i] + d.wb[u] * tb.t3[u] + tb.wb[u] * d.t3[u]);
        d.rm[i] = r;
    }
}
__global__ void k_sadv_adj(TapeBig tb, const double* __restrict__ ust, const double* __restrict__ vst, LinBig l,
                           LP p) {
    const long P2 = p.plane2;
    const long n = P2 * (p.nz - 1);
    const int nz = p.nz;
    LGRID(t, n) {
        const long i = P2 + t;
        const int j = (int)(i / P2);
        // the scalar advection is the only reader of the padded theta
        // gradients and the first (in reverse order) of the padded velocities:
        // their adjoints on levels 1..nz-1 are written, not accumulated
        const double r = l.rm[i];
        l.ub[i] = r * tb.t1[i];
        l.vb[i] = r * tb.t2[i];
        l.t1[i] = (tb.ub[i] + ust[j]) * r;
        l.t2[i] = (tb.vb[i] + vst[j]) * r;
        double aw = 0.0, at = 0.0;
        if (j >= 2) {
            const double s = l.rm[i - P2] + ((j <= nz - 1) ? r : 0.0);
            aw = 0.5 * tb.t3[i] * s;
            at = 0.5 * tb.wb[i] * s;
        }
        l.wb[i] = aw;
        l.t3[i] = at;
    }
}
void sadv_tl(les_ctx* c, const Tape& t, LinBufs& b) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_sadv_tl, lgrid(p.plane2 * (p.nz + 1)), dim3(256), 0, c->stream, tape_big(t),
                       (const double*)c->ust, (const double*)c->vst, lin_big(b), p);
}
void sadv_adj(les_ctx* c, const Tape& t, LinBufs& b) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_sadv_adj, lgrid(p.plane2 * (p.nz - 1)), dim3(256), 0, c->stream, tape_big(t),
                       (const double*)c->ust, (const double*)c->vst, lin_big(b), p);
    clear_levels2(c, b, b.v[L_RM], 1, p.nz - 1);
}

// ---- SGS scalar fluxes (010_scalars.f90:513-594) -------------------------------------
__device__ __forceinline__ double nu_avg(const double* nu, long i, int k, long plane, int nz) {
    if (k == 1) return nu[i];
    if (k == nz - 1) return nu[i + plane];
    return 0.5 * (nu[i] + nu[i + plane]);
}
__global__ void k_sflux_tl(const double* __restrict__ nu, const double* __restrict__ tx,
                           const double* __restrict__ ty, const double* __restrict__ tz,
                           const double* __restrict__ dnu, const double* __restrict__ dtx,
                           const double* __restrict__ dty, const double* __restrict__ dtz, double* fx, double* fy,
                           double* fz, LP p, double iPr) {
    const long n = p.plane * (p.nz + 1);
    LGRID(i, n) {
        const int k = (int)(i / p.plane);
        double a = 0.0, b = 0.0, z = 0.0;
        if (k >= 1 && k <= p.nz - 1) {
            const double na = nu_avg(nu, i, k, p.plane, p.nz), dna = nu_avg(dnu, i, k, p.plane, p.nz);
            a = iPr * (dna * tx[i] + na * dtx[i]);
            b = iPr * (dna * ty[i] + na * dty[i]);
        }
        if (k >= 2) z = iPr * (dnu[i] * tz[i] + nu[i] * dtz[i]);
        fx[i] = a; fy[i] = b; fz[i] = z;
    }
}
__global__ void k_sflux_adj(const double* __restrict__ nu, const double* __restrict__ tx,
                            const double* __restrict__ ty, const double* __restrict__ tz, double* lnu, double* ltx,
                            double* lty, double* ltz, const double* __restrict__ lfx, const double* __restrict__ lfy,
                            const double* __restrict__ lfz, LP p, double iPr) {
    const long n = p.plane * p.nz;
    const int nz = p.nz;
    LGRID(t, n) {
        const long i = p.plane + t;
        const int j = (int)(i / p.plane);
        const long dn = i - p.plane;
        double a = 0.0;
        if (j == 1) a += tx[i] * lfx[i] + ty[i] * lfy[i];
        if (j >= 2 && j <= nz - 2) a += 0.5 * (tx[i] * lfx[i] + ty[i] * lfy[i]);
        if (j - 1 >= 2 && j - 1 <= nz - 2) a += 0.5 * (tx[dn] * lfx[dn] + ty[dn] * lfy[dn]);
        if (j == nz) a += tx[dn] * lfx[dn] + ty[dn] * lfy[dn];
        if (j >= 2) a += tz[i] * lfz[i];
        // first (in reverse order) of the adjoints of Nu and of the theta
        // gradients: written on levels 1..nz, not accumulated
        lnu[i] = iPr * a;
        double ax = 0.0, ay = 0.0, az = 0.0;
        if (j <= nz - 1) {
            const double na = nu_avg(nu, i, j, p.plane, nz);
            ax = iPr * na * lfx[i];
            ay = iPr * na * lfy[i];
            if (j >= 2) az = iPr * nu[i] * lfz[i];
        }
        ltx[i] = ax;
        lty[i] = ay;
        ltz[i] = az;
    }
}
void sflux_tl(les_ctx* c, const Tape& t, LinBufs& b) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_sflux_tl, lgrid(p.plane * (p.nz + 1)), dim3(256), 0, c->stream, (const double*)t.nu,
                       (const double*)t.dT[0], (const double*)t.dT[1], (const double*)t.dT[2],
                       (const double*)b.v[L_NU], (const double*)b.v[L_TX], (const double*)b.v[L_TY],
                       (const double*)b.v[L_TZ], b.v[L_FX], b.v[L_FY], b.v[L_FZ], p, 1.0 / c->prm.Pr);
}
void sflux_adj(les_ctx* c, const Tape& t, LinBufs& b) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_sflux_adj, lgrid(p.plane * p.nz), dim3(256), 0, c->stream, (const double*)t.nu,
                       (const double*)t.dT[0], (const double*)t.dT[1], (const double*)t.dT[2], b.v[L_NU], b.v[L_TX],
                       b.v[L_TY], b.v[L_TZ], (const double*)b.v[L_FX], (const double*)b.v[L_FY],
                       (const double*)b.v[L_FZ], p, 1.0 / c->prm.Pr);
    // the flux adjoints need no clearing: k_sdiv_adj writes them
}

// ---- bulk surface flux of lbc = 0 (010_scalars.f90:359-394) -------------------------
// q = (T_s - theta(1)) kappa ustar / (log_zos - psi_h), with theta(1) the
// filtered theta, ustar the wall model's per-point value and psi_h obukhov's;
// sgs_t3(:,:,1) = q and the z flux at level 1 is -q.  lbc = 1 prescribes q.
__global__ void k_sfc_tl(const double* __restrict__ obk, ObkArgs oa, WallArgs wa, SfcArgs sa,
                         const double* __restrict__ uv1f, const double* __restrict__ th1,
                         const double* __restrict__ dth1, const double* __restrict__ dus,
                         const double* __restrict__ dpsh, double* dq, double* fz1, long plane) {
    const ObkPt o = obk_point(obk, oa);
    LGRID(i, plane) {
        double d = 0.0;
        if (sa.bulk) {
            const WallPt w = wall_point(uv1f[i], uv1f[plane + i], o, wa);
            const double den = sa.log_zos - o.ph;
            const double dT = sa.T_s - th1[i];
            const double q = dT * VONK * w.us / den;
            d = VONK / den * (-dth1[i] * w.us + dT * dus[i]) + q / den * dpsh[i];
        }
        dq[i] = d;
        fz1[i] = -1.0 * d;
    }
}
__global__ void k_sfc_adj(const double* __restrict__ obk, ObkArgs oa, WallArgs wa, SfcArgs sa,
                          const double* __restrict__ uv1f, const double* __restrict__ th1, double* lth1,
                          double* lus, double* lpsh, double* lq, double* lfz1, long plane) {
    const ObkPt o = obk_point(obk, oa);
    LGRID(i, plane) {
        if (sa.bulk) {
            const WallPt w = wall_point(uv1f[i], uv1f[plane + i], o, wa);
            const double den = sa.log_zos - o.ph;
            const double dT = sa.T_s - th1[i];
            const double q = dT * VONK * w.us / den;
            const double l = lq[i] - lfz1[i];
            lth1[i] = lth1[i] - VONK / den * w.us * l;
            lus[i] = lus[i] + VONK / den * dT * l;
            lpsh[i] = lpsh[i] + q / den * l;
        }
        lq[i] = 0.0;
        lfz1[i] = 0.0;
    }
}
void sfc_tl(les_ctx* c, const Tape& t, LinBufs& b, const ObkArgs& oa, const WallArgs& wa, const SfcArgs& sa) {
    const long np = (long)c->plane;
    hipLaunchKernelGGL(k_sfc_tl, lgrid(np), dim3(256), 0, c->stream, (const double*)t.obk, oa, wa, sa,
                       (const double*)t.uv1f, (const double*)(t.thf + np), (const double*)(b.v[L_TH] + np),
                       (const double*)b.v[L_DUS], (const double*)b.v[L_PSH], b.v[L_Q], b.v[L_FZ] + np, np);
}
void sfc_adj(les_ctx* c, const Tape& t, LinBufs& b, const ObkArgs& oa, const WallArgs& wa, const SfcArgs& sa) {
    const long np = (long)c->plane;
    hipLaunchKernelGGL(k_sfc_adj, lgrid(np), dim3(256), 0, c->stream, (const double*)t.obk, oa, wa, sa,
                       (const double*)t.uv1f, (const double*)(t.thf + np), b.v[L_TH] + np, b.v[L_DUS],
                       b.v[L_PSH], b.v[L_Q], b.v[L_FZ] + np, np);
}

// RT(k) = -adv + t1 + t2 + (fz(k+1) - fz(k))/dz on 1..nz-1; zero elsewhere
__global__ void k_sdiv_tl(double* rt, const double* __restrict__ adv, const double* __restrict__ t1,
                          const double* __restrict__ t2, const double* __restrict__ fz, LP p) {
    const long n = p.plane * (p.nz + 1);
    LGRID(i, n) {
        const int k = (int)(i / p.plane);
        rt[i] = (k >= 1 && k <= p.nz - 1) ? (-1.0 * adv[i] + t1[i] + t2[i] + (fz[i + p.plane] - fz[i]) / p.dz) : 0.0;
    }
}
// t1 = ddx(lrt), t2 = ddy(lrt)
__global__ void k_sdiv_adj(double* lrt, double* ladv, double* lfx, double* lfy, double* lfz,
                           const double* __restrict__ t1, const double* __restrict__ t2, LP p) {
    const long n = p.plane * p.nz;
    const int nz = p.nz;
    LGRID(t, n) {
        const long i = p.plane + t;
        const int j = (int)(i / p.plane);
        // the scalar RHS is the only reader of the advection term and the
        // fluxes: their adjoints are written on levels 1..nz
        const bool in = j <= nz - 1;
        ladv[i] = in ? -lrt[i] : 0.0;
        lfx[i] = in ? -t1[i] : 0.0;
        lfy[i] = in ? -t2[i] : 0.0;
        lfz[i] = (((j >= 2) ? lrt[i - p.plane] : 0.0) - (in ? lrt[i] : 0.0)) / p.dz;
    }
}
void sdiv_tl(les_ctx* c, LinBufs& b, const double* t1, const double* t2) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_sdiv_tl, lgrid(p.plane * (p.nz + 1)), dim3(256), 0, c->stream, b.v[L_RT],
                       (const double*)b.v[L_ADV], t1, t2, (const double*)b.v[L_FZ], p);
}
void sdiv_adj(les_ctx* c, LinBufs& b, const double* t1, const double* t2) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_sdiv_adj, lgrid(p.plane * p.nz), dim3(256), 0, c->stream, b.v[L_RT], b.v[L_ADV], b.v[L_FX],
                       b.v[L_FY], b.v[L_FZ], t1, t2, p);
    clear_levels(c, b, b.v[L_RT], 1, p.nz - 1);
}

// ---- calc_beta (010_scalars.f90:214-258) ----------------------------------------------
__global__ void k_beta_tl(double* db, const double* __restrict__ th, const double* __restrict__ m,
                          const double* __restrict__ dth, const double* __restrict__ dm, LP p, double g) {
    const long n = p.plane * (p.nz + 1);
    LGRID(i, n) {
        const int k = (int)(i / p.plane);
        double r = 0.0;
        if (k >= 2 && k <= p.nz - 1) {
            const long d = i - p.plane;
            r = 0.5 * g * (dth[i] / m[k] - th[i] * dm[k] / (m[k] * m[k]) + dth[d] / m[k - 1] -
                           th[d] * dm[k - 1] / (m[k - 1] * m[k - 1]));
        }
        db[i] = r;
    }
}
// per level j (1..nz-1): sum over the plane of theta(j) (lb(j) + lb(j+1)), lb
// restricted to levels 2..nz-1
__global__ __launch_bounds__(256) void k_beta_adj_red(const double* __restrict__ th, const double* __restrict__ lb,
                                                       LP p, double* __restrict__ out) {
    __shared__ double sh[256];
    const int j = 1 + blockIdx.x;
    const bool a = (j >= 2 && j <= p.nz - 1), b = (j + 1 <= p.nz - 1);
    double s = 0.0;
    for (long q = threadIdx.x; q < p.plane; q += 256) {
        const long i = (long)j * p.plane + q;
        const double l = (a ? lb[i] : 0.0) + (b ? lb[i + p.plane] : 0.0);
        s += th[i] * l;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = sh[0];
}
__global__ void k_beta_adj_pt(double* lth, double* lb, const double* __restrict__ m, const double* __restrict__ red,
                              LP p, double g, double nn) {
    const long n = p.plane * (p.nz - 1);
    LGRID(t, n) {
        const long i = p.plane + t;
        const int j = (int)(i / p.plane);
        const double l = ((j >= 2) ? lb[i] : 0.0) + ((j + 1 <= p.nz - 1) ? lb[i + p.plane] : 0.0);
        lth[i] += 0.5 * g * l / m[j] - 0.5 * g * red[j] / (m[j] * m[j] * nn);
    }
}
void beta_tl(les_ctx* c, const Tape& t, LinBufs& b, double g) {
    const LP p = lin_params(c);
    level_means(c->stream, b.v[L_TH], c->plane, 0, (int)c->nlev, (double)(c->nx * c->ny), b.small);
    hipLaunchKernelGGL(k_beta_tl, lgrid(p.plane * (p.nz + 1)), dim3(256), 0, c->stream, b.v[L_BETA],
                       (const double*)t.thf, (const double*)t.tmean, (const double*)b.v[L_TH],
                       (const double*)b.small, p, g);
}
void beta_adj(les_ctx* c, const Tape& t, LinBufs& b, double g) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_beta_adj_red, dim3(p.nz - 1), dim3(256), 0, c->stream, (const double*)t.thf,
                       (const double*)b.v[L_BETA], p, b.small);
    hipLaunchKernelGGL(k_beta_adj_pt, lgrid(p.plane * (p.nz - 1)), dim3(256), 0, c->stream, b.v[L_TH],
                       b.v[L_BETA], (const double*)t.tmean, (const double*)b.small, p, g,
                       (double)(c->nx * c->ny));
}

// ---- AB2 updates, RHS assembly, projection ----------------------------------------------
// x(k) += dt (1.5 r(k) - 0.5 rf(k)), k = 1..nz-1 (adv: lr += 1.5 dt lx, lrf += -0.5 dt lx)
__global__ void k_ab2_tl(double* x, const double* __restrict__ r, const double* __restrict__ rf, LP p) {
    const long n = p.plane * (p.nz - 1);
    LGRID(t, n) {
        const long i = p.plane + t;
        x[i] = x[i] + p.dt * (1.5 * r[i] - 0.5 * rf[i]);
    }
}
__global__ void k_ab2_adj(const double* __restrict__ lx, double* lr, double* lrf, LP p) {
    const long n = p.plane * (p.nz - 1);
    LGRID(t, n) {
        const long i = p.plane + t;
        lr[i] += 1.5 * p.dt * lx[i];
        lrf[i] += -0.5 * p.dt * lx[i];
    }
}
void ab2_tl(les_ctx* c, double* x, const double* r, const double* rf) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_ab2_tl, lgrid(p.plane * (p.nz - 1)), dim3(256), 0, c->stream, x, r, rf, p);
}
void ab2_adj(les_ctx* c, const double* lx, double* lr, double* lrf) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_ab2_adj, lgrid(p.plane * (p.nz - 1)), dim3(256), 0, c->stream, lx, lr, lrf, p);
}

__global__ void k_rhs_tl(LinRhs a, LP p, RhsFlags f, const double* __restrict__ sponge) {
    const long n = p.plane * (p.nz + 1);
    LGRID(i, n) {
        const int k = (int)(i / p.plane);
        if (k == 0 || k == p.nz) { a.rx[i] = 0.0; a.ry[i] = 0.0; a.rz[i] = 0.0; continue; }
        double rx = -a.cx[i] - a.dx[i], ry = -a.cy[i] - a.dy[i], rz = -a.cz[i] - a.dz[i];
        if (f.buoy) rz = rz + a.beta[i];
        if (f.cori) {
            rx = rx + f.coriol * a.v[i];
            ry = ry - f.coriol * a.u[i];
        }
        if (f.damp) rz = rz - sponge[k] * a.w[i];
        a.rx[i] = rx; a.ry[i] = ry; a.rz[i] = rz;
    }
}
__global__ void k_rhs_adj(LinRhs a, LP p, RhsFlags f, const double* __restrict__ sponge) {
    const long n = p.plane * (p.nz - 1);
    LGRID(t, n) {
        const long i = p.plane + t;
        const int k = (int)(i / p.plane);
        const double lx = a.rx[i], ly = a.ry[i], lz = a.rz[i];
        // the RHS is the only reader of the convective term, DIVX, DIVY and
        // beta: their adjoints are written here, not accumulated
        a.cx[i] = -lx; a.dx[i] = -lx;
        a.cy[i] = -ly; a.dy[i] = -ly;
        a.cz[i] = -lz; a.dz[i] += -lz;
        a.beta[i] = f.buoy ? lz : 0.0;
        if (f.cori) {
            a.v[i] += f.coriol * lx;
            a.u[i] += -f.coriol * ly;
        }
        if (f.damp) a.w[i] += -sponge[k] * lz;
        a.rx[i] = 0.0; a.ry[i] = 0.0; a.rz[i] = 0.0;
    }
}
void rhs_tl(les_ctx* c, const LinRhs& a, const RhsFlags& f) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_rhs_tl, lgrid(p.plane * (p.nz + 1)), dim3(256), 0, c->stream, a, p, f,
                       (const double*)c->sponge);
}
void rhs_adj(les_ctx* c, const LinRhs& a, const RhsFlags& f) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_rhs_adj, lgrid(p.plane * (p.nz - 1)), dim3(256), 0, c->stream, a, p, f,
                       (const double*)c->sponge);
}

// r(k) -= dp(k), k = 1..nz-1 (adj: ldp -= lr)
__global__ void k_sub_tl(double* r, const double* __restrict__ d, LP p) {
    const long n = p.plane * (p.nz - 1);
    LGRID(t, n) {
        const long i = p.plane + t;
        r[i] = r[i] - d[i];
    }
}
__global__ void k_sub_adj(const double* __restrict__ lr, double* ld, LP p) {
    const long n = p.plane * (p.nz - 1);
    LGRID(t, n) {
        const long i = p.plane + t;
        ld[i] += -lr[i];
    }
}
void sub_tl(les_ctx* c, double* r, const double* d) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_sub_tl, lgrid(p.plane * (p.nz - 1)), dim3(256), 0, c->stream, r, d, p);
}
void sub_adj(les_ctx* c, const double* lr, double* ld) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_sub_adj, lgrid(p.plane * (p.nz - 1)), dim3(256), 0, c->stream, lr, ld, p);
}

// project (110_project.f90:16-62): u, v on 1..nz-1, w on 2..nz-1, then the
// stress-free top copy u(nz-1) = u(nz-2), w(1) = w(nz) = 0
__global__ void k_project_tl(double* u, double* v, double* w, const double* __restrict__ px,
                             const double* __restrict__ py, const double* __restrict__ pz, LP p) {
    const long n = p.plane * (p.nz - 1);
    LGRID(t, n) {
        const long i = p.plane + t;
        const int k = (int)(i / p.plane);
        u[i] = u[i] + p.dt * (-1.5 * px[i]);
        v[i] = v[i] + p.dt * (-1.5 * py[i]);
        if (k >= 2) w[i] = w[i] + p.dt * (-1.5 * pz[i]);
    }
}
__global__ void k_project_bc_tl(double* u, double* v, double* w, LP p, int sf) {
    LGRID(q, p.plane) {
        const long top = (long)p.nz * p.plane + q;
        if (sf) {
            u[top - p.plane] = u[top - 2 * p.plane];
            v[top - p.plane] = v[top - 2 * p.plane];
        }
        w[top] = 0.0;
        w[p.plane + q] = 0.0;
        u[q] = 0.0; v[q] = 0.0; w[q] = 0.0;
        u[top] = 0.0; v[top] = 0.0;
    }
}
__global__ void k_project_bc_adj(double* lu, double* lv, double* lw, LP p, int sf) {
    LGRID(q, p.plane) {
        const long top = (long)p.nz * p.plane + q;
        if (sf) {
            lu[top - 2 * p.plane] += lu[top - p.plane];
            lv[top - 2 * p.plane] += lv[top - p.plane];
            lu[top - p.plane] = 0.0;
            lv[top - p.plane] = 0.0;
        }
        lw[top] = 0.0;
        lw[p.plane + q] = 0.0;
        lu[q] = 0.0; lv[q] = 0.0; lw[q] = 0.0;
        lu[top] = 0.0; lv[top] = 0.0;
    }
}
__global__ void k_project_adj(const double* __restrict__ lu, const double* __restrict__ lv,
                              const double* __restrict__ lw, double* lpx, double* lpy, double* lpz, LP p) {
    const long n = p.plane * (p.nz - 1);
    LGRID(t, n) {
        const long i = p.plane + t;
        const int k = (int)(i / p.plane);
        lpx[i] += -1.5 * p.dt * lu[i];
        lpy[i] += -1.5 * p.dt * lv[i];
        if (k >= 2) lpz[i] += -1.5 * p.dt * lw[i];
    }
}
void project_tl(les_ctx* c, LinBufs& b, int sf) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_project_tl, lgrid(p.plane * (p.nz - 1)), dim3(256), 0, c->stream, b.v[L_U], b.v[L_V],
                       b.v[L_W], (const double*)b.v[L_DPX], (const double*)b.v[L_DPY], (const double*)b.v[L_DPZ], p);
    hipLaunchKernelGGL(k_project_bc_tl, lgrid(p.plane), dim3(256), 0, c->stream, b.v[L_U], b.v[L_V], b.v[L_W], p,
                       sf);
}
void project_adj(les_ctx* c, LinBufs& b, int sf) {
    const LP p = lin_params(c);
    hipLaunchKernelGGL(k_project_bc_adj, lgrid(p.plane), dim3(256), 0, c->stream, b.v[L_U], b.v[L_V], b.v[L_W], p,
                       sf);
    hipLaunchKernelGGL(k_project_adj, lgrid(p.plane * (p.nz - 1)), dim3(256), 0, c->stream, (const double*)b.v[L_U],
                       (const double*)b.v[L_V], (const double*)b.v[L_W], b.v[L_DPX], b.v[L_DPY], b.v[L_DPZ], p);
}

// ---- step_scalar (010_scalars.f90:624-671); theta(nz) is overwritten by the
// boundary ghosts of project, so only 1..nz-1 carry a tangent -----------------
void step_scalar_tl(les_ctx* c, LinBufs& b) { ab2_tl(c, b.v[L_TH], b.v[L_RT], b.v[L_RTF]); }
void step_scalar_adj(les_ctx* c, LinBufs& b) { ab2_adj(c, b.v[L_TH], b.v[L_RT], b.v[L_RTF]); }

// ---- pressure adjoint ---------------------------------------------------------------
// lp(j), j = 0..nz-1, from p itself, dp/dz = (p(k) - p(k-1))/dz (k =
// 1..nz-1), dp/dx = ddx p and dp/dy = ddy p (1..nz-1): lp = lP +
// dpdz^T(lpz) - tx - ty; tx = ddx(lpx), ty = ddy(lpy)
__global__ void k_press_adj_p(double* lp, const double* __restrict__ lP, const double* __restrict__ lpz,
                              const double* __restrict__ tx, const double* __restrict__ ty, LP p) {
    const long n = p.plane * (p.nz + 1);
    LGRID(i, n) {
        const int j = (int)(i / p.plane);
        double a = lP[i];
        if (j >= 1 && j <= p.nz - 1) a += lpz[i] / p.dz - tx[i] - ty[i];
        if (j + 1 >= 1 && j + 1 <= p.nz - 1) a -= lpz[i + p.plane] / p.dz;
        lp[i] = (j <= p.nz - 1) ? a : 0.0;
    }
}

constexpr int LTD = 8;  // levels per load batch of the transposed TDMA

// Transposed TDMA on the spectrum G = F(lp) (levels 0..nz-1): lambda of Hx,
// Hy (levels 1..nz-1; 0 and nz zero) and Hz (0..nz).  The wave-mode matrix
// is symmetric, so its factorisation solves the transposed system.
__global__ __launch_bounds__(128) void k_tdma_adj(const double* __restrict__ G, double* lHx, double* lHy,
                                                   double* lHz, double* dsc, const double* __restrict__ kx2d,
                                                   const double* __restrict__ ky2d, const double* __restrict__ k2a,
                                                   const unsigned char* __restrict__ nn,
                                                   const double* __restrict__ cb, const double* __restrict__ cc,
                                                   long ns, int nz, double dz) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * ns) return;
    const long q = t >> 1;
    const int comp = (int)(t & 1);
    const long s2 = 2 * ns;
    const long oxy = t ^ 1;
    // levels 0 and nz of Hx, Hy: each lane clears its own component; levels
    // 1..nz-1 of component c are written by the lane of component c^1
    lHx[t] = 0.0; lHy[t] = 0.0;
    lHx[nz * s2 + t] = 0.0; lHy[nz * s2 + t] = 0.0;
    if (!nn[q]) {
        for (int k = 0; k <= nz; ++k) { lHx[k * s2 + t] = 0.0; lHy[k * s2 + t] = 0.0; lHz[k * s2 + t] = 0.0; }
        return;
    }
    const double k2 = k2a[q], kx = kx2d[q], ky = ky2d[q];
    const double g0 = G[t];
    if (k2 > 0.0) {
        const double idz = 1.0 / dz;
        const double sgn = comp ? 1.0 : -1.0;
        // loads issued LTD levels ahead of each sweep's dependency chain
        double dprev = 0.0;
        for (int k0 = 1; k0 <= nz - 1; k0 += LTD) {
            double g[LTD], bb[LTD];
#pragma unroll
            for (int d = 0; d < LTD; ++d) {
                const int k = k0 + d;
                if (k <= nz - 1) {
                    g[d] = G[k * s2 + t];
                    bb[d] = cb[k * ns + q];
                }
            }
#pragma unroll
            for (int d = 0; d < LTD; ++d) {
                const int k = k0 + d;
                if (k <= nz - 1) {
                    const double a = (k == 1) ? 0.0 : idz * idz;
                    const double gk = g[d] + ((k == 1) ? g0 : 0.0);
                    const double dv = (gk - a * dprev) / bb[d];
                    dsc[k * s2 + t] = dv;
                    dprev = dv;
                }
            }
        }
        double rn = dprev;          // lambda_r at level nz-1
        lHz[nz * s2 + t] = rn * idz;
        lHx[(nz - 1) * s2 + oxy] = sgn * kx * rn;
        lHy[(nz - 1) * s2 + oxy] = sgn * ky * rn;
        for (int k0 = nz - 2; k0 >= 1; k0 -= LTD) {
            double dv[LTD], cv[LTD];
#pragma unroll
            for (int d = 0; d < LTD; ++d) {
                const int k = k0 - d;
                if (k >= 1) {
                    dv[d] = dsc[k * s2 + t];
                    cv[d] = cc[k * ns + q];
                }
            }
#pragma unroll
            for (int d = 0; d < LTD; ++d) {
                const int k = k0 - d;
                if (k >= 1) {
                    const double r = dv[d] - cv[d] * rn;
                    lHz[(k + 1) * s2 + t] = (r - rn) * idz;
                    lHx[k * s2 + oxy] = sgn * kx * r;
                    lHy[k * s2 + oxy] = sgn * ky * r;
                    rn = r;
                }
            }
        }
        lHz[s2 + t] = -rn * idz;                  // level 1: -lambda_r(1)/dz
        lHz[t] = rn * idz - dz * g0;              // level 0: lambda_r(1)/dz - dz g0
    } else {
        // mean mode: p(k) = dz sum_{j=2..k} Hz(j), p(0) = -dz Hz(0)
        double acc = 0.0;
        lHz[nz * s2 + t] = 0.0;
        for (int k = nz - 1; k >= 2; --k) {
            acc += G[k * s2 + t];
            lHz[k * s2 + t] = dz * acc;
        }
        lHz[s2 + t] = 0.0;
        lHz[t] = -dz * g0;
        for (int k = 1; k <= nz - 1; ++k) { lHx[k * s2 + oxy] = 0.0; lHy[k * s2 + oxy] = 0.0; }
    }
}

// lu(k) += s lrx(k), lv, lw likewise on 1..nz-1; ldivtz(1) += cst lrz(0),
// ldivtz(nz) += cst lrz(nz)
__global__ void k_press_adj_out(const double* __restrict__ lrx, const double* __restrict__ lry,
                                const double* __restrict__ lrz, double* lu, double* lv, double* lw, double* ldz,
                                LP p, double s, double cst) {
    const long n = p.plane * (p.nz + 1);
    LGRID(i, n) {
        const int k = (int)(i / p.plane);
        if (k >= 1 && k <= p.nz - 1) {
            lu[i] += s * lrx[i];
            lv[i] += s * lry[i];
            lw[i] += s * lrz[i];
        } else if (k == 0) {
            ldz[i + p.plane] += cst * lrz[i];
        } else {
            ldz[i] += cst * lrz[i];
        }
    }
}

void press_adj(les_ctx* c, LinBufs& b) {
    const LP p = lin_params(c);
    const int nz = c->nz;
    const long ns = (long)c->ns;
    double* tx = c->work3[0];
    double* ty = c->work3[1];
    double* lp = c->work3[2];
    ddx_f<double, double>(c, tx, b.v[L_DPX], 0, (int)c->nlev);
    ddy_f<double, double>(c, ty, b.v[L_DPY], 0, (int)c->nlev);
    hipLaunchKernelGGL(k_press_adj_p, lgrid(p.plane * (nz + 1)), dim3(256), 0, c->stream, lp,
                       (const double*)b.v[L_P], (const double*)b.v[L_DPZ], (const double*)tx, (const double*)ty, p);
    cx<double>* G = c->spec3[3];
    x_r2c<double>(c->stream, lp, G, (long)nz * c->ny, c->nx, c->lhx, c->tw_x, 1.0);
    yline_spec<double>(c->stream, G, G, nz, c->lhx, ns, c->tw_y, SP_FWD, nullptr);
    cx<double>*Hx = c->spec3[0], *Hy = c->spec3[1], *Hz = c->spec3[2];
    double* dsc = b.dsc;
    const int nb = (int)((2 * ns + 127) / 128);
    hipLaunchKernelGGL(k_tdma_adj, dim3(nb), dim3(128), 0, c->stream, reinterpret_cast<const double*>(G),
                       reinterpret_cast<double*>(Hx), reinterpret_cast<double*>(Hy), reinterpret_cast<double*>(Hz),
                       dsc, (const double*)c->kx2d, (const double*)c->ky2d, (const double*)c->k2,
                       (const unsigned char*)c->notnyq, (const double*)c->coef_b, (const double*)c->coef_c, ns, nz,
                       c->dz);
    double *rx = c->work3[3], *ry = c->work3[4], *rz = c->work3[5];
    cx<double>* H[3] = {Hx, Hy, Hz};
    double* R[3] = {rx, ry, rz};
    for (int m = 0; m < 3; ++m) {
        yline_spec<double>(c->stream, H[m], H[m], nz + 1, c->lhx, ns, c->tw_y, SP_INV, nullptr);
        x_c2r<double>(c->stream, H[m], R[m], (long)(nz + 1) * c->ny, c->lhx, c->nx, c->tw_x, 0, 0.0);
    }
    const double cst = 1.0 / (double)(c->nx * c->ny);
    hipLaunchKernelGGL(k_press_adj_out, lgrid(p.plane * (nz + 1)), dim3(256), 0, c->stream, (const double*)rx,
                       (const double*)ry, (const double*)rz, b.v[L_U], b.v[L_V], b.v[L_W], b.v[L_DVZ], p,
                       cst / (1.5 * p.dt), cst);
    clear_levels(c, b, b.v[L_DPX], 1, nz - 1);
    clear_levels(c, b, b.v[L_DPY], 1, nz - 1);
    clear_levels(c, b, b.v[L_DPZ], 1, nz - 1);
    clear_levels(c, b, b.v[L_P], 0, nz - 1);
}

}  // namespace les
