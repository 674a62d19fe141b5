// Pressure Poisson solve: one Thomas (TDMA) system per (kx, ky) of the half
// spectrum.  The coefficients are real, so the real and imaginary parts of a
// system are independent: lane 2q+c solves component c of wavenumber q.
// Lanes are wavenumber-major, so at every level the lanes of a wavefront
// touch one contiguous run of complex values.
//
// Restates press_stag_array's right-hand side (109_press_stag_array.f90:201-225)
// and tdmaSolver_pipelined (:244-319) for the whole column at once: the
// reference pipelines the same recurrences over nyt stages of blocking
// messages between z-slabs; here each lane walks its column from the wall
// (level 1) to the top (level nzt-1) and back.
//   forward  d'_k = (r_k - a_k d'_{k-1}) / b'_k,   d'_0 = H_z(0)
//   backward p_{nzt-1} = d'_{nzt-1};  p_k = d'_k - c'_k p_{k+1}
//   surface  p_0 = p_1 - dz d'_0
// The mean mode (k^2 = 0, not Nyquist) has a = -1/dz, b = 1/dz, c = 0 above
// the wall, i.e. p_1 = 0, p_k = p_{k-1} + dz H_z(k).  Nyquist modes are zero.
//
// The column walk is a dependency chain, so loads are issued a batch of
// levels ahead of the arithmetic.
//
// The right-hand sides H may be FP32 (mixed precision); the recurrences, the
// forward-sweep store d' and the mean mode are FP64 in both precisions.  With
// hzm != null the mean mode reads its z right-hand side from hzm (FP64 plane
// sums, level 0..nz) instead of the transform's (0,0) coefficient.
#include "les_internal.hpp"

namespace les {

constexpr int TDF = 24;  // forward-sweep levels per load batch
constexpr int TDB = 32;  // backward-sweep levels per load batch
constexpr int TDL = 48;  // top levels whose d' stays in LDS (48 KB per 128-lane block:
                         // three blocks fit a CU, so 2 ns lanes still run in one round)

// Coefficients of one system (109_press_stag_array.f90:4-103 as
// init_spectral forms them) and its right-hand-side factors.
struct TdmaCoef {
    int kt;
    bool wave;
    double sgn, kx, ky, idz, tab1, b_edge, b_int, m_a, m_b, h0;
};

// Level k of the forward recurrence: b' = b - a c'_{k-1}, c' = c / b'
// (init_spectral's factorisation, bitwise), d' = (r - a d'_{k-1}) / b'.
// Returns d'_k; the carried values are scalars of the caller.
__device__ __forceinline__ double tdma_level(const TdmaCoef C, int k, double hxv, double hyv, double hzv,
                                             double& hzk, double& dprev, double& cprev) {
    const bool first = (k == 1), top = (k == C.kt);
    const double at = first ? 0.0 : (C.wave ? C.tab1 : C.m_a);
    const double bt = C.wave ? ((first || top) ? C.b_edge : C.b_int) : C.m_b;
    const double ct = (C.wave && (first || !top)) ? C.tab1 : 0.0;
    const double bp = bt - at * cprev;
    cprev = ct / bp;
    const double dh = first ? (hzv - hzk) + C.h0 : hzv - hzk;
    const double rw = C.sgn * (C.kx * hxv + C.ky * hyv) + dh * C.idz;
    const double r = C.wave ? rw : (first ? 0.0 : hzk);
    const double a = first ? 0.0 : (C.wave ? C.idz * C.idz : -C.idz);
    const double dv = (r - a * dprev) / bp;
    dprev = dv;
    hzk = hzv;
    return dv;
}

// dsc holds d' (FP64); out receives p.  For H of type double both may be the
// same buffer: each element is read by its own lane before it is overwritten.
//
// With one lane per system there are only 2 ns lanes (about four waves per
// CU at 256^2), so a sweep waits on one memory latency per load batch: the
// batches are long, and every load of a batch is unconditional (level
// indices clamped to the column) so that no branch splits them.  The
// factorised diagonal b' is formed in the forward sweep from the same
// operations as init_spectral's table (bitwise equal), which saves reading
// that table; c' is read from its table in the backward sweep.
template <typename TH>
__global__ __launch_bounds__(128) void k_tdma(const TH* __restrict__ Hx, const TH* __restrict__ Hy,
                                               const TH* __restrict__ Hz, double* dsc, TH* out,
                                               const double* __restrict__ hzm,
                                               const double* __restrict__ kx2d, const double* __restrict__ ky2d,
                                               const double* __restrict__ k2a, const unsigned char* __restrict__ nn,
                                               const double* __restrict__ cc, long ns, int nz, double dz) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * ns) return;
    const long q = t >> 1;
    const int comp = (int)(t & 1);     // 0: real part, 1: imaginary part
    const long s2 = 2 * ns;            // level stride in doubles
    if (!nn[q]) {
        for (int k = 0; k <= nz - 1; ++k) out[k * s2 + t] = (TH)0.0;
        return;
    }
    const int kt = nz - 1;             // top row of the system
    // d' of levels kl..kt stays in LDS (the backward sweep reads it first)
    __shared__ double dl[TDL][128];
    const int kl = max(kt - TDL + 1, 1);
    const double k2 = k2a[q], kx = kx2d[q], ky = ky2d[q];
    const bool wave = k2 > 0.0;
    const double idz = 1.0 / dz;
    // coefficients as init_spectral forms them (109_press_stag_array.f90:4-103)
    const double tab1 = 1.0 / (dz * dz), tab2 = 2.0 / (dz * dz);
    const double b_edge = -(k2 + tab1), b_int = -(k2 + tab2);
    const double m_a = -1.0 / dz, m_b = 1.0 / dz;
    // i (kx Hx + ky Hy): real part -(kx Hx.im + ky Hy.im), imag part kx Hx.re + ky Hy.re
    const long oxy = t ^ 1;            // the other component of the same element
    const double sgn = comp ? 1.0 : -1.0;
    const bool m64 = (hzm != nullptr) && (t == 0);
    const double h0 = m64 ? hzm[0] : (double)Hz[t];
    double hzk = m64 ? hzm[1] : (double)Hz[s2 + t];
    double dprev = h0, cprev = 0.0;
    const TdmaCoef C{kt, wave, sgn, kx, ky, idz, tab1, b_edge, b_int, m_a, m_b, h0};
    // ---- forward sweep: full batches, then one partial batch ----
    int k0 = 1;
    for (; k0 + TDF - 1 <= kt; k0 += TDF) {
        TH hx[TDF], hy[TDF];
        double hz[TDF];
#pragma unroll
        for (int d = 0; d < TDF; ++d) {
            const long o = (long)(k0 + d) * s2;
            hx[d] = Hx[o + oxy];
            hy[d] = Hy[o + oxy];
            hz[d] = (double)Hz[o + s2 + t];
        }
        if (m64) {  // the FP64 mean-mode column (one lane)
#pragma unroll
            for (int d = 0; d < TDF; ++d) hz[d] = hzm[k0 + d + 1];
        }
#pragma unroll
        for (int d = 0; d < TDF; ++d) {
            const int k = k0 + d;
            const double v = tdma_level(C, k, (double)hx[d], (double)hy[d], hz[d], hzk, dprev, cprev);
            if (k >= kl) dl[k - kl][threadIdx.x] = v;
            else dsc[k * s2 + t] = v;
        }
    }
    if (k0 <= kt) {
        TH hx[TDF], hy[TDF];
        double hz[TDF];
#pragma unroll
        for (int d = 0; d < TDF; ++d) {
            const long o = (long)min(k0 + d, kt) * s2;
            hx[d] = Hx[o + oxy];
            hy[d] = Hy[o + oxy];
            hz[d] = (double)Hz[o + s2 + t];
        }
        if (m64) {
#pragma unroll
            for (int d = 0; d < TDF; ++d) hz[d] = hzm[min(k0 + d, kt) + 1];
        }
#pragma unroll
        for (int d = 0; d < TDF; ++d) {
            const int k = k0 + d;
            if (k <= kt) {
                const double v = tdma_level(C, k, (double)hx[d], (double)hy[d], hz[d], hzk, dprev, cprev);
                if (k >= kl) dl[k - kl][threadIdx.x] = v;
                else dsc[k * s2 + t] = v;
            }
        }
    }
    // ---- backward sweep ----
    double pn = dprev;  // level nz-1 (c'(nz-1) = 0)
    if (kt >= 1) out[kt * s2 + t] = (TH)pn;
    k0 = kt - 1;
    for (; k0 - TDB + 1 >= 1; k0 -= TDB) {
        double dv[TDB], c[TDB];
#pragma unroll
        for (int d = 0; d < TDB; ++d) {
            const int k = k0 - d;
            if (k >= kl) dv[d] = dl[k - kl][threadIdx.x];
            else dv[d] = dsc[k * s2 + t];
            c[d] = cc[k * ns + q];
        }
#pragma unroll
        for (int d = 0; d < TDB; ++d) {
            pn = dv[d] - c[d] * pn;
            out[(k0 - d) * s2 + t] = (TH)pn;
        }
    }
    if (k0 >= 1) {
        double dv[TDB], c[TDB];
#pragma unroll
        for (int d = 0; d < TDB; ++d) {
            const int k = max(k0 - d, 1);
            if (k >= kl) dv[d] = dl[k - kl][threadIdx.x];
            else dv[d] = dsc[k * s2 + t];
            c[d] = cc[k * ns + q];
        }
#pragma unroll
        for (int d = 0; d < TDB; ++d) {
            if (k0 - d >= 1) {
                pn = dv[d] - c[d] * pn;
                out[(k0 - d) * s2 + t] = (TH)pn;
            }
        }
    }
    out[t] = (TH)(pn - dz * h0);
}

template <typename T>
void tdma_press(les_ctx* c, const cx<T>* Hx, const cx<T>* Hy, const cx<T>* Hz, cx<T>* ph, double* dsc,
                const double* hzm) {
    const long ns = (long)c->ns;
    const int nb = (int)((2 * ns + 127) / 128);
    // Hx, Hy (levels 1..nz-1) and Hz (0..nz) read, p_hat (0..nz-1) written
    FamScope fs(c->stream, FAM_TDMA,
                (double)ns * 2.0 * sizeof(T) * (double)(2 * (c->nz - 1) + (c->nz + 1) + c->nz) +
                    (hzm ? 8.0 * (c->nz + 1) : 0.0));
    hipLaunchKernelGGL(k_tdma<T>, dim3(nb), dim3(128), 0, c->stream, reinterpret_cast<const T*>(Hx),
                       reinterpret_cast<const T*>(Hy), reinterpret_cast<const T*>(Hz), dsc,
                       reinterpret_cast<T*>(ph), hzm, c->kx2d, c->ky2d, c->k2, c->notnyq, c->coef_c, ns,
                       c->nz, c->dz);
}
template void tdma_press<double>(les_ctx*, const cx<double>*, const cx<double>*, const cx<double>*, cx<double>*,
                                 double*, const double*);
template void tdma_press<float>(les_ctx*, const cx<float>*, const cx<float>*, const cx<float>*, cx<float>*,
                                double*, const double*);

}  // namespace les
