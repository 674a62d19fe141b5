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
// The column walk is a dependency chain, so loads are issued D levels at a
// time ahead of the arithmetic (one memory latency per D levels).
//
// The right-hand sides H may be FP32 (mixed precision); the recurrences, the
// forward-sweep store d' and the mean mode are FP64 in both precisions.  With
// hzm != null the mean mode reads its z right-hand side from hzm (FP64 plane
// sums, level 0..nz) instead of the transform's (0,0) coefficient.
#include "les_internal.hpp"

namespace les {

constexpr int TD = 8;  // levels per load batch

// dsc holds d' (FP64); out receives p.  For H of type double both may be the
// same buffer: each element is read by its own lane before it is overwritten.
template <typename TH>
__global__ __launch_bounds__(128) void k_tdma(const TH* __restrict__ Hx, const TH* __restrict__ Hy,
                                               const TH* __restrict__ Hz, double* dsc, TH* out,
                                               const double* __restrict__ hzm,
                                               const double* __restrict__ kx2d, const double* __restrict__ ky2d,
                                               const double* __restrict__ k2a, const unsigned char* __restrict__ nn,
                                               const double* __restrict__ cb, const double* __restrict__ cc,
                                               long ns, int nz, double dz) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * ns) return;
    const long q = t >> 1;
    const int comp = (int)(t & 1);     // 0: real part, 1: imaginary part
    const long s2 = 2 * ns;            // level stride in doubles
    if (!nn[q]) {
        for (int k = 0; k <= nz - 1; ++k) out[k * s2 + t] = (TH)0.0;
        return;
    }
    const double k2 = k2a[q], kx = kx2d[q], ky = ky2d[q];
    const bool wave = k2 > 0.0;
    const double idz = 1.0 / dz;
    // i (kx Hx + ky Hy): real part -(kx Hx.im + ky Hy.im), imag part kx Hx.re + ky Hy.re
    const long oxy = t ^ 1;            // the other component of the same element
    const double sgn = comp ? 1.0 : -1.0;
    const bool m64 = (hzm != nullptr) && (t == 0);
    const double h0 = m64 ? hzm[0] : (double)Hz[t];
    double hzk = m64 ? hzm[1] : (double)Hz[s2 + t];
    double dprev = h0;
    // ---- forward sweep ----
    for (int k0 = 1; k0 <= nz - 1; k0 += TD) {
        double hx[TD], hy[TD], hzn[TD], b[TD];
#pragma unroll
        for (int d = 0; d < TD; ++d) {
            const int k = k0 + d;
            if (k <= nz - 1) {
                const long o = k * s2;
                hx[d] = Hx[o + oxy];
                hy[d] = Hy[o + oxy];
                hzn[d] = m64 ? hzm[k + 1] : (double)Hz[o + s2 + t];
                b[d] = cb[k * ns + q];
            }
        }
#pragma unroll
        for (int d = 0; d < TD; ++d) {
            const int k = k0 + d;
            if (k <= nz - 1) {
                double r, a;
                if (wave) {
                    double dh = hzn[d] - hzk;
                    if (k == 1) dh += h0;
                    r = sgn * (kx * hx[d] + ky * hy[d]) + dh * idz;
                    a = (k == 1) ? 0.0 : idz * idz;
                } else {
                    r = (k == 1) ? 0.0 : hzk;
                    a = (k == 1) ? 0.0 : -idz;
                }
                const double dv = (r - a * dprev) / b[d];
                dsc[k * s2 + t] = dv;
                dprev = dv;
                hzk = hzn[d];
            }
        }
    }
    // ---- backward sweep ----
    double pn = dprev;  // level nz-1 (c'(nz-1) = 0)
    if (nz - 1 >= 1) out[(nz - 1) * s2 + t] = (TH)pn;
    for (int k0 = nz - 2; k0 >= 1; k0 -= TD) {
        double dv[TD], c[TD];
#pragma unroll
        for (int d = 0; d < TD; ++d) {
            const int k = k0 - d;
            if (k >= 1) {
                dv[d] = dsc[k * s2 + t];
                c[d] = cc[k * ns + q];
            }
        }
#pragma unroll
        for (int d = 0; d < TD; ++d) {
            const int k = k0 - d;
            if (k >= 1) {
                pn = dv[d] - c[d] * pn;
                out[k * s2 + t] = (TH)pn;
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
    hipLaunchKernelGGL(k_tdma<T>, dim3(nb), dim3(128), 0, c->stream, reinterpret_cast<const T*>(Hx),
                       reinterpret_cast<const T*>(Hy), reinterpret_cast<const T*>(Hz), dsc,
                       reinterpret_cast<T*>(ph), hzm, c->kx2d, c->ky2d, c->k2, c->notnyq, c->coef_b, c->coef_c, ns,
                       c->nz, c->dz);
}
template void tdma_press<double>(les_ctx*, const cx<double>*, const cx<double>*, const cx<double>*, cx<double>*,
                                 double*, const double*);
template void tdma_press<float>(les_ctx*, const cx<float>*, const cx<float>*, const cx<float>*, cx<float>*,
                                double*, const double*);

}  // namespace les
