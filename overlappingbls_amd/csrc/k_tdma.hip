// Pressure Poisson solve: one Thomas (TDMA) system per (kx, ky) of the half
// spectrum, one lane per system, wavenumber-major so that at every level the
// lanes of a wavefront touch consecutive complex values.
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
#include "les_internal.hpp"

namespace les {

__global__ __launch_bounds__(256) void k_tdma(const cx<double>* __restrict__ Hx, const cx<double>* __restrict__ Hy,
                                               const cx<double>* __restrict__ Hz, cx<double>* __restrict__ ph,
                                               const double* __restrict__ kx2d, const double* __restrict__ ky2d,
                                               const double* __restrict__ k2a, const unsigned char* __restrict__ nn,
                                               const double* __restrict__ cb, const double* __restrict__ cc,
                                               long ns, int nz, double dz) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= ns) return;
    if (!nn[q]) {
        for (int k = 0; k <= nz - 1; ++k) ph[k * ns + q] = {0.0, 0.0};
        return;
    }
    const double k2 = k2a[q], kx = kx2d[q], ky = ky2d[q];
    const bool wave = k2 > 0.0;
    const double idz2 = 1.0 / (dz * dz);
    const cx<double> h0 = Hz[q];
    cx<double> hzk = Hz[ns + q];
    cx<double> dprev = h0;
    // forward sweep, k = 1 .. nz-1; d' stored in ph
    for (int k = 1; k <= nz - 1; ++k) {
        const long o = k * ns + q;
        const cx<double> hzp = Hz[o + ns];
        cx<double> r;
        double a;
        if (wave) {
            const cx<double> hx = Hx[o], hy = Hy[o];
            // i (kx Hx + ky Hy) + (Hz(k+1) - Hz(k) [+ Hz(0) at the wall]) / dz
            const double sr = kx * hx.x + ky * hy.x, si = kx * hx.y + ky * hy.y;
            cx<double> dh = {hzp.x - hzk.x, hzp.y - hzk.y};
            if (k == 1) { dh.x += h0.x; dh.y += h0.y; }
            r = {-si + dh.x / dz, sr + dh.y / dz};
            a = (k == 1) ? 0.0 : idz2;
        } else {
            r = (k == 1) ? cx<double>{0.0, 0.0} : hzk;
            a = (k == 1) ? 0.0 : -1.0 / dz;
        }
        const double b = cb[o];
        const cx<double> d = {(r.x - a * dprev.x) / b, (r.y - a * dprev.y) / b};
        ph[o] = d;
        dprev = d;
        hzk = hzp;
    }
    // backward sweep
    cx<double> pn = dprev;  // level nz-1 (c'(nz-1) = 0)
    for (int k = nz - 2; k >= 1; --k) {
        const long o = k * ns + q;
        const cx<double> d = ph[o];
        const double c = cc[o];
        pn = {d.x - c * pn.x, d.y - c * pn.y};
        ph[o] = pn;
    }
    ph[q] = {pn.x - dz * h0.x, pn.y - dz * h0.y};
}

void tdma_press(les_ctx* c, const cx<double>* Hx, const cx<double>* Hy, const cx<double>* Hz, cx<double>* ph) {
    const long ns = (long)c->ns;
    const int nb = (int)((ns + 255) / 256);
    hipLaunchKernelGGL(k_tdma, dim3(nb), dim3(256), 0, c->stream, Hx, Hy, Hz, ph, c->kx2d, c->ky2d, c->k2,
                       c->notnyq, c->coef_b, c->coef_c, ns, c->nz, c->dz);
}

}  // namespace les
