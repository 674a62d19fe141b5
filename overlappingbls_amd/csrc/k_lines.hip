// Line-transform kernels: every spectral operation of the step is a
// separable pass along x (contiguous rows) or y (strided columns), staged in
// LDS.  Real lines are packed two to a complex line (row/column pairs), which
// is exact for any real-linear Fourier multiplier m(k) with m(-k) = conj m(k):
// derivative i k, Nyquist removal, 3/2 zero-padding and its inverse.
//
// Reference operations restated (all unnormalised FFTW conventions):
//   ddx / ddy               004_derivatives.f90:18-81
//   filter_data             006_test_filter.f90:82-105
//   zero_padding / unpad    108_01_padd.f90:73-197
//   fftn_mpi / ifftn_mpi    203_fftn_mpi.f90:4-67  (x r2c + y c2c, half spectrum)
//   test_filter             006_test_filter.f90:19-37
#include "les_internal.hpp"

namespace les {

constexpr int NT = 256;
constexpr int EPT = 8;

template <typename T>
__device__ __forceinline__ cx<T> line_op(const cx<T>* lin, int k, int nin, int nout, int op, T kfac) {
    switch (op) {
        case OP_DERIV: {
            const int n = nin;
            if (2 * k == n) return {T(0), T(0)};
            const int m = (2 * k < n) ? k : k - n;
            const T f = (T)m * kfac / (T)n;
            const cx<T> z = lin[k];
            return {-f * z.y, f * z.x};
        }
        case OP_NYQ: {
            const int n = nin;
            if (2 * k == n) return {T(0), T(0)};
            const T s = T(1) / (T)n;
            const cx<T> z = lin[k];
            return {z.x * s, z.y * s};
        }
        case OP_PAD: {  // nin = n, nout = 3n/2
            const int h = nin / 2;
            int src = -1;
            if (k < h) src = k;
            else if (k > nout - h) src = k - nout + nin;
            if (src < 0) return {T(0), T(0)};
            const T s = T(1) / (T)nin;
            const cx<T> z = lin[src];
            return {z.x * s, z.y * s};
        }
        default: {  // OP_UNPAD: nin = 3n/2, nout = n
            const int h = nout / 2;
            int src = -1;
            if (k < h) src = k;
            else if (k > h) src = k - nout + nin;
            if (src < 0) return {T(0), T(0)};
            const T s = T(1) / (T)nin;
            const cx<T> z = lin[src];
            return {z.x * s, z.y * s};
        }
    }
}

// Rewrite P lines of length nin (stride nin) into P lines of length nout
// (stride nout) through registers, applying the spectral op.
template <typename T>
__device__ __forceinline__ void apply_line_op(cx<T>* buf, int P, int nin, int nout, int op, T kfac) {
    cx<T> r[EPT];
    const int tot = P * nout;
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
        const int e = threadIdx.x + m * NT;
        if (e < tot) {
            const int line = e / nout, k = e - line * nout;
            r[m] = line_op(buf + line * nin, k, nin, nout, op, kfac);
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
        const int e = threadIdx.x + m * NT;
        if (e < tot) buf[e] = r[m];
    }
    __syncthreads();
}

template <typename T, typename TI>
__global__ __launch_bounds__(NT) void k_xline_real(const TI* __restrict__ in, T* __restrict__ out, long nrows,
                                                   int ld_in, int ld_out, FftPlan pin, FftPlan pout,
                                                   const cx<T>* __restrict__ twin_g,
                                                   const cx<T>* __restrict__ twout_g, int op, T kfac, int P) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    const int nin = pin.n, nout = pout.n;
    const int L = nin > nout ? nin : nout;
    // derivative in a lower precision than the input: subtract the line's
    // first element before the conversion (see SHIFT_IN of the specialised kernels)
    const bool shift = (op == OP_DERIV) && (sizeof(TI) > sizeof(T));
    cx<T>* twin = buf + P * L;
    cx<T>* twout = twin + nin;
    load_tw(twin, twin_g, nin);
    load_tw(twout, twout_g, nout);
    const long row0 = (long)blockIdx.x * 2 * P;
    for (int e = threadIdx.x; e < 2 * P * nin; e += NT) {
        const int rl = e / nin, col = e - rl * nin;
        const long gr = row0 + rl;
        T v = T(0);
        if (gr < nrows)
            v = shift ? (T)(in[gr * ld_in + col] - in[gr * ld_in]) : (T)in[gr * ld_in + col];
        reinterpret_cast<T*>(&buf[(rl >> 1) * nin + col])[rl & 1] = v;
    }
    __syncthreads();
    fft_lines<T, EPT>(buf, pin, P, nin, twin, -1);
    apply_line_op<T>(buf, P, nin, nout, op, kfac);
    fft_lines<T, EPT>(buf, pout, P, nout, twout, +1);
    for (int e = threadIdx.x; e < 2 * P * nout; e += NT) {
        const int rl = e / nout, col = e - rl * nout;
        const long gr = row0 + rl;
        if (gr < nrows) out[gr * ld_out + col] = reinterpret_cast<const T*>(&buf[(rl >> 1) * nout + col])[rl & 1];
    }
}

template <typename T, typename TI>
__global__ __launch_bounds__(NT) void k_yline_real(const TI* __restrict__ in, T* __restrict__ out, int ncols,
                                                   int ld_in, int ld_out, long ps_in, long ps_out, FftPlan pin,
                                                   FftPlan pout, const cx<T>* __restrict__ twin_g,
                                                   const cx<T>* __restrict__ twout_g, int op, T kfac, int C) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    const int nin = pin.n, nout = pout.n;
    const int L = nin > nout ? nin : nout;
    // derivative in a lower precision than the input: subtract the line's
    // first element before the conversion (see SHIFT_IN of the specialised kernels)
    const bool shift = (op == OP_DERIV) && (sizeof(TI) > sizeof(T));
    cx<T>* twin = buf + C * L;
    cx<T>* twout = twin + nin;
    load_tw(twin, twin_g, nin);
    load_tw(twout, twout_g, nout);
    const int x0 = blockIdx.x * 2 * C;
    const TI* src = in + (long)blockIdx.y * ps_in;
    T* dst = out + (long)blockIdx.y * ps_out;
    const int w2 = 2 * C;
    for (int e = threadIdx.x; e < nin * w2; e += NT) {
        const int y = e / w2, cl = e - y * w2;
        const int x = x0 + cl;
        T v = T(0);
        if (x < ncols) v = shift ? (T)(src[(long)y * ld_in + x] - src[x]) : (T)src[(long)y * ld_in + x];
        reinterpret_cast<T*>(&buf[(cl >> 1) * nin + y])[cl & 1] = v;
    }
    __syncthreads();
    fft_lines<T, EPT>(buf, pin, C, nin, twin, -1);
    apply_line_op<T>(buf, C, nin, nout, op, kfac);
    fft_lines<T, EPT>(buf, pout, C, nout, twout, +1);
    // pad / unpad with kfac != 0 accumulate: out += kfac * result
    const bool acc = (op == OP_PAD || op == OP_UNPAD) && kfac != T(0);
    for (int e = threadIdx.x; e < nout * w2; e += NT) {
        const int y = e / w2, cl = e - y * w2;
        const int x = x0 + cl;
        if (x >= ncols) continue;
        const T v = reinterpret_cast<const T*>(&buf[(cl >> 1) * nout + y])[cl & 1];
        T* o = dst + (long)y * ld_out + x;
        *o = acc ? *o + kfac * v : v;
    }
}

// x r2c of fftn_mpi (203_fftn_mpi.f90:21-25): rows -> half spectrum, x
// Nyquist zeroed, times `scale`.
template <typename T, typename TI>
__global__ __launch_bounds__(NT) void k_x_r2c(const TI* __restrict__ in, cx<T>* __restrict__ out, long nrows,
                                              int ld_in, int ldc, FftPlan pl, const cx<T>* __restrict__ tw_g,
                                              T scale, int P, int kst) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    const int n = pl.n, lh = kst;
    cx<T>* tw = buf + P * n;
    load_tw(tw, tw_g, n);
    const long row0 = (long)blockIdx.x * 2 * P;
    for (int e = threadIdx.x; e < 2 * P * n; e += NT) {
        const int rl = e / n, col = e - rl * n;
        const long gr = row0 + rl;
        const T v = (gr < nrows) ? (T)in[gr * ld_in + col] : T(0);
        reinterpret_cast<T*>(&buf[(rl >> 1) * n + col])[rl & 1] = v;
    }
    __syncthreads();
    fft_lines<T, EPT>(buf, pl, P, n, tw, -1);
    for (int e = threadIdx.x; e < 2 * P * lh; e += NT) {
        const int rl = e / lh, k = e - rl * lh;
        const long gr = row0 + rl;
        if (gr >= nrows) continue;
        cx<T> r = {T(0), T(0)};
        if (2 * k != n) {
            const cx<T> zk = buf[(rl >> 1) * n + k];
            const cx<T> zc = buf[(rl >> 1) * n + ((n - k) % n)];
            if ((rl & 1) == 0) r = {(zk.x + zc.x) * T(0.5), (zk.y - zc.y) * T(0.5)};
            else r = {(zk.y + zc.y) * T(0.5), -(zk.x - zc.x) * T(0.5)};
            r.x *= scale;
            r.y *= scale;
        }
        out[gr * ldc + k] = r;
    }
}

// x c2r of ifftn_mpi (203_fftn_mpi.f90:62-64): half spectrum rows -> real
// rows, imaginary parts of kx = 0 and kx = n/2 ignored; deriv applies i kx.
template <typename T>
__global__ __launch_bounds__(NT) void k_x_c2r(const cx<T>* __restrict__ in, T* __restrict__ out, long nrows,
                                              int ldc, int ld_out, FftPlan pl, const cx<T>* __restrict__ tw_g,
                                              int deriv, T kfac, int P, int kst) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    const int n = pl.n;
    cx<T>* tw = buf + P * n;
    load_tw(tw, tw_g, n);
    const long row0 = (long)blockIdx.x * 2 * P;
    for (int e = threadIdx.x; e < P * n; e += NT) {
        const int line = e / n, k = e - line * n;
        const int kk = (2 * k <= n) ? k : n - k;
        const long ga = row0 + 2 * line, gb = ga + 1;
        cx<T> a = (ga < nrows && kk < kst) ? in[ga * ldc + kk] : cx<T>{T(0), T(0)};
        cx<T> b = (gb < nrows && kk < kst) ? in[gb * ldc + kk] : cx<T>{T(0), T(0)};
        if (kk == 0 || 2 * kk == n) { a.y = T(0); b.y = T(0); }
        if (deriv) {
            const T f = (2 * kk == n) ? T(0) : (T)kk * kfac;
            a = {-f * a.y, f * a.x};
            b = {-f * b.y, f * b.x};
        }
        if (2 * k > n) { a.y = -a.y; b.y = -b.y; }
        buf[e] = {a.x - b.y, a.y + b.x};
    }
    __syncthreads();
    fft_lines<T, EPT>(buf, pl, P, n, tw, +1);
    for (int e = threadIdx.x; e < 2 * P * n; e += NT) {
        const int rl = e / n, col = e - rl * n;
        const long gr = row0 + rl;
        if (gr < nrows) out[gr * ld_out + col] = reinterpret_cast<const T*>(&buf[(rl >> 1) * n + col])[rl & 1];
    }
}

// complex y-lines of the half spectrum (203_fftn_mpi.f90:29-33, :52-54).
template <typename T>
__global__ __launch_bounds__(NT) void k_yline_spec(const cx<T>* __restrict__ in, cx<T>* __restrict__ out,
                                                   int ncols, int ld, int ldG, int mc, long ps, FftPlan pl,
                                                   const cx<T>* __restrict__ tw_g, int op,
                                                   const double* __restrict__ G, int C) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    const int ny = pl.n;
    cx<T>* tw = buf + C * ny;
    load_tw(tw, tw_g, ny);
    const int x0 = blockIdx.x * C;
    const cx<T>* src = in + (long)blockIdx.y * ps;
    cx<T>* dst = out + (long)blockIdx.y * ps;
    auto keep = [&](int y) { return y < mc || y > ny - mc; };
    for (int e = threadIdx.x; e < ny * C; e += NT) {
        const int y = e / C, cl = e - y * C;
        const int x = x0 + cl;
        const bool rd = x < ncols && (op != SP_MULINV || keep(y));
        buf[cl * ny + y] = rd ? src[(long)y * ld + x] : cx<T>{T(0), T(0)};
    }
    __syncthreads();
    if (op == SP_FWD || op == SP_FILT) {
        fft_lines<T, EPT>(buf, pl, C, ny, tw, -1);
        for (int cl = threadIdx.x; cl < C; cl += NT) buf[cl * ny + ny / 2] = {T(0), T(0)};
        __syncthreads();
    }
    if (op == SP_FILT || op == SP_MULINV) {
        for (int e = threadIdx.x; e < ny * C; e += NT) {
            const int cl = e / ny, y = e - cl * ny;
            const int x = x0 + cl;
            const T g = (x < ncols) ? (T)G[(long)y * ldG + x] : T(0);
            buf[e].x *= g;
            buf[e].y *= g;
        }
        __syncthreads();
    }
    if (op != SP_FWD) fft_lines<T, EPT>(buf, pl, C, ny, tw, +1);
    for (int e = threadIdx.x; e < ny * C; e += NT) {
        const int y = e / C, cl = e - y * C;
        const int x = x0 + cl;
        if (x < ncols && (op != SP_FWD || keep(y))) dst[(long)y * ld + x] = buf[cl * ny + y];
    }
}

// ---- host launchers -------------------------------------------------------
template <typename T>
static const cx<T>* twp(const Twiddle& t);
template <>
const cx<double>* twp<double>(const Twiddle& t) { return t.d; }
template <>
const cx<float>* twp<float>(const Twiddle& t) { return t.f; }

static inline int lines_per_block(int L) {
    int P = (EPT * NT) / L;
    return P < 1 ? 1 : P;
}

int g_use_ct = 1;

static int line_fam(LineOp op, bool y) {
    switch (op) {
        case OP_DERIV: return y ? FAM_YDERIV : FAM_XDERIV;
        case OP_NYQ: return y ? FAM_YFILTER : FAM_XFILTER;
        case OP_PAD: return y ? FAM_YPAD : FAM_XPAD;
        default: return y ? FAM_YUNPAD : FAM_XUNPAD;
    }
}

template <typename T, typename TI>
void xline_real(hipStream_t s, const TI* in, T* out, long nrows, int ld_in, int ld_out, const Twiddle& tin,
                const Twiddle& tout, LineOp op, double kfac) {
    FamScope fs(s, line_fam(op, false), (double)nrows * ((double)tin.n * sizeof(TI) + (double)tout.n * sizeof(T)));
    if (g_use_ct && xline_real_ct<T, TI>(s, in, out, nrows, ld_in, ld_out, tin, tout, op, kfac)) return;
    const int L = tin.n > tout.n ? tin.n : tout.n;
    const int P = lines_per_block(L);
    const long nb = (nrows + 2 * P - 1) / (2 * P);
    const size_t lds = sizeof(cx<T>) * ((size_t)P * L + tin.n + tout.n);
    hipLaunchKernelGGL((k_xline_real<T, TI>), dim3((unsigned)nb), dim3(NT), lds, s, in, out, nrows, ld_in, ld_out,
                       tin.plan, tout.plan, twp<T>(tin), twp<T>(tout), (int)op, (T)kfac, P);
}

template <typename T, typename TI>
void yline_real(hipStream_t s, const TI* in, T* out, int nplanes, int ncols, int ld_in, int ld_out, long ps_in,
                long ps_out, const Twiddle& tin, const Twiddle& tout, LineOp op, double kfac) {
    FamScope fs(s, line_fam(op, true),
                (double)nplanes * ncols * ((double)tin.n * sizeof(TI) + (double)tout.n * sizeof(T)));
    if (g_use_ct && yline_real_ct<T, TI>(s, in, out, nplanes, ncols, ld_in, ld_out, ps_in, ps_out, tin, tout, op, kfac)) return;
    const int L = tin.n > tout.n ? tin.n : tout.n;
    const int C = lines_per_block(L);
    const int nbx = (ncols + 2 * C - 1) / (2 * C);
    const size_t lds = sizeof(cx<T>) * ((size_t)C * L + tin.n + tout.n);
    hipLaunchKernelGGL((k_yline_real<T, TI>), dim3(nbx, nplanes), dim3(NT), lds, s, in, out, ncols, ld_in, ld_out, ps_in,
                       ps_out, tin.plan, tout.plan, twp<T>(tin), twp<T>(tout), (int)op, (T)kfac, C);
}

template <typename T, typename TI>
void x_r2c(hipStream_t s, const TI* in, cx<T>* out, long nrows, int ld_in, int ldc, const Twiddle& tw,
           double scale, int kst) {
    if (kst < 0) kst = tw.n / 2 + 1;
    FamScope fs(s, FAM_XR2C, (double)nrows * ((double)tw.n * sizeof(TI) + 2.0 * kst * sizeof(T)));
    if (g_use_ct && x_r2c_ct<T, TI>(s, in, out, nrows, ld_in, ldc, tw, scale, kst)) return;
    const int P = lines_per_block(tw.n);
    const long nb = (nrows + 2 * P - 1) / (2 * P);
    const size_t lds = sizeof(cx<T>) * ((size_t)P * tw.n + tw.n);
    hipLaunchKernelGGL((k_x_r2c<T, TI>), dim3((unsigned)nb), dim3(NT), lds, s, in, out, nrows, ld_in, ldc, tw.plan,
                       twp<T>(tw), (T)scale, P, kst);
}

template <typename T>
void x_c2r(hipStream_t s, const cx<T>* in, T* out, long nrows, int ldc, int ld_out, const Twiddle& tw, int deriv,
           double kfac, int kst) {
    if (kst < 0) kst = tw.n / 2 + 1;
    FamScope fs(s, FAM_XC2R, (double)nrows * (2.0 * kst * sizeof(T) + (double)tw.n * sizeof(T)));
    if (g_use_ct && x_c2r_ct<T>(s, in, out, nrows, ldc, ld_out, tw, deriv, kfac, kst)) return;
    const int P = lines_per_block(tw.n);
    const long nb = (nrows + 2 * P - 1) / (2 * P);
    const size_t lds = sizeof(cx<T>) * ((size_t)P * tw.n + tw.n);
    hipLaunchKernelGGL(k_x_c2r<T>, dim3((unsigned)nb), dim3(NT), lds, s, in, out, nrows, ldc, ld_out, tw.plan,
                       twp<T>(tw), deriv, (T)kfac, P, kst);
}

template <typename T>
void yline_spec(hipStream_t s, const cx<T>* in, cx<T>* out, int nplanes, int ncols, long ps, const Twiddle& tw,
                SpecOp op, const double* G, int ld, int ldG, int mc) {
    if (ld < 0) ld = ncols;
    if (ldG < 0) ldG = ld;
    if (mc < 0) mc = tw.n;
    // rows read and written: the pruned forward stores, and the pruned
    // multiply-inverse reads, only the rows inside the filter support
    const double kept = (2 * mc - 1 < tw.n) ? (double)(2 * mc - 1) : (double)tw.n;
    const double rin = (op == SP_MULINV) ? kept : (double)tw.n, rout = (op == SP_FWD) ? kept : (double)tw.n;
    FamScope fs(s, FAM_YSPEC, (double)nplanes * ncols * 2.0 * sizeof(T) * (rin + rout));
    if (g_use_ct && yline_spec_ct<T>(s, in, out, nplanes, ncols, ps, tw, op, G, ld, ldG, mc)) return;
    const int C = lines_per_block(tw.n);
    const int nbx = (ncols + C - 1) / C;
    const size_t lds = sizeof(cx<T>) * ((size_t)C * tw.n + tw.n);
    hipLaunchKernelGGL(k_yline_spec<T>, dim3(nbx, nplanes), dim3(NT), lds, s, in, out, ncols, ld, ldG, mc, ps,
                       tw.plan, twp<T>(tw), (int)op, G, C);
}

#define LES_INST_LINES(T, TI)                                                                                   \
    template void xline_real<T, TI>(hipStream_t, const TI*, T*, long, int, int, const Twiddle&, const Twiddle&,      \
                                    LineOp, double);                                                               \
    template void yline_real<T, TI>(hipStream_t, const TI*, T*, int, int, int, int, long, long, const Twiddle&,      \
                                    const Twiddle&, LineOp, double);                                               \
    template void x_r2c<T, TI>(hipStream_t, const TI*, cx<T>*, long, int, int, const Twiddle&, double, int);
LES_INST_LINES(double, double)
LES_INST_LINES(float, float)
LES_INST_LINES(float, double)
#undef LES_INST_LINES
template void x_c2r<double>(hipStream_t, const cx<double>*, double*, long, int, int, const Twiddle&, int, double,
                            int);
template void x_c2r<float>(hipStream_t, const cx<float>*, float*, long, int, int, const Twiddle&, int, double, int);
template void yline_spec<double>(hipStream_t, const cx<double>*, cx<double>*, int, int, long, const Twiddle&,
                                 SpecOp, const double*, int, int, int);
template void yline_spec<float>(hipStream_t, const cx<float>*, cx<float>*, int, int, long, const Twiddle&,
                                SpecOp, const double*, int, int, int);

}  // namespace les
