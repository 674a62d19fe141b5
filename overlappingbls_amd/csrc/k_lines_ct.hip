// Compile-time specialised line-transform kernels for the production line
// lengths (64..768).  Same operations as k_lines.hip (which keeps a generic
// runtime-length path for every other size): a separable pass along x or y
// staged in LDS, real lines packed two to a complex line.  The y-kernels pad
// the LDS line stride by one element so that the column-tile loads and
// stores spread over the LDS banks.
#include "fft_lds_ct.hpp"
#include "les_internal.hpp"

namespace les {

namespace {

constexpr int NTC = 256;   // threads per block
constexpr int EPTC = 8;    // complex elements per thread per block

constexpr int lines_for(int L) { return (EPTC * NTC) / L > 0 ? (EPTC * NTC) / L : 1; }

// Column-tile width of the y-kernels: whole 128-byte row segments (8 complex
// lines of FP64 = 16 real columns, 16 of FP32), halved while the tile would
// exceed 64 KB of LDS.
template <typename T>
constexpr int ycols(int LS) {
    int c = (sizeof(T) == 8) ? 8 : 16;
    while (c > 1 && c * LS * (int)sizeof(cx<T>) > 65536 - 2 * 768 * (int)sizeof(cx<T>)) c /= 2;
    return c;
}

template <typename T, int NIN, int NOUT, int OP>
__device__ __forceinline__ cx<T> op_ct(const cx<T>* lin, int k, T kfac) {
    if constexpr (OP == OP_DERIV) {
        if (2 * k == NIN) return {T(0), T(0)};
        const int m = (2 * k < NIN) ? k : k - NIN;
        const T f = (T)m * kfac / (T)NIN;
        const cx<T> z = lin[k];
        return {-f * z.y, f * z.x};
    } else if constexpr (OP == OP_NYQ) {
        if (2 * k == NIN) return {T(0), T(0)};
        const T s = T(1) / (T)NIN;
        const cx<T> z = lin[k];
        return {z.x * s, z.y * s};
    } else if constexpr (OP == OP_PAD) {
        constexpr int h = NIN / 2;
        int src = -1;
        if (k < h) src = k;
        else if (k > NOUT - h) src = k - NOUT + NIN;
        if (src < 0) return {T(0), T(0)};
        const T s = T(1) / (T)NIN;
        const cx<T> z = lin[src];
        return {z.x * s, z.y * s};
    } else {  // OP_UNPAD
        constexpr int h = NOUT / 2;
        int src = -1;
        if (k < h) src = k;
        else if (k > h) src = k - NOUT + NIN;
        if (src < 0) return {T(0), T(0)};
        const T s = T(1) / (T)NIN;
        const cx<T> z = lin[src];
        return {z.x * s, z.y * s};
    }
}

// P lines of length NIN (stride LSI) -> P lines of length NOUT (stride LSO)
template <typename T, int NIN, int NOUT, int P, int LSI, int LSO, int OP>
__device__ __forceinline__ void apply_op_ct(cx<T>* buf, T kfac) {
    constexpr int tot = P * NOUT;
    constexpr int M = (tot + NTC - 1) / NTC;
    cx<T> r[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int e = threadIdx.x + m * NTC;
        if ((tot % NTC == 0) || e < tot) {
            const int line = e / NOUT, k = e - line * NOUT;
            r[m] = op_ct<T, NIN, NOUT, OP>(buf + line * LSI, k, kfac);
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int e = threadIdx.x + m * NTC;
        if ((tot % NTC == 0) || e < tot) {
            const int line = e / NOUT, k = e - line * NOUT;
            buf[line * LSO + k] = r[m];
        }
    }
    __syncthreads();
}

template <typename T, int N>
__device__ __forceinline__ void load_tw_ct(cx<T>* lds, const cx<T>* g) {
    for (int m = threadIdx.x; m < N; m += NTC) lds[m] = g[m];
}

// A derivative computed in a lower precision than its input is taken of the
// line minus its first element (the derivative of a constant is zero), so the
// conversion rounds the fluctuation about the line, not the mean flow.
template <typename T, typename TI, int OP>
constexpr bool SHIFT_IN = (OP == OP_DERIV) && (sizeof(TI) > sizeof(T));

template <typename T, typename TI, int NIN, int NOUT, int OP>
__global__ __launch_bounds__(NTC) void kc_xline(const TI* __restrict__ in, T* __restrict__ out, long nrows,
                                                int ld_in, int ld_out, const cx<T>* __restrict__ twin_g,
                                                const cx<T>* __restrict__ twout_g, T kfac) {
    constexpr int L = NIN > NOUT ? NIN : NOUT;
    constexpr int P = lines_for(L);
    __shared__ cx<T> buf[P * L];
    __shared__ cx<T> twi[NIN];
    __shared__ cx<T> two[NOUT];
    load_tw_ct<T, NIN>(twi, twin_g);
    load_tw_ct<T, NOUT>(two, twout_g);
    const long row0 = (long)blockIdx.x * 2 * P;
#pragma unroll 4
    for (int e = threadIdx.x; e < 2 * P * NIN; e += NTC) {
        const int rl = e / NIN, col = e - rl * NIN;
        const long gr = row0 + rl;
        T v = T(0);
        if (gr < nrows) {
            if constexpr (SHIFT_IN<T, TI, OP>) v = (T)(in[gr * ld_in + col] - in[gr * ld_in]);
            else v = (T)in[gr * ld_in + col];
        }
        reinterpret_cast<T*>(&buf[(rl >> 1) * NIN + col])[rl & 1] = v;
    }
    __syncthreads();
    fft_lines_ct<T, NIN, NIN, P, -1, NTC>(buf, twi);
    apply_op_ct<T, NIN, NOUT, P, NIN, NOUT, OP>(buf, kfac);
    fft_lines_ct<T, NOUT, NOUT, P, +1, NTC>(buf, two);
#pragma unroll 4
    for (int e = threadIdx.x; e < 2 * P * NOUT; e += NTC) {
        const int rl = e / NOUT, col = e - rl * NOUT;
        const long gr = row0 + rl;
        if (gr < nrows) out[gr * ld_out + col] = reinterpret_cast<const T*>(&buf[(rl >> 1) * NOUT + col])[rl & 1];
    }
}

template <typename T, typename TI, int NIN, int NOUT, int OP>
__global__ __launch_bounds__(NTC) void kc_yline(const TI* __restrict__ in, T* __restrict__ out, int ncols, int ld_in,
                                                int ld_out, long ps_in, long ps_out, const cx<T>* __restrict__ twin_g,
                                                const cx<T>* __restrict__ twout_g, T kfac) {
    constexpr int L = NIN > NOUT ? NIN : NOUT;
    constexpr int C = ycols<T>(L + 1);
    constexpr int LSI = NIN + 1, LSO = NOUT + 1;
    __shared__ cx<T> buf[C * (L + 1)];
    __shared__ cx<T> twi[NIN];
    __shared__ cx<T> two[NOUT];
    load_tw_ct<T, NIN>(twi, twin_g);
    load_tw_ct<T, NOUT>(two, twout_g);
    const int x0 = blockIdx.x * 2 * C;
    const TI* src = in + (long)blockIdx.y * ps_in;
    T* dst = out + (long)blockIdx.y * ps_out;
    constexpr int w2 = 2 * C;
#pragma unroll 4
    for (int e = threadIdx.x; e < NIN * w2; e += NTC) {
        const int y = e / w2, cl = e - y * w2;
        const int x = x0 + cl;
        T v = T(0);
        if (x < ncols) {
            if constexpr (SHIFT_IN<T, TI, OP>) v = (T)(src[(long)y * ld_in + x] - src[x]);
            else v = (T)src[(long)y * ld_in + x];
        }
        reinterpret_cast<T*>(&buf[(cl >> 1) * LSI + y])[cl & 1] = v;
    }
    __syncthreads();
    fft_lines_ct<T, NIN, LSI, C, -1, NTC>(buf, twi);
    apply_op_ct<T, NIN, NOUT, C, LSI, LSO, OP>(buf, kfac);
    fft_lines_ct<T, NOUT, LSO, C, +1, NTC>(buf, two);
#pragma unroll 4
    for (int e = threadIdx.x; e < NOUT * w2; e += NTC) {
        const int y = e / w2, cl = e - y * w2;
        const int x = x0 + cl;
        if (x < ncols) dst[(long)y * ld_out + x] = reinterpret_cast<const T*>(&buf[(cl >> 1) * LSO + y])[cl & 1];
    }
}

template <typename T, typename TI, int N>
__global__ __launch_bounds__(NTC) void kc_x_r2c(const TI* __restrict__ in, cx<T>* __restrict__ out, long nrows,
                                                int ld_in, int ldc, const cx<T>* __restrict__ tw_g, T scale) {
    constexpr int P = lines_for(N);
    constexpr int LH = N / 2 + 1;
    __shared__ cx<T> buf[P * N];
    __shared__ cx<T> tw[N];
    load_tw_ct<T, N>(tw, tw_g);
    const long row0 = (long)blockIdx.x * 2 * P;
#pragma unroll 4
    for (int e = threadIdx.x; e < 2 * P * N; e += NTC) {
        const int rl = e / N, col = e - rl * N;
        const long gr = row0 + rl;
        const T v = (gr < nrows) ? (T)in[gr * ld_in + col] : T(0);
        reinterpret_cast<T*>(&buf[(rl >> 1) * N + col])[rl & 1] = v;
    }
    __syncthreads();
    fft_lines_ct<T, N, N, P, -1, NTC>(buf, tw);
    for (int e = threadIdx.x; e < 2 * P * LH; e += NTC) {
        const int rl = e / LH, k = e - rl * LH;
        const long gr = row0 + rl;
        if (gr >= nrows) continue;
        cx<T> r = {T(0), T(0)};
        if (2 * k != N) {
            const cx<T> zk = buf[(rl >> 1) * N + k];
            const cx<T> zc = buf[(rl >> 1) * N + ((N - k) % N)];
            if ((rl & 1) == 0) r = {(zk.x + zc.x) * T(0.5), (zk.y - zc.y) * T(0.5)};
            else r = {(zk.y + zc.y) * T(0.5), -(zk.x - zc.x) * T(0.5)};
            r.x *= scale;
            r.y *= scale;
        }
        out[gr * ldc + k] = r;
    }
}

template <typename T, int N, int DERIV>
__global__ __launch_bounds__(NTC) void kc_x_c2r(const cx<T>* __restrict__ in, T* __restrict__ out, long nrows,
                                                int ldc, int ld_out, const cx<T>* __restrict__ tw_g, T kfac) {
    constexpr int P = lines_for(N);
    __shared__ cx<T> buf[P * N];
    __shared__ cx<T> tw[N];
    load_tw_ct<T, N>(tw, tw_g);
    const long row0 = (long)blockIdx.x * 2 * P;
    for (int e = threadIdx.x; e < P * N; e += NTC) {
        const int line = e / N, k = e - line * N;
        const int kk = (2 * k <= N) ? k : N - k;
        const long ga = row0 + 2 * line, gb = ga + 1;
        cx<T> a = (ga < nrows) ? in[ga * ldc + kk] : cx<T>{T(0), T(0)};
        cx<T> b = (gb < nrows) ? in[gb * ldc + kk] : cx<T>{T(0), T(0)};
        if (kk == 0 || 2 * kk == N) { a.y = T(0); b.y = T(0); }
        if constexpr (DERIV) {
            const T f = (2 * kk == N) ? T(0) : (T)kk * kfac;
            a = {-f * a.y, f * a.x};
            b = {-f * b.y, f * b.x};
        }
        if (2 * k > N) { a.y = -a.y; b.y = -b.y; }
        buf[e] = {a.x - b.y, a.y + b.x};
    }
    __syncthreads();
    fft_lines_ct<T, N, N, P, +1, NTC>(buf, tw);
#pragma unroll 4
    for (int e = threadIdx.x; e < 2 * P * N; e += NTC) {
        const int rl = e / N, col = e - rl * N;
        const long gr = row0 + rl;
        if (gr < nrows) out[gr * ld_out + col] = reinterpret_cast<const T*>(&buf[(rl >> 1) * N + col])[rl & 1];
    }
}

template <typename T, int N, int OP>
__global__ __launch_bounds__(NTC) void kc_yspec(const cx<T>* __restrict__ in, cx<T>* __restrict__ out, int lhx,
                                                long ps, const cx<T>* __restrict__ tw_g,
                                                const double* __restrict__ G) {
    constexpr int LS = N + 1;
    constexpr int C = ycols<T>(LS);
    __shared__ cx<T> buf[C * LS];
    __shared__ cx<T> tw[N];
    load_tw_ct<T, N>(tw, tw_g);
    const int x0 = blockIdx.x * C;
    const cx<T>* src = in + (long)blockIdx.y * ps;
    cx<T>* dst = out + (long)blockIdx.y * ps;
#pragma unroll 4
    for (int e = threadIdx.x; e < N * C; e += NTC) {
        const int y = e / C, cl = e - y * C;
        const int x = x0 + cl;
        buf[cl * LS + y] = (x < lhx) ? src[(long)y * lhx + x] : cx<T>{T(0), T(0)};
    }
    __syncthreads();
    if constexpr (OP == SP_FWD || OP == SP_FILT) {
        fft_lines_ct<T, N, LS, C, -1, NTC>(buf, tw);
        if (threadIdx.x < C) buf[threadIdx.x * LS + N / 2] = {T(0), T(0)};
        __syncthreads();
    }
    if constexpr (OP == SP_FILT || OP == SP_MULINV) {
        for (int e = threadIdx.x; e < N * C; e += NTC) {
            const int y = e / C, cl = e - y * C;
            const int x = x0 + cl;
            const T g = (x < lhx) ? (T)G[(long)y * lhx + x] : T(0);
            buf[cl * LS + y].x *= g;
            buf[cl * LS + y].y *= g;
        }
        __syncthreads();
    }
    if constexpr (OP != SP_FWD) fft_lines_ct<T, N, LS, C, +1, NTC>(buf, tw);
#pragma unroll 4
    for (int e = threadIdx.x; e < N * C; e += NTC) {
        const int y = e / C, cl = e - y * C;
        const int x = x0 + cl;
        if (x < lhx) dst[(long)y * lhx + x] = buf[cl * LS + y];
    }
}

template <typename T>
const cx<T>* twp_ct(const Twiddle& t);
template <>
const cx<double>* twp_ct<double>(const Twiddle& t) { return t.d; }
template <>
const cx<float>* twp_ct<float>(const Twiddle& t) { return t.f; }

// ---- dispatch helpers ------------------------------------------------------
template <typename T, typename TI, int NIN, int NOUT, int OP>
void launch_x(hipStream_t s, const TI* in, T* out, long nrows, int ld_in, int ld_out, const Twiddle& tin,
              const Twiddle& tout, double kfac) {
    constexpr int P = lines_for(NIN > NOUT ? NIN : NOUT);
    const long nb = (nrows + 2 * P - 1) / (2 * P);
    hipLaunchKernelGGL((kc_xline<T, TI, NIN, NOUT, OP>), dim3((unsigned)nb), dim3(NTC), 0, s, in, out, nrows, ld_in,
                       ld_out, twp_ct<T>(tin), twp_ct<T>(tout), (T)kfac);
}

template <typename T, typename TI, int NIN, int NOUT, int OP>
void launch_y(hipStream_t s, const TI* in, T* out, int nplanes, int ncols, int ld_in, int ld_out, long ps_in,
              long ps_out, const Twiddle& tin, const Twiddle& tout, double kfac) {
    constexpr int C = ycols<T>((NIN > NOUT ? NIN : NOUT) + 1);
    const int nbx = (ncols + 2 * C - 1) / (2 * C);
    hipLaunchKernelGGL((kc_yline<T, TI, NIN, NOUT, OP>), dim3(nbx, nplanes), dim3(NTC), 0, s, in, out, ncols, ld_in,
                       ld_out, ps_in, ps_out, twp_ct<T>(tin), twp_ct<T>(tout), (T)kfac);
}

#define LES_SIZES(X) X(64) X(128) X(256) X(512)

template <typename T, typename TI>
bool line_dispatch_x(hipStream_t s, const TI* in, T* out, long nrows, int ld_in, int ld_out, const Twiddle& tin,
                     const Twiddle& tout, LineOp op, double kfac) {
    const int a = tin.n, b = tout.n;
#define XCASE(n)                                                                                    \
    if (a == n && b == n && op == OP_DERIV) { launch_x<T, TI, n, n, OP_DERIV>(s, in, out, nrows, ld_in, ld_out, tin, tout, kfac); return true; } \
    if (a == n && b == n && op == OP_NYQ) { launch_x<T, TI, n, n, OP_NYQ>(s, in, out, nrows, ld_in, ld_out, tin, tout, kfac); return true; }     \
    if (a == n && b == 3 * n / 2 && op == OP_PAD) { launch_x<T, TI, n, 3 * n / 2, OP_PAD>(s, in, out, nrows, ld_in, ld_out, tin, tout, kfac); return true; } \
    if (a == 3 * n / 2 && b == n && op == OP_UNPAD) { launch_x<T, TI, 3 * n / 2, n, OP_UNPAD>(s, in, out, nrows, ld_in, ld_out, tin, tout, kfac); return true; }
    LES_SIZES(XCASE)
#undef XCASE
    return false;
}

template <typename T, typename TI>
bool line_dispatch_y(hipStream_t s, const TI* in, T* out, int nplanes, int ncols, int ld_in, int ld_out, long ps_in,
                     long ps_out, const Twiddle& tin, const Twiddle& tout, LineOp op, double kfac) {
    const int a = tin.n, b = tout.n;
#define YCASE(n)                                                                                    \
    if (a == n && b == n && op == OP_DERIV) { launch_y<T, TI, n, n, OP_DERIV>(s, in, out, nplanes, ncols, ld_in, ld_out, ps_in, ps_out, tin, tout, kfac); return true; } \
    if (a == n && b == n && op == OP_NYQ) { launch_y<T, TI, n, n, OP_NYQ>(s, in, out, nplanes, ncols, ld_in, ld_out, ps_in, ps_out, tin, tout, kfac); return true; }     \
    if (a == n && b == 3 * n / 2 && op == OP_PAD) { launch_y<T, TI, n, 3 * n / 2, OP_PAD>(s, in, out, nplanes, ncols, ld_in, ld_out, ps_in, ps_out, tin, tout, kfac); return true; } \
    if (a == 3 * n / 2 && b == n && op == OP_UNPAD) { launch_y<T, TI, 3 * n / 2, n, OP_UNPAD>(s, in, out, nplanes, ncols, ld_in, ld_out, ps_in, ps_out, tin, tout, kfac); return true; }
    LES_SIZES(YCASE)
#undef YCASE
    return false;
}

template <typename T, typename TI, int N>
void launch_r2c(hipStream_t s, const TI* in, cx<T>* out, long nrows, int ld_in, int ldc, const Twiddle& tw,
                double scale) {
    constexpr int P = lines_for(N);
    const long nb = (nrows + 2 * P - 1) / (2 * P);
    hipLaunchKernelGGL((kc_x_r2c<T, TI, N>), dim3((unsigned)nb), dim3(NTC), 0, s, in, out, nrows, ld_in, ldc,
                       twp_ct<T>(tw), (T)scale);
}

template <typename T, int N, int D>
void launch_c2r(hipStream_t s, const cx<T>* in, T* out, long nrows, int ldc, int ld_out, const Twiddle& tw,
                double kfac) {
    constexpr int P = lines_for(N);
    const long nb = (nrows + 2 * P - 1) / (2 * P);
    hipLaunchKernelGGL((kc_x_c2r<T, N, D>), dim3((unsigned)nb), dim3(NTC), 0, s, in, out, nrows, ldc, ld_out,
                       twp_ct<T>(tw), (T)kfac);
}

template <typename T, int N, int OP>
void launch_spec(hipStream_t s, const cx<T>* in, cx<T>* out, int nplanes, int lhx, long ps, const Twiddle& tw,
                 const double* G) {
    constexpr int C = ycols<T>(N + 1);
    const int nbx = (lhx + C - 1) / C;
    hipLaunchKernelGGL((kc_yspec<T, N, OP>), dim3(nbx, nplanes), dim3(NTC), 0, s, in, out, lhx, ps, twp_ct<T>(tw),
                       G);
}

template <typename T, typename TI>
bool r2c_dispatch(hipStream_t s, const TI* in, cx<T>* out, long nrows, int ld_in, int ldc, const Twiddle& tw,
                  double scale) {
#define RCASE(NN_) \
    if (tw.n == NN_) { launch_r2c<T, TI, NN_>(s, in, out, nrows, ld_in, ldc, tw, scale); return true; }
    LES_SIZES(RCASE)
#undef RCASE
    return false;
}

template <typename T>
bool c2r_dispatch(hipStream_t s, const cx<T>* in, T* out, long nrows, int ldc, int ld_out, const Twiddle& tw,
                  int deriv, double kfac) {
#define CCASE(NN_)                                                                               \
    if (tw.n == NN_) {                                                                           \
        if (deriv) launch_c2r<T, NN_, 1>(s, in, out, nrows, ldc, ld_out, tw, kfac);              \
        else launch_c2r<T, NN_, 0>(s, in, out, nrows, ldc, ld_out, tw, kfac);                    \
        return true;                                                                           \
    }
    LES_SIZES(CCASE)
#undef CCASE
    return false;
}

template <typename T>
bool spec_dispatch(hipStream_t s, const cx<T>* in, cx<T>* out, int nplanes, int lhx, long ps, const Twiddle& tw,
                   SpecOp op, const double* G) {
#define SCASE(NN_)                                                                                          \
    if (tw.n == NN_) {                                                                                      \
        switch (op) {                                                                                     \
            case SP_FWD: launch_spec<T, NN_, SP_FWD>(s, in, out, nplanes, lhx, ps, tw, G); break;           \
            case SP_INV: launch_spec<T, NN_, SP_INV>(s, in, out, nplanes, lhx, ps, tw, G); break;           \
            case SP_FILT: launch_spec<T, NN_, SP_FILT>(s, in, out, nplanes, lhx, ps, tw, G); break;         \
            default: launch_spec<T, NN_, SP_MULINV>(s, in, out, nplanes, lhx, ps, tw, G); break;            \
        }                                                                                                 \
        return true;                                                                                      \
    }
    LES_SIZES(SCASE)
#undef SCASE
    return false;
}

}  // namespace

// exported to k_lines.hip
template <typename T, typename TI>
bool xline_real_ct(hipStream_t s, const TI* in, T* out, long nrows, int ld_in, int ld_out, const Twiddle& tin,
                   const Twiddle& tout, LineOp op, double kfac) {
    return line_dispatch_x<T, TI>(s, in, out, nrows, ld_in, ld_out, tin, tout, op, kfac);
}
template <typename T, typename TI>
bool yline_real_ct(hipStream_t s, const TI* in, T* out, int nplanes, int ncols, int ld_in, int ld_out, long ps_in,
                   long ps_out, const Twiddle& tin, const Twiddle& tout, LineOp op, double kfac) {
    return line_dispatch_y<T, TI>(s, in, out, nplanes, ncols, ld_in, ld_out, ps_in, ps_out, tin, tout, op, kfac);
}
template <typename T, typename TI>
bool x_r2c_ct(hipStream_t s, const TI* in, cx<T>* out, long nrows, int ld_in, int ldc, const Twiddle& tw,
              double scale) {
    return r2c_dispatch<T, TI>(s, in, out, nrows, ld_in, ldc, tw, scale);
}
template <typename T>
bool x_c2r_ct(hipStream_t s, const cx<T>* in, T* out, long nrows, int ldc, int ld_out, const Twiddle& tw, int deriv,
              double kfac) {
    return c2r_dispatch<T>(s, in, out, nrows, ldc, ld_out, tw, deriv, kfac);
}
template <typename T>
bool yline_spec_ct(hipStream_t s, const cx<T>* in, cx<T>* out, int nplanes, int lhx, long ps, const Twiddle& tw,
                   SpecOp op, const double* G) {
    return spec_dispatch<T>(s, in, out, nplanes, lhx, ps, tw, op, G);
}

#define INST_IO(T, TI)                                                                                       \
    template bool xline_real_ct<T, TI>(hipStream_t, const TI*, T*, long, int, int, const Twiddle&, const Twiddle&, \
                                       LineOp, double);                                                          \
    template bool yline_real_ct<T, TI>(hipStream_t, const TI*, T*, int, int, int, int, long, long, const Twiddle&, \
                                       const Twiddle&, LineOp, double);                                          \
    template bool x_r2c_ct<T, TI>(hipStream_t, const TI*, cx<T>*, long, int, int, const Twiddle&, double);
INST_IO(double, double)
INST_IO(float, float)
INST_IO(float, double)
#undef INST_IO
template bool x_c2r_ct<double>(hipStream_t, const cx<double>*, double*, long, int, int, const Twiddle&, int, double);
template bool x_c2r_ct<float>(hipStream_t, const cx<float>*, float*, long, int, int, const Twiddle&, int, double);
template bool yline_spec_ct<double>(hipStream_t, const cx<double>*, cx<double>*, int, int, long, const Twiddle&,
                                    SpecOp, const double*);
template bool yline_spec_ct<float>(hipStream_t, const cx<float>*, cx<float>*, int, int, long, const Twiddle&,
                                   SpecOp, const double*);

}  // namespace les
