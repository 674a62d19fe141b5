// LDS-staged batched complex FFT for CDNA4 (gfx950), wave64.
//
// A workgroup of NT threads transforms `nlines` complex lines of length n
// held in LDS (line stride ls elements).  The length factors as a runtime
// sequence of radix-4, radix-2 and radix-3 Stockham stages; each stage reads
// its butterfly inputs into registers, synchronises, and writes the outputs
// back to the same LDS buffer (autosort, natural order out).  Twiddles come
// from an LDS table tw[m] = exp(-2 pi i m / n), conjugated for the inverse.
//
// Convention (FFTW's): sign -1 forward, +1 backward, unnormalised.
#pragma once
#include <hip/hip_runtime.h>

namespace les {

template <typename T>
struct cx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }

struct FftPlan {
    int n;
    int nst;
    int radix[16];
};

inline FftPlan make_plan(int n) {
    FftPlan p{};
    p.n = n;
    int m = n;
    while (m > 1) {
        int r = (m % 4 == 0) ? 4 : (m % 2 == 0) ? 2 : 3;
        p.radix[p.nst++] = r;
        m /= r;
    }
    return p;
}

// One Stockham stage of compile-time radix R.  EPT bounds the elements per
// thread (nlines * n <= EPT * blockDim.x).
template <typename T, int R, int EPT>
__device__ __forceinline__ void fft_stage(cx<T>* buf, int n, int nlines, int ls, int Ns,
                                          const cx<T>* tw, int sign) {
    constexpr int PB = (EPT + R - 1) / R;
    const int nR = n / R;
    const int total = nlines * nR;
    const int tstep = n / (Ns * R);
    const int nt = blockDim.x;
    cx<T> v[PB][R];
    int dst[PB];
#pragma unroll
    for (int b = 0; b < PB; ++b) {
        const int jb = threadIdx.x + b * nt;
        dst[b] = -1;
        if (jb < total) {
            const int line = jb / nR;
            const int j = jb - line * nR;
            const int k = j % Ns;
            const cx<T>* src = buf + line * ls;
            cx<T> a[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                cx<T> t = src[j + r * nR];
                if (r > 0 && k > 0) {
                    cx<T> w = tw[(r * k * tstep) % n];
                    if (sign > 0) w.y = -w.y;
                    t = cmul(t, w);
                }
                a[r] = t;
            }
            if constexpr (R == 2) {
                v[b][0] = cadd(a[0], a[1]);
                v[b][1] = csub(a[0], a[1]);
            } else if constexpr (R == 4) {
                cx<T> t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
                cx<T> t2 = cadd(a[1], a[3]), t3 = csub(a[1], a[3]);
                // (sign) i * t3
                cx<T> it3 = (sign < 0) ? cx<T>{t3.y, -t3.x} : cx<T>{-t3.y, t3.x};
                v[b][0] = cadd(t0, t2);
                v[b][1] = cadd(t1, it3);
                v[b][2] = csub(t0, t2);
                v[b][3] = csub(t1, it3);
            } else {  // R == 3
                const T s3 = (T)0.86602540378443864676372317075294;
                cx<T> t1 = cadd(a[1], a[2]);
                cx<T> t2 = csub(a[1], a[2]);
                cx<T> m1 = {a[0].x - (T)0.5 * t1.x, a[0].y - (T)0.5 * t1.y};
                const T ss = (sign < 0) ? -s3 : s3;
                cx<T> m2 = {-ss * t2.y, ss * t2.x};
                v[b][0] = cadd(a[0], t1);
                v[b][1] = cadd(m1, m2);
                v[b][2] = csub(m1, m2);
            }
            dst[b] = line * ls + (j / Ns) * Ns * R + k;
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < PB; ++b) {
        if (dst[b] >= 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) buf[dst[b] + r * Ns] = v[b][r];
        }
    }
    __syncthreads();
}

template <typename T, int EPT>
__device__ __forceinline__ void fft_lines(cx<T>* buf, const FftPlan& p, int nlines, int ls,
                                          const cx<T>* tw, int sign) {
    int Ns = 1;
    for (int s = 0; s < p.nst; ++s) {
        const int R = p.radix[s];
        if (R == 4) fft_stage<T, 4, EPT>(buf, p.n, nlines, ls, Ns, tw, sign);
        else if (R == 2) fft_stage<T, 2, EPT>(buf, p.n, nlines, ls, Ns, tw, sign);
        else fft_stage<T, 3, EPT>(buf, p.n, nlines, ls, Ns, tw, sign);
        Ns *= R;
    }
}

// Load the twiddle table for length n from global memory into LDS.
template <typename T>
__device__ __forceinline__ void load_tw(cx<T>* lds, const cx<T>* g, int n) {
    for (int m = threadIdx.x; m < n; m += blockDim.x) lds[m] = g[m];
}

}  // namespace les
