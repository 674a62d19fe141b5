// Compile-time specialised LDS FFT: the line length N, the LDS line stride
// LS, the number of lines NL and the direction are template parameters, so
// every Stockham stage unrolls with constant radix, span and index maps.
// Same conventions as fft_lds.hpp (FFTW signs, unnormalised).
#pragma once
#include "fft_lds.hpp"

namespace les {

constexpr int ct_radix(int n) { return (n % 4 == 0) ? 4 : (n % 2 == 0) ? 2 : 3; }

constexpr int ct_nstages(int n) {
    int c = 0;
    while (n > 1) {
        n /= ct_radix(n);
        ++c;
    }
    return c;
}

// radix of stage s
constexpr int ct_radix_at(int n, int s) {
    for (int i = 0; i < s; ++i) n /= ct_radix(n);
    return ct_radix(n);
}

// span Ns before stage s (product of the radices of stages 0..s-1)
constexpr int ct_span_at(int n, int s) {
    int p = 1;
    for (int i = 0; i < s; ++i) {
        const int r = ct_radix(n);
        p *= r;
        n /= r;
    }
    return p;
}

template <typename T, int N, int R, int NS, int LS, int NL, int SIGN, int NTH>
__device__ __forceinline__ void ct_stage(cx<T>* buf, const cx<T>* tw) {
    constexpr int nR = N / R;
    constexpr int total = NL * nR;
    constexpr int tstep = N / (NS * R);
    constexpr int PB = (total + NTH - 1) / NTH;
    cx<T> v[PB][R];
    int dst[PB];
#pragma unroll
    for (int b = 0; b < PB; ++b) {
        const int jb = threadIdx.x + b * NTH;
        dst[b] = -1;
        if ((total % NTH == 0) || jb < total) {
            const int line = jb / nR;
            const int j = jb - line * nR;
            const int k = j % NS;
            const cx<T>* src = buf + line * LS;
            cx<T> a[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                cx<T> t = src[j + r * nR];
                if (r > 0 && NS > 1) {
                    cx<T> w = tw[r * k * tstep];
                    if (SIGN > 0) w.y = -w.y;
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
                cx<T> it3 = (SIGN < 0) ? cx<T>{t3.y, -t3.x} : cx<T>{-t3.y, t3.x};
                v[b][0] = cadd(t0, t2);
                v[b][1] = cadd(t1, it3);
                v[b][2] = csub(t0, t2);
                v[b][3] = csub(t1, it3);
            } else {
                const T s3 = (T)0.86602540378443864676372317075294;
                cx<T> t1 = cadd(a[1], a[2]);
                cx<T> t2 = csub(a[1], a[2]);
                cx<T> m1 = {a[0].x - (T)0.5 * t1.x, a[0].y - (T)0.5 * t1.y};
                const T ss = (SIGN < 0) ? -s3 : s3;
                cx<T> m2 = {-ss * t2.y, ss * t2.x};
                v[b][0] = cadd(a[0], t1);
                v[b][1] = cadd(m1, m2);
                v[b][2] = csub(m1, m2);
            }
            dst[b] = line * LS + (j / NS) * NS * R + k;
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < PB; ++b) {
        if (dst[b] >= 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) buf[dst[b] + r * NS] = v[b][r];
        }
    }
    __syncthreads();
}

template <typename T, int N, int S, int LS, int NL, int SIGN, int NTH>
__device__ __forceinline__ void ct_stages(cx<T>* buf, const cx<T>* tw) {
    if constexpr (S < ct_nstages(N)) {
        ct_stage<T, N, ct_radix_at(N, S), ct_span_at(N, S), LS, NL, SIGN, NTH>(buf, tw);
        ct_stages<T, N, S + 1, LS, NL, SIGN, NTH>(buf, tw);
    }
}

// NL lines of length N at stride LS, in place
template <typename T, int N, int LS, int NL, int SIGN, int NTH>
__device__ __forceinline__ void fft_lines_ct(cx<T>* buf, const cx<T>* tw) {
    ct_stages<T, N, 0, LS, NL, SIGN, NTH>(buf, tw);
}

}  // namespace les
