// Register-radix line FFT building blocks.
//
// A line of length N is transformed in 2-3 Stockham stages of radix 8..24,
// each butterfly computed entirely in registers (a radix-16 butterfly is a
// 4 x 4 Cooley-Tukey DFT with compile-time twiddles).  Between stages the
// line is exchanged through LDS once, so a length-256 transform costs one LDS
// round trip instead of the four of a radix-4 Stockham.  The first stage
// reads its inputs straight from global memory and the last stage writes its
// outputs straight to global memory: both access patterns are unit-stride
// across the butterflies of a stage.
//
// Conventions as fft_lds.hpp: SIGN = -1 forward (FFTW_FORWARD), +1 backward,
// unnormalised.  The twiddle table tw[m] = exp(-2 pi i m / N) (forward).
#pragma once
#include <type_traits>
#include "fft_lds.hpp"

namespace les {

// ---- compile-time loops and roots of unity ---------------------------------
template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

struct ce_cs {
    double c, s;
};

// Double-double arithmetic for the compile-time roots (correctly rounded
// constants rather than a few ulp from a plain double series).
struct dd {
    double hi, lo;
};
constexpr dd dd_two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
constexpr dd dd_add(dd a, dd b) {
    const dd s = dd_two_sum(a.hi, b.hi);
    return dd_two_sum(s.hi, s.lo + a.lo + b.lo);
}
constexpr dd dd_split(double a) {
    const double c = 134217729.0 * a;  // 2^27 + 1
    const double h = c - (c - a);
    return {h, a - h};
}
constexpr dd dd_two_prod(double a, double b) {
    const double p = a * b;
    const dd as = dd_split(a), bs = dd_split(b);
    const double e = ((as.hi * bs.hi - p) + as.hi * bs.lo + as.lo * bs.hi) + as.lo * bs.lo;
    return {p, e};
}
constexpr dd dd_mul(dd a, dd b) {
    const dd p = dd_two_prod(a.hi, b.hi);
    return dd_two_sum(p.hi, p.lo + a.hi * b.lo + a.lo * b.hi);
}
constexpr dd dd_div_int(dd a, double d) {  // a / d, d a small integer
    const double q1 = a.hi / d;
    const dd r = dd_add(a, dd_two_prod(-q1, d));
    const double q2 = r.hi / d;
    return dd_two_sum(q1, q2);
}

// 2 pi in double-double
constexpr dd CE_2PI = {6.283185307179586232, 2.449293598294706e-16};

// Taylor series on |x| <= pi/4, in double-double
constexpr dd ce_cos_taylor(dd x) {
    const dd x2 = dd_mul(x, x);
    dd t = {1.0, 0.0}, s = {1.0, 0.0};
    for (int n = 1; n < 16; ++n) {
        t = dd_div_int(dd_mul(t, x2), -(2.0 * n - 1.0) * (2.0 * n));
        s = dd_add(s, t);
    }
    return s;
}
constexpr dd ce_sin_taylor(dd x) {
    const dd x2 = dd_mul(x, x);
    dd t = x, s = x;
    for (int n = 1; n < 16; ++n) {
        t = dd_div_int(dd_mul(t, x2), -(2.0 * n) * (2.0 * n + 1.0));
        s = dd_add(s, t);
    }
    return s;
}

// cos and sin of 2 pi p / q, by exact rational reduction to an angle in
// [0, pi/4] (octant symmetries) followed by the Taylor series.
constexpr ce_cs ce_cossin(long p, long q) {
    p %= q;
    if (p < 0) p += q;
    if (2 * p > q) {  // 2 pi - a
        const ce_cs r = ce_cossin(q - p, q);
        return {r.c, -r.s};
    }
    if (4 * p > q) {  // pi - a, a = 2 pi (q - 2p) / (2q)
        const ce_cs r = ce_cossin(q - 2 * p, 2 * q);
        return {-r.c, r.s};
    }
    if (8 * p > q) {  // pi/2 - a, a = 2 pi (q - 4p) / (4q)
        const ce_cs r = ce_cossin(q - 4 * p, 4 * q);
        return {r.s, r.c};
    }
    const dd x = dd_div_int(dd_mul(CE_2PI, dd{(double)p, 0.0}), (double)q);
    const dd c = ce_cos_taylor(x), s = ce_sin_taylor(x);
    return {c.hi + c.lo, s.hi + s.lo};
}

// exp(SIGN 2 pi i m / R) as a constant of type T
template <typename T, int M, int R, int SIGN>
__device__ __forceinline__ cx<T> root() {
    constexpr ce_cs v = ce_cossin(M, R);
    return {(T)v.c, (T)(SIGN * v.s)};
}

// multiply by exp(SIGN 2 pi i M / R), exact for the quarter turns
template <typename T, int M, int R, int SIGN>
__device__ __forceinline__ cx<T> rot(cx<T> a) {
    constexpr int m = ((M % R) + R) % R;
    if constexpr (m == 0) {
        return a;
    } else if constexpr (4 * m == R) {  // SIGN i
        return (SIGN > 0) ? cx<T>{-a.y, a.x} : cx<T>{a.y, -a.x};
    } else if constexpr (2 * m == R) {
        return {-a.x, -a.y};
    } else if constexpr (4 * m == 3 * R) {  // -SIGN i
        return (SIGN > 0) ? cx<T>{a.y, -a.x} : cx<T>{-a.y, a.x};
    } else {
        return cmul(a, root<T, m, R, SIGN>());
    }
}

// ---- in-register DFT of radix R (natural order in and out) ------------------
template <typename T, int R, int SIGN>
__device__ __forceinline__ void dft(cx<T>* x) {
    if constexpr (R == 1) {
        return;
    } else if constexpr (R == 2) {
        const cx<T> a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    } else if constexpr (R == 3) {
        constexpr T s3 = (T)0.86602540378443864676372317075294;
        const cx<T> t1 = cadd(x[1], x[2]);
        const cx<T> t2 = csub(x[1], x[2]);
        const cx<T> m1 = {x[0].x - (T)0.5 * t1.x, x[0].y - (T)0.5 * t1.y};
        const T ss = (SIGN < 0) ? -s3 : s3;
        const cx<T> m2 = {-ss * t2.y, ss * t2.x};
        x[0] = cadd(x[0], t1);
        x[1] = cadd(m1, m2);
        x[2] = csub(m1, m2);
    } else if constexpr (R == 4) {
        const cx<T> t0 = cadd(x[0], x[2]), t1 = csub(x[0], x[2]);
        const cx<T> t2 = cadd(x[1], x[3]), t3 = csub(x[1], x[3]);
        const cx<T> it3 = (SIGN < 0) ? cx<T>{t3.y, -t3.x} : cx<T>{-t3.y, t3.x};
        x[0] = cadd(t0, t2);
        x[1] = cadd(t1, it3);
        x[2] = csub(t0, t2);
        x[3] = csub(t1, it3);
    } else {
        // R = A B: A-point DFTs over the stride-B subsequences, twiddles
        // W_R^{b k1}, then B-point DFTs; X[k1 + A k2].
        constexpr int A = (R % 4 == 0) ? 4 : (R % 2 == 0) ? 2 : 3;
        constexpr int B = R / A;
        cx<T> y[R];
        static_for<0, B>([&](auto bc) {
            constexpr int b = decltype(bc)::value;
            cx<T> s[A];
            static_for<0, A>([&](auto ac) {
                constexpr int a = decltype(ac)::value;
                s[a] = x[b + B * a];
            });
            dft<T, A, SIGN>(s);
            static_for<0, A>([&](auto kc) {
                constexpr int k1 = decltype(kc)::value;
                y[k1 * B + b] = rot<T, b * k1, R, SIGN>(s[k1]);
            });
        });
        static_for<0, A>([&](auto kc) {
            constexpr int k1 = decltype(kc)::value;
            dft<T, B, SIGN>(y + k1 * B);
            static_for<0, B>([&](auto k2c) {
                constexpr int k2 = decltype(k2c)::value;
                x[k1 + A * k2] = y[k1 * B + k2];
            });
        });
    }
}

// ---- stage plans for the production line lengths ----------------------------
// Radices of the forward transform; the inverse uses them in reverse order, so
// that the last forward stage and the first inverse stage share the butterfly
// layout (the spectral multiplier of a same-length op stays in registers).
constexpr int rg_nst(int n) {
    switch (n) {
        case 512: case 768: return 3;
        default: return 2;
    }
}
constexpr int rg_radix(int n, int s) {
    switch (n) {
        case 32: return s == 0 ? 8 : 4;
        case 64: return 8;
        case 96: return s == 0 ? 8 : 12;
        case 128: return s == 0 ? 16 : 8;
        case 192: return s == 0 ? 16 : 12;
        case 256: return 16;
        case 384: return s == 0 ? 16 : 24;
        case 512: return 8;
        case 768: return s == 0 ? 12 : 8;
        default: return 0;
    }
}
constexpr bool rg_supported(int n) {
    return n == 64 || n == 96 || n == 128 || n == 192 || n == 256 || n == 384 || n == 512 || n == 768;
}
// radix of stage s of the transform in direction SIGN (inverse = reversed plan)
constexpr int rg_r(int n, int s, int sign) { return sign < 0 ? rg_radix(n, s) : rg_radix(n, rg_nst(n) - 1 - s); }
// span (product of the radices of the earlier stages)
constexpr int rg_ns(int n, int s, int sign) {
    int p = 1;
    for (int i = 0; i < s; ++i) p *= rg_r(n, i, sign);
    return p;
}
constexpr int rg_maxbf(int n) {
    int m = 0;
    for (int s = 0; s < rg_nst(n); ++s) m = (n / rg_radix(n, s) > m) ? n / rg_radix(n, s) : m;
    return m;
}

// LDS line layout: one padding element per 16 (bank spread of the stride-R
// butterfly writes) and an odd line stride.
__device__ __forceinline__ int rg_pad(int p) { return p + (p >> 4); }
constexpr int rg_ls(int n) { return n + n / 16 + 1; }

}  // namespace les
