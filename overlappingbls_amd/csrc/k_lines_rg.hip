// Line-transform kernels for the production line lengths (64..768) built on
// the register-radix FFT of fft_rg.hpp.  Same operations as the generic
// runtime-length kernels of k_lines.hip:
//   x-lines  : real rows packed two to a complex line
//   y-lines  : real column pairs packed into one complex line (column tiles,
//              unit-stride row segments in global memory)
//   r2c/c2r  : x half-spectrum transforms of fftn_mpi / ifftn_mpi
//   yspec    : complex y-lines of the half spectrum
// The first stage of each transform reads global memory directly, the last
// stage writes it directly; stages exchange through LDS.
#include "fft_rg.hpp"
#include "les_lag.hpp"

namespace les {

namespace {

constexpr int round64(int n) { return ((n + 63) / 64) * 64; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }

// A first-stage accessor in two parts: raw(line, position) reads global memory
// in the input precision, cv(line, value) converts to the transform input.
// Stage::load issues every raw read before the first conversion; fused into
// one accessor, the compiler converts each value as it arrives and waits on
// each load in turn, which halves the bandwidth of FP64-input kernels.
template <typename RAW, typename CV>
struct SplitLoad {
    RAW raw;
    CV cv;
};
template <typename RAW, typename CV>
__device__ __forceinline__ SplitLoad<RAW, CV> split_load(const RAW& raw, const CV& cv) {
    return {raw, cv};
}

// ---- one Stockham stage: butterflies distributed over the block --------------
// Butterfly (line, b), b < N/R: inputs at b + r N/R, twiddles W_{NS R}^{r (b mod NS)},
// outputs at (b / NS) NS R + b mod NS + r NS.  YM: lines vary fastest across
// threads (y-lines: neighbouring threads read neighbouring columns).
template <typename T, int N, int S, int SIGN, int NL, int NTH, bool YM>
struct Stage {
    static constexpr int R = rg_r(N, S, SIGN);
    static constexpr int NS = rg_ns(N, S, SIGN);
    static constexpr int NB = N / R;
    static constexpr int TOT = NL * NB;
    static constexpr int PB = (TOT + NTH - 1) / NTH;
    cx<T> v[PB][R];

    __device__ __forceinline__ static bool bf(int q, int& line, int& b) {
        const int jb = (int)threadIdx.x + q * NTH;
        if (TOT % NTH != 0 && jb >= TOT) return false;
        if constexpr (YM) {
            line = jb % NL;
            b = jb / NL;
        } else {
            line = jb / NB;
            b = jb - line * NB;
        }
        return true;
    }

    // inputs from LDS (line stride LS, padded index)
    template <int LS>
    __device__ __forceinline__ void read(const cx<T>* buf) {
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            int line, b;
            if (!bf(q, line, b)) continue;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = buf[line * LS + rg_pad(b + r * NB)];
        }
    }

    // inputs from a functor ld(line, position)
    template <typename LD>
    __device__ __forceinline__ void load(const LD& ld) {
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            int line, b;
            if (!bf(q, line, b)) continue;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = ld(line, b + r * NB);
        }
    }

    // inputs from a split accessor: all reads first, then the conversions
    template <typename RAW, typename CV>
    __device__ __forceinline__ void load(const SplitLoad<RAW, CV>& ld) {
        using P = decltype(ld.raw(0, 0));
        P t[PB][R];
        // threads without a butterfly in this pass read line 0 (no branch
        // around the reads) and discard it
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            int line = 0, b = 0;
            if (!bf(q, line, b)) line = b = 0;
#pragma unroll
            for (int r = 0; r < R; ++r) t[q][r] = ld.raw(line, b + r * NB);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            int line, b;
            if (!bf(q, line, b)) continue;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = ld.cv(line, t[q][r]);
        }
    }

    __device__ __forceinline__ void compute(const cx<T>* tw) {
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            int line, b;
            if (!bf(q, line, b)) continue;
            if constexpr (NS > 1) {
                constexpr int tstep = N / (NS * R);
                const int k = b % NS;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    cx<T> w = tw[r * k * tstep];
                    if (SIGN > 0) w.y = -w.y;
                    v[q][r] = cmul(v[q][r], w);
                }
            }
            dft<T, R, SIGN>(v[q]);
        }
    }

    template <int LS>
    __device__ __forceinline__ void write(cx<T>* buf) const {
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            int line, b;
            if (!bf(q, line, b)) continue;
            const int base = (b / NS) * NS * R + b % NS;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[line * LS + rg_pad(base + r * NS)] = v[q][r];
        }
    }

    // last stage: outputs at b + r NB, to a functor st(line, position, value)
    template <typename ST>
    __device__ __forceinline__ void store(const ST& st) const {
        static_assert(NS * R == N, "store() is for the last stage");
#pragma unroll
        for (int q = 0; q < PB; ++q) {
            int line, b;
            if (!bf(q, line, b)) continue;
#pragma unroll
            for (int r = 0; r < R; ++r) st(line, b + r * NB, v[q][r]);
        }
    }
};

// Stages 1 .. NST-2 (LDS to LDS) of a transform whose stage 0 output is in LDS.
template <typename T, int N, int SIGN, int NL, int NTH, bool YM, int LS>
__device__ __forceinline__ void middle_stages(cx<T>* buf, const cx<T>* tw) {
    if constexpr (rg_nst(N) == 3) {
        Stage<T, N, 1, SIGN, NL, NTH, YM> s1;
        s1.template read<LS>(buf);
        s1.compute(tw);
        __syncthreads();
        s1.template write<LS>(buf);
        __syncthreads();
    }
}

// The pad kernels read their twiddle tables through the vector cache instead
// of staging them in LDS: the freed LDS admits more blocks per CU (measured
// at 256 -> 384: FP32 y pass 78 -> 55 us, FP64 x pass 99 -> 89 us).  The
// unpad and same-length kernels are faster with the tables in LDS.
template <int NIN, int NOUT>
constexpr bool tw_global() {
    return NIN < NOUT;
}

template <typename T>
__device__ __forceinline__ void load_twr(cx<T>* lds, const cx<T>* g, int n) {
    for (int m = threadIdx.x; m < n; m += blockDim.x) lds[m] = g[m];
}

// spectral multiplier of the same-length ops at output position k of NIN
template <typename T, int N, int OP>
__device__ __forceinline__ cx<T> same_len_op(cx<T> z, int k, T kfac) {
    if (2 * k == N) return {T(0), T(0)};
    if constexpr (OP == OP_DERIV) {
        const int m = (2 * k < N) ? k : k - N;
        const T f = (T)m * kfac / (T)N;
        return {-f * z.y, f * z.x};
    } else {
        const T s = T(1) / (T)N;
        return {z.x * s, z.y * s};
    }
}

// source position in the NIN spectrum of output position k of NOUT (pad /
// unpad), or -1 for a zero
template <int NIN, int NOUT, int OP>
__device__ __forceinline__ int resize_src(int k) {
    if constexpr (OP == OP_PAD) {
        constexpr int h = NIN / 2;
        if (k < h) return k;
        if (k > NOUT - h) return k - NOUT + NIN;
        return -1;
    } else {
        constexpr int h = NOUT / 2;
        if (k < h) return k;
        if (k > h) return k - NOUT + NIN;
        return -1;
    }
}

// FFT(NIN) -> op -> IFFT(NOUT) on NL lines; ld/st are the global accessors.
// LDS_IN: ld reads the line buffer itself (inputs staged there), so the
// first stage's reads complete before it overwrites the buffer.
template <typename T, int NIN, int NOUT, int OP, int NL, int NTH, bool YM, bool LDS_IN = false, typename LD,
          typename ST>
__device__ __forceinline__ void line_pipeline(const LD& ld, const ST& st, cx<T>* buf, const cx<T>* twi,
                                              const cx<T>* two, T kfac) {
    constexpr int LS = rg_ls(cmax(NIN, NOUT));
    constexpr int SF = rg_nst(NIN);
    constexpr int SI = rg_nst(NOUT);
    // forward
    Stage<T, NIN, 0, -1, NL, NTH, YM> f0;
    f0.load(ld);
    f0.compute(twi);
    if constexpr (LDS_IN) __syncthreads();
    f0.template write<LS>(buf);
    __syncthreads();
    middle_stages<T, NIN, -1, NL, NTH, YM, LS>(buf, twi);
    Stage<T, NIN, SF - 1, -1, NL, NTH, YM> fl;
    fl.template read<LS>(buf);
    fl.compute(twi);
    // op and first inverse stage
    Stage<T, NOUT, 0, +1, NL, NTH, YM> i0;
    if constexpr (NIN == NOUT) {
        static_assert(decltype(fl)::R == decltype(i0)::R, "reversed plans share the butterfly layout");
#pragma unroll
        for (int q = 0; q < decltype(i0)::PB; ++q) {
            int line, b;
            if (!decltype(i0)::bf(q, line, b)) continue;
#pragma unroll
            for (int r = 0; r < decltype(i0)::R; ++r)
                i0.v[q][r] = same_len_op<T, NIN, OP>(fl.v[q][r], b + r * decltype(i0)::NB, kfac);
        }
        i0.compute(two);
        __syncthreads();
    } else {
        __syncthreads();
        fl.store([&](int line, int p, cx<T> z) { buf[line * LS + rg_pad(p)] = z; });
        __syncthreads();
        const T s = T(1) / (T)NIN;
        i0.load([&](int line, int p) {
            const int src = resize_src<NIN, NOUT, OP>(p);
            if (src < 0) return cx<T>{T(0), T(0)};
            const cx<T> z = buf[line * LS + rg_pad(src)];
            return cx<T>{z.x * s, z.y * s};
        });
        i0.compute(two);
        __syncthreads();
    }
    i0.template write<LS>(buf);
    __syncthreads();
    middle_stages<T, NOUT, +1, NL, NTH, YM, LS>(buf, two);
    Stage<T, NOUT, SI - 1, +1, NL, NTH, YM> il;
    il.template read<LS>(buf);
    il.compute(two);
    il.store(st);
}

// A derivative computed in a lower precision than its input is taken of the
// line minus its first element (the derivative of a constant is zero), so the
// conversion rounds the fluctuation about the line, not the mean flow.
template <typename T, typename TI, int OP>
constexpr bool SHIFT_IN = (OP == OP_DERIV) && (sizeof(TI) > sizeof(T));

template <typename T>
struct pair_t;
template <>
struct pair_t<double> {
    using type = double2;
};
template <>
struct pair_t<float> {
    using type = float2;
};

// lines per block and threads per block of a line kernel: y-lines take
// 128-byte row segments (8 FP64 or 16 FP32 column pairs), x-lines enough lines
// for 128 threads; halved while the block would exceed 512 threads or 64 KB
// of LDS (line buffers and twiddle tables).
template <typename T, int NIN, int NOUT, bool YM>
constexpr int rg_lines() {
    constexpr int mb = cmax(rg_maxbf(NIN), rg_maxbf(NOUT));
    int nl = YM ? (sizeof(T) == 8 ? 8 : 16) : cmax(1, 128 / mb);
    while (nl > 1 && (nl * mb > 512 ||
                      (nl * rg_ls(cmax(NIN, NOUT)) + NIN + NOUT) * (int)sizeof(cx<T>) > 65536))
        nl /= 2;
    return nl;
}
template <typename T, int NIN, int NOUT, bool YM>
constexpr int rg_threads() {
    return round64(rg_lines<T, NIN, NOUT, YM>() * cmax(rg_maxbf(NIN), rg_maxbf(NOUT)));
}

template <typename T, typename TI, int NIN, int NOUT, int OP>
__global__ __launch_bounds__((rg_threads<T, NIN, NOUT, false>())) void kr_xline(
    const TI* __restrict__ in, T* __restrict__ out, long nrows, int ld_in, int ld_out,
    const cx<T>* __restrict__ twin_g, const cx<T>* __restrict__ twout_g, T kfac) {
    constexpr int NL = rg_lines<T, NIN, NOUT, false>();
    constexpr int NTH = rg_threads<T, NIN, NOUT, false>();
    constexpr int LS = rg_ls(cmax(NIN, NOUT));
    __shared__ cx<T> buf[NL * LS];
    constexpr bool TWG = tw_global<NIN, NOUT>();
    __shared__ cx<T> twl[TWG ? 1 : NIN];
    __shared__ cx<T> twl_o[(TWG || NIN == NOUT) ? 1 : NOUT];
    const cx<T>* twi = twin_g;
    const cx<T>* twop = twout_g;
    if constexpr (!TWG) {
        load_twr(twl, twin_g, NIN);
        if (NIN != NOUT) load_twr(twl_o, twout_g, NOUT);
        twi = twl;
        twop = (NIN == NOUT) ? twl : twl_o;
    }
    const long row0 = (long)blockIdx.x * 2 * NL;
    // first element of each of the block's rows (the shift of a reduced-
    // precision derivative), staged once instead of reloaded per element
    __shared__ TI first[SHIFT_IN<T, TI, OP> ? 2 * NL : 1];
    if constexpr (SHIFT_IN<T, TI, OP>) {
        for (int t = threadIdx.x; t < 2 * NL; t += NTH) {
            const long r = row0 + t;
            first[t] = (r < nrows) ? in[r * ld_in] : TI(0);
        }
    }
    using PI = typename pair_t<TI>::type;
    // rows 2 line, 2 line + 1 at position p (zero past the last row).  The
    // reads are unconditional, from the last row when past it, so that no
    // branch separates them.
    auto raw = [&](int line, int p) {
        const long ra = row0 + 2 * line, rb = ra + 1;
        const TI a = in[(ra < nrows ? ra : nrows - 1) * ld_in + p];
        const TI b = in[(rb < nrows ? rb : nrows - 1) * ld_in + p];
        return PI{(ra < nrows) ? a : TI(0), (rb < nrows) ? b : TI(0)};
    };
    // rows past the last have first[] = 0, so they stay zero
    auto cv = [&](int line, PI w) {
        if constexpr (SHIFT_IN<T, TI, OP>) return cx<T>{(T)(w.x - first[2 * line]), (T)(w.y - first[2 * line + 1])};
        else return cx<T>{(T)w.x, (T)w.y};
    };
    auto st = [&](int line, int p, cx<T> z) {
        const long ra = row0 + 2 * line, rb = ra + 1;
        if (ra < nrows) out[ra * ld_out + p] = z.x;
        if (rb < nrows) out[rb * ld_out + p] = z.y;
    };
    __syncthreads();
    if constexpr (sizeof(TI) > sizeof(T)) {
        line_pipeline<T, NIN, NOUT, OP, NL, NTH, false>(split_load(raw, cv), st, buf, twi, twop, kfac);
    } else {
        auto ld = [&](int line, int p) { return cv(line, raw(line, p)); };
        line_pipeline<T, NIN, NOUT, OP, NL, NTH, false>(ld, st, buf, twi, twop, kfac);
    }
}

template <typename T, typename TI, int NIN, int NOUT, int OP>
__global__ __launch_bounds__((rg_threads<T, NIN, NOUT, true>())) void kr_yline(
    const TI* __restrict__ in, T* __restrict__ out, int ncols, int ld_in, int ld_out, long ps_in, long ps_out,
    const cx<T>* __restrict__ twin_g, const cx<T>* __restrict__ twout_g, T kfac) {
    constexpr int NL = rg_lines<T, NIN, NOUT, true>();
    constexpr int NTH = rg_threads<T, NIN, NOUT, true>();
    constexpr int LS = rg_ls(cmax(NIN, NOUT));
    __shared__ cx<T> buf[NL * LS];
    constexpr bool TWG = tw_global<NIN, NOUT>();
    __shared__ cx<T> twl[TWG ? 1 : NIN];
    __shared__ cx<T> twl_o[(TWG || NIN == NOUT) ? 1 : NOUT];
    const cx<T>* twi = twin_g;
    const cx<T>* twop = twout_g;
    if constexpr (!TWG) {
        load_twr(twl, twin_g, NIN);
        if (NIN != NOUT) load_twr(twl_o, twout_g, NOUT);
        twi = twl;
        twop = (NIN == NOUT) ? twl : twl_o;
    }
    const int x0 = blockIdx.x * 2 * NL;
    const TI* src = in + (long)blockIdx.y * ps_in;
    T* dst = out + (long)blockIdx.y * ps_out;
    using PI = typename pair_t<TI>::type;
    using PO = typename pair_t<T>::type;
    // row 0 of each column pair (the shift of a reduced-precision derivative)
    __shared__ PI first[SHIFT_IN<T, TI, OP> ? NL : 1];
    if constexpr (SHIFT_IN<T, TI, OP>) {
        for (int t = threadIdx.x; t < NL; t += NTH) {
            const int x = x0 + 2 * t;
            first[t] = (x < ncols) ? *reinterpret_cast<const PI*>(src + x) : PI{TI(0), TI(0)};
        }
    }
    // column pair x, x + 1 at row p (zero past the last column; the read is
    // unconditional, from column 0 when past it, so that no branch separates
    // the reads)
    auto raw = [&](int line, int p) {
        const int x = x0 + 2 * line;
        const PI v = *reinterpret_cast<const PI*>(src + (long)p * ld_in + (x < ncols ? x : 0));
        return (x < ncols) ? v : PI{TI(0), TI(0)};
    };
    auto cv = [&](int line, PI v) {
        if constexpr (SHIFT_IN<T, TI, OP>) {
            const PI v0 = first[line];
            return cx<T>{(T)(v.x - v0.x), (T)(v.y - v0.y)};
        } else {
            return cx<T>{(T)v.x, (T)v.y};
        }
    };
    // pad / unpad with kfac != 0 accumulate: out += kfac * result (the
    // transposes of the 3/2-grid maps in the adjoint step)
    auto st = [&](int line, int p, cx<T> z) {
        const int x = x0 + 2 * line;
        if (x >= ncols) return;
        PO* o = reinterpret_cast<PO*>(dst + (long)p * ld_out + x);
        if constexpr (OP == OP_PAD || OP == OP_UNPAD) {
            if (kfac != T(0)) {
                const PO a = *o;
                *o = PO{a.x + kfac * z.x, a.y + kfac * z.y};
                return;
            }
        }
        *o = PO{z.x, z.y};
    };
    __syncthreads();
    if constexpr (sizeof(TI) > sizeof(T)) {
        line_pipeline<T, NIN, NOUT, OP, NL, NTH, true>(split_load(raw, cv), st, buf, twi, twop, kfac);
    } else {
        auto ld = [&](int line, int p) { return cv(line, raw(line, p)); };
        line_pipeline<T, NIN, NOUT, OP, NL, NTH, true>(ld, st, buf, twi, twop, kfac);
    }
}

// The x r2c of NL lines (two real rows each): ld(line, p) gives the packed
// inputs (LDS_IN: read from buf itself), row(rl) the output row of local row
// rl (null: none); the first LH coefficients are stored, times scale.
template <typename T, int N, int NL, int NTH, bool LDS_IN, typename LD, typename ROW>
__device__ __forceinline__ void r2c_body(const LD& ld, const ROW& row, int LH, cx<T>* buf, const cx<T>* tw,
                                         T scale) {
    constexpr int LS = rg_ls(N);
    constexpr int SF = rg_nst(N);
    Stage<T, N, 0, -1, NL, NTH, false> f0;
    f0.load(ld);
    f0.compute(tw);
    if constexpr (LDS_IN) __syncthreads();
    f0.template write<LS>(buf);
    __syncthreads();
    middle_stages<T, N, -1, NL, NTH, false, LS>(buf, tw);
    Stage<T, N, SF - 1, -1, NL, NTH, false> fl;
    fl.template read<LS>(buf);
    fl.compute(tw);
    __syncthreads();
    fl.store([&](int line, int p, cx<T> z) { buf[line * LS + rg_pad(p)] = z; });
    __syncthreads();
    for (int e = threadIdx.x; e < 2 * NL * LH; e += NTH) {
        const int rl = e / LH, k = e - rl * LH;
        cx<T>* o = row(rl);
        if (!o) continue;
        cx<T> r = {T(0), T(0)};
        if (2 * k != N) {
            const cx<T> zk = buf[(rl >> 1) * LS + rg_pad(k)];
            const cx<T> zc = buf[(rl >> 1) * LS + rg_pad((N - k) % N)];
            if ((rl & 1) == 0) r = {(zk.x + zc.x) * T(0.5), (zk.y - zc.y) * T(0.5)};
            else r = {(zk.y + zc.y) * T(0.5), -(zk.x - zc.x) * T(0.5)};
            r.x *= scale;
            r.y *= scale;
        }
        o[k] = r;
    }
}

// x r2c of fftn_mpi (FFTW r2c: k = 0..N/2; the x Nyquist is zeroed as in
// fftn_mpi), two real rows per complex line, split by Hermitian symmetry.
template <typename T, typename TI, int N>
__global__ __launch_bounds__((rg_threads<T, N, N, false>())) void kr_x_r2c(const TI* __restrict__ in,
                                                                         cx<T>* __restrict__ out, long nrows,
                                                                         int ld_in, int ldc, int kst,
                                                                         const cx<T>* __restrict__ tw_g, T scale) {
    constexpr int NL = rg_lines<T, N, N, false>();
    constexpr int NTH = rg_threads<T, N, N, false>();
    constexpr int LS = rg_ls(N);
    const int LH = kst;  // stored columns (at most N / 2 + 1)
    __shared__ cx<T> buf[NL * LS];
    __shared__ cx<T> tw[N];
    load_twr(tw, tw_g, N);
    const long row0 = (long)blockIdx.x * 2 * NL;
    __syncthreads();
    r2c_body<T, N, NL, NTH, false>(
        [&](int line, int p) {
            const long ra = row0 + 2 * line, rb = ra + 1;
            return cx<T>{(ra < nrows) ? (T)in[ra * ld_in + p] : T(0), (rb < nrows) ? (T)in[rb * ld_in + p] : T(0)};
        },
        [&](int rl) { return (row0 + rl < nrows) ? out + (row0 + rl) * ldc : (cx<T>*)nullptr; }, LH, buf, tw, scale);
}

// x c2r of ifftn_mpi (FFTW c2r: imaginary parts of the DC and Nyquist
// coefficients ignored); DERIV applies i kx (kx Nyquist = 0) first.
template <typename T, int N, int DERIV>
__global__ __launch_bounds__((rg_threads<T, N, N, false>())) void kr_x_c2r(const cx<T>* __restrict__ in,
                                                                         T* __restrict__ out, long nrows, int ldc,
                                                                         int kst, int ld_out,
                                                                         const cx<T>* __restrict__ tw_g, T kfac) {
    constexpr int NL = rg_lines<T, N, N, false>();
    constexpr int NTH = rg_threads<T, N, N, false>();
    constexpr int LS = rg_ls(N);
    constexpr int SI = rg_nst(N);
    __shared__ cx<T> buf[NL * LS];
    __shared__ cx<T> tw[N];
    load_twr(tw, tw_g, N);
    const long row0 = (long)blockIdx.x * 2 * NL;
    __syncthreads();
    Stage<T, N, 0, +1, NL, NTH, false> i0;
    i0.load([&](int line, int k) {
        const int kk = (2 * k <= N) ? k : N - k;
        const long ga = row0 + 2 * line, gb = ga + 1;
        cx<T> a = (ga < nrows && kk < kst) ? in[ga * ldc + kk] : cx<T>{T(0), T(0)};
        cx<T> b = (gb < nrows && kk < kst) ? in[gb * ldc + kk] : cx<T>{T(0), T(0)};
        if (kk == 0 || 2 * kk == N) {
            a.y = T(0);
            b.y = T(0);
        }
        if constexpr (DERIV) {
            const T f = (2 * kk == N) ? T(0) : (T)kk * kfac;
            a = {-f * a.y, f * a.x};
            b = {-f * b.y, f * b.x};
        }
        if (2 * k > N) {
            a.y = -a.y;
            b.y = -b.y;
        }
        return cx<T>{a.x - b.y, a.y + b.x};
    });
    i0.compute(tw);
    i0.template write<LS>(buf);
    __syncthreads();
    middle_stages<T, N, +1, NL, NTH, false, LS>(buf, tw);
    Stage<T, N, SI - 1, +1, NL, NTH, false> il;
    il.template read<LS>(buf);
    il.compute(tw);
    il.store([&](int line, int p, cx<T> z) {
        const long ra = row0 + 2 * line, rb = ra + 1;
        if (ra < nrows) out[ra * ld_out + p] = z.x;
        if (rb < nrows) out[rb * ld_out + p] = z.y;
    });
}

// complex y-lines of the half spectrum (ncols columns of row stride ld, N
// rows per plane); SP_FWD stores and SP_MULINV reads only the rows with
// min(p, N - p) < mc
template <typename T, int N, int OP>
__global__ __launch_bounds__((rg_threads<T, N, N, true>())) void kr_yspec(const cx<T>* __restrict__ in,
                                                                        cx<T>* __restrict__ out, int ncols, int ld,
                                                                        int ldG, int mc, long ps,
                                                                        const cx<T>* __restrict__ tw_g,
                                                                        const double* __restrict__ G) {
    // one complex column per line: 8 (FP64) / 16 (FP32) columns = 128 bytes
    constexpr int NL = rg_lines<T, N, N, true>();
    constexpr int NTH = rg_threads<T, N, N, true>();
    constexpr int LS = rg_ls(N);
    constexpr int S = rg_nst(N);
    __shared__ cx<T> buf[NL * LS];
    __shared__ cx<T> tw[N];
    load_twr(tw, tw_g, N);
    const int x0 = blockIdx.x * NL;
    const cx<T>* src = in + (long)blockIdx.y * ps;
    cx<T>* dst = out + (long)blockIdx.y * ps;
    const double* Gp = G;
    auto keep = [&](int p) { return p < mc || p > N - mc; };
    auto gload = [&](int line, int p) {
        const int x = x0 + line;
        return (x < ncols) ? src[(long)p * ld + x] : cx<T>{T(0), T(0)};
    };
    auto gstore = [&](int line, int p, cx<T> z) {
        const int x = x0 + line;
        if (x < ncols) dst[(long)p * ld + x] = z;
    };
    __syncthreads();
    if constexpr (OP == SP_FWD || OP == SP_FILT) {
        Stage<T, N, 0, -1, NL, NTH, true> f0;
        f0.load(gload);
        f0.compute(tw);
        f0.template write<LS>(buf);
        __syncthreads();
        middle_stages<T, N, -1, NL, NTH, true, LS>(buf, tw);
        Stage<T, N, S - 1, -1, NL, NTH, true> fl;
        fl.template read<LS>(buf);
        fl.compute(tw);
        if constexpr (OP == SP_FWD) {
            fl.store([&](int line, int p, cx<T> z) {
                if (keep(p)) gstore(line, p, (2 * p == N) ? cx<T>{T(0), T(0)} : z);
            });
            return;
        } else {
            // Nyquist row zeroed, times G, then the inverse from registers
            Stage<T, N, 0, +1, NL, NTH, true> i0;
#pragma unroll
            for (int q = 0; q < decltype(i0)::PB; ++q) {
                int line, b;
                if (!decltype(i0)::bf(q, line, b)) continue;
                const int x = x0 + line;
#pragma unroll
                for (int r = 0; r < decltype(i0)::R; ++r) {
                    const int p = b + r * decltype(i0)::NB;
                    const T g = (x < ncols && 2 * p != N) ? (T)Gp[(long)p * ldG + x] : T(0);
                    i0.v[q][r] = {fl.v[q][r].x * g, fl.v[q][r].y * g};
                }
            }
            i0.compute(tw);
            __syncthreads();
            i0.template write<LS>(buf);
        }
    } else {
        Stage<T, N, 0, +1, NL, NTH, true> i0;
        if constexpr (OP == SP_MULINV) {
            i0.load([&](int line, int p) {
                const int x = x0 + line;
                if (x >= ncols || !keep(p)) return cx<T>{T(0), T(0)};
                const cx<T> z = src[(long)p * ld + x];
                const T g = (T)Gp[(long)p * ldG + x];
                return cx<T>{z.x * g, z.y * g};
            });
        } else {
            i0.load(gload);
        }
        i0.compute(tw);
        i0.template write<LS>(buf);
    }
    __syncthreads();
    middle_stages<T, N, +1, NL, NTH, true, LS>(buf, tw);
    Stage<T, N, S - 1, +1, NL, NTH, true> il;
    il.template read<LS>(buf);
    il.compute(tw);
    il.store(gstore);
}

// Workgroups are dealt to the 8 XCDs round robin (workgroup b to XCD b % 8),
// each XCD with its own L2.  Kernels whose NC consecutive work items read the
// same inputs place them on one XCD: work item (group rg, item c) runs as
// workgroup (rg % 8) + 8 (c + NC (rg / 8)), so the shared inputs come from
// that XCD's L2.  Launch NC * 8 * ceil(ngroups / 8) workgroups; those past
// the last group return at once.
constexpr int N_XCD = 8;
template <int NC>
__device__ __forceinline__ void xcd_item(long& rg, int& c) {
    const long b = blockIdx.x, q = b / N_XCD;
    c = (int)(q % NC);
    rg = (q / NC) * N_XCD + b % N_XCD;
}
template <int NC>
inline long xcd_blocks(long ngroups) {
    return (long)NC * N_XCD * ((ngroups + N_XCD - 1) / N_XCD);
}

// ---- first passes of pads / unpads whose input is a pointwise function ----
// The 3/2-grid products of convec (cross_product, 108_convec.f90:103-202)
// and of the scalar advection (010_scalars.f90:466-484) feed an unpad; the
// vorticity (108_convec.f90:44-69) feeds a pad.  Each fused kernel evaluates
// the function for its block's lines straight into the line buffer (every
// value as the separate kernel k_cross / k_scalar_adv / k_vort computes and
// rounds it), then runs the x pass from LDS: the product or vorticity field
// is never written to and re-read from memory.
//
// A block takes one component (consecutive blocks of an XCD take the
// components of the same rows, so the shared inputs are re-read from cache)
// and one group of FusedWork: line l is the row pair of its column on level
// k0 + 1 + l.  The evaluation is branch
// free (every input loaded, the wall-level formula chosen by a select), so
// the loads of all of a thread's elements can be in flight together.
template <typename T, int COMP>
__device__ __forceinline__ double fused_value(const ConvIn<T>& a, int k, long i, long P2, int) {
    if constexpr (COMP == 3) {  // scalar advection
        const T ub = a.ub[i], vb = a.vb[i], wb = a.wb[i], wp = a.wb[i + P2];
        const T sx = a.b1[i], sy = a.b2[i], sz = a.b3[i], szp = a.b3[i + P2];
        const double ust = a.ust[k], vst = a.vst[k];
        const double w1 = (ub + ust) * sx + (vb + vst) * sy + (0.5 * wp) * szp;
        const double wg = (ub + ust) * sx + (vb + vst) * sy + ((double)wb * sz + (double)wp * szp) / 2.0;
        return (k == 1) ? w1 : wg;
    } else if constexpr (COMP == 0) {
        const T vb = a.vb[i], b3 = a.b3[i], wp = a.wb[i + P2], b2p = a.b2[i + P2], wb = a.wb[i], b2 = a.b2[i];
        const double vst = a.vst[k];
        const double w1 = (vb + vst) * (-b3) + 0.5 * wp * b2p;
        const double wg = (vb + vst) * (-b3) + 0.5 * (wp * (b2p) + wb * (b2));
        return (k == 1) ? w1 : wg;
    } else if constexpr (COMP == 1) {
        const T ub = a.ub[i], b3 = a.b3[i], wp = a.wb[i + P2], b1p = a.b1[i + P2], wb = a.wb[i], b1 = a.b1[i];
        const double ust = a.ust[k];
        const double w1 = (ub + ust) * (b3) + 0.5 * wp * (-b1p);
        const double wg = (ub + ust) * (b3) + 0.5 * (wp * (-b1p) + wb * (-b1));
        return (k == 1) ? w1 : wg;
    } else {
        const T ub = a.ub[i], um = a.ub[i - P2], b2 = a.b2[i], vb = a.vb[i], vm = a.vb[i - P2], b1 = a.b1[i];
        const double wg = 0.5 * ((ub + um + a.ust[k] + a.ust[k - 1]) * (-b2) + (vb + vm + a.vst[k] + a.vst[k - 1]) * (b1));
        return (k == 1) ? 0.0 : wg;
    }
}
template <typename T, int COMP>
__device__ __forceinline__ double fused_value(const VortIn<T>& g, int k, long i, long P, int nz) {
    if constexpr (COMP == 2) {
        return (double)g.vx[i] - (double)g.uy[i];
    } else {
        // neighbours clamped into the column (levels 1..nz are evaluated;
        // level nz + 1 does not exist)
        const long ip = (k < nz) ? i + P : i, im = i - P;
        if constexpr (COMP == 0) {
            const double wy = g.wy[i], wyp = g.wy[ip], wym = g.wy[im], vz = g.vz[i];
            const double w1 = 0.5 * (wy + wyp) - vz, wn = 0.5 * (wy + wym) - vz, wg = wy - vz;
            return (k == 1) ? w1 : ((k == nz) ? wn : wg);
        } else {
            const double wx = g.wx[i], wxp = g.wx[ip], wxm = g.wx[im], uz = g.uz[i];
            const double w1 = uz - 0.5 * (wx + wxp), wn = uz - 0.5 * (wx + wxm), wg = uz - wx;
            return (k == 1) ? w1 : ((k == nz) ? wn : wg);
        }
    }
}

// inputs of a block's NL lines into the line buffer (component COMP): line l
// is the row pair (rows 2 col, 2 col + 1) of level 1 + k0 + l; lines past the
// last of the nlz levels read the last level and are not stored
template <typename T, int NIN, int NL, int NTH, int LS, int COMP, typename IN>
__device__ __forceinline__ void fused_inputs(const IN& a, cx<T>* buf, int k0, int col, int nlz, long plane_in,
                                             int nz) {
    static_assert(NIN % 64 == 0 && NTH % 64 == 0, "a wave's run of positions lies in one line");
    constexpr int PER = (NL * NIN + NTH - 1) / NTH;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int e = (int)threadIdx.x + j * NTH;
        if (NL * NIN % NTH != 0 && e >= NL * NIN) break;
        // a wave's run of positions lies in one line: the line and its level
        // are wave-uniform (scalar arithmetic)
        const int l = __builtin_amdgcn_readfirstlane(e / NIN), p = e - l * NIN;
        const int kl = k0 + l;
        const int k = 1 + ((kl < nlz) ? kl : nlz - 1);
        const long ia = (long)k * plane_in + (long)(2 * col) * NIN + p;
        const double va = fused_value<T, COMP>(a, k, ia, plane_in, nz);
        const double vb = fused_value<T, COMP>(a, k, ia + NIN, plane_in, nz);
        buf[l * LS + rg_pad(p)] = (kl < nlz) ? cx<T>{(T)va, (T)vb} : cx<T>{T(0), T(0)};
    }
}

// Work of the fused kernels: a group is NL consecutive levels (a chunk) of
// one row pair (a column), so the level-neighbour rows that the products
// read (w, the vorticity at k + 1; u, v at k - 1; the vorticity's w
// gradients at k +- 1) are rows of the same group.  The chunks of a column
// are consecutive groups on one XCD (xcd_item), so the rows at the chunk
// boundaries come from that XCD's L2.  Columns are padded to a multiple of
// N_XCD; groups of padding columns return at once.
template <int NL>
struct FusedWork {
    int ycols, nlz, nch;
    __host__ __device__ FusedWork(long nrp, int ny_in)
        : ycols(ny_in / 2), nlz((int)(nrp / (ny_in / 2))), nch((nlz + NL - 1) / NL) {}
    __host__ long groups() const { return (long)nch * (((long)ycols + N_XCD - 1) / N_XCD) * N_XCD; }
    __device__ void item(long rg, int& col, int& k0) const {
        const long q = rg / N_XCD;
        col = (int)((q / nch) * N_XCD + rg % N_XCD);
        k0 = (int)(q % nch) * NL;
    }
};

// NC components per row pair; comp0 is the first component's index (3 for
// the scalar advection).  ny_in rows per level, plane_in values per plane;
// nrp = levels * ny_in / 2 row pairs.
template <typename T, int NIN, int NOUT, int OP, int NC, typename IN>
__global__ __launch_bounds__((rg_threads<T, NIN, NOUT, false>())) void kr_fused_x(
    IN a, int comp0, T* o0, T* o1, T* o2, long nrp, int ny_in, long plane_in, int nz,
    const cx<T>* __restrict__ twin_g, const cx<T>* __restrict__ twout_g) {
    constexpr int NL = rg_lines<T, NIN, NOUT, false>();
    constexpr int NTH = rg_threads<T, NIN, NOUT, false>();
    constexpr int LS = rg_ls(cmax(NIN, NOUT));
    constexpr bool TWG = tw_global<NIN, NOUT>();
    __shared__ cx<T> buf[NL * LS];
    __shared__ cx<T> twl[TWG ? 1 : NIN];
    __shared__ cx<T> twl_o[(TWG || NIN == NOUT) ? 1 : NOUT];
    const cx<T>* twi = twin_g;
    const cx<T>* twop = twout_g;
    if constexpr (!TWG) {
        load_twr(twl, twin_g, NIN);
        if (NIN != NOUT) load_twr(twl_o, twout_g, NOUT);
        twi = twl;
        twop = (NIN == NOUT) ? twl : twl_o;
    }
    long rg;
    int cb;
    xcd_item<NC>(rg, cb);
    const FusedWork<NL> wk(nrp, ny_in);
    int col, k0;
    wk.item(rg, col, k0);
    if (col >= wk.ycols) return;  // whole workgroup: a padding column
    const int comp = comp0 + cb;
    if (comp == 0) fused_inputs<T, NIN, NL, NTH, LS, 0>(a, buf, k0, col, wk.nlz, plane_in, nz);
    else if (comp == 1) fused_inputs<T, NIN, NL, NTH, LS, 1>(a, buf, k0, col, wk.nlz, plane_in, nz);
    else if (comp == 2) fused_inputs<T, NIN, NL, NTH, LS, 2>(a, buf, k0, col, wk.nlz, plane_in, nz);
    else fused_inputs<T, NIN, NL, NTH, LS, 3>(a, buf, k0, col, wk.nlz, plane_in, nz);
    __syncthreads();
    T* o = (cb == 0) ? o0 : ((cb == 1) ? o1 : o2);
    auto ld = [&](int line, int p) { return buf[line * LS + rg_pad(p)]; };
    auto st = [&](int line, int p, cx<T> z) {
        const int kl = k0 + line;
        if (kl >= wk.nlz) return;
        const long r = (long)kl * (2 * wk.ycols) + 2 * col;  // first row of the pair
        o[r * NOUT + p] = z.x;
        o[(r + 1) * NOUT + p] = z.y;
    };
    line_pipeline<T, NIN, NOUT, OP, NL, NTH, false, true>(ld, st, buf, twi, twop, T(0));
}

// ---- lagrange_Sdep: filtered fields contracted as they are transformed back ----
// (106_03_lagrange_Sdep.f90:132-331).  A workgroup takes one row (level k,
// row y) of all 21 test-filtered fields: the c2r of each row as a complex
// transform of half the length, the 21 real rows to LDS, then the
// contraction at each of the N points.  The 21 filtered fields are never
// written and re-read.
//  SCALE 2: L_ij M_ij and M_ij M_ij of the 2 delta filter to lm, mm (FP64).
//  SCALE 4: Q_ij N_ij and N_ij N_ij of the 4 delta filter, then with lm, mm
//           the relaxation of F_LM .. F_NN and Cs^2 (lag_update).
// The half-length c2r: with X the half spectrum (X[k] = 0 for k >= kst) of
// a real row x of length N, H = N/2 and E_k = exp(2 pi i k / N),
//   Z[k] = (X[k] + conj X[H-k]) + i E_k (X[k] - conj X[H-k]),  k < H,
// has the unnormalised inverse transform z[m] = x[2m] + i x[2m+1].  One row
// per workgroup needs 21 lines of H points of LDS (about 46 KB in FP64), so
// three workgroups share a CU.
constexpr int LAG_NF = 21;
template <typename T, int N>
constexpr int lag_threads() {
    return round64(LAG_NF * rg_maxbf(N / 2));
}

template <typename T, int N, int SCALE>
__global__ __launch_bounds__((lag_threads<T, N>())) void kr_lag_contract(const cx<T>* __restrict__ spt, int ldc,
                                                                        int kst, int nrows, int ny, long plane,
                                                                        double* lm, double* mm, LagUpdate u,
                                                                        double cst,
                                                                        const cx<T>* __restrict__ tw_g) {
    constexpr int H = N / 2;
    constexpr int NL = LAG_NF;
    constexpr int NTH = lag_threads<T, N>();
    constexpr int LS = rg_ls(H);
    static_assert(rg_nst(H) == 2, "two stages: load, LDS exchange, store");
    static_assert(LAG_NF * N * (int)sizeof(T) <= NL * LS * (int)sizeof(cx<T>), "rows fit the line buffer");
    __shared__ cx<T> buf[NL * LS];
    __shared__ cx<T> tw[H];  // exp(-2 pi i j / H)
    for (int m = threadIdx.x; m < H; m += NTH) tw[m] = tw_g[2 * m];
    const int row = blockIdx.x;  // level 1 + row / ny, row y = row % ny
    __syncthreads();
    Stage<T, H, 0, +1, NL, NTH, false> i0;
    i0.load([&](int line, int k) {
        const cx<T>* X = spt + ((long)line * nrows + row) * ldc;
        const int j = H - k;
        cx<T> a = (k < kst) ? X[k] : cx<T>{T(0), T(0)};
        cx<T> b = (j < kst) ? X[j] : cx<T>{T(0), T(0)};
        if (k == 0) {  // X[0] and X[H] are real
            a.y = T(0);
            b.y = T(0);
        }
        const cx<T> e = tw_g[k];  // conj(E_k)
        const T sx = a.x + b.x, sy = a.y - b.y;  // X[k] + conj X[H-k]
        const T dx = a.x - b.x, dy = a.y + b.y;  // X[k] - conj X[H-k]
        const T px = dx * e.x + dy * e.y, py = dy * e.x - dx * e.y;  // E_k (X[k] - conj X[H-k])
        return cx<T>{sx - py, sy + px};
    });
    i0.compute(tw);
    i0.template write<LS>(buf);
    __syncthreads();
    Stage<T, H, 1, +1, NL, NTH, false> il;
    il.template read<LS>(buf);
    il.compute(tw);
    __syncthreads();  // every line read: the buffer now takes the real rows
    T* rows = reinterpret_cast<T*>(buf);  // [field][x]
    il.store([&](int line, int m, cx<T> z) {
        rows[line * N + 2 * m] = z.x;
        rows[line * N + 2 * m + 1] = z.y;
    });
    __syncthreads();
    const int k = 1 + row / ny;
    const long row0 = (long)k * plane + (long)(row - (k - 1) * ny) * N;  // field index of the row's first point
    for (int x = threadIdx.x; x < N; x += NTH) {
        const long i = row0 + x;
        auto v = [&](int f) { return (double)rows[f * N + x]; };
        double a, b;
        if constexpr (SCALE == 2) {
            lag_scale(v, cst, 4.0, a, b);
            lm[i] = a;
            mm[i] = b;
        } else {
            lag_scale(v, cst, 16.0, a, b);
            lag_update(u, i, k, lm[i], mm[i], a, b);
        }
    }
}

template <typename T>
const cx<T>* twp(const Twiddle& t);
template <>
const cx<double>* twp<double>(const Twiddle& t) { return t.d; }
template <>
const cx<float>* twp<float>(const Twiddle& t) { return t.f; }

// ---- dispatch --------------------------------------------------------------
template <typename T, typename TI, int NIN, int NOUT, int OP>
void launch_x(hipStream_t s, const TI* in, T* out, long nrows, int ld_in, int ld_out, const Twiddle& tin,
              const Twiddle& tout, double kfac) {
    constexpr int NL = rg_lines<T, NIN, NOUT, false>();
    constexpr int NTH = rg_threads<T, NIN, NOUT, false>();
    const long nb = (nrows + 2 * NL - 1) / (2 * NL);
    hipLaunchKernelGGL((kr_xline<T, TI, NIN, NOUT, OP>), dim3((unsigned)nb), dim3(NTH), 0, s, in, out, nrows, ld_in,
                       ld_out, twp<T>(tin), twp<T>(tout), (T)kfac);
}

template <typename T, typename TI, int NIN, int NOUT, int OP>
void launch_y(hipStream_t s, const TI* in, T* out, int nplanes, int ncols, int ld_in, int ld_out, long ps_in,
              long ps_out, const Twiddle& tin, const Twiddle& tout, double kfac) {
    constexpr int NL = rg_lines<T, NIN, NOUT, true>();
    constexpr int NTH = rg_threads<T, NIN, NOUT, true>();
    const int nbx = (ncols + 2 * NL - 1) / (2 * NL);
    hipLaunchKernelGGL((kr_yline<T, TI, NIN, NOUT, OP>), dim3(nbx, nplanes), dim3(NTH), 0, s, in, out, ncols, ld_in,
                       ld_out, ps_in, ps_out, twp<T>(tin), twp<T>(tout), (T)kfac);
}

#define LES_SIZES(X) X(64) X(128) X(256) X(512)

// the y-kernels read column pairs as one aligned vector: even column count,
// row stride and plane stride
template <typename TI, typename T>
bool pairs_aligned(const TI* in, const T* out, int ncols, int ld_in, int ld_out, long ps_in, long ps_out) {
    return ncols % 2 == 0 && ld_in % 2 == 0 && ld_out % 2 == 0 && ps_in % 2 == 0 && ps_out % 2 == 0 &&
           (uintptr_t)in % (2 * sizeof(TI)) == 0 && (uintptr_t)out % (2 * sizeof(T)) == 0;
}

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
    if (!pairs_aligned(in, out, ncols, ld_in, ld_out, ps_in, ps_out)) return false;
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
                double scale, int kst) {
    constexpr int NL = rg_lines<T, N, N, false>();
    constexpr int NTH = rg_threads<T, N, N, false>();
    const long nb = (nrows + 2 * NL - 1) / (2 * NL);
    hipLaunchKernelGGL((kr_x_r2c<T, TI, N>), dim3((unsigned)nb), dim3(NTH), 0, s, in, out, nrows, ld_in, ldc, kst,
                       twp<T>(tw), (T)scale);
}

template <typename T, int N, int D>
void launch_c2r(hipStream_t s, const cx<T>* in, T* out, long nrows, int ldc, int ld_out, const Twiddle& tw,
                double kfac, int kst) {
    constexpr int NL = rg_lines<T, N, N, false>();
    constexpr int NTH = rg_threads<T, N, N, false>();
    const long nb = (nrows + 2 * NL - 1) / (2 * NL);
    hipLaunchKernelGGL((kr_x_c2r<T, N, D>), dim3((unsigned)nb), dim3(NTH), 0, s, in, out, nrows, ldc, kst, ld_out,
                       twp<T>(tw), (T)kfac);
}

template <typename T, int N, int OP>
void launch_spec(hipStream_t s, const cx<T>* in, cx<T>* out, int nplanes, int ncols, long ps, const Twiddle& tw,
                 const double* G, int ld, int ldG, int mc) {
    constexpr int NL = rg_lines<T, N, N, true>();
    constexpr int NTH = rg_threads<T, N, N, true>();
    const int nbx = (ncols + NL - 1) / NL;
    hipLaunchKernelGGL((kr_yspec<T, N, OP>), dim3(nbx, nplanes), dim3(NTH), 0, s, in, out, ncols, ld, ldG, mc, ps,
                       twp<T>(tw), G);
}

template <typename T, typename TI>
bool r2c_dispatch(hipStream_t s, const TI* in, cx<T>* out, long nrows, int ld_in, int ldc, const Twiddle& tw,
                  double scale, int kst) {
#define RCASE(NN_) \
    if (tw.n == NN_) { launch_r2c<T, TI, NN_>(s, in, out, nrows, ld_in, ldc, tw, scale, kst); return true; }
    LES_SIZES(RCASE)
#undef RCASE
    return false;
}

template <typename T>
bool c2r_dispatch(hipStream_t s, const cx<T>* in, T* out, long nrows, int ldc, int ld_out, const Twiddle& tw,
                  int deriv, double kfac, int kst) {
#define CCASE(NN_)                                                                               \
    if (tw.n == NN_) {                                                                           \
        if (deriv) launch_c2r<T, NN_, 1>(s, in, out, nrows, ldc, ld_out, tw, kfac, kst);         \
        else launch_c2r<T, NN_, 0>(s, in, out, nrows, ldc, ld_out, tw, kfac, kst);               \
        return true;                                                                           \
    }
    LES_SIZES(CCASE)
#undef CCASE
    return false;
}

template <typename T>
bool spec_dispatch(hipStream_t s, const cx<T>* in, cx<T>* out, int nplanes, int ncols, long ps, const Twiddle& tw,
                   SpecOp op, const double* G, int ld, int ldG, int mc) {
#define SCASE(NN_)                                                                                          \
    if (tw.n == NN_) {                                                                                      \
        switch (op) {                                                                                     \
            case SP_FWD: launch_spec<T, NN_, SP_FWD>(s, in, out, nplanes, ncols, ps, tw, G, ld, ldG, mc); break;  \
            case SP_INV: launch_spec<T, NN_, SP_INV>(s, in, out, nplanes, ncols, ps, tw, G, ld, ldG, mc); break;  \
            case SP_FILT: launch_spec<T, NN_, SP_FILT>(s, in, out, nplanes, ncols, ps, tw, G, ld, ldG, mc); break; \
            default: launch_spec<T, NN_, SP_MULINV>(s, in, out, nplanes, ncols, ps, tw, G, ld, ldG, mc); break;   \
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
              double scale, int kst) {
    return r2c_dispatch<T, TI>(s, in, out, nrows, ld_in, ldc, tw, scale, kst);
}
template <typename T>
bool x_c2r_ct(hipStream_t s, const cx<T>* in, T* out, long nrows, int ldc, int ld_out, const Twiddle& tw, int deriv,
              double kfac, int kst) {
    return c2r_dispatch<T>(s, in, out, nrows, ldc, ld_out, tw, deriv, kfac, kst);
}
template <typename T>
bool yline_spec_ct(hipStream_t s, const cx<T>* in, cx<T>* out, int nplanes, int ncols, long ps, const Twiddle& tw,
                   SpecOp op, const double* G, int ld, int ldG, int mc) {
    return spec_dispatch<T>(s, in, out, nplanes, ncols, ps, tw, op, G, ld, ldG, mc);
}

template <typename T, int NIN, int NOUT, int OP, int NC, typename IN>
void launch_fused(hipStream_t s, const IN& a, int comp0, T* o0, T* o1, T* o2, long nrp, int ny_in, long plane_in,
                  int nz, const Twiddle& tin, const Twiddle& tout) {
    constexpr int NL = rg_lines<T, NIN, NOUT, false>();
    constexpr int NTH = rg_threads<T, NIN, NOUT, false>();
    const long nb = xcd_blocks<NC>(FusedWork<NL>(nrp, ny_in).groups());
    hipLaunchKernelGGL((kr_fused_x<T, NIN, NOUT, OP, NC, IN>), dim3((unsigned)nb), dim3(NTH), 0, s, a, comp0, o0, o1,
                       o2, nrp, ny_in, plane_in, nz, twp<T>(tin), twp<T>(tout));
}

template <typename T>
bool cross_unpad_x(hipStream_t s, const ConvIn<T>& a, int scalar, T* o0, T* o1, T* o2, int nz, int ny2,
                   const Twiddle& tin, const Twiddle& tout) {
    const long plane2 = (long)tin.n * ny2;
    const long nrp = (long)(nz - 1) * ny2 / 2;
    if (ny2 % 2 != 0 || tin.n * 2 != tout.n * 3 || (tout.n != 128 && tout.n != 256 && tout.n != 512)) return false;
    // six padded inputs (levels 1..nz) read, one or three x-unpadded outputs
    // (levels 1..nz-1) written
    FamScope fs(s, scalar ? FAM_SCALAR_UNPAD_X : FAM_CROSS_UNPAD_X,
                sizeof(T) * (6.0 * nz * plane2 + (scalar ? 1.0 : 3.0) * (nz - 1) * ny2 * (double)tout.n));
#define FCASE(NN_)                                                                                          \
    if (tin.n == 3 * NN_ / 2 && tout.n == NN_ && ny2 % 2 == 0) {                                              \
        if (scalar)                                                                                       \
            launch_fused<T, 3 * NN_ / 2, NN_, OP_UNPAD, 1>(s, a, 3, o0, o1, o2, nrp, ny2, plane2, nz, tin, tout); \
        else                                                                                              \
            launch_fused<T, 3 * NN_ / 2, NN_, OP_UNPAD, 3>(s, a, 0, o0, o1, o2, nrp, ny2, plane2, nz, tin, tout); \
        return true;                                                                                      \
    }
    FCASE(128) FCASE(256) FCASE(512)  // (96-point rows would split waves across lines)
#undef FCASE
    return false;
}

template <typename T>
bool vort_pad_x(hipStream_t s, const VortIn<T>& g, T* o0, T* o1, T* o2, int nz, int ny, const Twiddle& tin,
                const Twiddle& tout) {
    const long plane = (long)tin.n * ny;
    const long nrp = (long)nz * ny / 2;
#define VOK(NN_) || (tin.n == NN_ && tout.n == 3 * NN_ / 2)
    if (ny % 2 != 0 || !(false LES_SIZES(VOK))) return false;
#undef VOK
    // six velocity gradients read, three x-padded vorticity components written
    FamScope fs(s, FAM_VORT_PAD_X, sizeof(T) * (6.0 * nz * plane + 3.0 * nz * ny * (double)tout.n));
#define VCASE(NN_)                                                                                           \
    if (tin.n == NN_ && tout.n == 3 * NN_ / 2 && ny % 2 == 0) {                                                \
        launch_fused<T, NN_, 3 * NN_ / 2, OP_PAD, 3>(s, g, 0, o0, o1, o2, nrp, ny, plane, nz, tin, tout);  \
        return true;                                                                                       \
    }
    LES_SIZES(VCASE)
#undef VCASE
    return false;
}
template <typename T>
bool lag_contract_x(hipStream_t s, int scale, const cx<T>* spt, int ldc, int kst, int nz, int ny, double* lm,
                    double* mm, const LagUpdate& u, double cst, const Twiddle& tw) {
    const int nrows = nz * ny;
    if (tw.n != 64 && tw.n != 128 && tw.n != 256) return false;
    // the 21 filtered spectra read; L M, M M (2 delta) or the F_* update and
    // Cs_opt2 (4 delta) written, FP64
    FamScope fs(s, FAM_LAG_CONTRACT_X,
                21.0 * nz * ny * ldc * 2.0 * sizeof(T) + (scale == 2 ? 2.0 : 7.0) * nz * ny * (double)tw.n * 8.0);
#define LCCASE(NN_)                                                                                         \
    if (tw.n == NN_) {                                                                                      \
        const dim3 g((unsigned)nrows), b(lag_threads<T, NN_>());                                            \
        if (scale == 2)                                                                                     \
            hipLaunchKernelGGL((kr_lag_contract<T, NN_, 2>), g, b, 0, s, spt, ldc, kst, nrows, ny,          \
                               (long)NN_ * ny, lm, mm, u, cst, twp<T>(tw));                                 \
        else                                                                                                \
            hipLaunchKernelGGL((kr_lag_contract<T, NN_, 4>), g, b, 0, s, spt, ldc, kst, nrows, ny,          \
                               (long)NN_ * ny, lm, mm, u, cst, twp<T>(tw));                                 \
        return true;                                                                                        \
    }
    LCCASE(64) LCCASE(128) LCCASE(256)
#undef LCCASE
    return false;
}
template bool lag_contract_x<double>(hipStream_t, int, const cx<double>*, int, int, int, int, double*, double*,
                                     const LagUpdate&, double, const Twiddle&);
template bool lag_contract_x<float>(hipStream_t, int, const cx<float>*, int, int, int, int, double*, double*,
                                    const LagUpdate&, double, const Twiddle&);
template bool cross_unpad_x<double>(hipStream_t, const ConvIn<double>&, int, double*, double*, double*, int, int,
                                    const Twiddle&, const Twiddle&);
template bool cross_unpad_x<float>(hipStream_t, const ConvIn<float>&, int, float*, float*, float*, int, int,
                                   const Twiddle&, const Twiddle&);
template bool vort_pad_x<double>(hipStream_t, const VortIn<double>&, double*, double*, double*, int, int,
                                 const Twiddle&, const Twiddle&);
template bool vort_pad_x<float>(hipStream_t, const VortIn<float>&, float*, float*, float*, int, int, const Twiddle&,
                                const Twiddle&);

#define INST_IO(T, TI)                                                                                       \
    template bool xline_real_ct<T, TI>(hipStream_t, const TI*, T*, long, int, int, const Twiddle&, const Twiddle&, \
                                       LineOp, double);                                                          \
    template bool yline_real_ct<T, TI>(hipStream_t, const TI*, T*, int, int, int, int, long, long, const Twiddle&, \
                                       const Twiddle&, LineOp, double);                                          \
    template bool x_r2c_ct<T, TI>(hipStream_t, const TI*, cx<T>*, long, int, int, const Twiddle&, double, int);
INST_IO(double, double)
INST_IO(float, float)
INST_IO(float, double)
#undef INST_IO
template bool x_c2r_ct<double>(hipStream_t, const cx<double>*, double*, long, int, int, const Twiddle&, int, double,
                               int);
template bool x_c2r_ct<float>(hipStream_t, const cx<float>*, float*, long, int, int, const Twiddle&, int, double,
                              int);
template bool yline_spec_ct<double>(hipStream_t, const cx<double>*, cx<double>*, int, int, long, const Twiddle&,
                                    SpecOp, const double*, int, int, int);
template bool yline_spec_ct<float>(hipStream_t, const cx<float>*, cx<float>*, int, int, long, const Twiddle&,
                                   SpecOp, const double*, int, int, int);

}  // namespace les
