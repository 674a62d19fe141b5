/*
 * ORACLE -- test infrastructure only (see orc_internal.h).
 *
 * Mixed-radix (2, 3, 4) complex FFT with FFTW's conventions, standing in for
 * the FFTW3 legacy API the reference calls (dfftw_execute_dft,
 * dfftw_execute_dft_r2c, dfftw_execute_dft_c2r; plans at 002_fft.f90:97-113).
 * FFTW is a third-party dependency absent from the reference tree with no
 * pinned version; its published definitions are restated:
 *   forward  X_k = sum_j x_j exp(-2 pi i jk/n)   (FFTW_FORWARD, sign -1)
 *   backward x_j = sum_k X_k exp(+2 pi i jk/n)   (FFTW_BACKWARD, sign +1)
 *   r2c keeps X_0..X_{n/2}; c2r assumes Hermitian symmetry and ignores the
 *   imaginary parts of X_0 and X_{n/2}.  All unnormalised.
 * Pinned in tests/test_oracle_fft.py against numpy.fft and a long-double
 * brute-force DFT.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc_internal.h"

int orc_plan_init(orc_plan* p, int n) {
    memset(p, 0, sizeof(*p));
    p->n = n;
    int m = n;
    while (m > 1) {
        int r;
#ifdef ORC_RADIX2
        /* alternative factorisation (test builds): FFTW chooses its plan at run
         * time, so the transform's summation order is not fixed by the
         * reference; this build measures the sensitivity to it */
        if (m % 2 == 0) r = 2;
#else
        if (m % 4 == 0) r = 4;
        else if (m % 2 == 0) r = 2;
#endif
        else if (m % 3 == 0) r = 3;
        else return -1;
        p->radix[p->nstage++] = r;
        m /= r;
    }
    p->tw_fwd = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    p->tw_bwd = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    for (int k = 0; k < n; ++k) {
        long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
        double cr = (double)cosl(a), si = (double)sinl(a);
        p->tw_fwd[k] = cr - I * si;
        p->tw_bwd[k] = cr + I * si;
    }
    return 0;
}

void orc_plan_free(orc_plan* p) {
    free(p->tw_fwd);
    free(p->tw_bwd);
    memset(p, 0, sizeof(*p));
}

/* Stockham autosort: stage of radix R maps x -> y with span Ns. */
void orc_fft(const orc_plan* p, cplx* x, cplx* work, int sign) {
    const int n = p->n;
    const cplx* tw = (sign < 0) ? p->tw_fwd : p->tw_bwd;
    cplx* src = x;
    cplx* dst = work;
    int Ns = 1;
    const double s3 = 0.86602540378443864676372317075294;  /* sin(2 pi/3) */
    for (int st = 0; st < p->nstage; ++st) {
        const int R = p->radix[st];
        const int nR = n / R;
        const int tstep = n / (Ns * R);
        for (int j = 0; j < nR; ++j) {
            const int k = j % Ns;
            cplx v[4];
            for (int r = 0; r < R; ++r) {
                cplx a = src[j + r * nR];
                if (r && k) a *= tw[(r * k * tstep) % n];
                v[r] = a;
            }
            cplx o[4];
            if (R == 2) {
                o[0] = v[0] + v[1];
                o[1] = v[0] - v[1];
            } else if (R == 4) {
                cplx t0 = v[0] + v[2], t1 = v[0] - v[2];
                cplx t2 = v[1] + v[3], t3 = v[1] - v[3];
                /* multiply t3 by (sign) i */
                cplx it3 = (sign < 0) ? (cimag(t3) - I * creal(t3)) : (-cimag(t3) + I * creal(t3));
                o[0] = t0 + t2;
                o[1] = t1 + it3;
                o[2] = t0 - t2;
                o[3] = t1 - it3;
            } else { /* R == 3 */
                cplx t1 = v[1] + v[2];
                cplx t2 = v[1] - v[2];
                cplx m1 = v[0] - 0.5 * t1;
                /* (sign) i sin(2pi/3) t2 */
                double ss = (sign < 0) ? -s3 : s3;
                cplx m2 = I * ss * t2;
                o[0] = v[0] + t1;
                o[1] = m1 + m2;
                o[2] = m1 - m2;
            }
            const int base = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; ++r) dst[base + r * Ns] = o[r];
        }
        cplx* t = src; src = dst; dst = t;
        Ns *= R;
    }
    if (src != x) memcpy(x, src, sizeof(cplx) * (size_t)n);
}

void orc_r2c(const orc_plan* p, const double* in, ptrdiff_t istride, cplx* out) {
    const int n = p->n;
    cplx buf[ORC_NMAX], work[ORC_NMAX];
    for (int j = 0; j < n; ++j) buf[j] = in[j * istride];
    orc_fft(p, buf, work, -1);
    for (int k = 0; k <= n / 2; ++k) out[k] = buf[k];
}

void orc_c2r(const orc_plan* p, const cplx* in, double* out, ptrdiff_t ostride) {
    const int n = p->n;
    cplx buf[ORC_NMAX], work[ORC_NMAX];
    buf[0] = creal(in[0]);
    for (int k = 1; k < (n + 1) / 2; ++k) {
        buf[k] = in[k];
        buf[n - k] = conj(in[k]);
    }
    if (n % 2 == 0) buf[n / 2] = creal(in[n / 2]);
    orc_fft(p, buf, work, +1);
    for (int j = 0; j < n; ++j) out[j * ostride] = creal(buf[j]);
}
