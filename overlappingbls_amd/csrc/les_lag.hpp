// lagrange_Sdep's contraction and relaxation at one point
// (106_03_lagrange_Sdep.f90:132-331), shared by k_lag_contract (k_point.hip)
// and the kernels that contract the test-filtered fields as they are
// transformed back (k_lines_rg.hip), so that both compute identical values.
#pragma once
#include "les_internal.hpp"

namespace les {

constexpr double LAG_CLIP = (double)1.e-24f;

__device__ __forceinline__ double lag_smag(double s11, double s12, double s13, double s22, double s23, double s33) {
    return sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33 + 2.0 * (s12 * s12 + s13 * s13 + s23 * s23)));
}

// One filter scale: v(j) returns the filtered field j at the point (0..2
// velocity, 3..8 u_i u_j in 11 12 13 22 23 33 order, 9..14 S_ij, 15..20
// |S| S_ij); f4 = 4 (2 delta, L and M) or 16 (4 delta, Q and N).  Returns
// L_ij M_ij (or Q_ij N_ij) in lm and M_ij M_ij (N_ij N_ij) in mm.
template <typename V>
__device__ __forceinline__ void lag_scale(const V& v, double cst, double f4, double& lm, double& mm) {
    const double ub = v(0), vb = v(1), wb = v(2);
    const double l11 = v(3) - ub * ub, l12 = v(4) - ub * vb, l13 = v(5) - ub * wb;
    const double l22 = v(6) - vb * vb, l23 = v(7) - vb * wb, l33 = v(8) - wb * wb;
    double sb[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) sb[m] = v(9 + m);
    const double S_bar = lag_smag(sb[0], sb[1], sb[2], sb[3], sb[4], sb[5]);
    double M[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) M[m] = cst * (v(15 + m) - f4 * S_bar * sb[m]);
    lm = l11 * M[0] + l22 * M[3] + l33 * M[5] + 2.0 * (l12 * M[1] + l13 * M[2] + l23 * M[4]);
    mm = M[0] * M[0] + M[3] * M[3] + M[5] * M[5] + 2.0 * (M[1] * M[1] + M[2] * M[2] + M[4] * M[4]);
}

__device__ __forceinline__ void lag_relax(double& F1, double& F2, double A, double B, double opftdelta,
                                          double lagran_dt) {
    double Tn = fmax(F1 * F2, LAG_CLIP);
    Tn = opftdelta * pow(Tn, -1.0 / 8.0);
    Tn = fmax(LAG_CLIP, Tn);
    const double dumfac = lagran_dt / Tn;
    const double epsi = dumfac / (1.0 + dumfac);
    F1 = epsi * A + (1.0 - epsi) * F1;
    F2 = epsi * B + (1.0 - epsi) * F2;
    F1 = fmax(LAG_CLIP, F1);
}

// The relaxation of F_LM .. F_NN and Cs^2 at point i of level k
struct LagUpdate {
    double *FLM, *FMM, *FQN, *FNN, *Cs;
    int init_lm, init_qn, nz;
    double opftdelta, lagran_dt, bexp;
};
__device__ __forceinline__ void lag_update(const LagUpdate& u, long i, int k, double LM, double MM, double QN,
                                           double NN) {
    double flm = u.FLM[i], fmm = u.FMM[i], fqn = u.FQN[i], fnn = u.FNN[i];
    if (u.init_lm) { fmm = MM; flm = 0.03 * MM; }
    lag_relax(flm, fmm, LM, MM, u.opftdelta, u.lagran_dt);
    const double cs2d = fmax(LAG_CLIP, flm / fmm);
    if (u.init_qn) { fnn = NN; fqn = 0.03 * NN; }
    lag_relax(fqn, fnn, QN, NN, u.opftdelta, u.lagran_dt);
    const double cs4d = fmax(LAG_CLIP, fqn / fnn);
    // bexp = log 2 / (log 4 - log 2) is 1 in double arithmetic, where
    // pow(x, 1) = x exactly
    double beta = (u.bexp == 1.0) ? cs4d / cs2d : pow(cs4d / cs2d, u.bexp);
    if (k == u.nz) beta = 1.0;
    const double betaclip = fmax(beta, 1.0 / (2.0 * 4.0));
    u.FLM[i] = flm; u.FMM[i] = fmm; u.FQN[i] = fqn; u.FNN[i] = fnn;
    u.Cs[i] = fmax(LAG_CLIP, cs2d / betaclip);
}

// The 2 delta (scale 2) or 4 delta (scale 4) filtered fields, as pruned half
// spectra spt (21 fields of nz ny rows, ldc per row, kst columns), transformed
// back and contracted in one pass (k_lines_rg.hip): scale 2 writes L M and
// M M to lm, mm; scale 4 reads them and applies lag_update.  False when the
// line length has no kernel (the caller then transforms and contracts apart).
template <typename T>
bool lag_contract_x(hipStream_t s, int scale, const cx<T>* spt, int ldc, int kst, int nz, int ny, double* lm,
                    double* mm, const LagUpdate& u, double cst, const Twiddle& tw);

}  // namespace les
