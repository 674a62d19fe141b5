// Transform building blocks on device fields, shared by the forward step
// (les_api.hip) and the linearised step (les_lin.hip).  Each enqueues line
// kernels on the context's stream.
#pragma once
#include "les_internal.hpp"

namespace les {

inline double* F(les_ctx* c, int id) { return c->f3[id]; }

// scratch memory (allocated for FP64) viewed in precision W
template <typename W>
inline W* scr(double* p) {
    return reinterpret_cast<W*>(p);
}
template <typename W>
inline cx<W>* scr(cx<double>* p) {
    return reinterpret_cast<cx<W>*>(p);
}

inline double kfx(const les_ctx* c) { return 2.0 * PI / c->prm.lx_tot; }
inline double kfy(const les_ctx* c) { return 2.0 * PI / c->prm.ly_tot; }

// d = ddx(f) / ddy(f) on levels lev0 .. lev0+nlev-1 (ddx / ddy, 004_derivatives.f90:4-80)
template <typename W, typename TI>
inline void ddx_f(les_ctx* c, W* d, const TI* f, int lev0, int nlev) {
    xline_real<W, TI>(c->stream, f + (size_t)lev0 * c->plane, d + (size_t)lev0 * c->plane, (long)nlev * c->ny, c->nx,
                      c->nx, c->tw_x, c->tw_x, OP_DERIV, kfx(c));
}
template <typename W, typename TI>
inline void ddy_f(les_ctx* c, W* d, const TI* f, int lev0, int nlev) {
    yline_real<W, TI>(c->stream, f + (size_t)lev0 * c->plane, d + (size_t)lev0 * c->plane, nlev, c->nx, c->nx, c->nx,
                      (long)c->plane, (long)c->plane, c->tw_y, c->tw_y, OP_DERIV, kfy(c));
}
// filter_data in place (Nyquist modes of x and y zeroed)
template <typename T>
inline void filter_f(les_ctx* c, T* f, int lev0, int nlev) {
    T* p = f + (size_t)lev0 * c->plane;
    xline_real<T>(c->stream, p, p, (long)nlev * c->ny, c->nx, c->nx, c->tw_x, c->tw_x, OP_NYQ, 0.0);
    yline_real<T>(c->stream, p, p, nlev, c->nx, c->nx, c->nx, (long)c->plane, (long)c->plane, c->tw_y, c->tw_y, OP_NYQ,
                  0.0);
}
// zero_padding of nlev planes: src (nx x ny planes) -> dst (nx2 x ny2 planes)
template <typename W, typename TI>
inline void pad_f(les_ctx* c, const TI* src, W* dst, int nlev) {
    W* tmp = scr<W>(c->work3[6]);  // >= nlev * ny * nx2
    xline_real<W, TI>(c->stream, src, tmp, (long)nlev * c->ny, c->nx, c->nx2, c->tw_x, c->tw_x2, OP_PAD, 0.0);
    yline_real<W>(c->stream, tmp, dst, nlev, c->nx2, c->nx2, c->nx2, (long)c->ny * c->nx2, (long)c->plane2, c->tw_y,
                  c->tw_y2, OP_PAD, 0.0);
}
// zero_unpadding of nlev planes: src (nx2 x ny2) -> dst (nx x ny)
template <typename W>
inline void unpad_f(les_ctx* c, const W* src, W* dst, int nlev) {
    W* tmp = scr<W>(c->work3[6]);  // >= nlev * ny2 * nx
    xline_real<W>(c->stream, src, tmp, (long)nlev * c->ny2, c->nx2, c->nx, c->tw_x2, c->tw_x, OP_UNPAD, 0.0);
    yline_real<W>(c->stream, tmp, dst, nlev, c->nx, c->nx, c->nx, (long)c->ny2 * c->nx, (long)c->plane, c->tw_y2,
                  c->tw_y, OP_UNPAD, 0.0);
}
// dst += acc * zero_padding(src) / zero_unpadding(src) (acc != 0): the
// transposes of unpad / pad in the adjoint step, accumulated by the y pass
inline void pad_acc_f(les_ctx* c, const double* src, double* dst, int nlev, double acc) {
    double* tmp = c->work3[6];
    xline_real<double>(c->stream, src, tmp, (long)nlev * c->ny, c->nx, c->nx2, c->tw_x, c->tw_x2, OP_PAD, 0.0);
    yline_real<double>(c->stream, tmp, dst, nlev, c->nx2, c->nx2, c->nx2, (long)c->ny * c->nx2, (long)c->plane2,
                       c->tw_y, c->tw_y2, OP_PAD, acc);
}
inline void unpad_acc_f(les_ctx* c, const double* src, double* dst, int nlev, double acc) {
    double* tmp = c->work3[6];
    xline_real<double>(c->stream, src, tmp, (long)nlev * c->ny2, c->nx2, c->nx, c->tw_x2, c->tw_x, OP_UNPAD, 0.0);
    yline_real<double>(c->stream, tmp, dst, nlev, c->nx, c->nx, c->nx, (long)c->ny2 * c->nx, (long)c->plane,
                       c->tw_y2, c->tw_y, OP_UNPAD, acc);
}
// test_filter of nplanes consecutive planes in place
template <typename T>
inline void test_filter_f(les_ctx* c, T* p, int nplanes, const double* G) {
    cx<T>* sp = scr<T>(c->spec3[2]);  // >= nplanes planes of ns
    x_r2c<T>(c->stream, p, sp, (long)nplanes * c->ny, c->nx, c->lhx, c->tw_x, 1.0);
    yline_spec<T>(c->stream, sp, sp, nplanes, c->lhx, (long)c->ns, c->tw_y, SP_FILT, G);
    x_c2r<T>(c->stream, sp, p, (long)nplanes * c->ny, c->lhx, c->nx, c->tw_x, 0, 0.0);
}

// Wait for both streams and return the first HIP error since the last call.
int finish(les_ctx* c);

}  // namespace les
