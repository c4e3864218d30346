// tersoff_terms.h -- device-side terms of the Tersoff silicon potential
// (J. Tersoff, PRB 38, 9902 (1988)) shared by the plain Tersoff kernels and
// the surrogate's per-atom fallback: parameters in the arithmetic type, the
// cut-off, the angular function g, the zeta exponential and the periodic
// minimum-image displacement.
#pragma once

#include "mdx_internal.h"

namespace mdx {
namespace ts {

__device__ inline double min_image(double d, double L, double invL) { return d - L * rint(d * invL); }

template <typename T>
struct Tp {   // parameters in the arithmetic type
    T A, B, l1, l2, l3, beta, n, c2, d2, h, R, D, rc;
    int32_t m;
};

template <typename T>
__device__ inline T t_exp(T x) {
    if constexpr (sizeof(T) == 4) return expf(x); else return exp(x);
}
template <typename T>
__device__ inline T t_pow(T x, T y) {
    if constexpr (sizeof(T) == 4) return powf(x, y); else return pow(x, y);
}

template <typename T>
__device__ inline void fcut(const Tp<T>& p, T r, T& fc, T& dfc) {
    if (r < p.R - p.D) { fc = 1; dfc = 0; return; }
    if (r > p.R + p.D) { fc = 0; dfc = 0; return; }
    const T x = (T)M_PI_2 * (r - p.R) / p.D;
    T sx, cx;
    if constexpr (sizeof(T) == 4) { sincosf(x, &sx, &cx); } else { sincos(x, &sx, &cx); }
    fc = (T)0.5 - (T)0.5 * sx;
    dfc = -(T)0.25 * (T)M_PI / p.D * cx;
}

template <typename T>
__device__ inline void gang(const Tp<T>& p, T cs, T& g, T& dg) {
    const T u = p.h - cs, den = p.d2 + u * u;
    if constexpr (sizeof(T) == 4) {
        // c^2/d^2 - c^2/den = c^2 u^2 / (d^2 den): for Si c^2/d^2 = 3.8e7 while g
        // is ~1e5, so the difference form would lose 6e-6 in FP32
        g = 1 + p.c2 * (u * u) / (p.d2 * den);
    } else {
        g = 1 + p.c2 / p.d2 - p.c2 / den;
    }
    dg = -2 * p.c2 * u / (den * den);   // dg/dcos
}

template <typename T>
__device__ inline void ex3(const Tp<T>& p, T x, T& e, T& de) {
    if (p.l3 == 0) { e = 1; de = 0; return; }
    if (p.m == 3) {
        const T l3 = p.l3 * p.l3 * p.l3;
        e = t_exp<T>(l3 * x * x * x);
        de = 3 * l3 * x * x * e;
    } else {
        e = t_exp<T>(p.l3 * x);
        de = p.l3 * e;
    }
}

struct Pos3 {
    const double* x;
    const double* y;
    const double* z;
    double L, invL;
    // displacement from atom a to atom b (minimum image), FP64
    __device__ void d(int32_t a, int32_t b, double out[3]) const {
        out[0] = min_image(x[b] - x[a], L, invL);
        out[1] = min_image(y[b] - y[a], L, invL);
        out[2] = min_image(z[b] - z[a], L, invL);
    }
};

template <typename T>
inline Tp<T> cast_params(const TersoffParams& q) {
    Tp<T> p;
    p.A = (T)q.A; p.B = (T)q.B; p.l1 = (T)q.lambda1; p.l2 = (T)q.lambda2; p.l3 = (T)q.lambda3;
    p.beta = (T)q.beta; p.n = (T)q.n; p.c2 = (T)q.c2; p.d2 = (T)q.d2; p.h = (T)q.h;
    p.R = (T)q.R; p.D = (T)q.D; p.rc = (T)(q.R + q.D); p.m = q.m;
    return p;
}

inline Pos3 pos3(const DevState& st, double L) { return Pos3{st.r[0], st.r[1], st.r[2], L, 1.0 / L}; }

// derivative pieces of one zeta term: bond (centre, partner) with unit vector
// u and length rij, third atom with unit vector w and length rik
template <typename T>
struct ZTerm {
    T cs, d_rij, d_rik, d_cs;
};
template <typename T>
__device__ inline ZTerm<T> zterm(const Tp<T>& p, const T u[3], T rij, const T w[3], T rik) {
    ZTerm<T> z;
    T fk, dfk, gg, dg, ee, dee;
    fcut(p, rik, fk, dfk);
    z.cs = u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
    gang(p, z.cs, gg, dg);
    ex3(p, rij - rik, ee, dee);
    z.d_rij = fk * gg * dee;
    z.d_rik = dfk * gg * ee - fk * gg * dee;
    z.d_cs = fk * dg * ee;
    return z;
}

template <typename T>
__device__ inline void unit(const double d[3], T& r, T u[3]) {
    const double rr = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    r = (T)rr;
    for (int q = 0; q < 3; q++) u[q] = (T)(d[q] / rr);
}

}  // namespace ts
}  // namespace mdx
