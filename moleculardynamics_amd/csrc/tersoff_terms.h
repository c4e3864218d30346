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

// FP32 (mixed precision) uses the hardware exp2, log2 and reciprocal (about
// 1 ulp each); FP64 the correctly rounded library functions.
template <typename T>
__device__ inline T t_exp(T x) {
    if constexpr (sizeof(T) == 4) return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); else return exp(x);
}
template <typename T>
__device__ inline T t_pow(T x, T y) {   // x > 0
    if constexpr (sizeof(T) == 4) return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); else return pow(x, y);
}
template <typename T>
__device__ inline T t_rcp(T x) {
    if constexpr (sizeof(T) == 4) return __builtin_amdgcn_rcpf(x); else return 1.0 / x;
}
template <typename T>
__device__ inline T t_div(T a, T b) {
    if constexpr (sizeof(T) == 4) return a * __builtin_amdgcn_rcpf(b); else return a / b;
}

template <typename T>
__device__ inline void fcut(const Tp<T>& p, T r, T& fc, T& dfc) {
    if (r < p.R - p.D) { fc = 1; dfc = 0; return; }
    if (r > p.R + p.D) { fc = 0; dfc = 0; return; }
    const T x = t_div<T>((T)M_PI_2 * (r - p.R), p.D);
    T sx, cx;
    if constexpr (sizeof(T) == 4) { sincosf(x, &sx, &cx); } else { sincos(x, &sx, &cx); }
    fc = (T)0.5 - (T)0.5 * sx;
    dfc = t_div<T>(-(T)0.25 * (T)M_PI, p.D) * cx;
}

template <typename T>
__device__ inline void gang(const Tp<T>& p, T cs, T& g, T& dg) {
    const T u = p.h - cs, den = p.d2 + u * u;
    if constexpr (sizeof(T) == 4) {
        // c^2/d^2 - c^2/den = c^2 u^2 / (d^2 den): for Si c^2/d^2 = 3.8e7 while g
        // is ~1e5, so the difference form would lose 6e-6 in FP32
        g = 1 + t_div<T>(p.c2 * (u * u), p.d2 * den);
    } else {
        g = 1 + p.c2 / p.d2 - p.c2 / den;
    }
    dg = t_div<T>(-2 * p.c2 * u, den * den);   // dg/dcos
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
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if constexpr (sizeof(T) == 4) {   // r^2 in FP64, the rest in FP32
        r = __builtin_sqrtf((float)r2);
        const float ir = __builtin_amdgcn_rcpf(r);
        for (int q = 0; q < 3; q++) u[q] = (float)d[q] * ir;
    } else {
        const double rr = sqrt(r2);
        r = rr;
        for (int q = 0; q < 3; q++) u[q] = d[q] / rr;
    }
}


// ---- one centre's terms: e_i = 1/2 sum_j V_ij and w_is = de_i/d(x_j - x_i) ----
// e_i depends only on the displacements from i to its neighbours, so its
// gradient is fully described by one vector per list slot; the force on atom
// a is then sum_s w_as - sum_{(i, s) -> a} w_is (a gather, no atomics).

// pair and bond-order part of bond (centre, partner at distance rij): returns
// the bond's energy and dE/dr at fixed zeta (de_r) and dE/dzeta (de_z)
template <typename T>
__device__ inline double bond_terms(const Tp<T>& p, T rij, T zeta, T& de_r, T& de_z) {
    T fc, dfc;
    fcut(p, rij, fc, dfc);
    const T fr = p.A * t_exp<T>(-p.l1 * rij), dfr = -p.l1 * fr;
    const T fa = -p.B * t_exp<T>(-p.l2 * rij), dfa = -p.l2 * fa;
    T bz = 1, dbz = 0;
    if (zeta > 0) {
        const T t = t_pow<T>(p.beta * zeta, p.n);
        bz = t_pow<T>(1 + t, (T)-0.5 / p.n);
        dbz = (T)-0.5 * t_div<T>(bz, 1 + t) * t_div<T>(t, zeta);
    }
    de_r = (T)0.5 * (dfc * (fr + bz * fa) + fc * (dfr + bz * dfa));
    de_z = (T)0.5 * fc * fa * dbz;
    return 0.5 * (double)fc * ((double)fr + (double)bz * (double)fa);
}

// Up to kS list entries held in registers; w[s] for s < cnt.
template <typename T, int kS>
__device__ inline double centre_reg(const Pos3& P, const Tp<T>& p, const int32_t* __restrict__ list, int32_t n,
                                    int32_t i, int32_t cnt, T (&w)[kS][3]) {
    T r[kS], ir[kS], u[kS][3];
    bool act[kS];
#pragma unroll
    for (int s = 0; s < kS; s++) {
        w[s][0] = w[s][1] = w[s][2] = 0;
        act[s] = false;
        r[s] = 1;
        ir[s] = 1;
        u[s][0] = u[s][1] = u[s][2] = 0;
        if (s < cnt) {
            double d[3];
            P.d(i, list[(int64_t)s * n + i], d);
            unit(d, r[s], u[s]);
            ir[s] = t_rcp<T>(r[s]);
            act[s] = r[s] < p.rc;
        }
    }
    double e = 0.0;
#pragma unroll
    for (int a = 0; a < kS; a++) {
        if (!act[a]) continue;
        T zeta = 0;
#pragma unroll
        for (int b = 0; b < kS; b++) {
            if (b == a || !act[b]) continue;
            T fk, dfk, gg, dg, ee, dee;
            fcut(p, r[b], fk, dfk);
            gang(p, u[a][0] * u[b][0] + u[a][1] * u[b][1] + u[a][2] * u[b][2], gg, dg);
            ex3(p, r[a] - r[b], ee, dee);
            zeta += fk * gg * ee;
        }
        T de_r, de_z;
        e += bond_terms(p, r[a], zeta, de_r, de_z);
#pragma unroll
        for (int q = 0; q < 3; q++) w[a][q] += de_r * u[a][q];
        if (de_z != 0) {
#pragma unroll
            for (int b = 0; b < kS; b++) {
                if (b == a || !act[b]) continue;
                const ZTerm<T> z = zterm(p, u[a], r[a], u[b], r[b]);
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    w[a][q] += de_z * (z.d_rij * u[a][q] + z.d_cs * (u[b][q] - z.cs * u[a][q]) * ir[a]);
                    w[b][q] += de_z * (z.d_rik * u[b][q] + z.d_cs * (u[a][q] - z.cs * u[b][q]) * ir[b]);
                }
            }
        }
    }
    return e;
}

// Any count: the list read from memory, w accumulated in place ([slot][n]
// vectors of 4 T, x y z used).
template <typename T, typename W4>
__device__ inline double centre_mem(const Pos3& P, const Tp<T>& p, const int32_t* __restrict__ list, int32_t n,
                                    int32_t i, int32_t cnt, W4* __restrict__ w4) {
    T* w = reinterpret_cast<T*>(w4);
    for (int32_t a = 0; a < cnt; a++)
        for (int q = 0; q < 4; q++) w[4 * ((int64_t)a * n + i) + q] = 0;
    double e = 0.0;
    for (int32_t a = 0; a < cnt; a++) {
        double dj[3];
        P.d(i, list[(int64_t)a * n + i], dj);
        T rij, u[3];
        unit(dj, rij, u);
        if (!(rij < p.rc)) continue;
        T zeta = 0;
        for (int32_t b = 0; b < cnt; b++) {
            if (b == a) continue;
            double dk[3];
            P.d(i, list[(int64_t)b * n + i], dk);
            T rik, v[3];
            unit(dk, rik, v);
            if (!(rik < p.rc)) continue;
            T fk, dfk, gg, dg, ee, dee;
            fcut(p, rik, fk, dfk);
            gang(p, u[0] * v[0] + u[1] * v[1] + u[2] * v[2], gg, dg);
            ex3(p, rij - rik, ee, dee);
            zeta += fk * gg * ee;
        }
        T de_r, de_z;
        e += bond_terms(p, rij, zeta, de_r, de_z);
        T gj[3] = {de_r * u[0], de_r * u[1], de_r * u[2]};
        if (de_z != 0) {
            for (int32_t b = 0; b < cnt; b++) {
                if (b == a) continue;
                double dk[3];
                P.d(i, list[(int64_t)b * n + i], dk);
                T rik, v[3];
                unit(dk, rik, v);
                if (!(rik < p.rc)) continue;
                const ZTerm<T> z = zterm(p, u, rij, v, rik);
                for (int q = 0; q < 3; q++) {
                    gj[q] += de_z * (z.d_rij * u[q] + z.d_cs * (v[q] - z.cs * u[q]) / rij);
                    w[4 * ((int64_t)b * n + i) + q] += de_z * (z.d_rik * v[q] + z.d_cs * (u[q] - z.cs * v[q]) / rik);
                }
            }
        }
        for (int q = 0; q < 3; q++) w[4 * ((int64_t)a * n + i) + q] += gj[q];
    }
    return e;
}

}  // namespace ts
}  // namespace mdx
