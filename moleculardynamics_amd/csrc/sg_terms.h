// sg_terms.h -- descriptor of the learned Tersoff surrogate, forward and
// reverse mode, in FP32 on registers.  Shared by the inference kernel
// (surrogate.hip) and the training kernels (sg_train.hip).
//
// For atom i and its neighbour slots s (x_j - x_i = dv_s, r_s < 3.0 A):
//   fc  Tersoff cut-off: 1 up to 2.4 A, 1/2 - 1/2 sin(pi/2 (r - 2.7)/0.3) to 3.0 A
//   G[k]        = sum_s exp(-eta (r_s - mu_k)^2) fc(r_s),   mu_k = 2 + k/31, k < 32,
//                 eta = 1/2 (31/1.5)^2
//   A_s0 = sum_{t != s} fc(r_t),  A_sl = sum_{t != s} fc(r_t) T_l(u_s . u_t)   (l = 1..3)
//   G[32+8m+q]  = sum_s psi_m(r_s) mono_q(A_s),  psi_m = exp(-(r - nu_m)^2 / 0.08) fc,
//                 nu = (2.25, 2.55), mono = (A1, A2, A3, A1^2, A1 A2, A2^2, A0 A1, A1^3)
// The reverse pass takes h = dE/dG and returns w_s = dE/d(dv_s) per slot.
#pragma once

#include "mdx_internal.h"

namespace mdx {
namespace sg {

constexpr int kNF = 48;      // descriptor length
constexpr int kNR = 32;      // radial part
constexpr int kNH = 64;      // hidden width
constexpr int kNTrain = kNH * kNF + kNH + kNH * kNH + kNH + kNH + 1;   // 7361 trainable
constexpr int kNBlob = kNTrain + 4 * kNF;                              // + mean, istd, lo, hi
// offsets in the parameter blob
constexpr int kOffW1 = 0, kOffB1 = kNH * kNF, kOffW2 = kOffB1 + kNH, kOffB2 = kOffW2 + kNH * kNH,
              kOffW3 = kOffB2 + kNH, kOffB3 = kOffW3 + kNH, kOffMean = kOffB3 + 1, kOffIstd = kOffMean + kNF,
              kOffLo = kOffIstd + kNF, kOffHi = kOffLo + kNF;

constexpr float kRc = 3.0f, kRin = 2.4f, kRmid = 2.7f, kRhalf = 0.3f;
constexpr float kEta = 0.5 * (31.0 / 1.5) * (31.0 / 1.5);
constexpr float kEtaB = 0.5 / (0.2 * 0.2);
constexpr float kLog2e = 1.4426950408889634f;
// radial Gaussians by recurrence from two anchor centres a = 8 and 24, each
// for the 16 centres around it:
//   e_{k+1} = e_k up_k, up_{k+1} = up_k q;  e_{k-1} = e_k dn_k, dn_{k-1} = dn_k q
//   up_a = exp(2 eta D x - eta D^2), dn_a = exp(-2 eta D x - eta D^2), q = exp(-2 eta D^2)
// with x = r - mu_a, D = 1/31.  An anchor underflows FP32 only when |x| > 0.64 A,
// and then every centre of its half (at most 8 D = 0.26 A away) is below 1e-13.
// A float32 emulation over r in [1.5, 3] A stays within 2.7e-5 relative on every
// term above 1e-8.
constexpr int kMid = 16;
constexpr float kDelta = 1.0f / 31.0f;
constexpr float kMuMid = 2.0 + 16.0 / 31.0;
constexpr float kQ = 0.6411803884299546f;
constexpr float kE2 = -308.0955409542875f;      // -eta log2(e)
constexpr float kE2U = 19.877131674470164f;     // 2 eta D log2(e)
constexpr float kE2C = -0.3205988979753252f;    // -eta D^2 log2(e)
constexpr float kEB2 = -18.03368801111204f;     // -eta_b log2(e)

__device__ inline float nu(int m) { return m == 0 ? 2.25f : 2.55f; }
__device__ inline float exp2_raw(float x) { return __builtin_amdgcn_exp2f(x); }

struct Nb {   // one neighbour slot; fc = dfc = 0 marks an empty slot
    float ux, uy, uz, r, ir, fc, dfc;
};

__device__ inline void cutoff(float r, float& fc, float& dfc) {
    if (r <= kRin) {
        fc = 1.0f;
        dfc = 0.0f;
    } else if (r >= kRc) {
        fc = 0.0f;
        dfc = 0.0f;
    } else {
        float sx, cx;
        sincosf((float)M_PI_2 * (1.0f / kRhalf) * (r - kRmid), &sx, &cx);
        fc = 0.5f - 0.5f * sx;
        dfc = -0.25f * (float)M_PI * (1.0f / kRhalf) * cx;
    }
}

// slot from a displacement (FP64 minimum image): r^2 in FP64 (the Gaussian
// tails amplify an error in r by 2 eta |r - mu|), the rest in FP32
__device__ inline Nb make_nb(double dx, double dy, double dz) {
    Nb b;
    const float x = (float)dx, y = (float)dy, z = (float)dz;
    b.r = __builtin_sqrtf((float)(dx * dx + dy * dy + dz * dz));
    b.ir = __builtin_amdgcn_rcpf(b.r);
    b.ux = x * b.ir;
    b.uy = y * b.ir;
    b.uz = z * b.ir;
    cutoff(b.r, b.fc, b.dfc);
    if (!(b.fc > 0.0f)) {   // outside rc: no contribution, no derivative
        b.fc = 0.0f;
        b.dfc = 0.0f;
    }
    return b;
}

__device__ inline Nb empty_nb() { return Nb{0.0f, 0.0f, 0.0f, 10.0f, 0.1f, 0.0f, 0.0f}; }

// calls f(k, e_k) for the 32 radial Gaussians of distance r
template <typename F>
__device__ inline void radial(float r, F&& f) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int ka = 8 + 16 * h;
        const float x = r - (h == 0 ? (float)(2.0 + 8.0 / 31.0) : (float)(2.0 + 24.0 / 31.0));
        const float e = exp2_raw(kE2 * x * x);
        float up = exp2_raw(fmaf(kE2U, x, kE2C)), dn = exp2_raw(fmaf(-kE2U, x, kE2C));
        f(ka, e);
        float eu = e, ed = e;
#pragma unroll
        for (int k = ka + 1; k < ka + 8; k++) {
            eu *= up;
            up *= kQ;
            f(k, eu);
        }
#pragma unroll
        for (int k = ka - 1; k >= ka - 8; k--) {
            ed *= dn;
            dn *= kQ;
            f(k, ed);
        }
    }
}

// bond-resolved angular moments A[s][0..3]
template <int kS>
__device__ inline void moments(const Nb (&nb)[kS], float (&A)[kS][4]) {
#pragma unroll
    for (int s = 0; s < kS; s++) A[s][0] = A[s][1] = A[s][2] = A[s][3] = 0.0f;
#pragma unroll
    for (int s = 0; s < kS; s++) {
#pragma unroll
        for (int t = s + 1; t < kS; t++) {
            if (nb[s].fc > 0.0f && nb[t].fc > 0.0f) {
                const float c = nb[s].ux * nb[t].ux + nb[s].uy * nb[t].uy + nb[s].uz * nb[t].uz;
                const float T2 = 2.0f * c * c - 1.0f, T3 = (4.0f * c * c - 3.0f) * c;
                A[s][0] += nb[t].fc;
                A[s][1] = fmaf(nb[t].fc, c, A[s][1]);
                A[s][2] = fmaf(nb[t].fc, T2, A[s][2]);
                A[s][3] = fmaf(nb[t].fc, T3, A[s][3]);
                A[t][0] += nb[s].fc;
                A[t][1] = fmaf(nb[s].fc, c, A[t][1]);
                A[t][2] = fmaf(nb[s].fc, T2, A[t][2]);
                A[t][3] = fmaf(nb[s].fc, T3, A[t][3]);
            }
        }
    }
}

__device__ inline void monomials(const float A[4], float (&mo)[8]) {
    mo[0] = A[1];
    mo[1] = A[2];
    mo[2] = A[3];
    mo[3] = A[1] * A[1];
    mo[4] = A[1] * A[2];
    mo[5] = A[2] * A[2];
    mo[6] = A[0] * A[1];
    mo[7] = A[1] * A[1] * A[1];
}

template <int kS>
__device__ inline void forward(const Nb (&nb)[kS], float (&G)[kNF]) {
#pragma unroll
    for (int f = 0; f < kNF; f++) G[f] = 0.0f;
#pragma unroll
    for (int s = 0; s < kS; s++) {
        if (nb[s].fc > 0.0f) {
            const float fc = nb[s].fc;
            radial(nb[s].r, [&](int k, float e) { G[k] = fmaf(fc, e, G[k]); });
        }
    }
    float A[kS][4];
    moments(nb, A);
#pragma unroll
    for (int s = 0; s < kS; s++) {
        if (nb[s].fc > 0.0f) {
            float mo[8];
            monomials(A[s], mo);
#pragma unroll
            for (int m = 0; m < 2; m++) {
                const float d = nb[s].r - nu(m);
                const float psi = exp2_raw(kEB2 * d * d) * nb[s].fc;
#pragma unroll
                for (int q = 0; q < 8; q++) G[kNR + 8 * m + q] = fmaf(psi, mo[q], G[kNR + 8 * m + q]);
            }
        }
    }
}

// w[s] = dE/d(dv_s) for h = dE/dG
template <int kS>
__device__ inline void backward(const Nb (&nb)[kS], const float (&h)[kNF], float (&w)[kS][3]) {
    float A[kS][4];
    moments(nb, A);
    float Lam[kS][4];
#pragma unroll
    for (int s = 0; s < kS; s++) {
        w[s][0] = w[s][1] = w[s][2] = 0.0f;
        Lam[s][0] = Lam[s][1] = Lam[s][2] = Lam[s][3] = 0.0f;
        if (nb[s].fc > 0.0f) {
            const float r = nb[s].r, fc = nb[s].fc, dfc = nb[s].dfc;
            // sum_k h_k e_k (dfc - 2 eta (r - mu_k) fc), r - mu_k = x - (k - 16) D
            float s0 = 0.0f, s1 = 0.0f;
            radial(r, [&](int k, float e) {
                const float he = h[k] * e;
                s0 += he;
                s1 = fmaf(he, (float)(k - kMid), s1);
            });
            const float x = r - kMuMid;
            float coef = dfc * s0 - 2.0f * kEta * fc * fmaf(x, s0, -kDelta * s1);
            float mo[8];
            monomials(A[s], mo);
            float P[8];
#pragma unroll
            for (int q = 0; q < 8; q++) P[q] = 0.0f;
#pragma unroll
            for (int m = 0; m < 2; m++) {
                const float d = r - nu(m);
                const float g = exp2_raw(kEB2 * d * d);
                const float psi = g * fc, dpsi = g * fmaf(-2.0f * kEtaB * d, fc, dfc);
                float Hm = 0.0f;
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    Hm = fmaf(h[kNR + 8 * m + q], mo[q], Hm);
                    P[q] = fmaf(psi, h[kNR + 8 * m + q], P[q]);
                }
                coef = fmaf(dpsi, Hm, coef);
            }
            const float A0 = A[s][0], A1 = A[s][1], A2 = A[s][2];
            Lam[s][0] = P[6] * A1;
            Lam[s][1] = P[0] + 2.0f * A1 * P[3] + A2 * P[4] + A0 * P[6] + 3.0f * A1 * A1 * P[7];
            Lam[s][2] = P[1] + A1 * P[4] + 2.0f * A2 * P[5];
            Lam[s][3] = P[2];
            w[s][0] = coef * nb[s].ux;
            w[s][1] = coef * nb[s].uy;
            w[s][2] = coef * nb[s].uz;
        }
    }
#pragma unroll
    for (int s = 0; s < kS; s++) {
#pragma unroll
        for (int t = s + 1; t < kS; t++) {
            if (nb[s].fc > 0.0f && nb[t].fc > 0.0f) {
                const float c = nb[s].ux * nb[t].ux + nb[s].uy * nb[t].uy + nb[s].uz * nb[t].uz;
                const float T2 = 2.0f * c * c - 1.0f, T3 = (4.0f * c * c - 3.0f) * c;
                const float D2 = 4.0f * c, D3 = 12.0f * c * c - 3.0f;
                const float as = Lam[s][1] + Lam[s][2] * D2 + Lam[s][3] * D3;
                const float at = Lam[t][1] + Lam[t][2] * D2 + Lam[t][3] * D3;
                const float bs = Lam[s][0] + Lam[s][1] * c + Lam[s][2] * T2 + Lam[s][3] * T3;
                const float bt = Lam[t][0] + Lam[t][1] * c + Lam[t][2] * T2 + Lam[t][3] * T3;
                const float k = nb[t].fc * as + nb[s].fc * at;
                const float ks = k * nb[s].ir, kt = k * nb[t].ir;
                const float es = nb[s].dfc * bt, et = nb[t].dfc * bs;
                w[s][0] += ks * (nb[t].ux - c * nb[s].ux) + es * nb[s].ux;
                w[s][1] += ks * (nb[t].uy - c * nb[s].uy) + es * nb[s].uy;
                w[s][2] += ks * (nb[t].uz - c * nb[s].uz) + es * nb[s].uz;
                w[t][0] += kt * (nb[s].ux - c * nb[t].ux) + et * nb[t].ux;
                w[t][1] += kt * (nb[s].uy - c * nb[t].uy) + et * nb[t].uy;
                w[t][2] += kt * (nb[s].uz - c * nb[t].uz) + et * nb[t].uz;
            }
        }
    }
}

__device__ inline float tanh_f(float x) {
    // 1 - 2 / (e^{2x} + 1) with the hardware exp2 and reciprocal: limits +-1 for large |x|
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(exp2_raw(2.0f * kLog2e * x) + 1.0f);
}

// float -> bfloat16 bits, round to nearest, ties to even (finite inputs)
__host__ __device__ inline uint32_t bf16_bits(float x) {
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

}  // namespace sg
}  // namespace mdx
