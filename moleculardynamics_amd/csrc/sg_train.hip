// sg_train.hip -- training of the learned Tersoff surrogate on the GPU.
//
// A training set is a list of periodic frames of the same box with their
// Tersoff forces and energies.  At creation the kernels find each atom's
// neighbours inside 3.0 A (at most kT), the descriptors G and their Jacobian
// J_s = dG/d(x_j - x_i) per slot (the reverse pass of sg_terms.h run once per
// descriptor component), and keep both in HBM:
//   G[f * N + a],  J[((s * 3 + d) * 48 + f) * N + a]     (a: atom of the set)
// A minibatch step (frames fr[0..nb)):
//   k_tr_fwd    network forward and dE/dG in FP32, w_is = J_is^T dE/dG
//   k_tr_force  F_i = sum_s w_is - w_{j,rev}; rho_i = dL/dF_i; squared residuals
//   k_tr_frame  frame energies, eps_f = dL/dE_i, energy residuals
//   k_tr_bwd    weight gradient of the loss: the directional derivative of
//               dE/dGs along u_i = istd * sum_s J_is (rho_i - rho_j), back-
//               propagated (double back-propagation written out), plus the
//               energy term; one wave per 64 atoms, per-block partial sums
//   k_tr_reduce deterministic sum of the partials (FP64), loss
//   k_tr_adam   Adam update of the FP32 weights
// Loss = mean over force components (F - F_ref)^2 + w_E mean over frames ((E - E_ref)/n)^2.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mdx_internal.h"
#include "sg_terms.h"

using namespace mdx;
using namespace mdx::sg;

namespace {

constexpr int kT = 8;            // neighbour slots per training atom
constexpr int kTB = 64;          // one wave per block
constexpr int kStride = 65;      // LDS row stride of per-lane vectors (conflict-free both ways)

struct TrainView {
    int32_t nf, na;
    int64_t N;
    double L;
    const int32_t* nbr;    // [s][N] frame-local neighbour index
    const int32_t* cnt;
    const uint8_t* rev;    // [s][N]
    const float* G;
    const float* J;
    const float* fref;     // [d][N]
    const double* eref;    // [nf]
};

__device__ inline double wrap(double d, double L) { return d - L * rint(d / L); }

__global__ void __launch_bounds__(256)
k_tr_nbr(const double* __restrict__ pos, int32_t nf, int32_t na, double L, int32_t* __restrict__ nbr,
         int32_t* __restrict__ cnt, int32_t* __restrict__ err) {
    const int64_t N = (int64_t)nf * na;
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    const int64_t base = a - a % na;
    const double x = pos[3 * a], y = pos[3 * a + 1], z = pos[3 * a + 2];
    int32_t k = 0;
    for (int32_t j = 0; j < na; j++) {
        if (base + j == a) continue;
        const double dx = wrap(pos[3 * (base + j)] - x, L), dy = wrap(pos[3 * (base + j) + 1] - y, L),
                     dz = wrap(pos[3 * (base + j) + 2] - z, L);
        if (dx * dx + dy * dy + dz * dz < (double)kRc * kRc) {
            if (k < kT) nbr[k * N + a] = j;
            k++;
        }
    }
    cnt[a] = k < kT ? k : kT;
    if (k > kT) atomicMax(err, k);
}

__global__ void __launch_bounds__(256)
k_tr_rev(int32_t nf, int32_t na, const int32_t* __restrict__ nbr, const int32_t* __restrict__ cnt,
         uint8_t* __restrict__ rev, int32_t* __restrict__ err) {
    const int64_t N = (int64_t)nf * na;
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    const int64_t base = a - a % na;
    const int32_t self = (int32_t)(a - base);
    for (int32_t s = 0; s < cnt[a]; s++) {
        const int64_t j = base + nbr[s * N + a];
        int32_t b = 0;
        while (b < cnt[j] && nbr[b * N + j] != self) b++;
        if (b == cnt[j]) {
            atomicOr(err + 1, 1);
            b = 0;
        }
        rev[s * N + a] = (uint8_t)b;
    }
}

__device__ inline void slots(const double* __restrict__ pos, const TrainView& v, int64_t a, Nb (&nb)[kT]) {
    const int64_t base = a - a % v.na;
    const int32_t c = v.cnt[a];
#pragma unroll
    for (int s = 0; s < kT; s++) {
        nb[s] = empty_nb();
        if (s < c) {
            const int64_t j = base + v.nbr[s * v.N + a];
            nb[s] = make_nb(wrap(pos[3 * j] - pos[3 * a], v.L), wrap(pos[3 * j + 1] - pos[3 * a + 1], v.L),
                            wrap(pos[3 * j + 2] - pos[3 * a + 2], v.L));
        }
    }
}

__global__ void __launch_bounds__(128)
k_tr_desc(const double* __restrict__ pos, TrainView v, float* __restrict__ G, float* __restrict__ J) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= v.N) return;
    Nb nb[kT];
    slots(pos, v, a, nb);
    float g[kNF];
    forward(nb, g);
    for (int f = 0; f < kNF; f++) G[f * v.N + a] = g[f];
    const int32_t c = v.cnt[a];
    for (int f = 0; f < kNF; f++) {
        float h[kNF];
#pragma unroll
        for (int q = 0; q < kNF; q++) h[q] = q == f ? 1.0f : 0.0f;
        float w[kT][3];
        backward(nb, h, w);
#pragma unroll
        for (int s = 0; s < kT; s++)
            if (s < c)
#pragma unroll
                for (int d = 0; d < 3; d++) J[((int64_t)(s * 3 + d) * kNF + f) * v.N + a] = w[s][d];
    }
}

// per-feature mean, std, min, max over all atoms: one block per feature
__global__ void __launch_bounds__(256)
k_tr_stats(const float* __restrict__ G, int64_t N, double* __restrict__ out) {
    const int f = blockIdx.x;
    double s = 0, s2 = 0, lo = INFINITY, hi = -INFINITY;
    for (int64_t a = threadIdx.x; a < N; a += blockDim.x) {
        const double x = G[f * N + a];
        s += x;
        s2 += x * x;
        lo = fmin(lo, x);
        hi = fmax(hi, x);
    }
    __shared__ double r[4][256];
    r[0][threadIdx.x] = s;
    r[1][threadIdx.x] = s2;
    r[2][threadIdx.x] = lo;
    r[3][threadIdx.x] = hi;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) {
            r[0][threadIdx.x] += r[0][threadIdx.x + k];
            r[1][threadIdx.x] += r[1][threadIdx.x + k];
            r[2][threadIdx.x] = fmin(r[2][threadIdx.x], r[2][threadIdx.x + k]);
            r[3][threadIdx.x] = fmax(r[3][threadIdx.x], r[3][threadIdx.x + k]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double mean = r[0][0] / (double)N;
        out[f] = mean;
        out[kNF + f] = sqrt(fmax(r[1][0] / (double)N - mean * mean, 0.0));
        out[2 * kNF + f] = r[2][0];
        out[3 * kNF + f] = r[3][0];
    }
}

// minibatch atom b -> atom of the set
__device__ inline int64_t set_atom(const TrainView& v, const int32_t* fr, int64_t b) {
    return (int64_t)fr[b / v.na] * v.na + b % v.na;
}

// LDS vectors [unit * kStride + lane]
#define VEC(base, unit) (base)[(unit) * kStride + lane]

// network forward from standardized inputs in sGs; leaves a1 in sA1 and
// returns E; fills sG2 with dE/dz2 and sG1 with dE/dz1 when grad is set
__device__ inline float net_forward(const float* __restrict__ P, float* sGs, float* sA1, float* sG2, float* sG1,
                                    int lane) {
    const float* W1 = P + kOffW1;
    const float* W2 = P + kOffW2;
    for (int h = 0; h < kNH; h++) {
        float z = P[kOffB1 + h];
#pragma unroll 16
        for (int f = 0; f < kNF; f++) z = fmaf(W1[h * kNF + f], VEC(sGs, f), z);
        VEC(sA1, h) = tanh_f(z);
    }
    float E = P[kOffB3];
    for (int h = 0; h < kNH; h++) {
        float z = P[kOffB2 + h];
#pragma unroll 16
        for (int k = 0; k < kNH; k++) z = fmaf(W2[h * kNH + k], VEC(sA1, k), z);
        const float a2 = tanh_f(z);
        E = fmaf(P[kOffW3 + h], a2, E);
        VEC(sG2, h) = P[kOffW3 + h] * (1.0f - a2 * a2);
    }
    for (int k = 0; k < kNH; k++) {
        float q = 0.0f;
#pragma unroll 16
        for (int h = 0; h < kNH; h++) q = fmaf(W2[h * kNH + k], VEC(sG2, h), q);
        const float a1 = VEC(sA1, k);
        VEC(sG1, k) = q * (1.0f - a1 * a1);
    }
    return E;
}

__global__ void __launch_bounds__(kTB)
k_tr_fwd(TrainView v, const int32_t* __restrict__ fr, int64_t M, const float* __restrict__ P,
         float* __restrict__ w, float* __restrict__ Eat) {
    extern __shared__ float sm[];
    const int lane = threadIdx.x;
    float* sGs = sm;
    float* sA1 = sGs + kNF * kStride;
    float* sG2 = sA1 + kNH * kStride;
    float* sG1 = sG2 + kNH * kStride;
    float* sH = sG1 + kNH * kStride;
    const int64_t b = (int64_t)blockIdx.x * kTB + lane;
    const bool live = b < M;
    const int64_t a = live ? set_atom(v, fr, b) : 0;
    for (int f = 0; f < kNF; f++)
        VEC(sGs, f) = live ? (v.G[f * v.N + a] - P[kOffMean + f]) * P[kOffIstd + f] : 0.0f;
    const float E = net_forward(P, sGs, sA1, sG2, sG1, lane);
    const float* W1 = P + kOffW1;
    for (int f = 0; f < kNF; f++) {
        float hs = 0.0f;
#pragma unroll 16
        for (int k = 0; k < kNH; k++) hs = fmaf(W1[k * kNF + f], VEC(sG1, k), hs);
        VEC(sH, f) = hs * P[kOffIstd + f];   // dE/dG
    }
    if (!live) return;
    Eat[b] = E;
    const int32_t c = v.cnt[a];
    for (int s = 0; s < c; s++)
        for (int d = 0; d < 3; d++) {
            float acc = 0.0f;
            const float* Jp = v.J + (int64_t)(s * 3 + d) * kNF * v.N + a;
            for (int f = 0; f < kNF; f++) acc = fmaf(Jp[f * v.N], VEC(sH, f), acc);
            w[(s * 3 + d) * M + b] = acc;
        }
}

__global__ void __launch_bounds__(256)
k_tr_force(TrainView v, const int32_t* __restrict__ fr, int64_t M, const float* __restrict__ w,
           float* __restrict__ rho, double* __restrict__ sse_part) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double sse = 0.0;
    if (b < M) {
        const int64_t a = set_atom(v, fr, b);
        const int64_t bbase = b - b % v.na;
        const int32_t c = v.cnt[a];
        double F[3] = {0.0, 0.0, 0.0};
        for (int s = 0; s < c; s++) {
            const int64_t jb = bbase + v.nbr[s * v.N + a];   // partner in the minibatch
            const int rb = v.rev[s * v.N + a];
            for (int d = 0; d < 3; d++) F[d] += (double)w[(s * 3 + d) * M + b] - (double)w[(rb * 3 + d) * M + jb];
        }
        for (int d = 0; d < 3; d++) {
            const double r = F[d] - (double)v.fref[d * v.N + a];
            sse += r * r;
            rho[d * M + b] = (float)(2.0 * r / (3.0 * (double)M));
        }
    }
    __shared__ double red[256];
    red[threadIdx.x] = sse;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) sse_part[blockIdx.x] = red[0];
}

// one block per minibatch frame
__global__ void __launch_bounds__(256)
k_tr_frame(TrainView v, const int32_t* __restrict__ fr, int32_t nb, const float* __restrict__ Eat, double w_energy,
           float* __restrict__ eps, double* __restrict__ eloss) {
    const int32_t q = blockIdx.x;
    double s = 0.0;
    for (int32_t i = threadIdx.x; i < v.na; i += blockDim.x) s += (double)Eat[(int64_t)q * v.na + i];
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double n = (double)v.na, r = (red[0] - v.eref[fr[q]]) / n;
        eps[q] = (float)(2.0 * w_energy * r / n / (double)nb);
        eloss[q] = r * r;
    }
}

__global__ void __launch_bounds__(kTB)
k_tr_bwd(TrainView v, const int32_t* __restrict__ fr, int64_t M, const float* __restrict__ P,
         const float* __restrict__ rho, const float* __restrict__ eps, float* __restrict__ part) {
    extern __shared__ float sm[];
    const int lane = threadIdx.x;
    float* sGs = sm;
    float* sU = sGs + kNF * kStride;
    float* sA1 = sU + kNF * kStride;
    float* sT1A = sA1 + kNH * kStride;
    float* sG2 = sT1A + kNH * kStride;
    float* sC2 = sG2 + kNH * kStride;
    float* sC1 = sC2 + kNH * kStride;
    float* sG1T = sC1 + kNH * kStride;
    float* sW3T = sG1T + kNH * kStride;
    const int64_t b = (int64_t)blockIdx.x * kTB + lane;
    const bool live = b < M;
    const int64_t a = live ? set_atom(v, fr, b) : 0;
    const float ep = live ? eps[b / v.na] : 0.0f;
    const float* W1 = P + kOffW1;
    const float* W2 = P + kOffW2;
    // inputs and the tangent u = istd * sum_s J_s (rho_i - rho_j)
    const int32_t c = live ? v.cnt[a] : 0;
    const int64_t bbase = b - b % v.na;
    float dr[kT][3];
#pragma unroll
    for (int s = 0; s < kT; s++)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            dr[s][d] = 0.0f;
            if (s < c) dr[s][d] = rho[d * M + b] - rho[d * M + bbase + v.nbr[s * v.N + a]];
        }
    for (int f = 0; f < kNF; f++) {
        VEC(sGs, f) = live ? (v.G[f * v.N + a] - P[kOffMean + f]) * P[kOffIstd + f] : 0.0f;
        float u = 0.0f;
#pragma unroll
        for (int s = 0; s < kT; s++)
            if (s < c)
#pragma unroll
                for (int d = 0; d < 3; d++) u = fmaf(v.J[((int64_t)(s * 3 + d) * kNF + f) * v.N + a], dr[s][d], u);
        VEC(sU, f) = u * P[kOffIstd + f];
    }
    // forward: a1; tangent t1a = (1 - a1^2) W1 u
    for (int h = 0; h < kNH; h++) {
        float z = P[kOffB1 + h], t = 0.0f;
#pragma unroll 16
        for (int f = 0; f < kNF; f++) {
            z = fmaf(W1[h * kNF + f], VEC(sGs, f), z);
            t = fmaf(W1[h * kNF + f], VEC(sU, f), t);
        }
        const float a1 = tanh_f(z);
        VEC(sA1, h) = a1;
        VEC(sT1A, h) = (1.0f - a1 * a1) * t;
    }
    // second layer: a2, t2 = W2 t1a; g2 = w3 d2; c2 = -2 a2 d2 w3 t2 + eps g2
    for (int h = 0; h < kNH; h++) {
        float z = P[kOffB2 + h], t2 = 0.0f;
#pragma unroll 16
        for (int k = 0; k < kNH; k++) {
            z = fmaf(W2[h * kNH + k], VEC(sA1, k), z);
            t2 = fmaf(W2[h * kNH + k], VEC(sT1A, k), t2);
        }
        const float a2 = tanh_f(z), d2 = 1.0f - a2 * a2, w3 = P[kOffW3 + h];
        const float g2 = w3 * d2;
        VEC(sG2, h) = g2;
        VEC(sC2, h) = live ? -2.0f * a2 * d2 * w3 * t2 + ep * g2 : 0.0f;
        VEC(sW3T, h) = live ? d2 * t2 + ep * a2 : 0.0f;
    }
    // first layer: qq = W2^T g2; c1 = (W2^T c2) d1 - 2 a1 qq t1a; g1t = qq d1
    for (int k = 0; k < kNH; k++) {
        float qq = 0.0f, cc = 0.0f;
#pragma unroll 16
        for (int h = 0; h < kNH; h++) {
            qq = fmaf(W2[h * kNH + k], VEC(sG2, h), qq);
            cc = fmaf(W2[h * kNH + k], VEC(sC2, h), cc);
        }
        const float a1 = VEC(sA1, k), d1 = 1.0f - a1 * a1;
        VEC(sC1, k) = live ? cc * d1 - 2.0f * a1 * qq * VEC(sT1A, k) : 0.0f;
        VEC(sG1T, k) = live ? qq * d1 : 0.0f;
    }
    __syncthreads();
    // lane t accumulates row t of W1 and W2 over the block's 64 atoms
    const int t = lane;
    float acc1[kNF], acc2[kNH];
#pragma unroll
    for (int f = 0; f < kNF; f++) acc1[f] = 0.0f;
#pragma unroll
    for (int k = 0; k < kNH; k++) acc2[k] = 0.0f;
    float sb1 = 0.0f, sb2 = 0.0f, sw3 = 0.0f, sb3 = 0.0f;
    for (int q = 0; q < kTB; q++) {
        const float g1t = sG1T[t * kStride + q], c1 = sC1[t * kStride + q];
        const float g2 = sG2[t * kStride + q], c2 = sC2[t * kStride + q];
#pragma unroll
        for (int f = 0; f < kNF; f++)
            acc1[f] = fmaf(g1t, sU[f * kStride + q], fmaf(c1, sGs[f * kStride + q], acc1[f]));
        // (padding atoms have u = 0, so t1a = c1 = c2 = g1t = 0 there)
#pragma unroll
        for (int k = 0; k < kNH; k++)
            acc2[k] = fmaf(g2, sT1A[k * kStride + q], fmaf(c2, sA1[k * kStride + q], acc2[k]));
        sb1 += c1;
        sb2 += c2;
        sw3 += sW3T[t * kStride + q];
    }
    float* out = part + (int64_t)blockIdx.x * kNTrain;
#pragma unroll
    for (int f = 0; f < kNF; f++) out[kOffW1 + t * kNF + f] = acc1[f];
    out[kOffB1 + t] = sb1;
#pragma unroll
    for (int k = 0; k < kNH; k++) out[kOffW2 + t * kNH + k] = acc2[k];
    out[kOffB2 + t] = sb2;
    out[kOffW3 + t] = sw3;
    if (t == 0) {
        for (int q = 0; q < kTB; q++)
            if ((int64_t)blockIdx.x * kTB + q < M) sb3 += eps[((int64_t)blockIdx.x * kTB + q) / v.na];
        out[kOffB3] = sb3;
    }
}

__global__ void __launch_bounds__(256)
k_tr_reduce(const float* __restrict__ part, int32_t nblocks, float* __restrict__ grad) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kNTrain) return;
    double s = 0.0;
    for (int32_t q = 0; q < nblocks; q++) s += (double)part[(int64_t)q * kNTrain + k];
    grad[k] = (float)s;
}

__global__ void __launch_bounds__(256)
k_tr_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1, float* __restrict__ m2,
          float lr, float c1, float c2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kNTrain) return;
    const float gk = g[k];
    m1[k] = 0.9f * m1[k] + 0.1f * gk;
    m2[k] = 0.999f * m2[k] + 0.001f * gk * gk;
    p[k] -= lr * (m1[k] * c1) / (sqrtf(m2[k] * c2) + 1e-8f);
}

}  // namespace

struct mdx_sg_trainer {
    int32_t nf = 0, na = 0;
    int64_t N = 0;
    double L = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    double* pos = nullptr;
    float* fref = nullptr;
    double* eref = nullptr;
    int32_t* nbr = nullptr;
    int32_t* cnt = nullptr;
    uint8_t* rev = nullptr;
    float* G = nullptr;
    float* J = nullptr;
    int32_t* derr = nullptr;
    float* params = nullptr;   // kNBlob
    float* grad = nullptr;
    float* m1 = nullptr;
    float* m2 = nullptr;
    int64_t adam_t = 0;
    // minibatch scratch
    int64_t cap_m = 0;
    int32_t* fr = nullptr;
    float* w = nullptr;
    float* Eat = nullptr;
    float* rho = nullptr;
    float* eps = nullptr;
    double* eloss = nullptr;
    double* sse = nullptr;
    float* part = nullptr;
    std::string err;

    TrainView view() const {
        return TrainView{nf, na, N, L, nbr, cnt, rev, G, J, fref, eref};
    }
};

namespace {

thread_local std::string g_terr;  // errors before a trainer exists, per calling thread

int tfail(mdx_sg_trainer* t, int code, const std::string& msg) {
    if (t) t->err = msg; else g_terr = msg;
    return code;
}

#define TCHK(t, expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) return tfail(t, MDX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int talloc(mdx_sg_trainer* t, T** p, size_t count) {
    if (hipMalloc((void**)p, sizeof(T) * (count ? count : 1)) != hipSuccess)
        return tfail(t, MDX_ERR_ALLOC, "device allocation failed");
    return MDX_OK;
}

constexpr size_t kFwdLds = (size_t)(kNF + 3 * kNH + kNF) * kStride * sizeof(float);
constexpr size_t kBwdLds = (size_t)(2 * kNF + 7 * kNH) * kStride * sizeof(float);

int ensure_batch(mdx_sg_trainer* t, const int32_t* frames, int32_t nb) {
    if (nb <= 0 || !frames) return tfail(t, MDX_ERR_ARG, "empty minibatch");
    for (int32_t q = 0; q < nb; q++)
        if (frames[q] < 0 || frames[q] >= t->nf) return tfail(t, MDX_ERR_ARG, "frame index out of range");
    const int64_t M = (int64_t)nb * t->na;
    if (M > t->cap_m) {
        hipFree(t->fr); hipFree(t->w); hipFree(t->Eat); hipFree(t->rho); hipFree(t->eps); hipFree(t->eloss);
        hipFree(t->sse); hipFree(t->part);
        t->fr = nullptr; t->w = nullptr; t->Eat = nullptr; t->rho = nullptr; t->eps = nullptr; t->eloss = nullptr;
        t->sse = nullptr; t->part = nullptr;
        const int64_t blocks = (M + kTB - 1) / kTB;
        int rc;
        if ((rc = talloc(t, &t->fr, (size_t)nb)) || (rc = talloc(t, &t->w, (size_t)(kT * 3 * M))) ||
            (rc = talloc(t, &t->Eat, (size_t)M)) || (rc = talloc(t, &t->rho, (size_t)(3 * M))) ||
            (rc = talloc(t, &t->eps, (size_t)nb)) || (rc = talloc(t, &t->eloss, (size_t)nb)) ||
            (rc = talloc(t, &t->sse, (size_t)((M + 255) / 256))) || (rc = talloc(t, &t->part, (size_t)(blocks * kNTrain)))) {
            t->cap_m = 0;
            return rc;
        }
        t->cap_m = M;
    }
    TCHK(t, hipMemcpyAsync(t->fr, frames, sizeof(int32_t) * nb, hipMemcpyHostToDevice, t->stream));
    return MDX_OK;
}

// forward + force + frame passes; loss parts to out3 (loss, force mse, energy mse per atom)
int forward_pass(mdx_sg_trainer* t, int32_t nb, double w_energy, double* out3) {
    const int64_t M = (int64_t)nb * t->na;
    const TrainView v = t->view();
    const int blocks = (int)((M + kTB - 1) / kTB);
    k_tr_fwd<<<blocks, kTB, kFwdLds, t->stream>>>(v, t->fr, M, t->params, t->w, t->Eat);
    TCHK(t, hipGetLastError());
    const int fb = (int)((M + 255) / 256);
    k_tr_force<<<fb, 256, 0, t->stream>>>(v, t->fr, M, t->w, t->rho, t->sse);
    TCHK(t, hipGetLastError());
    k_tr_frame<<<nb, 256, 0, t->stream>>>(v, t->fr, nb, t->Eat, w_energy, t->eps, t->eloss);
    TCHK(t, hipGetLastError());
    if (out3) {
        std::vector<double> sse(fb), el(nb);
        TCHK(t, hipMemcpyAsync(sse.data(), t->sse, sizeof(double) * fb, hipMemcpyDeviceToHost, t->stream));
        TCHK(t, hipMemcpyAsync(el.data(), t->eloss, sizeof(double) * nb, hipMemcpyDeviceToHost, t->stream));
        TCHK(t, hipStreamSynchronize(t->stream));
        double fs = 0, es = 0;
        for (double x : sse) fs += x;
        for (double x : el) es += x;
        out3[1] = fs / (3.0 * (double)M);
        out3[2] = es / nb;
        out3[0] = out3[1] + w_energy * out3[2];
    }
    return MDX_OK;
}

int gradient_pass(mdx_sg_trainer* t, int32_t nb) {
    const int64_t M = (int64_t)nb * t->na;
    const int blocks = (int)((M + kTB - 1) / kTB);
    k_tr_bwd<<<blocks, kTB, kBwdLds, t->stream>>>(t->view(), t->fr, M, t->params, t->rho, t->eps, t->part);
    TCHK(t, hipGetLastError());
    k_tr_reduce<<<(kNTrain + 255) / 256, 256, 0, t->stream>>>(t->part, blocks, t->grad);
    TCHK(t, hipGetLastError());
    return MDX_OK;
}

}  // namespace

extern "C" {

const char* mdx_sg_trainer_last_error(const mdx_sg_trainer* t) { return t ? t->err.c_str() : g_terr.c_str(); }

int mdx_sg_trainer_destroy(mdx_sg_trainer* t) {
    if (!t) return MDX_OK;
    hipSetDevice(t->device);
    if (t->stream) hipStreamSynchronize(t->stream);
    for (void* p : {(void*)t->pos, (void*)t->fref, (void*)t->eref, (void*)t->nbr, (void*)t->cnt, (void*)t->rev,
                    (void*)t->G, (void*)t->J, (void*)t->derr, (void*)t->params, (void*)t->grad, (void*)t->m1,
                    (void*)t->m2, (void*)t->fr, (void*)t->w, (void*)t->Eat, (void*)t->rho, (void*)t->eps,
                    (void*)t->eloss, (void*)t->sse, (void*)t->part})
        if (p) hipFree(p);
    if (t->stream) hipStreamDestroy(t->stream);
    delete t;
    return MDX_OK;
}

int mdx_sg_trainer_create(int32_t n_frames, int32_t n_atoms, double box_length, const double* pos3,
                          const double* force3, const double* energy, mdx_sg_trainer** out) {
    if (!out) return tfail(nullptr, MDX_ERR_ARG, "out is null");
    *out = nullptr;
    if (n_frames <= 0 || n_atoms <= 1 || !pos3 || !force3 || !energy)
        return tfail(nullptr, MDX_ERR_ARG, "need frames, atoms, positions, forces and energies");
    if (!(box_length >= 2.0 * kRc)) return tfail(nullptr, MDX_ERR_ARG, "box_length must be at least 6.0 A");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return tfail(nullptr, MDX_ERR_HIP, "no HIP device available");
    mdx_sg_trainer* t = new mdx_sg_trainer();
    t->nf = n_frames;
    t->na = n_atoms;
    t->N = (int64_t)n_frames * n_atoms;
    t->L = box_length;
    hipGetDevice(&t->device);
    auto bail = [&](int rc) {
        std::string msg = t->err;
        mdx_sg_trainer_destroy(t);
        g_terr = msg;
        return rc;
    };
    if (hipStreamCreate(&t->stream) != hipSuccess) return bail(tfail(t, MDX_ERR_HIP, "stream creation failed"));
    const int64_t N = t->N;
    int rc;
    if ((rc = talloc(t, &t->pos, (size_t)(3 * N))) || (rc = talloc(t, &t->fref, (size_t)(3 * N))) ||
        (rc = talloc(t, &t->eref, (size_t)n_frames)) || (rc = talloc(t, &t->nbr, (size_t)(kT * N))) ||
        (rc = talloc(t, &t->cnt, (size_t)N)) || (rc = talloc(t, &t->rev, (size_t)(kT * N))) ||
        (rc = talloc(t, &t->G, (size_t)(kNF * N))) || (rc = talloc(t, &t->J, (size_t)(kT * 3 * kNF) * (size_t)N)) ||
        (rc = talloc(t, &t->derr, 2)) || (rc = talloc(t, &t->params, kNBlob)) || (rc = talloc(t, &t->grad, kNTrain)) ||
        (rc = talloc(t, &t->m1, kNTrain)) || (rc = talloc(t, &t->m2, kNTrain)))
        return bail(rc);
    // forces in [d][N] layout
    std::vector<float> fsoa((size_t)(3 * N));
    for (int64_t a = 0; a < N; a++)
        for (int d = 0; d < 3; d++) fsoa[(size_t)(d * N + a)] = (float)force3[3 * a + d];
    std::vector<double> wrapped(pos3, pos3 + 3 * N);
    for (double& x : wrapped) x -= box_length * std::floor(x / box_length);
    hipStream_t s = t->stream;
    if (hipMemcpyAsync(t->pos, wrapped.data(), sizeof(double) * 3 * N, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(t->fref, fsoa.data(), sizeof(float) * 3 * N, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(t->eref, energy, sizeof(double) * n_frames, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemsetAsync(t->derr, 0, 2 * sizeof(int32_t), s) != hipSuccess ||
        hipMemsetAsync(t->params, 0, sizeof(float) * kNBlob, s) != hipSuccess ||
        hipMemsetAsync(t->m1, 0, sizeof(float) * kNTrain, s) != hipSuccess ||
        hipMemsetAsync(t->m2, 0, sizeof(float) * kNTrain, s) != hipSuccess)
        return bail(tfail(t, MDX_ERR_HIP, "upload of the training set failed"));
    const int gb = (int)((N + 255) / 256);
    k_tr_nbr<<<gb, 256, 0, s>>>(t->pos, n_frames, n_atoms, box_length, t->nbr, t->cnt, t->derr);
    k_tr_rev<<<gb, 256, 0, s>>>(n_frames, n_atoms, t->nbr, t->cnt, t->rev, t->derr);
    int32_t e2[2];
    if (hipGetLastError() != hipSuccess || hipMemcpyAsync(e2, t->derr, sizeof e2, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return bail(tfail(t, MDX_ERR_HIP, "neighbour search failed"));
    if (e2[0] > 0)
        return bail(tfail(t, MDX_ERR_CAPACITY, "an atom has " + std::to_string(e2[0]) + " neighbours inside 3.0 A (at most " +
                                                std::to_string(kT) + ")"));
    if (e2[1]) return bail(tfail(t, MDX_ERR_STATE, "asymmetric neighbour lists"));
    k_tr_desc<<<(int)((N + 127) / 128), 128, 0, s>>>(t->pos, t->view(), t->G, t->J);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        return bail(tfail(t, MDX_ERR_HIP, "descriptor pass failed"));
    if (hipFuncSetAttribute((const void*)k_tr_fwd, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFwdLds) != hipSuccess ||
        hipFuncSetAttribute((const void*)k_tr_bwd, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kBwdLds) != hipSuccess)
        return bail(tfail(t, MDX_ERR_HIP, "LDS attribute failed"));
    *out = t;
    return MDX_OK;
}

int mdx_sg_trainer_feature_stats(mdx_sg_trainer* t, double* out4x48) {
    if (!t || !out4x48) return tfail(t, MDX_ERR_ARG, "null argument");
    double* d;
    int rc;
    if ((rc = talloc(t, &d, 4 * kNF))) return rc;
    k_tr_stats<<<kNF, 256, 0, t->stream>>>(t->G, t->N, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out4x48, d, sizeof(double) * 4 * kNF, hipMemcpyDeviceToHost, t->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    hipFree(d);
    if (e != hipSuccess) return tfail(t, MDX_ERR_HIP, hipGetErrorString(e));
    return MDX_OK;
}

int mdx_sg_trainer_descriptors(mdx_sg_trainer* t, int32_t frame, float* out) {
    if (!t || !out || frame < 0 || frame >= t->nf) return tfail(t, MDX_ERR_ARG, "bad argument");
    for (int f = 0; f < kNF; f++)
        TCHK(t, hipMemcpyAsync(out + (size_t)f * t->na, t->G + f * t->N + (int64_t)frame * t->na,
                               sizeof(float) * t->na, hipMemcpyDeviceToHost, t->stream));
    TCHK(t, hipStreamSynchronize(t->stream));
    // [f][atom] -> [atom][f]
    std::vector<float> tmp(out, out + (size_t)kNF * t->na);
    for (int32_t a = 0; a < t->na; a++)
        for (int f = 0; f < kNF; f++) out[(size_t)a * kNF + f] = tmp[(size_t)f * t->na + a];
    return MDX_OK;
}

int mdx_sg_trainer_set_params(mdx_sg_trainer* t, const float* params, int64_t len) {
    if (!t || !params || len != kNBlob) return tfail(t, MDX_ERR_ARG, "parameter blob must hold MDX_SG_NPARAM floats");
    TCHK(t, hipMemcpyAsync(t->params, params, sizeof(float) * kNBlob, hipMemcpyHostToDevice, t->stream));
    TCHK(t, hipMemsetAsync(t->m1, 0, sizeof(float) * kNTrain, t->stream));
    TCHK(t, hipMemsetAsync(t->m2, 0, sizeof(float) * kNTrain, t->stream));
    TCHK(t, hipStreamSynchronize(t->stream));
    t->adam_t = 0;
    return MDX_OK;
}

int mdx_sg_trainer_get_params(mdx_sg_trainer* t, float* params, int64_t len) {
    if (!t || !params || len != kNBlob) return tfail(t, MDX_ERR_ARG, "parameter blob must hold MDX_SG_NPARAM floats");
    TCHK(t, hipMemcpyAsync(params, t->params, sizeof(float) * kNBlob, hipMemcpyDeviceToHost, t->stream));
    TCHK(t, hipStreamSynchronize(t->stream));
    return MDX_OK;
}

int mdx_sg_trainer_gradient(mdx_sg_trainer* t, const int32_t* frames, int32_t n, double w_energy, float* grad,
                            double* out3) {
    if (!t) return tfail(nullptr, MDX_ERR_ARG, "null trainer");
    int rc;
    if ((rc = ensure_batch(t, frames, n)) || (rc = forward_pass(t, n, w_energy, out3))) return rc;
    if (grad) {
        if ((rc = gradient_pass(t, n))) return rc;
        TCHK(t, hipMemcpyAsync(grad, t->grad, sizeof(float) * kNTrain, hipMemcpyDeviceToHost, t->stream));
    }
    TCHK(t, hipStreamSynchronize(t->stream));
    return MDX_OK;
}

int mdx_sg_trainer_step(mdx_sg_trainer* t, const int32_t* frames, int32_t n, double w_energy, double lr,
                        double* out3) {
    if (!t) return tfail(nullptr, MDX_ERR_ARG, "null trainer");
    int rc;
    if ((rc = ensure_batch(t, frames, n)) || (rc = forward_pass(t, n, w_energy, out3)) || (rc = gradient_pass(t, n)))
        return rc;
    t->adam_t++;
    const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)t->adam_t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)t->adam_t)));
    k_tr_adam<<<(kNTrain + 255) / 256, 256, 0, t->stream>>>(t->params, t->grad, t->m1, t->m2, (float)lr, c1, c2);
    TCHK(t, hipGetLastError());
    if (!out3) TCHK(t, hipStreamSynchronize(t->stream));
    return MDX_OK;
}

}  // extern "C"
