// gear.hip -- Gear 5th-order predictor-corrector in FP64 on SoA streams
// (cpu/common/Device.cpp:17-134, coefficients common/Constants.h:64-67).
//
// Each kernel is one pass over unit-stride FP64 arrays; the arithmetic of
// every component is written in the reference's operation order so the
// device result equals the FP64 oracle bit for bit (the library is compiled
// with -ffp-contract=off).  The periodic-box path fuses the corrector of
// step t with the predictor of step t+1 and tracks the largest displacement
// since the last list build for the rebuild decision.
#include "mdx_internal.h"

namespace mdx {

namespace {

constexpr int kBlock = 256;

__device__ inline void predict_atom(const DevState& st, int32_t i, const GearCoeffs& gc) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double r = st.r[d][i], v = st.v[d][i], a = st.a[d][i], b = st.b[d][i], c = st.c[d][i];
        st.r[d][i] = r + gc.k[0] * v + gc.k[1] * a + gc.k[2] * b + gc.k[3] * c;
        st.v[d][i] = v + gc.k[0] * a + gc.k[1] * b + gc.k[2] * c;
        st.a[d][i] = a + gc.k[0] * b + gc.k[1] * c;
        st.b[d][i] = b + gc.k[0] * c;
    }
}

// clear (may be null): a word zeroed by the first thread (the corrector's
// wall-hit word, so no separate fill runs before the corrector)
__global__ void __launch_bounds__(kBlock)
k_predict(DevState st, int32_t n, GearCoeffs gc, int32_t* __restrict__ clear) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (clear && i == 0) *clear = 0;
    if (i >= n) return;
    predict_atom(st, i, gc);
}

// corrector of one component; returns the corrected position
__device__ inline void correct1(double f, double mass, double scale, const GearCoeffs& gc,
                                double& r, double& v, double& a, double& b, double& c) {
    double anew = f / mass * scale;
    double corr = anew - a;
    r = r + gc.kc[0] * corr;
    v = v + gc.kc[1] * corr;
    a = anew;
    b = b + gc.kc[2] * corr;
    c = c + gc.kc[3] * corr;
}

// Device.cpp:44-120: corrector, then the walls at +-npos; true when atom i
// hit a wall
__device__ inline bool correct_walled_atom(const DevState& st, int32_t i, const GearCoeffs& gc, const MassTable& m,
                                           double npos, float eloss) {
    bool h = false;
    const double mass = m.mass[st.type[i] & (kTypes - 1)];
    double R[3], V[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double r = st.r[d][i], v = st.v[d][i], a = st.a[d][i], b = st.b[d][i], c = st.c[d][i];
        correct1(st.f[d][i], mass, m.scale, gc, r, v, a, b, c);
        R[d] = r; V[d] = v;
        st.a[d][i] = a; st.b[d][i] = b; st.c[d][i] = c;
    }
    if (R[0] > npos && R[1] > 0) { h = true; V[0] = -V[0] * (double)eloss; R[0] = npos; }
    if (R[0] > npos && R[1] < 0) { h = true; R[0] = -npos; }
    if (R[0] < -npos) { h = true; V[0] = -V[0]; R[0] = -npos; }
#pragma unroll
    for (int d = 1; d < 3; d++)
        if (fabs(R[d]) > npos) { h = true; V[d] = -V[d]; R[d] = npos * (R[d] / fabs(R[d])); }
#pragma unroll
    for (int d = 0; d < 3; d++) { st.r[d][i] = R[d]; st.v[d][i] = V[d]; }
    return h;
}

__global__ void __launch_bounds__(kBlock)
k_correct_walled(DevState st, int32_t n, GearCoeffs gc, MassTable m, double npos, float eloss,
                 int32_t* __restrict__ hit) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool h = i < n && correct_walled_atom(st, i, gc, m, npos, eloss);
    if (__any(h) && (threadIdx.x & (kWave - 1)) == 0) atomicOr(hit, 1);
}

__device__ inline void accelerations_atom(const DevState& st, int32_t i, const MassTable& m) {
    const double mass = m.mass[st.type[i] & (kTypes - 1)];
#pragma unroll
    for (int d = 0; d < 3; d++) st.a[d][i] = st.f[d][i] / mass * m.scale;
}

__global__ void __launch_bounds__(kBlock)
k_accelerations(DevState st, int32_t n, MassTable m, const int32_t* __restrict__ gate) {
    if (gate && *gate == 0) return;
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    accelerations_atom(st, i, m);
}

// ---- sweep members, all in one launch (blockIdx.y = member) ----

// the predictor; block 0 of each member clears the member's wall-hit word for
// the corrector that follows
__global__ void __launch_bounds__(kBlock)
k_sweep_predict(const MemberView* __restrict__ views, int32_t n, GearCoeffs gc) {
    const MemberView& v = views[blockIdx.y];
    if (blockIdx.x == 0 && threadIdx.x == 0) *v.hit = 0;
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    predict_atom(v.st, i, gc);
}

__global__ void __launch_bounds__(kBlock)
k_sweep_correct(const MemberView* __restrict__ views, int32_t n, GearCoeffs gc, MassTable m, double npos,
                float eloss) {
    const MemberView& v = views[blockIdx.y];
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool h = i < n && correct_walled_atom(v.st, i, gc, m, npos, eloss);
    if (__any(h) && (threadIdx.x & (kWave - 1)) == 0) atomicOr(v.hit, 1);
}

// a = F / m for the members whose corrector hit a wall
__global__ void __launch_bounds__(kBlock)
k_sweep_accelerations(const MemberView* __restrict__ views, int32_t n, MassTable m) {
    const MemberView& v = views[blockIdx.y];
    if (*v.hit == 0) return;
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    accelerations_atom(v.st, i, m);
}

// Corrector of step t then predictor of step t+1 on each component, and the
// squared displacement of the predicted position from the list-build
// position, max-reduced over the wave into *maxdisp2_bits (a non-negative
// double's bits order like the value).
__global__ void __launch_bounds__(kBlock)
k_correct_predict(DevState st, int32_t n, GearCoeffs gc, MassTable m, const double* __restrict__ r0x,
                  const double* __restrict__ r0y, const double* __restrict__ r0z,
                  unsigned long long* __restrict__ maxdisp2_bits) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double d2 = 0.0;
    if (i < n) {
        const double mass = m.mass[st.type[i] & (kTypes - 1)];
        const double* r0[3] = {r0x, r0y, r0z};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double r = st.r[d][i], v = st.v[d][i], a = st.a[d][i], b = st.b[d][i], c = st.c[d][i];
            correct1(st.f[d][i], mass, m.scale, gc, r, v, a, b, c);
            double rp = r + gc.k[0] * v + gc.k[1] * a + gc.k[2] * b + gc.k[3] * c;
            double vp = v + gc.k[0] * a + gc.k[1] * b + gc.k[2] * c;
            double ap = a + gc.k[0] * b + gc.k[1] * c;
            double bp = b + gc.k[0] * c;
            st.r[d][i] = rp; st.v[d][i] = vp; st.a[d][i] = ap; st.b[d][i] = bp; st.c[d][i] = c;
            double dd = rp - r0[d][i];
            d2 += dd * dd;
        }
    }
    unsigned long long bits = __double_as_longlong(d2);
    for (int o = kWave / 2; o > 0; o >>= 1) {
        unsigned long long other = __shfl_xor(bits, o, kWave);
        bits = other > bits ? other : bits;
    }
    if ((threadIdx.x & (kWave - 1)) == 0) atomicMax(maxdisp2_bits, bits);
}

__global__ void __launch_bounds__(kBlock)
k_predict_track(DevState st, int32_t n, GearCoeffs gc, const double* __restrict__ r0x,
                const double* __restrict__ r0y, const double* __restrict__ r0z,
                unsigned long long* __restrict__ flags, int32_t spread) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double d2 = 0.0, s2 = 0.0;
    if (i < n) {
        const double* r0[3] = {r0x, r0y, r0z};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double r = st.r[d][i], v = st.v[d][i], a = st.a[d][i], b = st.b[d][i], c = st.c[d][i];
            double rp = r + gc.k[0] * v + gc.k[1] * a + gc.k[2] * b + gc.k[3] * c;
            st.r[d][i] = rp;
            st.v[d][i] = v + gc.k[0] * a + gc.k[1] * b + gc.k[2] * c;
            st.a[d][i] = a + gc.k[0] * b + gc.k[1] * c;
            st.b[d][i] = b + gc.k[0] * c;
            double dd = rp - r0[d][i], ds = rp - r;
            d2 += dd * dd;
            s2 += ds * ds;
        }
    }
    // wave shuffles, then LDS over the block; one atomic per block, spread
    // over flag_spread() addresses so blocks do not serialise on one word
    for (int o = kWave / 2; o > 0; o >>= 1) {
        d2 = fmax(d2, __shfl_xor(d2, o, kWave));
        s2 = fmax(s2, __shfl_xor(s2, o, kWave));
    }
    __shared__ double red[kBlock / kWave][2];
    if ((threadIdx.x & (kWave - 1)) == 0) {
        red[threadIdx.x / kWave][0] = d2;
        red[threadIdx.x / kWave][1] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kBlock / kWave; q++) {
            d2 = fmax(d2, red[q][0]);
            s2 = fmax(s2, red[q][1]);
        }
        unsigned long long* f = flags + 3 * (blockIdx.x % spread);
        atomicMax(&f[0], (unsigned long long)__double_as_longlong(d2));
        atomicMax(&f[1], (unsigned long long)__double_as_longlong(s2));
    }
}

// Periodic box, forces from a separate kernel: corrector of step t, then
// (unless final) the predictor of step t+1 with the displacement from the
// list-build positions and the single-step move max-reduced into the spread
// flags, as k_predict_track does.
// With kGather the force is first gathered from the per-slot gradients (and
// written to st.f on the final step of a run) and atoms farther than sqrt(hard2) from the
// list-build positions are counted as stale, as the bond-order kernels do.
template <bool kFinal, bool kGather>
__global__ void __launch_bounds__(kBlock)
k_correct_predict_track(DevState st, int32_t n, GearCoeffs gc, MassTable m, const double* __restrict__ r0x,
                        const double* __restrict__ r0y, const double* __restrict__ r0z,
                        unsigned long long* __restrict__ flags, int32_t spread, SlotGather g, double hard2,
                        unsigned long long* __restrict__ clear, FlagPublish pub) {
    if (blockIdx.x == 0) {
        if (clear)   // the next step's flag words
            for (int32_t t = threadIdx.x; t < 3 * spread; t += blockDim.x) clear[t] = 0;
        hand_over(pub, 3 * spread);   // the previous step's words to the host (pub.host non-null)
    }
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double d2 = 0.0, s2 = 0.0;
    const double* r0[3] = {r0x, r0y, r0z};
    double F[3] = {0.0, 0.0, 0.0};
    if constexpr (kGather) {
        bool stale = false;
        if (i < n) {
            const int32_t cnt = g.count[i];
            for (int32_t s = 0; s < cnt; s++) {
                const int32_t j = g.list[(int64_t)s * n + i];
                const int32_t b = g.rev[(int64_t)s * n + i];
                if (g.wd) {
                    const double4 u = g.wd[(int64_t)s * n + i], v = g.wd[(int64_t)b * n + j];
                    F[0] += u.x - v.x;
                    F[1] += u.y - v.y;
                    F[2] += u.z - v.z;
                } else {
                    const float4 u = g.w[(int64_t)s * n + i], v = g.w[(int64_t)b * n + j];
                    F[0] += (double)u.x - (double)v.x;
                    F[1] += (double)u.y - (double)v.y;
                    F[2] += (double)u.z - (double)v.z;
                }
            }
            double e2 = 0.0;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                if constexpr (kFinal) st.f[d][i] = F[d];   // read back only after a run
                const double e = st.r[d][i] - r0[d][i];
                e2 += e * e;
            }
            stale = e2 > hard2;
        }
        const unsigned long long nst = __popcll(__ballot(stale));
        if ((threadIdx.x & (kWave - 1)) == 0 && nst)
            atomicAdd(&flags[3 * ((blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave) % spread) + 2], nst);
    }
    if (i < n) {
        const double mass = m.mass[st.type[i] & (kTypes - 1)];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double r = st.r[d][i], v = st.v[d][i], a = st.a[d][i], b = st.b[d][i], c = st.c[d][i];
            correct1(kGather ? F[d] : st.f[d][i], mass, m.scale, gc, r, v, a, b, c);
            if constexpr (!kFinal) {
                const double rp = r + gc.k[0] * v + gc.k[1] * a + gc.k[2] * b + gc.k[3] * c;
                v = v + gc.k[0] * a + gc.k[1] * b + gc.k[2] * c;
                a = a + gc.k[0] * b + gc.k[1] * c;
                b = b + gc.k[0] * c;
                const double dd = rp - r0[d][i], ds = rp - r;
                d2 += dd * dd;
                s2 += ds * ds;
                r = rp;
            }
            st.r[d][i] = r; st.v[d][i] = v; st.a[d][i] = a; st.b[d][i] = b; st.c[d][i] = c;
        }
    }
    if constexpr (!kFinal) {
        for (int o = kWave / 2; o > 0; o >>= 1) {
            d2 = fmax(d2, __shfl_xor(d2, o, kWave));
            s2 = fmax(s2, __shfl_xor(s2, o, kWave));
        }
        __shared__ double red[kBlock / kWave][2];
        if ((threadIdx.x & (kWave - 1)) == 0) {
            red[threadIdx.x / kWave][0] = d2;
            red[threadIdx.x / kWave][1] = s2;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int q = 1; q < kBlock / kWave; q++) {
                d2 = fmax(d2, red[q][0]);
                s2 = fmax(s2, red[q][1]);
            }
            unsigned long long* f = flags + 3 * (blockIdx.x % spread);
            atomicMax(&f[0], (unsigned long long)__double_as_longlong(d2));
            atomicMax(&f[1], (unsigned long long)__double_as_longlong(s2));
        }
    }
}

}  // namespace

hipError_t launch_correct_predict_track(hipStream_t s, const DevState& st, int32_t n, const GearCoeffs& gc,
                                        const MassTable& m, double* const rref[3], uint64_t* flags, bool final_step,
                                        uint64_t* clear_next) {
    if (n <= 0) return hipSuccess;
    auto* f = reinterpret_cast<unsigned long long*>(flags);
    auto* c = reinterpret_cast<unsigned long long*>(clear_next);
    const SlotGather none{nullptr, nullptr, nullptr, nullptr, nullptr};
    if (final_step)
        k_correct_predict_track<true, false><<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(
            st, n, gc, m, rref[0], rref[1], rref[2], f, flag_spread(), none, 0.0, c, FlagPublish{});
    else
        k_correct_predict_track<false, false><<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(
            st, n, gc, m, rref[0], rref[1], rref[2], f, flag_spread(), none, 0.0, c, FlagPublish{});
    return hipGetLastError();
}

hipError_t launch_gather_correct_predict_track(hipStream_t s, const DevState& st, int32_t n, const GearCoeffs& gc,
                                               const MassTable& m, double* const rref[3], uint64_t* flags,
                                               bool final_step, const SlotGather& g, double hard2,
                                               uint64_t* clear_next, const FlagPublish& pub) {
    if (n <= 0) return hipSuccess;
    if (pub.host && !pub.prev) return hipErrorInvalidValue;
    auto* f = reinterpret_cast<unsigned long long*>(flags);
    auto* c = reinterpret_cast<unsigned long long*>(clear_next);
    if (final_step)
        k_correct_predict_track<true, true><<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(
            st, n, gc, m, rref[0], rref[1], rref[2], f, flag_spread(), g, hard2, c, pub);
    else
        k_correct_predict_track<false, true><<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(
            st, n, gc, m, rref[0], rref[1], rref[2], f, flag_spread(), g, hard2, c, pub);
    return hipGetLastError();
}

hipError_t launch_predict_track(hipStream_t s, const DevState& st, int32_t n, const GearCoeffs& gc,
                                double* const rref[3], uint64_t* flags) {
    if (n <= 0) return hipSuccess;
    k_predict_track<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(
        st, n, gc, rref[0], rref[1], rref[2], reinterpret_cast<unsigned long long*>(flags), flag_spread());
    return hipGetLastError();
}

// Device.cpp:19-22 and :48-56: c1 = dt, c2 = c1 dt/2, c3 = c2 dt/3,
// c4 = c3 dt/4 in real (double) from the float DT; corrector factors from the
// float Gear constants.
GearCoeffs gear_coeffs(float dtf) {
    GearCoeffs g;
    double dt = (double)dtf;
    g.k[0] = dt;
    g.k[1] = g.k[0] * dt / 2.0f;
    g.k[2] = g.k[1] * dt / 3.0f;
    g.k[3] = g.k[2] * dt / 4.0f;
    g.kc[0] = 0.15833333333333333333333333333333f * g.k[1];
    g.kc[1] = 0.75f * g.k[1] / g.k[0];
    g.kc[2] = 0.5f * g.k[1] / g.k[2];
    g.kc[3] = 0.08333333333333333333333333333333f * g.k[1] / g.k[3];
    return g;
}

static inline dim3 grid_of(int32_t n) { return dim3((n + kBlock - 1) / kBlock); }

hipError_t launch_predict(hipStream_t s, const DevState& st, int32_t n, const GearCoeffs& gc, int32_t* clear) {
    if (n <= 0) return hipSuccess;
    k_predict<<<grid_of(n), kBlock, 0, s>>>(st, n, gc, clear);
    return hipGetLastError();
}

hipError_t launch_correct_walled(hipStream_t s, const DevState& st, int32_t n, const GearCoeffs& gc,
                                 const MassTable& m, double npos, float eloss, int32_t* hit) {
    if (n <= 0) return hipSuccess;
    k_correct_walled<<<grid_of(n), kBlock, 0, s>>>(st, n, gc, m, npos, eloss, hit);
    return hipGetLastError();
}

hipError_t launch_accelerations(hipStream_t s, const DevState& st, int32_t n, const MassTable& m,
                                const int32_t* gate) {
    if (n <= 0) return hipSuccess;
    k_accelerations<<<grid_of(n), kBlock, 0, s>>>(st, n, m, gate);
    return hipGetLastError();
}

hipError_t launch_sweep_predict(hipStream_t s, const MemberView* views, int32_t members, int32_t n,
                                const GearCoeffs& gc) {
    if (n <= 0 || members <= 0) return hipSuccess;
    k_sweep_predict<<<dim3((n + kBlock - 1) / kBlock, members), kBlock, 0, s>>>(views, n, gc);
    return hipGetLastError();
}

hipError_t launch_sweep_correct(hipStream_t s, const MemberView* views, int32_t members, int32_t n,
                                const GearCoeffs& gc, const MassTable& m, double npos, float eloss) {
    if (n <= 0 || members <= 0) return hipSuccess;
    k_sweep_correct<<<dim3((n + kBlock - 1) / kBlock, members), kBlock, 0, s>>>(views, n, gc, m, npos, eloss);
    return hipGetLastError();
}

hipError_t launch_sweep_accelerations(hipStream_t s, const MemberView* views, int32_t members, int32_t n,
                                      const MassTable& m) {
    if (n <= 0 || members <= 0) return hipSuccess;
    k_sweep_accelerations<<<dim3((n + kBlock - 1) / kBlock, members), kBlock, 0, s>>>(views, n, m);
    return hipGetLastError();
}

hipError_t launch_correct_predict(hipStream_t s, const DevState& st, int32_t n, const GearCoeffs& gc,
                                  const MassTable& m, double* const rref[3], uint32_t* maxdisp2_bits) {
    if (n <= 0) return hipSuccess;
    k_correct_predict<<<grid_of(n), kBlock, 0, s>>>(st, n, gc, m, rref[0], rref[1], rref[2],
                                                    reinterpret_cast<unsigned long long*>(maxdisp2_bits));
    return hipGetLastError();
}

}  // namespace mdx
