// reduce.hip -- FP64 reductions for the diagnostics of
// common/DebugCalculations.cpp:46-124 and the list schedule of
// cpu/cpumd/Host.cpp:245-271.  Sums use fixed block partitions and a fixed
// final order (wave shuffles, then one block over the partials), so results
// are deterministic run to run; maxima use integer atomics on the bits of
// non-negative doubles.
#include "mdx_internal.h"

namespace mdx {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 1024;

__device__ inline double wave_sum(double x) {
    for (int o = kWave / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
    return x;
}

// block-wide sum of K values; result valid in thread 0
template <int K> __device__ inline void block_sum(double (&v)[K]) {
    __shared__ double part[kBlock / kWave][K];
    int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = wave_sum(v[k]);
        if (lane == 0) part[w][k] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            double s = 0;
            for (int q = 0; q < kBlock / kWave; q++) s += part[q][k];
            v[k] = s;
        }
    }
    __syncthreads();
}

// pass 1: per block {sum m v^2, sum m v, sum m, sum m r}
__global__ void __launch_bounds__(kBlock)
k_kin1(DevState st, int32_t n, MassTable m, double* __restrict__ part) {
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        double mass = m.mass[st.type[i] & (kTypes - 1)];
        double vx = st.v[0][i], vy = st.v[1][i], vz = st.v[2][i];
        v[0] += mass * (vx * vx + vy * vy + vz * vz);
        v[1] += mass * vx;
        v[2] += mass * vy;
        v[3] += mass * vz;
        v[4] += mass;
        v[5] += mass * st.r[0][i];
        v[6] += mass * st.r[1][i];
        v[7] += mass * st.r[2][i];
    }
    block_sum<8>(v);
    if (threadIdx.x == 0)
        for (int k = 0; k < 8; k++) part[blockIdx.x * 8 + k] = v[k];
}

// pass 2 (one block): totals into tot[0..7]
__global__ void __launch_bounds__(kBlock)
k_kin2(const double* __restrict__ part, int32_t nb, double* __restrict__ tot) {
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int32_t b = threadIdx.x; b < nb; b += kBlock)
        for (int k = 0; k < 8; k++) v[k] += part[b * 8 + k];
    block_sum<8>(v);
    if (threadIdx.x == 0)
        for (int k = 0; k < 8; k++) tot[k] = v[k];
}

// pass 3: per block sum m |v - v_cm|^2
__global__ void __launch_bounds__(kBlock)
k_kin3(DevState st, int32_t n, MassTable m, const double* __restrict__ tot, double* __restrict__ part) {
    double cx = tot[1] / tot[4], cy = tot[2] / tot[4], cz = tot[3] / tot[4];
    double v[1] = {0};
    for (int32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        double mass = m.mass[st.type[i] & (kTypes - 1)];
        double ux = st.v[0][i] - cx, uy = st.v[1][i] - cy, uz = st.v[2][i] - cz;
        v[0] += mass * (ux * ux + uy * uy + uz * uz);
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

// pass 4 (one block): out8 = {EK, T, COM, P}
__global__ void __launch_bounds__(kBlock)
k_kin4(const double* __restrict__ part, int32_t nb, const double* __restrict__ tot, int32_t n,
       double kboltz, double* __restrict__ out) {
    double v[1] = {0};
    for (int32_t b = threadIdx.x; b < nb; b += kBlock) v[0] += part[b];
    block_sum<1>(v);
    if (threadIdx.x == 0) {
        out[0] = tot[0] / 2;
        out[1] = v[0] / (3.0 * (double)n * kboltz);
        out[2] = tot[5] / tot[4];
        out[3] = tot[6] / tot[4];
        out[4] = tot[7] / tot[4];
        out[5] = tot[1];
        out[6] = tot[2];
        out[7] = tot[3];
    }
}

__global__ void __launch_bounds__(kBlock)
k_sum1(const double* __restrict__ x, int32_t n, double* __restrict__ part) {
    double v[1] = {0};
    for (int32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) v[0] += x[i];
    block_sum<1>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

__global__ void __launch_bounds__(kBlock)
k_sum2(const double* __restrict__ part, int32_t nb, double* __restrict__ out) {
    double v[1] = {0};
    for (int32_t b = threadIdx.x; b < nb; b += kBlock) v[0] += part[b];
    block_sum<1>(v);
    if (threadIdx.x == 0) *out = v[0];
}

// max |v|^2 and max |coordinate| as bit patterns (both non-negative); one
// workgroup (kOne) writes them, several max-reduce into words cleared before
template <bool kOne>
__global__ void __launch_bounds__(kBlock)
k_ready(DevState st, int32_t n, unsigned long long* __restrict__ out) {
    double v2 = 0, c = 0;
    for (int32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        double vx = st.v[0][i], vy = st.v[1][i], vz = st.v[2][i];
        v2 = fmax(v2, vx * vx + vy * vy + vz * vz);
        c = fmax(c, fmax(fabs(st.r[0][i]), fmax(fabs(st.r[1][i]), fabs(st.r[2][i]))));
    }
    unsigned long long a = __double_as_longlong(v2), b = __double_as_longlong(c);
    for (int o = kWave / 2; o > 0; o >>= 1) {
        unsigned long long x = __shfl_xor(a, o, kWave), y = __shfl_xor(b, o, kWave);
        a = x > a ? x : a;
        b = y > b ? y : b;
    }
    if constexpr (kOne) {
        __shared__ unsigned long long wm[kBlock / kWave][2];
        if ((threadIdx.x & (kWave - 1)) == 0) {
            wm[threadIdx.x / kWave][0] = a;
            wm[threadIdx.x / kWave][1] = b;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < kBlock / kWave; w++) {
                a = wm[w][0] > a ? wm[w][0] : a;
                b = wm[w][1] > b ? wm[w][1] : b;
            }
            out[0] = a;
            out[1] = b;
        }
    } else if ((threadIdx.x & (kWave - 1)) == 0) {
        atomicMax(&out[0], a);
        atomicMax(&out[1], b);
    }
}

// w[i] = r_s . F_s of slot s = map[i] (or i), the virial of an isolated
// system atom by atom
__global__ void __launch_bounds__(kBlock)
k_rf(DevState pos, DevState frc, const int32_t* __restrict__ map, int32_t n, double* __restrict__ w) {
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int32_t s = map ? map[i] : i;
    w[i] = pos.r[0][s] * frc.f[0][s] + pos.r[1][s] * frc.f[1][s] + pos.r[2][s] * frc.f[2][s];
}

inline int32_t blocks_for(int32_t n) {
    int32_t b = (n + kBlock - 1) / kBlock;
    return b < 1 ? 1 : (b > kMaxBlocks ? kMaxBlocks : b);
}

}  // namespace

int32_t kinetics_blocks(int32_t n) { return blocks_for(n); }

// partials: at least 8 * kinetics_blocks(n) + 8 doubles
hipError_t launch_kinetics(hipStream_t s, const DevState& st, int32_t n, const MassTable& m,
                           double* partials, int32_t nb, double* out8, double kboltz) {
    double* tot = partials + 8 * nb;
    k_kin1<<<nb, kBlock, 0, s>>>(st, n, m, partials);
    k_kin2<<<1, kBlock, 0, s>>>(partials, nb, tot);
    k_kin3<<<nb, kBlock, 0, s>>>(st, n, m, tot, partials);
    k_kin4<<<1, kBlock, 0, s>>>(partials, nb, tot, n, kboltz, out8);
    return hipGetLastError();
}

hipError_t launch_sum(hipStream_t s, const double* x, int32_t n, double* partials, int32_t nb,
                      double* out) {
    k_sum1<<<nb, kBlock, 0, s>>>(x, n, partials);
    k_sum2<<<1, kBlock, 0, s>>>(partials, nb, out);
    return hipGetLastError();
}

hipError_t launch_rf_virial(hipStream_t s, const DevState& pos, const DevState& frc, const int32_t* map, int32_t n,
                            double* w) {
    if (n <= 0) return hipSuccess;
    k_rf<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(pos, frc, map, n, w);
    return hipGetLastError();
}

// up to this many atoms one workgroup reduces them (no clearing, no atomics)
constexpr int32_t kReadyOneBlock = 65536;

hipError_t launch_ready_list(hipStream_t s, const DevState& st, int32_t n, uint64_t* out2_bits) {
    auto* out = reinterpret_cast<unsigned long long*>(out2_bits);
    if (n <= kReadyOneBlock) {
        k_ready<true><<<1, kBlock, 0, s>>>(st, n, out);
    } else {
        hipError_t e = hipMemsetAsync(out2_bits, 0, 2 * sizeof(uint64_t), s);
        if (e != hipSuccess) return e;
        k_ready<false><<<blocks_for(n), kBlock, 0, s>>>(st, n, out);
    }
    return hipGetLastError();
}

}  // namespace mdx
