// surrogate.hip -- learned surrogate of the Tersoff silicon bond-order force
// in the periodic box: it takes the place of the bond-order kernels
// (tersoff.hip; in the reference, manyBodyForces, cpu/common/BondOrderFuncs.h:283-397)
// for the atoms whose environment the network covers, and plain Tersoff
// per-atom terms for the others.
//
//   E = sum_i E_i,  E_i = NN((G_i - mean) * istd)            (covered atoms)
//                   E_i = 1/2 sum_j V_ij (Tersoff)            (flagged atoms)
//   NN(x) = w3 . tanh(W2 tanh(W1 x + b1) + b2) + b3           48 -> 64 -> 64 -> 1
//
// An atom is flagged when it has more than kS list entries or any of its
// standardized descriptors lies outside the training range [lo, hi].
//
// k_sg_eval, one thread per atom, 4 waves of 64 atoms per block:
//   1. descriptors G_i in FP32 registers (sg_terms.h), standardized, stored
//      as BF16 rows in LDS;
//   2. the network and its gradient on MFMA (v_mfma_f32_16x16x32_bf16, FP32
//      accumulation), 16 atoms per tile, hidden units kept in the MFMA
//      result layout between layers: the 16 values a lane holds after one
//      layer are exactly the 16 K values it feeds the next, because the
//      weights are stored with their input columns permuted to match
//      (sg_pack below);
//   3. dE/dG back to one atom per lane through LDS, then the reverse pass of
//      the descriptor gives w_is = dE_i/d(x_j - x_i) per list slot.
// k_sg_gather: F_i = sum_s w_is - sum_s w_{j_s, rev_s} in FP64, no atomics.
#include "mdx_internal.h"
#include "sg_terms.h"
#include "tersoff_terms.h"

namespace mdx {

namespace {

using namespace ts;
using namespace sg;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;            // 4 waves, 64 atoms each
constexpr int kS = 6;                  // list entries held in registers
constexpr int kFrags = 30;             // 16 B MFMA A-operand fragments per lane
constexpr int kFragL1 = 0, kFragL2 = 8, kFragB2 = 16, kFragB1 = 24;
constexpr int kHRow = kNF + 1;         // staged row: dE/dG (48) and E
constexpr int kRegion = kWave * kHRow * 4;       // bytes per wave
constexpr int kGOff = kRegion - kWave * 128;     // BF16 descriptor rows (128 B) at the region's end
constexpr int kLds = kFrags * kWave * 16 + (kBlock / kWave) * kRegion;
static_assert(kGOff >= 0 && kRegion % 16 == 0 && kGOff % 16 == 0, "staging layout");
// small vectors: b1, b2, w3 (64 each), b3, mean, istd, lo, hi (48 each)
constexpr int kVB1 = 0, kVB2 = 64, kVW3 = 128, kVB3 = 192, kVMean = 193, kVIstd = 241, kVLo = 289, kVHi = 337;
constexpr int kVec = 385;

__device__ inline bf16x8 as_bf16x8(uint4 v) {
    bf16x8 r;
    __builtin_memcpy(&r, &v, 16);
    return r;
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// two floats -> two BF16 (round to nearest even, v_cvt_pk_bf16_f32), a low
__device__ inline uint32_t pack2(float a, float b) {
    const f32x2 v = {a, b};
    const bf16x2 h = __builtin_convertvector(v, bf16x2);
    uint32_t u;
    __builtin_memcpy(&u, &h, 4);
    return u;
}

// 8 values a[2ks + t/4][t%4], t = 0..7, as one BF16 B operand
__device__ inline bf16x8 operand(const float (&a)[4][4], int ks) {
    uint4 v;
    v.x = pack2(a[2 * ks][0], a[2 * ks][1]);
    v.y = pack2(a[2 * ks][2], a[2 * ks][3]);
    v.z = pack2(a[2 * ks + 1][0], a[2 * ks + 1][1]);
    v.w = pack2(a[2 * ks + 1][2], a[2 * ks + 1][3]);
    return as_bf16x8(v);
}

__global__ void __launch_bounds__(kBlock)
k_sg_eval(Pos3 P, const int32_t* __restrict__ list, const int32_t* __restrict__ count, int32_t n,
          const uint4* __restrict__ frag, const float* __restrict__ vec, Tp<float> tp, float4* __restrict__ w,
          double* __restrict__ e_atom, int32_t* __restrict__ nflag) {
    extern __shared__ __align__(16) unsigned char lds[];
    uint4* sfrag = reinterpret_cast<uint4*>(lds);
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
    unsigned char* region = lds + kFrags * kWave * 16 + wv * kRegion;
    for (int q = tid; q < kFrags * kWave; q += kBlock) sfrag[q] = frag[q];

    // 1. descriptors of this lane's atom
    const int32_t i = blockIdx.x * kBlock + tid;
    const bool live = i < n;
    const int32_t cnt = live ? count[i] : 0;
    bool flag = cnt > kS;
    Nb nb[kS];
#pragma unroll
    for (int s = 0; s < kS; s++) {
        nb[s] = empty_nb();
        if (!flag && s < cnt) {
            double d[3];
            P.d(i, list[(int64_t)s * n + i], d);
            nb[s] = make_nb(d[0], d[1], d[2]);
        }
    }
    float G[kNF];
    forward(nb, G);
    bool out = false;   // a standardized value outside [lo, hi] (or NaN); no branches
#pragma unroll
    for (int f = 0; f < kNF; f++) {
        const float gs = (G[f] - vec[kVMean + f]) * vec[kVIstd + f];
        out |= !(gs >= vec[kVLo + f]) | !(gs <= vec[kVHi + f]);
        G[f] = gs;
    }
    flag |= out;
    if (flag || !live) {
#pragma unroll
        for (int f = 0; f < kNF; f++) G[f] = 0.0f;
    }
    {   // BF16 row, 16 B chunks swizzled by (row & 7); features 48..63 are zero
        uint4* grow = reinterpret_cast<uint4*>(region + kGOff) + lane * 8;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            uint4 v = {0u, 0u, 0u, 0u};
            if (c < 6) {
                v.x = pack2(G[8 * c + 0], G[8 * c + 1]);
                v.y = pack2(G[8 * c + 2], G[8 * c + 3]);
                v.z = pack2(G[8 * c + 4], G[8 * c + 5]);
                v.w = pack2(G[8 * c + 6], G[8 * c + 7]);
            }
            grow[c ^ (lane & 7)] = v;
        }
    }
    __syncthreads();

    // 2. network on MFMA tiles of 16 atoms: lane = 16 g + nl serves atom
    //    nl of the tile and hidden rows 4 g .. 4 g + 3 of each 16-row block
    {
        const int g = lane >> 4, nl = lane & 15;
        float b1v[4][4], b2v[4][4], w3v[4][4], isv[3][4];
#pragma unroll
        for (int mt = 0; mt < 4; mt++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                b1v[mt][r] = vec[kVB1 + mt * 16 + 4 * g + r];
                b2v[mt][r] = vec[kVB2 + mt * 16 + 4 * g + r];
                w3v[mt][r] = vec[kVW3 + mt * 16 + 4 * g + r];
                if (mt < 3) isv[mt][r] = vec[kVIstd + mt * 16 + 4 * g + r];
            }
        const float b3 = vec[kVB3];
        const uint4* gbase = reinterpret_cast<const uint4*>(region + kGOff);
        float* hbase = reinterpret_cast<float*>(region);
        for (int tile = 0; tile < kWave / 16; tile++) {
            const int row = tile * 16 + nl;
            const bf16x8 x0 = as_bf16x8(gbase[row * 8 + ((0 * 4 + g) ^ (row & 7))]);
            const bf16x8 x1 = as_bf16x8(gbase[row * 8 + ((1 * 4 + g) ^ (row & 7))]);
            float a1[4][4], a2[4][4], t[4][4];
#pragma unroll
            for (int mt = 0; mt < 4; mt++) {
                f32x4 z = {b1v[mt][0], b1v[mt][1], b1v[mt][2], b1v[mt][3]};
                z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(sfrag[(kFragL1 + 2 * mt) * kWave + lane]), x0, z, 0, 0, 0);
                z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(sfrag[(kFragL1 + 2 * mt + 1) * kWave + lane]), x1, z, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; r++) a1[mt][r] = tanh_f(z[r]);
            }
            const bf16x8 y0 = operand(a1, 0), y1 = operand(a1, 1);
            float e = 0.0f;
#pragma unroll
            for (int mt = 0; mt < 4; mt++) {
                f32x4 z = {b2v[mt][0], b2v[mt][1], b2v[mt][2], b2v[mt][3]};
                z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(sfrag[(kFragL2 + 2 * mt) * kWave + lane]), y0, z, 0, 0, 0);
                z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(sfrag[(kFragL2 + 2 * mt + 1) * kWave + lane]), y1, z, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    a2[mt][r] = tanh_f(z[r]);
                    e = fmaf(w3v[mt][r], a2[mt][r], e);
                    t[mt][r] = w3v[mt][r] * (1.0f - a2[mt][r] * a2[mt][r]);   // dE/dz2
                }
            }
            e += __shfl_xor(e, 16);
            e += __shfl_xor(e, 32);
            const bf16x8 v0 = operand(t, 0), v1 = operand(t, 1);
#pragma unroll
            for (int mt = 0; mt < 4; mt++) {
                f32x4 q = {0.0f, 0.0f, 0.0f, 0.0f};
                q = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(sfrag[(kFragB2 + 2 * mt) * kWave + lane]), v0, q, 0, 0, 0);
                q = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(sfrag[(kFragB2 + 2 * mt + 1) * kWave + lane]), v1, q, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; r++) t[mt][r] = q[r] * (1.0f - a1[mt][r] * a1[mt][r]);   // dE/dz1
            }
            const bf16x8 u0 = operand(t, 0), u1 = operand(t, 1);
#pragma unroll
            for (int mt = 0; mt < 3; mt++) {
                f32x4 hh = {0.0f, 0.0f, 0.0f, 0.0f};
                hh = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(sfrag[(kFragB1 + 2 * mt) * kWave + lane]), u0, hh, 0, 0, 0);
                hh = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(sfrag[(kFragB1 + 2 * mt + 1) * kWave + lane]), u1, hh, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; r++) hbase[row * kHRow + mt * 16 + 4 * g + r] = hh[r] * isv[mt][r];
            }
            if (g == 0) hbase[row * kHRow + kNF] = e + b3;
        }
    }
    __syncthreads();

    // 3. reverse pass of the descriptor: per-slot gradients
    if (live) {
        const float* hr = reinterpret_cast<const float*>(region) + lane * kHRow;
        double e;
        if (!flag) {
            float h[kNF];
#pragma unroll
            for (int f = 0; f < kNF; f++) h[f] = hr[f];
            e = hr[kNF];
            float ws[kS][3];
            backward(nb, h, ws);
#pragma unroll
            for (int s = 0; s < kS; s++)
                if (s < cnt) w[(int64_t)s * n + i] = make_float4(ws[s][0], ws[s][1], ws[s][2], 0.0f);
        } else if (cnt <= kS) {   // plain Tersoff for this centre (tersoff_terms.h)
            float wr[kS][3];
            e = centre_reg<float, kS>(P, tp, list, n, i, cnt, wr);
#pragma unroll
            for (int s = 0; s < kS; s++)
                if (s < cnt) w[(int64_t)s * n + i] = make_float4(wr[s][0], wr[s][1], wr[s][2], 0.0f);
        } else {
            e = centre_mem<float>(P, tp, list, n, i, cnt, w);
        }
        if (e_atom) e_atom[i] = e;
    }
    const unsigned long long nf = __popcll(__ballot(live && flag));
    if (lane == 0 && nf) atomicAdd(nflag, (int32_t)nf);
}

__global__ void __launch_bounds__(kBlock)
k_sg_rev(const int32_t* __restrict__ list, const int32_t* __restrict__ count, int32_t n, uint8_t* __restrict__ rev,
         int32_t* __restrict__ err) {
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int32_t cnt = count[i];
    for (int32_t a = 0; a < cnt; a++) {
        const int32_t j = list[(int64_t)a * n + i];
        const int32_t cj = count[j];
        int32_t b = 0;
        while (b < cj && list[(int64_t)b * n + j] != i) b++;
        if (b == cj) {
            atomicOr(err, 1);
            b = 0;
        }
        rev[(int64_t)a * n + i] = (uint8_t)b;
    }
}

// forces from per-slot vectors: FP32 (w) or FP64 (wd, when non-null)
__global__ void __launch_bounds__(kBlock)
k_sg_gather(Pos3 P, const int32_t* __restrict__ list, const int32_t* __restrict__ count,
            const uint8_t* __restrict__ rev, int32_t n, const float4* __restrict__ w,
            const double4* __restrict__ wd, double* __restrict__ fx,
            double* __restrict__ fy, double* __restrict__ fz, const double* __restrict__ r0x,
            const double* __restrict__ r0y, const double* __restrict__ r0z, double hard2,
            unsigned long long* __restrict__ flags, int32_t spread) {
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    bool stale = false;
    if (i < n) {
        const int32_t cnt = count[i];
        double F[3] = {0.0, 0.0, 0.0};
        for (int32_t a = 0; a < cnt; a++) {
            const int32_t j = list[(int64_t)a * n + i];
            const int32_t b = rev[(int64_t)a * n + i];
            if (wd) {
                const double4 u = wd[(int64_t)a * n + i], v = wd[(int64_t)b * n + j];
                F[0] += u.x - v.x;
                F[1] += u.y - v.y;
                F[2] += u.z - v.z;
            } else {
                const float4 u = w[(int64_t)a * n + i], v = w[(int64_t)b * n + j];
                F[0] += (double)u.x - (double)v.x;
                F[1] += (double)u.y - (double)v.y;
                F[2] += (double)u.z - (double)v.z;
            }
        }
        fx[i] = F[0];
        fy[i] = F[1];
        fz[i] = F[2];
        if (r0x) {
            const double e0 = P.x[i] - r0x[i], e1 = P.y[i] - r0y[i], e2 = P.z[i] - r0z[i];
            stale = e0 * e0 + e1 * e1 + e2 * e2 > hard2;
        }
    }
    if (flags) {
        const unsigned long long nst = __popcll(__ballot(stale));
        if ((threadIdx.x & (kWave - 1)) == 0 && nst)
            atomicAdd(&flags[3 * ((blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave) % spread) + 2], nst);
    }
}

// per-atom virial of a sum of atom energies e_i(x_j - x_i) held as per-slot
// gradients w_is = de_i/d(x_j - x_i): W_i = -sum_s (x_j - x_i) . w_is (FP64)
__global__ void __launch_bounds__(kBlock)
k_slot_virial(Pos3 P, const int32_t* __restrict__ list, const int32_t* __restrict__ count, int32_t n,
              const float4* __restrict__ w, const double4* __restrict__ wd, double* __restrict__ w_atom) {
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int32_t cnt = count[i];
    double v = 0.0;
    for (int32_t a = 0; a < cnt; a++) {
        const int32_t j = list[(int64_t)a * n + i];
        double d[3];
        P.d(i, j, d);
        if (wd) {
            const double4 u = wd[(int64_t)a * n + i];
            v -= d[0] * u.x + d[1] * u.y + d[2] * u.z;
        } else {
            const float4 u = w[(int64_t)a * n + i];
            v -= d[0] * (double)u.x + d[1] * (double)u.y + d[2] * (double)u.z;
        }
    }
    w_atom[i] = v;
}

inline int grid_for(int32_t n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

hipError_t launch_slot_virial(hipStream_t s, const DevState& st, double L, const int32_t* list, const int32_t* count,
                              int32_t n, const float4* w, const double4* wd, double* w_atom) {
    if (n <= 0) return hipSuccess;
    k_slot_virial<<<grid_for(n), kBlock, 0, s>>>(pos3(st, L), list, count, n, w, wd, w_atom);
    return hipGetLastError();
}

// Host-side packing of the FP32 parameter blob into the kernel's BF16 MFMA
// fragments (fragment f, lane l: 8 values) and the small FP32 vectors.
//   k index kappa = 32 ks + 8 g + t (g = l / 16, t < 8) carries hidden unit
//   p(kappa) = 16 (2 ks + t / 4) + 4 g + t % 4: the unit that lane holds
//   after the previous layer.
//   L1[mt][ks]: W1[16 mt + l%16][kappa]             (features; 0 past 48)
//   L2[mt][ks]: W2[16 mt + l%16][p(kappa)]
//   B2[mt][ks]: W2[p(kappa)][16 mt + l%16]           (W2^T)
//   B1[mt][ks]: W1[p(kappa)][16 mt + l%16], mt < 3   (W1^T)
void sg_pack(const float* blob, uint16_t* frag, float* vec) {
    auto at = [&](int f, int l, int t) -> uint16_t& { return frag[((size_t)f * kWave + l) * 8 + t]; };
    const float* W1 = blob + kOffW1;
    const float* W2 = blob + kOffW2;
    for (int l = 0; l < kWave; l++) {
        const int g = l >> 4, col = l & 15;
        for (int ks = 0; ks < 2; ks++)
            for (int t = 0; t < 8; t++) {
                const int kappa = 32 * ks + 8 * g + t;
                const int p = 16 * (2 * ks + t / 4) + 4 * g + t % 4;
                for (int mt = 0; mt < 4; mt++) {
                    const int row = 16 * mt + col;
                    at(kFragL1 + 2 * mt + ks, l, t) = (uint16_t)(kappa < kNF ? bf16_bits(W1[row * kNF + kappa]) : 0u);
                    at(kFragL2 + 2 * mt + ks, l, t) = (uint16_t)bf16_bits(W2[row * kNH + p]);
                    at(kFragB2 + 2 * mt + ks, l, t) = (uint16_t)bf16_bits(W2[p * kNH + row]);
                    if (mt < 3) at(kFragB1 + 2 * mt + ks, l, t) = (uint16_t)bf16_bits(W1[p * kNF + row]);
                }
            }
    }
    for (int k = 0; k < kNH; k++) {
        vec[kVB1 + k] = blob[kOffB1 + k];
        vec[kVB2 + k] = blob[kOffB2 + k];
        vec[kVW3 + k] = blob[kOffW3 + k];
    }
    vec[kVB3] = blob[kOffB3];
    for (int f = 0; f < kNF; f++) {
        vec[kVMean + f] = blob[kOffMean + f];
        vec[kVIstd + f] = blob[kOffIstd + f];
        vec[kVLo + f] = blob[kOffLo + f];
        vec[kVHi + f] = blob[kOffHi + f];
    }
}

size_t sg_frag_u16() { return (size_t)kFrags * kWave * 8; }
size_t sg_vec_len() { return kVec; }
int32_t sg_slots() { return kS; }

hipError_t launch_sg_rev(hipStream_t s, const int32_t* list, const int32_t* count, int32_t n, uint8_t* rev,
                         int32_t* err) {
    if (n <= 0) return hipSuccess;
    k_sg_rev<<<grid_for(n), kBlock, 0, s>>>(list, count, n, rev, err);
    return hipGetLastError();
}

hipError_t launch_sg_eval(hipStream_t s, const DevState& st, double L, const TersoffParams& p, const int32_t* list,
                          const int32_t* count, int32_t n, const uint16_t* frag, const float* vec, float4* w,
                          double* e_atom, int32_t* nflag) {
    if (n <= 0) return hipSuccess;
    // once per process, thread-safe
    static const hipError_t attr =
        hipFuncSetAttribute((const void*)k_sg_eval, hipFuncAttributeMaxDynamicSharedMemorySize, kLds);
    if (attr != hipSuccess) return attr;
    k_sg_eval<<<grid_for(n), kBlock, kLds, s>>>(pos3(st, L), list, count, n, reinterpret_cast<const uint4*>(frag),
                                                vec, cast_params<float>(p), w, e_atom, nflag);
    return hipGetLastError();
}

hipError_t launch_sg_forces(hipStream_t s, const DevState& st, double L, const TersoffParams& p, const int32_t* list,
                            const int32_t* count, const uint8_t* rev, int32_t n, const uint16_t* frag,
                            const float* vec, float4* w, double* e_atom, int32_t* nflag, double* const rref[3],
                            double hard2, uint64_t* flags, const DevState& out) {
    if (n <= 0) return hipSuccess;
    hipError_t e = launch_sg_eval(s, st, L, p, list, count, n, frag, vec, w, e_atom, nflag);
    if (e != hipSuccess) return e;
    return launch_slot_gather(s, st, L, list, count, rev, n, w, nullptr, rref, hard2, flags, out);
}

hipError_t launch_slot_gather(hipStream_t s, const DevState& st, double L, const int32_t* list, const int32_t* count,
                              const uint8_t* rev, int32_t n, const float4* w, const double4* wd,
                              double* const rref[3], double hard2, uint64_t* flags, const DevState& out) {
    if (n <= 0) return hipSuccess;
    const double* r0[3] = {rref ? rref[0] : nullptr, rref ? rref[1] : nullptr, rref ? rref[2] : nullptr};
    k_sg_gather<<<grid_for(n), kBlock, 0, s>>>(pos3(st, L), list, count, rev, n, w, wd, out.f[0], out.f[1], out.f[2],
                                               r0[0], r0[1], r0[2], hard2,
                                               reinterpret_cast<unsigned long long*>(flags), flag_spread());
    return hipGetLastError();
}

}  // namespace mdx
