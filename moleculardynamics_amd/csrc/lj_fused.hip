// lj_fused.hip -- periodic-box performance path: the step in one kernel.
//
//  k_nlist (at each cell-list rebuild): for every atom, the candidates of the
//    27-cell neighbourhood within the list radius rs, in the reference's
//    candidate order (cells in (dx,dy,dz) lexicographic order, ascending slot
//    within a cell).  Entries are slot | image << 27, image = 9(wx+1) +
//    3(wy+1) + (wz+1) for the periodic wrap w = -1, 0, +1 of the neighbour
//    cell per axis (shift w L); the list is stored
//    column-major ([k][atom]) so a wave's loads of entry k are coalesced.
//  k_step (every step): LJ force of atom i over its list (FP64 displacement,
//    FP32 12-6 force law, FP64 accumulation), then the Gear corrector of step
//    t and the predictor of step t+1, all in registers.  Positions are
//    double-buffered (read r(t) from A, write r(t+1) to B) so no thread
//    reads a position another thread has already advanced; the force never
//    goes to HBM unless asked for.
// HBM per atom and step: Gear state r,v,a,b,c read and written (240 B), the
// list-build reference position (24 B) and the list entries (4 B each).
#include "mdx_internal.h"

namespace mdx {

namespace {

constexpr int kBlock = 256;
constexpr int kImageShift = 27;
constexpr uint32_t kSlotMask = (1u << kImageShift) - 1;
constexpr uint32_t kNoShift = 13;  // image code of (0,0,0)

__global__ void __launch_bounds__(kBlock)
k_nlist(const double* __restrict__ rx, const double* __restrict__ ry, const double* __restrict__ rz,
        const int32_t* __restrict__ cell_of, GridView g, double rl2, int32_t n, int32_t cap,
        uint32_t* __restrict__ list, int32_t* __restrict__ count, int32_t* __restrict__ overflow) {
    int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const double xi = rx[s], yi = ry[s], zi = rz[s];
    const int32_t own = cell_of[s];
    const int32_t cz = own % g.g[2], cy = (own / g.g[2]) % g.g[1], cx = own / (g.g[2] * g.g[1]);
    int32_t k = 0;
    for (int dx = -1; dx <= 1; dx++) {
        int32_t x = cx + dx, wx = 0;   // w*: periodic wrap of the neighbour cell (-1, 0, +1)
        if (x < 0) { x += g.g[0]; wx = -1; } else if (x >= g.g[0]) { x -= g.g[0]; wx = 1; }
        const double sx = wx * g.L;
        for (int dy = -1; dy <= 1; dy++) {
            int32_t y = cy + dy, wy = 0;
            if (y < 0) { y += g.g[1]; wy = -1; } else if (y >= g.g[1]) { y -= g.g[1]; wy = 1; }
            const double sy = wy * g.L;
            for (int dz = -1; dz <= 1; dz++) {
                int32_t z = cz + dz, wz = 0;
                if (z < 0) { z += g.g[2]; wz = -1; } else if (z >= g.g[2]) { z -= g.g[2]; wz = 1; }
                const double sz = wz * g.L;
                const uint32_t code = (uint32_t)(9 * (wx + 1) + 3 * (wy + 1) + (wz + 1));
                const int32_t c = (x * g.g[1] + y) * g.g[2] + z;
                const int32_t pe = g.start[c + 1];
                for (int32_t p = g.start[c]; p < pe; p++) {
                    if (p == s) continue;
                    const double ddx = (xi - rx[p]) - sx, ddy = (yi - ry[p]) - sy, ddz = (zi - rz[p]) - sz;
                    const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                    if (r2 >= rl2) continue;
                    if (k < cap) list[(int64_t)k * n + s] = (uint32_t)p | (code << kImageShift);
                    k++;
                }
            }
        }
    }
    count[s] = k < cap ? k : cap;
    if (k > cap) atomicMax(overflow, k);
}

struct StepArgs {
    DevState st;            // st.r = positions r(t) (read), Gear state v,a,b,c (read/write)
    double* rn[3];          // r(t+1) out
    const double* r0[3];    // positions at the last list build
    const uint32_t* list;
    const int32_t* count;
    double L;
    double rc2;
    float inv_sig2f, coeff_f;
    double sig2, coeff;
    double mass, scale;
    GearCoeffs gc;
    double hard2;           // list validity: displacement^2 must stay below this
    unsigned long long* flags;  // [0] max disp^2 (bits), [1] max step^2 (bits), [2] stale count
};

__device__ inline double image(uint32_t code, int axis, double L) {
    // code = 9 ix + 3 iy + iz with i* in {0,1,2}; shift = (i* - 1) L
    const uint32_t q = axis == 0 ? code / 9 : (axis == 1 ? (code / 3) % 3 : code % 3);
    return ((double)q - 1.0) * L;
}

template <bool kFp64, bool kFinal, bool kForceOut>
__global__ void __launch_bounds__(kBlock)
k_step(StepArgs a, int32_t n) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double d2 = 0.0, mv2 = 0.0;  // displacement since the build, move in this step
    bool stale = false;
    if (i < n) {
        const double xi = a.st.r[0][i], yi = a.st.r[1][i], zi = a.st.r[2][i];
        double fx = 0, fy = 0, fz = 0;
        const int32_t cnt = a.count[i];
        const uint32_t* e = a.list + i;
        for (int32_t k = 0; k < cnt; k++) {
            const uint32_t ent = e[(int64_t)k * n];
            const uint32_t j = ent & kSlotMask, code = ent >> kImageShift;
            double ddx = xi - a.st.r[0][j], ddy = yi - a.st.r[1][j], ddz = zi - a.st.r[2][j];
            if (code != kNoShift) {
                ddx -= image(code, 0, a.L);
                ddy -= image(code, 1, a.L);
                ddz -= image(code, 2, a.L);
            }
            const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
            if (r2 >= a.rc2) continue;
            if constexpr (kFp64) {
                const double q = r2 / a.sig2;
                const double iq4 = 1.0 / (q * q * q * q);
                const double iq7 = iq4 / (q * q * q);
                const double fr = (2 * iq7 - iq4) * a.coeff;
                fx += ddx * fr; fy += ddy * fr; fz += ddz * fr;
            } else {
                // s2 = (sig/r)^2; F/r = coeff (sig/r)^8 (2 (sig/r)^6 - 1)
                const float s2 = __builtin_amdgcn_rcpf((float)r2 * a.inv_sig2f);
                const float s4 = s2 * s2;
                const float s6 = s4 * s2;
                const float s8 = s4 * s4;
                const double fr = (double)(s8 * __builtin_fmaf(2.0f, s6, -1.0f) * a.coeff_f);
                fx = __builtin_fma(fr, ddx, fx);
                fy = __builtin_fma(fr, ddy, fy);
                fz = __builtin_fma(fr, ddz, fz);
            }
        }
        const double f[3] = {fx, fy, fz};
        const double rt[3] = {xi, yi, zi};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double r = rt[d], v = a.st.v[d][i], ac = a.st.a[d][i], b = a.st.b[d][i], c = a.st.c[d][i];
            // corrector (cpu/common/Device.cpp:58-84)
            const double anew = f[d] / a.mass * a.scale;
            const double corr = anew - ac;
            r = r + a.gc.kc[0] * corr;
            v = v + a.gc.kc[1] * corr;
            ac = anew;
            b = b + a.gc.kc[2] * corr;
            c = c + a.gc.kc[3] * corr;
            if constexpr (!kFinal) {
                // predictor of the next step (cpu/common/Device.cpp:24-38)
                const double rp = r + a.gc.k[0] * v + a.gc.k[1] * ac + a.gc.k[2] * b + a.gc.k[3] * c;
                v = v + a.gc.k[0] * ac + a.gc.k[1] * b + a.gc.k[2] * c;
                ac = ac + a.gc.k[0] * b + a.gc.k[1] * c;
                b = b + a.gc.k[0] * c;
                const double dd = rp - a.r0[d][i], ds = rp - rt[d];
                d2 += dd * dd;
                mv2 += ds * ds;
                r = rp;
            }
            a.rn[d][i] = r;
            a.st.v[d][i] = v;
            a.st.a[d][i] = ac;
            a.st.b[d][i] = b;
            a.st.c[d][i] = c;
            if constexpr (kForceOut) a.st.f[d][i] = f[d];
        }
        // the force just used was valid only if r(t) was within the list margin
        const double e0 = xi - a.r0[0][i], e1 = yi - a.r0[1][i], e2 = zi - a.r0[2][i];
        stale = e0 * e0 + e1 * e1 + e2 * e2 >= a.hard2;
    }
    if constexpr (!kFinal) {
        unsigned long long bd = __double_as_longlong(d2), bs = __double_as_longlong(mv2);
        for (int o = kWave / 2; o > 0; o >>= 1) {
            unsigned long long x = __shfl_xor(bd, o, kWave), y = __shfl_xor(bs, o, kWave);
            bd = x > bd ? x : bd;
            bs = y > bs ? y : bs;
        }
        if ((threadIdx.x & (kWave - 1)) == 0) {
            atomicMax(&a.flags[0], bd);
            atomicMax(&a.flags[1], bs);
        }
    }
    if (__any(stale) && (threadIdx.x & (kWave - 1)) == 0) atomicAdd(&a.flags[2], 1ull);
}

}  // namespace

hipError_t launch_nlist(hipStream_t s, const DevState& st, const int32_t* cell_of, const GridView& g,
                        double rl2, int32_t n, int32_t cap, uint32_t* list, int32_t* count,
                        int32_t* overflow) {
    if (n <= 0) return hipSuccess;
    k_nlist<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(st.r[0], st.r[1], st.r[2], cell_of, g, rl2, n, cap,
                                                        list, count, overflow);
    return hipGetLastError();
}

hipError_t launch_step(hipStream_t s, const DevState& st, double* const rn[3], double* const r0[3],
                       const uint32_t* list, const int32_t* count, const LjTable& t, const MassTable& m,
                       const GearCoeffs& gc, double L, double hard2, int32_t n, int32_t precision,
                       bool final_step, bool force_out, uint64_t* flags) {
    if (n <= 0) return hipSuccess;
    StepArgs a;
    a.st = st;
    for (int d = 0; d < 3; d++) { a.rn[d] = rn[d]; a.r0[d] = r0[d]; }
    a.list = list;
    a.count = count;
    a.L = L;
    a.rc2 = t.rc2;
    a.inv_sig2f = t.inv_sig2f[0];
    a.coeff_f = t.coeff_f[0];
    a.sig2 = t.sig2[0];
    a.coeff = t.coeff[0];
    a.mass = m.mass[0];
    a.scale = m.scale;
    a.gc = gc;
    a.hard2 = hard2;
    a.flags = reinterpret_cast<unsigned long long*>(flags);
    dim3 grid((n + kBlock - 1) / kBlock);
    const bool fp64 = precision == MDX_PREC_FP64;
#define MDX_STEP(F64, FIN, FO) k_step<F64, FIN, FO><<<grid, kBlock, 0, s>>>(a, n)
    if (fp64) {
        if (final_step) { if (force_out) MDX_STEP(true, true, true); else MDX_STEP(true, true, false); }
        else { if (force_out) MDX_STEP(true, false, true); else MDX_STEP(true, false, false); }
    } else {
        if (final_step) { if (force_out) MDX_STEP(false, true, true); else MDX_STEP(false, true, false); }
        else { if (force_out) MDX_STEP(false, false, true); else MDX_STEP(false, false, false); }
    }
#undef MDX_STEP
    return hipGetLastError();
}

}  // namespace mdx
