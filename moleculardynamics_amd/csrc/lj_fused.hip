// lj_fused.hip -- periodic-box performance path: LDS-staged neighbourhoods and
// the whole MD step in one kernel.
//
// Work unit: a "segment" = K consecutive cells of one z-column (x, y fixed).
// Its atoms are one contiguous slot range (slots are cell-sorted, z fastest).
// The segment's neighbourhood -- the 9 columns (x+dx, y+dy), dx, dy in
// {-1,0,1}, each over z0-1 .. z1 -- is staged into LDS as FP64 positions
// already moved to the periodic image seen from the segment, in the order
// column (dx,dy lexicographic), then z, then slot; so every atom's 27-cell
// candidate set is 9 contiguous LDS ranges, visited in the reference's
// candidate order (cells in (dx,dy,dz) lexicographic order, ascending index
// within a cell: common/NeighborList.cpp:99-136).
//
//  k_nlist_lds (at each rebuild): each atom keeps the LDS indices of its
//    candidates within the list radius rs (16-bit entries, column-major
//    [k][slot] so a wave's loads are coalesced).
//  k_step_lds (every step): LJ force over the list (FP64 displacement, FP32
//    12-6 force law in the mixed mode, FP64 accumulation), then the Gear
//    corrector of step t and the predictor of step t+1, in registers.
//    Positions are double-buffered: r(t) is read (and staged) from A and
//    r(t+1) written to B, so no workgroup reads a position another one has
//    already advanced.
// Workgroup-to-segment mapping is XCD-aware: workgroups are dealt to the 8
// XCDs round robin, so block b takes segment (b % 8) * per + b / 8 and each
// XCD's L2 serves one contiguous slab of the box.
#include "mdx_internal.h"

namespace mdx {

namespace {

constexpr int kBlock = 256;
constexpr int kXcd = 8;
constexpr int kLdsAtoms = 3072;   // staged atoms per segment at most (3 x 24 KiB of FP64)
constexpr int kBatch = 8;         // list entries loaded ahead per thread
constexpr int kMaxK = 14;         // cells per segment
constexpr int kSpread = 64;       // flag atomics are spread over this many addresses

struct SegGrid {
    int32_t n[3];        // cells per axis
    int32_t nseg;        // segments per column
    int32_t total;       // n[0] * n[1] * nseg
    int32_t per_xcd;     // ceil(total / 8)
    double L;
    const int32_t* start;
};

// z range [z0, z1) of segment s of a column with nz cells
__device__ inline void seg_z(int32_t s, int32_t nseg, int32_t nz, int32_t& z0, int32_t& z1) {
    z0 = (int32_t)(((int64_t)s * nz) / nseg);
    z1 = (int32_t)(((int64_t)(s + 1) * nz) / nseg);
}

__device__ inline int32_t logical_segment(const SegGrid& g) {
    const int32_t b = blockIdx.x;
    return (b % kXcd) * g.per_xcd + b / kXcd;
}

// LDS header shared by both kernels; the staged FP64 positions x, y, z of
// `cap` atoms each live in a separate LDS array (Pos), addressed through
// pointers derived from the __shared__ array itself so accesses stay ds_*.
struct Pos {
    double* x;
    double* y;
    double* z;
    int32_t cap;
};
struct Stage {
    int32_t colbase[9];                 // LDS start of each column
    int32_t cellpre[9][kMaxK + 3];      // per column, prefix counts over the K+2 cells
    int32_t total;
    double red[kBlock / kWave][3];
};

// Stage the neighbourhood of segment (cx, cy, z0..z1).  Returns false (for
// the whole block) when it does not fit in LDS.
__device__ inline bool stage(Stage& S, const Pos& P, const SegGrid& g, int32_t cx, int32_t cy, int32_t z0,
                             int32_t z1, const double* __restrict__ rx, const double* __restrict__ ry,
                             const double* __restrict__ rz) {
    const int32_t K2 = z1 - z0 + 2;    // cells per column in the window
    const int32_t tid = threadIdx.x;
    if (tid < 9 * K2) {                // per-cell counts (9 columns x K2 cells)
        const int32_t col = tid / K2, o = tid % K2;
        int32_t x = cx + col / 3 - 1, y = cy + col % 3 - 1, z = z0 - 1 + o;
        x = x < 0 ? x + g.n[0] : (x >= g.n[0] ? x - g.n[0] : x);
        y = y < 0 ? y + g.n[1] : (y >= g.n[1] ? y - g.n[1] : y);
        z = z < 0 ? z + g.n[2] : (z >= g.n[2] ? z - g.n[2] : z);
        const int32_t c = (x * g.n[1] + y) * g.n[2] + z;
        S.cellpre[col][o + 1] = g.start[c + 1] - g.start[c];
    }
    __syncthreads();
    if (tid < 9) {
        S.cellpre[tid][0] = 0;
        for (int32_t o = 1; o <= K2; o++) S.cellpre[tid][o] += S.cellpre[tid][o - 1];
    }
    __syncthreads();
    if (tid == 0) {
        int32_t acc = 0;
        for (int32_t col = 0; col < 9; col++) {
            S.colbase[col] = acc;
            acc += S.cellpre[col][K2];
        }
        S.total = acc;
    }
    __syncthreads();
    if (S.total > P.cap) return false;
    // Copy column by column.  A column's window z0-1 .. z1 is one contiguous
    // slot range unless it wraps in z; then its first (or last) cell comes
    // from the other end of the column with a shift of -L (or +L).
    for (int32_t col = 0; col < 9; col++) {
        int32_t x = cx + col / 3 - 1, y = cy + col % 3 - 1;
        double sx = 0, sy = 0;
        if (x < 0) { x += g.n[0]; sx = -g.L; } else if (x >= g.n[0]) { x -= g.n[0]; sx = g.L; }
        if (y < 0) { y += g.n[1]; sy = -g.L; } else if (y >= g.n[1]) { y -= g.n[1]; sy = g.L; }
        const int32_t cb = (x * g.n[1] + y) * g.n[2];
        const int32_t za = z0 - 1, zb = z1;   // window cells, inclusive
        // pieces: [za] wrapped, [max(za,0) .. min(zb,nz-1)] contiguous, [zb] wrapped
        int32_t dst = S.colbase[col];
        if (za < 0) {
            const int32_t c = cb + g.n[2] - 1, b = g.start[c], e = g.start[c + 1];
            for (int32_t q = tid; q < e - b; q += kBlock) {
                P.x[dst + q] = rx[b + q] + sx;
                P.y[dst + q] = ry[b + q] + sy;
                P.z[dst + q] = rz[b + q] - g.L;
            }
            dst += e - b;
        }
        {
            const int32_t lo = za < 0 ? 0 : za, hi = zb >= g.n[2] ? g.n[2] - 1 : zb;
            const int32_t b = g.start[cb + lo], e = g.start[cb + hi + 1];
            for (int32_t q = tid; q < e - b; q += kBlock) {
                P.x[dst + q] = rx[b + q] + sx;
                P.y[dst + q] = ry[b + q] + sy;
                P.z[dst + q] = rz[b + q];
            }
            dst += e - b;
        }
        if (zb >= g.n[2]) {
            const int32_t b = g.start[cb], e = g.start[cb + 1];
            for (int32_t q = tid; q < e - b; q += kBlock) {
                P.x[dst + q] = rx[b + q] + sx;
                P.y[dst + q] = ry[b + q] + sy;
                P.z[dst + q] = rz[b + q] + g.L;
            }
        }
    }
    __syncthreads();
    return true;
}

__global__ void __launch_bounds__(kBlock)
k_nlist_lds(SegGrid g, const double* __restrict__ rx, const double* __restrict__ ry,
            const double* __restrict__ rz, double rl2, int32_t n, int32_t cap, uint16_t* __restrict__ list,
            int32_t* __restrict__ count, int32_t* __restrict__ err) {
    __shared__ Stage S;
    __shared__ double buf[3 * kLdsAtoms];
    const int32_t seg = logical_segment(g);
    if (seg >= g.total) return;
    const Pos P{buf, buf + kLdsAtoms, buf + 2 * kLdsAtoms, kLdsAtoms};
    const int32_t col = seg / g.nseg, szi = seg % g.nseg;
    const int32_t cx = col / g.n[1], cy = col % g.n[1];
    int32_t z0, z1;
    seg_z(szi, g.nseg, g.n[2], z0, z1);
    if (!stage(S, P, g, cx, cy, z0, z1, rx, ry, rz)) {
        if (threadIdx.x == 0) atomicMax(&err[0], S.total);
        return;
    }
    if (threadIdx.x == 0) atomicMax(&err[2], S.total);   // largest staging, sizes the step kernel's LDS
    const int32_t cbase = (cx * g.n[1] + cy) * g.n[2];
    const int32_t hbeg = g.start[cbase + z0], hend = g.start[cbase + z1];
    for (int32_t s = hbeg + threadIdx.x; s < hend; s += kBlock) {
        // window cell o (1..K) of the atom and its own LDS index
        const int32_t qc = s - hbeg;
        int32_t o = 1;
        while (S.cellpre[4][o + 1] - S.cellpre[4][1] <= qc) o++;
        const int32_t self = S.colbase[4] + S.cellpre[4][1] + qc;
        const double xi = P.x[self], yi = P.y[self], zi = P.z[self];
        int32_t k = 0;
        for (int32_t c = 0; c < 9; c++) {
            const int32_t b = S.colbase[c] + S.cellpre[c][o - 1], e = S.colbase[c] + S.cellpre[c][o + 2];
            for (int32_t p0 = b; p0 < e; p0 += kBatch) {
                bool hit[kBatch];
#pragma unroll
                for (int u = 0; u < kBatch; u++) {
                    const int32_t p = p0 + u < e ? p0 + u : b;
                    const double dx = xi - P.x[p], dy = yi - P.y[p], dz = zi - P.z[p];
                    hit[u] = (p0 + u < e) && (p0 + u != self) && (dx * dx + dy * dy + dz * dz < rl2);
                }
#pragma unroll
                for (int u = 0; u < kBatch; u++) {
                    if (hit[u]) {
                        if (k < cap) list[(int64_t)k * n + s] = (uint16_t)(p0 + u);
                        k++;
                    }
                }
            }
        }
        count[s] = k < cap ? k : cap;
        if (k > cap) atomicMax(&err[1], k);
    }
}

struct StepArgs {
    SegGrid g;
    DevState st;            // st.r = positions r(t) (read), Gear state v,a,b,c (read/write)
    double* rn[3];          // r(t+1) out
    const double* r0[3];    // positions at the last list build
    const uint16_t* list;
    const int32_t* count;
    double rc2;
    float inv_sig2f, coeff_f;
    double sig2, coeff;
    double mass, scale;
    GearCoeffs gc;
    double hard2;                // list validity: displacement^2 must stay below this
    unsigned long long* flags;   // kSpread x {max disp^2 bits, max step^2 bits, stale count}
    int32_t n;
    int32_t lds_cap;             // staged atoms the dynamic LDS holds
};

template <bool kFp64, bool kFinal, bool kForceOut>
__global__ void __launch_bounds__(kBlock)
k_step_lds(StepArgs a) {
    __shared__ Stage S;
    extern __shared__ double dyn[];
    const SegGrid& g = a.g;
    const int32_t seg = logical_segment(g);
    if (seg >= g.total) return;
    const Pos P{dyn, dyn + a.lds_cap, dyn + 2 * a.lds_cap, a.lds_cap};
    const int32_t col = seg / g.nseg, szi = seg % g.nseg;
    const int32_t cx = col / g.n[1], cy = col % g.n[1];
    int32_t z0, z1;
    seg_z(szi, g.nseg, g.n[2], z0, z1);
    unsigned long long* fl = a.flags + 3 * (blockIdx.x % kSpread);
    if (!stage(S, P, g, cx, cy, z0, z1, a.st.r[0], a.st.r[1], a.st.r[2])) {
        // cannot happen after a successful list build on the same cells; report it as stale
        if (threadIdx.x == 0) atomicAdd(&fl[2], 1ull << 32);
        return;
    }
    const int32_t cbase = (cx * g.n[1] + cy) * g.n[2];
    const int32_t hbeg = g.start[cbase + z0], hend = g.start[cbase + z1];
    double d2max = 0.0, mv2max = 0.0, stale = 0.0;
    const int32_t n = a.n;
    for (int32_t i = hbeg + threadIdx.x; i < hend; i += kBlock) {
        const double xi = a.st.r[0][i], yi = a.st.r[1][i], zi = a.st.r[2][i];
        double fx = 0, fy = 0, fz = 0;
        const int32_t cnt = a.count[i];
        const uint16_t* e = a.list + i;
        // list entries are loaded kBatch ahead so several loads are in flight;
        // the list has 2 kBatch spare rows, so the look-ahead stays in bounds
        int32_t jn[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++) jn[u] = e[(int64_t)u * n];
        for (int32_t k0 = 0; k0 < cnt; k0 += kBatch) {
            int32_t jc[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; u++) {
                jc[u] = jn[u];
                jn[u] = e[(int64_t)(k0 + kBatch + u) * n];
            }
            // straight-line body: every entry of the batch is evaluated and
            // masked (select, so a 1/0 of a padding entry never reaches the sum),
            // letting the 8 LDS loads and FP64 chains overlap
#pragma unroll
            for (int u = 0; u < kBatch; u++) {
                const bool valid = k0 + u < cnt;
                const int32_t j = valid ? jc[u] : 0;
                const double ddx = xi - P.x[j], ddy = yi - P.y[j], ddz = zi - P.z[j];
                if constexpr (kFp64) {
                    const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                    const double q = r2 / a.sig2;
                    const double iq4 = 1.0 / (q * q * q * q);
                    const double iq7 = iq4 / (q * q * q);
                    const double fr = (valid && r2 < a.rc2) ? (2 * iq7 - iq4) * a.coeff : 0.0;
                    fx += ddx * fr; fy += ddy * fr; fz += ddz * fr;
                } else {
                    const double r2 = __builtin_fma(ddx, ddx, __builtin_fma(ddy, ddy, ddz * ddz));
                    // w = (sig/r)^2; F/r = coeff w^4 (2 w^3 - 1)
                    const float w = __builtin_amdgcn_rcpf((float)r2 * a.inv_sig2f);
                    const float w2 = w * w;
                    const float w3 = w2 * w;
                    const float w4 = w2 * w2;
                    const float frf = w4 * __builtin_fmaf(2.0f, w3, -1.0f) * a.coeff_f;
                    const double fr = (valid && r2 < a.rc2) ? (double)frf : 0.0;
                    fx = __builtin_fma(fr, ddx, fx);
                    fy = __builtin_fma(fr, ddy, fy);
                    fz = __builtin_fma(fr, ddz, fz);
                }
            }
        }
        const double f[3] = {fx, fy, fz};
        const double rt[3] = {xi, yi, zi};
        double d2 = 0.0, mv2 = 0.0;
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
        // the force just used is valid only if r(t) was within the list margin
        const double e0 = xi - a.r0[0][i], e1 = yi - a.r0[1][i], e2 = zi - a.r0[2][i];
        if (e0 * e0 + e1 * e1 + e2 * e2 >= a.hard2) stale += 1.0;
        d2max = fmax(d2max, d2);
        mv2max = fmax(mv2max, mv2);
    }
    // block reduction (wave shuffles, then LDS), one spread atomic per block
    for (int o = kWave / 2; o > 0; o >>= 1) {
        d2max = fmax(d2max, __shfl_xor(d2max, o, kWave));
        mv2max = fmax(mv2max, __shfl_xor(mv2max, o, kWave));
        stale += __shfl_xor(stale, o, kWave);
    }
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        S.red[w][0] = d2max;
        S.red[w][1] = mv2max;
        S.red[w][2] = stale;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kBlock / kWave; q++) {
            d2max = fmax(d2max, S.red[q][0]);
            mv2max = fmax(mv2max, S.red[q][1]);
            stale += S.red[q][2];
        }
        if (!kFinal) {
            atomicMax(&fl[0], (unsigned long long)__double_as_longlong(d2max));
            atomicMax(&fl[1], (unsigned long long)__double_as_longlong(mv2max));
        }
        if (stale > 0) atomicAdd(&fl[2], (unsigned long long)stale);
    }
}

SegGrid make_grid(const GridView& v, int32_t nseg) {
    SegGrid g;
    for (int d = 0; d < 3; d++) g.n[d] = v.g[d];
    g.nseg = nseg;
    g.total = v.g[0] * v.g[1] * nseg;
    g.per_xcd = (g.total + kXcd - 1) / kXcd;
    g.L = v.L;
    g.start = v.start;
    return g;
}

}  // namespace

int32_t lds_segments_per_column(int32_t nz, double atoms_per_cell, int32_t kmax) {
    // about 230 home atoms per 256-thread workgroup, at most kmax cells per segment
    kmax = kmax < kMaxK ? kmax : kMaxK;
    kmax = kmax < 1 ? 1 : kmax;
    const double apc = atoms_per_cell > 0 ? atoms_per_cell : 1.0;
    int32_t nseg = (int32_t)(nz * apc / 230.0 + 0.5);
    const int32_t lo = (nz + kmax - 1) / kmax;   // segments no longer than kmax cells
    nseg = nseg < lo ? lo : nseg;
    return nseg > nz ? nz : nseg;
}

int32_t lds_capacity() { return kLdsAtoms; }
int32_t list_spare_rows() { return 2 * kBatch; }
int32_t flag_spread() { return kSpread; }

hipError_t launch_nlist_lds(hipStream_t s, const DevState& st, const GridView& v, int32_t nseg, double rl2,
                            int32_t n, int32_t cap, uint16_t* list, int32_t* count, int32_t* err) {
    if (n <= 0) return hipSuccess;
    SegGrid g = make_grid(v, nseg);
    k_nlist_lds<<<g.per_xcd * kXcd, kBlock, 0, s>>>(g, st.r[0], st.r[1], st.r[2], rl2, n, cap, list, count, err);
    return hipGetLastError();
}

hipError_t launch_step_lds(hipStream_t s, const DevState& st, const GridView& v, int32_t nseg,
                           double* const rn[3], double* const r0[3], const uint16_t* list, const int32_t* count,
                           const LjTable& t, const MassTable& m, const GearCoeffs& gc, double hard2, int32_t n,
                           int32_t precision, bool final_step, bool force_out, uint64_t* flags, int32_t lds_atoms) {
    if (n <= 0) return hipSuccess;
    static bool attr_set = false;
    if (!attr_set) {
        // allow dynamic LDS above the 64 KiB default for the largest staging
        const int bytes = 3 * kLdsAtoms * (int)sizeof(double);
        (void)hipFuncSetAttribute((const void*)k_step_lds<true, true, true>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        (void)hipFuncSetAttribute((const void*)k_step_lds<true, true, false>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        (void)hipFuncSetAttribute((const void*)k_step_lds<true, false, true>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        (void)hipFuncSetAttribute((const void*)k_step_lds<true, false, false>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        (void)hipFuncSetAttribute((const void*)k_step_lds<false, true, true>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        (void)hipFuncSetAttribute((const void*)k_step_lds<false, true, false>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        (void)hipFuncSetAttribute((const void*)k_step_lds<false, false, true>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        (void)hipFuncSetAttribute((const void*)k_step_lds<false, false, false>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        attr_set = true;
    }
    if (lds_atoms < 1 || lds_atoms > kLdsAtoms) return hipErrorInvalidValue;
    StepArgs a;
    a.g = make_grid(v, nseg);
    a.st = st;
    for (int d = 0; d < 3; d++) { a.rn[d] = rn[d]; a.r0[d] = r0[d]; }
    a.list = list;
    a.count = count;
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
    a.n = n;
    a.lds_cap = lds_atoms;
    const size_t dyn = 3 * (size_t)lds_atoms * sizeof(double);
    dim3 grid(a.g.per_xcd * kXcd);
    const bool fp64 = precision == MDX_PREC_FP64;
#define MDX_STEP(F64, FIN, FO) k_step_lds<F64, FIN, FO><<<grid, kBlock, dyn, s>>>(a)
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
