// replicas.hip -- batched independent periodic Lennard-Jones replicas, the
// whole batch advanced by one graph launch per two steps.
//
// The reference runs independent systems one after another: its vxcm sweep
// (cpu/cpumd/Host.cpp:75-216) restarts the same step loop for every velocity
// offset.  Here R replicas of different sizes and boxes share one set of
// device arrays and every kernel covers all of them:
//
//  * layout: replica r owns atoms [off[r], off[r] + n[r]) of the SoA state;
//    off[r] is a multiple of 64, so every wave belongs to exactly one replica
//    (padding atoms are inactive);
//  * lists are rebuilt per replica, gated by a device flag, so one static
//    graph serves all steps: k_rep_sort (one workgroup per replica: cell
//    histogram and scan in LDS, scatter, each cell sorted by atom index --
//    deterministic -- and a cell-ordered FP32 copy of the wrapped positions),
//    then k_rep_list (compact list within rs from the 27 neighbouring cells'
//    contiguous position ranges, or from the whole replica when its box holds
//    fewer than 3 cells per axis; minimum image);
//  * k_rep_step: LJ force over the list (minimum image), Gear corrector of
//    step t and predictor of step t+1; positions are double-buffered; the
//    largest displacement of r(t+1) from the list-build positions is
//    max-reduced per replica;
//  * staged replicas (at least 3 cells per axis and a neighbourhood that fits
//    LDS) take the single box's LDS path instead: at each rebuild their state
//    is permuted into the new cell order with positions wrapped into the box
//    (k_rep_perm_*), their lists are built and their steps run by
//    k_nlist_lds / k_step_lds of lj_fused.hip, one workgroup per z-segment of
//    a cell column; an atom-id map keeps the caller's order.  They occupy the
//    slots below ubeg;
//  * small replicas (boxes narrower than 3 cells, up to kSmallMax atoms) sit
//    from slot sfirst on and keep no list: k_rep_small steps them over all
//    pairs staged in LDS;
//  * the gather path (k_rep_list after the first build, k_rep_step) runs over
//    the slots [ubeg, sfirst) only, what neither of the above holds;
//  * k_rep_decide: a replica rebuilds before the next step exactly when that
//    displacement reached (rs - rc)/2 -- the decision reads the positions the
//    next step will use, so it needs no margin for lag and no host round trip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mdx_internal.h"

using namespace mdx;

namespace {

constexpr int kBlock = 256;
constexpr int kSortBlock = 1024;
constexpr int kMaxNc = 10;                               // cells per axis at most (larger cells stay correct)
constexpr int kMaxCells = kMaxNc * kMaxNc * kMaxNc;
constexpr int kMaxReplica = 16384;                       // atoms per replica (the LDS sort holds them all)
// the paths of a replica (mdx_batch_staged reports them)
constexpr uint8_t kGather = 0, kStaged = 1, kSmall = 2;
constexpr int kSmallMax = 2048;   // atoms of a small replica (its FP64 positions, 48 KiB of LDS)

struct RepView {
    const int32_t* rep_of;     // replica of each atom slot (-1: padding)
    const int32_t* off;        // R + 1 atom offsets
    const int32_t* nat;        // R atom counts
    const int32_t* coff;       // R + 1 offsets of the per-replica cell-start arrays (ncell + 1 entries each)
    const int32_t* nc;         // R cells per axis (0: one cell, all pairs)
    const double* L;           // R box edges
    const int32_t* rebuild;    // R flags
};

__device__ inline double mi(double d, double L, double invL) { return d - L * rint(d * invL); }

// into [0, L) up to rounding at the ends (cell_index clamps those)
__device__ inline double wrap(double x, double L, double invL) { return x - L * floor(x * invL); }

__device__ inline int32_t cell_index(double y, double cells_per_length, int32_t nc) {   // y wrapped
    int32_t c = (int32_t)(y * cells_per_length);
    return c >= nc ? nc - 1 : (c < 0 ? 0 : c);
}

// One workgroup per replica that needs a new list: counting sort of its atoms
// into cells (histogram, scan and scatter in LDS), each cell then ordered by
// atom index (an atom's rank among its cell's atoms), and a cell-ordered copy
// (x, y, z wrapped into [0, L) as FP32, atom index in w).  tpos is scratch for
// the unsorted copy.  Every atom's cell stays in LDS (16 bits; the LDS holds
// 6 bytes per atom of the batch's largest replica, cap), so after the first
// pass no loop waits on a global read per iteration; the global loads of a
// thread's atoms are issued kBatch at a time.
constexpr int kPer = kMaxReplica / kSortBlock;   // atoms per thread at most
constexpr int kBatch = 8;
__global__ void __launch_bounds__(kSortBlock)
k_rep_sort(RepView v, const double* __restrict__ rx, const double* __restrict__ ry, const double* __restrict__ rz,
           int32_t* __restrict__ cell_of, int32_t* __restrict__ cstart, int32_t* __restrict__ cstart_g,
           int32_t* __restrict__ sidx, float4* __restrict__ tpos, float4* __restrict__ spos,
           const uint8_t* __restrict__ staged, int32_t cap) {
    const int32_t r = blockIdx.x;
    if (!v.rebuild[r]) return;
    __shared__ int32_t cnt[kMaxCells + 1];
    __shared__ int32_t beg[kMaxCells + 1];
    __shared__ int32_t wsum[kSortBlock / kWave];
    extern __shared__ int32_t sort_lds[];
    int32_t* lidx = sort_lds;                                      // atoms in scatter order
    uint16_t* cell = reinterpret_cast<uint16_t*>(sort_lds + cap);  // cell of each atom
    const int32_t tid = threadIdx.x;
    const int32_t n = v.nat[r], base = v.off[r], nc = v.nc[r], cb = v.coff[r];
    const int32_t ncl = nc ? nc * nc * nc : 1;
    const double L = v.L[r], invL = 1.0 / L, cpl = nc * invL;
    for (int32_t c = tid; c < ncl; c += kSortBlock) cnt[c] = 0;
    __syncthreads();
    for (int32_t k0 = 0; k0 < kPer; k0 += kBatch) {
        if (tid + k0 * kSortBlock >= n) break;
        double px[kBatch], py[kBatch], pz[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; k++) {
            const int32_t i = tid + (k0 + k) * kSortBlock;
            if (i < n) { px[k] = rx[base + i]; py[k] = ry[base + i]; pz[k] = rz[base + i]; }
        }
#pragma unroll
        for (int k = 0; k < kBatch; k++) {
            const int32_t i = tid + (k0 + k) * kSortBlock;
            if (i >= n) continue;
            const double x = wrap(px[k], L, invL), y = wrap(py[k], L, invL), z = wrap(pz[k], L, invL);
            const int32_t c =
                nc ? (cell_index(x, cpl, nc) * nc + cell_index(y, cpl, nc)) * nc + cell_index(z, cpl, nc) : 0;
            cell_of[base + i] = c;
            cell[i] = (uint16_t)c;
            tpos[base + i] = make_float4((float)x, (float)y, (float)z, __int_as_float(i));
            atomicAdd(&cnt[c], 1);
        }
    }
    __syncthreads();
    // exclusive scan over the cells: contiguous chunks per thread, the chunk
    // sums scanned within each wave (shuffles) and across the waves (LDS)
    const int32_t chunk = (ncl + kSortBlock - 1) / kSortBlock;
    const int32_t lo = tid * chunk < ncl ? tid * chunk : ncl, hi = lo + chunk < ncl ? lo + chunk : ncl;
    int32_t sum = 0;
    for (int32_t c = lo; c < hi; c++) sum += cnt[c];
    const int32_t lane = tid & (kWave - 1), wv = tid / kWave;
    int32_t inc = sum;
    for (int o = 1; o < kWave; o <<= 1) {
        const int32_t y = __shfl_up(inc, o, kWave);
        if (lane >= o) inc += y;
    }
    if (lane == kWave - 1) wsum[wv] = inc;
    __syncthreads();
    int32_t run = inc - sum;
    for (int32_t w = 0; w < wv; w++) run += wsum[w];
    for (int32_t c = lo; c < hi; c++) {
        const int32_t x = cnt[c];
        beg[c] = run;
        cnt[c] = run;   // becomes the fill cursor
        cstart[cb + c] = run;
        cstart_g[cb + c] = base + run;
        run += x;
    }
    if (tid == 0) { beg[ncl] = n; cstart[cb + ncl] = n; cstart_g[cb + ncl] = base + n; }
    __syncthreads();
    if (nc == 0) {
        // one cell: ascending index order is the identity
        for (int32_t i = tid; i < n; i += kSortBlock) sidx[base + i] = i;
    } else {
        for (int32_t i = tid; i < n; i += kSortBlock) lidx[atomicAdd(&cnt[cell[i]], 1)] = i;
        __syncthreads();
        // an atom's slot: its cell's start plus the number of the cell's atoms
        // with a smaller index (independent LDS reads, ~20-50 per atom)
        for (int32_t p = tid; p < n; p += kSortBlock) {
            const int32_t i = lidx[p], c = cell[i];
            const int32_t b = beg[c], e = beg[c + 1];
            int32_t rank = 0;
            for (int32_t q = b; q < e; q++) rank += lidx[q] < i ? 1 : 0;
            sidx[base + b + rank] = i;
        }
    }
    __syncthreads();   // this workgroup's sidx writes are visible to it
    // a staged replica is permuted into this order next (k_rep_perm_*): its
    // records name the new slots and its cells follow the new order
    const bool perm = nc != 0 && staged[r];
    for (int32_t k0 = 0; k0 < kPer; k0 += kBatch) {
        if (tid + k0 * kSortBlock >= n) break;
        int32_t at[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; k++) {
            const int32_t p = tid + (k0 + k) * kSortBlock;
            at[k] = p < n ? sidx[base + p] : 0;
        }
#pragma unroll
        for (int k = 0; k < kBatch; k++) {
            const int32_t p = tid + (k0 + k) * kSortBlock;
            if (p >= n) continue;
            float4 w = tpos[base + at[k]];
            if (perm) {
                w.w = __int_as_float(p);
                cell_of[base + p] = cell[at[k]];
            }
            spos[base + p] = w;
        }
    }
}

// The segments of the staged replicas that rebuild, compacted in replica
// order into wl[0 .. *count): the list build's work.  One workgroup takes the
// replicas in rounds of kSortBlock (a block scan of their segment counts).
__global__ void __launch_bounds__(kSortBlock)
k_rep_worklist(int32_t R, const int32_t* __restrict__ rebuild, const int32_t* __restrict__ sbeg,
               int32_t* __restrict__ wl, int32_t* __restrict__ count) {
    __shared__ int32_t wsum[kSortBlock / kWave];
    __shared__ int32_t carry;
    const int32_t tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int32_t r0 = 0; r0 < R; r0 += kSortBlock) {
        const int32_t r = r0 + tid;
        const int32_t m = r < R && rebuild[r] ? sbeg[r + 1] - sbeg[r] : 0;
        int32_t inc = m;
        for (int o = 1; o < kWave; o <<= 1) {
            const int32_t y = __shfl_up(inc, o, kWave);
            if (lane >= o) inc += y;
        }
        if (lane == kWave - 1) wsum[wv] = inc;
        __syncthreads();
        int32_t run = carry + inc - m;
        for (int32_t w = 0; w < wv; w++) run += wsum[w];
        for (int32_t k = 0; k < m; k++) wl[run + k] = sbeg[r] + k;
        __syncthreads();   // carry and wsum read by every thread
        if (tid == kSortBlock - 1) carry = run + m;
        __syncthreads();
    }
    if (tid == 0) *count = carry;
}

// The state of a rebuilt staged replica into the new cell order: slot p
// takes old slot sidx[p] (15 components: positions of the current buffer,
// v, a, b, c), first gathered into tmp, then copied back with the positions
// wrapped into the box; the list-build positions r0 and the atom-id map
// follow.
struct RepPerm {
    double* comp[15];   // x y z of r, v, a, b, c
    int32_t* atom;
    double* r0[3];
};

__global__ void __launch_bounds__(kBlock)
k_rep_perm_gather(RepView v, const uint8_t* __restrict__ staged, const int32_t* __restrict__ sidx, RepPerm pm,
                  double* __restrict__ tmp, int32_t* __restrict__ tmp_atom, int32_t ntot) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntot) return;
    const int32_t r = v.rep_of[i];
    if (r < 0 || !v.rebuild[r] || !staged[r] || v.nc[r] == 0) return;
    const int32_t src = v.off[r] + sidx[i];
#pragma unroll
    for (int k = 0; k < 15; k++) tmp[(int64_t)k * ntot + i] = pm.comp[k][src];
    tmp_atom[i] = pm.atom[src];
}

__global__ void __launch_bounds__(kBlock)
k_rep_perm_scatter(RepView v, const uint8_t* __restrict__ staged, RepPerm pm, const double* __restrict__ tmp,
                   const int32_t* __restrict__ tmp_atom, int32_t ntot) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntot) return;
    const int32_t r = v.rep_of[i];
    if (r < 0 || !v.rebuild[r] || !staged[r] || v.nc[r] == 0) return;
    // positions wrapped as k_rep_sort wrapped them to bin them: the staging
    // adds each cell column's image shift to the stored value
    const double L = v.L[r], invL = 1.0 / L;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double x = wrap(tmp[(int64_t)d * ntot + i], L, invL);
        pm.comp[d][i] = x;
        pm.r0[d][i] = x;
    }
#pragma unroll
    for (int k = 3; k < 15; k++) pm.comp[k][i] = tmp[(int64_t)k * ntot + i];
    pm.atom[i] = tmp_atom[i];
}

// Compact list of each atom of a flagged replica in slots [first, last), a
// quad of 4 lanes per atom:
// candidates from the 27 cells around it (or the whole replica), the lanes
// testing interleaved candidates and numbering their hits lane-major, so the
// list order is deterministic.  The radius test runs on the FP32 copy with rs
// widened by 1e-4, so the list holds every pair within rs.
__global__ void __launch_bounds__(kBlock)
k_rep_list(RepView v, const double* __restrict__ rx, const double* __restrict__ ry, const double* __restrict__ rz,
           double* __restrict__ r0x, double* __restrict__ r0y, double* __restrict__ r0z, int32_t ntot,
           const int32_t* __restrict__ cell_of, const int32_t* __restrict__ cstart, const float4* __restrict__ spos,
           float rs2, int32_t cap, int32_t* __restrict__ list, int32_t* __restrict__ count, int32_t* __restrict__ err,
           const int32_t* __restrict__ gate, int32_t first, int32_t last, int32_t write_r0) {
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t i = first + (t >> 2), q = t & 3;
    if (i >= last) return;   // whole quads
    const int32_t r = v.rep_of[i];
    if (r < 0 || !gate[r]) return;
    const int32_t base = v.off[r], nc = v.nc[r], cb = v.coff[r];
    const double L = v.L[r];
    const double x = rx[i], y = ry[i], z = rz[i];
    if (q == 0 && write_r0) {
        r0x[i] = x;
        r0y[i] = y;
        r0z[i] = z;
    }
    const float Lf = (float)L, iLf = 1.0f / Lf;
    const double invL = 1.0 / L;
    const float xi = (float)wrap(x, L, invL), yi = (float)wrap(y, L, invL), zi = (float)wrap(z, L, invL);
    const int32_t li = i - base;
    int32_t k = 0;   // hits so far, the same in the 4 lanes
    auto scan = [&](int32_t c) {
        const int32_t b = cstart[cb + c], e = cstart[cb + c + 1];
        for (int32_t p0 = b; p0 < e; p0 += 16) {
            uint32_t mask = 0;
            int32_t js[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int32_t p = p0 + 4 * u + q;
                const float4 w = spos[base + (p < e ? p : b)];
                js[u] = __float_as_int(w.w);
                float dx = xi - w.x, dy = yi - w.y, dz = zi - w.z;
                dx -= Lf * rintf(dx * iLf);
                dy -= Lf * rintf(dy * iLf);
                dz -= Lf * rintf(dz * iLf);
                const bool hit = (p < e) & (js[u] != li) & (dx * dx + dy * dy + dz * dz < rs2);
                mask |= (uint32_t)hit << u;
            }
            int32_t total;
            int32_t kk = k + quad_excl(__popc(mask), q, total);
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if ((mask >> u) & 1u) {
                    if (kk < cap) list[(int64_t)kk * ntot + i] = base + js[u];
                    kk++;
                }
            }
            k += total;
        }
    };
    if (nc == 0) {
        scan(0);
    } else {
        const int32_t c = cell_of[i];
        const int32_t cz = c % nc, cy = (c / nc) % nc, cx = c / (nc * nc);
        for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dz = -1; dz <= 1; dz++)
                    scan((((cx + dx + nc) % nc) * nc + (cy + dy + nc) % nc) * nc + (cz + dz + nc) % nc);
    }
    if (q == 0) {
        count[i] = k < cap ? k : cap;
        if (k > cap) atomicMax(&err[1], k);
    }
}

struct RepStep {
    RepView v;
    const double* r[3];   // r(t), read
    double* rn[3];        // r(t+1), written
    double* vel[3];
    double* acc[3];
    double* b3[3];
    double* c3[3];
    double* f[3];
    const double* r0[3];
    const double4* rq;    // r(t) again, one 32-byte record per atom (the neighbour gathers)
    double4* rqn;         // r(t+1) as records
    const int32_t* list;
    const int32_t* count;
    int32_t ntot;
    double rc2, sig2, coeff;
    float inv_sig2f, coeff_f;
    double mass, scale;
    GearCoeffs gc;
    unsigned long long* dmax;   // R: max displacement^2 bits of r(t+1)
    int32_t first, last;        // k_rep_step: its slots (the gather path's replicas)
};

// kMode: 0 = corrector + next predictor, 1 = corrector only (last step),
// 2 = forces and a = F/m only (start of a run); atoms in slots [a.first, a.last).  A quad of 4 lanes per atom:
// lane q sums list entries q, q + 4, ...; the quad's sums are added by DPP and
// lane q < 3 integrates component q (the state arrays hold the 3 components of
// a vector in one block of 3 x ntot).
template <bool kFp64, int kMode>
__global__ void __launch_bounds__(kBlock)
k_rep_step(RepStep a) {
    constexpr bool kFinal = kMode != 0;
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t i = a.first + (t >> 2), q = t & 3;
    const int32_t r = i < a.last ? a.v.rep_of[i] : -1;
    double d2 = 0.0;
    if (r >= 0) {
        const double L = a.v.L[r], invL = 1.0 / L;
        const double4 pi = a.rq[i];
        const double xi = pi.x, yi = pi.y, zi = pi.z;
        double fx = 0, fy = 0, fz = 0;
        const int32_t cnt = a.count[i];
        const int32_t* lp = a.list + i;
        int32_t jn = q < cnt ? lp[(int64_t)q * a.ntot] : 0;
        for (int32_t k = q; k < cnt; k += 4) {
            const int32_t j = jn;
            // the next entry is loaded one iteration ahead; the empty asm keeps
            // the compiler from moving the load to the top of the next iteration
            uint64_t pa = reinterpret_cast<uint64_t>(lp + (int64_t)(k + 4 < cnt ? k + 4 : k) * a.ntot);
            asm volatile("" : "+v"(pa));
            jn = load_global_i32(pa);
            const double4 pj = a.rq[j];
            const double dx = mi(xi - pj.x, L, invL), dy = mi(yi - pj.y, L, invL), dz = mi(zi - pj.z, L, invL);
            if constexpr (kFp64) {
                const double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 < a.rc2) {
                    const double qq = r2 / a.sig2;
                    const double iq4 = 1.0 / (qq * qq * qq * qq);
                    const double iq7 = iq4 / (qq * qq * qq);
                    const double fr = (2 * iq7 - iq4) * a.coeff;
                    fx += dx * fr; fy += dy * fr; fz += dz * fr;
                }
            } else {
                const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
                const float w = __builtin_amdgcn_rcpf((float)r2 * a.inv_sig2f);
                const float w2 = w * w, w3 = w2 * w;
                const float frf = r2 < a.rc2 ? w2 * w2 * __builtin_fmaf(2.0f, w3, -1.0f) * a.coeff_f : 0.0f;
                const double fr = (double)frf;
                fx = __builtin_fma(fr, dx, fx);
                fy = __builtin_fma(fr, dy, fy);
                fz = __builtin_fma(fr, dz, fz);
            }
        }
        fx = quad_sum(fx);
        fy = quad_sum(fy);
        fz = quad_sum(fz);
        if (q < 3) {
            const double F = q == 0 ? fx : (q == 1 ? fy : fz);
            const int64_t o = (int64_t)q * a.ntot + i;   // component q of atom i
            if constexpr (kMode == 2) {
                a.acc[0][o] = F / a.mass * a.scale;
                a.f[0][o] = F;
            } else {
                double rr = a.r[0][o], v = a.vel[0][o], ac = a.acc[0][o], b = a.b3[0][o], c = a.c3[0][o];
                // corrector (cpu/common/Device.cpp:58-84)
                const double anew = F / a.mass * a.scale;
                const double corr = anew - ac;
                rr = rr + a.gc.kc[0] * corr;
                v = v + a.gc.kc[1] * corr;
                ac = anew;
                b = b + a.gc.kc[2] * corr;
                c = c + a.gc.kc[3] * corr;
                if constexpr (!kFinal) {
                    // predictor of the next step (cpu/common/Device.cpp:24-38)
                    const double rp = rr + a.gc.k[0] * v + a.gc.k[1] * ac + a.gc.k[2] * b + a.gc.k[3] * c;
                    v = v + a.gc.k[0] * ac + a.gc.k[1] * b + a.gc.k[2] * c;
                    ac = ac + a.gc.k[0] * b + a.gc.k[1] * c;
                    b = b + a.gc.k[0] * c;
                    const double dd = rp - a.r0[0][o];
                    d2 = dd * dd;
                    rr = rp;
                }
                a.rn[0][o] = rr;
                reinterpret_cast<double*>(a.rqn + i)[q] = rr;
                a.vel[0][o] = v;
                a.acc[0][o] = ac;
                a.b3[0][o] = b;
                a.c3[0][o] = c;
                a.f[0][o] = F;
            }
        }
        if constexpr (!kFinal) d2 = quad_sum(d2);
    }
    if constexpr (!kFinal) {
        // a wave (16 atoms) never spans two replicas (offsets are multiples of 64)
        for (int o = kWave / 2; o > 0; o >>= 1) d2 = fmax(d2, __shfl_xor(d2, o, kWave));
        const int32_t r0 = __shfl(r, 0, kWave);
        if ((threadIdx.x & (kWave - 1)) == 0 && r0 >= 0)
            atomicMax(&a.dmax[r0], (unsigned long long)__double_as_longlong(d2));
    }
}

// Replicas narrower than 3 cells of rs per axis (all pairs), one workgroup per
// kSmallAtoms of a replica's atoms (block table: replica, first atom): the
// replica's positions r(t) staged in LDS, 8 lanes per atom testing every
// other atom by minimum image (lane q takes atoms q, q + 8, ...; the groups of
// a wave read the same 8 atoms, which the LDS broadcasts; the two quads' sums
// are added by one shuffle, the same in all 8 lanes), then the Gear step of
// k_rep_step.  The pairs within rc are those of the list path;
// only the order of the sum differs.  No list is kept: a replica that rebuilds
// records its list-build positions here (its rebuild decision, count and sort
// still run, for the energies' lists and the build counts).
constexpr int kSmallLanes = 8;
constexpr int kSmallAtoms = kBlock / kSmallLanes;
template <bool kFp64, bool kFinal>
__global__ void __launch_bounds__(kBlock)
k_rep_small(RepStep a, const int2* __restrict__ blocks) {
    extern __shared__ double stg[];   // x y z of each atom of the replica
    const int2 bk = blocks[blockIdx.x];
    const int32_t r = bk.x;
    const int32_t base = a.v.off[r], n = a.v.nat[r];
    const double L = a.v.L[r], invL = 1.0 / L;
    // staging: a thread's atoms' loads (4 atoms, 3 components) go out together
    for (int32_t j0 = 0; j0 < n; j0 += 4 * kBlock) {
        double bx[4], by[4], bz[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int32_t j = j0 + u * kBlock + (int32_t)threadIdx.x;
            if (j < n) { bx[u] = a.r[0][base + j]; by[u] = a.r[1][base + j]; bz[u] = a.r[2][base + j]; }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int32_t j = j0 + u * kBlock + (int32_t)threadIdx.x;
            if (j < n) { stg[3 * j] = bx[u]; stg[3 * j + 1] = by[u]; stg[3 * j + 2] = bz[u]; }
        }
    }
    __syncthreads();
    const int32_t li = bk.y + (int32_t)threadIdx.x / kSmallLanes, q = threadIdx.x & (kSmallLanes - 1);
    const bool on = li < n;   // whole groups
    double fx = 0, fy = 0, fz = 0;
    if (on) {
        const double xi = stg[3 * li], yi = stg[3 * li + 1], zi = stg[3 * li + 2];
        // the atom itself fails the cutoff test by a select (no branch, so the
        // loop unrolls and its LDS reads overlap)
#pragma unroll 4
        for (int32_t j = q; j < n; j += kSmallLanes) {
            const double dx = mi(xi - stg[3 * j], L, invL), dy = mi(yi - stg[3 * j + 1], L, invL),
                         dz = mi(zi - stg[3 * j + 2], L, invL);
            const double rc2 = j == li ? -1.0 : a.rc2;
            if constexpr (kFp64) {
                const double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 < rc2) {
                    const double qq = r2 / a.sig2;
                    const double iq4 = 1.0 / (qq * qq * qq * qq);
                    const double iq7 = iq4 / (qq * qq * qq);
                    const double fr = (2 * iq7 - iq4) * a.coeff;
                    fx += dx * fr; fy += dy * fr; fz += dz * fr;
                }
            } else {
                const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
                const float w = __builtin_amdgcn_rcpf((float)r2 * a.inv_sig2f);
                const float w2 = w * w, w3 = w2 * w;
                const float frf = r2 < rc2 ? w2 * w2 * __builtin_fmaf(2.0f, w3, -1.0f) * a.coeff_f : 0.0f;
                const double fr = (double)frf;
                fx = __builtin_fma(fr, dx, fx);
                fy = __builtin_fma(fr, dy, fy);
                fz = __builtin_fma(fr, dz, fz);
            }
        }
    }
    fx = quad_sum(fx);
    fy = quad_sum(fy);
    fz = quad_sum(fz);
    fx += __shfl_xor(fx, 4, kWave);
    fy += __shfl_xor(fy, 4, kWave);
    fz += __shfl_xor(fz, 4, kWave);
    double d2 = 0.0;
    if (on && q < 3) {
        const double F = q == 0 ? fx : (q == 1 ? fy : fz);
        const int64_t o = (int64_t)q * a.ntot + base + li;   // component q of the atom
        const double rt = stg[3 * li + q];
        double r0 = rt;
        if (a.v.rebuild[r]) const_cast<double*>(a.r0[0])[o] = rt;   // this step's positions are the build's
        else r0 = a.r0[0][o];
        double rr = rt, v = a.vel[0][o], ac = a.acc[0][o], b = a.b3[0][o], c = a.c3[0][o];
        // corrector (cpu/common/Device.cpp:58-84)
        const double anew = F / a.mass * a.scale;
        const double corr = anew - ac;
        rr = rr + a.gc.kc[0] * corr;
        v = v + a.gc.kc[1] * corr;
        ac = anew;
        b = b + a.gc.kc[2] * corr;
        c = c + a.gc.kc[3] * corr;
        if constexpr (!kFinal) {
            // predictor of the next step (cpu/common/Device.cpp:24-38)
            const double rp = rr + a.gc.k[0] * v + a.gc.k[1] * ac + a.gc.k[2] * b + a.gc.k[3] * c;
            v = v + a.gc.k[0] * ac + a.gc.k[1] * b + a.gc.k[2] * c;
            ac = ac + a.gc.k[0] * b + a.gc.k[1] * c;
            b = b + a.gc.k[0] * c;
            const double dd = rp - r0;
            d2 = dd * dd;
            rr = rp;
        }
        a.rn[0][o] = rr;
        a.vel[0][o] = v;
        a.acc[0][o] = ac;
        a.b3[0][o] = b;
        a.c3[0][o] = c;
        a.f[0][o] = F;
    }
    if constexpr (!kFinal) {
        d2 = quad_sum(d2);
        for (int o = kWave / 2; o > 0; o >>= 1) d2 = fmax(d2, __shfl_xor(d2, o, kWave));
        if ((threadIdx.x & (kWave - 1)) == 0) atomicMax(&a.dmax[r], (unsigned long long)__double_as_longlong(d2));
    }
}

// predictor (cpu/common/Device.cpp:17-39) with the per-replica largest
// displacement of the predicted positions from the list-build positions
__global__ void __launch_bounds__(kBlock)
k_rep_predict(RepStep a) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t r = i < a.ntot ? a.v.rep_of[i] : -1;
    double d2 = 0.0;
    if (r >= 0) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double* rp = const_cast<double*>(a.r[d]);
            const double x = rp[i], v = a.vel[d][i], ac = a.acc[d][i], b = a.b3[d][i], c = a.c3[d][i];
            const double xn = x + a.gc.k[0] * v + a.gc.k[1] * ac + a.gc.k[2] * b + a.gc.k[3] * c;
            rp[i] = xn;
            reinterpret_cast<double*>(const_cast<double4*>(a.rq) + i)[d] = xn;
            a.vel[d][i] = v + a.gc.k[0] * ac + a.gc.k[1] * b + a.gc.k[2] * c;
            a.acc[d][i] = ac + a.gc.k[0] * b + a.gc.k[1] * c;
            a.b3[d][i] = b + a.gc.k[0] * c;
            const double dd = xn - a.r0[d][i];
            d2 += dd * dd;
        }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) d2 = fmax(d2, __shfl_xor(d2, o, kWave));
    const int32_t r0 = __shfl(r, 0, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0 && r0 >= 0) atomicMax(&a.dmax[r0], (unsigned long long)__double_as_longlong(d2));
}

// a replica rebuilds before the next step when r(t+1) moved (rs - rc)/2 from
// its list-build positions; clears the accumulator for the next step.  With
// staged replicas (sbeg non-null) it also writes the next list build's work
// list, as k_rep_worklist does.  One workgroup, the replicas in rounds.
__global__ void __launch_bounds__(kSortBlock)
k_rep_decide(int32_t R, unsigned long long* __restrict__ dmax, int32_t* __restrict__ rebuild,
             long long* __restrict__ builds, double margin2, const int32_t* __restrict__ sbeg,
             int32_t* __restrict__ wl, int32_t* __restrict__ count) {
    __shared__ int32_t wsum[kSortBlock / kWave];
    __shared__ int32_t carry;
    const int32_t tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int32_t r0 = 0; r0 < R; r0 += kSortBlock) {
        const int32_t r = r0 + tid;
        int32_t f = 0;
        if (r < R) {
            const double d2 = __longlong_as_double((long long)dmax[r]);
            f = d2 >= margin2 ? 1 : 0;
            rebuild[r] = f;
            builds[r] += f;
            dmax[r] = 0ull;
        }
        if (!sbeg) continue;   // uniform
        const int32_t m = f ? sbeg[r + 1] - sbeg[r] : 0;
        int32_t inc = m;
        for (int o = 1; o < kWave; o <<= 1) {
            const int32_t y = __shfl_up(inc, o, kWave);
            if (lane >= o) inc += y;
        }
        if (lane == kWave - 1) wsum[wv] = inc;
        __syncthreads();
        int32_t run = carry + inc - m;
        for (int32_t w = 0; w < wv; w++) run += wsum[w];
        for (int32_t k = 0; k < m; k++) wl[run + k] = sbeg[r] + k;
        __syncthreads();   // carry and wsum read by every thread
        if (tid == kSortBlock - 1) carry = run + m;
        __syncthreads();
    }
    if (sbeg && tid == 0) *count = carry;
}

// the position records of one buffer from its three component arrays
__global__ void k_rep_records(const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                              double4* __restrict__ rq, int32_t ntot) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ntot) rq[i] = make_double4(x[i], y[i], z[i], 0.0);
}

__global__ void k_rep_set(int32_t R, int32_t* __restrict__ rebuild, int32_t v) {
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) rebuild[r] = v;
}

// per-replica kinetic energy, shifted-LJ potential energy and virial
// (sum over pairs of r_ij . F_ij) into out[3 r + {0, 1, 2}], FP64: every wave
// lies inside one replica (offsets are multiples of 64), so each wave sums its
// lanes by shuffles and adds the three sums once
__global__ void __launch_bounds__(kBlock)
k_rep_energy(RepStep a, double shift, double e4, double* __restrict__ out) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t r = i < a.ntot ? a.v.rep_of[i] : -1;
    double ek = 0, u = 0, w = 0;
    if (r >= 0) {
        const double L = a.v.L[r], invL = 1.0 / L;
        const double xi = a.r[0][i], yi = a.r[1][i], zi = a.r[2][i];
        const int32_t cnt = a.count[i];
        for (int32_t k = 0; k < cnt; k++) {
            const int32_t j = a.list[(int64_t)k * a.ntot + i];
            const double dx = mi(xi - a.r[0][j], L, invL), dy = mi(yi - a.r[1][j], L, invL),
                         dz = mi(zi - a.r[2][j], L, invL);
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < a.rc2) {
                const double q = a.sig2 / r2, q3 = q * q * q;
                u += 0.5 * (e4 * q3 * (q3 - 1) - shift);
                // F/r = 24 eps/sig^2 (2 q^7 - q^4), half the pair's r . F on each side
                w += 0.5 * a.coeff * q3 * q * (2 * q3 - 1) * r2;
            }
        }
        ek = 0.5 * a.mass * (a.vel[0][i] * a.vel[0][i] + a.vel[1][i] * a.vel[1][i] + a.vel[2][i] * a.vel[2][i]);
    }
    int32_t rw = r;
    for (int o = kWave / 2; o > 0; o >>= 1) {
        ek += __shfl_xor(ek, o, kWave);
        u += __shfl_xor(u, o, kWave);
        w += __shfl_xor(w, o, kWave);
        rw = max(rw, __shfl_xor(rw, o, kWave));
    }
    if ((threadIdx.x & (kWave - 1)) == 0 && rw >= 0) {
        atomicAdd(&out[3 * rw], ek);
        atomicAdd(&out[3 * rw + 1], u);
        atomicAdd(&out[3 * rw + 2], w);
    }
}

}  // namespace

struct mdx_batch {
    mdx_config cfg{};
    int32_t R = 0, ntot = 0, ncell_tot = 0;
    hipStream_t stream = nullptr;
    std::vector<int32_t> off, nat, coff, nc;
    std::vector<uint8_t> path;     // per replica: kGather, kStaged or kSmall
    std::vector<double> L;
    std::vector<void*> allocs;
    int32_t *d_rep_of = nullptr, *d_off = nullptr, *d_nat = nullptr, *d_coff = nullptr, *d_nc = nullptr;
    double* d_L = nullptr;
    int32_t *d_rebuild = nullptr, *d_cell_of = nullptr, *d_cstart = nullptr, *d_sidx = nullptr;
    float4* d_spos = nullptr;
    float4* d_tpos = nullptr;
    double4* d_rq[2] = {nullptr, nullptr};   // position records of buffers A (r) and B (rn)
    int32_t *d_list = nullptr, *d_count = nullptr, *d_err = nullptr;
    int32_t cap = 0;
    mdx_shape shape{};                    // launch shape (segments' fill, staging aim)
    // staged replicas (the single box's LDS path, see the file header)
    uint8_t* d_staged = nullptr;          // R flags
    int32_t* d_staged_i = nullptr;        // staged and small replicas (int32 lists only for the energies)
    int32_t* d_seg = nullptr;             // (replica, column, z0, z1) per segment
    int32_t* d_sbeg = nullptr;            // R + 1: first segment of each replica
    int32_t* d_wl = nullptr;              // the list build's work list and its length
    int32_t* d_wcount = nullptr;
    int32_t nseg_blocks = 0;
    int32_t* d_cstart_g = nullptr;        // cell starts as global slots
    int32_t* d_atom = nullptr;            // replica-local atom id of each slot
    double* d_ptmp = nullptr;             // permutation scratch, 15 x ntot
    int32_t* d_patom = nullptr;
    uint16_t* d_rs_list = nullptr;        // LDS-index lists of the staged replicas
    int32_t* d_hdr = nullptr;             // the list build's segment headers, read by the step kernel
    int32_t* d_rs_count = nullptr;
    int32_t* d_rs_err = nullptr;          // staging overflow, list overflow, largest staging
    int32_t rs_cap = 0, rs_lds = 0;
    int32_t n_unstaged = 0;               // replicas on the gather path,
    int32_t ubeg = 0, sfirst = 0;         // in slots [ubeg, sfirst); the small replicas from sfirst on
    int2* d_small_blocks = nullptr;       // k_rep_small's workgroups: (replica, first atom)
    int32_t sort_cap = 0;                 // atoms of the largest replica (k_rep_sort's LDS)
    int32_t nsmall_blocks = 0, small_lds = 0;
    long long* d_builds = nullptr;
    unsigned long long* d_dmax = nullptr;
    double* r[3] = {}; double* rn[3] = {}; double* v[3] = {}; double* a[3] = {}; double* b[3] = {}; double* c[3] = {};
    double* f[3] = {}; double* r0[3] = {};
    double* d_energy = nullptr;
    GearCoeffs gc{};
    bool initialized = false, flip = false;
    hipGraphExec_t graph = nullptr;  // two steps (A -> B, B -> A)
    bool graph_flip = false;
    bool timing = false;
    std::vector<std::pair<const char*, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<std::pair<std::string, std::pair<int64_t, double>>> times;
    std::string err;
};

namespace {

thread_local std::string g_batch_err;  // errors before a batch exists, per calling thread

int bfail(mdx_batch* b, int code, const std::string& m) {
    if (b) b->err = m; else g_batch_err = m;
    return code;
}

#define BCHK(b, expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return bfail(b, MDX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T> int balloc(mdx_batch* b, T** p, size_t n) {
    *p = nullptr;
    if (hipMalloc((void**)p, (n ? n : 1) * sizeof(T)) != hipSuccess) return bfail(b, MDX_ERR_ALLOC, "hipMalloc failed");
    b->allocs.push_back(*p);
    return MDX_OK;
}

int balloc3(mdx_batch* b, double* p[3], int32_t n) {
    int rc = balloc(b, &p[0], 3 * (size_t)n);
    if (rc) return rc;
    p[1] = p[0] + n;
    p[2] = p[0] + 2 * (size_t)n;
    return MDX_OK;
}

RepView view(const mdx_batch* b) {
    return RepView{b->d_rep_of, b->d_off, b->d_nat, b->d_coff, b->d_nc, b->d_L, b->d_rebuild};
}

RepStep step_args(const mdx_batch* b, bool flip) {
    RepStep s{};
    s.v = view(b);
    for (int d = 0; d < 3; d++) {
        s.r[d] = flip ? b->rn[d] : b->r[d];
        s.rn[d] = flip ? b->r[d] : b->rn[d];
        s.vel[d] = b->v[d]; s.acc[d] = b->a[d]; s.b3[d] = b->b[d]; s.c3[d] = b->c[d]; s.f[d] = b->f[d];
        s.r0[d] = b->r0[d];
    }
    s.rq = flip ? b->d_rq[1] : b->d_rq[0];
    s.rqn = flip ? b->d_rq[0] : b->d_rq[1];
    s.list = b->d_list;
    s.count = b->d_count;
    s.ntot = b->ntot;
    const mdx_config& c = b->cfg;
    const double sg = c.lj_sigma > 0 ? c.lj_sigma : 1.0, ep = c.lj_eps > 0 ? c.lj_eps : 1.0;
    s.rc2 = (double)c.lj_rcut * (double)c.lj_rcut;
    s.sig2 = sg * sg;
    s.coeff = 24.0 * ep / (sg * sg);
    s.inv_sig2f = (float)(1.0 / (sg * sg));
    s.coeff_f = (float)s.coeff;
    s.mass = c.mass > 0 ? c.mass : 1.0;
    s.scale = c.accel_scale > 0 ? c.accel_scale : 1.0;
    s.gc = b->gc;
    s.dmax = b->d_dmax;
    s.first = 0;
    s.last = b->ntot;
    return s;
}

RepSegs rep_segs(const mdx_batch* b) {
    return RepSegs{b->d_seg, b->nseg_blocks, b->d_nc, b->d_L, b->d_cstart_g, b->d_coff};
}

RepPerm rep_perm(const mdx_batch* b, bool flip) {
    RepPerm pm{};
    double* const* cur = flip ? b->rn : b->r;
    double* const* arr[5] = {cur, b->v, b->a, b->b, b->c};
    for (int k = 0; k < 5; k++)
        for (int d = 0; d < 3; d++) pm.comp[3 * k + d] = arr[k][d];
    pm.atom = b->d_atom;
    for (int d = 0; d < 3; d++) pm.r0[d] = b->r0[d];
    return pm;
}

inline int grid_for(int64_t n) { return (int)std::max<int64_t>(1, (n + kBlock - 1) / kBlock); }

template <class F> hipError_t btimed(mdx_batch* b, const char* name, F&& launch) {
    if (!b->timing) return launch();
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    hipEventRecord(e0, b->stream);
    hipError_t e = launch();
    hipEventRecord(e1, b->stream);
    b->pending.push_back({name, {e0, e1}});
    return e;
}
#define BLAUNCH(b, name, expr) BCHK(b, btimed(b, name, [&]() { (expr); return hipGetLastError(); }))

// the int32 lists of k_rep_list for the replicas gate[] names, in slots
// [first, last); write_r0 records the list-build positions
int enqueue_int_lists(mdx_batch* b, bool flip, const int32_t* gate, int32_t first, int32_t last, int32_t write_r0) {
    if (first >= last) return MDX_OK;
    hipStream_t s = b->stream;
    const RepView v = view(b);
    double* const* hp = flip ? b->rn : b->r;
    const float rs = (float)b->cfg.lj_rs + 1e-4f;
    BLAUNCH(b, "rep_list", (k_rep_list<<<grid_for(4 * (int64_t)(last - first)), kBlock, 0, s>>>(
                               v, hp[0], hp[1], hp[2], b->r0[0], b->r0[1], b->r0[2], b->ntot, b->d_cell_of,
                               b->d_cstart, b->d_spos, rs * rs, b->cap, b->d_list, b->d_count, b->d_err, gate, first,
                               last, write_r0)));
    return MDX_OK;
}

// gated list rebuild of the flagged replicas; at the start of a run (all
// flagged) the staged replicas get int32 lists too, for the first forces
int enqueue_rebuild(mdx_batch* b, bool flip, bool start = false) {
    hipStream_t s = b->stream;
    const RepView v = view(b);
    double* const* hp = flip ? b->rn : b->r;
    static const bool attr_set = [] {   // once per process, thread-safe
        hipFuncSetAttribute(reinterpret_cast<const void*>(k_rep_sort), hipFuncAttributeMaxDynamicSharedMemorySize,
                            kMaxReplica * (int)(sizeof(int32_t) + sizeof(uint16_t)));
        return true;
    }();
    (void)attr_set;
    BLAUNCH(b, "rep_sort", (k_rep_sort<<<b->R, kSortBlock, (sizeof(int32_t) + sizeof(uint16_t)) * b->sort_cap, s>>>(
                               v, hp[0], hp[1], hp[2], b->d_cell_of, b->d_cstart, b->d_cstart_g, b->d_sidx, b->d_tpos,
                               b->d_spos, b->d_staged, b->sort_cap)));
    const bool staged = b->nseg_blocks > 0;
    if (staged) {
        const RepPerm pm = rep_perm(b, flip);
        BLAUNCH(b, "rep_perm", (k_rep_perm_gather<<<grid_for(b->ntot), kBlock, 0, s>>>(v, b->d_staged, b->d_sidx, pm,
                                                                                        b->d_ptmp, b->d_patom, b->ntot)));
        BLAUNCH(b, "rep_perm", (k_rep_perm_scatter<<<grid_for(b->ntot), kBlock, 0, s>>>(v, b->d_staged, pm, b->d_ptmp,
                                                                                         b->d_patom, b->ntot)));
    }
    // the gather path's replicas lie in [ubeg, sfirst); the first forces of a
    // run come from the int32 lists of every replica
    int rc = start ? enqueue_int_lists(b, flip, b->d_rebuild, 0, b->ntot, 1)
                   : enqueue_int_lists(b, flip, b->d_rebuild, b->ubeg, b->sfirst, 1);
    if (rc) return rc;
    if (staged) {
        const double rl2 = (double)b->cfg.lj_rs * (double)b->cfg.lj_rs;
        // after the first build k_rep_decide writes the work list with the flags
        if (start)
            BLAUNCH(b, "rep_worklist", (k_rep_worklist<<<1, kSortBlock, 0, s>>>(b->R, b->d_rebuild, b->d_sbeg,
                                                                                 b->d_wl, b->d_wcount)));
        BLAUNCH(b, "rep_nlist", launch_rep_nlist(s, rep_segs(b), b->d_wl, b->d_wcount, hp, rl2, b->rs_cap,
                                                 b->d_rs_list, b->d_rs_count, b->d_rs_err, lds_capacity(),
                                                 b->d_hdr));
    }
    return MDX_OK;
}

int enqueue_decide(mdx_batch* b) {
    const double margin = 0.5 * ((double)b->cfg.lj_rs - (double)b->cfg.lj_rcut);
    const bool staged = b->nseg_blocks > 0;
    BLAUNCH(b, "rep_decide", (k_rep_decide<<<1, kSortBlock, 0, b->stream>>>(
                                 b->R, b->d_dmax, b->d_rebuild, b->d_builds, margin * margin,
                                 staged ? b->d_sbeg : nullptr, b->d_wl, b->d_wcount)));
    return MDX_OK;
}

int enqueue_step(mdx_batch* b, bool flip, bool fin) {
    hipStream_t s = b->stream;
    int rc = enqueue_rebuild(b, flip);
    if (rc) return rc;
    const RepStep a = step_args(b, flip);
    const bool fp64 = b->cfg.precision == MDX_PREC_FP64;
    if (b->nseg_blocks > 0) {
        DevState st{};
        double* const* cur = flip ? b->rn : b->r;
        for (int d = 0; d < 3; d++) {
            st.r[d] = cur[d]; st.v[d] = b->v[d]; st.a[d] = b->a[d]; st.b[d] = b->b[d]; st.c[d] = b->c[d];
            st.f[d] = b->f[d];
        }
        double* const* nxt = flip ? b->r : b->rn;
        BLAUNCH(b, "rep_step_lds", launch_rep_step(s, rep_segs(b), st, nxt, b->r0, b->d_rs_list, b->d_rs_count,
                                                   b->rs_cap, a.rc2, a.sig2, a.coeff, a.mass, a.scale, b->gc, b->ntot,
                                                   b->cfg.precision, fin, b->d_dmax, b->d_rs_err, b->rs_lds,
                                                   b->d_hdr));
    }
    if (b->nsmall_blocks > 0) {
        const size_t dyn = (size_t)b->small_lds;
        if (fin) {
            if (fp64) BLAUNCH(b, "rep_small", (k_rep_small<true, true><<<b->nsmall_blocks, kBlock, dyn, s>>>(a, b->d_small_blocks)));
            else BLAUNCH(b, "rep_small", (k_rep_small<false, true><<<b->nsmall_blocks, kBlock, dyn, s>>>(a, b->d_small_blocks)));
        } else {
            if (fp64) BLAUNCH(b, "rep_small", (k_rep_small<true, false><<<b->nsmall_blocks, kBlock, dyn, s>>>(a, b->d_small_blocks)));
            else BLAUNCH(b, "rep_small", (k_rep_small<false, false><<<b->nsmall_blocks, kBlock, dyn, s>>>(a, b->d_small_blocks)));
        }
    }
    if (b->n_unstaged > 0) {
        RepStep ag = a;
        ag.first = b->ubeg;
        ag.last = b->sfirst;
        const int g = grid_for(4 * (int64_t)(b->sfirst - b->ubeg));
        if (fin) {
            if (fp64) BLAUNCH(b, "rep_step", (k_rep_step<true, 1><<<g, kBlock, 0, s>>>(ag)));
            else BLAUNCH(b, "rep_step", (k_rep_step<false, 1><<<g, kBlock, 0, s>>>(ag)));
        } else {
            if (fp64) BLAUNCH(b, "rep_step", (k_rep_step<true, 0><<<g, kBlock, 0, s>>>(ag)));
            else BLAUNCH(b, "rep_step", (k_rep_step<false, 0><<<g, kBlock, 0, s>>>(ag)));
        }
    }
    return fin ? MDX_OK : enqueue_decide(b);
}

int check_err(mdx_batch* b) {
    int32_t e[2], f[3] = {0, 0, 0};
    BCHK(b, hipMemcpyAsync(e, b->d_err, sizeof e, hipMemcpyDeviceToHost, b->stream));
    if (b->nseg_blocks > 0) BCHK(b, hipMemcpyAsync(f, b->d_rs_err, sizeof f, hipMemcpyDeviceToHost, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    if (e[1]) return bfail(b, MDX_ERR_CAPACITY, "a neighbour list exceeded max_neighbors (longest " + std::to_string(e[1]) + ")");
    if (f[1]) return bfail(b, MDX_ERR_CAPACITY, "a neighbour list exceeded max_neighbors (longest " + std::to_string(f[1]) + ")");
    if (f[0]) return bfail(b, MDX_ERR_CAPACITY, "a replica column's neighbourhood exceeds the LDS staging (" +
                                                    std::to_string(f[0]) + " atoms)");
    // the staged step's LDS: the largest staging so far (f[2], from every list
    // build) with a fifth to grow; sized again, with a new graph, once that
    // staging comes within a tenth of it (a step staging more than it holds
    // fails the run with the error above)
    if (b->nseg_blocks > 0 && (b->rs_lds == 0 || 1.1 * f[2] > b->rs_lds)) {
        b->rs_lds = std::min<int32_t>(lds_capacity(), ((int32_t)(1.2 * f[2]) + 64 + 63) / 64 * 64);
        if (b->graph) {
            hipGraphExecDestroy(b->graph);
            b->graph = nullptr;
        }
    }
    return MDX_OK;
}

}  // namespace

extern "C" {

int mdx_batch_create(const mdx_config* common, int32_t n_replicas, const int32_t* n_atoms, const double* box_length,
                     mdx_batch** out) {
    if (!out) return bfail(nullptr, MDX_ERR_ARG, "null output");
    *out = nullptr;
    if (!common || n_replicas <= 0 || !n_atoms || !box_length) return bfail(nullptr, MDX_ERR_ARG, "bad arguments");
    const mdx_config& c = *common;
    if (!(c.lj_rcut > 0 && c.lj_rs >= c.lj_rcut)) return bfail(nullptr, MDX_ERR_ARG, "need 0 < lj_rcut <= lj_rs");
    if (c.precision != MDX_PREC_MIXED && c.precision != MDX_PREC_FP64) return bfail(nullptr, MDX_ERR_ARG, "unknown precision");
    mdx_batch* b = new mdx_batch;
    b->cfg = c;
    b->R = n_replicas;
    {
        // the launch shape for the batch's total atom count from the tuning table
        int64_t total = 0;
        for (int32_t r = 0; r < n_replicas; r++) total += n_atoms[r] > 0 ? n_atoms[r] : 0;
        std::string why;
        const int rc = tuned_shape("replicas", total, &b->shape, why);
        if (rc) {
            delete b;
            return bfail(nullptr, rc, why);
        }
    }
    const double stage_aim = b->shape.stage_aim;
    int64_t coff = 0;
    for (int32_t r = 0; r < n_replicas; r++) {
        const double L = box_length[r];
        if (n_atoms[r] <= 0 || n_atoms[r] > kMaxReplica || !(L >= 2.0 * c.lj_rs)) {
            delete b;
            return bfail(nullptr, MDX_ERR_ARG, "replica " + std::to_string(r) + ": need 0 < n <= " +
                                                   std::to_string(kMaxReplica) + " and box >= 2 rs");
        }
        b->nat.push_back(n_atoms[r]);
        b->L.push_back(L);
        int32_t nc = (int32_t)std::floor(L / (double)c.lj_rs);
        nc = nc < 3 ? 0 : (nc > kMaxNc ? kMaxNc : nc);
        b->nc.push_back(nc);
        b->coff.push_back((int32_t)coff);
        coff += (nc ? (int64_t)nc * nc * nc : 1) + 1;
    }
    // staged replicas: at least 3 cells per axis, the z-columns cut into
    // segments of about fill x 128 home atoms (as the single box cuts them),
    // shorter where their neighbourhood, 9 columns of K + 2 cells, would stage
    // more than the shape's stage_aim atoms; a replica whose segments cannot fit the LDS
    // staging with a fifth to spare stays on the gather path.  Boxes narrower
    // than 3 cells of up to kSmallMax atoms take k_rep_small.
    std::vector<uint8_t> staged(n_replicas, 0), path(n_replicas, kGather);
    std::vector<int32_t> staged_i(n_replicas, 0), seg, sbeg(n_replicas + 1, 0);
    for (int32_t r = 0; r < n_replicas; r++) {
        const int32_t nc = b->nc[r];
        sbeg[r] = (int32_t)(seg.size() / 4);
        if (nc == 0) {
            if (n_atoms[r] <= kSmallMax) {
                path[r] = kSmall;
                staged_i[r] = 1;
            }
            continue;
        }
        const double per_cell = (double)n_atoms[r] / ((double)nc * nc * nc);
        int32_t nseg = lds_segments_per_column(nc, per_cell, nc, 512, b->shape.fill);
        while (nseg < nc && 9.0 * ((nc + nseg - 1) / nseg + 2) * per_cell > stage_aim) nseg++;
        const int32_t K = (nc + nseg - 1) / nseg;
        if (1.2 * 9.0 * (K + 2) * per_cell + 64 > lds_capacity()) continue;
        staged[r] = 1;
        staged_i[r] = 1;
        path[r] = kStaged;
        for (int32_t col = 0; col < nc * nc; col++)
            for (int32_t k = 0; k < nseg; k++) {
                seg.push_back(r);
                seg.push_back(col);
                seg.push_back(k * nc / nseg);
                seg.push_back((k + 1) * nc / nseg);
            }
    }
    // layout: the staged replicas first, the gather path's from ubeg on, the
    // small ones from sfirst on
    b->off.assign(n_replicas + 1, 0);
    int64_t off = 0, ubeg = 0, sfirst = 0;
    std::vector<int2> small_blocks;
    int32_t small_max = 0;
    for (const uint8_t kind : {kStaged, kGather, kSmall}) {
        if (kind == kGather) ubeg = off;
        if (kind == kSmall) sfirst = off;
        for (int32_t r = 0; r < n_replicas; r++)
            if (path[r] == kind) {
                b->off[r] = (int32_t)std::min<int64_t>(off, INT32_MAX);
                off += ((int64_t)n_atoms[r] + kWave - 1) / kWave * kWave;
                if (kind == kSmall) {
                    for (int32_t a0 = 0; a0 < n_atoms[r]; a0 += kSmallAtoms) small_blocks.push_back(make_int2(r, a0));
                    small_max = std::max(small_max, n_atoms[r]);
                }
            }
    }
    if (off > INT32_MAX / 2 || coff > (int64_t(1) << 27)) {
        delete b;
        return bfail(nullptr, MDX_ERR_ARG, "batch too large");
    }
    b->off[n_replicas] = (int32_t)off;
    b->sort_cap = *std::max_element(n_atoms, n_atoms + n_replicas);
    b->ubeg = (int32_t)ubeg;
    b->sfirst = (int32_t)sfirst;
    b->nsmall_blocks = (int32_t)small_blocks.size();
    b->small_lds = 3 * (int32_t)sizeof(double) * small_max;
    b->coff.push_back((int32_t)coff);
    b->ntot = (int32_t)off;
    b->ncell_tot = (int32_t)coff;
    auto bail = [&](int rc) { mdx_batch_destroy(b); return rc; };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return bail(bfail(nullptr, MDX_ERR_HIP, "stream"));
    // a list capacity from the densest replica's mean count inside rs, with slack
    double rho_max = 0;
    for (int32_t r = 0; r < n_replicas; r++) rho_max = std::max(rho_max, n_atoms[r] / std::pow(b->L[r], 3));
    b->cap = c.max_neighbors > 0 ? c.max_neighbors
                                 : (int32_t)(1.5 * 4.0 / 3.0 * M_PI * std::pow((double)c.lj_rs, 3) * rho_max) + 32;
    std::vector<int32_t> rep_of(b->ntot, -1);
    for (int32_t r = 0; r < n_replicas; r++)
        for (int32_t i = 0; i < n_atoms[r]; i++) rep_of[b->off[r] + i] = r;
    const int32_t n = b->ntot;
    int rc;
    if ((rc = balloc(b, &b->d_rep_of, n)) || (rc = balloc(b, &b->d_off, n_replicas + 1)) ||
        (rc = balloc(b, &b->d_nat, n_replicas)) || (rc = balloc(b, &b->d_coff, n_replicas + 1)) ||
        (rc = balloc(b, &b->d_nc, n_replicas)) || (rc = balloc(b, &b->d_L, n_replicas)) ||
        (rc = balloc(b, &b->d_rebuild, n_replicas)) || (rc = balloc(b, &b->d_builds, n_replicas)) ||
        (rc = balloc(b, &b->d_dmax, n_replicas)) || (rc = balloc(b, &b->d_cell_of, n)) ||
        (rc = balloc(b, &b->d_cstart, b->ncell_tot)) || (rc = balloc(b, &b->d_sidx, n)) ||
        (rc = balloc(b, &b->d_spos, n)) || (rc = balloc(b, &b->d_tpos, n)) ||
        (rc = balloc(b, &b->d_rq[0], n)) || (rc = balloc(b, &b->d_rq[1], n)) ||
        (rc = balloc(b, &b->d_list, (size_t)b->cap * n)) || (rc = balloc(b, &b->d_count, n)) ||
        (rc = balloc(b, &b->d_err, 2)) ||
        (rc = balloc(b, &b->d_energy, 3 * (size_t)n_replicas)) ||
        (rc = balloc3(b, b->r, n)) || (rc = balloc3(b, b->rn, n)) || (rc = balloc3(b, b->v, n)) ||
        (rc = balloc3(b, b->a, n)) || (rc = balloc3(b, b->b, n)) || (rc = balloc3(b, b->c, n)) ||
        (rc = balloc3(b, b->f, n)) || (rc = balloc3(b, b->r0, n))) {
        g_batch_err = b->err;
        return bail(rc);
    }
    b->nseg_blocks = (int32_t)(seg.size() / 4);
    sbeg[n_replicas] = b->nseg_blocks;
    b->path = path;
    for (int32_t r = 0; r < n_replicas; r++) b->n_unstaged += path[r] == kGather ? 1 : 0;
    if (b->nsmall_blocks > 0 &&
        ((rc = balloc(b, &b->d_small_blocks, small_blocks.size())) ||
         hipMemcpyAsync(b->d_small_blocks, small_blocks.data(), sizeof(int2) * small_blocks.size(),
                        hipMemcpyHostToDevice, b->stream) != hipSuccess)) {
        g_batch_err = b->err;
        return bail(rc ? rc : bfail(nullptr, MDX_ERR_HIP, "batch upload failed"));
    }
    if (b->nseg_blocks > 0) {
        b->rs_cap = (b->cap + list_cap_multiple() - 1) / list_cap_multiple() * list_cap_multiple();
        if ((rc = balloc(b, &b->d_staged, n_replicas)) ||
            (rc = balloc(b, &b->d_seg, seg.size())) || (rc = balloc(b, &b->d_cstart_g, b->ncell_tot)) ||
            (rc = balloc(b, &b->d_sbeg, sbeg.size())) || (rc = balloc(b, &b->d_wl, b->nseg_blocks)) ||
            (rc = balloc(b, &b->d_wcount, 1)) ||
            (rc = balloc(b, &b->d_ptmp, 15 * (size_t)n)) || (rc = balloc(b, &b->d_patom, n)) ||
            (rc = balloc(b, &b->d_rs_list, (size_t)list_entries(n, b->rs_cap))) ||
            (rc = balloc(b, &b->d_hdr, (size_t)b->nseg_blocks * seg_header_words())) ||
            (rc = balloc(b, &b->d_rs_count, n)) || (rc = balloc(b, &b->d_rs_err, 4))) {
            g_batch_err = b->err;
            return bail(rc);
        }
    } else if ((rc = balloc(b, &b->d_staged, n_replicas)) || (rc = balloc(b, &b->d_cstart_g, b->ncell_tot))) {
        g_batch_err = b->err;
        return bail(rc);
    }
    if ((rc = balloc(b, &b->d_atom, n)) || (rc = balloc(b, &b->d_staged_i, n_replicas))) {
        g_batch_err = b->err;
        return bail(rc);
    }
    hipStream_t s = b->stream;
    bool staged_ok = hipMemcpyAsync(b->d_staged, staged.data(), n_replicas, hipMemcpyHostToDevice, s) == hipSuccess &&
                     hipMemcpyAsync(b->d_staged_i, staged_i.data(), sizeof(int32_t) * n_replicas,
                                    hipMemcpyHostToDevice, s) == hipSuccess;
    if (b->nseg_blocks > 0)
        staged_ok = staged_ok &&
                    hipMemcpyAsync(b->d_seg, seg.data(), sizeof(int32_t) * seg.size(), hipMemcpyHostToDevice, s) ==
                        hipSuccess &&
                    hipMemcpyAsync(b->d_sbeg, sbeg.data(), sizeof(int32_t) * sbeg.size(), hipMemcpyHostToDevice, s) ==
                        hipSuccess &&
                    hipMemsetAsync(b->d_rs_err, 0, 4 * sizeof(int32_t), s) == hipSuccess &&
                    hipMemsetAsync(b->d_rs_count, 0, sizeof(int32_t) * n, s) == hipSuccess;
    if (!staged_ok) return bail(bfail(nullptr, MDX_ERR_HIP, "batch upload failed"));
    bool ok = hipMemcpyAsync(b->d_rep_of, rep_of.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(b->d_off, b->off.data(), sizeof(int32_t) * (n_replicas + 1), hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(b->d_nat, b->nat.data(), sizeof(int32_t) * n_replicas, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(b->d_coff, b->coff.data(), sizeof(int32_t) * (n_replicas + 1), hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(b->d_nc, b->nc.data(), sizeof(int32_t) * n_replicas, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(b->d_L, b->L.data(), sizeof(double) * n_replicas, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemsetAsync(b->d_builds, 0, sizeof(long long) * n_replicas, s) == hipSuccess &&
              hipMemsetAsync(b->d_dmax, 0, sizeof(unsigned long long) * n_replicas, s) == hipSuccess &&
              hipMemsetAsync(b->d_err, 0, 2 * sizeof(int32_t), s) == hipSuccess &&
              hipMemsetAsync(b->d_count, 0, sizeof(int32_t) * n, s) == hipSuccess &&
              hipStreamSynchronize(s) == hipSuccess;
    if (!ok) return bail(bfail(nullptr, MDX_ERR_HIP, "batch upload failed"));
    b->gc = gear_coeffs(c.dt);
    *out = b;
    return MDX_OK;
}

int mdx_batch_destroy(mdx_batch* b) {
    if (!b) return MDX_OK;
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->graph) hipGraphExecDestroy(b->graph);
    for (void* p : b->allocs) hipFree(p);
    for (auto& p : b->pending) { hipEventDestroy(p.second.first); hipEventDestroy(p.second.second); }
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return MDX_OK;
}

const char* mdx_batch_last_error(const mdx_batch* b) { return b ? b->err.c_str() : g_batch_err.c_str(); }

int mdx_batch_set_state(mdx_batch* b, const double* pos3, const double* vel3) {
    if (!b || !pos3 || !vel3) return bfail(b, MDX_ERR_ARG, "pos3 and vel3 are required");
    const int32_t n = b->ntot;
    std::vector<double> buf(3 * (size_t)n, 0.0), vb(3 * (size_t)n, 0.0);
    std::vector<int32_t> atom(n, -1);
    int64_t src = 0;
    for (int32_t r = 0; r < b->R; r++)
        for (int32_t i = 0; i < b->nat[r]; i++, src++) {
            atom[b->off[r] + i] = i;
            for (int d = 0; d < 3; d++) {
                buf[(size_t)d * n + b->off[r] + i] = pos3[3 * src + d];
                vb[(size_t)d * n + b->off[r] + i] = vel3[3 * src + d];
            }
        }
    hipStream_t s = b->stream;
    BCHK(b, hipMemcpyAsync(b->d_atom, atom.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
    // the staged step's LDS is measured again at the first run; its graph goes
    b->rs_lds = 0;
    if (b->graph) {
        hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    // a fresh start: the state is (re)loaded into buffer A, Gear history cleared
    BCHK(b, hipMemcpyAsync(b->r[0], buf.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
    BCHK(b, hipMemcpyAsync(b->v[0], vb.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
    for (double* p : {b->a[0], b->b[0], b->c[0], b->f[0]}) BCHK(b, hipMemsetAsync(p, 0, sizeof(double) * 3 * n, s));
    BCHK(b, hipStreamSynchronize(s));
    b->flip = false;
    b->initialized = false;
    return MDX_OK;
}

int mdx_batch_get_state(mdx_batch* b, double* pos3, double* vel3, double* force3) {
    if (!b) return bfail(b, MDX_ERR_ARG, "null batch");
    const int32_t n = b->ntot;
    hipStream_t s = b->stream;
    // staged replicas keep their atoms in the cell order of their last list build
    std::vector<int32_t> atom(n);
    BCHK(b, hipMemcpyAsync(atom.data(), b->d_atom, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    BCHK(b, hipStreamSynchronize(s));
    auto get = [&](double* const* src, double* dst) -> int {
        if (!dst) return MDX_OK;
        std::vector<double> buf(3 * (size_t)n);
        BCHK(b, hipMemcpyAsync(buf.data(), src[0], sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s));
        BCHK(b, hipStreamSynchronize(s));
        int64_t q0 = 0;
        for (int32_t r = 0; r < b->R; r++) {
            for (int32_t p = 0; p < b->nat[r]; p++) {
                const int64_t q = q0 + atom[b->off[r] + p];
                for (int d = 0; d < 3; d++) dst[3 * q + d] = buf[(size_t)d * n + b->off[r] + p];
            }
            q0 += b->nat[r];
        }
        return MDX_OK;
    };
    int rc;
    if ((rc = get(b->flip ? b->rn : b->r, pos3)) || (rc = get(b->v, vel3)) || (rc = get(b->f, force3))) return rc;
    return MDX_OK;
}

// The first call after set_state builds all lists and computes a = F/m.  Each
// call then runs the predictor of its first step and nsteps steps of [gated
// rebuild] + step kernel + decision (the last one corrector-only, leaving the
// corrected state).  Without timing, pairs of steps replay one captured graph.
int mdx_batch_run(mdx_batch* b, int32_t nsteps) {
    if (!b) return bfail(b, MDX_ERR_ARG, "null batch");
    if (nsteps < 0) return bfail(b, MDX_ERR_ARG, "nsteps must be >= 0");
    hipStream_t s = b->stream;
    int rc;
    if (!b->initialized) {
        // first call after set_state: lists for all replicas, forces, a = F/m
        // (cpu/cpumd/Host.cpp:132-138)
        BLAUNCH(b, "rep_set", (k_rep_set<<<(b->R + 255) / 256, 256, 0, s>>>(b->R, b->d_rebuild, 1)));
        if ((rc = enqueue_rebuild(b, b->flip, true))) return rc;
        if ((rc = check_err(b))) return rc;
        double* const* cur = b->flip ? b->rn : b->r;
        BLAUNCH(b, "rep_records", (k_rep_records<<<grid_for(b->ntot), kBlock, 0, s>>>(cur[0], cur[1], cur[2],
                                                                                       b->d_rq[b->flip ? 1 : 0], b->ntot)));
        RepStep a = step_args(b, b->flip);   // the first forces of every replica, from the int32 lists
        if (b->cfg.precision == MDX_PREC_FP64)
            BLAUNCH(b, "rep_force0", (k_rep_step<true, 2><<<grid_for(4 * (int64_t)b->ntot), kBlock, 0, s>>>(a)));
        else
            BLAUNCH(b, "rep_force0", (k_rep_step<false, 2><<<grid_for(4 * (int64_t)b->ntot), kBlock, 0, s>>>(a)));
        b->initialized = true;
    }
    if (nsteps == 0) { BCHK(b, hipStreamSynchronize(s)); return MDX_OK; }
    // predictor of this call's first step (in place), then the exact rebuild decision
    BLAUNCH(b, "rep_predict", (k_rep_predict<<<grid_for(b->ntot), kBlock, 0, s>>>(step_args(b, b->flip))));
    if ((rc = enqueue_decide(b))) return rc;
    int32_t k = 0;
    const int32_t body = nsteps - 1;   // the last step is the corrector-only variant
    if (!b->timing) {
        if (!b->graph || b->graph_flip != b->flip) {
            if (b->graph) hipGraphExecDestroy(b->graph);
            b->graph = nullptr;
            hipGraph_t g;
            BCHK(b, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int r1 = enqueue_step(b, b->flip, false);
            int r2 = r1 ? r1 : enqueue_step(b, !b->flip, false);
            hipError_t e = hipStreamEndCapture(s, &g);
            if (r1 || r2) return r1 ? r1 : r2;
            if (e != hipSuccess) return bfail(b, MDX_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (e != hipSuccess) return bfail(b, MDX_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
            b->graph_flip = b->flip;
        }
        for (; k + 2 <= body; k += 2) BCHK(b, hipGraphLaunch(b->graph, s));
    }
    for (; k < body; k++) {
        if ((rc = enqueue_step(b, b->flip, false))) return rc;
        b->flip = !b->flip;
    }
    if ((rc = enqueue_step(b, b->flip, true))) return rc;
    b->flip = !b->flip;
    // the state is corrected r(t); the next call continues with the predictor
    return check_err(b);
}

int mdx_batch_build_counts(mdx_batch* b, int64_t* builds) {
    if (!b || !builds) return bfail(b, MDX_ERR_ARG, "null argument");
    std::vector<long long> h(b->R);
    BCHK(b, hipMemcpyAsync(h.data(), b->d_builds, sizeof(long long) * b->R, hipMemcpyDeviceToHost, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    for (int32_t r = 0; r < b->R; r++) builds[r] = h[r];
    return MDX_OK;
}

int mdx_batch_staged(mdx_batch* b, int32_t* path) {
    if (!b || !path) return bfail(b, MDX_ERR_ARG, "null argument");
    for (int32_t r = 0; r < b->R; r++) path[r] = b->path[r];
    return MDX_OK;
}

// per replica {kinetic, potential, virial} into out3 (3 R doubles)
static int batch_energies(mdx_batch* b, double* out3) {
    hipStream_t s = b->stream;
    BCHK(b, hipMemsetAsync(b->d_energy, 0, sizeof(double) * 3 * b->R, s));
    if (b->nseg_blocks > 0 || b->nsmall_blocks > 0) {
        // int32 lists of the staged and small replicas from their last build (the
        // list-build positions stay as they are, so the rebuild schedule is unchanged)
        int rc = enqueue_int_lists(b, b->flip, b->d_staged_i, 0, b->ntot, 0);
        if (rc) return rc;
    }
    const RepStep a = step_args(b, b->flip);
    const mdx_config& c = b->cfg;
    const double sg = c.lj_sigma > 0 ? c.lj_sigma : 1.0, ep = c.lj_eps > 0 ? c.lj_eps : 1.0;
    const double qc = a.rc2 / (sg * sg), iq6 = 1.0 / (qc * qc * qc);
    BLAUNCH(b, "rep_energy", (k_rep_energy<<<grid_for(b->ntot), kBlock, 0, s>>>(a, 4 * ep * iq6 * (iq6 - 1), 4 * ep,
                                                                                 b->d_energy)));
    BCHK(b, hipMemcpyAsync(out3, b->d_energy, sizeof(double) * 3 * b->R, hipMemcpyDeviceToHost, s));
    BCHK(b, hipStreamSynchronize(s));
    return MDX_OK;
}

int mdx_batch_energies(mdx_batch* b, double* out2) {
    if (!b || !out2) return bfail(b, MDX_ERR_ARG, "null argument");
    std::vector<double> e3(3 * (size_t)b->R);
    int rc = batch_energies(b, e3.data());
    if (rc) return rc;
    for (int32_t r = 0; r < b->R; r++) {
        out2[2 * r] = e3[3 * r];
        out2[2 * r + 1] = e3[3 * r + 1];
    }
    return MDX_OK;
}

int mdx_batch_virials(mdx_batch* b, double* out) {
    if (!b || !out) return bfail(b, MDX_ERR_ARG, "null argument");
    std::vector<double> e3(3 * (size_t)b->R);
    int rc = batch_energies(b, e3.data());
    if (rc) return rc;
    for (int32_t r = 0; r < b->R; r++) out[r] = e3[3 * r + 2];
    return MDX_OK;
}

int mdx_batch_set_timing(mdx_batch* b, int32_t enabled) {
    if (!b) return bfail(b, MDX_ERR_ARG, "null batch");
    b->timing = enabled != 0;
    return MDX_OK;
}

int mdx_batch_kernel_times(mdx_batch* b, char* names, int64_t* counts, double* total_ms, int32_t cap, int32_t* n) {
    if (!b) return bfail(b, MDX_ERR_ARG, "null batch");
    BCHK(b, hipStreamSynchronize(b->stream));
    for (auto& p : b->pending) {
        float ms = 0;
        hipEventElapsedTime(&ms, p.second.first, p.second.second);
        auto it = std::find_if(b->times.begin(), b->times.end(), [&](auto& t) { return t.first == p.first; });
        if (it == b->times.end()) { b->times.push_back({p.first, {0, 0.0}}); it = b->times.end() - 1; }
        it->second.first += 1;
        it->second.second += ms;
        hipEventDestroy(p.second.first);
        hipEventDestroy(p.second.second);
    }
    b->pending.clear();
    int32_t k = 0;
    for (auto& t : b->times) {
        if (k >= cap) break;
        std::strncpy(names + 48 * k, t.first.c_str(), 47);
        names[48 * k + 47] = 0;
        counts[k] = t.second.first;
        total_ms[k] = t.second.second;
        k++;
    }
    if (n) *n = k;
    b->times.clear();
    return MDX_OK;
}

int mdx_batch_synchronize(mdx_batch* b) {
    if (!b) return bfail(b, MDX_ERR_ARG, "null batch");
    BCHK(b, hipStreamSynchronize(b->stream));
    return MDX_OK;
}

int mdx_batch_launch_shape(mdx_batch* b, mdx_shape* out) {
    if (!b || !out) return bfail(b, MDX_ERR_ARG, "null argument");
    *out = b->shape;
    return MDX_OK;
}

}  // extern "C"
