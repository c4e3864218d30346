// cells.hip -- cell-list build on gfx950: atom binning, integer counting sort
// into cell order, and the state permutation that makes neighbour cells
// contiguous slot ranges.  Restates the set semantics of
// common/NeighborList.cpp:14-156 (cells of edge rs, 27-cell stencil,
// ascending index within a cell) with dense-grid integer work instead of the
// reference's hashed buckets.
#include "mdx_internal.h"

namespace mdx {

namespace {

constexpr int kBlock = 256;
constexpr int kScanItems = 8;                      // items per thread in the scan
constexpr int kScanTile = kBlock * kScanItems;     // items per block

inline int grid_for(int64_t n, int block = kBlock) {
    int64_t g = (n + block - 1) / block;
    return (int)(g < 1 ? 1 : g);
}

__device__ inline int32_t cell_coord_walled(double x, double npos, double h) {
    return (int32_t)((x + npos) / h);
}

// Periodic wrap into [0, L); identical in binning and permutation so an
// atom's stored position always lies in the cell it was binned into.
__device__ inline double wrap_periodic(double x, double L) {
    double y = x - L * floor(x / L);
    return y >= L ? y - L : y;
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v, int32_t* __restrict__ zero,
                           int32_t nzero) {
    if (blockIdx.x == 0 && (int32_t)threadIdx.x < nzero) zero[threadIdx.x] = 0;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// Runs of equal keys among the lanes of a wave (the active lanes must be a
// prefix of the wave, as at the tail of a 1-D grid).  Consecutive slots are
// mostly in one cell, so one atomic per run replaces one per item.
struct Run {
    int32_t leader;   // first lane of this lane's run
    int32_t offset;   // lane - leader
    int32_t len;      // lanes in the run
};
__device__ inline Run wave_runs(int32_t key, bool active) {
    const int32_t lane = threadIdx.x & (kWave - 1);
    const int32_t prev = __shfl_up(key, 1, kWave);
    const bool head = active && (lane == 0 || prev != key);
    const uint64_t heads = __ballot(head);
    const int32_t nact = __popcll(__ballot(active));
    const uint64_t upto = (2ull << lane) - 1;          // lanes 0..lane (all 64 at lane 63)
    Run r;
    r.leader = 63 - __clzll(heads & upto);
    const uint64_t later = heads & ~upto;
    const int32_t end = later ? __ffsll((unsigned long long)later) - 1 : nact;
    r.offset = lane - r.leader;
    r.len = end - r.leader;
    return r;
}

// One thread per item; item q lives in slot item_slot[q] (or slot q).
__global__ void k_bin(DevState st, const int32_t* __restrict__ item_slot, int32_t items, GridView g,
                      int32_t* __restrict__ cell_out, int32_t* __restrict__ counts,
                      int32_t* __restrict__ err) {
    const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = q < items;
    const int32_t s = active ? (item_slot ? item_slot[q] : q) : 0;
    int32_t c[3];
    bool ok = true;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double x = active ? st.r[d][s] : 0.0;
        // range-check in floating point first so NaN or huge values never
        // reach the integer conversion
        if (g.periodic) {
            double y = wrap_periodic(x, g.L) / g.h;
            ok = ok && (y >= 0.0 && y < (double)g.g[d] + 1.0);
            int32_t k = ok ? (int32_t)y : 0;
            c[d] = k >= g.g[d] ? g.g[d] - 1 : k;
        } else {
            double y = (x + g.npos) / g.h;
            ok = ok && (y > -1.0 && y < (double)g.g[d]);
            c[d] = ok ? (int32_t)y : 0;
            ok = ok && (c[d] >= 0 && c[d] < g.g[d]);
        }
    }
    // an atom outside the grid is reported and counted in cell 0
    const int32_t cell = ok ? (c[0] * g.g[1] + c[1]) * g.g[2] + c[2] : 0;
    if (active && !ok) atomicOr(err, 1);
    const Run r = wave_runs(cell, active);
    if (active) {
        cell_out[q] = cell;
        if (r.offset == 0) atomicAdd(&counts[cell], r.len);
    }
}

__device__ inline int32_t wave_incl_scan(int32_t x) {
    int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        int32_t y = __shfl_up(x, o, kWave);
        if (lane >= o) x += y;
    }
    return x;
}

// exclusive scan over a 256-thread block; *total receives the block sum
__device__ inline int32_t block_excl_scan(int32_t x, int32_t* total) {
    __shared__ int32_t wsum[kBlock / kWave];
    int32_t incl = wave_incl_scan(x);
    int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    int32_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kBlock / kWave; i++) {
        tot += wsum[i];
        if (i < w) off += wsum[i];
    }
    __syncthreads();
    *total = tot;
    return off + incl - x;
}

// pass 1: block-local exclusive prefix of kScanTile items, block totals out
__global__ void k_scan_local(const int32_t* __restrict__ in, int32_t* __restrict__ out, int64_t n,
                             int32_t* __restrict__ block_sums) {
    int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int32_t v[kScanItems];
    int32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        sum += v[k];
    }
    int32_t total;
    int32_t run = block_excl_scan(sum, &total);
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// pass 2: one block scans the block totals in place; out[n] = grand total
__global__ void k_scan_sums(int32_t* __restrict__ sums, int64_t nb, int32_t* __restrict__ out,
                            int64_t n) {
    int32_t carry = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += kBlock) {
        int64_t i = b0 + threadIdx.x;
        int32_t x = i < nb ? sums[i] : 0;
        int32_t total;
        int32_t ex = block_excl_scan(x, &total);
        if (i < nb) sums[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) out[n] = carry;
}

// pass 3: add block offsets
__global__ void k_scan_add(int32_t* __restrict__ out, int64_t n, const int32_t* __restrict__ sums) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += sums[i / kScanTile];
}

// Items into their cells' ranges (order inside a cell is settled by
// k_sort_cells); one atomic per run of equal cells in a wave.
__global__ void k_scatter(const int32_t* __restrict__ cell_of, const int32_t* __restrict__ start,
                          int32_t* __restrict__ fill, int32_t* __restrict__ sorted_items,
                          int32_t items) {
    const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = q < items;
    const int32_t c = active ? cell_of[q] : -1;
    const Run r = wave_runs(c, active);
    int32_t base = 0;
    if (active && r.offset == 0) base = atomicAdd(&fill[c], r.len);
    base = __shfl(base, r.leader, kWave);
    if (active) sorted_items[start[c] + base + r.offset] = q;
}

// Ascending key order inside each cell: one wave per cell, each lane's
// position is the number of smaller keys in the cell (keys are distinct).
constexpr int kCellsPerBlock = kBlock / kWave;
__global__ void __launch_bounds__(kBlock)
k_sort_cells(const int32_t* __restrict__ start, int64_t ncell, const int32_t* __restrict__ in,
             int32_t* __restrict__ out, const int32_t* __restrict__ key_of) {
    const int64_t c = (int64_t)blockIdx.x * kCellsPerBlock + threadIdx.x / kWave;
    if (c >= ncell) return;   // whole waves
    const int32_t lane = threadIdx.x & (kWave - 1);
    const int32_t b = start[c], cnt = start[c + 1] - b;
    if (cnt <= kWave) {
        const int32_t it = lane < cnt ? in[b + lane] : 0;
        const int32_t key = lane < cnt ? (key_of ? key_of[it] : it) : INT32_MAX;
        int32_t rank = 0;
        for (int32_t l = 0; l < cnt; l++) rank += __builtin_amdgcn_readlane(key, l) < key ? 1 : 0;
        if (lane < cnt) out[b + rank] = it;
    } else {
        for (int32_t p = lane; p < cnt; p += kWave) {
            const int32_t it = in[b + p];
            const int32_t key = key_of ? key_of[it] : it;
            int32_t rank = 0;
            for (int32_t l = 0; l < cnt; l++) {
                const int32_t o = in[b + l];
                rank += (key_of ? key_of[o] : o) < key ? 1 : 0;
            }
            out[b + rank] = it;
        }
    }
}

// The same order for sparse grids (a few items per cell): one thread per cell
// ranks its items by comparing keys.
__global__ void __launch_bounds__(kBlock)
k_sort_cells_sparse(const int32_t* __restrict__ start, int64_t ncell, const int32_t* __restrict__ in,
                    int32_t* __restrict__ out, const int32_t* __restrict__ key_of) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= ncell) return;
    const int32_t b = start[c], cnt = start[c + 1] - b;
    for (int32_t p = 0; p < cnt; p++) {
        const int32_t it = in[b + p];
        const int32_t key = key_of ? key_of[it] : it;
        int32_t rank = 0;
        for (int32_t q = 0; q < cnt; q++) {
            const int32_t o = in[b + q];
            rank += (key_of ? key_of[o] : o) < key ? 1 : 0;
        }
        out[b + rank] = it;
    }
}

__global__ void k_permute(DevState src, DevState dst, const int32_t* __restrict__ sorted_src,
                          const int32_t* __restrict__ cell_of_src, int32_t* __restrict__ cell_of_dst,
                          int32_t* __restrict__ slot_of_id, double* rr0, double* rr1, double* rr2,
                          int32_t n, int32_t wrap, double L) {
    int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int32_t s = sorted_src[k];
    double* rref[3] = {rr0, rr1, rr2};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double x = src.r[d][s];
        if (wrap) x = wrap_periodic(x, L);
        dst.r[d][k] = x;
        if (rref[d]) rref[d][k] = x;
        dst.v[d][k] = src.v[d][s];
        dst.a[d][k] = src.a[d][s];
        dst.b[d][k] = src.b[d][s];
        dst.c[d][k] = src.c[d][s];
        dst.f[d][k] = src.f[d][s];
    }
    dst.type[k] = src.type[s];
    int32_t id = src.id[s];
    dst.id[k] = id;
    cell_of_dst[k] = cell_of_src[s];
    slot_of_id[id] = k;
}

__global__ void k_max_count(const int32_t* __restrict__ counts, int64_t ncell, int32_t* out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int32_t m = 0;
    for (; i < ncell; i += stride) m = max(m, counts[i]);
    for (int o = kWave / 2; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0) atomicMax(out, m);
}

__device__ inline void decode_cell(int32_t c, const GridView& g, int32_t& x, int32_t& y, int32_t& z) {
    z = c % g.g[2];
    y = (c / g.g[2]) % g.g[1];
    x = c / (g.g[2] * g.g[1]);
}

// neighbour cell (x+dx, y+dy, z+dz): returns -1 when skipped (walled bound)
__device__ inline int32_t neighbour_cell(const GridView& g, int32_t x, int32_t y, int32_t z) {
    if (g.periodic) {
        x = x < 0 ? x + g.g[0] : (x >= g.g[0] ? x - g.g[0] : x);
        y = y < 0 ? y + g.g[1] : (y >= g.g[1] ? y - g.g[1] : y);
        z = z < 0 ? z + g.g[2] : (z >= g.g[2] ? z - g.g[2] : z);
    } else if (x < 0 || y < 0 || z < 0 || x >= g.cells || y >= g.cells || z >= g.cells) {
        return -1;
    }
    return (x * g.g[1] + y) * g.g[2] + z;
}

__global__ void k_row_lengths(GridView g, const int32_t* __restrict__ cell_by_item, int32_t items,
                              int32_t* __restrict__ max_len) {
    int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t len = 0;
    if (q < items) {
        int32_t cx, cy, cz;
        int32_t own = cell_by_item[q];
        decode_cell(own, g, cx, cy, cz);
        for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dz = -1; dz <= 1; dz++) {
                    int32_t c = neighbour_cell(g, cx + dx, cy + dy, cz + dz);
                    if (c < 0) continue;
                    len += g.start[c + 1] - g.start[c];
                    if (c == own) len -= 1;
                }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) len = max(len, __shfl_xor(len, o, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0) atomicMax(max_len, len);
}

__global__ void k_emit_list(GridView g, const int32_t* __restrict__ cell_by_item,
                            const int32_t* __restrict__ member_ids, int32_t items, int32_t nls,
                            int32_t* __restrict__ out) {
    int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= items) return;
    int32_t* row = out + (int64_t)q * nls;
    int32_t k = 0, cx, cy, cz;
    decode_cell(cell_by_item[q], g, cx, cy, cz);
    for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dz = -1; dz <= 1; dz++) {
                int32_t c = neighbour_cell(g, cx + dx, cy + dy, cz + dz);
                if (c < 0) continue;
                for (int32_t p = g.start[c]; p < g.start[c + 1]; p++) {
                    int32_t m = member_ids[p];
                    if (m != q && k < nls) row[k++] = m;
                }
            }
    for (; k < nls; k++) row[k] = -1;
}

// move the whole state between caller order (index = id) and slot order
// (to_slots: dst[idx[i]] = src[i]; otherwise dst[i] = src[idx[i]])
__global__ void k_reorder(DevState src, DevState dst, const int32_t* __restrict__ idx, int32_t n,
                          int32_t to_slots) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t si = to_slots ? i : idx[i];
    int32_t di = to_slots ? idx[i] : i;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        dst.r[d][di] = src.r[d][si];
        dst.v[d][di] = src.v[d][si];
        dst.a[d][di] = src.a[d][si];
        dst.b[d][di] = src.b[d][si];
        dst.c[d][di] = src.c[d][si];
        dst.f[d][di] = src.f[d][si];
    }
    dst.type[di] = src.type[si];
    dst.id[di] = src.id[si];
}

__global__ void k_gather_i32(const int32_t* __restrict__ src, const int32_t* __restrict__ idx,
                             int32_t* __restrict__ dst, int32_t n) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

__global__ void k_scatter3(const double* __restrict__ s0, const double* __restrict__ s1,
                           const double* __restrict__ s2, const int32_t* __restrict__ idx, double* __restrict__ d0,
                           double* __restrict__ d1, double* __restrict__ d2, int32_t n) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t k = idx[i];
    d0[k] = s0[i];
    d1[k] = s1[i];
    d2[k] = s2[i];
}

// ---- the whole build of a small walled grid in one workgroup ----
// The kernels above run a grid build as ~15 launches, which for the walled
// box's few thousand atoms and few thousand cells costs launch latency, not
// work.  Here one workgroup of kOneBlock threads does every phase in LDS,
// separated by barriers, and writes the same outputs: the same cells, prefix
// sums, ascending keys inside each cell, permutation and sizes.  The
// histogram lives in LDS (global atomics from one workgroup would queue on a
// few hundred addresses); after the scan it holds each cell's start and,
// after the scatter, each cell's end, which the later phases read as the
// cells' ranges.
constexpr int kOneBlock = 1024;

__device__ inline int32_t one_block_excl_scan(int32_t x, int32_t* wsum, int32_t* total) {
    const int32_t incl = wave_incl_scan(x);
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    if (lane == kWave - 1) wsum[w] = incl;
    __syncthreads();
    int32_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kOneBlock / kWave; i++) {
        tot += wsum[i];
        if (i < w) off += wsum[i];
    }
    __syncthreads();
    *total = tot;
    return off + incl - x;
}

__global__ void __launch_bounds__(kOneBlock)
k_small_build(SmallBuild b) {
    __shared__ int32_t cell_l[kSmallBuildItems];   // cell of item q
    __shared__ int32_t item_l[kSmallBuildItems];   // scattered items, cell-ordered ranges
    __shared__ int32_t key_l[kSmallBuildItems];    // their keys
    __shared__ int32_t sorted_l[kSmallBuildItems]; // items in cell order, ascending key inside a cell
    __shared__ int32_t cnt_l[kSmallBuildCells + 1];  // histogram -> cell starts -> cell ends
    __shared__ int32_t wsum[kOneBlock / kWave];
    __shared__ int32_t wmax[2][kOneBlock / kWave];
    __shared__ int32_t bad_any;
    const int32_t tid = threadIdx.x;
    const GridView& g = b.g;
    const int32_t ncell = (int32_t)b.ncell;
    if (tid == 0) bad_any = 0;
    for (int32_t c = tid; c <= ncell; c += kOneBlock) cnt_l[c] = 0;
    __syncthreads();
    // bin (k_bin): an atom outside the grid is reported and counted in cell 0
    int32_t bad = 0;
    for (int32_t q = tid; q < b.items; q += kOneBlock) {
        const int32_t sl = b.item_slot ? b.item_slot[q] : q;
        int32_t c[3];
        bool ok = true;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double y = (b.src.r[d][sl] + g.npos) / g.h;
            ok = ok && (y > -1.0 && y < (double)g.g[d]);
            c[d] = ok ? (int32_t)y : 0;
            ok = ok && (c[d] >= 0 && c[d] < g.g[d]);
        }
        const int32_t cell = ok ? (c[0] * g.g[1] + c[1]) * g.g[2] + c[2] : 0;
        bad |= ok ? 0 : 1;
        cell_l[q] = cell;
        b.cell_out[q] = cell;
        atomicAdd(&cnt_l[cell], 1);
    }
    if (bad) atomicOr(&bad_any, 1);
    __syncthreads();
    // exclusive scan (each thread a contiguous run of cells) into start and
    // into cnt_l, which then holds each cell's start
    int32_t* start = const_cast<int32_t*>(g.start);
    {
        const int32_t per = (ncell + kOneBlock - 1) / kOneBlock;
        const int32_t c0 = tid * per, c1 = c0 + per < ncell ? c0 + per : ncell;
        int32_t sum = 0;
        for (int32_t c = c0; c < c1; c++) sum += cnt_l[c];
        int32_t total;
        int32_t run = one_block_excl_scan(sum, wsum, &total);
        for (int32_t c = c0; c < c1; c++) {
            const int32_t k = cnt_l[c];
            start[c] = run;
            cnt_l[c] = run;
            run += k;
        }
        if (tid == 0) start[ncell] = total;
    }
    __syncthreads();
    // scatter (any order inside a cell) with each item's key; cnt_l[c]
    // becomes the end of cell c, i.e. the start of cell c + 1
    for (int32_t q = tid; q < b.items; q += kOneBlock) {
        const int32_t pos = atomicAdd(&cnt_l[cell_l[q]], 1);
        item_l[pos] = q;
        key_l[pos] = b.key_of ? b.key_of[q] : q;
    }
    __syncthreads();
    // each item's rank among its cell's keys (keys are distinct) is its place
    for (int32_t p = tid; p < b.items; p += kOneBlock) {
        const int32_t q = item_l[p], key = key_l[p];
        const int32_t c = cell_l[q];
        const int32_t lo = c > 0 ? cnt_l[c - 1] : 0, hi = cnt_l[c];
        int32_t rank = 0;
        for (int32_t o = lo; o < hi; o++) rank += key_l[o] < key ? 1 : 0;
        b.sorted[lo + rank] = q;
        sorted_l[lo + rank] = q;
    }
    __syncthreads();
    // the state in the new slot order (k_permute), when asked for
    if (b.dst.r[0]) {
        for (int32_t k = tid; k < b.items; k += kOneBlock) {
            const int32_t sl = sorted_l[k];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                b.dst.r[d][k] = b.src.r[d][sl];
                b.dst.v[d][k] = b.src.v[d][sl];
                b.dst.a[d][k] = b.src.a[d][sl];
                b.dst.b[d][k] = b.src.b[d][sl];
                b.dst.c[d][k] = b.src.c[d][sl];
                b.dst.f[d][k] = b.src.f[d][sl];
            }
            b.dst.type[k] = b.src.type[sl];
            const int32_t id = b.src.id[sl];
            b.dst.id[k] = id;
            b.cell_of_dst[k] = cell_l[sl];
            b.slot_of_id[id] = k;
        }
    }
    // sizes: the largest cell and the longest row (27 cells, self excluded,
    // cells past `cells` skipped as in k_row_lengths)
    auto count_of = [&](int32_t c) { return cnt_l[c] - (c > 0 ? cnt_l[c - 1] : 0); };
    int32_t mc = 0, ml = 0;
    for (int32_t c = tid; c < ncell; c += kOneBlock) mc = max(mc, count_of(c));
    for (int32_t q = tid; q < b.items; q += kOneBlock) {
        int32_t cx, cy, cz;
        const int32_t own = cell_l[q];
        decode_cell(own, g, cx, cy, cz);
        int32_t len = 0;
        for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dz = -1; dz <= 1; dz++) {
                    const int32_t c = neighbour_cell(g, cx + dx, cy + dy, cz + dz);
                    if (c < 0) continue;
                    len += count_of(c);
                    if (c == own) len -= 1;
                }
        ml = max(ml, len);
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        mc = max(mc, __shfl_xor(mc, o, kWave));
        ml = max(ml, __shfl_xor(ml, o, kWave));
    }
    if ((tid & (kWave - 1)) == 0) {
        wmax[0][tid / kWave] = mc;
        wmax[1][tid / kWave] = ml;
    }
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < kOneBlock / kWave; w++) {
            mc = max(mc, wmax[0][w]);
            ml = max(ml, wmax[1][w]);
        }
        b.words[0] = bad_any;
        b.words[1] = mc;
        b.words[2] = ml;
    }
}

}  // namespace

bool small_build_fits(int32_t items, int64_t ncell) {
    return items <= kSmallBuildItems && ncell <= kSmallBuildCells;
}

hipError_t launch_small_build(hipStream_t s, const SmallBuild& b) {
    if (!small_build_fits(b.items, b.ncell)) return hipErrorInvalidValue;
    k_small_build<<<1, kOneBlock, 0, s>>>(b);
    return hipGetLastError();
}

hipError_t launch_scatter3(hipStream_t s, double* const src[3], const int32_t* idx, double* const dst[3], int32_t n) {
    if (n <= 0) return hipSuccess;
    k_scatter3<<<grid_for(n), kBlock, 0, s>>>(src[0], src[1], src[2], idx, dst[0], dst[1], dst[2], n);
    return hipGetLastError();
}

hipError_t launch_fill_i32(hipStream_t s, int32_t* p, int64_t n, int32_t v, int32_t* zero, int32_t nzero) {
    if (nzero < 0 || nzero > kBlock || (nzero > 0 && !zero)) return hipErrorInvalidValue;
    if (n <= 0 && nzero == 0) return hipSuccess;
    int g = n > 0 ? grid_for(n) : 1;
    if (g > 65536) g = 65536;
    k_fill_i32<<<g, kBlock, 0, s>>>(p, n, v, zero, nzero);
    return hipGetLastError();
}

hipError_t launch_bin(hipStream_t s, const DevState& st, const int32_t* item_slot, int32_t items,
                      const GridView& g, int32_t* cell_out, int32_t* counts, int32_t* err) {
    if (items <= 0) return hipSuccess;
    k_bin<<<grid_for(items), kBlock, 0, s>>>(st, item_slot, items, g, cell_out, counts, err);
    return hipGetLastError();
}

int64_t scan_block_sums_needed(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

hipError_t launch_exclusive_scan(hipStream_t s, const int32_t* in, int32_t* out, int64_t n,
                                 int32_t* block_sums, int64_t cap) {
    int64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb < 1) nb = 1;
    if (nb > cap) return hipErrorInvalidValue;
    k_scan_local<<<(unsigned)nb, kBlock, 0, s>>>(in, out, n, block_sums);
    k_scan_sums<<<1, kBlock, 0, s>>>(block_sums, nb, out, n);
    k_scan_add<<<grid_for(n), kBlock, 0, s>>>(out, n, block_sums);
    return hipGetLastError();
}

hipError_t launch_scatter(hipStream_t s, const int32_t* cell_of, const int32_t* start,
                          int32_t* fill, int32_t* sorted_items, int32_t items) {
    if (items <= 0) return hipSuccess;
    k_scatter<<<grid_for(items), kBlock, 0, s>>>(cell_of, start, fill, sorted_items, items);
    return hipGetLastError();
}

hipError_t launch_sort_cells(hipStream_t s, const int32_t* start, int64_t ncell, const int32_t* items_in,
                             int32_t* items_out, const int32_t* key_of_item, int64_t items) {
    if (ncell <= 0) return hipSuccess;
    if (items < 4 * ncell)   // sparse: a thread per cell
        k_sort_cells_sparse<<<grid_for(ncell, kBlock), kBlock, 0, s>>>(start, ncell, items_in, items_out, key_of_item);
    else
        k_sort_cells<<<grid_for(ncell, kCellsPerBlock), kBlock, 0, s>>>(start, ncell, items_in, items_out, key_of_item);
    return hipGetLastError();
}

hipError_t launch_permute(hipStream_t s, const DevState& src, const DevState& dst,
                          const int32_t* sorted_src, const int32_t* cell_of_src,
                          int32_t* cell_of_dst, int32_t* slot_of_id, double* const rref[3],
                          int32_t n, int32_t wrap, double L) {
    if (n <= 0) return hipSuccess;
    k_permute<<<grid_for(n), kBlock, 0, s>>>(src, dst, sorted_src, cell_of_src, cell_of_dst,
                                             slot_of_id, rref[0], rref[1], rref[2], n, wrap, L);
    return hipGetLastError();
}

hipError_t launch_max_count(hipStream_t s, const int32_t* counts, int64_t ncell, int32_t* out) {
    int g = grid_for(ncell);
    if (g > 4096) g = 4096;
    k_max_count<<<g, kBlock, 0, s>>>(counts, ncell, out);
    return hipGetLastError();
}

hipError_t launch_row_lengths(hipStream_t s, const GridView& g, const int32_t* cell_by_item,
                              int32_t items, int32_t* max_len) {
    if (items <= 0) return hipSuccess;
    k_row_lengths<<<grid_for(items), kBlock, 0, s>>>(g, cell_by_item, items, max_len);
    return hipGetLastError();
}

hipError_t launch_emit_list(hipStream_t s, const GridView& g, const int32_t* cell_by_item,
                            const int32_t* member_ids, int32_t items, int32_t nlistsize,
                            int32_t* out) {
    if (items <= 0) return hipSuccess;
    k_emit_list<<<grid_for(items), kBlock, 0, s>>>(g, cell_by_item, member_ids, items, nlistsize, out);
    return hipGetLastError();
}

hipError_t launch_reorder(hipStream_t s, const DevState& src, const DevState& dst,
                          const int32_t* idx, int32_t n, int32_t to_slots) {
    if (n <= 0) return hipSuccess;
    k_reorder<<<grid_for(n), kBlock, 0, s>>>(src, dst, idx, n, to_slots);
    return hipGetLastError();
}

hipError_t launch_gather_i32(hipStream_t s, const int32_t* src, const int32_t* idx, int32_t* dst,
                             int32_t n) {
    if (n <= 0) return hipSuccess;
    k_gather_i32<<<grid_for(n), kBlock, 0, s>>>(src, idx, dst, n);
    return hipGetLastError();
}

}  // namespace mdx
