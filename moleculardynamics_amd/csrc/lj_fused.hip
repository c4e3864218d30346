// lj_fused.hip -- periodic-box performance path: LDS-staged neighbourhoods and
// the whole MD step in one kernel.
//
// Work unit: a "segment" = K consecutive cells of one z-column (x, y fixed).
// Its atoms are one contiguous slot range (slots are cell-sorted, z fastest).
// The segment's neighbourhood -- the 9 columns (x+dx, y+dy), dx, dy in
// {-1,0,1}, each over z0-1 .. z1 -- is staged into LDS as FP64 positions
// already moved to the periodic image seen from the segment, in the order
// column (dx,dy lexicographic), then z, then slot; so every atom's 27-cell
// candidate set is 9 contiguous LDS ranges (cells in (dx,dy,dz)
// lexicographic order, ascending index within a cell, the candidate order of
// common/NeighborList.cpp:99-136).
//
// Each atom is served by a quad of 4 lanes.  A workgroup of 512 lanes holds
// 128 atoms per pass, so a segment of about 110 atoms and its ~1400 staged
// neighbours (34 KiB of LDS) leave room for 4 workgroups (32 waves) per CU:
// enough waves to cover the HBM and LDS latencies of the gathers.
//
//  k_nlist_lds (at each rebuild): the quad scans the atom's candidates, 32 at
//    a time, and keeps those within the list radius rs.  Entries are 16-bit
//    LDS indices, numbered in the quad's scan order and grouped 16 to a row:
//    entry k is entry k % 4 of lane (k / 4) % 4's 8-byte chunk in row k / 16
//    (rows of 16 waste less on the padded last row than rows of 32: an LJ
//    liquid atom has 70-85 entries, 5 or 6 rows).  The hits go to a four-row
//    ring in LDS first, and each lane writes its 8-byte chunk of a row once
//    the row is complete (the last row padded with the dummy atom).  Atoms are
//    grouped 16 to a block; block B holds its rows c = 0 .. NC-1 one after the
//    other, each as [16 atoms][4 lanes] 8-byte chunks (512 B), so a wave's
//    chunk loads cover one or two contiguous runs and an atom's rows are
//    512 B apart.
//  k_step_lds (every step): LJ force over the list (FP64 displacement, FP32
//    12-6 force law in the mixed mode, FP64 accumulation), summed over the
//    quad, then the Gear corrector of step t and the predictor of step t+1,
//    lane q handling component q.  Positions are double-buffered: r(t) is read
//    (and staged) from A and r(t+1) written to B, so no workgroup reads a
//    position another one has already advanced.  The staging layout of a
//    segment (column starts, copy pieces, home range) changes only with the
//    cells, so the list build writes it to a per-segment header table and the
//    step kernel loads it instead of deriving it from the cell starts.
// Workgroup-to-segment mapping is XCD-aware: workgroups are dealt to the 8
// XCDs round robin, so block b takes segment (b % 8) * per + b / 8 and each
// XCD's L2 serves one contiguous slab of the box.
#include "mdx_internal.h"

namespace mdx {

namespace {

constexpr int kBlock = 512;
constexpr int kLanes = 4;                  // lanes per atom
constexpr int kApp = kBlock / kLanes;      // atoms per pass (the replicas' workgroups)
constexpr int kChunk = 4;                  // 16-bit entries per 8-byte chunk
constexpr int kScan = 8;                   // candidates per lane per scan step
constexpr int kScanSlack = kLanes * kScan; // staged slots past the last atom the scan may read
constexpr int kListBlock = 16;             // atoms per list block
constexpr int kChunkRow = kListBlock * kLanes * kChunk;   // 16-bit entries in one chunk of a block (256)
constexpr int kRowEntries = kLanes * kChunk;               // entries in one row of an atom (16)
// The list build's LDS ring per quad, four rows: a scan step adds at most
// kLanes x kScan = 32 hits to a partly filled row, so it completes at most two
// rows and starts a third.
constexpr int kRing = 4 * kRowEntries;
static_assert(kLanes * kScan <= 2 * kRowEntries, "a scan step completes at most two rows");
constexpr int kXcd = 8;
constexpr int kLdsAtoms = 3072;            // staged atoms per segment at most (3 x 24 KiB of FP64)
constexpr int kMaxK = 14;                  // cells per segment
constexpr int kSpread = 64;                // flag atomics are spread over this many addresses
constexpr int kCopyBatch = 4;              // staging: independent loads in flight per thread

struct SegGrid {
    int32_t n[3];        // cells per axis
    int32_t nseg;        // segments per column
    int32_t total;       // n[0] * n[1] * nseg
    int32_t per_xcd;     // ceil(total / 8)
    float rnseg, rny;    // 1 / nseg, 1 / n[1]
    int32_t small;       // total and nseg * n[2] below 2^22: div_small applies
    double L;
    const int32_t* start;
};

// a / b for 0 <= a < 2^22 and b >= 1, given rb ~ 1 / b to a few ulp: the
// float quotient is then within 1 of the true one, and two integer tests fix
// it up exactly (an integer division by a run-time value costs ~40 VALU)
__device__ inline int32_t div_small(int32_t a, int32_t b, float rb) {
    int32_t q = (int32_t)((float)a * rb);
    q += (q + 1) * b <= a ? 1 : 0;
    q -= q * b > a ? 1 : 0;
    return q;
}

// segment number -> column (cx, cy) and z range [z0, z1); the workgroup
// mapping is XCD-aware (block b is dealt to XCD b % 8)
__device__ inline int32_t seg_geometry(const SegGrid& g, int32_t& cx, int32_t& cy, int32_t& z0, int32_t& z1) {
    const int32_t b = blockIdx.x;
    const int32_t seg = (b % kXcd) * g.per_xcd + b / kXcd;
    if (seg >= g.total) return -1;
    int32_t col, szi;
    if (g.small) {
        col = div_small(seg, g.nseg, g.rnseg);
        szi = seg - col * g.nseg;
        cx = div_small(col, g.n[1], g.rny);
        z0 = div_small(szi * g.n[2], g.nseg, g.rnseg);
        z1 = div_small((szi + 1) * g.n[2], g.nseg, g.rnseg);
    } else {
        col = seg / g.nseg;
        szi = seg % g.nseg;
        cx = col / g.n[1];
        z0 = (int32_t)(((int64_t)szi * g.n[2]) / g.nseg);
        z1 = (int32_t)(((int64_t)(szi + 1) * g.n[2]) / g.nseg);
    }
    cy = col - cx * g.n[1];
    return seg;
}

// Batched replicas (replicas.hip): segment b covers cells z0 = seg[4b+2] to
// z1 = seg[4b+3] (exclusive) of z-column seg[4b+1] of replica seg[4b]; the
// replica's cell starts are global slots.  The step kernel takes every
// segment, the list build the work list wl[0 .. *wcount) of the segments of
// the replicas that rebuild.
struct RepGeo {
    const int32_t* seg;
    int32_t nblocks;         // segments in the table
    const int32_t* nc;       // cells per axis of each replica (>= 3 here)
    const double* L;
    const int32_t* cstart;   // per replica nc^3 + 1 global slots, at coff[r]
    const int32_t* coff;
    const int32_t* wl;       // list build: segment numbers
    const int32_t* wcount;
};

// the step kernel's segment of this block: the 8 XCDs take contiguous eighths
// of the table (a replica's segments share their neighbour columns in L2)
__device__ inline int32_t rep_block(const RepGeo& rg) {
    const int32_t per = (rg.nblocks + kXcd - 1) / kXcd;
    const int32_t b = (int32_t)(blockIdx.x % kXcd) * per + (int32_t)(blockIdx.x / kXcd);
    return b < rg.nblocks ? b : -1;
}

__device__ inline int32_t rep_geometry(const RepGeo& rg, int32_t b, SegGrid& g, int32_t& cx, int32_t& cy,
                                       int32_t& z0, int32_t& z1) {
    const int4 sg = reinterpret_cast<const int4*>(rg.seg)[b];
    const int32_t r = sg.x;
    const int32_t m = rg.nc[r], col = sg.y;
    g.n[0] = g.n[1] = g.n[2] = m;
    g.nseg = 1;
    g.total = m * m;
    g.per_xcd = 0;
    g.rnseg = 1.0f;
    g.rny = 1.0f / (float)m;
    g.small = 1;
    g.L = rg.L[r];
    g.start = rg.cstart + rg.coff[r];
    cx = col / m;
    cy = col - cx * m;
    z0 = sg.z;
    z1 = sg.w;
    return r;
}

// Staged FP64 positions of up to `cap` atoms, x y z side by side (24 bytes an
// atom, so one address serves the three reads), addressed through a pointer
// derived from the __shared__ array itself so accesses stay ds_*.  The slot
// after the last staged atom holds a far-away dummy atom: list padding points
// at it, and its distance fails every cutoff test.
constexpr double kFar = 1e100;
struct Pos {
    double* p;
    int32_t cap;
    __device__ double x(int32_t i) const { return p[3 * i]; }
    __device__ double y(int32_t i) const { return p[3 * i + 1]; }
    __device__ double z(int32_t i) const { return p[3 * i + 2]; }
    __device__ void set(int32_t i, double x, double y, double z) const {
        p[3 * i] = x;
        p[3 * i + 1] = y;
        p[3 * i + 2] = z;
    }
};
// The list build's staging: FP32 positions relative to the segment's corner
// (12 bytes an atom).  Within a segment's neighbourhood the coordinates stay
// below ~30 sigma, so a distance is good to ~1e-5 sigma^2 in r^2; the build
// tests against a radius that covers that, keeping the list a superset.
struct PosF {
    float* p;
    int32_t cap;
    __device__ float x(uint32_t i) const { return p[3 * i]; }
    __device__ float y(uint32_t i) const { return p[3 * i + 1]; }
    __device__ float z(uint32_t i) const { return p[3 * i + 2]; }
    __device__ void set(int32_t i, double x, double y, double z) const {
        p[3 * i] = (float)x;
        p[3 * i + 1] = (float)y;
        p[3 * i + 2] = (float)z;
    }
};
// The part of a segment's staging layout the step kernel needs.  The list
// build writes it to a table (one per segment, kHdrWords words) and the step
// kernels load it from there: the cells do not change between builds.
struct SegHdr {
    int32_t colbase[9];                 // LDS start of each column
    int32_t plen[9][3];                 // per column, lengths of the 3 copy pieces
    int32_t psrc[9][3];                 // per column, first slot of each piece
    int32_t total;                      // staged atoms
    int32_t self0;                      // LDS index of the first home atom
    int32_t hbeg, hend;                 // slot range of the home atoms
    int32_t pad;
};
constexpr int kHdrWords = 68;
static_assert(sizeof(SegHdr) == kHdrWords * sizeof(int32_t), "segment header layout");
// LDS header shared by both kernels
struct Stage {
    SegHdr h;
    int32_t cellpre[9][kMaxK + 3];      // per column, prefix counts over the K+2 cells (list build)
    int32_t cstart[9][kMaxK + 2];       // per column and window cell, first slot (list build)
    unsigned long long red[3];          // step kernel: max disp^2 bits, max step^2 bits, stale count
};
// The step kernel's LDS header: the segment header and the block's flag
// reduction (without the list build's cell tables, 1.2 KiB less per workgroup)
struct StepStage {
    SegHdr h;
    unsigned long long red[3];          // max disp^2 bits, max step^2 bits, stale count
};

// The staging layout of segment (cx, cy, z0..z1): per column its LDS start,
// the prefix counts of its window cells and the slot ranges of its copy
// pieces (the whole block takes part; S is complete on return).
__device__ inline void stage_header(Stage& S, const SegGrid& g, int32_t cx, int32_t cy, int32_t z0, int32_t z1) {
    const int32_t K2 = z1 - z0 + 2;    // cells per column in the window
    const int32_t tid = threadIdx.x;
    if (tid < 3) S.red[tid] = 0;
    if (tid < 9 * K2) {                // per-cell counts and first slots (9 columns x K2 cells)
        const int32_t col = div_small(tid, K2, __builtin_amdgcn_rcpf((float)K2)), o = tid - col * K2;
        int32_t x = cx + col / 3 - 1, y = cy + col % 3 - 1, z = z0 - 1 + o;
        x = x < 0 ? x + g.n[0] : (x >= g.n[0] ? x - g.n[0] : x);
        y = y < 0 ? y + g.n[1] : (y >= g.n[1] ? y - g.n[1] : y);
        z = z < 0 ? z + g.n[2] : (z >= g.n[2] ? z - g.n[2] : z);
        const int32_t c = (x * g.n[1] + y) * g.n[2] + z;
        const int32_t b = g.start[c], e = g.start[c + 1];
        S.cellpre[col][o + 1] = e - b;
        S.cstart[col][o] = b;
    }
    __syncthreads();
    if (tid < 9) {
        // prefix over the window, then the copy pieces: the window z0-1 .. z1 is
        // one contiguous slot range except for a first (last) cell that wraps in z
        S.cellpre[tid][0] = 0;
        for (int32_t o = 1; o <= K2; o++) S.cellpre[tid][o] += S.cellpre[tid][o - 1];
        const bool wlo = z0 == 0, whi = z1 == g.n[2];
        const int32_t olo = wlo ? 1 : 0, ohi = whi ? K2 - 2 : K2 - 1;
        S.h.plen[tid][0] = wlo ? S.cellpre[tid][1] : 0;
        S.h.psrc[tid][0] = S.cstart[tid][0];
        S.h.plen[tid][1] = S.cellpre[tid][ohi + 1] - S.cellpre[tid][olo];
        S.h.psrc[tid][1] = S.cstart[tid][olo];
        S.h.plen[tid][2] = whi ? S.cellpre[tid][K2] - S.cellpre[tid][K2 - 1] : 0;
        S.h.psrc[tid][2] = S.cstart[tid][K2 - 1];
    }
    __syncthreads();
    if (tid == 0) {
        int32_t acc = 0;
        for (int32_t col = 0; col < 9; col++) {
            S.h.colbase[col] = acc;
            acc += S.cellpre[col][K2];
        }
        S.h.total = acc;
        // the home atoms: window cells 1 .. K2 - 2 of the centre column (z0 .. z1 - 1)
        S.h.self0 = S.h.colbase[4] + S.cellpre[4][1];
        S.h.hbeg = S.cstart[4][1];
        S.h.hend = S.cstart[4][1] + (S.cellpre[4][K2 - 1] - S.cellpre[4][1]);
        S.h.pad = 0;
    }
    __syncthreads();
}

// the slot of staged atom w (LDS index) of a segment whose header is S
__device__ inline int32_t staged_slot(const Stage& S, int32_t w) {
    int32_t col = 0;
    while (col < 8 && S.h.colbase[col + 1] <= w) col++;
    w -= S.h.colbase[col];
    const int32_t l0 = S.h.plen[col][0], l1 = S.h.plen[col][1];
    return w < l0 ? S.h.psrc[col][0] + w : (w < l0 + l1 ? S.h.psrc[col][1] + (w - l0) : S.h.psrc[col][2] + (w - l0 - l1));
}

// Fill the staging of segment (cx, cy, ...) whose header is S: positions
// minus the origin o (zero for the FP64 staging: the values are then exactly
// the images).  Returns false (for the whole block) when it does not fit in LDS.
template <int kB, typename StageT, typename PosT>
__device__ inline bool stage_fill(StageT& S, const PosT& P, const SegGrid& g, int32_t cx, int32_t cy,
                                  const double* __restrict__ rx, const double* __restrict__ ry,
                                  const double* __restrict__ rz, double ox = 0.0, double oy = 0.0, double oz = 0.0) {
    const int32_t tid = threadIdx.x;
    if (S.h.total + 1 > P.cap) return false;   // + the dummy atom
    // copy: kColThreads threads per column, kCopyBatch independent loads each
    constexpr int32_t kColThreads = kB / 9;
    if (tid < 9 * kColThreads) {
        const int32_t col = tid / kColThreads, l = tid % kColThreads;
        int32_t x = cx + col / 3 - 1, y = cy + col % 3 - 1;
        const double sx = x < 0 ? -g.L : (x >= g.n[0] ? g.L : 0.0);
        const double sy = y < 0 ? -g.L : (y >= g.n[1] ? g.L : 0.0);
        const int32_t l0 = S.h.plen[col][0], l1 = S.h.plen[col][1], l2 = S.h.plen[col][2];
        const int32_t s0 = S.h.psrc[col][0], s1 = S.h.psrc[col][1], s2 = S.h.psrc[col][2];
        const int32_t len = l0 + l1 + l2, dst = S.h.colbase[col];
        for (int32_t w0 = l; w0 < len; w0 += kCopyBatch * kColThreads) {
            double bx[kCopyBatch], by[kCopyBatch], bz[kCopyBatch];
            // every load goes out, the ones past the column's end clamped to
            // its last atom (read, never stored); the empty asm takes all of
            // them as inputs, so they are issued together and one wait covers
            // the batch (the loads are of read-only memory, which a memory
            // clobber alone would not hold back)
            int32_t src[kCopyBatch];
            double sz[kCopyBatch];
#pragma unroll
            for (int u = 0; u < kCopyBatch; u++) {
                const int32_t w = min(w0 + u * kColThreads, len - 1);
                const bool p0 = w < l0, p1 = w < l0 + l1;
                src[u] = p0 ? s0 + w : (p1 ? s1 + (w - l0) : s2 + (w - l0 - l1));
                sz[u] = p0 ? -g.L : (p1 ? 0.0 : g.L);
            }
#pragma unroll
            for (int u = 0; u < kCopyBatch; u++) {
                bx[u] = rx[src[u]];
                by[u] = ry[src[u]];
                bz[u] = rz[src[u]];
            }
            static_assert(kCopyBatch == 4, "the asm below names four atoms");
            asm volatile("" ::"v"(bx[0]), "v"(bx[1]), "v"(bx[2]), "v"(bx[3]), "v"(by[0]), "v"(by[1]), "v"(by[2]),
                         "v"(by[3]), "v"(bz[0]), "v"(bz[1]), "v"(bz[2]), "v"(bz[3]));
#pragma unroll
            for (int u = 0; u < kCopyBatch; u++) {
                bx[u] = bx[u] + sx - ox;
                by[u] = by[u] + sy - oy;
                bz[u] = bz[u] + sz[u] - oz;
            }
#pragma unroll
            for (int u = 0; u < kCopyBatch; u++) {
                const int32_t w = w0 + u * kColThreads;
                if (w < len) P.set(dst + w, bx[u], by[u], bz[u]);
            }
        }
    }
    if (tid == kB - 1) P.set(S.h.total, kFar, kFar, kFar);
    __syncthreads();
    return true;
}


// window cell (1..K) of home atom number qc of the segment
__device__ inline int32_t home_cell(const Stage& S, int32_t qc) {
    int32_t o = 1;
    while (S.cellpre[4][o + 1] - S.cellpre[4][1] <= qc) o++;
    return o;
}

// LDS bytes ahead of the list build's row rings: the FP32 staging of lds_cap
// atoms and the scan's slack, rounded up to 16 bytes
__host__ __device__ constexpr int32_t ring_offset(int32_t lds_cap) {
    return (3 * (lds_cap + kScanSlack) * (int32_t)sizeof(float) + 15) / 16 * 16;
}

// the lists of one segment's home atoms (the whole block of kB threads)
template <int kB>
__device__ __forceinline__ void nlist_segment(Stage& S, double* dyn, const SegGrid& g, int32_t cx, int32_t cy,
                                              int32_t z0, int32_t z1, const double* __restrict__ rx,
                                              const double* __restrict__ ry, const double* __restrict__ rz,
                                              double rl2, int32_t cap, uint16_t* __restrict__ list,
                                              int32_t* __restrict__ count, int32_t* __restrict__ err,
                                              int32_t lds_cap, int32_t* __restrict__ hdr_out) {
    const int32_t nchunk = cap / kRowEntries + 1;   // rows per atom with the spare
    const int32_t rows = cap / kRowEntries;         // rows that hold entries
    const PosF P{reinterpret_cast<float*>(dyn), lds_cap};
    stage_header(S, g, cx, cy, z0, z1);
    // this segment's header for the step kernels
    if (threadIdx.x < kHdrWords) hdr_out[threadIdx.x] = reinterpret_cast<const int32_t*>(&S.h)[threadIdx.x];
    if (!stage_fill<kB>(S, P, g, cx, cy, rx, ry, rz, cx * (g.L / g.n[0]), cy * (g.L / g.n[1]), z0 * (g.L / g.n[2]))) {
        if (threadIdx.x == 0) atomicMax(&err[0], S.h.total + 1);
        return;
    }
    const float rlf = (float)(rl2 + 1e-4);   // superset of the FP64 test (see PosF)
    // largest staging (with the dummy atom), sizes the step kernel's LDS
    if (threadIdx.x == 0) atomicMax(&err[2], S.h.total + 1);
    const int32_t hbeg = S.h.hbeg, hend = S.h.hend;
    const int32_t q = threadIdx.x & (kLanes - 1);
    // this quad's four-row ring; lane q's chunk of a row is 8 bytes at entry 4q
    uint16_t* const ring = reinterpret_cast<uint16_t*>(reinterpret_cast<char*>(dyn) + ring_offset(lds_cap)) +
                           (threadIdx.x / kLanes) * kRing;
    const uint16_t dummy = (uint16_t)S.h.total;
    for (int32_t base = hbeg; base < hend; base += kB / kLanes) {
        const int32_t s = base + threadIdx.x / kLanes;
        if (s >= hend) continue;             // whole quads drop out together
        const int32_t qc = s - hbeg;
        const int32_t o = home_cell(S, qc);
        const int32_t self = S.h.colbase[4] + S.cellpre[4][1] + qc;
        const float xi = P.x(self), yi = P.y(self), zi = P.z(self);
        // lane q's chunk of row r of atom s
        uint2* const la = reinterpret_cast<uint2*>(list + ((int64_t)(s / kListBlock) * nchunk * kListBlock + s % kListBlock) *
                                                              kRowEntries) + q;
        // row r of the ring to lane q's chunk in the list (rows past the
        // capacity are dropped; the count reports the overflow).  The ring was
        // written by the quad's lanes: LDS operations of one wavefront take
        // effect in program order, and the asm keeps the compiler from moving
        // the read above those writes.
        auto flush = [&](int32_t r) {
            asm volatile("" ::: "memory");
            if (r < rows)
                la[(int64_t)r * (kChunkRow / kChunk)] =
                    *reinterpret_cast<const uint2*>(ring + (r & (kRing / kRowEntries - 1)) * kRowEntries + q * kChunk);
            asm volatile("" ::: "memory");
        };
        int32_t k = 0;                        // hits so far, the same in all 4 lanes
        for (int32_t c = 0; c < 9; c++) {
            const int32_t b = S.h.colbase[c] + S.cellpre[c][o - 1], e = S.h.colbase[c] + S.cellpre[c][o + 2];
            for (int32_t p0 = b; p0 < e; p0 += kLanes * kScan) {
                // lane q tests candidates p0 + kLanes u + q (the list radius test
                // may round differently from the reference's; the list only has to
                // hold every pair that can come within rc before the next build).
                // All 8 are read first, from one base address; reads past e land
                // in the staging's slack (slots < cap + kScanSlack, so a 16-bit
                // index) and are masked.
                float cx8[kScan], cy8[kScan], cz8[kScan];
                const float* pb = P.p + 3 * (uint32_t)(uint16_t)(p0 + q);
#pragma unroll
                for (int u = 0; u < kScan; u++) {
                    cx8[u] = pb[3 * kLanes * u];
                    cy8[u] = pb[3 * kLanes * u + 1];
                    cz8[u] = pb[3 * kLanes * u + 2];
                }
                uint32_t mask = 0;
#pragma unroll
                for (int u = 0; u < kScan; u++) {
                    const float dx = xi - cx8[u], dy = yi - cy8[u], dz = zi - cz8[u];
                    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                    mask |= (uint32_t)(r2 < rlf) << u;
                }
                // candidate u is p0 + kLanes u + q: keep u below the range end,
                // and in the home column (c = 4, the same for the whole
                // wavefront) drop the atom itself
                const int32_t nv = min(max((e - p0 - q + kLanes - 1) >> 2, 0), kScan);
                mask &= (1u << nv) - 1u;
                if (c == 4) {
                    const int32_t ds = self - p0 - q;
                    if (ds >= 0 && ds < kLanes * kScan && ds % kLanes == 0) mask &= ~(1u << (ds / kLanes));
                }
                // hits are numbered lane-major within the quad; at most 32 of
                // them, so they complete at most two rows
                const int32_t h = __popc(mask);
                int32_t total;
                int32_t kk = k + quad_excl(h, q, total);
                for (uint32_t mb = mask; mb; mb &= mb - 1, kk++)
                    ring[kk & (kRing - 1)] = (uint16_t)(p0 + kLanes * __builtin_ctz(mb) + q);
                const int32_t kn = k + total;
                // rows r0 .. r1 - 1 are complete (unsigned: the counts are never
                // negative, and the division becomes a shift)
                const int32_t r0 = (int32_t)((uint32_t)k / kRowEntries), r1 = (int32_t)((uint32_t)kn / kRowEntries);
                if (r1 != r0) {
                    flush(r0);
                    if (r1 != r0 + 1) flush(r0 + 1);
                }
                k = kn;
            }
        }
        const int32_t kc = k < cap ? k : cap;
        if (q == 0) {
            count[s] = kc;
            if (k > cap) atomicMax(&err[1], k);
        }
        // the last row: the dummy atom after the entries, so the step kernel
        // evaluates whole chunks without a validity test
        if (k % kRowEntries != 0) {
            const int32_t r = k / kRowEntries, used = k % kRowEntries;
            uint16_t* const row = ring + (r & (kRing / kRowEntries - 1)) * kRowEntries + q * kChunk;
#pragma unroll
            for (int u = 0; u < kChunk; u++)
                if (q * kChunk + u >= used) row[u] = dummy;
            flush(r);
        }
    }
}

// The single box: one segment per block.  Replicas: a fixed grid over the work
// list, each XCD taking a contiguous eighth of it.
template <bool kRep, int kB>
__global__ void __launch_bounds__(kB, 8)
k_nlist_lds(SegGrid g, RepGeo rg, const double* __restrict__ rx, const double* __restrict__ ry,
            const double* __restrict__ rz, double rl2, int32_t cap, uint16_t* __restrict__ list,
            int32_t* __restrict__ count, int32_t* __restrict__ err, int32_t lds_cap, int32_t* __restrict__ hdr) {
    __shared__ Stage S;
    extern __shared__ double dyn[];
    int32_t cx, cy, z0, z1;
    if constexpr (kRep) {
        const int32_t nw = *rg.wcount;
        const int32_t per_x = (int32_t)gridDim.x / kXcd, x = (int32_t)(blockIdx.x % kXcd);
        const int32_t w1 = (int32_t)((int64_t)nw * (x + 1) / kXcd);
        for (int32_t w = (int32_t)((int64_t)nw * x / kXcd) + (int32_t)(blockIdx.x / kXcd); w < w1; w += per_x) {
            const int32_t b = rg.wl[w];
            rep_geometry(rg, b, g, cx, cy, z0, z1);
            nlist_segment<kB>(S, dyn, g, cx, cy, z0, z1, rx, ry, rz, rl2, cap, list, count, err, lds_cap,
                              hdr + (int64_t)b * kHdrWords);
            __syncthreads();   // the staging is reused by the next segment
        }
    } else {
        const int32_t seg = seg_geometry(g, cx, cy, z0, z1);
        if (seg < 0) return;
        nlist_segment<kB>(S, dyn, g, cx, cy, z0, z1, rx, ry, rz, rl2, cap, list, count, err, lds_cap,
                          hdr + (int64_t)seg * kHdrWords);
    }
}

// The lists as caller atom indices (parity exports): entry kk of home slot i
// becomes id[slot of its staged atom] in row id[i] of rows (width ints a
// row); a count above width is reported in *err.  One block per segment.
template <bool kRep>
__global__ void __launch_bounds__(kBlock)
k_nlist_decode(SegGrid g, RepGeo rg, const uint16_t* __restrict__ list, const int32_t* __restrict__ count,
               int32_t nchunk, const int32_t* __restrict__ id, int32_t width, int32_t* __restrict__ rows,
               int32_t* __restrict__ counts, int32_t* __restrict__ err) {
    __shared__ Stage S;
    int32_t cx, cy, z0, z1;
    if constexpr (kRep) {
        if ((int32_t)blockIdx.x >= rg.nblocks) return;
        rep_geometry(rg, blockIdx.x, g, cx, cy, z0, z1);
    } else if (seg_geometry(g, cx, cy, z0, z1) < 0) {
        return;
    }
    stage_header(S, g, cx, cy, z0, z1);
    const int32_t cbase = (cx * g.n[1] + cy) * g.n[2];
    const int32_t hbeg = g.start[cbase + z0], hend = g.start[cbase + z1];
    for (int32_t i = hbeg + (int32_t)threadIdx.x; i < hend; i += kBlock) {
        const uint16_t* la = list + ((int64_t)(i / kListBlock) * nchunk * kListBlock + i % kListBlock) * (kLanes * kChunk);
        const int32_t c = count[i], me = id[i];
        counts[me] = c;
        if (c > width) {
            atomicMax(err, c);
            continue;
        }
        for (int32_t kk = 0; kk < c; kk++) {
            const int32_t w = la[(kk / kRowEntries) * kChunkRow + kk % kRowEntries];
            // an index past the staged atoms (the padding's dummy) inside the count would be a bug: -2
            rows[(int64_t)me * width + kk] = w < S.h.total ? id[staged_slot(S, w)] : -2;
        }
    }
}

struct StepArgs {
    SegGrid g;
    DevState st;            // st.r = positions r(t) (read), Gear state v,a,b,c (read/write)
    double* rn[3];          // r(t+1) out
    const double* r0[3];    // positions at the last list build
    const uint16_t* list;
    const int32_t* count;
    const int32_t* hdr;     // segment headers (kHdrWords each) written by the list build
    int32_t nchunk;         // list chunks per lane (capacity / 16 + the spare)
    double rc2;
    float f12, f6;          // F/r = x^4 (f12 x^3 + f6), x = 1/r^2: 48 eps sig^12, -24 eps sig^6
    double sig2, coeff;
    double mass, scale;
    GearCoeffs gc;
    double hard2;                // list validity: displacement^2 must stay below this
    unsigned long long* flags;   // kSpread x {max disp^2 bits, max step^2 bits, stale count}
    unsigned long long* clear;   // the next step's flag words, zeroed by block 0 (may be null)
    int64_t cs;                  // component stride: component d of a vector at [0] + d * cs
    int32_t n;
    int32_t lds_cap;             // staged atoms the dynamic LDS holds
    RepGeo rg;                   // batched replicas: geometry,
    unsigned long long* dmax;    // the replicas' largest displacement^2 bits of r(t+1),
    int32_t* err;                // and a staging overflow word
    FlagPublish pub;             // the single box's flag hand-over
};

__device__ inline double pick3(int32_t q, double a, double b, double c) { return q == 0 ? a : (q == 1 ? b : c); }

// An 8-byte chunk of the pair list, loaded with the non-temporal hint: the list
// streams through once per step (about 170 MB at 1M atoms), and plain loads
// would push out of the XCD's L2 the positions that the segments of the
// neighbouring columns stage again soon after (md_step 175 -> 168 us at 1M atoms).
// The pointer is cast to the global address space: the step kernel hides
// the look-ahead address from the compiler, which then no longer knows it is
// global and would emit a flat load, whose completion the waits can only
// track together with the LDS reads (so every pair would wait for the list).
__device__ inline uint2 load_list_chunk(const uint2* p) {
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    typedef __attribute__((address_space(1))) const u32x2 global_u32x2;
    const u32x2 t = __builtin_nontemporal_load((global_u32x2*)reinterpret_cast<uintptr_t>(p));
    return make_uint2(t.x, t.y);
}

// 8 waves per SIMD (64 VGPRs) for the mixed body; the FP64 body's divisions
// need more registers, so it runs at 4
template <bool kRep, bool kFp64, bool kFinal, bool kForceOut, int kB>
__global__ void __launch_bounds__(kB, kFp64 ? 4 : 8)
k_step_lds(StepArgs a) {
    __shared__ StepStage S;
    extern __shared__ double dyn[];
    SegGrid g = a.g;
    // the single box's list gate: set when the previous launch's flags say
    // that r(t) has left the list's margin (every wavefront loads the
    // kSpread displacement words, one a lane, and tests them by a ballot)
    unsigned long long gw = 0;
    if constexpr (!kRep) gw = a.pub.prev[3 * (threadIdx.x & (kWave - 1))];
    if (blockIdx.x == 0) {
        if (a.clear)
            for (int32_t t = threadIdx.x; t < 3 * kSpread; t += kB) a.clear[t] = 0;
        if constexpr (!kRep) {
            hand_over(a.pub, 3 * kSpread);
            const bool halt = !a.pub.open && __ballot(gw >= a.pub.gate_bits) != 0;   // every lane votes
            if (threadIdx.x == 0 && halt) {
                // a halted step's own words: the next launch halts, the host sees why
                a.flags[0] = kHaltBits;
                a.flags[2] = kHaltMark;
            }
        }
    }
    int32_t cx, cy, z0, z1, rep = 0, seg = 0;
    if constexpr (kRep) {
        seg = rep_block(a.rg);   // the segment table's entry
        if (seg < 0) return;
        rep = rep_geometry(a.rg, seg, g, cx, cy, z0, z1);
    } else if ((seg = seg_geometry(g, cx, cy, z0, z1)) < 0) {
        return;
    }
    const Pos P{dyn, a.lds_cap};
    unsigned long long* fl = kRep ? nullptr : a.flags + 3 * (blockIdx.x % kSpread);
    // this segment's header from the last list build, loaded with the gate's words
    int32_t hw = 0;
    if (threadIdx.x < kHdrWords) hw = a.hdr[(int64_t)seg * kHdrWords + threadIdx.x];
    if constexpr (!kRep) {
        // the empty asm keeps the header load above the test, so one wait
        // covers both
        asm volatile("" ::"v"(hw));
        // the list does not cover r(t): the step runs as a no-op (workgroup 0
        // has marked its flag words)
        if (!a.pub.open && __ballot(gw >= a.pub.gate_bits) != 0) return;
    }
    if (threadIdx.x < kHdrWords) reinterpret_cast<int32_t*>(&S.h)[threadIdx.x] = hw;
    if (threadIdx.x < 3) S.red[threadIdx.x] = 0;
    __syncthreads();
    const int32_t hbeg = S.h.hbeg, hend = S.h.hend;
    const int32_t q = threadIdx.x & (kLanes - 1);
    // the first pass's list count and first chunk, loaded before the staging so
    // that their latency passes while it runs
    constexpr int32_t cstride = kListBlock * kLanes;   // chunks between a lane's rows
    auto list_of = [&](int32_t i) {
        return reinterpret_cast<const uint2*>(a.list) +
               ((int64_t)(i / kListBlock) * a.nchunk * kListBlock + i % kListBlock) * kLanes + q;
    };
    const int32_t i_first = hbeg + (int32_t)threadIdx.x / kLanes;
    int32_t count_first = 0;
    uint2 chunk_first = make_uint2(0u, 0u);
    if (i_first < hend) {
        count_first = a.count[i_first];
        chunk_first = load_list_chunk(list_of(i_first));
    }
    if (!stage_fill<kB>(S, P, g, cx, cy, a.st.r[0], a.st.r[1], a.st.r[2])) {
        // cannot happen after a successful list build on the same cells: the
        // single box reports it as stale, the replicas through their error word
        if (threadIdx.x == 0) {
            if constexpr (kRep) atomicMax(&a.err[0], S.h.total + 1);
            else atomicAdd(&fl[2], 1ull << 32);
        }
        return;
    }
    double d2max = 0.0, mv2max = 0.0, stale = 0.0;
    for (int32_t base = hbeg; base < hend; base += kB / kLanes) {
        const int32_t i = base + threadIdx.x / kLanes;
        if (i >= hend) continue;              // whole quads drop out together
        // r(t) of the home atom: its staged copy (centre column, never shifted)
        const int32_t self = S.h.self0 + (i - hbeg);
        const double xi = P.x(self), yi = P.y(self), zi = P.z(self);
        double fx = 0, fy = 0, fz = 0;
        const bool first = base == hbeg;
        const int32_t m = first ? count_first : a.count[i];   // the atom's entries, 16 to a row
        // lane q takes its 8-byte chunk of every row; rows are 16 x 4 chunks
        // apart; the next row's chunk is loaded while the current one is
        // evaluated (the list has one spare row, so the look-ahead stays in
        // bounds); the last row is padded with the dummy atom
        const uint2* lp = list_of(i);
        uint2 nx = first ? chunk_first : load_list_chunk(lp);
        for (int32_t k0 = 0; k0 < m; k0 += kRowEntries) {
            const uint2 cu = nx;
            lp += cstride;
            // the empty asm hides that this address is the next iteration's
            // chunk, so the compiler keeps the load here (one chunk ahead)
            // instead of merging it into a load at the top of the next iteration
            uint64_t pa = reinterpret_cast<uint64_t>(lp);
            asm volatile("" : "+v"(pa));
            nx = load_list_chunk(reinterpret_cast<const uint2*>(pa));
            const uint32_t wv[2] = {cu.x, cu.y};
            // straight-line body over the whole chunk: entries past the lane's
            // count are the dummy atom, which fails the cutoff test (a select,
            // so the 1/0 of its force law never reaches the sum)
#pragma unroll
            for (int u = 0; u < kChunk; u++) {
                const int32_t j = (int32_t)((wv[u >> 1] >> (16 * (u & 1))) & 0xffffu);
                const double ddx = xi - P.x(j), ddy = yi - P.y(j), ddz = zi - P.z(j);
                if constexpr (kFp64) {
                    const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                    const double qq = r2 / a.sig2;
                    const double iq4 = 1.0 / (qq * qq * qq * qq);
                    const double iq7 = iq4 / (qq * qq * qq);
                    const double fr = r2 < a.rc2 ? (2 * iq7 - iq4) * a.coeff : 0.0;
                    fx += ddx * fr; fy += ddy * fr; fz += ddz * fr;
                } else {
                    const double r2 = __builtin_fma(ddx, ddx, __builtin_fma(ddy, ddy, ddz * ddz));
                    const float x = __builtin_amdgcn_rcpf((float)r2);
                    const float x2 = x * x;
                    const float x3 = x2 * x;
                    // the cutoff masks x^4 (one 32-bit select) before the product
                    const float x4 = r2 < a.rc2 ? x2 * x2 : 0.0f;
                    const double fr = (double)(x4 * __builtin_fmaf(a.f12, x3, a.f6));
                    fx = __builtin_fma(fr, ddx, fx);
                    fy = __builtin_fma(fr, ddy, fy);
                    fz = __builtin_fma(fr, ddz, fz);
                }
            }
        }
        fx = quad_sum(fx);
        fy = quad_sum(fy);
        fz = quad_sum(fz);
        // lane q < 3 integrates component q
        double d2 = 0.0, mv2 = 0.0, e2 = 0.0;
        if (q < 3) {
            const double fd = pick3(q, fx, fy, fz);
            const double rt = pick3(q, xi, yi, zi);
            const int64_t o = q * a.cs + i;
            const double r0 = a.r0[0][o];
            double r = rt, v = a.st.v[0][o], ac = a.st.a[0][o], b = a.st.b[0][o], c = a.st.c[0][o];
            // corrector (cpu/common/Device.cpp:58-84)
            const double anew = fd / a.mass * a.scale;
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
                const double dd = rp - r0, ds = rp - rt;
                d2 = dd * dd;
                mv2 = ds * ds;
                r = rp;
            }
            a.rn[0][o] = r;
            a.st.v[0][o] = v;
            a.st.a[0][o] = ac;
            a.st.b[0][o] = b;
            a.st.c[0][o] = c;
            if constexpr (kForceOut) a.st.f[0][o] = fd;
            // the force just used is valid only if r(t) was within the list margin
            const double e = rt - r0;
            e2 = e * e;
        }
        d2 = quad_sum(d2);
        mv2 = quad_sum(mv2);
        e2 = quad_sum(e2);
        if (q == 0 && e2 > a.hard2) stale += 1.0;
        d2max = fmax(d2max, d2);
        mv2max = fmax(mv2max, mv2);
    }
    // block reduction (non-negative doubles order like their bits): the
    // wave's largest values by shuffles across its quads (the lanes of a quad
    // hold equal values), one LDS atomic per wave, then one spread global
    // atomic per block (the single box) or one into the replica's
    // displacement accumulator (the exact rebuild decision of replicas.hip;
    // stale atoms cannot occur there).  An atomic issued by the 16 quad
    // leaders of a wave would run as a loop over them, one lane at a time.
    if constexpr (!kFinal) {
        unsigned long long b0 = (unsigned long long)__double_as_longlong(d2max);
        unsigned long long b1 = (unsigned long long)__double_as_longlong(mv2max);
        for (int o = kLanes; o < kWave; o <<= 1) {
            const unsigned long long c0 = __shfl_xor(b0, o, kWave), c1 = __shfl_xor(b1, o, kWave);
            b0 = c0 > b0 ? c0 : b0;
            b1 = c1 > b1 ? c1 : b1;
        }
        if ((threadIdx.x & (kWave - 1)) == 0) {
            atomicMax(&S.red[0], b0);
            atomicMax(&S.red[1], b1);
        }
    }
    if (q == 0 && stale > 0) atomicAdd(&S.red[2], (unsigned long long)stale);
    __syncthreads();
    if (threadIdx.x == 0) {
        if constexpr (kRep) {
            if (!kFinal) atomicMax(&a.dmax[rep], S.red[0]);
        } else {
            if (!kFinal) {
                atomicMax(&fl[0], S.red[0]);
                atomicMax(&fl[1], S.red[1]);
            }
            if (S.red[2] > 0) atomicAdd(&fl[2], S.red[2]);
        }
    }
}

SegGrid make_grid(const GridView& v, int32_t nseg) {
    SegGrid g;
    for (int d = 0; d < 3; d++) g.n[d] = v.g[d];
    g.nseg = nseg;
    g.total = v.g[0] * v.g[1] * nseg;
    g.per_xcd = (g.total + kXcd - 1) / kXcd;
    g.rnseg = 1.0f / (float)nseg;
    g.rny = 1.0f / (float)v.g[1];
    g.small = g.total < (1 << 22) && (int64_t)nseg * v.g[2] < (1 << 22);
    g.L = v.L;
    g.start = v.start;
    return g;
}

template <typename K>
void allow_lds(K kernel) {
    // dynamic LDS above the 64 KiB default, up to the largest staging (the
    // step kernel's FP64 one; the list build's FP32 staging and row rings
    // stay below it)
    static_assert(ring_offset(kLdsAtoms) + 1024 / kLanes * kRing * 2 <= 3 * kLdsAtoms * 8, "list build LDS");
    (void)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              3 * kLdsAtoms * (int)sizeof(double));
}

}  // namespace

int32_t lds_segments_per_column(int32_t nz, double atoms_per_cell, int32_t kmax, int32_t block, double fill) {
    // about fill x block / 4 home atoms per segment (a fraction of one pass of
    // the workgroup's quads), at most kmax cells per segment
    kmax = kmax < kMaxK ? kmax : kMaxK;
    kmax = kmax < 1 ? 1 : kmax;
    const double apc = atoms_per_cell > 0 ? atoms_per_cell : 1.0;
    int32_t nseg = (int32_t)std::ceil(nz * apc / (fill * (block / kLanes)));
    const int32_t lo = (nz + kmax - 1) / kmax;   // segments no longer than kmax cells
    nseg = nseg < lo ? lo : nseg;
    return nseg > nz ? nz : nseg;
}

int32_t lds_capacity() { return kLdsAtoms; }
int64_t list_entries(int32_t n, int32_t cap) {
    const int64_t blocks = ((int64_t)n + kListBlock - 1) / kListBlock;
    return blocks * kListBlock * (cap + kLanes * kChunk);
}
int32_t list_cap_multiple() { return kLanes * kChunk; }
int32_t flag_spread() { return kSpread; }

int32_t seg_header_words() { return kHdrWords; }
int64_t seg_count(const GridView& v, int32_t nseg) { return (int64_t)v.g[0] * v.g[1] * nseg; }

hipError_t launch_nlist_lds(hipStream_t s, const DevState& st, const GridView& v, int32_t nseg, double rl2,
                            int32_t n, int32_t cap, uint16_t* list, int32_t* count, int32_t* err,
                            int32_t lds_atoms, int32_t block, int32_t* hdr) {
    if (n <= 0) return hipSuccess;
    if (cap % (kLanes * kChunk) != 0) return hipErrorInvalidValue;
    if (lds_atoms < 1 || lds_atoms > kLdsAtoms) return hipErrorInvalidValue;
    if (block != 512 && block != 1024) return hipErrorInvalidValue;
    static const bool attr_set = [] {  // once per process, thread-safe
        allow_lds(k_nlist_lds<false, 512>);
        allow_lds(k_nlist_lds<false, 1024>);
        return true;
    }();
    (void)attr_set;
    SegGrid g = make_grid(v, nseg);
    const size_t dyn = ring_offset(lds_atoms) + (size_t)(block / kLanes) * kRing * sizeof(uint16_t);
    if (block == 1024)
        k_nlist_lds<false, 1024><<<g.per_xcd * kXcd, 1024, dyn, s>>>(g, RepGeo{}, st.r[0], st.r[1], st.r[2], rl2,
                                                                     cap, list, count, err, lds_atoms, hdr);
    else
        k_nlist_lds<false, 512><<<g.per_xcd * kXcd, 512, dyn, s>>>(g, RepGeo{}, st.r[0], st.r[1], st.r[2], rl2,
                                                                   cap, list, count, err, lds_atoms, hdr);
    return hipGetLastError();
}

hipError_t launch_step_lds(hipStream_t s, const DevState& st, const GridView& v, int32_t nseg,
                           double* const rn[3], double* const r0[3], const uint16_t* list, const int32_t* count,
                           int32_t cap, const LjTable& t, const MassTable& m, const GearCoeffs& gc, double hard2, int32_t n,
                           int64_t cs, int32_t precision, bool final_step, bool force_out, uint64_t* flags,
                           int32_t lds_atoms, uint64_t* clear_next, FlagPublish pub, int32_t block,
                           const int32_t* hdr) {
    if (n <= 0) return hipSuccess;
    static const bool attr_set = [] {  // once per process, thread-safe
        allow_lds(k_step_lds<false, true, true, true, 512>);
        allow_lds(k_step_lds<false, true, false, false, 512>);
        allow_lds(k_step_lds<false, false, true, true, 512>);
        allow_lds(k_step_lds<false, false, false, false, 512>);
        allow_lds(k_step_lds<false, true, true, true, 1024>);
        allow_lds(k_step_lds<false, true, false, false, 1024>);
        allow_lds(k_step_lds<false, false, true, true, 1024>);
        allow_lds(k_step_lds<false, false, false, false, 1024>);
        return true;
    }();
    (void)attr_set;
    if (lds_atoms < 1 || lds_atoms > kLdsAtoms) return hipErrorInvalidValue;
    if (block != 512 && block != 1024) return hipErrorInvalidValue;
    if (force_out != final_step) return hipErrorInvalidValue;   // the force is written on the last step
    StepArgs a{};
    a.g = make_grid(v, nseg);
    a.st = st;
    for (int d = 0; d < 3; d++) { a.rn[d] = rn[d]; a.r0[d] = r0[d]; }
    a.list = list;
    a.count = count;
    a.hdr = hdr;
    if (cap % (kLanes * kChunk) != 0) return hipErrorInvalidValue;
    a.nchunk = cap / (kLanes * kChunk) + 1;
    a.rc2 = t.rc2;
    // coeff = 24 eps / sig^2: 48 eps sig^12 = 2 coeff sig^14, -24 eps sig^6 = -coeff sig^8
    const double s4 = t.sig2[0] * t.sig2[0] * t.sig2[0] * t.sig2[0];
    a.f12 = (float)(2.0 * t.coeff[0] * s4 * t.sig2[0] * t.sig2[0] * t.sig2[0]);
    a.f6 = (float)(-t.coeff[0] * s4);
    a.sig2 = t.sig2[0];
    a.coeff = t.coeff[0];
    a.mass = m.mass[0];
    a.scale = m.scale;
    a.gc = gc;
    a.hard2 = hard2;
    a.flags = reinterpret_cast<unsigned long long*>(flags);
    a.clear = reinterpret_cast<unsigned long long*>(clear_next);
    a.cs = cs;
    for (int d = 1; d < 3; d++)
        if (st.v[d] != st.v[0] + d * cs || st.r[d] != st.r[0] + d * cs || rn[d] != rn[0] + d * cs ||
            r0[d] != r0[0] + d * cs)
            return hipErrorInvalidValue;
    a.n = n;
    a.lds_cap = lds_atoms;
    if (!pub.prev) return hipErrorInvalidValue;   // the single box's step kernels read their gate
    a.pub = pub;
    const size_t dyn = 3 * (size_t)lds_atoms * sizeof(double);
    dim3 grid(a.g.per_xcd * kXcd);
    const bool fp64 = precision == MDX_PREC_FP64;
#define MDX_STEP(F64, FIN)                                                               \
    do {                                                                                 \
        if (block == 1024) k_step_lds<false, F64, FIN, FIN, 1024><<<grid, 1024, dyn, s>>>(a); \
        else k_step_lds<false, F64, FIN, FIN, 512><<<grid, 512, dyn, s>>>(a);              \
    } while (0)
    if (fp64) {
        if (final_step) MDX_STEP(true, true); else MDX_STEP(true, false);
    } else {
        if (final_step) MDX_STEP(false, true); else MDX_STEP(false, false);
    }
#undef MDX_STEP
    return hipGetLastError();
}

hipError_t launch_nlist_decode(hipStream_t s, const GridView& v, int32_t nseg, const uint16_t* list,
                               const int32_t* count, int32_t cap, const int32_t* id, int32_t width, int32_t* rows,
                               int32_t* counts, int32_t* err) {
    if (cap % (kLanes * kChunk) != 0) return hipErrorInvalidValue;
    SegGrid g = make_grid(v, nseg);
    k_nlist_decode<false><<<g.per_xcd * kXcd, kBlock, 0, s>>>(g, RepGeo{}, list, count, cap / (kLanes * kChunk) + 1,
                                                              id, width, rows, counts, err);
    return hipGetLastError();
}

// ---- batched replicas: one workgroup per z-segment of each staged replica ----

namespace {
RepGeo rep_geo(const RepSegs& rs, const int32_t* wl, const int32_t* wcount) {
    return RepGeo{rs.seg, rs.nblocks, rs.nc, rs.L, rs.cstart, rs.coff, wl, wcount};
}
}  // namespace

hipError_t launch_rep_nlist(hipStream_t s, const RepSegs& rs, const int32_t* wl, const int32_t* wcount,
                            double* const r[3], double rl2, int32_t cap, uint16_t* list, int32_t* count,
                            int32_t* err, int32_t lds_atoms, int32_t* hdr) {
    if (rs.nblocks <= 0) return hipSuccess;
    if (cap % (kLanes * kChunk) != 0) return hipErrorInvalidValue;
    if (lds_atoms < 1 || lds_atoms > kLdsAtoms) return hipErrorInvalidValue;
    static const bool attr_set = (allow_lds(k_nlist_lds<true, kBlock>), true);  // once per process, thread-safe
    (void)attr_set;
    const size_t dyn = ring_offset(lds_atoms) + (size_t)(kBlock / kLanes) * kRing * sizeof(uint16_t);
    // a grid of resident blocks (4 a CU at the full staging), a multiple of 8
    const int32_t all = (rs.nblocks + kXcd - 1) / kXcd * kXcd;
    const int32_t grid = all < 1024 ? all : 1024;
    k_nlist_lds<true, kBlock><<<grid, kBlock, dyn, s>>>(SegGrid{}, rep_geo(rs, wl, wcount), r[0], r[1], r[2], rl2, cap, list,
                                                count, err, lds_atoms, hdr);
    return hipGetLastError();
}

hipError_t launch_rep_step(hipStream_t s, const RepSegs& rs, const DevState& st, double* const rn[3],
                           double* const r0[3], const uint16_t* list, const int32_t* count, int32_t cap, double rc2,
                           double sig2, double coeff, double mass, double scale, const GearCoeffs& gc, int64_t cs,
                           int32_t precision, bool final_step, unsigned long long* dmax, int32_t* err,
                           int32_t lds_atoms, const int32_t* hdr) {
    if (rs.nblocks <= 0) return hipSuccess;
    static const bool attr_set = [] {  // once per process, thread-safe
        allow_lds(k_step_lds<true, true, true, true, kBlock>);
        allow_lds(k_step_lds<true, true, false, false, kBlock>);
        allow_lds(k_step_lds<true, false, true, true, kBlock>);
        allow_lds(k_step_lds<true, false, false, false, kBlock>);
        return true;
    }();
    (void)attr_set;
    if (lds_atoms < 1 || lds_atoms > kLdsAtoms) return hipErrorInvalidValue;
    if (cap % (kLanes * kChunk) != 0) return hipErrorInvalidValue;
    for (int d = 1; d < 3; d++)
        if (st.v[d] != st.v[0] + d * cs || st.r[d] != st.r[0] + d * cs || rn[d] != rn[0] + d * cs ||
            r0[d] != r0[0] + d * cs || st.a[d] != st.a[0] + d * cs || st.b[d] != st.b[0] + d * cs ||
            st.c[d] != st.c[0] + d * cs || st.f[d] != st.f[0] + d * cs)
            return hipErrorInvalidValue;
    StepArgs a{};
    a.st = st;
    for (int d = 0; d < 3; d++) { a.rn[d] = rn[d]; a.r0[d] = r0[d]; }
    a.list = list;
    a.count = count;
    a.hdr = hdr;
    a.nchunk = cap / (kLanes * kChunk) + 1;
    a.rc2 = rc2;
    // coeff = 24 eps / sig^2: 48 eps sig^12 = 2 coeff sig^14, -24 eps sig^6 = -coeff sig^8
    const double s4 = sig2 * sig2 * sig2 * sig2;
    a.f12 = (float)(2.0 * coeff * s4 * sig2 * sig2 * sig2);
    a.f6 = (float)(-coeff * s4);
    a.sig2 = sig2;
    a.coeff = coeff;
    a.mass = mass;
    a.scale = scale;
    a.gc = gc;
    a.hard2 = INFINITY;   // the replicas' rebuild decision is exact
    a.cs = cs;
    a.lds_cap = lds_atoms;
    a.rg = rep_geo(rs, nullptr, nullptr);
    a.dmax = dmax;
    a.err = err;
    const size_t dyn = 3 * (size_t)lds_atoms * sizeof(double);
    const bool fp64 = precision == MDX_PREC_FP64;
    const dim3 grid((rs.nblocks + kXcd - 1) / kXcd * kXcd);   // rep_block: XCD-major order
    // a step: corrector and the next predictor; the last step of a run:
    // corrector only, with the force written for read-back
    if (fp64) {
        if (final_step) k_step_lds<true, true, true, true, kBlock><<<grid, kBlock, dyn, s>>>(a);
        else k_step_lds<true, true, false, false, kBlock><<<grid, kBlock, dyn, s>>>(a);
    } else {
        if (final_step) k_step_lds<true, false, true, true, kBlock><<<grid, kBlock, dyn, s>>>(a);
        else k_step_lds<true, false, false, false, kBlock><<<grid, kBlock, dyn, s>>>(a);
    }
    return hipGetLastError();
}

}  // namespace mdx
