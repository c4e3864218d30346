// system.hip -- host side of the C-ABI (include/mdx.h): system lifetime,
// state transfer, list builds, the device-resident step loop and the
// diagnostics.  All device work of one system runs on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "mdx_internal.h"

using namespace mdx;

namespace {

constexpr int kNe = 2;

struct Grid {
    GridView view{};
    int32_t* count = nullptr;  // ncell (histogram, then fill cursor)
    int32_t* start = nullptr;  // ncell + 1
    int64_t ncell = 0, cap = 0;
    int32_t nlistsize = 0, nlargestsize = 0, largest = 0;
    int32_t next_build = 0;    // walled schedule (cpu/cpumd/Host.cpp:288-290)
    bool built = false;
};

struct EventPair {
    const char* name;
    hipEvent_t a, b;
};

}  // namespace

struct mdx_system {
    mdx_config cfg{};
    int32_t n = 0;
    hipStream_t stream = nullptr;
    int device = 0;
    DevState st{}, alt{};
    std::vector<void*> allocs;
    int32_t* slot_of_id = nullptr;
    int32_t* cell_of = nullptr;      // primary grid cell per slot
    int32_t* cell_of_alt = nullptr;
    int32_t* sorted = nullptr;       // counting-sort output (items)
    int32_t* cell_by_id = nullptr;   // scratch for list emission
    double* rref[3] = {nullptr, nullptr, nullptr};
    Grid lj, mb;
    int32_t* mb_cell_of = nullptr;   // bond grid cell per id (< n_bond)
    int32_t* mb_ids = nullptr;       // bond grid members (ids), cell-sorted
    int32_t* scan_sums = nullptr;
    int64_t scan_cap = 0;
    int32_t* xi_j = nullptr;
    double* xi_v = nullptr;
    int64_t xi_alloc = 0;            // entries allocated (n_bond * cap)
    int32_t xi_cap = 0;
    double* partials = nullptr;
    int32_t nb = 1;
    double* dscal = nullptr;         // [0..7] kinetics, [8..9] sums
    int32_t* dint = nullptr;         // [0] wall hit, [1] bin error, [2] max count, [3] max row,
                                     // [4..6] list-build error words, [8..9] pair-list export
    bool words_zeroed = false;       // dint[1..6] zeroed by this rebuild's cell clear, not yet used
    uint64_t* dbits = nullptr;       // [0] max disp^2, [1] max v^2, [2] max |coord|
    double* e_atom = nullptr;        // 4n scratch: energies (LJ, bond order), then virials (LJ, bond order)
    void* pinned = nullptr;          // 256 bytes
    LjTable ljt{};
    MassTable mt{};
    BondRefParams bp{};
    GearCoeffs gc{};
    double vmax = 0, box_size_list = 0;
    bool have_ready = false;
    bool initialized = false;
    bool predicted = false;          // periodic loop: predictor of the next step applied
    int32_t t = 0;
    int64_t builds = 0;
    double disp_limit2 = 0;          // periodic: rebuild at max displacement^2 >= this
    // periodic fused path
    double* rnext[3] = {nullptr, nullptr, nullptr};  // r(t+1) written by the step kernel
    uint16_t* nlist = nullptr;       // chunked per-lane neighbour list of LDS indices
    int32_t* ncount = nullptr;
    int32_t nl_cap = 0;
    int32_t nseg = 0;                // segments per cell column
    int32_t seg_kmax = 1 << 20;      // cap on cells per segment (lowered if staging overflows)
    int32_t lds_atoms = 0;           // largest segment staging of the last build
    int32_t build_lds = 0;           // staging capacity the list build runs with
    int32_t box_block = 512;         // threads a workgroup of the LDS kernels (512 or 1024)
    int32_t* seg_hdr = nullptr;      // the list build's per-segment headers for the step kernels
    int64_t seg_hdr_cap = 0;         // words allocated
    mdx_shape shape{};               // launch shape (box_block and the segments' fill come from it)
    // periodic Tersoff silicon: compact lists [k][slot] and per-bond derivatives
    TersoffParams tp{};
    int32_t* ts_list = nullptr;
    int32_t* ts_count = nullptr;
    double4* ts_wd = nullptr;        // FP64 Tersoff: de_i/d(x_j - x_i) per list slot, [slot][n]
    int32_t ts_cap = 0;
    bool ts_built = false;
    // learned surrogate (MDX_BOND_SURROGATE) over the same compact lists
    uint16_t* sg_frag = nullptr;     // BF16 MFMA weight fragments
    float* sg_vec = nullptr;         // biases, output weights, feature scaling and range
    float4* sg_w = nullptr;          // dE_i/d(x_j - x_i) per list slot, [slot][n] (x, y, z, 0); mixed Tersoff too
    uint8_t* sg_rev = nullptr;       // slot of i in the list of its partner, [slot][n]
    int32_t* sg_nflag = nullptr;     // atoms evaluated by the Tersoff fallback (cumulative)
    bool sg_set = false;
    int64_t sg_evals = 0;            // force evaluations
    static constexpr int kRing = 3;
    static constexpr int kFlagWords = 3 * 64;  // flag_spread() x {disp^2, step^2, stale}
    uint64_t* dflags = nullptr;      // kRing x kFlagWords
    // pinned: kRing slots of kFlagWords copied from dflags and the sequence
    // number of the step whose words they are (written by the step kernels)
    static constexpr int kHostWords = kFlagWords + 1;
    uint64_t* hflags = nullptr;
    uint64_t pub_seq = 0;            // last sequence number handed out
    hipEvent_t ev[kRing] = {nullptr, nullptr, nullptr};
    bool owns_stream = true;         // false for the members of a sweep (the sweep's stream)
    bool sweep_member = false;       // destroyed with its sweep only
    bool timing = false;
    std::vector<EventPair> pending;
    std::vector<hipEvent_t> free_events;
    std::map<std::string, std::pair<int64_t, double>> times;
    std::string err;
};

namespace {

thread_local std::string g_err;  // errors before a system exists, per calling thread

int fail(mdx_system* s, int code, const std::string& msg) {
    if (s) s->err = msg; else g_err = msg;
    return code;
}

#define HIPCHK(sys, expr)                                                                    \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(sys, MDX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T> int dalloc(mdx_system* s, T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(s, MDX_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    s->allocs.push_back(*p);
    return MDX_OK;
}

// three components of one per-atom vector quantity in one block: component d
// at p[0] + d * component_stride(n), so a kernel can address component d by an
// offset instead of selecting a pointer
int64_t component_stride(int n) { return ((int64_t)n + 63) / 64 * 64; }
int dalloc3(mdx_system* s, double* p[3], int n) {
    const int64_t cs = component_stride(n);
    int rc = dalloc(s, &p[0], 3 * (size_t)cs);
    if (rc) return rc;
    p[1] = p[0] + cs;
    p[2] = p[0] + 2 * cs;
    return MDX_OK;
}

void dfree(mdx_system* s, void* p) {
    if (!p) return;
    auto it = std::find(s->allocs.begin(), s->allocs.end(), p);
    if (it != s->allocs.end()) s->allocs.erase(it);
    hipFree(p);
}

hipEvent_t get_event(mdx_system* s) {
    if (!s->free_events.empty()) {
        hipEvent_t e = s->free_events.back();
        s->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
}

// Launch wrapper: brackets the launch with events while timing is on.
template <class F> hipError_t timed(mdx_system* s, const char* name, F&& launch) {
    if (!s->timing) return launch();
    EventPair p{name, get_event(s), get_event(s)};
    hipEventRecord(p.a, s->stream);
    hipError_t e = launch();
    hipEventRecord(p.b, s->stream);
    s->pending.push_back(p);
    return e;
}

#define LAUNCH(sys, name, expr) HIPCHK(sys, timed(sys, name, [&]() { return (expr); }))

int harvest(mdx_system* s) {
    HIPCHK(s, hipStreamSynchronize(s->stream));
    for (auto& p : s->pending) {
        float ms = 0;
        hipEventElapsedTime(&ms, p.a, p.b);
        auto& slot = s->times[p.name];
        slot.first += 1;
        slot.second += ms;
        s->free_events.push_back(p.a);
        s->free_events.push_back(p.b);
    }
    s->pending.clear();
    return MDX_OK;
}

// ---- parameter tables ----
void setup_tables(mdx_system* s) {
    const mdx_config& c = s->cfg;
    LjTable& t = s->ljt;
    std::memset(&t, 0, sizeof t);
    MassTable& m = s->mt;
    m.scale = c.accel_scale > 0 ? c.accel_scale : 1.0;
    if (c.lj_table == MDX_LJ_REFERENCE) {
        // common/Constants.h:36-43, 74-76; cpu/common/LennardJones.h:63-88
        const double sig[3] = {(double)3.015f, (double)2.835f, (double)2.72f};  // partner N, O, Ne
        const double eps[3] = {(double)0.328e-4f, (double)0.447e-4f, (double)0.391e-4f};
        for (int ti = 0; ti < kTypes; ti++)
            for (int tj = 0; tj < kTypes; tj++) {
                int k = ti * kTypes + tj;
                if (ti > 2 || tj > 2) continue;
                if (ti != kNe && tj != kNe) continue;
                int sel = ti == kNe ? tj : ti;
                double sg = sig[sel], ep = eps[sel];
                t.sig2[k] = sg * sg;
                t.coeff[k] = 24.0f * ep / (sg * sg);
                t.e4[k] = 4 * ep;
                t.shift[k] = 0;
                t.active |= 1u << k;
            }
        const double mass[3] = {(double)14.00674f, (double)15.9994f, (double)20.1797f};
        for (int q = 0; q < kTypes; q++) m.mass[q] = q < 3 ? mass[q] : mass[2];
        // listSettings.rcutsq is a float (common/NeighborList.h:31, cpu/cpumd/Host.cpp:59)
        t.rc2 = (double)(float)std::pow((double)c.lj_rcut, 2);
    } else {
        double sg = c.lj_sigma > 0 ? c.lj_sigma : 1.0, ep = c.lj_eps > 0 ? c.lj_eps : 1.0;
        double rc2 = (double)c.lj_rcut * (double)c.lj_rcut;
        double qc = rc2 / (sg * sg), iq6 = 1.0 / (qc * qc * qc);
        for (int k = 0; k < kTypes * kTypes; k++) {
            t.sig2[k] = sg * sg;
            t.coeff[k] = 24.0 * ep / (sg * sg);
            t.e4[k] = 4 * ep;
            t.shift[k] = c.box == MDX_BOX_PERIODIC ? 4 * ep * iq6 * (iq6 - 1) : 0.0;
            t.active |= 1u << k;
        }
        for (int q = 0; q < kTypes; q++) m.mass[q] = c.mass > 0 ? c.mass : 1.0;
        t.rc2 = rc2;
    }
    for (int k = 0; k < kTypes * kTypes; k++) {
        t.sig2f[k] = (float)t.sig2[k];
        t.inv_sig2f[k] = t.sig2[k] > 0 ? (float)(1.0 / t.sig2[k]) : 0.f;
        t.coeff_f[k] = (float)t.coeff[k];
        t.e4_f[k] = (float)t.e4[k];
    }
    t.rc2f = (float)t.rc2;
    // reference bond order (common/Constants.h:36-61, common/Getters.cpp:24-78):
    // kinds 0 = N-N, 1 = O-O, 2 = N-O
    BondRefParams& b = s->bp;
    const float re[3] = {1.097f, 1.2075f, 1.15f}, de[3] = {941.4e-4f, 493.55309e-4f, 626.734e-4f};
    const float be[3] = {2.71009f, 2.2478f, 2.24068f}, r0v[3] = {3.72f, 3.31f, 3.51f};
    for (int k = 0; k < 3; k++) {
        b.re[k] = re[k]; b.de[k] = de[k]; b.beta[k] = be[k]; b.r0v[k] = r0v[k];
    }
    b.rcut2 = (double)(float)std::pow((double)c.mb_rcut, 2);
    b.pair_cut2 = (double)(225 * 3.31f * 3.31f);
    b.pair_cut = (double)(15 * 3.31f);
    s->gc = gear_coeffs(c.dt);
}

int alloc_state(mdx_system* s, DevState& d) {
    int n = s->n;
    int rc;
    if ((rc = dalloc3(s, d.r, n)) || (rc = dalloc3(s, d.v, n)) || (rc = dalloc3(s, d.a, n)) ||
        (rc = dalloc3(s, d.b, n)) || (rc = dalloc3(s, d.c, n)) || (rc = dalloc3(s, d.f, n)))
        return rc;
    if ((rc = dalloc(s, &d.type, n)) || (rc = dalloc(s, &d.id, n))) return rc;
    return MDX_OK;
}

int ensure_grid(mdx_system* s, Grid& g, int64_t ncell) {
    if (ncell > (int64_t(1) << 27))
        return fail(s, MDX_ERR_GRID, "cell grid of " + std::to_string(ncell) + " cells exceeds 2^27");
    if (ncell > g.cap) {
        dfree(s, g.count);
        dfree(s, g.start);
        int64_t cap = std::max<int64_t>(ncell, 2 * g.cap);
        if (cap > (int64_t(1) << 27)) cap = ncell;
        int rc;
        if ((rc = dalloc(s, &g.count, cap)) || (rc = dalloc(s, &g.start, cap + 1))) return rc;
        g.cap = cap;
    }
    int64_t need = scan_block_sums_needed(ncell + 1);
    if (need > s->scan_cap) {
        dfree(s, s->scan_sums);
        int rc = dalloc(s, &s->scan_sums, need);
        if (rc) return rc;
        s->scan_cap = need;
    }
    g.ncell = ncell;
    g.view.start = g.start;
    return MDX_OK;
}

// grid geometry: walled uses the reference's rule (common/NeighborList.cpp:21-27),
// periodic floor(L/rs) cells of edge L/n
int set_geometry(mdx_system* s, Grid& g, float rs, double box_size_list) {
    GridView& v = g.view;
    if (!(rs > 0)) return fail(s, MDX_ERR_ARG, "list radius rs must be positive");
    if (s->cfg.box == MDX_BOX_PERIODIC) {
        double L = s->cfg.box_length;
        int32_t nc = (int32_t)std::floor(L / (double)rs);
        if (nc < 3) return fail(s, MDX_ERR_ARG, "periodic box must hold at least 3 cells of edge rs per axis");
        v.g[0] = v.g[1] = v.g[2] = nc;
        v.cells = nc;
        v.periodic = 1;
        v.L = L;
        v.h = L / nc;
        v.npos = 0;
    } else {
        int32_t cells = (int32_t)(box_size_list / (double)rs + 0.99f);
        if (cells < 1) cells = 1;
        v.g[0] = v.g[1] = v.g[2] = cells + 1;
        v.cells = cells;
        v.periodic = 0;
        v.L = 0;
        v.h = (double)rs;
        v.npos = box_size_list / 2;
    }
    return ensure_grid(s, g, (int64_t)v.g[0] * v.g[1] * v.g[2]);
}

int read_ints(mdx_system* s, int32_t* host, const int32_t* dev, int count) {
    HIPCHK(s, hipMemcpyAsync(s->pinned, dev, sizeof(int32_t) * count, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    std::memcpy(host, s->pinned, sizeof(int32_t) * count);
    return MDX_OK;
}

// Byte offsets in the pinned buffer of read-backs that a sweep's batched
// rebuild collects from several members before one synchronisation.
constexpr int kPinReady = 128;   // readyList: max v^2 bits, max |coord| bits
constexpr int kPinLj = 160;      // LJ list sizes (3 words)
constexpr int kPinMb = 176;      // bond-order list sizes (3 words)

// reference list sizes: nlistsize = min(27 * largest, n - 1) (NeighborList.cpp:88-90).
// One read-back brings the binning's out-of-box flag (dint[1]; an atom outside
// the grid is counted in cell 0, so the kernels before stay in bounds) with
// the two sizes.  sizes_launch queues the kernels and the copy into the
// pinned buffer at byte offset po; sizes_finish reads it after a
// synchronisation.
int sizes_launch(mdx_system* s, Grid& g, const int32_t* cell_by_item, int32_t items, int po) {
    HIPCHK(s, hipMemsetAsync(s->dint + 2, 0, 2 * sizeof(int32_t), s->stream));
    LAUNCH(s, "list_max_count", launch_max_count(s->stream, g.count, g.ncell, s->dint + 2));
    LAUNCH(s, "list_row_lengths", launch_row_lengths(s->stream, g.view, cell_by_item, items, s->dint + 3));
    HIPCHK(s, hipMemcpyAsync((char*)s->pinned + po, s->dint + 1, 3 * sizeof(int32_t), hipMemcpyDeviceToHost,
                             s->stream));
    return MDX_OK;
}

int sizes_finish(mdx_system* s, Grid& g, int32_t items, int po) {
    int32_t h[3];
    std::memcpy(h, (char*)s->pinned + po, sizeof h);
    if (h[0]) return fail(s, MDX_ERR_ARG, "an atom lies outside the list box (box_size_list too small)");
    g.largest = h[1];
    int64_t nls = 27 * (int64_t)h[1];
    if (nls > items - 1) nls = items - 1;
    if (nls < 0) nls = 0;
    g.nlistsize = (int32_t)nls;
    g.nlargestsize = h[2];
    return MDX_OK;
}

int finish_sizes(mdx_system* s, Grid& g, const int32_t* cell_by_item, int32_t items) {
    int rc = sizes_launch(s, g, cell_by_item, items, 0);
    if (rc) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return sizes_finish(s, g, items, 0);
}

// The sizes words a one-workgroup build wrote (dint[1..3]) into the pinned
// buffer at po, then (po < 0) read at once.
int small_sizes(mdx_system* s, Grid& g, int32_t items, int po) {
    const int at = po < 0 ? 0 : po;
    HIPCHK(s, hipMemcpyAsync((char*)s->pinned + at, s->dint + 1, 3 * sizeof(int32_t), hipMemcpyDeviceToHost,
                             s->stream));
    if (po >= 0) return MDX_OK;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return sizes_finish(s, g, items, 0);
}

// Primary grid: bins every atom, counting-sorts slots into cell order (ascending
// id within a cell) and permutes the whole state into that order.  With
// sizes_po >= 0 the sizes are only queued (sizes_launch at that offset).
int build_primary(mdx_system* s, Grid& g, float rs, double box_size_list, bool sizes, int sizes_po = -1) {
    int rc = set_geometry(s, g, rs, box_size_list);
    if (rc) return rc;
    hipStream_t st = s->stream;
    const int32_t n = s->n;
    if (s->cfg.box == MDX_BOX_WALLED && small_build_fits(n, g.ncell)) {
        // few atoms and cells: the whole build in one workgroup
        SmallBuild b{};
        b.src = s->st;
        b.dst = s->alt;
        b.items = n;
        b.g = g.view;
        b.ncell = g.ncell;
        b.cell_out = s->cell_of;
        b.sorted = s->sorted;
        b.key_of = s->st.id;
        b.cell_of_dst = s->cell_of_alt;
        b.slot_of_id = s->slot_of_id;
        b.words = s->dint + 1;
        LAUNCH(s, "cells_small_build", launch_small_build(st, b));
        std::swap(s->st, s->alt);
        std::swap(s->cell_of, s->cell_of_alt);
        g.built = true;
        s->builds++;
        return sizes ? small_sizes(s, g, n, sizes_po) : MDX_OK;
    }
    // the first clear also zeroes the binning's out-of-grid word (dint[1]) and
    // the list builds' error words (dint[2..6]), in place of their memsets
    LAUNCH(s, "cells_clear", launch_fill_i32(st, g.count, g.ncell, 0, s->dint + 1, 6));
    s->words_zeroed = true;
    LAUNCH(s, "cells_bin", launch_bin(st, s->st, nullptr, n, g.view, s->cell_of, g.count, s->dint + 1));
    LAUNCH(s, "cells_scan", launch_exclusive_scan(st, g.count, g.start, g.ncell, s->scan_sums, s->scan_cap));
    LAUNCH(s, "cells_clear", launch_fill_i32(st, g.count, g.ncell, 0));
    // scatter into scratch (cell_by_id is free until the sizes pass), then
    // order each cell by id into sorted
    LAUNCH(s, "cells_scatter", launch_scatter(st, s->cell_of, g.start, g.count, s->cell_by_id, n));
    LAUNCH(s, "cells_sort", launch_sort_cells(st, g.start, g.ncell, s->cell_by_id, s->sorted, s->st.id, n));
    const bool periodic = s->cfg.box == MDX_BOX_PERIODIC;
    double* none[3] = {nullptr, nullptr, nullptr};
    LAUNCH(s, "cells_permute",
           launch_permute(st, s->st, s->alt, s->sorted, s->cell_of, s->cell_of_alt, s->slot_of_id,
                          periodic ? s->rref : none, n, periodic ? 1 : 0, s->cfg.box_length));
    std::swap(s->st, s->alt);
    std::swap(s->cell_of, s->cell_of_alt);
    // the histogram is needed again for the sizes; rebuild it from start[]
    g.built = true;
    s->builds++;
    if (sizes) {
        // counts = start[c+1] - start[c]; recompute by binning again
        LAUNCH(s, "cells_clear", launch_fill_i32(st, g.count, g.ncell, 0));
        LAUNCH(s, "cells_bin", launch_bin(st, s->st, nullptr, n, g.view, s->cell_by_id, g.count, s->dint + 1));
        return sizes_po >= 0 ? sizes_launch(s, g, s->cell_of, n, sizes_po) : finish_sizes(s, g, s->cell_of, n);
    }
    return MDX_OK;
}

// the xi table of the bond-order force: one entry per candidate within the
// cutoff, at most the row length (after the bond-order list's sizes)
int bond_walled_finish(mdx_system* s) {
    Grid& g = s->mb;
    int32_t cap = std::max(g.nlargestsize, 1);
    int64_t need = (int64_t)s->cfg.n_bond * cap;
    if (need > s->xi_alloc) {
        dfree(s, s->xi_j);
        dfree(s, s->xi_v);
        int rc;
        if ((rc = dalloc(s, &s->xi_j, need)) || (rc = dalloc(s, &s->xi_v, 2 * need))) return rc;
        s->xi_alloc = need;
    }
    s->xi_cap = cap;
    g.built = true;
    return MDX_OK;
}

// Bond-order grid of the walled box: the first n_bond atoms (by id) binned with
// the bond list radius; members are ids, the state is not permuted.  With
// sizes_po >= 0 the sizes are only queued; sizes_finish and bond_walled_finish
// complete the build after a synchronisation.
int build_bond_walled(mdx_system* s, double box_size_list, int sizes_po = -1) {
    Grid& g = s->mb;
    int rc = set_geometry(s, g, s->cfg.mb_rs, box_size_list);
    if (rc) return rc;
    hipStream_t st = s->stream;
    const int32_t nb = s->cfg.n_bond;
    if (small_build_fits(nb, g.ncell)) {
        // few atoms and cells: the whole build in one workgroup
        SmallBuild b{};
        b.src = s->st;
        b.item_slot = s->slot_of_id;
        b.items = nb;
        b.g = g.view;
        b.ncell = g.ncell;
        b.cell_out = s->mb_cell_of;
        b.sorted = s->mb_ids;
        b.words = s->dint + 1;
        LAUNCH(s, "cells_small_build", launch_small_build(st, b));
        if (sizes_po >= 0) return small_sizes(s, g, nb, sizes_po);
        if ((rc = small_sizes(s, g, nb, -1))) return rc;
        return bond_walled_finish(s);
    }
    // the first clear also zeroes the binning's out-of-grid word (dint[1])
    LAUNCH(s, "cells_clear", launch_fill_i32(st, g.count, g.ncell, 0, s->dint + 1, 1));
    LAUNCH(s, "cells_bin", launch_bin(st, s->st, s->slot_of_id, nb, g.view, s->mb_cell_of, g.count, s->dint + 1));
    LAUNCH(s, "cells_scan", launch_exclusive_scan(st, g.count, g.start, g.ncell, s->scan_sums, s->scan_cap));
    LAUNCH(s, "cells_clear", launch_fill_i32(st, g.count, g.ncell, 0));
    LAUNCH(s, "cells_scatter", launch_scatter(st, s->mb_cell_of, g.start, g.count, s->cell_by_id, nb));
    LAUNCH(s, "cells_sort", launch_sort_cells(st, g.start, g.ncell, s->cell_by_id, s->mb_ids, nullptr, nb));
    // rebuild the histogram for the sizes (the out-of-box flag is read with them)
    LAUNCH(s, "cells_clear", launch_fill_i32(st, g.count, g.ncell, 0));
    LAUNCH(s, "cells_bin", launch_bin(st, s->st, s->slot_of_id, nb, g.view, s->cell_by_id, g.count, s->dint + 1));
    if (sizes_po >= 0) return sizes_launch(s, g, s->mb_cell_of, nb, sizes_po);
    if ((rc = finish_sizes(s, g, s->mb_cell_of, nb))) return rc;
    return bond_walled_finish(s);
}

// readyList (cpu/cpumd/Host.cpp:245-271) on the device: the kernel and the
// copy of its two words into the pinned buffer at byte offset po, then (after
// a synchronisation) vmax and boxSizeList from them
int ready_launch(mdx_system* s, int po) {
    LAUNCH(s, "ready_list", launch_ready_list(s->stream, s->st, s->n, s->dbits + 1));
    HIPCHK(s, hipMemcpyAsync((char*)s->pinned + po, s->dbits + 1, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                             s->stream));
    return MDX_OK;
}

void ready_finish(mdx_system* s, int po) {
    double v2, c;
    std::memcpy(&v2, (char*)s->pinned + po, 8);
    std::memcpy(&c, (char*)s->pinned + po + 8, 8);
    s->vmax = std::sqrt(v2);
    s->box_size_list = 2 * c;
    s->have_ready = true;
}

int ready_list(mdx_system* s) {
    int rc = ready_launch(s, 0);
    if (rc) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    ready_finish(s, 0);
    return MDX_OK;
}

// cpu/cpumd/Host.cpp:288-290 (INT32_MAX: never again)
int32_t next_build(float rs, float rc, double vmax, float dt, int32_t t) {
    float maxnlmove = (rs - rc) * 0.5f;
    double q = ((double)maxnlmove / vmax) / std::fabs((double)dt);
    if (!(q < 2147483647.0 - (double)t)) return INT32_MAX;
    return (int32_t)q + t;
}

bool is_tersoff(const mdx_system* s) { return s->cfg.box == MDX_BOX_PERIODIC && s->cfg.use_bond; }
bool is_surrogate(const mdx_system* s) { return is_tersoff(s) && s->cfg.bond_form == MDX_BOND_SURROGATE; }
// plain Tersoff in FP64 keeps FP64 per-slot vectors; every other Tersoff-family path FP32
bool ts_wide(const mdx_system* s) { return !is_surrogate(s) && s->cfg.precision == MDX_PREC_FP64; }

// Tersoff compact lists over the primary grid (cells of edge mb_rs); grows the
// capacity and builds again when a list overflows
int ts_build_list(mdx_system* s) {
    const int n = s->n;
    for (;;) {
        if (!s->ts_list) {
            int rc;
            const size_t e = (size_t)s->ts_cap * n;
            if ((rc = dalloc(s, &s->ts_list, e)) || (rc = dalloc(s, &s->sg_rev, e))) return rc;
            if (ts_wide(s)) {
                if ((rc = dalloc(s, &s->ts_wd, e))) return rc;
            } else if ((rc = dalloc(s, &s->sg_w, e))) {
                return rc;
            }
        }
        if (s->ts_cap > 255) return fail(s, MDX_ERR_CAPACITY, "Tersoff lists hold at most 255 neighbours");
        // the list, then the slot of each atom in its partners' lists (for the
        // force gather); one read-back covers both: dint[4] the longest list
        // when one overflowed, dint[5] a partner list without the atom
        // (on the first attempt the cell clear of this rebuild zeroed them)
        if (!s->words_zeroed) HIPCHK(s, hipMemsetAsync(s->dint + 4, 0, 2 * sizeof(int32_t), s->stream));
        s->words_zeroed = false;
        LAUNCH(s, "ts_list", launch_ts_list(s->stream, s->st, s->cell_of, s->lj.view, (double)s->cfg.mb_rs, n,
                                            s->ts_cap, s->ts_list, s->ts_count, s->dint + 4));
        LAUNCH(s, "sg_rev", launch_sg_rev(s->stream, s->ts_list, s->ts_count, n, s->sg_rev, s->dint + 5));
        int32_t e[2];
        int rc = read_ints(s, e, s->dint + 4, 2);
        if (rc) return rc;
        if (e[0] == 0) {
            if (e[1]) return fail(s, MDX_ERR_STATE, "asymmetric neighbour lists");
            break;
        }
        const int32_t longest = e[0];
        dfree(s, s->ts_list);
        dfree(s, s->ts_wd);
        dfree(s, s->sg_w);
        dfree(s, s->sg_rev);
        s->ts_list = nullptr;
        s->ts_wd = nullptr;
        s->sg_w = nullptr;
        s->sg_rev = nullptr;
        s->ts_cap = longest + longest / 4 + 4;
    }
    s->ts_built = true;
    return MDX_OK;
}

int build_list(mdx_system* s, int32_t which, double box, bool sizes) {
    const mdx_config& c = s->cfg;
    int rc;
    if (c.box == MDX_BOX_PERIODIC) {
        float rs = c.use_lj ? c.lj_rs : c.mb_rs;
        if ((rc = build_primary(s, s->lj, rs, 0.0, sizes))) return rc;
        return is_tersoff(s) ? ts_build_list(s) : MDX_OK;
    }
    if (!s->have_ready && (rc = ready_list(s))) return rc;
    if (which == MDX_LIST_LJ) {
        if (!c.use_lj) return fail(s, MDX_ERR_STATE, "UseLennardJones is off");
        rc = build_primary(s, s->lj, c.lj_rs, box, sizes);
        if (rc) return rc;
        s->lj.next_build = next_build(c.lj_rs, c.lj_rcut, s->vmax, c.dt, s->t);
        return MDX_OK;
    }
    if (!c.use_bond) return fail(s, MDX_ERR_STATE, "UseManyBody is off");
    rc = build_bond_walled(s, box);
    if (rc) return rc;
    s->mb.next_build = next_build(c.mb_rs, c.mb_rcut, s->vmax, c.dt, s->t);
    return MDX_OK;
}

int lj_forces(mdx_system* s, double* e_atom, const int32_t* gate, const DevState* target = nullptr,
              double* w_atom = nullptr) {
    if (!s->lj.built) return fail(s, MDX_ERR_STATE, "Lennard-Jones list not built");
    const DevState& d = target ? *target : s->st;
    // per-atom energies come with per-atom virials (scratch after the energies when not asked for)
    if (e_atom && !w_atom) w_atom = s->e_atom + 2 * (size_t)s->n;
    LAUNCH(s, "lj_forces", launch_lj(s->stream, d, s->cell_of, s->lj.view, s->ljt, s->n,
                                     s->cfg.precision, e_atom, gate, w_atom));
    return MDX_OK;
}

// assign: -1 the reference's rule (add to the LJ force when LJ is on), 1 assign
int bond_forces(mdx_system* s, double* e_atom, const int32_t* gate, const DevState* target = nullptr,
                int assign_rule = -1) {
    const mdx_config& c = s->cfg;
    if (is_surrogate(s)) {
        if (!s->ts_built) return fail(s, MDX_ERR_STATE, "Tersoff list not built");
        if (!s->sg_set) return fail(s, MDX_ERR_STATE, "surrogate network not set (mdx_set_surrogate)");
        const DevState& d = target ? *target : s->st;
        LAUNCH(s, "sg_forces", launch_sg_forces(s->stream, s->st, c.box_length, s->tp, s->ts_list, s->ts_count,
                                                s->sg_rev, s->n, s->sg_frag, s->sg_vec, s->sg_w, e_atom,
                                                s->sg_nflag, nullptr, 0.0, nullptr, d));
        s->sg_evals++;
        return MDX_OK;
    }
    if (is_tersoff(s)) {
        if (!s->ts_built) return fail(s, MDX_ERR_STATE, "Tersoff list not built");
        const DevState& d = target ? *target : s->st;
        LAUNCH(s, "ts_atoms", launch_ts_atoms(s->stream, s->st, c.box_length, s->tp, c.precision, s->ts_list,
                                              s->ts_count, s->n, s->sg_w, s->ts_wd, e_atom));
        LAUNCH(s, "ts_gather", launch_slot_gather(s->stream, s->st, c.box_length, s->ts_list, s->ts_count, s->sg_rev,
                                                  s->n, s->sg_w, s->ts_wd, nullptr, 0.0, nullptr, d));
        return MDX_OK;
    }
    if (c.box == MDX_BOX_PERIODIC || c.bond_form != MDX_BOND_REFERENCE)
        return fail(s, MDX_ERR_ARG, "this bond-order form is not available in this box");
    if (!s->mb.built) return fail(s, MDX_ERR_STATE, "bond-order list not built");
    const DevState& d = target ? *target : s->st;
    const int assign = assign_rule >= 0 ? assign_rule : (c.use_lj ? 0 : 1);  // BondOrderFuncs.h:325-336
    LAUNCH(s, "bond_forces",
           launch_bond_ref(s->stream, d, s->slot_of_id, s->mb_cell_of, s->mb_ids, s->mb.view, s->bp,
                           c.n_bond, s->xi_cap, c.precision, assign, s->xi_j, s->xi_v, e_atom, gate,
                           std::min<int64_t>(27 * (int64_t)s->mb.largest, c.n_bond)));
    return MDX_OK;
}

int forces(mdx_system* s, const int32_t* gate) {
    int rc;
    if (s->cfg.use_lj && (rc = lj_forces(s, nullptr, gate))) return rc;
    if (s->cfg.use_bond && (rc = bond_forces(s, nullptr, gate))) return rc;
    return MDX_OK;
}

double npos_of(const mdx_system* s) {
    return s->cfg.box == MDX_BOX_WALLED ? s->cfg.wall : INFINITY;
}

// walled box: the reference loop body (cpu/cpumd/Host.cpp:141-195)
int run_walled(mdx_system* s, int32_t nsteps) {
    const mdx_config& c = s->cfg;
    int rc;
    if (!s->initialized) {
        if ((c.use_lj && !s->lj.built) || (c.use_bond && !s->mb.built)) {
            if ((rc = ready_list(s))) return rc;
            if (c.use_lj && !s->lj.built && (rc = build_list(s, MDX_LIST_LJ, s->box_size_list, true))) return rc;
            if (c.use_bond && !s->mb.built && (rc = build_list(s, MDX_LIST_BOND, s->box_size_list, true))) return rc;
        }
        if ((rc = forces(s, nullptr))) return rc;
        LAUNCH(s, "accelerations", launch_accelerations(s->stream, s->st, s->n, s->mt, nullptr));
        s->initialized = true;
    }
    for (int32_t k = 0; k < nsteps; k++) {
        bool lj_due = c.use_lj ? s->lj.next_build == s->t : s->t == 0;
        bool mb_due = c.use_bond ? s->mb.next_build == s->t : s->t == 0;
        if ((c.use_lj && lj_due) || (c.use_bond && mb_due)) {
            if ((rc = ready_list(s))) return rc;
            if (c.use_lj && lj_due && (rc = build_list(s, MDX_LIST_LJ, s->box_size_list, true))) return rc;
            if (c.use_bond && mb_due && (rc = build_list(s, MDX_LIST_BOND, s->box_size_list, true))) return rc;
        }
        // the predictor clears the wall-hit word the corrector sets
        LAUNCH(s, "predict", launch_predict(s->stream, s->st, s->n, s->gc, s->dint));
        if ((rc = forces(s, nullptr))) return rc;
        LAUNCH(s, "correct", launch_correct_walled(s->stream, s->st, s->n, s->gc, s->mt, c.wall,
                                                   c.energy_loss, s->dint));
        // wall hit: forces again and a = F/m (Host.cpp:168-178), predicated on the device flag
        if ((rc = forces(s, s->dint))) return rc;
        LAUNCH(s, "accelerations", launch_accelerations(s->stream, s->st, s->n, s->mt, s->dint));
        s->t++;
    }
    return MDX_OK;
}

// Periodic box, cell grid + LDS-index neighbour lists of radius rs.
int rebuild_periodic(mdx_system* s) {
    const mdx_config& c = s->cfg;
    int rc = build_list(s, MDX_LIST_LJ, 0.0, false);
    if (rc) return rc;
    const int n = s->n;
    const double rl2 = (double)c.lj_rs * (double)c.lj_rs;
    const GridView& v = s->lj.view;
    const double per_cell = (double)n / ((double)v.g[0] * v.g[1] * v.g[2]);
    const int32_t cm = list_cap_multiple();
    if (!s->nlist) {
        // first guess from the density: 1.5 x the mean count inside rs, plus slack
        double rho = n / (c.box_length * c.box_length * c.box_length);
        double est = 4.0 / 3.0 * M_PI * std::pow((double)c.lj_rs, 3) * rho;
        int32_t cap = c.max_neighbors > 0 ? c.max_neighbors : (int32_t)(1.5 * est) + 32;
        cap = (cap + cm - 1) / cm * cm;
        if ((rc = dalloc(s, &s->nlist, (size_t)list_entries(n, cap))) || (rc = dalloc(s, &s->ncount, n)))
            return rc;
        s->nl_cap = cap;
    }
    auto round64 = [](double a) { return ((int32_t)std::ceil(a) + 63) / 64 * 64; };
    for (;;) {
        s->nseg = lds_segments_per_column(v.g[2], per_cell, s->seg_kmax, s->box_block, s->shape.fill);
        const int64_t hw = seg_count(v, s->nseg) * seg_header_words();
        if (hw > s->seg_hdr_cap) {
            dfree(s, s->seg_hdr);
            s->seg_hdr = nullptr;
            if ((rc = dalloc(s, &s->seg_hdr, (size_t)hw))) return rc;
            s->seg_hdr_cap = hw;
        }
        if (s->build_lds == 0) {
            // staging estimate: 9 columns of K + 2 cells at 1.3 x the mean occupancy
            const int32_t kcells = (v.g[2] + s->nseg - 1) / s->nseg;
            s->build_lds = std::min(lds_capacity(), round64(9.0 * (kcells + 2) * per_cell * 1.3 + 64));
        }
        // (on the first attempt the cell clear of this rebuild zeroed them)
        if (!s->words_zeroed) HIPCHK(s, hipMemsetAsync(s->dint + 4, 0, 3 * sizeof(int32_t), s->stream));
        s->words_zeroed = false;
        LAUNCH(s, "nlist_build", launch_nlist_lds(s->stream, s->st, v, s->nseg, rl2, n, s->nl_cap, s->nlist,
                                                  s->ncount, s->dint + 4, s->build_lds, s->box_block, s->seg_hdr));
        int32_t e[3];
        if ((rc = read_ints(s, e, s->dint + 4, 3))) return rc;
        s->lds_atoms = std::min(lds_capacity(), round64(e[2]));
        if (e[0] > 0) {
            // a segment's neighbourhood did not fit: more LDS, else shorter segments
            if (s->build_lds < lds_capacity()) {
                s->build_lds = std::min(lds_capacity(), round64(1.1 * e[0]));
                continue;
            }
            int32_t k = (v.g[2] + s->nseg - 1) / s->nseg;
            if (k <= 1) return fail(s, MDX_ERR_CAPACITY, "cell neighbourhood exceeds the LDS staging capacity");
            s->seg_kmax = k - 1;
            continue;
        }
        if (e[1] == 0) break;
        // a neighbour list exceeded its capacity: grow and build again
        dfree(s, s->nlist);
        s->nlist = nullptr;
        int32_t cap = (e[1] + e[1] / 4 + cm - 1) / cm * cm;
        if ((rc = dalloc(s, &s->nlist, (size_t)list_entries(n, cap)))) return rc;
        s->nl_cap = cap;
    }
    return MDX_OK;
}

double bits_to_double(uint64_t b) {
    double d;
    std::memcpy(&d, &b, 8);
    return d;
}

// Periodic box.  Lennard-Jones runs the fused path: one kernel per step
// computes the LJ force of r(t) from the neighbour list, corrects step t and
// predicts step t+1.  The rebuild decision is exact and made on the device:
// the last workgroup of step t's kernel finds the largest displacement of
// r(t+1) since the build and sets the list gate when it reaches the margin
// (rs - rc)/2; step t+1's kernel then runs as a no-op, as does the one queued
// after it.  The host keeps two launches in flight and reads each step's flags
// as they arrive; on a halted step it rebuilds the lists from that step's
// positions and launches it again.  Tersoff silicon runs its two force kernels
// and a corrector / next-predictor kernel per step with the decision on the
// host, one step behind the device: it reads the flags of r(k-1) before
// launching step k and rebuilds when the largest displacement D plus 1.2 x
// the largest single-step move S reaches 0.9 of the margin.  Every force
// kernel re-checks its own positions against the full margin and counts
// violations, which turn into MDX_ERR_LIST_STALE.
int run_periodic(mdx_system* s, int32_t nsteps) {
    const mdx_config& c = s->cfg;
    int rc;
    const int n = s->n;
    const bool ts = is_tersoff(s);
    const double margin = ts ? 0.5 * ((double)c.mb_rs - (s->tp.R + s->tp.D))
                             : 0.5 * ((double)c.lj_rs - (double)c.lj_rcut);
    const double hard2 = margin * margin;
    hipStream_t st = s->stream;
    auto rebuild_lists = [&]() { return ts ? build_list(s, MDX_LIST_BOND, 0.0, false) : rebuild_periodic(s); };
    if (!s->initialized) {
        if ((rc = rebuild_lists())) return rc;
        if ((rc = ts ? bond_forces(s, nullptr, nullptr) : lj_forces(s, nullptr, nullptr))) return rc;  // Host.cpp:132-138
        LAUNCH(s, "accelerations", launch_accelerations(st, s->st, n, s->mt, nullptr));
        s->initialized = true;
    }
    if (nsteps <= 0) return MDX_OK;
    const int R = mdx_system::kRing;
    const int W = mdx_system::kFlagWords;
    const int HW = mdx_system::kHostWords;
    auto ring = [&](int k) { return ((k % R) + R) % R; };
    // the stale count in step k's flags is of that step's own positions
    const int32_t t0 = s->t;
    auto stale_error = [&](int k) {
        return fail(s, MDX_ERR_LIST_STALE, "an atom moved past the neighbour-list margin: step " +
                                               std::to_string((int64_t)t0 + k) + " used a stale list");
    };
    auto slot_flags = [&](int k) { return s->dflags + W * (size_t)ring(k); };
    auto host_flags = [&](int k) { return s->hflags + HW * (size_t)ring(k); };
    // reduce the spread flags: max displacement^2, max step^2, stale count
    auto reduce = [&](const uint64_t* h, double& D, double& S, uint64_t& stale) {
        double d2 = 0, s2 = 0;
        stale = 0;
        for (int q = 0; q < W / 3; q++) {
            d2 = std::max(d2, bits_to_double(h[3 * q]));
            s2 = std::max(s2, bits_to_double(h[3 * q + 1]));
            stale += h[3 * q + 2];
        }
        D = std::sqrt(d2);
        S = std::sqrt(s2);
    };
    // Each step's last kernel (the LJ step kernel; the Tersoff and surrogate
    // gather + Gear kernel) hands the previous step's flag words to their host
    // slot and writes a sequence number after them (FlagPublish); the last
    // step's and the predictor's flags are copied and marked with an event.
    uint64_t expect[mdx_system::kRing] = {0, 0, 0};   // per slot: sequence number awaited (0: event)
    uint64_t gate_bits;
    std::memcpy(&gate_bits, &hard2, 8);
    // step k's hand-over: the words of step k - 1 when the host has yet to read them
    // (open: step k is the first on its list, which needs no gate)
    auto publish = [&](int k, bool to_host, bool open) {
        FlagPublish p{reinterpret_cast<const unsigned long long*>(slot_flags(k - 1)), nullptr, 0,
                      (unsigned long long)gate_bits, open ? 1 : 0};
        if (to_host) {
            expect[ring(k - 1)] = ++s->pub_seq;
            p.host = reinterpret_cast<unsigned long long*>(host_flags(k - 1));
            p.seq = (unsigned long long)s->pub_seq;
        }
        return p;
    };
    // the device slots cleared of the halted launches' marks
    auto clear_slots = [&]() -> int {
        HIPCHK(s, hipMemsetAsync(s->dflags, 0, (size_t)R * W * sizeof(uint64_t), st));
        return MDX_OK;
    };
    auto post = [&](int k) -> int {  // copy step flags to the host and mark them
        HIPCHK(s, hipMemcpyAsync(host_flags(k), slot_flags(k), W * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        HIPCHK(s, hipEventRecord(s->ev[ring(k)], st));
        expect[ring(k)] = 0;
        return MDX_OK;
    };
    auto wait_flags = [&](int k) -> int {  // the host slot of step k holds its words
        const int r = ring(k);
        if (expect[r] == 0) {
            HIPCHK(s, hipEventSynchronize(s->ev[r]));
            return MDX_OK;
        }
        const uint64_t* seq = host_flags(k) + W;   // written by the device; read with acquire
        for (uint32_t spin = 1; __atomic_load_n(seq, __ATOMIC_ACQUIRE) != expect[r]; spin++) {
            if ((spin & 255u) != 0) continue;
            // every 256 polls: a stream that has finished (or failed) ends the wait
            const hipError_t e = hipStreamQuery(st);
            if (e == hipErrorNotReady) continue;
            if (e != hipSuccess) return fail(s, MDX_ERR_HIP, std::string("step: ") + hipGetErrorString(e));
            if (__atomic_load_n(seq, __ATOMIC_ACQUIRE) != expect[r])
                return fail(s, MDX_ERR_HIP, "a step kernel finished without handing over its flags");
        }
        return MDX_OK;
    };
    // flags of r(0) come from the first predictor (index -1)
    // the predictor's slot and step 0's slot start cleared; from then on each
    // step's last kernel clears the next step's slot (the launch before it
    // has read that slot, two steps back in the ring, and copied it)
    HIPCHK(s, hipMemsetAsync(slot_flags(-1), 0, W * sizeof(uint64_t), st));
    HIPCHK(s, hipMemsetAsync(slot_flags(0), 0, W * sizeof(uint64_t), st));
    LAUNCH(s, "predict", launch_predict_track(st, s->st, n, s->gc, s->rref, slot_flags(-1)));
    if ((rc = post(-1))) return rc;
    if (!ts) {
        // r(0) is checked here, exactly; every later step by its kernel's gate
        if ((rc = wait_flags(-1))) return rc;
        double D, S;
        uint64_t stale;
        reduce(host_flags(-1), D, S, stale);
        if (stale != 0) return stale_error(-1);
        // the gate's test, on the same bits: largest displacement^2 >= margin^2
        uint64_t d2_bits = 0;
        for (int q = 0; q < W / 3; q++) d2_bits = std::max<uint64_t>(d2_bits, host_flags(-1)[3 * q]);
        if (d2_bits >= gate_bits && ((rc = rebuild_lists()) || (rc = clear_slots()))) return rc;
        int32_t k = 0;        // next step to launch
        int32_t checked = 0;  // steps before this one have run with a valid list
        int32_t fresh = 0;    // the first step on the current list (r(0) was checked above)
        // with timing on: the record of each launch in flight, so that a
        // halted one is counted as "md_step_halted" instead of "md_step"
        int64_t rec[4] = {-1, -1, -1, -1};
        while (checked < nsteps) {
            if (k < nsteps && k - checked < 2) {
                const bool fin = k + 1 == nsteps;
                LAUNCH(s, "md_step", launch_step_lds(st, s->st, s->lj.view, s->nseg, s->rnext, s->rref, s->nlist,
                                                     s->ncount, s->nl_cap, s->ljt, s->mt, s->gc, hard2, n,
                                                     component_stride(n), c.precision, fin, fin, slot_flags(k),
                                                     s->lds_atoms, slot_flags(k + 1), publish(k, k - 1 >= checked, k == fresh),
                                                     s->box_block, s->seg_hdr));
                rec[k % 4] = s->timing ? (int64_t)s->pending.size() - 1 : -1;
                if (fin && (rc = post(k))) return rc;   // no launch follows to hand these over
                for (int d = 0; d < 3; d++) std::swap(s->st.r[d], s->rnext[d]);
                s->t++;
                k++;
                continue;
            }
            if ((rc = wait_flags(checked))) return rc;
            reduce(host_flags(checked), D, S, stale);
            if (stale >= kHaltMark) {
                // steps checked .. k-1 ran as no-ops: r(checked) is where the
                // first of them left it; undo their position swaps and steps
                const int32_t undone = k - checked;
                for (int32_t j = checked; j < k; j++)
                    if (rec[j % 4] >= 0) s->pending[rec[j % 4]].name = "md_step_halted";
                if (undone % 2 != 0)
                    for (int d = 0; d < 3; d++) std::swap(s->st.r[d], s->rnext[d]);
                s->t -= undone;
                k = checked;
                // the halted launches marked their flag words; the gate opens for the new list
                if ((rc = rebuild_lists()) || (rc = clear_slots())) return rc;
                fresh = checked;
                continue;
            }
            if (stale != 0) return stale_error(checked);
            checked++;
        }
        HIPCHK(s, hipStreamSynchronize(st));
        return MDX_OK;
    }
    int last_build = -1;   // step index whose positions were used for the last build
    // step k hands over step k - 1's words (the predictor's, k = 0, were copied)
    auto handover = [&](int32_t k) { return publish(k, k >= 1, false); };
    for (int32_t k = 0; k < nsteps; k++) {
        // Flags index f describes r(f+1) (f = -1: the predictor's r(0)).  Step k
        // uses r(k); the host reads index k-2, i.e. r(k-1), while launch k-1 runs.
        const int fk = k >= 1 ? k - 2 : -1;
        if ((rc = wait_flags(fk))) return rc;
        double D, S;
        uint64_t stale;
        reduce(host_flags(fk), D, S, stale);
        if (stale != 0) return stale_error(fk);
        // rebuild at 90% of the margin, estimating r(k) as one more step of at most
        // 1.2 S away; the force kernels' own check of the full margin (stale count)
        // catches anything the estimate misses
        const double trigger = 0.9 * margin;
        bool rebuild = false;
        if (k == 0)
            rebuild = D >= trigger;                // exact: these flags are of r(0)
        else if (fk >= last_build)
            rebuild = D + 1.2 * S >= trigger;
        else
            // the flags predate the last build and r(k) is one step from it: only
            // a step as long as the trigger (every step when rs = rc) forces a build
            rebuild = 1.2 * S >= trigger;
        if (rebuild) {
            if ((rc = rebuild_lists())) return rc;
            last_build = k;
        }
        const bool fin = k + 1 == nsteps;
        if (is_surrogate(s)) {
            LAUNCH(s, "sg_eval", launch_sg_eval(st, s->st, c.box_length, s->tp, s->ts_list, s->ts_count, n,
                                                s->sg_frag, s->sg_vec, s->sg_w, nullptr, s->sg_nflag));
            s->sg_evals++;
            const SlotGather g{s->ts_list, s->ts_count, s->sg_rev, s->sg_w, nullptr};
            LAUNCH(s, "gather_gear_step", launch_gather_correct_predict_track(st, s->st, n, s->gc, s->mt, s->rref,
                                                                              slot_flags(k), fin, g, hard2,
                                                                              slot_flags(k + 1), handover(k)));
        } else {
            LAUNCH(s, "ts_atoms", launch_ts_atoms(st, s->st, c.box_length, s->tp, c.precision, s->ts_list, s->ts_count,
                                                  n, s->sg_w, s->ts_wd, nullptr));
            const SlotGather g{s->ts_list, s->ts_count, s->sg_rev, s->sg_w, s->ts_wd};
            LAUNCH(s, "gather_gear_step", launch_gather_correct_predict_track(st, s->st, n, s->gc, s->mt, s->rref,
                                                                              slot_flags(k), fin, g, hard2,
                                                                              slot_flags(k + 1), handover(k)));
        }
        if (fin && (rc = post(k))) return rc;   // no launch follows to hand these over
        s->t++;
    }
    // stale check of the last steps
    HIPCHK(s, hipStreamSynchronize(st));
    for (int k = std::max(0, nsteps - R); k < nsteps; k++) {
        double D, S;
        uint64_t stale;
        reduce(host_flags(k), D, S, stale);
        if (stale != 0) return stale_error(k);
    }
    return MDX_OK;
}

bool valid_config(const mdx_config* c, std::string& why) {
    if (!c) { why = "null config"; return false; }
    if (c->n_atoms <= 0) { why = "n_atoms must be positive"; return false; }
    if (c->n_bond < 0 || c->n_bond > c->n_atoms) { why = "n_bond out of range"; return false; }
    if (c->box != MDX_BOX_WALLED && c->box != MDX_BOX_PERIODIC) { why = "unknown box"; return false; }
    if (c->precision != MDX_PREC_MIXED && c->precision != MDX_PREC_FP64) { why = "unknown precision"; return false; }
    if (c->lj_table != MDX_LJ_REFERENCE && c->lj_table != MDX_LJ_UNIFORM) { why = "unknown LJ table"; return false; }
    if (c->bond_form != MDX_BOND_REFERENCE && c->bond_form != MDX_BOND_TERSOFF && c->bond_form != MDX_BOND_SURROGATE) {
        why = "unknown bond form";
        return false;
    }
    if (c->use_lj && !(c->lj_rcut > 0 && c->lj_rs >= c->lj_rcut)) { why = "need 0 < LennardJonesRCUT <= LennardJonesRS"; return false; }
    if (c->use_bond && !(c->mb_rcut > 0 && c->mb_rs >= c->mb_rcut)) { why = "need 0 < ManyBodyRCUT <= ManyBodyRS"; return false; }
    if (c->box == MDX_BOX_PERIODIC) {
        if (!(c->box_length > 0)) { why = "periodic box needs box_length > 0"; return false; }
        if (c->use_lj && c->use_bond) { why = "periodic box runs one potential at a time"; return false; }
        if (c->use_bond && c->bond_form == MDX_BOND_REFERENCE) { why = "periodic bond order is Tersoff or its surrogate"; return false; }
        if (c->use_bond && c->n_bond != c->n_atoms) { why = "periodic Tersoff needs n_bond == n_atoms"; return false; }
        if (c->use_bond && !(c->mb_rs >= 3.0f)) { why = "Tersoff list radius mb_rs must be >= R + D = 3.0"; return false; }
    } else if (c->use_bond && c->bond_form != MDX_BOND_REFERENCE) {
        why = "walled box bond order is the reference form";
        return false;
    }
    return true;
}

}  // namespace

// A vxcm sweep run as one batch: members walled systems of one configuration,
// all on the sweep's stream; each step kernel is launched once for every
// member (blockIdx.y = member) over a device table of their views.
struct mdx_sweep {
    mdx_config cfg{};
    hipStream_t stream = nullptr;
    std::vector<mdx_system*> sys;
    MemberView* dviews = nullptr;    // device table
    MemberView* hviews = nullptr;    // pinned staging of the table
    std::string err;
};

namespace {

thread_local std::string g_sweep_err;   // errors before a sweep exists

int sweep_fail(mdx_sweep* w, int code, const std::string& msg) {
    if (w) w->err = msg; else g_sweep_err = msg;
    return code;
}

// a member's error, reported by the sweep
int member_fail(mdx_sweep* w, int32_t m, int code) {
    return sweep_fail(w, code, "member " + std::to_string(m) + ": " + w->sys[m]->err);
}

MemberView member_view(const mdx_system* s) {
    MemberView v{};
    v.st = s->st;
    v.cell_of = s->cell_of;
    v.lj = s->lj.view;
    v.slot_of_id = s->slot_of_id;
    v.mb_cell_of = s->mb_cell_of;
    v.mb_ids = s->mb_ids;
    v.mb = s->mb.view;
    v.xi_j = s->xi_j;
    v.xi_v = s->xi_v;
    v.xi_cap = s->xi_cap;
    const int32_t nb = s->cfg.n_bond;
    v.bond_wave = bond_ref_wave(nb, std::min<int64_t>(27 * (int64_t)s->mb.largest, nb), s->xi_cap) ? 1 : 0;
    v.hit = s->dint;
    return v;
}

// Write the views of members ms to the device table.  The pinned staging is
// rewritten only after a synchronisation of the stream (a rebuild's
// read-backs or the end of a run), so the previous copy has completed.
int upload_views(mdx_sweep* w, const std::vector<int32_t>& ms) {
    if (ms.empty()) return MDX_OK;
    for (int32_t m : ms) w->hviews[m] = member_view(w->sys[m]);
    const size_t bytes = sizeof(MemberView) * w->sys.size();
    if (hipMemcpyAsync(w->dviews, w->hviews, bytes, hipMemcpyHostToDevice, w->stream) != hipSuccess)
        return sweep_fail(w, MDX_ERR_HIP, "view table upload failed");
    return MDX_OK;
}

// The rebuilds of every member due at its current step (cpu/cpumd/Host.cpp:143-152),
// with two synchronisations in all instead of three per member: every
// member's readyList, one wait; every list build up to its sizes, one wait;
// then the sizes, the next build steps and the xi tables on the host.
int sweep_rebuild(mdx_sweep* w, const std::vector<int32_t>& due, const std::vector<uint8_t>& lj_due,
                  const std::vector<uint8_t>& mb_due) {
    const mdx_config& c = w->cfg;
    int rc;
    for (int32_t m : due)
        if ((rc = ready_launch(w->sys[m], kPinReady))) return member_fail(w, m, rc);
    if (hipStreamSynchronize(w->stream) != hipSuccess) return sweep_fail(w, MDX_ERR_HIP, "readyList failed");
    for (int32_t m : due) {
        mdx_system* s = w->sys[m];
        ready_finish(s, kPinReady);
        if (lj_due[m] && (rc = build_primary(s, s->lj, c.lj_rs, s->box_size_list, true, kPinLj)))
            return member_fail(w, m, rc);
        if (mb_due[m] && (rc = build_bond_walled(s, s->box_size_list, kPinMb))) return member_fail(w, m, rc);
    }
    if (hipStreamSynchronize(w->stream) != hipSuccess) return sweep_fail(w, MDX_ERR_HIP, "list build failed");
    for (int32_t m : due) {
        mdx_system* s = w->sys[m];
        if (lj_due[m]) {
            if ((rc = sizes_finish(s, s->lj, s->n, kPinLj))) return member_fail(w, m, rc);
            s->lj.next_build = next_build(c.lj_rs, c.lj_rcut, s->vmax, c.dt, s->t);
        }
        if (mb_due[m]) {
            if ((rc = sizes_finish(s, s->mb, c.n_bond, kPinMb)) || (rc = bond_walled_finish(s)))
                return member_fail(w, m, rc);
            s->mb.next_build = next_build(c.mb_rs, c.mb_rcut, s->vmax, c.dt, s->t);
        }
    }
    return upload_views(w, due);
}

// nsteps of the reference loop body (cpu/cpumd/Host.cpp:141-195) for every
// member: rebuilds when due, predict, forces, corrector with walls, and the
// forces and a = F/m again for the members whose corrector hit a wall
int sweep_run(mdx_sweep* w, int32_t nsteps) {
    const mdx_config& c = w->cfg;
    const int32_t R = (int32_t)w->sys.size();
    mdx_system* s0 = w->sys[0];
    int rc;
    std::vector<int32_t> all(R);
    for (int32_t m = 0; m < R; m++) {
        all[m] = m;
        // first call: the member's missing lists, its forces and a = F/m (Host.cpp:132-138)
        if (!w->sys[m]->initialized && (rc = run_walled(w->sys[m], 0))) return member_fail(w, m, rc);
    }
    if ((rc = upload_views(w, all))) return rc;
    const int32_t assign = c.use_lj ? 0 : 1;   // BondOrderFuncs.h:325-336
    std::vector<int32_t> due;
    std::vector<uint8_t> lj_due(R), mb_due(R);
    for (int32_t k = 0; k < nsteps; k++) {
        due.clear();
        for (int32_t m = 0; m < R; m++) {
            const mdx_system* s = w->sys[m];
            lj_due[m] = c.use_lj && s->lj.next_build == s->t;
            mb_due[m] = c.use_bond && s->mb.next_build == s->t;
            if (lj_due[m] || mb_due[m]) due.push_back(m);
        }
        if (!due.empty() && (rc = sweep_rebuild(w, due, lj_due, mb_due))) return rc;
        bool any_wave = false, any_thread = false;
        for (int32_t m = 0; m < R; m++) (w->hviews[m].bond_wave ? any_wave : any_thread) = true;
        for (int pass = 0; pass < 2; pass++) {
            const bool gated = pass == 1;   // the recompute after a wall hit (Host.cpp:168-178)
            if (!gated) {
                LAUNCH(s0, "sweep_predict", launch_sweep_predict(w->stream, w->dviews, R, s0->n, s0->gc));
            }
            if (c.use_lj) {
                LAUNCH(s0, "sweep_lj", launch_sweep_lj(w->stream, w->dviews, R, s0->n, s0->ljt, c.precision, gated));
            }
            if (c.use_bond) {
                LAUNCH(s0, "sweep_bond", launch_sweep_bond(w->stream, w->dviews, R, s0->bp, c.n_bond, c.precision,
                                                           assign, gated, any_wave, any_thread));
            }
            if (!gated) {
                LAUNCH(s0, "sweep_correct", launch_sweep_correct(w->stream, w->dviews, R, s0->n, s0->gc, s0->mt,
                                                                 c.wall, c.energy_loss));
            } else {
                LAUNCH(s0, "sweep_accelerations", launch_sweep_accelerations(w->stream, w->dviews, R, s0->n, s0->mt));
            }
        }
        for (mdx_system* s : w->sys) s->t++;
    }
    if (hipStreamSynchronize(w->stream) != hipSuccess) return sweep_fail(w, MDX_ERR_HIP, "sweep step failed");
    return MDX_OK;
}

}  // namespace

extern "C" {

const char* mdx_last_error(const mdx_system* s) { return s ? s->err.c_str() : g_err.c_str(); }

// mdx_create on a stream of its own (shared == nullptr) or on the caller's
static int create_system(const mdx_config* cfg, hipStream_t shared, mdx_system** out) {
    if (!out) return fail(nullptr, MDX_ERR_ARG, "out is null");
    *out = nullptr;
    std::string why;
    if (!valid_config(cfg, why)) return fail(nullptr, MDX_ERR_ARG, why);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(nullptr, MDX_ERR_HIP, "no HIP device available");
    mdx_system* s = new mdx_system();
    s->cfg = *cfg;
    s->n = cfg->n_atoms;
    hipGetDevice(&s->device);
    int rc = MDX_OK;
    auto bail = [&](int code) {
        mdx_destroy(s);
        return code;
    };
    if (shared) {
        s->stream = shared;
        s->owns_stream = false;
    } else if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        return bail(fail(nullptr, MDX_ERR_HIP, "hipStreamCreate failed"));
    }
    const int n = s->n;
    if ((rc = alloc_state(s, s->st)) || (rc = alloc_state(s, s->alt))) { g_err = s->err; return bail(rc); }
    if ((rc = dalloc(s, &s->slot_of_id, n)) || (rc = dalloc(s, &s->cell_of, n)) ||
        (rc = dalloc(s, &s->cell_of_alt, n)) || (rc = dalloc(s, &s->sorted, n)) ||
        (rc = dalloc(s, &s->cell_by_id, n)) || (rc = dalloc(s, &s->mb_cell_of, std::max(cfg->n_bond, 1))) ||
        (rc = dalloc(s, &s->mb_ids, std::max(cfg->n_bond, 1))) || (rc = dalloc3(s, s->rref, n)) ||
        (rc = dalloc(s, &s->e_atom, 4 * (size_t)n)) || (rc = dalloc(s, &s->dscal, 16)) ||
        (rc = dalloc(s, &s->dint, 16)) || (rc = dalloc(s, &s->dbits, 4))) {
        g_err = s->err;
        return bail(rc);
    }
    // launch shape of the periodic LJ box's LDS kernels: the tuning table's
    s->shape = default_shape();
    if (cfg->box == MDX_BOX_PERIODIC && cfg->use_lj) {
        std::string why;
        if ((rc = tuned_shape("lj", n, &s->shape, why))) {
            g_err = why;
            return bail(rc);
        }
        s->box_block = s->shape.block;
    }
    s->nb = kinetics_blocks(n);
    if ((rc = dalloc(s, &s->partials, 8 * (size_t)s->nb + 16)) ||
        (rc = dalloc(s, &s->dflags, (size_t)mdx_system::kFlagWords * mdx_system::kRing))) {
        g_err = s->err;
        return bail(rc);
    }
    if (cfg->box == MDX_BOX_PERIODIC)
        if ((rc = dalloc3(s, s->rnext, n))) { g_err = s->err; return bail(rc); }
    if (cfg->box == MDX_BOX_PERIODIC && cfg->use_bond) {
        s->tp = tersoff_si_params();
        s->ts_cap = cfg->max_neighbors > 0 ? cfg->max_neighbors : 32;
        if ((rc = dalloc(s, &s->ts_count, n))) { g_err = s->err; return bail(rc); }
        if (cfg->bond_form == MDX_BOND_SURROGATE &&
            ((rc = dalloc(s, &s->sg_frag, sg_frag_u16())) || (rc = dalloc(s, &s->sg_vec, sg_vec_len())) ||
             (rc = dalloc(s, &s->sg_nflag, 1)))) {
            g_err = s->err;
            return bail(rc);
        }
        if (s->sg_nflag) hipMemsetAsync(s->sg_nflag, 0, sizeof(int32_t), s->stream);
    }
    if (hipHostMalloc(&s->pinned, 256, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&s->hflags, sizeof(uint64_t) * mdx_system::kHostWords * mdx_system::kRing,
                      hipHostMallocDefault) != hipSuccess)
        return bail(fail(nullptr, MDX_ERR_ALLOC, "hipHostMalloc failed"));
    std::memset(s->hflags, 0, sizeof(uint64_t) * mdx_system::kHostWords * mdx_system::kRing);
    for (int k = 0; k < mdx_system::kRing; k++)
        if (hipEventCreateWithFlags(&s->ev[k], hipEventDisableTiming) != hipSuccess)
            return bail(fail(nullptr, MDX_ERR_HIP, "hipEventCreate failed"));
    setup_tables(s);
    hipMemsetAsync(s->dint, 0, 8 * sizeof(int32_t), s->stream);
    hipMemsetAsync(s->dbits, 0, 4 * sizeof(uint64_t), s->stream);
    if (hipStreamSynchronize(s->stream) != hipSuccess) return bail(fail(nullptr, MDX_ERR_HIP, "init failed"));
    *out = s;
    return MDX_OK;
}

int mdx_create(const mdx_config* cfg, mdx_system** out) { return create_system(cfg, nullptr, out); }

static void destroy_system(mdx_system* s) {
    if (s->stream) hipStreamSynchronize(s->stream);
    for (auto& p : s->pending) { hipEventDestroy(p.a); hipEventDestroy(p.b); }
    for (auto e : s->free_events) hipEventDestroy(e);
    for (void* p : s->allocs) hipFree(p);
    for (auto e : s->ev)
        if (e) hipEventDestroy(e);
    if (s->pinned) hipHostFree(s->pinned);
    if (s->hflags) hipHostFree(s->hflags);
    if (s->stream && s->owns_stream) hipStreamDestroy(s->stream);
    delete s;
}

int mdx_destroy(mdx_system* s) {
    if (!s) return MDX_OK;
    if (s->sweep_member) return fail(s, MDX_ERR_STATE, "a sweep member is destroyed with its sweep (mdx_sweep_destroy)");
    destroy_system(s);
    return MDX_OK;
}

int mdx_set_state(mdx_system* s, const double* pos4, const double* vel3, const double* acc3,
                  const double* b3, const double* c3) {
    if (!s || !pos4 || !vel3) return fail(s, MDX_ERR_ARG, "pos4 and vel3 are required");
    const int n = s->n;
    std::vector<double> buf(3 * (size_t)n);
    std::vector<int32_t> type(n), id(n);
    for (int i = 0; i < n; i++) {
        double w = pos4[4 * i + 3];
        int t = (int)w;
        if (t < 0 || t >= kTypes) return fail(s, MDX_ERR_ARG, "species code out of range at atom " + std::to_string(i));
        if (s->cfg.use_bond && s->cfg.box == MDX_BOX_WALLED && i < s->cfg.n_bond && t == kNe)
            return fail(s, MDX_ERR_ARG, "bond-order atom " + std::to_string(i) + " is Ne");
        type[i] = t;
        id[i] = i;
    }
    hipStream_t st = s->stream;
    // stage in caller order into alt, then scatter into slot order
    auto put = [&](double* const dst[3], const double* src, int stride) -> int {
        for (int d = 0; d < 3; d++) {
            for (int i = 0; i < n; i++) buf[(size_t)d * n + i] = src ? src[(size_t)stride * i + d] : 0.0;
            HIPCHK(s, hipMemcpyAsync(dst[d], buf.data() + (size_t)d * n, sizeof(double) * n, hipMemcpyHostToDevice, st));
        }
        HIPCHK(s, hipStreamSynchronize(st));
        return MDX_OK;
    };
    int rc;
    if ((rc = put(s->alt.r, pos4, 4)) || (rc = put(s->alt.v, vel3, 3)) || (rc = put(s->alt.a, acc3, 3)) ||
        (rc = put(s->alt.b, b3, 3)) || (rc = put(s->alt.c, c3, 3)) || (rc = put(s->alt.f, nullptr, 3)))
        return rc;
    HIPCHK(s, hipMemcpyAsync(s->alt.type, type.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
    HIPCHK(s, hipMemcpyAsync(s->alt.id, id.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
    if (!s->lj.built) {
        // identity slots
        HIPCHK(s, hipMemcpyAsync(s->slot_of_id, id.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
    }
    LAUNCH(s, "state_in", launch_reorder(st, s->alt, s->st, s->slot_of_id, n, 1));
    HIPCHK(s, hipStreamSynchronize(st));
    s->have_ready = false;
    s->initialized = false;
    s->t = 0;
    return MDX_OK;
}

int mdx_get_state(mdx_system* s, double* pos4, double* vel3, double* acc3, double* b3, double* c3,
                  double* force3) {
    if (!s) return fail(s, MDX_ERR_ARG, "null system");
    const int n = s->n;
    hipStream_t st = s->stream;
    LAUNCH(s, "state_out", launch_reorder(st, s->st, s->alt, s->slot_of_id, n, 0));
    std::vector<double> buf((size_t)n);
    auto get = [&](double* const src[3], double* dst, int stride) -> int {
        if (!dst) return MDX_OK;
        for (int d = 0; d < 3; d++) {
            HIPCHK(s, hipMemcpyAsync(buf.data(), src[d], sizeof(double) * n, hipMemcpyDeviceToHost, st));
            HIPCHK(s, hipStreamSynchronize(st));
            for (int i = 0; i < n; i++) dst[(size_t)stride * i + d] = buf[i];
        }
        return MDX_OK;
    };
    int rc;
    if ((rc = get(s->alt.r, pos4, 4)) || (rc = get(s->alt.v, vel3, 3)) || (rc = get(s->alt.a, acc3, 3)) ||
        (rc = get(s->alt.b, b3, 3)) || (rc = get(s->alt.c, c3, 3)) || (rc = get(s->alt.f, force3, 3)))
        return rc;
    if (pos4) {
        std::vector<int32_t> type(n);
        HIPCHK(s, hipMemcpy(type.data(), s->alt.type, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; i++) pos4[4 * i + 3] = (double)type[i];
    }
    return MDX_OK;
}

int mdx_set_forces(mdx_system* s, const double* force3) {
    if (!s || !force3) return fail(s, MDX_ERR_ARG, "force3 is required");
    const int n = s->n;
    hipStream_t st = s->stream;
    std::vector<double> buf(3 * (size_t)n);
    for (int d = 0; d < 3; d++)
        for (int i = 0; i < n; i++) buf[(size_t)d * n + i] = force3[3 * (size_t)i + d];
    for (int d = 0; d < 3; d++)
        HIPCHK(s, hipMemcpyAsync(s->alt.f[d], buf.data() + (size_t)d * n, sizeof(double) * n, hipMemcpyHostToDevice,
                                 st));
    LAUNCH(s, "forces_in", launch_scatter3(st, s->alt.f, s->slot_of_id, s->st.f, n));
    HIPCHK(s, hipStreamSynchronize(st));
    return MDX_OK;
}

int mdx_set_surrogate(mdx_system* s, const float* params, int64_t len) {
    if (!s || !params) return fail(s, MDX_ERR_ARG, "params is required");
    if (s->cfg.bond_form != MDX_BOND_SURROGATE || !is_tersoff(s))
        return fail(s, MDX_ERR_STATE, "the system's bond form is not MDX_BOND_SURROGATE");
    if (len != MDX_SG_NPARAM) return fail(s, MDX_ERR_ARG, "parameter blob must hold MDX_SG_NPARAM floats");
    for (int64_t k = 0; k < len; k++)
        if (!std::isfinite(params[k]) && !(k >= MDX_SG_NTRAIN + 2 * MDX_SG_NFEAT))   // lo / hi may be infinite
            return fail(s, MDX_ERR_ARG, "non-finite network parameter " + std::to_string(k));
    std::vector<uint16_t> frag(sg_frag_u16());
    std::vector<float> vec(sg_vec_len());
    sg_pack(params, frag.data(), vec.data());
    HIPCHK(s, hipMemcpyAsync(s->sg_frag, frag.data(), sizeof(uint16_t) * frag.size(), hipMemcpyHostToDevice, s->stream));
    HIPCHK(s, hipMemcpyAsync(s->sg_vec, vec.data(), sizeof(float) * vec.size(), hipMemcpyHostToDevice, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    s->sg_set = true;
    return MDX_OK;
}

int mdx_surrogate_counts(mdx_system* s, int64_t* out2) {
    if (!s || !out2) return fail(s, MDX_ERR_ARG, "null argument");
    if (!is_surrogate(s)) return fail(s, MDX_ERR_STATE, "the system's bond form is not MDX_BOND_SURROGATE");
    int32_t nf;
    int rc = read_ints(s, &nf, s->sg_nflag, 1);
    if (rc) return rc;
    out2[0] = nf;
    out2[1] = s->sg_evals * (int64_t)s->n;
    return MDX_OK;
}

int mdx_ready_list(mdx_system* s, double* vmax, double* box_size_list) {
    if (!s) return fail(s, MDX_ERR_ARG, "null system");
    int rc = ready_list(s);
    if (rc) return rc;
    if (vmax) *vmax = s->vmax;
    if (box_size_list) *box_size_list = s->box_size_list;
    return MDX_OK;
}

int mdx_build_list(mdx_system* s, int32_t which, double box_size_list, int32_t* nlistsize,
                   int32_t* nlargestsize) {
    if (!s) return fail(s, MDX_ERR_ARG, "null system");
    if (which != MDX_LIST_LJ && which != MDX_LIST_BOND) return fail(s, MDX_ERR_ARG, "unknown list");
    int rc = build_list(s, which, box_size_list, true);
    if (rc) return rc;
    Grid& g = (which == MDX_LIST_LJ || s->cfg.box == MDX_BOX_PERIODIC) ? s->lj : s->mb;
    if (nlistsize) *nlistsize = g.nlistsize;
    if (nlargestsize) *nlargestsize = g.nlargestsize;
    return MDX_OK;
}

int mdx_copy_neighbor_list(mdx_system* s, int32_t which, int32_t* out, int64_t capacity) {
    if (!s || !out) return fail(s, MDX_ERR_ARG, "null argument");
    const bool primary = which == MDX_LIST_LJ || s->cfg.box == MDX_BOX_PERIODIC;
    Grid& g = primary ? s->lj : s->mb;
    if (!g.built) return fail(s, MDX_ERR_STATE, "list not built");
    const int32_t items = primary ? s->n : s->cfg.n_bond;
    const int64_t need = (int64_t)items * g.nlistsize;
    if (capacity < need) return fail(s, MDX_ERR_CAPACITY, "output capacity below n * nlistsize");
    if (need == 0) return MDX_OK;
    hipStream_t st = s->stream;
    const int32_t* cell_by_item;
    const int32_t* members;
    if (primary) {
        LAUNCH(s, "list_gather", launch_gather_i32(st, s->cell_of, s->slot_of_id, s->cell_by_id, items));
        cell_by_item = s->cell_by_id;
        members = s->st.id;
    } else {
        cell_by_item = s->mb_cell_of;
        members = s->mb_ids;
    }
    int32_t* dout;
    HIPCHK(s, hipMalloc((void**)&dout, sizeof(int32_t) * need));
    hipError_t e = launch_emit_list(st, g.view, cell_by_item, members, items, g.nlistsize, dout);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, sizeof(int32_t) * need, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hipFree(dout);
    if (e != hipSuccess) return fail(s, MDX_ERR_HIP, std::string("emit list: ") + hipGetErrorString(e));
    return MDX_OK;
}

int mdx_copy_pair_list(mdx_system* s, int32_t* out, int64_t capacity, int32_t* counts, int32_t* width,
                       double* build_pos3) {
    if (!s || !width) return fail(s, MDX_ERR_ARG, "null argument");
    if (s->cfg.box != MDX_BOX_PERIODIC)
        return fail(s, MDX_ERR_ARG, "the walled box's kernels read the candidate rows (mdx_copy_neighbor_list)");
    const bool ts = is_tersoff(s);
    if (ts ? !s->ts_built : !(s->nlist && s->initialized))
        return fail(s, MDX_ERR_STATE, "no pair list built (mdx_run builds it)");
    const int32_t n = s->n;
    hipStream_t st = s->stream;
    const int32_t* cnt = ts ? s->ts_count : s->ncount;
    HIPCHK(s, hipMemsetAsync(s->dint + 8, 0, 2 * sizeof(int32_t), st));
    LAUNCH(s, "list_max_count", launch_max_count(st, cnt, n, s->dint + 8));
    int32_t w;
    int rc = read_ints(s, &w, s->dint + 8, 1);
    if (rc) return rc;
    *width = w;
    if (build_pos3) {
        // the list-build positions, slot order -> caller order (alt.f is scratch here)
        LAUNCH(s, "list_positions", launch_scatter3(st, s->rref, s->st.id, s->alt.f, n));
        std::vector<double> buf((size_t)n);
        for (int d = 0; d < 3; d++) {
            HIPCHK(s, hipMemcpyAsync(buf.data(), s->alt.f[d], sizeof(double) * n, hipMemcpyDeviceToHost, st));
            HIPCHK(s, hipStreamSynchronize(st));
            for (int i = 0; i < n; i++) build_pos3[3 * (size_t)i + d] = buf[i];
        }
    }
    if (!out) return MDX_OK;
    if (!counts) return fail(s, MDX_ERR_ARG, "counts is required with out");
    const int64_t need = (int64_t)n * w;
    if (capacity < need) return fail(s, MDX_ERR_CAPACITY, "output capacity below n * width");
    int32_t* drows = nullptr;
    int32_t* dcounts = nullptr;
    HIPCHK(s, hipMalloc((void**)&drows, sizeof(int32_t) * std::max<int64_t>(need, 1)));
    hipError_t e = hipMalloc((void**)&dcounts, sizeof(int32_t) * n);
    if (e == hipSuccess) e = hipMemsetAsync(drows, 0xff, sizeof(int32_t) * std::max<int64_t>(need, 1), st);
    if (e == hipSuccess)
        e = ts ? launch_compact_decode(st, s->ts_list, s->ts_count, n, s->st.id, w, drows, dcounts, s->dint + 9)
               : launch_nlist_decode(st, s->lj.view, s->nseg, s->nlist, s->ncount, s->nl_cap, s->st.id, w, drows,
                                     dcounts, s->dint + 9);
    if (e == hipSuccess && need > 0)
        e = hipMemcpyAsync(out, drows, sizeof(int32_t) * need, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(counts, dcounts, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hipFree(drows);
    hipFree(dcounts);
    if (e != hipSuccess) return fail(s, MDX_ERR_HIP, std::string("pair list: ") + hipGetErrorString(e));
    int32_t over;
    if ((rc = read_ints(s, &over, s->dint + 9, 1))) return rc;
    if (over) return fail(s, MDX_ERR_STATE, "a list count changed during the copy");
    return MDX_OK;
}

int mdx_predict(mdx_system* s) {
    if (!s) return fail(s, MDX_ERR_ARG, "null system");
    LAUNCH(s, "predict", launch_predict(s->stream, s->st, s->n, s->gc));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return MDX_OK;
}

int mdx_lj_forces(mdx_system* s) {
    if (!s) return fail(s, MDX_ERR_ARG, "null system");
    if (!s->cfg.use_lj) return fail(s, MDX_ERR_STATE, "UseLennardJones is off");
    int rc = lj_forces(s, nullptr, nullptr);
    if (rc) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return MDX_OK;
}

int mdx_bond_forces(mdx_system* s) {
    if (!s) return fail(s, MDX_ERR_ARG, "null system");
    if (!s->cfg.use_bond) return fail(s, MDX_ERR_STATE, "UseManyBody is off");
    int rc = bond_forces(s, nullptr, nullptr);
    if (rc) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return MDX_OK;
}

int mdx_correct(mdx_system* s, int32_t* wall_hit) {
    if (!s) return fail(s, MDX_ERR_ARG, "null system");
    HIPCHK(s, hipMemsetAsync(s->dint, 0, sizeof(int32_t), s->stream));
    LAUNCH(s, "correct", launch_correct_walled(s->stream, s->st, s->n, s->gc, s->mt, npos_of(s),
                                               s->cfg.energy_loss, s->dint));
    int32_t h;
    int rc = read_ints(s, &h, s->dint, 1);
    if (rc) return rc;
    if (wall_hit) *wall_hit = h;
    return MDX_OK;
}

int mdx_accelerations(mdx_system* s) {
    if (!s) return fail(s, MDX_ERR_ARG, "null system");
    LAUNCH(s, "accelerations", launch_accelerations(s->stream, s->st, s->n, s->mt, nullptr));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return MDX_OK;
}

int mdx_run(mdx_system* s, int32_t nsteps) {
    if (!s) return fail(s, MDX_ERR_ARG, "null system");
    if (nsteps < 0) return fail(s, MDX_ERR_ARG, "nsteps must be >= 0");
    int rc = s->cfg.box == MDX_BOX_PERIODIC ? run_periodic(s, nsteps) : run_walled(s, nsteps);
    if (rc) return rc;
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return MDX_OK;
}

int mdx_build_count(mdx_system* s, int64_t* builds) {
    if (!s || !builds) return fail(s, MDX_ERR_ARG, "null argument");
    *builds = s->builds;
    return MDX_OK;
}

int mdx_kinetics(mdx_system* s, double kboltz, double* out8) {
    if (!s || !out8) return fail(s, MDX_ERR_ARG, "null argument");
    LAUNCH(s, "kinetics", launch_kinetics(s->stream, s->st, s->n, s->mt, s->partials, s->nb, s->dscal, kboltz));
    HIPCHK(s, hipMemcpyAsync(s->pinned, s->dscal, 8 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    std::memcpy(out8, s->pinned, 8 * sizeof(double));
    return MDX_OK;
}

int mdx_potential(mdx_system* s, double* out2) {
    if (!s || !out2) return fail(s, MDX_ERR_ARG, "null argument");
    out2[0] = out2[1] = 0;
    // forces of this evaluation go to scratch so the stored force is untouched
    int rc;
    hipStream_t st = s->stream;
    // (a build permutes the state, so it must precede taking the state pointers)
    if (s->cfg.box == MDX_BOX_PERIODIC && !s->lj.built && (rc = build_list(s, MDX_LIST_LJ, 0.0, false)))
        return rc;
    DevState scratch = s->st;
    for (int d = 0; d < 3; d++) scratch.f[d] = s->alt.f[d];
    if (s->cfg.use_lj) {
        if ((rc = lj_forces(s, s->e_atom, nullptr, &scratch))) return rc;
        LAUNCH(s, "energy_sum", launch_sum(st, s->e_atom, s->n, s->partials, s->nb, s->dscal + 8));
    }
    if (s->cfg.use_bond) {
        if ((rc = bond_forces(s, s->e_atom + s->n, nullptr, &scratch))) return rc;
        LAUNCH(s, "energy_sum", launch_sum(st, s->e_atom + s->n, s->cfg.n_bond, s->partials, s->nb, s->dscal + 9));
    }
    HIPCHK(s, hipMemcpyAsync(s->pinned, s->dscal + 8, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(s, hipStreamSynchronize(st));
    double v[2];
    std::memcpy(v, s->pinned, sizeof v);
    if (s->cfg.use_lj) out2[0] = v[0];
    if (s->cfg.use_bond) out2[1] = v[1];
    return MDX_OK;
}

int mdx_potential_virial(mdx_system* s, double* out4) {
    if (!s || !out4) return fail(s, MDX_ERR_ARG, "null argument");
    for (int k = 0; k < 4; k++) out4[k] = 0;
    int rc;
    hipStream_t st = s->stream;
    const int32_t n = s->n;
    // (a build permutes the state, so it must precede taking the state pointers)
    if (s->cfg.box == MDX_BOX_PERIODIC && !s->lj.built && (rc = build_list(s, MDX_LIST_LJ, 0.0, false)))
        return rc;
    double* e_lj = s->e_atom;
    double* e_b = s->e_atom + n;
    double* w_lj = s->e_atom + 2 * (size_t)n;
    double* w_b = s->e_atom + 3 * (size_t)n;
    DevState scratch = s->st;
    for (int d = 0; d < 3; d++) scratch.f[d] = s->alt.f[d];
    if (s->cfg.use_lj) {
        if ((rc = lj_forces(s, e_lj, nullptr, &scratch, w_lj))) return rc;
        LAUNCH(s, "energy_sum", launch_sum(st, e_lj, n, s->partials, s->nb, s->dscal + 8));
        LAUNCH(s, "energy_sum", launch_sum(st, w_lj, n, s->partials, s->nb, s->dscal + 10));
    }
    if (s->cfg.use_bond) {
        if (is_tersoff(s)) {
            // the force evaluation leaves its per-slot gradients behind
            if ((rc = bond_forces(s, e_b, nullptr, &scratch))) return rc;
            LAUNCH(s, "slot_virial", launch_slot_virial(st, s->st, s->cfg.box_length, s->ts_list, s->ts_count, n,
                                                        s->sg_w, ts_wide(s) ? s->ts_wd : nullptr, w_b));
            LAUNCH(s, "energy_sum", launch_sum(st, w_b, n, s->partials, s->nb, s->dscal + 11));
        } else {
            // the walled box is isolated: the bond-order virial is sum r_i . F_i
            // over the bond-order atoms, from their bond forces alone
            DevState bonly = s->st;
            for (int d = 0; d < 3; d++) bonly.f[d] = s->alt.a[d];
            if ((rc = bond_forces(s, e_b, nullptr, &bonly, 1))) return rc;
            LAUNCH(s, "rf_virial", launch_rf_virial(st, s->st, bonly, s->slot_of_id, s->cfg.n_bond, w_b));
            LAUNCH(s, "energy_sum", launch_sum(st, w_b, s->cfg.n_bond, s->partials, s->nb, s->dscal + 11));
        }
        LAUNCH(s, "energy_sum", launch_sum(st, e_b, s->cfg.n_bond, s->partials, s->nb, s->dscal + 9));
    }
    HIPCHK(s, hipMemcpyAsync(s->pinned, s->dscal + 8, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(s, hipStreamSynchronize(st));
    double v[4];
    std::memcpy(v, s->pinned, sizeof v);
    if (s->cfg.use_lj) { out4[0] = v[0]; out4[2] = v[2]; }
    if (s->cfg.use_bond) { out4[1] = v[1]; out4[3] = v[3]; }
    return MDX_OK;
}

int mdx_potential_allpairs(mdx_system* s, double* out2) {
    if (!s || !out2) return fail(s, MDX_ERR_ARG, "null argument");
    if (s->cfg.box != MDX_BOX_WALLED) return fail(s, MDX_ERR_ARG, "all-pairs energy is defined for the walled box");
    hipStream_t st = s->stream;
    LAUNCH(s, "allpairs", launch_allpairs(st, s->st, s->n, s->cfg.n_bond, s->ljt, s->bp, s->e_atom));
    LAUNCH(s, "energy_sum", launch_sum(st, s->e_atom, s->n, s->partials, s->nb, s->dscal + 8));
    LAUNCH(s, "energy_sum", launch_sum(st, s->e_atom + s->n, s->n, s->partials, s->nb, s->dscal + 9));
    HIPCHK(s, hipMemcpyAsync(s->pinned, s->dscal + 8, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(s, hipStreamSynchronize(st));
    std::memcpy(out2, s->pinned, 2 * sizeof(double));
    return MDX_OK;
}

int mdx_set_timing(mdx_system* s, int32_t enabled) {
    if (!s) return fail(s, MDX_ERR_ARG, "null system");
    int rc = harvest(s);
    if (rc) return rc;
    s->times.clear();
    s->timing = enabled != 0;
    return MDX_OK;
}

int mdx_kernel_times(mdx_system* s, char* names, int64_t* counts, double* total_ms, int32_t cap,
                     int32_t* n) {
    if (!s || !n) return fail(s, MDX_ERR_ARG, "null argument");
    int rc = harvest(s);
    if (rc) return rc;
    int32_t k = 0;
    for (auto& kv : s->times) {
        if (k >= cap) break;
        if (names) {
            std::strncpy(names + 48 * k, kv.first.c_str(), 47);
            names[48 * k + 47] = 0;
        }
        if (counts) counts[k] = kv.second.first;
        if (total_ms) total_ms[k] = kv.second.second;
        k++;
    }
    *n = k;
    return MDX_OK;
}

int mdx_synchronize(mdx_system* s) {
    if (!s) return fail(s, MDX_ERR_ARG, "null system");
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return MDX_OK;
}

int mdx_launch_shape(mdx_system* s, mdx_shape* out) {
    if (!s || !out) return fail(s, MDX_ERR_ARG, "null argument");
    *out = s->shape;
    return MDX_OK;
}

const char* mdx_sweep_last_error(const mdx_sweep* w) { return w ? w->err.c_str() : g_sweep_err.c_str(); }

int mdx_sweep_create(const mdx_config* cfg, int32_t members, mdx_sweep** out) {
    if (!out) return sweep_fail(nullptr, MDX_ERR_ARG, "out is null");
    *out = nullptr;
    if (members < 1) return sweep_fail(nullptr, MDX_ERR_ARG, "a sweep needs at least one member");
    if (members > 65535) return sweep_fail(nullptr, MDX_ERR_ARG, "a sweep holds at most 65535 members");
    std::string why;
    if (!valid_config(cfg, why)) return sweep_fail(nullptr, MDX_ERR_ARG, why);
    if (cfg->box != MDX_BOX_WALLED) return sweep_fail(nullptr, MDX_ERR_ARG, "sweep members are walled systems");
    mdx_sweep* w = new mdx_sweep();
    w->cfg = *cfg;
    auto bail = [&](int code, const std::string& msg) {
        mdx_sweep_destroy(w);
        return sweep_fail(nullptr, code, msg);
    };
    if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(MDX_ERR_HIP, "hipStreamCreate failed");
    for (int32_t m = 0; m < members; m++) {
        mdx_system* s = nullptr;
        int rc = create_system(cfg, w->stream, &s);
        if (rc) return bail(rc, "member " + std::to_string(m) + ": " + g_err);
        s->sweep_member = true;
        w->sys.push_back(s);
    }
    const size_t bytes = sizeof(MemberView) * members;
    if (hipMalloc((void**)&w->dviews, bytes) != hipSuccess) return bail(MDX_ERR_ALLOC, "hipMalloc failed");
    if (hipHostMalloc((void**)&w->hviews, bytes, hipHostMallocDefault) != hipSuccess)
        return bail(MDX_ERR_ALLOC, "hipHostMalloc failed");
    std::memset(w->hviews, 0, bytes);
    *out = w;
    return MDX_OK;
}

int mdx_sweep_destroy(mdx_sweep* w) {
    if (!w) return MDX_OK;
    if (w->stream) hipStreamSynchronize(w->stream);
    for (mdx_system* s : w->sys) destroy_system(s);
    if (w->dviews) hipFree(w->dviews);
    if (w->hviews) hipHostFree(w->hviews);
    if (w->stream) hipStreamDestroy(w->stream);
    delete w;
    return MDX_OK;
}

int mdx_sweep_member(mdx_sweep* w, int32_t member, mdx_system** out) {
    if (!w || !out) return sweep_fail(w, MDX_ERR_ARG, "null argument");
    if (member < 0 || member >= (int32_t)w->sys.size()) return sweep_fail(w, MDX_ERR_ARG, "member out of range");
    *out = w->sys[member];
    return MDX_OK;
}

int mdx_sweep_run(mdx_sweep* w, int32_t nsteps) {
    if (!w) return sweep_fail(w, MDX_ERR_ARG, "null sweep");
    if (nsteps < 0) return sweep_fail(w, MDX_ERR_ARG, "nsteps must be >= 0");
    return sweep_run(w, nsteps);
}

int mdx_sweep_set_timing(mdx_sweep* w, int32_t enabled) {
    if (!w) return sweep_fail(w, MDX_ERR_ARG, "null sweep");
    for (mdx_system* s : w->sys) {
        int rc = mdx_set_timing(s, enabled);
        if (rc) return sweep_fail(w, rc, s->err);
    }
    return MDX_OK;
}

}  // extern "C"
