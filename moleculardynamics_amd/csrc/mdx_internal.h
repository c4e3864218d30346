// mdx_internal.h -- device-side data layout and kernel entry points of the
// gfx950 engine.  Not part of the C-ABI (include/mdx.h is).
//
// Layout in HBM: every per-atom quantity is a separate FP64 array (SoA) in
// "slot" order.  Slots are the cell-sorted order produced by the last build
// of the primary cell grid: the atoms of cell c occupy slots
// [start[c], start[c+1]) in ascending original index.  id[slot] is the
// caller's atom index and slot_of_id[] its inverse.  Kernels stream the SoA
// arrays with unit stride; neighbour cells are contiguous slot ranges.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/mdx.h"

namespace mdx {

constexpr int kWave = 64;

// Quads of 4 consecutive lanes serving one atom.  All four lanes must be
// active: a DPP move reading a disabled lane gets a stale value.
template <int kCtrl>
__device__ inline double quad_perm(double v) {   // kCtrl: DPP quad permutation
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), kCtrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), kCtrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
// (v0 + v1) + (v2 + v3) in every lane, the same bits in all four:
// permutations [1,0,3,2] = 0xB1, [2,3,0,1] = 0x4E
__device__ inline double quad_sum(double v) {
    v += quad_perm<0xB1>(v);
    v += quad_perm<0x4E>(v);
    return v;
}
// exclusive prefix of h over the quad in lane order (q = lane & 3) and the
// quad total; the empty asm keeps each permutation ahead of the per-lane
// select, executed by all four lanes
__device__ inline int32_t quad_excl(int32_t h, int32_t q, int32_t& total) {
    int32_t t1 = __builtin_amdgcn_mov_dpp(h, 0x90, 0xF, 0xF, false);   // [0,0,1,2]
    asm volatile("" : "+v"(t1));
    int32_t incl = h + (q >= 1 ? t1 : 0);
    int32_t t2 = __builtin_amdgcn_mov_dpp(incl, 0x44, 0xF, 0xF, false);   // [0,1,0,1]
    asm volatile("" : "+v"(t2));
    incl += q >= 2 ? t2 : 0;
    total = __builtin_amdgcn_mov_dpp(incl, 0xFF, 0xF, 0xF, false);     // [3,3,3,3]
    return incl - h;
}

// A 32-bit load from global memory at an address the compiler cannot see
// through (the look-ahead loads hide their address behind an empty asm so
// they stay one iteration ahead).  The cast keeps the load a global one: a
// flat load's completion is tracked together with the LDS reads', so every
// LDS wait after it would wait for this load too.
__device__ inline int32_t load_global_i32(uint64_t addr) {
    return *(__attribute__((address_space(1))) const int32_t*)(uintptr_t)addr;
}

// lane l's value of v in every lane (l uniform across the wavefront)
__device__ inline double lane_double(double v, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}

// The per-step flag hand-over of the periodic LJ step kernel (k_step_lds), in
// place of a device-to-host copy after each launch.  A launch reads the flag
// words of the launch before it, which are complete since that launch has
// ended: every wavefront takes its list gate from them, and workgroup 0
// copies them to a slot of pinned host memory and then writes the sequence
// number after them, so a host that sees the number sees the words.  The
// step kernel's workgroups thus end right after their flag atomics, with no
// wait for their stores and no arrival count (the last arrival of a count
// at the end of each workgroup handing the words over kept every
// workgroup's LDS the time of a store drain and a returning atomic longer).
// The last launch's words are copied by the host.
//
// The list gate: a step kernel whose previous flag words hold a largest
// displacement^2 since the list build (words 3q) at or above gate_bits
// (non-negative doubles order like their bits), unless it is the first
// launch on a new list (open), runs as a no-op ("halted"):
// no state is written, and workgroup 0 marks the step's own flag words with
// kHaltBits (so the launch queued after it halts too) and kHaltMark in the
// first stale word (so the host sees it).
constexpr unsigned long long kHaltMark = 1ull << 62;
constexpr unsigned long long kHaltBits = 0x7ff0000000000000ull;   // +inf
struct FlagPublish {
    const unsigned long long* prev;   // the previous launch's flag words
    unsigned long long* host;         // their host slot, then the sequence number (null: none)
    unsigned long long seq;
    unsigned long long gate_bits;
    int32_t open;                     // 1: the list was built from r(t), no gate
};

// workgroup 0, every thread: prev's words to the host slot, then the number
__device__ inline void hand_over(const FlagPublish& p, int32_t words) {
    if (!p.host) return;
    for (int32_t w = threadIdx.x; w < words; w += blockDim.x)
        __hip_atomic_store(p.host + w, p.prev[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __builtin_amdgcn_s_waitcnt(0);   // the words have been written
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(p.host + words, p.seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Per-atom SoA state (slot order).
struct DevState {
    double* r[3];
    double* v[3];
    double* a[3];
    double* b[3];
    double* c[3];
    double* f[3];
    int32_t* type;
    int32_t* id;
};

// Cell grid as kernels see it.
struct GridView {
    int32_t g[3];      // storage extent per axis
    int32_t cells;     // neighbour cells with index >= cells are skipped (walled)
    int32_t periodic;  // 1: wrap with image shift
    double L;          // periodic edge
    double h;          // cell edge (walled: rs as double)
    double npos;       // walled: half of box_size_list
    const int32_t* start;  // ncell + 1 prefix offsets
};

// Pair table for Lennard-Jones, indexed by ti*kTypes + tj.
constexpr int kTypes = 4;
struct LjTable {
    float sig2f[kTypes * kTypes];
    float inv_sig2f[kTypes * kTypes];
    float coeff_f[kTypes * kTypes];   // 24 eps / sig^2
    float e4_f[kTypes * kTypes];      // 4 eps
    double sig2[kTypes * kTypes];
    double coeff[kTypes * kTypes];
    double e4[kTypes * kTypes];
    double shift[kTypes * kTypes];    // energy at rc (0 for the walled box)
    uint32_t active;                  // bit ti*kTypes+tj
    double rc2;
    float rc2f;
};

struct MassTable {
    double mass[kTypes];
    double scale;   // a = F / mass * scale
};

// Reference bond-order parameter table (common/Constants.h, common/Getters.cpp),
// indexed by [same-species N, same-species O, unlike] via pair_kind().
struct BondRefParams {
    double re[3], de[3], beta[3], r0v[3];
    double rcut2;        // list cutoff squared (mb rcut^2)
    double pair_cut2;    // CUTOFF_RIJSQ
    double pair_cut;     // CUTOFF_RIJ
};

// Tersoff silicon: c2 = c^2, d2 = d^2; zeta term exp((lambda3 (r_ij - r_ik))^m)
struct TersoffParams {
    double A, B, lambda1, lambda2, lambda3, beta, n, c2, d2, h, R, D;
    int32_t m;
};

// One member of a sweep (mdx_sweep_*): a walled system whose step kernels run
// in one launch with the other members' (blockIdx.y = member).  The host
// rewrites a member's view after each of its list builds.
struct MemberView {
    DevState st;
    const int32_t* cell_of;      // LJ grid cell per slot
    GridView lj;
    const int32_t* slot_of_id;
    const int32_t* mb_cell_of;   // bond-order grid cell per id
    const int32_t* mb_ids;       // bond-order grid members (ids), cell-sorted
    GridView mb;
    int32_t* xi_j;
    double* xi_v;
    int32_t xi_cap;              // xi table entries per bond-order atom
    int32_t bond_wave;           // 1: the bond-order force of this member runs one wavefront per atom
    int32_t* hit;                // wall-hit word of the member's corrector
    int32_t pad;
};

// ---- host-side launch wrappers (one translation unit each) ----
// cells.hip
// p[0..n) = v; block 0 also zeroes zero[0..nzero) (nzero <= 256), which saves
// a memset's launch for a few flag words
hipError_t launch_fill_i32(hipStream_t s, int32_t* p, int64_t n, int32_t v, int32_t* zero = nullptr,
                           int32_t nzero = 0);
hipError_t launch_bin(hipStream_t s, const DevState& st, const int32_t* item_slot, int32_t items,
                      const GridView& g, int32_t* cell_out, int32_t* counts, int32_t* err);
hipError_t launch_exclusive_scan(hipStream_t s, const int32_t* in, int32_t* out, int64_t n,
                                 int32_t* block_sums, int64_t block_sums_cap);
int64_t scan_block_sums_needed(int64_t n);
hipError_t launch_scatter(hipStream_t s, const int32_t* cell_of, const int32_t* start,
                          int32_t* fill, int32_t* sorted_items, int32_t items);
// items_in grouped by cell -> items_out with ascending key inside each cell
// (key_of_item null: the item is its own key)
hipError_t launch_sort_cells(hipStream_t s, const int32_t* start, int64_t ncell, const int32_t* items_in,
                             int32_t* items_out, const int32_t* key_of_item,
                             int64_t items);
hipError_t launch_permute(hipStream_t s, const DevState& src, const DevState& dst,
                          const int32_t* sorted_src, const int32_t* cell_of_src,
                          int32_t* cell_of_dst, int32_t* slot_of_id, double* const rref[3],
                          int32_t n, int32_t wrap, double L);
hipError_t launch_max_count(hipStream_t s, const int32_t* counts, int64_t ncell, int32_t* out);
// A walled grid build of few items and cells in one workgroup (one launch):
// the bin, scan, scatter, in-cell sort, state permutation and sizes of the
// kernels above, with the same results.
struct SmallBuild {
    DevState src;              // positions to bin (and the state to permute)
    DevState dst;              // the permuted state (dst.r[0] null: no permutation)
    const int32_t* item_slot;  // slot of item q (null: item q is slot q)
    int32_t items;
    GridView g;                // the grid; g.start (ncell + 1) is written
    int64_t ncell;
    int32_t* cell_out;         // the cell of each item (item order)
    int32_t* sorted;           // items in cell order, ascending key inside a cell
    const int32_t* key_of;     // key of item q (null: q itself)
    int32_t* cell_of_dst;      // with dst: the cells in the new slot order
    int32_t* slot_of_id;       // with dst: the slot of each atom id
    int32_t* words;            // [0] out-of-box flag, [1] largest cell, [2] longest row
};
constexpr int32_t kSmallBuildItems = 6144;
constexpr int64_t kSmallBuildCells = 12287;
bool small_build_fits(int32_t items, int64_t ncell);
hipError_t launch_small_build(hipStream_t s, const SmallBuild& b);
hipError_t launch_row_lengths(hipStream_t s, const GridView& g, const int32_t* cell_by_item,
                              int32_t items, int32_t* max_len);
hipError_t launch_emit_list(hipStream_t s, const GridView& g, const int32_t* cell_by_item,
                            const int32_t* member_ids, int32_t items, int32_t nlistsize,
                            int32_t* out);
hipError_t launch_gather_i32(hipStream_t s, const int32_t* src, const int32_t* idx, int32_t* dst,
                             int32_t n);
// dst[d][idx[i]] = src[d][i] for the three components
hipError_t launch_scatter3(hipStream_t s, double* const src[3], const int32_t* idx, double* const dst[3], int32_t n);
// to_slots: dst[idx[i]] = src[i] (caller order -> slots); else dst[i] = src[idx[i]]
hipError_t launch_reorder(hipStream_t s, const DevState& src, const DevState& dst,
                          const int32_t* idx, int32_t n, int32_t to_slots);

// lj.hip
// gate: when non-null the kernel does nothing unless *gate != 0 (device-side
// predicate for the force recompute after a wall hit)
// with energy_per_atom the per-atom virial (half of r_ij . F_ij over the
// atom's pairs) goes to virial_per_atom, which must then be non-null
hipError_t launch_lj(hipStream_t s, const DevState& st, const int32_t* cell_of, const GridView& g,
                     const LjTable& t, int32_t n, int32_t precision, double* energy_per_atom,
                     const int32_t* gate, double* virial_per_atom = nullptr);
// the LJ force of every sweep member (the kernel launch_lj picks for a walled
// box of n atoms); gated: only the members whose wall-hit word is set
hipError_t launch_sweep_lj(hipStream_t s, const MemberView* views, int32_t members, int32_t n, const LjTable& t,
                           int32_t precision, bool gated);

// gear.hip
struct GearCoeffs {
    double k[4];    // dt, dt^2/2, dt^3/6, dt^4/24 (as the reference forms them)
    double kc[4];   // corrector: GEAR1 k1, GEAR2 k1/k0, GEAR3 k1/k2, GEAR4 k1/k3
};
GearCoeffs gear_coeffs(float dt);
// clear (may be null): a word the kernel zeroes (the corrector's wall-hit word)
hipError_t launch_predict(hipStream_t s, const DevState& st, int32_t n, const GearCoeffs& gc, int32_t* clear = nullptr);
hipError_t launch_correct_walled(hipStream_t s, const DevState& st, int32_t n, const GearCoeffs& gc,
                                 const MassTable& m, double npos, float eloss, int32_t* hit);
hipError_t launch_accelerations(hipStream_t s, const DevState& st, int32_t n, const MassTable& m,
                                const int32_t* gate);
hipError_t launch_correct_predict(hipStream_t s, const DevState& st, int32_t n, const GearCoeffs& gc,
                                  const MassTable& m, double* const rref[3], uint32_t* maxdisp2_bits);
// the same kernels over every member of a sweep (views: device array of
// `members` MemberViews, n atoms each).  The predictor clears each member's
// wall-hit word; the corrector sets it; the accelerations (a = F/m) run only
// for the members whose word is set.
hipError_t launch_sweep_predict(hipStream_t s, const MemberView* views, int32_t members, int32_t n,
                                const GearCoeffs& gc);
hipError_t launch_sweep_correct(hipStream_t s, const MemberView* views, int32_t members, int32_t n,
                                const GearCoeffs& gc, const MassTable& m, double npos, float eloss);
hipError_t launch_sweep_accelerations(hipStream_t s, const MemberView* views, int32_t members, int32_t n,
                                      const MassTable& m);

// predictor that also max-reduces |r - rref|^2 and the squared step
// |r_new - r_old|^2 (bits of non-negative doubles) into the spread flag
// layout of the step kernel: flags[3 q + {0, 1}], q = block % flag_spread()
hipError_t launch_predict_track(hipStream_t s, const DevState& st, int32_t n, const GearCoeffs& gc,
                                double* const rref[3], uint64_t* flags);
// corrector of step t and (unless final_step) predictor of step t+1 with the
// same flag tracking, in place (forces computed by a separate kernel)
// clear_next (may be null): the next step's flag words, zeroed by the kernel
hipError_t launch_correct_predict_track(hipStream_t s, const DevState& st, int32_t n, const GearCoeffs& gc,
                                        const MassTable& m, double* const rref[3], uint64_t* flags, bool final_step,
                                        uint64_t* clear_next);
// Forces from per-slot gradients (the surrogate): F_i = sum_s w[s][d][i] -
// w[rev[s][i]][d][list[s][i]] in FP64, written to st.f.
struct SlotGather {
    const int32_t* list;
    const int32_t* count;
    const uint8_t* rev;
    const float4* w;    // [slot][atom], x y z used
    const double4* wd;  // the same in FP64 (used instead of w when non-null)
};
// the gather, the stale count of st's positions against rref (sqrt(hard2)),
// then launch_correct_predict_track's corrector and predictor, in one kernel;
// workgroup 0 hands pub.prev's flag words over to pub.host first (no gate)
hipError_t launch_gather_correct_predict_track(hipStream_t s, const DevState& st, int32_t n, const GearCoeffs& gc,
                                               const MassTable& m, double* const rref[3], uint64_t* flags,
                                               bool final_step, const SlotGather& g, double hard2,
                                               uint64_t* clear_next, const FlagPublish& pub);

// lj_fused.hip: LDS-staged segment kernels of the periodic box
// segments per cell column for workgroups of `block` threads (4 lanes an atom)
int32_t lds_segments_per_column(int32_t nz, double atoms_per_cell, int32_t kmax, int32_t block, double fill);
// launch shapes (tuning.hip): the built-in one, and the tuning table's for a
// system of kind ("lj" or "replicas") of this many atoms on the current device
mdx_shape default_shape();
int tuned_shape(const char* kind, int64_t atoms, mdx_shape* out, std::string& err);
int32_t lds_capacity();
int32_t flag_spread();
int64_t list_entries(int32_t n, int32_t cap);   // 16-bit entries to allocate for the list
int32_t list_cap_multiple(); // the list capacity (entries per atom) must be a multiple of this
// err[0]: staged atoms of a segment exceeded lds_atoms; err[1]: longest list above
// cap; err[2]: the largest staging (atoms), which sizes the step kernel's LDS
hipError_t launch_nlist_lds(hipStream_t s, const DevState& st, const GridView& v, int32_t nseg, double rl2,
                            int32_t n, int32_t cap, uint16_t* list, int32_t* count, int32_t* err,
                            int32_t lds_atoms, int32_t block, int32_t* hdr);
// the list build's per-segment headers, read by the step kernels: words per
// segment, and the segments of a grid cut nseg times per column
int32_t seg_header_words();
int64_t seg_count(const GridView& v, int32_t nseg);
// st, rn and r0 hold each vector's three components in one block, component d
// at [0] + d * cs; clear_next (may be null): the next step's flag words, zeroed;
// pub: the flag hand-over (blocks is set here); block: 512 or 1024 threads a
// workgroup (the segments must have been cut for it); the force is written on
// the final step only (force_out == final_step)
hipError_t launch_step_lds(hipStream_t s, const DevState& st, const GridView& v, int32_t nseg,
                           double* const rn[3], double* const r0[3], const uint16_t* list, const int32_t* count,
                           int32_t cap, const LjTable& t, const MassTable& m, const GearCoeffs& gc, double hard2,
                           int32_t n, int64_t cs, int32_t precision, bool final_step, bool force_out, uint64_t* flags,
                           int32_t lds_atoms, uint64_t* clear_next, FlagPublish pub, int32_t block,
                           const int32_t* hdr);

// the lists of launch_nlist_lds as caller indices: row id[i] of rows (width
// entries) holds atom i's partners in list order, counts[id[i]] their number;
// *err receives a count above width
hipError_t launch_nlist_decode(hipStream_t s, const GridView& v, int32_t nseg, const uint16_t* list,
                               const int32_t* count, int32_t cap, const int32_t* id, int32_t width, int32_t* rows,
                               int32_t* counts, int32_t* err);
// the same for compact [k][slot] lists of slots (Tersoff)
hipError_t launch_compact_decode(hipStream_t s, const int32_t* list, const int32_t* count, int32_t n,
                                 const int32_t* id, int32_t width, int32_t* rows, int32_t* counts, int32_t* err);

// tersoff.hip: Tersoff silicon in the periodic box (compact [k][slot] lists)
TersoffParams tersoff_si_params();
// list of the atoms within rs of each slot; err: longest list when above cap
hipError_t launch_ts_list(hipStream_t s, const DevState& st, const int32_t* cell_of, const GridView& g, double rs,
                          int32_t n, int32_t cap, int32_t* list, int32_t* count, int32_t* err);
// e_i and the per-slot vectors w_is = de_i/d(x_j - x_i) of every atom: FP32
// terms into w (MDX_PREC_MIXED) or FP64 into wd (MDX_PREC_FP64); e_atom may be null
hipError_t launch_ts_atoms(hipStream_t s, const DevState& st, double L, const TersoffParams& p, int32_t precision,
                           const int32_t* list, const int32_t* count, int32_t n, float4* w, double4* wd,
                           double* e_atom);

// surrogate.hip: learned surrogate of Tersoff silicon over the same lists
// FP32 blob (MDX_SG_NPARAM) -> BF16 MFMA fragments and FP32 vectors
void sg_pack(const float* blob, uint16_t* frag, float* vec);
size_t sg_frag_u16();
size_t sg_vec_len();
int32_t sg_slots();
// rev[k][i]: the slot of i in the list of its k-th partner; err set if absent
hipError_t launch_sg_rev(hipStream_t s, const int32_t* list, const int32_t* count, int32_t n, uint8_t* rev,
                         int32_t* err);
// forces of st's positions into out.f (and per-atom energies into e_atom if
// non-null); nflag counts the atoms that took the Tersoff fallback; stale
// atoms against rref are counted into the spread flags
// F = the gather of per-slot vectors (w, or wd when non-null) into out.f;
// atoms farther than sqrt(hard2) from rref are counted as stale into flags
hipError_t launch_slot_gather(hipStream_t s, const DevState& st, double L, const int32_t* list, const int32_t* count,
                              const uint8_t* rev, int32_t n, const float4* w, const double4* wd,
                              double* const rref[3], double hard2, uint64_t* flags, const DevState& out);
// descriptors, network and per-slot gradients only (w), for the fused step
hipError_t launch_sg_eval(hipStream_t s, const DevState& st, double L, const TersoffParams& p, const int32_t* list,
                          const int32_t* count, int32_t n, const uint16_t* frag, const float* vec, float4* w,
                          double* e_atom, int32_t* nflag);
hipError_t launch_sg_forces(hipStream_t s, const DevState& st, double L, const TersoffParams& p, const int32_t* list,
                            const int32_t* count, const uint8_t* rev, int32_t n, const uint16_t* frag,
                            const float* vec, float4* w, double* e_atom, int32_t* nflag, double* const rref[3],
                            double hard2, uint64_t* flags, const DevState& out);

// lj_fused.hip, batched replicas: block b serves cells seg[4b+2] .. seg[4b+3]
// (exclusive) of z-column seg[4b+1] of replica seg[4b] (every replica here has
// nc >= 3 cells per axis); cstart holds the replicas' cell starts as global
// slots, nc^3 + 1 per replica at coff[r]
struct RepSegs {
    const int32_t* seg;   // 4 per block, 16-byte aligned
    int32_t nblocks;
    const int32_t* nc;
    const double* L;
    const int32_t* cstart;
    const int32_t* coff;
};
// the lists of the segments wl[0 .. *wcount) (a fixed grid loops over them)
hipError_t launch_rep_nlist(hipStream_t s, const RepSegs& rs, const int32_t* wl, const int32_t* wcount,
                            double* const r[3], double rl2, int32_t cap, uint16_t* list, int32_t* count,
                            int32_t* err, int32_t lds_atoms, int32_t* hdr);
hipError_t launch_rep_step(hipStream_t s, const RepSegs& rs, const DevState& st, double* const rn[3],
                           double* const r0[3], const uint16_t* list, const int32_t* count, int32_t cap, double rc2,
                           double sig2, double coeff, double mass, double scale, const GearCoeffs& gc, int64_t cs,
                           int32_t precision, bool final_step, unsigned long long* dmax, int32_t* err,
                           int32_t lds_atoms, const int32_t* hdr);

// bond_ref.hip
hipError_t launch_bond_ref(hipStream_t s, const DevState& st, const int32_t* slot_of_id,
                           const int32_t* mb_cell_of, const int32_t* mb_ids, const GridView& g,
                           const BondRefParams& p, int32_t n_bond, int32_t nlistsize,
                           int32_t precision, int32_t assign, int32_t* xi_j, double* xi_v,
                           double* energy_per_atom, const int32_t* gate, int64_t row_bound);
// whether launch_bond_ref runs the wavefront-per-atom kernels for these sizes
bool bond_ref_wave(int32_t n_bond, int64_t row_bound, int32_t cap);
// the bond-order force of every sweep member, each on the kernels
// launch_bond_ref would pick for it (MemberView::bond_wave); launches only
// the kinds named by any_wave / any_thread; gated as launch_sweep_lj
hipError_t launch_sweep_bond(hipStream_t s, const MemberView* views, int32_t members, const BondRefParams& p,
                             int32_t n_bond, int32_t precision, int32_t assign, bool gated, bool any_wave,
                             bool any_thread);

// reduce.hip
hipError_t launch_kinetics(hipStream_t s, const DevState& st, int32_t n, const MassTable& m,
                           double* partials, int32_t nblocks, double* out8, double kboltz);
int32_t kinetics_blocks(int32_t n);
hipError_t launch_sum(hipStream_t s, const double* x, int32_t n, double* partials, int32_t nblocks,
                      double* out);
// out2_bits: max |v|^2 and max |coordinate| (bits of non-negative doubles); clears them itself
hipError_t launch_ready_list(hipStream_t s, const DevState& st, int32_t n, uint64_t* out2_bits);
// w[i] = pos.r_s . frc.f_s for slot s = map[i] (map null: s = i), i < n
hipError_t launch_rf_virial(hipStream_t s, const DevState& pos, const DevState& frc, const int32_t* map, int32_t n,
                            double* w);
// per-atom virial from per-slot gradients w_is = de_i/d(x_j - x_i):
// w_atom[i] = -sum_s (x_j - x_i) . w_is (FP32 w, or FP64 wd when non-null)
hipError_t launch_slot_virial(hipStream_t s, const DevState& st, double L, const int32_t* list, const int32_t* count,
                              int32_t n, const float4* w, const double4* wd, double* w_atom);
hipError_t launch_allpairs(hipStream_t s, const DevState& st, int32_t n, int32_t n_bond,
                           const LjTable& lj, const BondRefParams& p, double* per_atom2);

}  // namespace mdx
