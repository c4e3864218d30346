// lj.hip -- Lennard-Jones pair force over the 27-cell neighbourhood of each
// atom's cell (cpu/common/LennardJones.h:12-93).
//
// Gather formulation: one thread per atom (slot order), no atomics, every
// atom's force summed in the reference's candidate order (neighbour cells in
// (dx,dy,dz) lexicographic order, ascending index within a cell), so the
// FP64 variant reproduces the oracle's sums operation for operation.
// The displacement and the cutoff test are FP64 in both variants; the mixed
// variant evaluates the 12-6 force law in FP32 from the FP64 distance and
// accumulates the per-atom force in FP64.
#include "mdx_internal.h"

namespace mdx {

namespace {

constexpr int kBlock = 256;

// the force on slot s (thread per atom)
template <bool kFp64, bool kEnergy>
__device__ inline void lj_atom(const DevState& st, const int32_t* __restrict__ cell_of, const GridView& g,
                               const LjTable& t, int32_t s, double* __restrict__ e_atom,
                               double* __restrict__ w_atom) {
    const double xi = st.r[0][s], yi = st.r[1][s], zi = st.r[2][s];
    const int32_t ti = st.type[s] & (kTypes - 1);
    const int32_t own = cell_of[s];
    const int32_t cz = own % g.g[2], cy = (own / g.g[2]) % g.g[1], cx = own / (g.g[2] * g.g[1]);
    double fx = 0, fy = 0, fz = 0, e = 0, w = 0;
    for (int dx = -1; dx <= 1; dx++) {
        int32_t x = cx + dx;
        double sx = 0.0;
        if (g.periodic) {
            if (x < 0) { x += g.g[0]; sx = -g.L; }
            else if (x >= g.g[0]) { x -= g.g[0]; sx = g.L; }
        } else if (x < 0 || x >= g.cells) continue;
        for (int dy = -1; dy <= 1; dy++) {
            int32_t y = cy + dy;
            double sy = 0.0;
            if (g.periodic) {
                if (y < 0) { y += g.g[1]; sy = -g.L; }
                else if (y >= g.g[1]) { y -= g.g[1]; sy = g.L; }
            } else if (y < 0 || y >= g.cells) continue;
            for (int dz = -1; dz <= 1; dz++) {
                int32_t z = cz + dz;
                double sz = 0.0;
                if (g.periodic) {
                    if (z < 0) { z += g.g[2]; sz = -g.L; }
                    else if (z >= g.g[2]) { z -= g.g[2]; sz = g.L; }
                } else if (z < 0 || z >= g.cells) continue;
                const int32_t c = (x * g.g[1] + y) * g.g[2] + z;
                const int32_t pe = g.start[c + 1];
                for (int32_t p = g.start[c]; p < pe; p++) {
                    if (p == s) continue;
                    const double ddx = (xi - st.r[0][p]) - sx;
                    const double ddy = (yi - st.r[1][p]) - sy;
                    const double ddz = (zi - st.r[2][p]) - sz;
                    const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                    if (r2 >= t.rc2) continue;
                    const int32_t k = ti * kTypes + (st.type[p] & (kTypes - 1));
                    if (!((t.active >> k) & 1u)) continue;
                    if constexpr (kFp64) {
                        const double q = r2 / t.sig2[k];
                        const double iq4 = 1.0 / (q * q * q * q);
                        const double iq7 = iq4 / (q * q * q);
                        const double fr = (2 * iq7 - iq4) * t.coeff[k];
                        fx += ddx * fr;
                        fy += ddy * fr;
                        fz += ddz * fr;
                        if constexpr (kEnergy) {
                            const double iq6 = 1.0 / (q * q * q);
                            e += 0.5 * (t.e4[k] * iq6 * (iq6 - 1) - t.shift[k]);
                            w += 0.5 * fr * r2;   // r_ij . F_ij, half on each side of the pair
                        }
                    } else {
                        // s2 = (sig/r)^2, s6 = (sig/r)^6, s8 = (sig/r)^8
                        const float s2 = __builtin_amdgcn_rcpf((float)r2 * t.inv_sig2f[k]);
                        const float s4 = s2 * s2;
                        const float s6 = s4 * s2;
                        const float s8 = s4 * s4;
                        // 24 eps/sig^2 [2 (sig/r)^14 - (sig/r)^8] = coeff s8 (2 s6 - 1)
                        const float fr = s8 * __builtin_fmaf(2.0f, s6, -1.0f) * t.coeff_f[k];
                        const double frd = (double)fr;
                        fx = __builtin_fma(frd, ddx, fx);
                        fy = __builtin_fma(frd, ddy, fy);
                        fz = __builtin_fma(frd, ddz, fz);
                        if constexpr (kEnergy) {
                            e += 0.5 * ((double)(t.e4_f[k] * s6 * (s6 - 1.0f)) - t.shift[k]);
                            w += 0.5 * frd * r2;
                        }
                    }
                }
            }
        }
    }
    st.f[0][s] = fx;
    st.f[1][s] = fy;
    st.f[2][s] = fz;
    if constexpr (kEnergy) {
        e_atom[s] = e;
        w_atom[s] = w;
    }
}

template <bool kFp64, bool kEnergy>
__global__ void __launch_bounds__(kBlock)
k_lj(DevState st, const int32_t* __restrict__ cell_of, GridView g, LjTable t, int32_t n,
     double* __restrict__ e_atom, double* __restrict__ w_atom, const int32_t* __restrict__ gate) {
    if (gate && *gate == 0) return;
    int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    lj_atom<kFp64, kEnergy>(st, cell_of, g, t, s, e_atom, w_atom);
}

// Walled box, few atoms: one wavefront per atom.  The 27 candidate cells of
// the reference's row are 9 z-runs of up to 3 consecutive cells, i.e. 9
// contiguous slot ranges in row order; the 64 lanes take 64 consecutive row
// entries at a time and test them against rc.  The pairs within rc are
// appended, in row order, to a 64-entry buffer of the wavefront in LDS
// (displacement, r^2, pair type): about one candidate in nine of a walled gas
// lies within rc, so the force law, with its three FP64 divisions, then runs
// over full wavefronts of pairs instead of over every batch of candidates.  A
// full (or the last) buffer is evaluated one pair a lane, and lanes 0-2 (3
// and 4: energy and virial) add their component of the terms in buffer
// order: the thread-per-atom kernel's additions, operation for operation.
constexpr int kWaveBlock = 256;

// the force on slot s, the whole wavefront (lane = its lane)
template <bool kFp64, bool kEnergy>
__device__ inline void lj_atom_wave(const DevState& st, const int32_t* __restrict__ cell_of, const GridView& g,
                                    const LjTable& t, int32_t s, int lane, double* __restrict__ e_atom,
                                    double* __restrict__ w_atom) {
    // per wavefront: dx, dy, dz, r^2 of the pending pairs; after their
    // evaluation rows 0-2 hold the force terms (FP64) or stay the
    // displacement with the FP32 law's F/r in row 3 (mixed); rows 4-5 the
    // energy and virial terms
    __shared__ double pend[kWaveBlock / 64][6][64];
    __shared__ int32_t pend_k[kWaveBlock / 64][64];
    double (*const B)[64] = pend[threadIdx.x >> 6];
    int32_t* const K = pend_k[threadIdx.x >> 6];
    const double xi = st.r[0][s], yi = st.r[1][s], zi = st.r[2][s];
    const int32_t ti = st.type[s] & (kTypes - 1);
    const int32_t own = cell_of[s];
    const int32_t cz = own % g.g[2], cy = (own / g.g[2]) % g.g[1], cx = own / (g.g[2] * g.g[1]);
    // lane r < 9: slot range of the column (cx + r/3 - 1, cy + r%3 - 1), z clipped to the grid
    int32_t rb = 0, rn = 0;
    if (lane < 9) {
        const int32_t x = cx + lane / 3 - 1, y = cy + lane % 3 - 1;
        if (x >= 0 && x < g.cells && y >= 0 && y < g.cells) {
            const int32_t z0 = max(cz - 1, 0), z1 = min(cz + 1, g.cells - 1);
            if (z0 <= z1) {
                const int32_t col = (x * g.g[1] + y) * g.g[2];
                rb = g.start[col + z0];
                rn = g.start[col + z1 + 1] - rb;
            }
        }
    }
    int32_t base[9], off[9], len = 0;
#pragma unroll
    for (int r = 0; r < 9; r++) {
        base[r] = __builtin_amdgcn_readlane(rb, r);
        off[r] = len;
        len += __builtin_amdgcn_readlane(rn, r);
    }
    // lane c < 3: force component c; 3: energy; 4: virial
    constexpr int kSums = kEnergy ? 5 : 3;
    const int row = lane < 3 ? lane : lane + 1;
    double acc = 0.0;
    // evaluate the first c pending pairs and add their terms in order
    auto flush = [&](int32_t c) {
        // LDS operations of one wavefront take effect in program order; the
        // empty asm keeps the compiler from moving reads of other lanes'
        // entries above their writes
        asm volatile("" ::: "memory");
        if (lane < c) {
            const double ddx = B[0][lane], ddy = B[1][lane], ddz = B[2][lane], r2 = B[3][lane];
            const int32_t k = K[lane];
            double te = 0.0, tw = 0.0;
            if constexpr (kFp64) {
                const double qq = r2 / t.sig2[k];
                const double iq4 = 1.0 / (qq * qq * qq * qq);
                const double iq7 = iq4 / (qq * qq * qq);
                const double fr = (2 * iq7 - iq4) * t.coeff[k];
                B[0][lane] = ddx * fr;
                B[1][lane] = ddy * fr;
                B[2][lane] = ddz * fr;
                if constexpr (kEnergy) {
                    const double iq6 = 1.0 / (qq * qq * qq);
                    te = 0.5 * (t.e4[k] * iq6 * (iq6 - 1) - t.shift[k]);
                    tw = 0.5 * fr * r2;
                }
            } else {
                const float s2 = __builtin_amdgcn_rcpf((float)r2 * t.inv_sig2f[k]);
                const float s4 = s2 * s2;
                const float s6 = s4 * s2;
                const float s8 = s4 * s4;
                const double frd = (double)(s8 * __builtin_fmaf(2.0f, s6, -1.0f) * t.coeff_f[k]);
                B[3][lane] = frd;
                if constexpr (kEnergy) {
                    te = 0.5 * ((double)(t.e4_f[k] * s6 * (s6 - 1.0f)) - t.shift[k]);
                    tw = 0.5 * frd * r2;
                }
            }
            if constexpr (kEnergy) {
                B[4][lane] = te;
                B[5][lane] = tw;
            }
        }
        asm volatile("" ::: "memory");
        if (lane < kSums) {
            // FP64: acc + term, written as fma(1, term, acc), which rounds the
            // same; mixed: fma(F/r, d, acc) for the force components
            for (int32_t l = 0; l < c; l++) {
                const double m = (!kFp64 && lane < 3) ? B[3][l] : 1.0;
                acc = __builtin_fma(m, B[row][l], acc);
            }
        }
        asm volatile("" ::: "memory");
    };
    int32_t pending = 0;
    for (int32_t q0 = 0; q0 < len; q0 += 64) {
        const int32_t q = q0 + lane;
        int32_t p = -1;
        if (q < len) {
            // the last range starting at or before q holds q (an empty range
            // shares its start with the next one)
#pragma unroll
            for (int r = 0; r < 9; r++)
                if (q >= off[r]) p = base[r] + (q - off[r]);
        }
        bool in = false;
        double ddx = 0, ddy = 0, ddz = 0, r2 = 0;
        int32_t k = 0;
        if (p >= 0 && p != s) {
            ddx = xi - st.r[0][p];
            ddy = yi - st.r[1][p];
            ddz = zi - st.r[2][p];
            r2 = ddx * ddx + ddy * ddy + ddz * ddz;
            k = ti * kTypes + (st.type[p] & (kTypes - 1));
            in = r2 < t.rc2 && ((t.active >> k) & 1u);
        }
        const uint64_t m = __ballot(in);
        const int32_t nin = (int32_t)__popcll(m);
        if (pending + nin > 64) {
            flush(pending);
            pending = 0;
        }
        if (in) {
            // this lane's place among the batch's pairs: the set lanes below it
            const int32_t at = pending + (int32_t)__builtin_amdgcn_mbcnt_hi(
                                             (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            B[0][at] = ddx;
            B[1][at] = ddy;
            B[2][at] = ddz;
            B[3][at] = r2;
            K[at] = k;
        }
        pending += nin;
    }
    if (pending > 0) flush(pending);
    if (lane < 3) st.f[lane][s] = acc;
    if constexpr (kEnergy) {
        if (lane == 3) e_atom[s] = acc;
        if (lane == 4) w_atom[s] = acc;
    }
}

template <bool kFp64, bool kEnergy>
__global__ void __launch_bounds__(kWaveBlock)
k_lj_wave(DevState st, const int32_t* __restrict__ cell_of, GridView g, LjTable t, int32_t n,
          double* __restrict__ e_atom, double* __restrict__ w_atom, const int32_t* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int32_t s = (int32_t)((blockIdx.x * kWaveBlock + threadIdx.x) >> 6);
    if (s >= n) return;
    lj_atom_wave<kFp64, kEnergy>(st, cell_of, g, t, s, threadIdx.x & 63, e_atom, w_atom);
}

// below this many atoms a walled box takes the wavefront-per-atom kernel
constexpr int32_t kWaveAtoms = 32768;

// sweep members in one launch (blockIdx.y = member); gated: only the members
// whose wall-hit word is set
template <bool kFp64>
__global__ void __launch_bounds__(kBlock)
k_sweep_lj(const MemberView* __restrict__ views, LjTable t, int32_t n, int32_t gated) {
    const MemberView& v = views[blockIdx.y];
    if (gated && *v.hit == 0) return;
    int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    lj_atom<kFp64, false>(v.st, v.cell_of, v.lj, t, s, nullptr, nullptr);
}

template <bool kFp64>
__global__ void __launch_bounds__(kWaveBlock)
k_sweep_lj_wave(const MemberView* __restrict__ views, LjTable t, int32_t n, int32_t gated) {
    const MemberView& v = views[blockIdx.y];
    if (gated && *v.hit == 0) return;
    const int32_t s = (int32_t)((blockIdx.x * kWaveBlock + threadIdx.x) >> 6);
    if (s >= n) return;
    lj_atom_wave<kFp64, false>(v.st, v.cell_of, v.lj, t, s, threadIdx.x & 63, nullptr, nullptr);
}

}  // namespace

hipError_t launch_lj(hipStream_t s, const DevState& st, const int32_t* cell_of, const GridView& g,
                     const LjTable& t, int32_t n, int32_t precision, double* energy_per_atom,
                     const int32_t* gate, double* virial_per_atom) {
    if (energy_per_atom && !virial_per_atom) return hipErrorInvalidValue;
    if (n <= 0) return hipSuccess;
    const bool fp64 = precision == MDX_PREC_FP64;
    if (!g.periodic && n <= kWaveAtoms) {
        dim3 wgrid((n + kWaveBlock / 64 - 1) / (kWaveBlock / 64));
        if (fp64 && energy_per_atom)
            k_lj_wave<true, true><<<wgrid, kWaveBlock, 0, s>>>(st, cell_of, g, t, n, energy_per_atom, virial_per_atom, gate);
        else if (fp64)
            k_lj_wave<true, false><<<wgrid, kWaveBlock, 0, s>>>(st, cell_of, g, t, n, nullptr, nullptr, gate);
        else if (energy_per_atom)
            k_lj_wave<false, true><<<wgrid, kWaveBlock, 0, s>>>(st, cell_of, g, t, n, energy_per_atom, virial_per_atom, gate);
        else
            k_lj_wave<false, false><<<wgrid, kWaveBlock, 0, s>>>(st, cell_of, g, t, n, nullptr, nullptr, gate);
        return hipGetLastError();
    }
    dim3 grid((n + kBlock - 1) / kBlock);
    if (fp64 && energy_per_atom)
        k_lj<true, true><<<grid, kBlock, 0, s>>>(st, cell_of, g, t, n, energy_per_atom, virial_per_atom, gate);
    else if (fp64)
        k_lj<true, false><<<grid, kBlock, 0, s>>>(st, cell_of, g, t, n, nullptr, nullptr, gate);
    else if (energy_per_atom)
        k_lj<false, true><<<grid, kBlock, 0, s>>>(st, cell_of, g, t, n, energy_per_atom, virial_per_atom, gate);
    else
        k_lj<false, false><<<grid, kBlock, 0, s>>>(st, cell_of, g, t, n, nullptr, nullptr, gate);
    return hipGetLastError();
}

hipError_t launch_sweep_lj(hipStream_t s, const MemberView* views, int32_t members, int32_t n, const LjTable& t,
                           int32_t precision, bool gated) {
    if (n <= 0 || members <= 0) return hipSuccess;
    const bool fp64 = precision == MDX_PREC_FP64;
    const int32_t gt = gated ? 1 : 0;
    if (n <= kWaveAtoms) {
        dim3 wgrid((n + kWaveBlock / 64 - 1) / (kWaveBlock / 64), members);
        if (fp64) k_sweep_lj_wave<true><<<wgrid, kWaveBlock, 0, s>>>(views, t, n, gt);
        else k_sweep_lj_wave<false><<<wgrid, kWaveBlock, 0, s>>>(views, t, n, gt);
        return hipGetLastError();
    }
    dim3 grid((n + kBlock - 1) / kBlock, members);
    if (fp64) k_sweep_lj<true><<<grid, kBlock, 0, s>>>(views, t, n, gt);
    else k_sweep_lj<false><<<grid, kBlock, 0, s>>>(views, t, n, gt);
    return hipGetLastError();
}

}  // namespace mdx
