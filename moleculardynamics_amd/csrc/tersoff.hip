// tersoff.hip -- Tersoff silicon (J. Tersoff, PRB 38, 9902 (1988)) in a
// periodic box: the bond-order benchmark system.  The reference has no
// periodic box or Tersoff form; its bond-order driver (manyBodyForces,
// cpu/common/BondOrderFuncs.h:283-397) keeps xi_ij in N x N scratch, which is
// replaced here by per-slot vectors of the compact neighbour list.
//
//   E = sum_i e_i,  e_i = 1/2 sum_{j != i} fc(r_ij) [A e^{-l1 r_ij} - b_ij B e^{-l2 r_ij}]
//   b_ij = (1 + beta^n zeta_ij^n)^{-1/(2n)}
//   zeta_ij = sum_{k != i,j} fc(r_ik) g(theta_ijk) exp(l3^m (r_ij - r_ik)^m)
//   g(theta) = 1 + c^2/d^2 - c^2/(d^2 + (h - cos theta)^2)
//   fc = 1 below R - D, 1/2 - 1/2 sin(pi/2 (r - R)/D) up to R + D, 0 beyond.
//
// e_i depends only on the displacements from i to its neighbours, so each
// atom evaluates its own bond terms once and leaves one vector per list slot,
// w_is = de_i/d(x_j - x_i).  Kernels, one thread per atom (slot order), no
// atomics:
//   k_ts_list:  compact list of the atoms within rs (27 cells, minimum image),
//               column-major [k][slot]; lists are symmetric.
//   k_ts_atoms: e_i and w_is of every atom (tersoff_terms.h: registers for up
//               to kS entries, the list in memory beyond).
// The force F_a = sum_s w_as - sum_s w_{j_s, rev} is gathered by the Gear
// step kernel (gear.hip) or by launch_slot_gather (surrogate.hip).
#include "mdx_internal.h"
#include "tersoff_terms.h"

namespace mdx {

namespace {

using namespace ts;

constexpr int kBlock = 256;

// a / b for 0 <= a < 2^22, b >= 1: float quotient, fixed up by one step
__device__ inline int32_t idiv(int32_t a, int32_t b) {
    int32_t q = (int32_t)((float)a * __builtin_amdgcn_rcpf((float)b));
    q += (q + 1) * b <= a ? 1 : 0;
    q -= q * b > a ? 1 : 0;
    return q;
}

// The 27 neighbour cells as 9 z-columns: cells z-1 .. z+1 of a column are one
// contiguous slot range, or two where the column wraps (the grid has at least
// 3 cells per axis).  The distance test is FP32 on the FP64 minimum-image
// difference, which is exactly antisymmetric, so the lists stay symmetric;
// rl2 is rs^2 enlarged by a few ulp, so the list holds every atom within rs.
__global__ void __launch_bounds__(kBlock)
k_ts_list(Pos3 P, const int32_t* __restrict__ cell_of, GridView g, float rl2, int32_t n, int32_t cap,
          int32_t* __restrict__ list, int32_t* __restrict__ count, int32_t* __restrict__ err, int32_t small) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t g1 = g.g[1], g2 = g.g[2];
    const int32_t c = cell_of[i];
    const int32_t t = small ? idiv(c, g2) : c / g2, cz = c - t * g2;
    const int32_t cx = small ? idiv(t, g1) : t / g1, cy = t - cx * g1;
    const double xi = P.x[i], yi = P.y[i], zi = P.z[i];
    int32_t k = 0;
    for (int dx = -1; dx <= 1; dx++) {
        int32_t x = cx + dx;
        x = x < 0 ? x + g.g[0] : (x >= g.g[0] ? x - g.g[0] : x);
        for (int dy = -1; dy <= 1; dy++) {
            int32_t y = cy + dy;
            y = y < 0 ? y + g1 : (y >= g1 ? y - g1 : y);
            const int32_t col = (x * g1 + y) * g2;
            int32_t rb[2], re[2];
            if (cz == 0) {
                rb[0] = g.start[col]; re[0] = g.start[col + 2];
                rb[1] = g.start[col + g2 - 1]; re[1] = g.start[col + g2];
            } else if (cz == g2 - 1) {
                rb[0] = g.start[col + g2 - 2]; re[0] = g.start[col + g2];
                rb[1] = g.start[col]; re[1] = g.start[col + 1];
            } else {
                rb[0] = g.start[col + cz - 1]; re[0] = g.start[col + cz + 2];
                rb[1] = re[1] = 0;
            }
#pragma unroll
            for (int r = 0; r < 2; r++)
                for (int32_t j = rb[r]; j < re[r]; j++) {
                    if (j == i) continue;
                    const float ex = (float)min_image(P.x[j] - xi, P.L, P.invL);
                    const float ey = (float)min_image(P.y[j] - yi, P.L, P.invL);
                    const float ez = (float)min_image(P.z[j] - zi, P.L, P.invL);
                    if (fmaf(ex, ex, fmaf(ey, ey, ez * ez)) < rl2) {
                        if (k < cap) list[(int64_t)k * n + i] = j;
                        k++;
                    }
                }
        }
    }
    count[i] = k < cap ? k : cap;
    if (k > cap) atomicMax(err, k);
}

constexpr int kS = 6;   // list entries held in registers

inline int grid_for(int32_t n) { return (n + kBlock - 1) / kBlock; }

template <typename T, typename W4>
__global__ void __launch_bounds__(kBlock)
k_ts_atoms(Pos3 P, Tp<T> p, const int32_t* __restrict__ list, const int32_t* __restrict__ count, int32_t n,
           W4* __restrict__ w, double* __restrict__ e_atom) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t cnt = count[i];
    double e;
    if (cnt <= kS) {
        T wr[kS][3];
        e = centre_reg<T, kS>(P, p, list, n, i, cnt, wr);
#pragma unroll
        for (int s = 0; s < kS; s++)
            if (s < cnt) {
                W4 v;
                v.x = wr[s][0];
                v.y = wr[s][1];
                v.z = wr[s][2];
                v.w = 0;
                w[(int64_t)s * n + i] = v;
            }
    } else {
        e = centre_mem<T>(P, p, list, n, i, cnt, w);
    }
    if (e_atom) e_atom[i] = e;
}

}  // namespace

TersoffParams tersoff_si_params() {
    TersoffParams p;
    p.A = 1830.8; p.B = 471.18; p.lambda1 = 2.4799; p.lambda2 = 1.7322; p.lambda3 = 0.0;
    p.beta = 1.1e-6; p.n = 0.78734;
    p.c2 = 100390.0 * 100390.0; p.d2 = 16.217 * 16.217; p.h = -0.59825;
    p.R = 2.7; p.D = 0.3; p.m = 3;
    return p;
}

hipError_t launch_ts_list(hipStream_t s, const DevState& st, const int32_t* cell_of, const GridView& g, double rs,
                          int32_t n, int32_t cap, int32_t* list, int32_t* count, int32_t* err) {
    if (n <= 0) return hipSuccess;
    const float rl2 = (float)(rs * rs * (1.0 + 1e-6));
    const int32_t small = (int64_t)g.g[0] * g.g[1] * g.g[2] < (1 << 22) ? 1 : 0;
    k_ts_list<<<grid_for(n), kBlock, 0, s>>>(pos3(st, g.L), cell_of, g, rl2, n, cap, list, count, err, small);
    return hipGetLastError();
}

// compact [k][slot] list of slots -> rows of caller indices (parity export)
__global__ void __launch_bounds__(kBlock)
k_compact_decode(const int32_t* __restrict__ list, const int32_t* __restrict__ count, int32_t n,
                 const int32_t* __restrict__ id, int32_t width, int32_t* __restrict__ rows,
                 int32_t* __restrict__ counts, int32_t* __restrict__ err) {
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int32_t c = count[i], me = id[i];
    counts[me] = c;
    if (c > width) {
        atomicMax(err, c);
        return;
    }
    for (int32_t k = 0; k < c; k++) rows[(int64_t)me * width + k] = id[list[(int64_t)k * n + i]];
}

hipError_t launch_compact_decode(hipStream_t s, const int32_t* list, const int32_t* count, int32_t n,
                                 const int32_t* id, int32_t width, int32_t* rows, int32_t* counts, int32_t* err) {
    if (n <= 0) return hipSuccess;
    k_compact_decode<<<grid_for(n), kBlock, 0, s>>>(list, count, n, id, width, rows, counts, err);
    return hipGetLastError();
}

hipError_t launch_ts_atoms(hipStream_t s, const DevState& st, double L, const TersoffParams& p, int32_t precision,
                           const int32_t* list, const int32_t* count, int32_t n, float4* w, double4* wd,
                           double* e_atom) {
    if (n <= 0) return hipSuccess;
    const Pos3 P = pos3(st, L);
    if (precision == MDX_PREC_FP64)
        k_ts_atoms<double, double4><<<grid_for(n), kBlock, 0, s>>>(P, cast_params<double>(p), list, count, n, wd, e_atom);
    else
        k_ts_atoms<float, float4><<<grid_for(n), kBlock, 0, s>>>(P, cast_params<float>(p), list, count, n, w, e_atom);
    return hipGetLastError();
}

}  // namespace mdx
