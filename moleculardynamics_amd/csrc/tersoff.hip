// tersoff.hip -- Tersoff silicon (J. Tersoff, PRB 38, 9902 (1988)) in a
// periodic box: the bond-order benchmark system.  The reference has no
// periodic box or Tersoff form; its bond-order driver (manyBodyForces,
// cpu/common/BondOrderFuncs.h:283-397) keeps xi_ij in N x N scratch, which is
// replaced here by per-bond arrays of the compact neighbour list.
//
//   E = 1/2 sum_i sum_{j != i} fc(r_ij) [A e^{-l1 r_ij} - b_ij B e^{-l2 r_ij}]
//   b_ij = (1 + beta^n zeta_ij^n)^{-1/(2n)}
//   zeta_ij = sum_{k != i,j} fc(r_ik) g(theta_ijk) exp(l3^m (r_ij - r_ik)^m)
//   g(theta) = 1 + c^2/d^2 - c^2/(d^2 + (h - cos theta)^2)
//   fc = 1 below R - D, 1/2 - 1/2 sin(pi/2 (r - R)/D) up to R + D, 0 beyond.
//
// Three kernels, one thread per atom (slot order), no atomics:
//   k_ts_list:  compact list of the atoms within rs (27 cells, minimum image),
//               column-major [k][slot]; lists are symmetric.
//   k_ts_bond:  per bond (i, slot a): zeta_ij, b_ij; stores dE/dr_ij at fixed
//               zeta and dE/dzeta_ij; per-atom energy; counts atoms that moved
//               past the list margin.
//   k_ts_force: the force on atom x gathered from every bond term that moves
//               x: as the centre i, as the partner j (bond m -> x) and as the
//               third atom k (bond m -> j, k = x), reading the stored bond
//               derivatives of x's neighbours.
#include "mdx_internal.h"
#include "tersoff_terms.h"

namespace mdx {

namespace {

using namespace ts;

constexpr int kBlock = 256;

__global__ void __launch_bounds__(kBlock)
k_ts_list(Pos3 P, const int32_t* __restrict__ cell_of, GridView g, double rs2, int32_t n, int32_t cap,
          int32_t* __restrict__ list, int32_t* __restrict__ count, int32_t* __restrict__ err) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t c = cell_of[i];
    const int32_t cz = c % g.g[2], cy = (c / g.g[2]) % g.g[1], cx = c / (g.g[1] * g.g[2]);
    int32_t k = 0;
    for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dz = -1; dz <= 1; dz++) {
                int32_t x = cx + dx, y = cy + dy, z = cz + dz;
                x = x < 0 ? x + g.g[0] : (x >= g.g[0] ? x - g.g[0] : x);
                y = y < 0 ? y + g.g[1] : (y >= g.g[1] ? y - g.g[1] : y);
                z = z < 0 ? z + g.g[2] : (z >= g.g[2] ? z - g.g[2] : z);
                const int32_t cc = (x * g.g[1] + y) * g.g[2] + z;
                for (int32_t j = g.start[cc]; j < g.start[cc + 1]; j++) {
                    if (j == i) continue;
                    double d[3];
                    P.d(i, j, d);
                    if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < rs2) {
                        if (k < cap) list[(int64_t)k * n + i] = j;
                        k++;
                    }
                }
            }
    count[i] = k < cap ? k : cap;
    if (k > cap) atomicMax(err, k);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_ts_bond(Pos3 P, Tp<T> p, const int32_t* __restrict__ list, const int32_t* __restrict__ count, int32_t n,
          double* __restrict__ dedr, double* __restrict__ dedz, double* __restrict__ e_atom,
          const double* __restrict__ r0x, const double* __restrict__ r0y, const double* __restrict__ r0z,
          double hard2, unsigned long long* __restrict__ flags, int32_t spread) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool stale = false;
    if (i < n) {
        const int32_t cnt = count[i];
        double e = 0;
        for (int32_t a = 0; a < cnt; a++) {
            const int32_t j = list[(int64_t)a * n + i];
            double dj[3];
            P.d(i, j, dj);
            const T rij = (T)sqrt(dj[0] * dj[0] + dj[1] * dj[1] + dj[2] * dj[2]);
            double de_r = 0, de_z = 0;
            if (rij < p.rc) {
                T zeta = 0;
                for (int32_t b = 0; b < cnt; b++) {
                    if (b == a) continue;
                    const int32_t kk = list[(int64_t)b * n + i];
                    double dk[3];
                    P.d(i, kk, dk);
                    const T rik = (T)sqrt(dk[0] * dk[0] + dk[1] * dk[1] + dk[2] * dk[2]);
                    if (!(rik < p.rc)) continue;
                    T fk, dfk, gg, dg, ee, dee;
                    fcut(p, rik, fk, dfk);
                    const T cs = (T)((dj[0] * dk[0] + dj[1] * dk[1] + dj[2] * dk[2]) / ((double)rij * (double)rik));
                    gang(p, cs, gg, dg);
                    ex3(p, rij - rik, ee, dee);
                    zeta += fk * gg * ee;
                }
                T fc, dfc;
                fcut(p, rij, fc, dfc);
                const T fr = p.A * t_exp<T>(-p.l1 * rij), dfr = -p.l1 * fr;
                const T fa = -p.B * t_exp<T>(-p.l2 * rij), dfa = -p.l2 * fa;
                T bz = 1, dbz = 0;
                if (zeta > 0) {
                    const T t = t_pow<T>(p.beta * zeta, p.n);
                    bz = t_pow<T>(1 + t, (T)-0.5 / p.n);
                    dbz = (T)-0.5 * bz / (1 + t) * t / zeta;
                }
                e += 0.5 * (double)fc * ((double)fr + (double)bz * (double)fa);
                de_r = 0.5 * (double)(dfc * (fr + bz * fa) + fc * (dfr + bz * dfa));
                de_z = 0.5 * (double)(fc * fa * dbz);
            }
            dedr[(int64_t)a * n + i] = de_r;
            dedz[(int64_t)a * n + i] = de_z;
        }
        if (e_atom) e_atom[i] = e;
        if (r0x) {
            const double e0 = P.x[i] - r0x[i], e1 = P.y[i] - r0y[i], e2 = P.z[i] - r0z[i];
            stale = e0 * e0 + e1 * e1 + e2 * e2 >= hard2;
        }
    }
    if (flags) {
        const unsigned long long nst = __popcll(__ballot(stale));
        if ((threadIdx.x & (kWave - 1)) == 0 && nst)
            atomicAdd(&flags[3 * ((blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave) % spread) + 2], nst);
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_ts_force(Pos3 P, Tp<T> p, const int32_t* __restrict__ list, const int32_t* __restrict__ count, int32_t n,
           const double* __restrict__ dedr, const double* __restrict__ dedz, double* __restrict__ fx,
           double* __restrict__ fy, double* __restrict__ fz) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int32_t cnt = count[x];
    double F[3] = {0, 0, 0};
    // (A) x is the centre of bond x -> j
    for (int32_t a = 0; a < cnt; a++) {
        const int32_t j = list[(int64_t)a * n + x];
        double dj[3];
        P.d(x, j, dj);
        T rij, u[3];
        unit(dj, rij, u);
        if (!(rij < p.rc)) continue;
        const double de_r = dedr[(int64_t)a * n + x], de_z = dedz[(int64_t)a * n + x];
        double gi[3];
        for (int q = 0; q < 3; q++) gi[q] = -de_r * (double)u[q];
        if (de_z != 0) {
            for (int32_t b = 0; b < cnt; b++) {
                if (b == a) continue;
                double dk[3];
                P.d(x, list[(int64_t)b * n + x], dk);
                T rik, w[3];
                unit(dk, rik, w);
                if (!(rik < p.rc)) continue;
                const ZTerm<T> z = zterm(p, u, rij, w, rik);
                for (int q = 0; q < 3; q++) {
                    const T dcs_dj = (w[q] - z.cs * u[q]) / rij;
                    const T dcs_dk = (u[q] - z.cs * w[q]) / rik;
                    const double gj = de_z * (double)(z.d_rij * u[q] + z.d_cs * dcs_dj);
                    const double gk = de_z * (double)(z.d_rik * w[q] + z.d_cs * dcs_dk);
                    gi[q] -= gj + gk;
                }
            }
        }
        for (int q = 0; q < 3; q++) F[q] -= gi[q];
    }
    // (B) x is the partner of bond m -> x; (C) x is the third atom of bond m -> j
    for (int32_t a = 0; a < cnt; a++) {
        const int32_t m = list[(int64_t)a * n + x];
        double dmx[3];
        P.d(m, x, dmx);
        T rmx, w[3];
        unit(dmx, rmx, w);
        if (!(rmx < p.rc)) continue;
        const int32_t cm = count[m];
        for (int32_t b = 0; b < cm; b++) {
            const int32_t j = list[(int64_t)b * n + m];
            if (j == x) {
                // bond m -> x: gj = dE/dr u + sum_k dE/dzeta (d_rij u + d_cs dcos/dx_j)
                const double de_r = dedr[(int64_t)b * n + m], de_z = dedz[(int64_t)b * n + m];
                double gj[3];
                for (int q = 0; q < 3; q++) gj[q] = de_r * (double)w[q];
                if (de_z != 0) {
                    for (int32_t c = 0; c < cm; c++) {
                        if (c == b) continue;
                        double dk[3];
                        P.d(m, list[(int64_t)c * n + m], dk);
                        T rik, v[3];
                        unit(dk, rik, v);
                        if (!(rik < p.rc)) continue;
                        const ZTerm<T> z = zterm(p, w, rmx, v, rik);
                        for (int q = 0; q < 3; q++) {
                            const T dcs_dj = (v[q] - z.cs * w[q]) / rmx;
                            gj[q] += de_z * (double)(z.d_rij * w[q] + z.d_cs * dcs_dj);
                        }
                    }
                }
                for (int q = 0; q < 3; q++) F[q] -= gj[q];
            } else {
                // bond m -> j with x as the third atom: gk = dE/dzeta (d_rik w + d_cs dcos/dx_k)
                const double de_z = dedz[(int64_t)b * n + m];
                if (de_z == 0) continue;
                double dmj[3];
                P.d(m, j, dmj);
                T rmj, u[3];
                unit(dmj, rmj, u);
                if (!(rmj < p.rc)) continue;
                const ZTerm<T> z = zterm(p, u, rmj, w, rmx);
                for (int q = 0; q < 3; q++) {
                    const T dcs_dk = (u[q] - z.cs * w[q]) / rmx;
                    F[q] -= de_z * (double)(z.d_rik * w[q] + z.d_cs * dcs_dk);
                }
            }
        }
    }
    fx[x] = F[0];
    fy[x] = F[1];
    fz[x] = F[2];
}

inline int grid_for(int32_t n) { return (n + kBlock - 1) / kBlock; }

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
    k_ts_list<<<grid_for(n), kBlock, 0, s>>>(pos3(st, g.L), cell_of, g, rs * rs, n, cap, list, count, err);
    return hipGetLastError();
}

hipError_t launch_ts_forces(hipStream_t s, const DevState& st, double L, const TersoffParams& p, int32_t precision,
                            const int32_t* list, const int32_t* count, int32_t n, double* dedr, double* dedz,
                            double* e_atom, double* const rref[3], double hard2, uint64_t* flags,
                            const DevState& out) {
    if (n <= 0) return hipSuccess;
    const Pos3 P = pos3(st, L);
    const double* r0[3] = {rref ? rref[0] : nullptr, rref ? rref[1] : nullptr, rref ? rref[2] : nullptr};
    auto* fl = reinterpret_cast<unsigned long long*>(flags);
    if (precision == MDX_PREC_FP64) {
        const Tp<double> q = cast_params<double>(p);
        k_ts_bond<double><<<grid_for(n), kBlock, 0, s>>>(P, q, list, count, n, dedr, dedz, e_atom, r0[0], r0[1],
                                                         r0[2], hard2, fl, flag_spread());
        k_ts_force<double><<<grid_for(n), kBlock, 0, s>>>(P, q, list, count, n, dedr, dedz, out.f[0], out.f[1],
                                                          out.f[2]);
    } else {
        const Tp<float> q = cast_params<float>(p);
        k_ts_bond<float><<<grid_for(n), kBlock, 0, s>>>(P, q, list, count, n, dedr, dedz, e_atom, r0[0], r0[1],
                                                        r0[2], hard2, fl, flag_spread());
        k_ts_force<float><<<grid_for(n), kBlock, 0, s>>>(P, q, list, count, n, dedr, dedz, out.f[0], out.f[1],
                                                         out.f[2]);
    }
    return hipGetLastError();
}

}  // namespace mdx
