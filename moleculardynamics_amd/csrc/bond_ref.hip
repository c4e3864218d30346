// bond_ref.hip -- the reference's N/O bond-order force
// (cpu/common/BondOrderFuncs.h:12-397 with the pair terms of
// common/arvdwfc.cpp:16-59 and the constants of common/Constants.h:36-79).
//
// Two gather kernels, one thread per bond-order atom, no atomics:
//  pass 1 (BondOrderFuncs.h:296-337): pair + bond-order force on i from every
//    candidate j within the cutoff; xi_ij / xi_ji are stored per (i, j) in a
//    compact per-atom table instead of the reference's N x N scratch.
//  pass 2 (BondOrderFuncs.h:339-393): three-body force on k from candidate
//    pairs (i, j) of k's row, reading the xi values of pass 1.
// Candidates are the 27-cell neighbourhood of the atom's cell in the
// bond-order grid, visited in the reference's row order.  The arithmetic
// type T is float (mixed precision) or double (FP64); differences of FP64
// positions are formed in FP64 before the conversion.
#include "mdx_internal.h"

namespace mdx {

namespace {

constexpr int kBlock = 128;
constexpr int kNe = 2;   // species code of Ne (common/Constants.h:15)

template <class T> struct K;
template <> struct K<double> {
    static constexpr double R0 = (double)1.3f, ALFA = 2.0, LAMDA3 = (double)1.33f;
    static constexpr double C = 25.0, D = (double)4.0417f, H = 0.0, GM = (double)0.3f, ALFAV = 3.0;
    static constexpr double ALFAVSQ = 9.0, E7 = (double)7.0e-8f, VDW = (double)1.0e-4f;
    static constexpr double COD = (double)(25.f / 4.0417f);
};
template <> struct K<float> {
    static constexpr float R0 = 1.3f, ALFA = 2.0f, LAMDA3 = 1.33f;
    static constexpr float C = 25.0f, D = 4.0417f, H = 0.0f, GM = 0.3f, ALFAV = 3.0f;
    static constexpr float ALFAVSQ = 9.0f, E7 = 7.0e-8f, VDW = 1.0e-4f;
    static constexpr float COD = 25.f / 4.0417f;
};
// vdW numerator constant 1 + 50 GM + GM^2 and 10 (GM - 1), float-evaluated as in arvdwfc.cpp:45
__device__ constexpr float kVdwP0 = 1.0f + 50.0f * 0.3f + 0.09f;
__device__ constexpr float kVdwK10 = 10.0f * (0.3f - 1.0f);

__device__ inline double d_exp(double x) { return exp(x); }
__device__ inline float d_exp(float x) { return expf(x); }
__device__ inline double d_tanh(double x) { return tanh(x); }
__device__ inline float d_tanh(float x) { return tanhf(x); }
__device__ inline double d_sqrt(double x) { return sqrt(x); }
__device__ inline float d_sqrt(float x) { return sqrtf(x); }

template <class T> __device__ inline void cutoff(T r, T& fc, T& dfc) {
    T t = d_tanh(K<T>::ALFA * (r - K<T>::R0));
    dfc = (T)-0.5 * K<T>::ALFA * ((T)1 - t * t);
    fc = (T)0.5 * ((T)1 - t);
}
// V_A with dva = -dV_A/dr; S = 2 so S De/(S-1) = 2 De and beta sqrt(2/S) = beta
template <class T> __device__ inline void attractive(T r, T de, T beta, T re, T& va, T& dva) {
    va = (T)2 * de * d_exp(-beta * (r - re));
    dva = beta * va;
}
// V_R = De exp(-2 beta (r - re)) + 7e-8 exp(-10 beta (r - re)), dvr = 2 beta V_R
template <class T> __device__ inline void repulsive(T r, T de, T beta, T re, T& vr, T& dvr) {
    vr = de * d_exp((T)-2 * beta * (r - re)) + K<T>::E7 * d_exp((T)-10 * beta * (r - re));
    dvr = (T)2 * beta * vr;
}
template <class T> __device__ inline void vanderwaals(T r, T r0v, T& vdw, T& dvdw) {
    T x = K<T>::ALFAV * (r - r0v);
    T ep = d_exp(x), em = K<T>::GM * d_exp(-x);
    T a = ep + em, ap = K<T>::ALFAV * (ep - em), app = K<T>::ALFAVSQ * a;
    T num = (T)kVdwP0 + (T)kVdwK10 * ap;
    dvdw = -K<T>::VDW * (((T)kVdwK10 * app * a - (T)2 * ap * num) / (a * a * a));
    vdw = K<T>::VDW * num / (a * a);
}
template <class T> __device__ inline void ang(T cs, T& g, T& dg) {
    T u = K<T>::H - cs;
    T den = K<T>::D + u * u;
    g = (T)1 + K<T>::COD - K<T>::C / den;
    dg = (T)-2 * K<T>::C * u / (den * den);
}

// pair kind: 0 = N-N, 1 = O-O, 2 = unlike (common/Getters.cpp:24-78)
__device__ inline int pair_kind(int ti, int tj) { return ti == tj ? (ti == 0 ? 0 : 1) : 2; }

template <class T> struct V3 { T x, y, z; };
template <class T> __device__ inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> __device__ inline V3<T> operator+(V3<T> a, V3<T> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> __device__ inline V3<T> operator-(V3<T> a, V3<T> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> __device__ inline V3<T> operator*(V3<T> a, T s) { return mk(a.x * s, a.y * s, a.z * s); }
template <class T> __device__ inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

struct Ctx {
    DevState st;
    const int32_t* slot_of_id;
    const int32_t* cell_of;   // bond-order grid cell per id
    const int32_t* ids;       // grid members (ids), cell-sorted
    GridView g;
};

// FP64 difference r_a - r_b of two atoms given by id, converted to T;
// *r2 (optional) receives the FP64 squared distance for cutoff tests
template <class T> __device__ inline V3<T> diff(const Ctx& c, int32_t a, int32_t b,
                                                 double* r2 = nullptr) {
    int32_t sa = c.slot_of_id[a], sb = c.slot_of_id[b];
    double dx = c.st.r[0][sa] - c.st.r[0][sb], dy = c.st.r[1][sa] - c.st.r[1][sb];
    double dz = c.st.r[2][sa] - c.st.r[2][sb];
    if (r2) *r2 = dx * dx + dy * dy + dz * dz;
    return mk((T)dx, (T)dy, (T)dz);
}

// visit the candidate row of atom `self` in the reference's order
template <class F> __device__ inline void for_row(const Ctx& c, int32_t self, F&& fn) {
    int32_t own = c.cell_of[self];
    int32_t cz = own % c.g.g[2], cy = (own / c.g.g[2]) % c.g.g[1], cx = own / (c.g.g[2] * c.g.g[1]);
    for (int dx = -1; dx <= 1; dx++) {
        int32_t x = cx + dx;
        if (x < 0 || x >= c.g.cells) continue;
        for (int dy = -1; dy <= 1; dy++) {
            int32_t y = cy + dy;
            if (y < 0 || y >= c.g.cells) continue;
            for (int dz = -1; dz <= 1; dz++) {
                int32_t z = cz + dz;
                if (z < 0 || z >= c.g.cells) continue;
                int32_t cell = (x * c.g.g[1] + y) * c.g.g[2] + z;
                for (int32_t p = c.g.start[cell]; p < c.g.start[cell + 1]; p++) {
                    int32_t m = c.ids[p];
                    if (m != self) fn(m);
                }
            }
        }
    }
}

__device__ inline int32_t type_of(const Ctx& c, int32_t id) { return c.st.type[c.slot_of_id[id]]; }

// pass 1 for bond-order atom i (thread per atom)
template <class T>
__device__ inline void bond_pass1_atom(const Ctx& c, const BondRefParams& p, int32_t i, int32_t cap, int32_t assign,
                                       int32_t* __restrict__ xi_j, double* __restrict__ xi_v,
                                       double* __restrict__ e_atom) {
    const int32_t ti = type_of(c, i);
    double Fx = 0, Fy = 0, Fz = 0, e = 0;
    int32_t nx = 0;
    for_row(c, i, [&](int32_t j) {
        double r2d;
        V3<T> ij = diff<T>(c, i, j, &r2d);
        if (r2d >= p.rcut2 || r2d > p.pair_cut2) return;
        const int32_t tj = type_of(c, j);
        if (tj == kNe) return;
        const T rij = d_sqrt((T)r2d);
        const int kd = pair_kind(ti, tj);
        T va = 0, dva = 0, vr = 0, dvr = 0, vdw, dvdw, bo = 0;
        V3<T> dbo = mk<T>(0, 0, 0);
        T fcut, dfcut;
        cutoff<T>(rij / (T)10, fcut, dfcut);
        if (fcut > 0) {
            const T re = (T)p.re[kd], de = (T)p.de[kd], be = (T)p.beta[kd];
            attractive<T>(rij, de, be, re, va, dva);
            repulsive<T>(rij, de, be, re, vr, dvr);
            // bondOrder1 (BondOrderFuncs.h:12-143)
            const V3<T> eij = ij * ((T)1 / rij);
            const T ex = d_exp(K<T>::LAMDA3 * (rij - re));
            T xi = 0, xj = 0;
            V3<T> dxi = mk<T>(0, 0, 0), dxj = mk<T>(0, 0, 0);
            for_row(c, i, [&](int32_t m) {
                if (m == j || type_of(c, m) == kNe) return;
                double rim2;
                V3<T> im = diff<T>(c, i, m, &rim2);
                if (rim2 >= p.rcut2) return;
                T rim = d_sqrt((T)rim2);
                V3<T> jm = diff<T>(c, j, m);
                T rjm = d_sqrt(dot(jm, jm));
                T fci, dfci, fcj, dfcj;
                cutoff<T>(rim, fci, dfci);
                cutoff<T>(rjm, fcj, dfcj);
                V3<T> eim = im * ((T)1 / rim);
                if (fci > 0) {
                    T cs = (rij * rij + rim * rim - rjm * rjm) / ((T)2 * rij * rim);
                    T g, dg;
                    ang<T>(cs, g, dg);
                    xi += fci * g * ex;
                    V3<T> dcs = (eim - eij * cs) * ((T)1 / rij) + (eij - eim * cs) * ((T)1 / rim);
                    dxi = dxi + eim * (dfci * g * ex) + dcs * (fci * dg * ex) + eij * (fci * g * K<T>::LAMDA3 * ex);
                }
                if (fcj > 0) {
                    T cs = (rij * rij + rjm * rjm - rim * rim) / ((T)2 * rij * rjm);
                    T g, dg;
                    ang<T>(cs, g, dg);
                    xj += fcj * g * ex;
                    V3<T> dcs = (jm * ((T)-1 / rjm) - eij * cs) * ((T)1 / rij);
                    dxj = dxj + dcs * (fcj * dg * ex) + eij * (fcj * g * K<T>::LAMDA3 * ex);
                }
            });
            T qi = (T)1 / (((T)1 + xi) * ((T)1 + xi)), qj = (T)1 / (((T)1 + xj) * ((T)1 + xj));
            bo = (T)0.5 * ((T)1 / ((T)1 + xi) + (T)1 / ((T)1 + xj));
            dbo = (dxi * qi + dxj * qj) * (T)0.5;
            if (nx < cap) {
                int64_t q = (int64_t)i * cap + nx;
                xi_j[q] = j;
                xi_v[2 * q] = (double)xi;
                xi_v[2 * q + 1] = (double)xj;
            }
            nx++;
        }
        vanderwaals<T>(rij, (T)p.r0v[kd], vdw, dvdw);
        T fij = dvr - bo * dva + dvdw;
        T s = fij / rij;
        Fx += (double)(ij.x * s - dbo.x * va);
        Fy += (double)(ij.y * s - dbo.y * va);
        Fz += (double)(ij.z * s - dbo.z * va);
        e += 0.5 * (double)((vr - bo * va) + vdw);
    });
    if (nx < cap) xi_j[(int64_t)i * cap + nx] = -1;
    const int32_t si = c.slot_of_id[i];
    if (assign) {
        c.st.f[0][si] = Fx; c.st.f[1][si] = Fy; c.st.f[2][si] = Fz;
    } else {
        c.st.f[0][si] += Fx; c.st.f[1][si] += Fy; c.st.f[2][si] += Fz;
    }
    if (e_atom) e_atom[i] = e;
}

template <class T>
__global__ void __launch_bounds__(kBlock)
k_bond_pass1(Ctx c, BondRefParams p, int32_t n_bond, int32_t cap, int32_t assign,
             int32_t* __restrict__ xi_j, double* __restrict__ xi_v, double* __restrict__ e_atom,
             const int32_t* __restrict__ gate) {
    if (gate && *gate == 0) return;
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_bond) return;
    bond_pass1_atom<T>(c, p, i, cap, assign, xi_j, xi_v, e_atom);
}

// pass 2 for bond-order atom k (thread per atom)
template <class T>
__device__ inline void bond_pass2_atom(const Ctx& c, const BondRefParams& p, int32_t k, int32_t cap,
                                       const int32_t* __restrict__ xi_j, const double* __restrict__ xi_v) {
    double Fx = 0, Fy = 0, Fz = 0;
    for_row(c, k, [&](int32_t i) {
        double rik2d;
        const V3<T> ik = diff<T>(c, k, i, &rik2d);
        if (rik2d >= p.rcut2) return;
        const T rik = d_sqrt((T)rik2d);
        const int32_t ti = type_of(c, i);
        for_row(c, k, [&](int32_t j) {
            if (j == i) return;
            double r2d;
            diff<T>(c, i, j, &r2d);
            if (r2d >= p.rcut2 || r2d > p.pair_cut2) return;
            const T rij = d_sqrt((T)r2d);
            const int kd = pair_kind(ti, type_of(c, j));
            const T re = (T)p.re[kd], de = (T)p.de[kd], be = (T)p.beta[kd];
            T va, dva;
            attractive<T>(rij, de, be, re, va, dva);
            T xi = 0, xj = 0;
            for (int32_t q = 0; q < cap; q++) {
                int32_t jj = xi_j[(int64_t)i * cap + q];
                if (jj < 0) break;
                if (jj == j) {
                    xi = (T)xi_v[2 * ((int64_t)i * cap + q)];
                    xj = (T)xi_v[2 * ((int64_t)i * cap + q) + 1];
                    break;
                }
            }
            // bondOrder2 (BondOrderFuncs.h:148-237): -db_ij/dx_k, halves included
            V3<T> kj = diff<T>(c, k, j);
            T rjk = d_sqrt(dot(kj, kj));
            V3<T> ejk = kj * ((T)1 / rjk);
            V3<T> eik = ik * ((T)1 / rik);
            V3<T> uij = diff<T>(c, j, i) * ((T)1 / rij);
            T ex = d_exp(K<T>::LAMDA3 * (rij - re));
            T fci, dfci, fcj, dfcj;
            cutoff<T>(rik, fci, dfci);
            cutoff<T>(rjk, fcj, dfcj);
            V3<T> d = mk<T>(0, 0, 0);
            if (fci > 0) {
                T cs = (rij * rij + rik * rik - rjk * rjk) / ((T)2 * rij * rik);
                T g, dg;
                ang<T>(cs, g, dg);
                V3<T> dcs = (uij - eik * cs) * ((T)1 / rik);
                V3<T> dx_ = eik * (dfci * g * ex) + dcs * (fci * dg * ex);
                d = d + dx_ * ((T)0.5 / (((T)1 + xi) * ((T)1 + xi)));
            }
            if (fcj > 0) {
                T cs = (rij * rij + rjk * rjk - rik * rik) / ((T)2 * rij * rjk);
                T g, dg;
                ang<T>(cs, g, dg);
                V3<T> dcs = (uij * (T)-1 - ejk * cs) * ((T)1 / rjk);
                V3<T> dx_ = ejk * (dfcj * g * ex) + dcs * (fcj * dg * ex);
                d = d + dx_ * ((T)0.5 / (((T)1 + xj) * ((T)1 + xj)));
            }
            Fx -= (double)((T)0.5 * d.x * va);
            Fy -= (double)((T)0.5 * d.y * va);
            Fz -= (double)((T)0.5 * d.z * va);
        });
    });
    const int32_t sk = c.slot_of_id[k];
    c.st.f[0][sk] += Fx;
    c.st.f[1][sk] += Fy;
    c.st.f[2][sk] += Fz;
}

template <class T>
__global__ void __launch_bounds__(kBlock)
k_bond_pass2(Ctx c, BondRefParams p, int32_t n_bond, int32_t cap, const int32_t* __restrict__ xi_j,
             const double* __restrict__ xi_v, const int32_t* __restrict__ gate) {
    if (gate && *gate == 0) return;
    int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_bond) return;
    bond_pass2_atom<T>(c, p, k, cap, xi_j, xi_v);
}

// ---- one wavefront per bond-order atom (few atoms, short rows) ----
//
// The atom's candidate row (9 contiguous runs of the cell-sorted member
// array, in the reference's row order) is staged once into LDS with FP64
// positions, so the nested loops over the row read LDS instead of chasing
// ids through memory.  The outer loops (j in pass 1, i in pass 2) run in row
// order over the ballot of the lanes whose entry qualifies; the inner loop
// over the row is spread over the lanes and summed by a fixed butterfly, so
// results are deterministic and differ from the thread-per-atom kernels only
// in the association of those inner sums.
constexpr int kWaveBlock = 256;
constexpr int kRowMax = 256;   // row entries staged per wavefront

struct RowEntry {
    double x, y, z;
    int32_t id;     // -1 for the atom itself (not a candidate)
    int32_t type;
};

template <class T> __device__ inline T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// stage the row of `self` into row[0..len), entries first, first + stride, ...
// (lane: the caller's lane in its wavefront); the host guarantees len <= kRowMax
__device__ inline int32_t stage_row(const Ctx& c, int32_t self, RowEntry* row, int lane, int32_t first,
                                    int32_t stride) {
    const int32_t own = c.cell_of[self];
    const int32_t cz = own % c.g.g[2], cy = (own / c.g.g[2]) % c.g.g[1], cx = own / (c.g.g[2] * c.g.g[1]);
    int32_t rb = 0, rn = 0;
    if (lane < 9) {
        const int32_t x = cx + lane / 3 - 1, y = cy + lane % 3 - 1;
        if (x >= 0 && x < c.g.cells && y >= 0 && y < c.g.cells) {
            const int32_t z0 = max(cz - 1, 0), z1 = min(cz + 1, c.g.cells - 1);
            if (z0 <= z1) {
                const int32_t col = (x * c.g.g[1] + y) * c.g.g[2];
                rb = c.g.start[col + z0];
                rn = c.g.start[col + z1 + 1] - rb;
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
    len = min(len, kRowMax);
    for (int32_t q = first; q < len; q += stride) {
        int32_t p = 0;
#pragma unroll
        for (int r = 0; r < 9; r++)
            if (q >= off[r]) p = base[r] + (q - off[r]);
        const int32_t m = c.ids[p];
        const int32_t sm = c.slot_of_id[m];
        row[q] = RowEntry{c.st.r[0][sm], c.st.r[1][sm], c.st.r[2][sm], m == self ? -1 : m, c.st.type[sm]};
    }
    return len;
}

// One workgroup of kAtomWaves wavefronts per bond-order atom: the row is
// staged once, and the atom's outer loop (over j in pass 1, over i in pass 2)
// is dealt to the wavefronts round robin in row order, so each wavefront's
// chain of dependent FP64 work is a quarter as long (with one wavefront per
// atom a sweep's few hundred bond-order atoms left the GPU waiting on them).
constexpr int kAtomWaves = kWaveBlock / 64;

// pass 1 of bond-order atom i by its workgroup, whose staged row is
// row[0..len).  Every wavefront walks the same candidates j in row order;
// candidate o is evaluated by wavefront o % kAtomWaves, which parks its force
// and energy terms in part[o] (and writes its xi table entry, numbered among
// the j's with fcut > 0 as the reference's loop numbers them).  Thread 0 then
// adds the terms in row order: the additions of the one-wavefront loop.
template <class T>
__device__ inline void bond_wave1_block(const Ctx& c, const BondRefParams& p, int32_t i, int32_t cap, int32_t assign,
                                        int32_t* __restrict__ xi_j, double* __restrict__ xi_v,
                                        double* __restrict__ e_atom, const RowEntry* row, int32_t len,
                                        double (*part)[4]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int32_t si = c.slot_of_id[i];
    const double px = c.st.r[0][si], py = c.st.r[1][si], pz = c.st.r[2][si];
    const int32_t ti = c.st.type[si];
    int32_t ord = 0, nx = 0;   // candidates so far; those with fcut > 0
    for (int32_t q0 = 0; q0 < len; q0 += 64) {
        bool cand = false, live = false;
        if (q0 + lane < len) {
            const RowEntry a = row[q0 + lane];
            const double dx = px - a.x, dy = py - a.y, dz = pz - a.z;
            const double r2d = dx * dx + dy * dy + dz * dz;
            cand = a.id >= 0 && !(r2d >= p.rcut2 || r2d > p.pair_cut2) && a.type != kNe;
            if (cand) {
                T fc, dfc;
                cutoff<T>(d_sqrt((T)r2d) / (T)10, fc, dfc);
                live = fc > 0;
            }
        }
        uint64_t mj = __ballot(cand);
        const uint64_t mf = __ballot(live);
        while (mj) {
            const int l = __builtin_ctzll(mj);
            mj &= mj - 1;
            const int32_t o = ord++;
            if (o % kAtomWaves != w) continue;
            const int32_t nxj = nx + __popcll(mf & ((1ull << l) - 1));   // this j's xi table entry
            const RowEntry bj = row[q0 + l];
            const int32_t j = bj.id;
            const double djx = px - bj.x, djy = py - bj.y, djz = pz - bj.z;
            const double r2d = djx * djx + djy * djy + djz * djz;
            const V3<T> ij = mk((T)djx, (T)djy, (T)djz);
            const T rij = d_sqrt((T)r2d);
            const int kd = pair_kind(ti, bj.type);
            T va = 0, dva = 0, vr = 0, dvr = 0, vdw, dvdw, bo = 0;
            V3<T> dbo = mk<T>(0, 0, 0);
            T fcut, dfcut;
            cutoff<T>(rij / (T)10, fcut, dfcut);
            if (fcut > 0) {
                const T re = (T)p.re[kd], de = (T)p.de[kd], be = (T)p.beta[kd];
                attractive<T>(rij, de, be, re, va, dva);
                repulsive<T>(rij, de, be, re, vr, dvr);
                const V3<T> eij = ij * ((T)1 / rij);
                const T ex = d_exp(K<T>::LAMDA3 * (rij - re));
                T xi = 0, xj = 0;
                V3<T> dxi = mk<T>(0, 0, 0), dxj = mk<T>(0, 0, 0);
                for (int32_t m0 = 0; m0 < len; m0 += 64) {
                    if (m0 + lane >= len) continue;
                    const RowEntry bm = row[m0 + lane];
                    if (bm.id < 0 || bm.id == j || bm.type == kNe) continue;
                    const double ax = px - bm.x, ay = py - bm.y, az = pz - bm.z;
                    const double rim2 = ax * ax + ay * ay + az * az;
                    if (rim2 >= p.rcut2) continue;
                    const V3<T> im = mk((T)ax, (T)ay, (T)az);
                    const T rim = d_sqrt((T)rim2);
                    const V3<T> jm = mk((T)(bj.x - bm.x), (T)(bj.y - bm.y), (T)(bj.z - bm.z));
                    const T rjm = d_sqrt(dot(jm, jm));
                    T fci, dfci, fcj, dfcj;
                    cutoff<T>(rim, fci, dfci);
                    cutoff<T>(rjm, fcj, dfcj);
                    const V3<T> eim = im * ((T)1 / rim);
                    if (fci > 0) {
                        T cs = (rij * rij + rim * rim - rjm * rjm) / ((T)2 * rij * rim);
                        T g, dg;
                        ang<T>(cs, g, dg);
                        xi += fci * g * ex;
                        V3<T> dcs = (eim - eij * cs) * ((T)1 / rij) + (eij - eim * cs) * ((T)1 / rim);
                        dxi = dxi + eim * (dfci * g * ex) + dcs * (fci * dg * ex) + eij * (fci * g * K<T>::LAMDA3 * ex);
                    }
                    if (fcj > 0) {
                        T cs = (rij * rij + rjm * rjm - rim * rim) / ((T)2 * rij * rjm);
                        T g, dg;
                        ang<T>(cs, g, dg);
                        xj += fcj * g * ex;
                        V3<T> dcs = (jm * ((T)-1 / rjm) - eij * cs) * ((T)1 / rij);
                        dxj = dxj + dcs * (fcj * dg * ex) + eij * (fcj * g * K<T>::LAMDA3 * ex);
                    }
                }
                xi = wave_sum(xi);
                xj = wave_sum(xj);
                dxi = mk(wave_sum(dxi.x), wave_sum(dxi.y), wave_sum(dxi.z));
                dxj = mk(wave_sum(dxj.x), wave_sum(dxj.y), wave_sum(dxj.z));
                T qi = (T)1 / (((T)1 + xi) * ((T)1 + xi)), qj = (T)1 / (((T)1 + xj) * ((T)1 + xj));
                bo = (T)0.5 * ((T)1 / ((T)1 + xi) + (T)1 / ((T)1 + xj));
                dbo = (dxi * qi + dxj * qj) * (T)0.5;
                if (nxj < cap && lane == 0) {
                    int64_t q = (int64_t)i * cap + nxj;
                    xi_j[q] = j;
                    xi_v[2 * q] = (double)xi;
                    xi_v[2 * q + 1] = (double)xj;
                }
            }
            vanderwaals<T>(rij, (T)p.r0v[kd], vdw, dvdw);
            T fij = dvr - bo * dva + dvdw;
            T sf = fij / rij;
            if (lane == 0) {
                part[o][0] = (double)(ij.x * sf - dbo.x * va);
                part[o][1] = (double)(ij.y * sf - dbo.y * va);
                part[o][2] = (double)(ij.z * sf - dbo.z * va);
                part[o][3] = 0.5 * (double)((vr - bo * va) + vdw);
            }
        }
        nx += __popcll(mf);
    }
    __syncthreads();   // every candidate's terms are parked
    if (threadIdx.x != 0) return;
    double Fx = 0, Fy = 0, Fz = 0, e = 0;
    for (int32_t o = 0; o < ord; o++) {
        Fx += part[o][0];
        Fy += part[o][1];
        Fz += part[o][2];
        e += part[o][3];
    }
    if (nx < cap) xi_j[(int64_t)i * cap + nx] = -1;
    if (assign) {
        c.st.f[0][si] = Fx; c.st.f[1][si] = Fy; c.st.f[2][si] = Fz;
    } else {
        c.st.f[0][si] += Fx; c.st.f[1][si] += Fy; c.st.f[2][si] += Fz;
    }
    if (e_atom) e_atom[i] = e;
}

template <class T>
__global__ void __launch_bounds__(kWaveBlock)
k_bond_wave1(Ctx c, BondRefParams p, int32_t n_bond, int32_t cap, int32_t assign,
             int32_t* __restrict__ xi_j, double* __restrict__ xi_v, double* __restrict__ e_atom,
             const int32_t* __restrict__ gate) {
    __shared__ RowEntry row[kRowMax];
    __shared__ double part[kRowMax][4];
    if (gate && *gate == 0) return;   // the same for every workgroup
    const int32_t i = (int32_t)blockIdx.x;   // one workgroup per bond-order atom
    if (i >= n_bond) return;
    const int32_t len = stage_row(c, i, row, threadIdx.x & 63, (int32_t)threadIdx.x, kWaveBlock);
    __syncthreads();
    bond_wave1_block<T>(c, p, i, cap, assign, xi_j, xi_v, e_atom, row, len, part);
}

// pass 2 of atom k by its workgroup: i runs over k's row in order, candidate
// o on wavefront o % kAtomWaves, j over k's row on the lanes; the xi table of
// i (at most 64 entries) is held one entry per lane.  Each wavefront sums its
// lanes' terms, and thread 0 adds the wavefronts' sums in wavefront order.
template <class T>
__device__ inline void bond_wave2_block(const Ctx& c, const BondRefParams& p, int32_t k, int32_t cap,
                                        const int32_t* __restrict__ xi_j, const double* __restrict__ xi_v,
                                        const RowEntry* row, int32_t len, double (*part)[4]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int32_t sk = c.slot_of_id[k];
    const double px = c.st.r[0][sk], py = c.st.r[1][sk], pz = c.st.r[2][sk];
    double Fx = 0, Fy = 0, Fz = 0;
    int32_t ord = 0;
    for (int32_t q0 = 0; q0 < len; q0 += 64) {
        bool cand = false;
        if (q0 + lane < len) {
            const RowEntry a = row[q0 + lane];
            const double dx = px - a.x, dy = py - a.y, dz = pz - a.z;
            cand = a.id >= 0 && dx * dx + dy * dy + dz * dz < p.rcut2;
        }
        uint64_t mi = __ballot(cand);
        while (mi) {
            const int l = __builtin_ctzll(mi);
            mi &= mi - 1;
            if (ord++ % kAtomWaves != w) continue;
            const RowEntry bi = row[q0 + l];
            const int32_t i = bi.id;
            const double kx = px - bi.x, ky = py - bi.y, kz = pz - bi.z;
            const V3<T> ik = mk((T)kx, (T)ky, (T)kz);
            const T rik = d_sqrt((T)(kx * kx + ky * ky + kz * kz));
            // i's xi table up to its first -1 (the reference scans no further)
            int32_t tj = -1;
            double t0 = 0, t1 = 0;
            if (lane < cap) {
                const int64_t q = (int64_t)i * cap + lane;
                tj = xi_j[q];
                t0 = xi_v[2 * q];
                t1 = xi_v[2 * q + 1];
            }
            const uint64_t ends = __ballot(lane < cap && tj < 0);
            const int nt = ends ? __builtin_ctzll(ends) : cap;
            for (int32_t m0 = 0; m0 < len; m0 += 64) {
                const bool have = m0 + lane < len;
                const RowEntry bj = row[have ? m0 + lane : 0];
                const int32_t j = have ? bj.id : -1;
                T xi = 0, xj = 0;
                for (int t = 0; t < nt; t++) {
                    const int32_t jj = __builtin_amdgcn_readlane(tj, t);
                    const double v0 = lane_double(t0, t), v1 = lane_double(t1, t);
                    if (jj == j) {
                        xi = (T)v0;
                        xj = (T)v1;
                    }
                }
                if (j < 0 || j == i) continue;
                const double ux = bi.x - bj.x, uy = bi.y - bj.y, uz = bi.z - bj.z;
                const double r2d = ux * ux + uy * uy + uz * uz;
                if (r2d >= p.rcut2 || r2d > p.pair_cut2) continue;
                const T rij = d_sqrt((T)r2d);
                const int kd = pair_kind(bi.type, bj.type);
                const T re = (T)p.re[kd], de = (T)p.de[kd], be = (T)p.beta[kd];
                T va, dva;
                attractive<T>(rij, de, be, re, va, dva);
                // bondOrder2 (BondOrderFuncs.h:148-237): -db_ij/dx_k, halves included
                const V3<T> kj = mk((T)(px - bj.x), (T)(py - bj.y), (T)(pz - bj.z));
                T rjk = d_sqrt(dot(kj, kj));
                V3<T> ejk = kj * ((T)1 / rjk);
                V3<T> eik = ik * ((T)1 / rik);
                V3<T> uij = mk((T)-ux, (T)-uy, (T)-uz) * ((T)1 / rij);
                T ex = d_exp(K<T>::LAMDA3 * (rij - re));
                T fci, dfci, fcj, dfcj;
                cutoff<T>(rik, fci, dfci);
                cutoff<T>(rjk, fcj, dfcj);
                V3<T> d = mk<T>(0, 0, 0);
                if (fci > 0) {
                    T cs = (rij * rij + rik * rik - rjk * rjk) / ((T)2 * rij * rik);
                    T g, dg;
                    ang<T>(cs, g, dg);
                    V3<T> dcs = (uij - eik * cs) * ((T)1 / rik);
                    V3<T> dx_ = eik * (dfci * g * ex) + dcs * (fci * dg * ex);
                    d = d + dx_ * ((T)0.5 / (((T)1 + xi) * ((T)1 + xi)));
                }
                if (fcj > 0) {
                    T cs = (rij * rij + rjk * rjk - rik * rik) / ((T)2 * rij * rjk);
                    T g, dg;
                    ang<T>(cs, g, dg);
                    V3<T> dcs = (uij * (T)-1 - ejk * cs) * ((T)1 / rjk);
                    V3<T> dx_ = ejk * (dfcj * g * ex) + dcs * (fcj * dg * ex);
                    d = d + dx_ * ((T)0.5 / (((T)1 + xj) * ((T)1 + xj)));
                }
                Fx -= (double)((T)0.5 * d.x * va);
                Fy -= (double)((T)0.5 * d.y * va);
                Fz -= (double)((T)0.5 * d.z * va);
            }
        }
    }
    Fx = wave_sum(Fx);
    Fy = wave_sum(Fy);
    Fz = wave_sum(Fz);
    if (lane == 0) {
        part[w][0] = Fx;
        part[w][1] = Fy;
        part[w][2] = Fz;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double Gx = 0, Gy = 0, Gz = 0;
    for (int q = 0; q < kAtomWaves; q++) {
        Gx += part[q][0];
        Gy += part[q][1];
        Gz += part[q][2];
    }
    c.st.f[0][sk] += Gx;
    c.st.f[1][sk] += Gy;
    c.st.f[2][sk] += Gz;
}

template <class T>
__global__ void __launch_bounds__(kWaveBlock)
k_bond_wave2(Ctx c, BondRefParams p, int32_t n_bond, int32_t cap, const int32_t* __restrict__ xi_j,
             const double* __restrict__ xi_v, const int32_t* __restrict__ gate) {
    __shared__ RowEntry row[kRowMax];
    __shared__ double part[kAtomWaves][4];
    if (gate && *gate == 0) return;
    const int32_t k = (int32_t)blockIdx.x;   // one workgroup per bond-order atom
    if (k >= n_bond) return;
    const int32_t len = stage_row(c, k, row, threadIdx.x & 63, (int32_t)threadIdx.x, kWaveBlock);
    __syncthreads();
    bond_wave2_block<T>(c, p, k, cap, xi_j, xi_v, row, len, part);
}

// below this many bond-order atoms (with rows and xi tables that fit) the
// wavefront-per-atom kernels run
constexpr int32_t kWaveAtoms = 32768;

// ---- sweep members in one launch (blockIdx.y = member) ----
// Each member runs on the kind of kernel launch_bond_ref picks for it
// (v.bond_wave); the blocks of the other kind return at once.  gated: only
// the members whose wall-hit word is set.
__device__ inline Ctx member_ctx(const MemberView& v) { return Ctx{v.st, v.slot_of_id, v.mb_cell_of, v.mb_ids, v.mb}; }

template <class T>
__global__ void __launch_bounds__(kBlock)
k_sweep_bond1(const MemberView* __restrict__ views, BondRefParams p, int32_t n_bond, int32_t assign, int32_t gated) {
    const MemberView& v = views[blockIdx.y];
    if (v.bond_wave || (gated && *v.hit == 0)) return;
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_bond) return;
    bond_pass1_atom<T>(member_ctx(v), p, i, v.xi_cap, assign, v.xi_j, v.xi_v, nullptr);
}

template <class T>
__global__ void __launch_bounds__(kBlock)
k_sweep_bond2(const MemberView* __restrict__ views, BondRefParams p, int32_t n_bond, int32_t gated) {
    const MemberView& v = views[blockIdx.y];
    if (v.bond_wave || (gated && *v.hit == 0)) return;
    int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_bond) return;
    bond_pass2_atom<T>(member_ctx(v), p, k, v.xi_cap, v.xi_j, v.xi_v);
}

template <class T>
__global__ void __launch_bounds__(kWaveBlock)
k_sweep_bond_wave1(const MemberView* __restrict__ views, BondRefParams p, int32_t n_bond, int32_t assign,
                   int32_t gated) {
    __shared__ RowEntry row[kRowMax];
    __shared__ double part[kRowMax][4];
    const MemberView& v = views[blockIdx.y];
    if (!v.bond_wave || (gated && *v.hit == 0)) return;   // the same for the whole workgroup
    const int32_t i = (int32_t)blockIdx.x;   // one workgroup per bond-order atom
    if (i >= n_bond) return;
    const Ctx c = member_ctx(v);
    const int32_t len = stage_row(c, i, row, threadIdx.x & 63, (int32_t)threadIdx.x, kWaveBlock);
    __syncthreads();
    bond_wave1_block<T>(c, p, i, v.xi_cap, assign, v.xi_j, v.xi_v, nullptr, row, len, part);
}

template <class T>
__global__ void __launch_bounds__(kWaveBlock)
k_sweep_bond_wave2(const MemberView* __restrict__ views, BondRefParams p, int32_t n_bond, int32_t gated) {
    __shared__ RowEntry row[kRowMax];
    __shared__ double part[kAtomWaves][4];
    const MemberView& v = views[blockIdx.y];
    if (!v.bond_wave || (gated && *v.hit == 0)) return;
    const int32_t k = (int32_t)blockIdx.x;   // one workgroup per bond-order atom
    if (k >= n_bond) return;
    const Ctx c = member_ctx(v);
    const int32_t len = stage_row(c, k, row, threadIdx.x & 63, (int32_t)threadIdx.x, kWaveBlock);
    __syncthreads();
    bond_wave2_block<T>(c, p, k, v.xi_cap, v.xi_j, v.xi_v, row, len, part);
}

// common/DebugCalculations.cpp:129-270: per-atom share of the brute-force
// potential.  Ordered pairs with a Ne atom add 2 eps [(s/r)^12 - (s/r)^6]
// (no cutoff); other pairs with r <= CUTOFF_RIJ add (V_R - b V_A + vdW)/2
// with xi over every non-Ne atom.  out[s] = LJ part, out[n + s] = bond part.
__global__ void __launch_bounds__(kBlock)
k_allpairs(DevState st, int32_t n, LjTable lj, BondRefParams p, double* __restrict__ out) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t ti = st.type[i];
    double elj = 0, emb = 0;
    for (int32_t j = 0; j < n; j++) {
        if (j == i) continue;
        const int32_t tj = st.type[j];
        double dx = st.r[0][i] - st.r[0][j], dy = st.r[1][i] - st.r[1][j], dz = st.r[2][i] - st.r[2][j];
        double r2 = dx * dx + dy * dy + dz * dz;
        if (ti == kNe || tj == kNe) {
            int k = (ti & (kTypes - 1)) * kTypes + (tj & (kTypes - 1));
            double s2 = lj.sig2[k] / r2, s6 = s2 * s2 * s2;
            elj += 0.5 * lj.e4[k] * (s6 * s6 - s6);
            continue;
        }
        double rij = sqrt(r2);
        if (rij > p.pair_cut) continue;
        const int kd = pair_kind(ti, tj);
        double va = 0, dva, vr = 0, dvr, vdw, dvdw, bo = 0, fcut, dfcut;
        cutoff<double>(rij / 10, fcut, dfcut);
        if (fcut > 0) {
            const double re = p.re[kd], de = p.de[kd], be = p.beta[kd];
            attractive<double>(rij, de, be, re, va, dva);
            repulsive<double>(rij, de, be, re, vr, dvr);
            double ex = exp(K<double>::LAMDA3 * (rij - re)), xi = 0, xj = 0;
            for (int32_t m = 0; m < n; m++) {
                if (st.type[m] == kNe || m == i || m == j) continue;
                double ax = st.r[0][i] - st.r[0][m], ay = st.r[1][i] - st.r[1][m], az = st.r[2][i] - st.r[2][m];
                double bx = st.r[0][j] - st.r[0][m], by = st.r[1][j] - st.r[1][m], bz = st.r[2][j] - st.r[2][m];
                double rim = sqrt(ax * ax + ay * ay + az * az), rjm = sqrt(bx * bx + by * by + bz * bz);
                double fci, dfci, fcj, dfcj, g, dg;
                cutoff<double>(rim, fci, dfci);
                cutoff<double>(rjm, fcj, dfcj);
                if (fci > 0) {
                    ang<double>((rij * rij + rim * rim - rjm * rjm) / (2.0 * rij * rim), g, dg);
                    xi += fci * g * ex;
                }
                if (fcj > 0) {
                    ang<double>((rij * rij + rjm * rjm - rim * rim) / (2.0 * rij * rjm), g, dg);
                    xj += fcj * g * ex;
                }
            }
            bo = 0.5 * (1.0 / (1.0 + xi) + 1.0 / (1.0 + xj));
        }
        vanderwaals<double>(rij, p.r0v[kd], vdw, dvdw);
        emb += 0.5 * ((vr - bo * va) + vdw);
    }
    out[i] = elj;
    out[n + i] = emb;
}

}  // namespace

hipError_t launch_allpairs(hipStream_t s, const DevState& st, int32_t n, int32_t n_bond,
                           const LjTable& lj, const BondRefParams& p, double* per_atom2) {
    (void)n_bond;
    if (n <= 0) return hipSuccess;
    k_allpairs<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(st, n, lj, p, per_atom2);
    return hipGetLastError();
}

bool bond_ref_wave(int32_t n_bond, int64_t row_bound, int32_t cap) {
    return n_bond <= kWaveAtoms && row_bound <= kRowMax && cap <= 64;
}

hipError_t launch_sweep_bond(hipStream_t s, const MemberView* views, int32_t members, const BondRefParams& p,
                             int32_t n_bond, int32_t precision, int32_t assign, bool gated, bool any_wave,
                             bool any_thread) {
    if (n_bond <= 0 || members <= 0) return hipSuccess;
    const int32_t gt = gated ? 1 : 0;
    const bool fp64 = precision == MDX_PREC_FP64;
    if (any_wave) {
        dim3 wgrid(n_bond, members);   // one workgroup per bond-order atom
        if (fp64) {
            k_sweep_bond_wave1<double><<<wgrid, kWaveBlock, 0, s>>>(views, p, n_bond, assign, gt);
            k_sweep_bond_wave2<double><<<wgrid, kWaveBlock, 0, s>>>(views, p, n_bond, gt);
        } else {
            k_sweep_bond_wave1<float><<<wgrid, kWaveBlock, 0, s>>>(views, p, n_bond, assign, gt);
            k_sweep_bond_wave2<float><<<wgrid, kWaveBlock, 0, s>>>(views, p, n_bond, gt);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (any_thread) {
        dim3 grid((n_bond + kBlock - 1) / kBlock, members);
        if (fp64) {
            k_sweep_bond1<double><<<grid, kBlock, 0, s>>>(views, p, n_bond, assign, gt);
            k_sweep_bond2<double><<<grid, kBlock, 0, s>>>(views, p, n_bond, gt);
        } else {
            k_sweep_bond1<float><<<grid, kBlock, 0, s>>>(views, p, n_bond, assign, gt);
            k_sweep_bond2<float><<<grid, kBlock, 0, s>>>(views, p, n_bond, gt);
        }
    }
    return hipGetLastError();
}

hipError_t launch_bond_ref(hipStream_t s, const DevState& st, const int32_t* slot_of_id,
                           const int32_t* mb_cell_of, const int32_t* mb_ids, const GridView& g,
                           const BondRefParams& p, int32_t n_bond, int32_t cap,
                           int32_t precision, int32_t assign, int32_t* xi_j, double* xi_v,
                           double* energy_per_atom, const int32_t* gate, int64_t row_bound) {
    if (n_bond <= 0) return hipSuccess;
    Ctx c{st, slot_of_id, mb_cell_of, mb_ids, g};
    if (bond_ref_wave(n_bond, row_bound, cap)) {
        dim3 wgrid(n_bond);   // one workgroup per bond-order atom
        if (precision == MDX_PREC_FP64) {
            k_bond_wave1<double><<<wgrid, kWaveBlock, 0, s>>>(c, p, n_bond, cap, assign, xi_j, xi_v,
                                                             energy_per_atom, gate);
            k_bond_wave2<double><<<wgrid, kWaveBlock, 0, s>>>(c, p, n_bond, cap, xi_j, xi_v, gate);
        } else {
            k_bond_wave1<float><<<wgrid, kWaveBlock, 0, s>>>(c, p, n_bond, cap, assign, xi_j, xi_v,
                                                            energy_per_atom, gate);
            k_bond_wave2<float><<<wgrid, kWaveBlock, 0, s>>>(c, p, n_bond, cap, xi_j, xi_v, gate);
        }
        return hipGetLastError();
    }
    dim3 grid((n_bond + kBlock - 1) / kBlock);
    if (precision == MDX_PREC_FP64) {
        k_bond_pass1<double><<<grid, kBlock, 0, s>>>(c, p, n_bond, cap, assign, xi_j, xi_v,
                                                    energy_per_atom, gate);
        k_bond_pass2<double><<<grid, kBlock, 0, s>>>(c, p, n_bond, cap, xi_j, xi_v, gate);
    } else {
        k_bond_pass1<float><<<grid, kBlock, 0, s>>>(c, p, n_bond, cap, assign, xi_j, xi_v,
                                                   energy_per_atom, gate);
        k_bond_pass2<float><<<grid, kBlock, 0, s>>>(c, p, n_bond, cap, xi_j, xi_v, gate);
    }
    return hipGetLastError();
}

}  // namespace mdx
