// ref_adapter_harness.cpp -- test driver: the cpumd step loop for one vxcm
// value (cpu/cpumd/Host.cpp:78-198) written against the reference's function
// signatures, linked with libmdx_ref.so (the engine behind those signatures).
// The parts of hostMain that stay in the caller (readyList and the rebuild
// schedule, Host.cpp:245-291) come from the oracle.  Nothing but the
// reference's own calls reaches the library.  The GPU tests run it and
// compare its final state with the oracle's own loop.
//
// usage: ref_adapter_harness IN OUT
//   IN:  int32 n_lj, n_mb, timesteps, use_lj, use_mb; float dt, lj_rs, lj_rcut,
//        mb_rs, mb_rcut, energy_loss (before the square root); then n x 4
//        doubles of positions (species in w) and n x 3 doubles of velocities
//   OUT: n x 4 doubles of positions, n x 3 of velocities, int32 list builds
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../include/mdx_reference.h"
#include "../oracle/oracle.h"

namespace {

void settings(listSettings& s, float rs, float rc) {   // cpu/cpumd/Host.cpp:50-61
    s.maxnlmove = (rs - rc) * 0.5f;
    s.maxnlmovesq = (float)std::pow(s.maxnlmove, 2);
    s.rcutsq = (float)std::pow(rc, 2);
    s.rcut = rc;
    s.rs = rs;
    s.nlistsize = s.nlargestsize = 0;
}

template <class T>
bool get(FILE* f, T* p, size_t n) { return std::fread(p, sizeof(T), n, f) == n; }

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
        return 2;
    }
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) return 2;
    int32_t hdr[5];
    float fl[6];
    if (!get(in, hdr, 5) || !get(in, fl, 6)) return 2;
    const int n_lj = hdr[0], n_mb = hdr[1], steps = hdr[2];
    const bool use_lj = hdr[3] != 0, use_mb = hdr[4] != 0;
    const double dt = (double)fl[0];
    const int n = n_lj + n_mb;
    std::vector<real4> r(n);
    std::vector<real3> v(n), f(n), a(n), b(n), c(n);
    if (!get(in, &r[0].x, 4 * (size_t)n) || !get(in, &v[0].x, 3 * (size_t)n)) return 2;
    std::fclose(in);

    const double box = std::pow((double)n, 1.0 / 3.0) * 52.8;   // Host.cpp:31
    const float eloss = std::sqrt(fl[5]);                         // Host.cpp:73
    listSettings lj{}, mb{};
    settings(lj, fl[1], fl[2]);
    settings(mb, fl[3], fl[4]);

    const double half = box / 2;                                  // Host.cpp:94-106 (vxcm = 0)
    for (int i = 0; i < n; i++) {
        double* R = &r[i].x;
        for (int d = 0; d < 3; d++)
            if (std::fabs(R[d]) > half) R[d] = half * (R[d] / std::fabs(R[d]));
    }
    const int buckets = 3 * n;
    int* ljList = nullptr;
    int* mbList = nullptr;
    int ljNext = 0, mbNext = 0, builds = 0;
    auto ready = [&](double& vmax, double& bsl) {
        orc_ref_ready_list(reinterpret_cast<const orc_real4*>(r.data()), reinterpret_cast<const orc_real3*>(v.data()),
                           n, &vmax, &bsl);
    };
    auto build = [&](listSettings& s, int** list, int count, int& next, double bsl, double vmax, int t) {
        delete[] *list;
        *list = buildNeighborList(r.data(), &s, bsl, buckets, count);
        next = orc_ref_next_build(s.maxnlmove, vmax, dt, t);
        builds++;
    };
    auto forces = [&]() {
        if (use_lj) lennardJonesForces(r.data(), f.data(), n, lj.nlargestsize, lj.nlistsize, ljList, lj.rcutsq);
        if (use_mb) manyBodyForces(r.data(), f.data(), n_mb, mb.nlargestsize, mb.nlistsize, mbList, mb.rcutsq, use_lj);
    };

    double vmax = 0, bsl = 0;
    ready(vmax, bsl);
    if (use_lj) build(lj, &ljList, n, ljNext, bsl, vmax, 0);
    if (use_mb) build(mb, &mbList, n_mb, mbNext, bsl, vmax, 0);
    forces();
    calculateAccelerations(r.data(), a.data(), f.data(), n);
    for (int t = 0; t < steps; t++) {
        if (ljNext == t || mbNext == t) {
            ready(vmax, bsl);
            if (use_lj && ljNext == t) build(lj, &ljList, n, ljNext, bsl, vmax, t);
            if (use_mb && mbNext == t) build(mb, &mbList, n_mb, mbNext, bsl, vmax, t);
        }
        predict(r.data(), v.data(), a.data(), b.data(), c.data(), dt, n);
        forces();
        bool hit = false;
        correct(r.data(), v.data(), f.data(), a.data(), b.data(), c.data(), dt, n, box, hit, eloss);
        if (hit) {
            forces();
            calculateAccelerations(r.data(), a.data(), f.data(), n);
        }
    }
    delete[] ljList;
    delete[] mbList;
    mdxRefRelease();

    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    std::fwrite(&r[0].x, sizeof(double), 4 * (size_t)n, out);
    std::fwrite(&v[0].x, sizeof(double), 3 * (size_t)n, out);
    int32_t nb = builds;
    std::fwrite(&nb, sizeof nb, 1, out);
    std::fclose(out);
    return 0;
}
