// ref_adapter.cpp -- the reference's free functions (include/mdx_reference.h)
// on top of the C-ABI (include/mdx.h).
//
// One device system serves all calls.  Its configuration is assembled from
// what the calls pass (particle counts and list settings from mdxRefSetup and
// buildNeighborList, dt from predict/correct, the wall position and
// restitution from correct); when a value changes the system is created again
// and the list builds done so far are replayed from their recorded positions,
// so the device lists always equal the ones the caller holds.  Every call
// uploads the caller's arrays, runs one entry point and downloads what that
// entry point changes.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "mdx.h"
#include "mdx_reference.h"

namespace {

[[noreturn]] void die(const char* what, const mdx_system* s) {
    std::fprintf(stderr, "mdx reference adapter: %s: %s\n", what, mdx_last_error(s));
    std::exit(EXIT_FAILURE);
}

struct BuildRecord {
    bool valid = false;
    std::vector<double> pos4;   // all atoms, at the time of the build
    double box = 0;
    const int* handed_out = nullptr;  // the list returned to the caller
};

struct Adapter {
    mdx_system* sys = nullptr;
    mdx_config have{};   // configuration of sys
    mdx_config want{};   // configuration the calls so far ask for
    bool setup = false;
    int n_all = 0, n_mb = 0;
    const listSettings* ls_lj = nullptr;
    const listSettings* ls_mb = nullptr;
    BuildRecord lj, mb;
    std::vector<double> zeros;

    Adapter() {
        std::memset(&have, 0, sizeof have);
        std::memset(&want, 0, sizeof want);
        want.box = MDX_BOX_WALLED;
        const char* p = std::getenv("MDX_REF_PRECISION");
        want.precision = (p && std::strcmp(p, "mixed") == 0) ? MDX_PREC_MIXED : MDX_PREC_FP64;
        want.lj_table = MDX_LJ_REFERENCE;
        want.bond_form = MDX_BOND_REFERENCE;
        want.accel_scale = 1.0;
    }
    ~Adapter() {
        if (sys) mdx_destroy(sys);
    }

    void require_setup() const {
        if (!setup) {
            std::fprintf(stderr, "mdx reference adapter: call mdxRefSetup before the first reference call\n");
            std::exit(EXIT_FAILURE);
        }
    }

    void check_n(int n, int expect, const char* fn) const {
        if (n != expect) {
            std::fprintf(stderr, "mdx reference adapter: %s called for %d atoms, set up for %d\n", fn, n, expect);
            std::exit(EXIT_FAILURE);
        }
    }

    // (re)create the device system when the wanted configuration changed,
    // then replay the list builds the caller still holds
    void ensure() {
        require_setup();
        // the caller's list settings objects hold the current radii
        if (ls_lj) { want.lj_rs = ls_lj->rs; want.lj_rcut = ls_lj->rcut; }
        if (ls_mb) { want.mb_rs = ls_mb->rs; want.mb_rcut = ls_mb->rcut; }
        if (sys && std::memcmp(&have, &want, sizeof want) == 0) return;
        if (sys) mdx_destroy(sys);
        sys = nullptr;
        if (mdx_create(&want, &sys) != MDX_OK) die("mdx_create", nullptr);
        have = want;
        for (int which : {MDX_LIST_LJ, MDX_LIST_BOND}) {
            BuildRecord& r = which == MDX_LIST_LJ ? lj : mb;
            if (!r.valid) continue;
            upload_pos(r.pos4.data());
            int32_t nls = 0, nlg = 0;
            if (mdx_build_list(sys, which, r.box, &nls, &nlg) != MDX_OK) die("list rebuild", sys);
        }
    }

    void upload(const real4* pos, const real3* vel, const real3* a, const real3* b, const real3* c) {
        const double* z = zeros.data();
        if (mdx_set_state(sys, &pos->x, vel ? &vel->x : z, a ? &a->x : z, b ? &b->x : z, c ? &c->x : z) != MDX_OK)
            die("mdx_set_state", sys);
    }
    void upload_pos(const double* pos4) {
        if (mdx_set_state(sys, pos4, zeros.data(), nullptr, nullptr, nullptr) != MDX_OK) die("mdx_set_state", sys);
    }
    void upload_forces(const real3* f) {
        if (mdx_set_forces(sys, &f->x) != MDX_OK) die("mdx_set_forces", sys);
    }
    void download(real4* pos, real3* vel, real3* a, real3* b, real3* c, real3* f) {
        if (mdx_get_state(sys, pos ? &pos->x : nullptr, vel ? &vel->x : nullptr, a ? &a->x : nullptr,
                          b ? &b->x : nullptr, c ? &c->x : nullptr, f ? &f->x : nullptr) != MDX_OK)
            die("mdx_get_state", sys);
    }
};

Adapter& ad() {
    static Adapter a;
    return a;
}

}  // namespace

void mdxRefSetup(int lennardJonesParticles, int manyBodyParticles, bool useLennardJones, bool useManyBody,
                 const listSettings* lennardJonesList, const listSettings* manyBodyList) {
    Adapter& A = ad();
    A.n_all = lennardJonesParticles + manyBodyParticles;
    A.n_mb = manyBodyParticles;
    A.ls_lj = lennardJonesList;
    A.ls_mb = manyBodyList;
    A.want.n_atoms = A.n_all;
    A.want.n_bond = useManyBody ? manyBodyParticles : 0;
    A.want.use_lj = useLennardJones ? 1 : 0;
    A.want.use_bond = useManyBody ? 1 : 0;
    A.zeros.assign(3 * (size_t)(A.n_all > 0 ? A.n_all : 1), 0.0);
    A.lj = BuildRecord();
    A.mb = BuildRecord();
    A.setup = true;
}

void mdxRefRelease() {
    Adapter& A = ad();
    if (A.sys) mdx_destroy(A.sys);
    A.sys = nullptr;
    std::memset(&A.have, 0, sizeof A.have);
}

int* buildNeighborList(real4* pos, listSettings* settings, real boxSize, int /*buckets*/, int n) {
    Adapter& A = ad();
    A.require_setup();
    const bool bond = settings == A.ls_mb && settings != A.ls_lj;
    if (!bond && settings != A.ls_lj) {
        std::fprintf(stderr, "mdx reference adapter: buildNeighborList with settings not passed to mdxRefSetup\n");
        std::exit(EXIT_FAILURE);
    }
    A.check_n(n, bond ? A.n_mb : A.n_all, "buildNeighborList");
    A.ensure();
    BuildRecord& r = bond ? A.mb : A.lj;
    r.pos4.assign(&pos->x, &pos->x + 4 * (size_t)A.n_all);
    r.box = boxSize;
    r.valid = true;
    A.upload_pos(r.pos4.data());
    const int which = bond ? MDX_LIST_BOND : MDX_LIST_LJ;
    int32_t nls = 0, nlg = 0;
    if (mdx_build_list(A.sys, which, boxSize, &nls, &nlg) != MDX_OK) die("mdx_build_list", A.sys);
    settings->nlistsize = nls;
    settings->nlargestsize = nlg;
    const size_t count = (size_t)n * (size_t)nls;
    int* out = new int[count > 0 ? count : 1];
    if (mdx_copy_neighbor_list(A.sys, which, out, (int64_t)count) != MDX_OK) die("mdx_copy_neighbor_list", A.sys);
    r.handed_out = out;
    return out;
}

void predict(real4* pos, real3* vel, real3* acc, real3* acc1, real3* acc2, real dt, int n) {
    Adapter& A = ad();
    A.require_setup();
    A.check_n(n, A.n_all, "predict");
    A.want.dt = (float)dt;
    A.ensure();
    A.upload(pos, vel, acc, acc1, acc2);
    if (mdx_predict(A.sys) != MDX_OK) die("mdx_predict", A.sys);
    A.download(pos, vel, acc, acc1, acc2, nullptr);
}

void correct(real4* pos, real3* vel, real3* force, real3* acc, real3* acc1, real3* acc2, real dt, int n, real wall,
             bool& hit, float energyLoss) {
    Adapter& A = ad();
    A.require_setup();
    A.check_n(n, A.n_all, "correct");
    A.want.dt = (float)dt;
    A.want.wall = wall;
    A.want.energy_loss = energyLoss;
    A.ensure();
    A.upload(pos, vel, acc, acc1, acc2);
    A.upload_forces(force);
    int32_t h = 0;
    if (mdx_correct(A.sys, &h) != MDX_OK) die("mdx_correct", A.sys);
    A.download(pos, vel, acc, acc1, acc2, nullptr);
    if (h) hit = true;   // the reference only ever sets the flag
}

void lennardJonesForces(real4* pos, real3* force, int n, int /*nlargestsize*/, int /*nlistsize*/, int* list,
                        real /*rcutsq*/) {
    Adapter& A = ad();
    A.require_setup();
    A.check_n(n, A.n_all, "lennardJonesForces");
    A.ensure();
    if (!A.lj.valid || list != A.lj.handed_out) {
        std::fprintf(stderr, "mdx reference adapter: lennardJonesForces needs the list of the last "
                             "buildNeighborList call\n");
        std::exit(EXIT_FAILURE);
    }
    A.upload(pos, nullptr, nullptr, nullptr, nullptr);
    if (mdx_lj_forces(A.sys) != MDX_OK) die("mdx_lj_forces", A.sys);
    A.download(nullptr, nullptr, nullptr, nullptr, nullptr, force);
}

void manyBodyForces(real4* pos, real3* force, int n, int /*nlargestsize*/, int /*nlistsize*/, int* list,
                    real /*rcutsq*/, bool addToLennardJones) {
    Adapter& A = ad();
    A.require_setup();
    A.check_n(n, A.n_mb, "manyBodyForces");
    if ((A.want.use_lj != 0) != addToLennardJones) {
        std::fprintf(stderr, "mdx reference adapter: manyBodyForces ljexist differs from mdxRefSetup\n");
        std::exit(EXIT_FAILURE);
    }
    A.ensure();
    if (!A.mb.valid || list != A.mb.handed_out) {
        std::fprintf(stderr, "mdx reference adapter: manyBodyForces needs the list of the last "
                             "buildNeighborList call\n");
        std::exit(EXIT_FAILURE);
    }
    A.upload(pos, nullptr, nullptr, nullptr, nullptr);
    // with Lennard-Jones on, the bond-order forces add to the caller's array;
    // otherwise they are assigned (atoms past the first n keep what they hold)
    A.upload_forces(force);
    if (mdx_bond_forces(A.sys) != MDX_OK) die("mdx_bond_forces", A.sys);
    A.download(nullptr, nullptr, nullptr, nullptr, nullptr, force);
}

void calculateAccelerations(real4* pos, real3* acc, real3* force, int n) {
    Adapter& A = ad();
    A.require_setup();
    A.check_n(n, A.n_all, "calculateAccelerations");
    A.ensure();
    A.upload(pos, nullptr, nullptr, nullptr, nullptr);
    A.upload_forces(force);
    if (mdx_accelerations(A.sys) != MDX_OK) die("mdx_accelerations", A.sys);
    A.download(nullptr, nullptr, acc, nullptr, nullptr, nullptr);
}
