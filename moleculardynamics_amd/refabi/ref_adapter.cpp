// ref_adapter.cpp -- the reference's free functions (include/mdx_reference.h)
// on top of the C-ABI (include/mdx.h), for the reference's hostMain unchanged.
//
// The reference's calls carry everything the device needs, spread over the
// first step of hostMain (cpu/cpumd/Host.cpp:117-138):
//  * buildNeighborList(pos, settings, box, buckets, n) builds a list of the
//    first n atoms with cells of edge settings->rs.  Its result depends only on
//    those arguments, not on which potential will use it, so it is served at
//    once by a list-building system of n atoms and that radius;
//  * lennardJonesForces(..., n, ..., list, ...) names the list of the
//    Lennard-Jones settings (the settings object whose last build returned
//    `list`) and the atom count; manyBodyForces(..., n, ..., list, ..., ljexist)
//    names the bond-order settings, ManyBodyParticles and whether Lennard-Jones
//    is on; predict, correct and calculateAccelerations give the atom count,
//    dt, the wall and the restitution.
// The force system is created from what is known when a call needs it, and
// created again (replaying the list builds the caller still holds) when a
// later call adds to it; that happens within the first step only.  Every
// call uploads the caller's arrays, runs one entry point and downloads what
// that entry point changes.
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

[[noreturn]] void die_msg(const char* msg) {
    std::fprintf(stderr, "mdx reference adapter: %s\n", msg);
    std::exit(EXIT_FAILURE);
}

// the last build of one listSettings object
struct ListRec {
    const listSettings* ls = nullptr;
    int n = 0;                       // atoms of the list (the first n of the arrays)
    std::vector<double> pos4;        // their positions at the build
    double box = 0;
    const int* handed_out = nullptr; // the list returned to the caller
    long version = 0;                // builds of this object so far
    int32_t nls = 0, nlg = 0;        // nlistsize, nlargestsize of the last build
};

// a walled system of n atoms that only builds candidate lists of radius rs
struct ListSys {
    mdx_config cfg{};
    mdx_system* sys = nullptr;
};

bool same_config(const mdx_config& a, const mdx_config& b) { return std::memcmp(&a, &b, sizeof a) == 0; }

struct Adapter {
    std::vector<ListRec> recs;
    std::vector<ListSys> listers;
    mdx_system* sys = nullptr;   // the force and integration system
    mdx_config have{};           // its configuration
    mdx_config want{};           // the configuration the calls so far ask for
    int n_all = 0;               // the largest atom count seen
    int lj_rec = -1, mb_rec = -1;  // records of the Lennard-Jones and bond-order settings, once known
    long lj_built = -1, mb_built = -1;   // the record versions sys holds lists of
    int precision = MDX_PREC_FP64;
    std::vector<double> zeros, pad;

    Adapter() {
        std::memset(&have, 0, sizeof have);
        std::memset(&want, 0, sizeof want);
        const char* p = std::getenv("MDX_REF_PRECISION");
        precision = (p && std::strcmp(p, "mixed") == 0) ? MDX_PREC_MIXED : MDX_PREC_FP64;
        want.box = MDX_BOX_WALLED;
        want.precision = precision;
        want.lj_table = MDX_LJ_REFERENCE;
        want.bond_form = MDX_BOND_REFERENCE;
        want.accel_scale = 1.0;
        want.energy_loss = 1.0f;
    }
    ~Adapter() { release(); }

    void release() {
        if (sys) mdx_destroy(sys);
        sys = nullptr;
        std::memset(&have, 0, sizeof have);
        lj_built = mb_built = -1;
        for (ListSys& l : listers)
            if (l.sys) mdx_destroy(l.sys);
        listers.clear();
    }

    void see_atoms(int n) {
        if (n > n_all) n_all = n;
        if (zeros.size() < 3 * (size_t)n_all) zeros.assign(3 * (size_t)n_all, 0.0);
    }

    int rec_of_settings(const listSettings* ls) {
        for (size_t k = 0; k < recs.size(); k++)
            if (recs[k].ls == ls) return (int)k;
        recs.push_back(ListRec());
        recs.back().ls = ls;
        return (int)recs.size() - 1;
    }

    int rec_of_list(const int* list, const char* fn) const {
        for (size_t k = 0; k < recs.size(); k++)
            if (recs[k].handed_out == list && list) return (int)k;
        std::fprintf(stderr, "mdx reference adapter: %s needs a list returned by the last buildNeighborList call "
                             "of its settings\n", fn);
        std::exit(EXIT_FAILURE);
    }

    mdx_system* lister(int n, float rs, float rc) {
        mdx_config c{};
        c.n_atoms = n;
        c.box = MDX_BOX_WALLED;
        c.precision = MDX_PREC_FP64;
        c.use_lj = 1;
        c.lj_table = MDX_LJ_REFERENCE;
        c.bond_form = MDX_BOND_REFERENCE;
        c.dt = 1.0f;
        c.lj_rs = rs;
        c.lj_rcut = rc;
        c.energy_loss = 1.0f;
        c.wall = 1.0;
        c.accel_scale = 1.0;
        for (ListSys& l : listers)
            if (same_config(l.cfg, c)) return l.sys;
        ListSys l;
        l.cfg = c;
        if (mdx_create(&c, &l.sys) != MDX_OK) die("mdx_create (list)", nullptr);
        listers.push_back(l);
        return l.sys;
    }

    // positions of a record for the whole system (atoms past its n at the origin)
    const double* padded(const ListRec& r) {
        pad.assign(4 * (size_t)n_all, 0.0);
        std::memcpy(pad.data(), r.pos4.data(), sizeof(double) * 4 * (size_t)r.n);
        return pad.data();
    }

    // the force system for what the calls have told so far, with the lists of
    // the records it uses as the caller holds them
    void ensure() {
        want.n_atoms = n_all;
        want.use_lj = lj_rec >= 0 ? 1 : 0;
        want.use_bond = mb_rec >= 0 ? 1 : 0;
        if (lj_rec >= 0) { want.lj_rs = recs[lj_rec].ls->rs; want.lj_rcut = recs[lj_rec].ls->rcut; }
        if (mb_rec >= 0) { want.mb_rs = recs[mb_rec].ls->rs; want.mb_rcut = recs[mb_rec].ls->rcut; }
        if (!sys || !same_config(have, want)) {
            if (sys) mdx_destroy(sys);
            sys = nullptr;
            if (mdx_create(&want, &sys) != MDX_OK) die("mdx_create", nullptr);
            have = want;
            lj_built = mb_built = -1;
        }
        for (int which : {MDX_LIST_LJ, MDX_LIST_BOND}) {
            const int k = which == MDX_LIST_LJ ? lj_rec : mb_rec;
            long& built = which == MDX_LIST_LJ ? lj_built : mb_built;
            if (k < 0 || built == recs[k].version) continue;
            if (mdx_set_state(sys, padded(recs[k]), zeros.data(), nullptr, nullptr, nullptr) != MDX_OK)
                die("mdx_set_state", sys);
            if (mdx_build_list(sys, which, recs[k].box, &recs[k].nls, &recs[k].nlg) != MDX_OK) die("list build", sys);
            built = recs[k].version;
        }
    }

    void upload(const real4* pos, const real3* vel, const real3* a, const real3* b, const real3* c) {
        const double* z = zeros.data();
        if (mdx_set_state(sys, &pos->x, vel ? &vel->x : z, a ? &a->x : z, b ? &b->x : z, c ? &c->x : z) != MDX_OK)
            die("mdx_set_state", sys);
    }
    void upload_forces(const real3* f) {
        if (mdx_set_forces(sys, &f->x) != MDX_OK) die("mdx_set_forces", sys);
    }
    void download(real4* pos, real3* vel, real3* a, real3* b, real3* c, real3* f) {
        if (mdx_get_state(sys, pos ? &pos->x : nullptr, vel ? &vel->x : nullptr, a ? &a->x : nullptr,
                          b ? &b->x : nullptr, c ? &c->x : nullptr, f ? &f->x : nullptr) != MDX_OK)
            die("mdx_get_state", sys);
    }
    void check_all(int n, const char* fn) {
        see_atoms(n);
        if (n != n_all) {
            std::fprintf(stderr, "mdx reference adapter: %s called for %d atoms after calls for %d\n", fn, n, n_all);
            std::exit(EXIT_FAILURE);
        }
    }
};

Adapter& ad() {
    static Adapter a;
    return a;
}

}  // namespace

void mdxRefRelease() { ad().release(); }

int* buildNeighborList(real4* pos, listSettings* settings, real boxSize, int /*buckets*/, int n) {
    Adapter& A = ad();
    if (!settings || n < 1) die_msg("buildNeighborList needs settings and at least one atom");
    A.see_atoms(n);
    const int k = A.rec_of_settings(settings);
    ListRec& r = A.recs[k];
    if (k == A.lj_rec && n != A.n_all) die_msg("buildNeighborList of the Lennard-Jones settings for a different atom count");
    r.n = n;
    r.pos4.assign(&pos->x, &pos->x + 4 * (size_t)n);
    r.box = boxSize;
    r.version++;
    mdx_system* from;
    int which;
    if (k == A.lj_rec || k == A.mb_rec) {
        // a list of a known role: built by the force system itself
        if (k == A.mb_rec) A.want.n_bond = n;
        A.ensure();
        from = A.sys;
        which = k == A.lj_rec ? MDX_LIST_LJ : MDX_LIST_BOND;
    } else {
        // before its first force call: built by a list system of these n atoms
        from = A.lister(n, settings->rs, settings->rcut);
        which = MDX_LIST_LJ;
        if (mdx_set_state(from, r.pos4.data(), A.zeros.data(), nullptr, nullptr, nullptr) != MDX_OK)
            die("mdx_set_state (list)", from);
        if (mdx_build_list(from, which, boxSize, &r.nls, &r.nlg) != MDX_OK) die("mdx_build_list", from);
    }
    settings->nlistsize = r.nls;
    settings->nlargestsize = r.nlg;
    const size_t count = (size_t)n * (size_t)r.nls;
    int* out = new int[count > 0 ? count : 1];
    if (mdx_copy_neighbor_list(from, which, out, (int64_t)count) != MDX_OK) die("mdx_copy_neighbor_list", from);
    r.handed_out = out;
    return out;
}

void predict(real4* pos, real3* vel, real3* acc, real3* acc1, real3* acc2, real dt, int n) {
    Adapter& A = ad();
    A.check_all(n, "predict");
    A.want.dt = (float)dt;
    A.ensure();
    A.upload(pos, vel, acc, acc1, acc2);
    if (mdx_predict(A.sys) != MDX_OK) die("mdx_predict", A.sys);
    A.download(pos, vel, acc, acc1, acc2, nullptr);
}

void correct(real4* pos, real3* vel, real3* force, real3* acc, real3* acc1, real3* acc2, real dt, int n, real wall,
             bool& hit, float energyLoss) {
    Adapter& A = ad();
    A.check_all(n, "correct");
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
    const int k = A.rec_of_list(list, "lennardJonesForces");
    if (A.lj_rec >= 0 && A.lj_rec != k) die_msg("lennardJonesForces with the list of other settings");
    if (A.recs[k].n != n) die_msg("lennardJonesForces for a different atom count than its list");
    A.check_all(n, "lennardJonesForces");
    A.lj_rec = k;
    A.ensure();
    A.upload(pos, nullptr, nullptr, nullptr, nullptr);
    if (mdx_lj_forces(A.sys) != MDX_OK) die("mdx_lj_forces", A.sys);
    A.download(nullptr, nullptr, nullptr, nullptr, nullptr, force);
}

void manyBodyForces(real4* pos, real3* force, int n, int /*nlargestsize*/, int /*nlistsize*/, int* list,
                    real /*rcutsq*/, bool addToLennardJones) {
    Adapter& A = ad();
    const int k = A.rec_of_list(list, "manyBodyForces");
    if (A.mb_rec >= 0 && A.mb_rec != k) die_msg("manyBodyForces with the list of other settings");
    if (A.recs[k].n != n) die_msg("manyBodyForces for a different atom count than its list");
    if (addToLennardJones != (A.lj_rec >= 0))
        die_msg("manyBodyForces adds to Lennard-Jones forces exactly when lennardJonesForces is called");
    A.see_atoms(n);
    A.mb_rec = k;
    A.want.n_bond = n;
    A.ensure();
    A.upload(pos, nullptr, nullptr, nullptr, nullptr);
    // with Lennard-Jones on, the bond-order forces add to the caller's array;
    // otherwise they are assigned (atoms past the first n keep what they hold)
    A.upload_forces(force);
    if (mdx_bond_forces(A.sys) != MDX_OK) die("mdx_bond_forces", A.sys);
    A.download(nullptr, nullptr, nullptr, nullptr, nullptr, force);
}

void calculateAccelerations(real4* pos, real3* acc, real3* force, int n) {
    Adapter& A = ad();
    A.check_all(n, "calculateAccelerations");
    A.ensure();
    A.upload(pos, nullptr, nullptr, nullptr, nullptr);
    A.upload_forces(force);
    if (mdx_accelerations(A.sys) != MDX_OK) die("mdx_accelerations", A.sys);
    A.download(nullptr, nullptr, acc, nullptr, nullptr, nullptr);
}
