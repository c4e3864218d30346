// cpumd.cpp -- hostMain of cpu/cpumd/Host.cpp:14-240 driving the gfx950
// engine through include/mdx.h.  Each vxcm member keeps the reference's order
// of events (initial lists, initial output, timer, initial forces, steps,
// in-loop outputs at printout == t, animation frames, final output, files);
// steps between two events run device-resident.
//
// The reference runs the sweep's members one after another
// (Host.cpp:75-216).  They are independent systems of one configuration, and
// one small system leaves most of the GPU idle, so here up to SweepBatch
// members form one mdx_sweep: every step kernel is launched once for all of
// them, each member keeping its own list schedule and wall flag.  Members of
// a batch share their events, so they advance in lockstep between them.
// Every member's text is printed, and its files are written, in sweep order:
// stdout and the files are those of the sequential sweep, except that "Time
// took" is the wall time of the member's batch.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>

#include "mdx.h"
#include "mdx_host.h"

namespace mdxh {

namespace {

constexpr double kBoltzmann = (double)8.314e-7f;  // common/DebugCalculations.cpp:122

// One member of the vxcm sweep: what it prints, what its files hold.
struct Member {
    float vxcm = 0;
    std::ostringstream out;  // stdout text of the member
    std::string err;         // the "Error: ..." line's text when the member failed
    std::vector<double> pos, vel;
    std::vector<Results> rv;
    std::vector<std::vector<float>> posVec, velVec;
    double elapsed = 0;  // ms
    bool ok = false;
};

bool report(mdx_system* s, Member& m, const char* what) {
    m.err = std::string(what) + ": " + mdx_last_error(s);
    return false;
}

// performOutput (common/DebugCalculations.cpp:14-41) on device reductions
bool perform_output(mdx_system* s, Member& m, int t, float dt, bool print, bool periodic) {
    double k[8], u[2];
    if (mdx_kinetics(s, periodic ? 1.0 : kBoltzmann, k) != MDX_OK) return false;
    if ((periodic ? mdx_potential(s, u) : mdx_potential_allpairs(s, u)) != MDX_OK) return false;
    Results r;
    r.eu = u[1] + u[0];
    r.ek = k[0];
    r.e = r.ek + r.eu;
    r.temperature = k[1];
    r.time = t * (double)dt;
    r.centerOfMassx = k[2]; r.centerOfMassy = k[3]; r.centerOfMassz = k[4];
    r.momentumx = k[5]; r.momentumy = k[6]; r.momentumz = k[7];
    m.rv.push_back(r);
    if (print)
        m.out << "t=" << r.time << "\tT=" << r.temperature << "\tE=" << r.e << "\tEK=" << r.ek
              << "\tEU=" << r.eu << std::endl;
    return true;
}

mdx_config make_config(const Configuration& c, int n, double boxSize) {
    const bool periodic = c.BoxLength > 0;
    mdx_config mc{};
    mc.n_atoms = n;
    mc.n_bond = c.ManyBodyParticles;
    mc.box = periodic ? MDX_BOX_PERIODIC : MDX_BOX_WALLED;
    mc.precision = c.Precision == 0 ? MDX_PREC_MIXED : MDX_PREC_FP64;
    mc.use_lj = c.useLennardJones;
    mc.use_bond = c.useManyBody;
    mc.lj_table = periodic ? MDX_LJ_UNIFORM : MDX_LJ_REFERENCE;
    mc.bond_form = periodic ? MDX_BOND_TERSOFF : MDX_BOND_REFERENCE;
    mc.dt = c.DT;
    mc.lj_rs = c.LennardJonesRS; mc.lj_rcut = c.LennardJonesRCUT;
    mc.mb_rs = c.ManyBodyRS; mc.mb_rcut = c.ManyBodyRCUT;
    mc.energy_loss = c.energyLoss;
    mc.wall = boxSize;
    mc.box_length = c.BoxLength;
    mc.lj_sigma = c.LJSigma; mc.lj_eps = c.LJEpsilon; mc.mass = c.Mass;
    mc.accel_scale = 1.0;
    return mc;
}

// The sweep loop's set-up for one vxcm on system s (Host.cpp:84-128): read
// the input, clamp it into +-boxSize/2, add vxcm, build the lists and write
// the initial output.
bool set_up_member(mdx_system* s, const Configuration& c, int n, double boxSize, Member& m) {
    const bool periodic = c.BoxLength > 0;
    m.pos.assign(4 * (size_t)n, 0.0);
    m.vel.assign(3 * (size_t)n, 0.0);
    double* pos = m.pos.data();
    double* vel = m.vel.data();
    if (!readInput(c, pos, vel, &m.err)) return false;
    for (int i = 0; i < n; i++) {  // Host.cpp:94-106
        for (int d = 0; d < 3 && !periodic; d++) {
            double& x = pos[4 * i + d];
            if (std::fabs(x) > boxSize / 2) x = (boxSize / 2) * (x / std::fabs(x));
        }
        vel[3 * i] += m.vxcm;
    }
    std::ostream& out = m.out;
    out << std::fixed << std::setprecision(6);
    out << "Before Simulation [VXCM = " << m.vxcm << "]:" << std::endl;
    if (mdx_set_state(s, pos, vel, nullptr, nullptr, nullptr) != MDX_OK) return report(s, m, "mdx_set_state");
    double vmax, bsl;
    if (mdx_ready_list(s, &vmax, &bsl) != MDX_OK) return report(s, m, "readyList");
    int32_t nls, nlg;
    if (c.useLennardJones && mdx_build_list(s, MDX_LIST_LJ, bsl, &nls, &nlg) != MDX_OK)
        return report(s, m, "buildList (Lennard-Jones)");
    if (c.useManyBody && !periodic && mdx_build_list(s, MDX_LIST_BOND, bsl, &nls, &nlg) != MDX_OK)
        return report(s, m, "buildList (many body)");
    if (!perform_output(s, m, 0, c.DT, true, periodic)) return report(s, m, "performOutput");
    out << std::endl;
    if (c.animts > -1) pushAnim(m.posVec, m.velVec, m.pos.data(), m.vel.data(), n);
    return true;
}

// The members of one batch through the sweep loop's body (Host.cpp:84-197)
// in lockstep: every member has the same events (in-loop outputs at
// printout == t, animation frames), so the steps between two events advance
// all of them at once (advance(k): k steps of every member).  Each member's
// "Time took" is the batch's wall time.  Returns false when a member failed
// (its err says why) or advance did (it keeps its own message).
template <class Advance>
bool run_batch(const std::vector<mdx_system*>& sys, const std::vector<Member*>& ms, const Configuration& c, int n,
               double boxSize, Advance&& advance) {
    const bool periodic = c.BoxLength > 0;
    const size_t R = sys.size();
    for (size_t k = 0; k < R; k++)
        if (!set_up_member(sys[k], c, n, boxSize, *ms[k])) return false;
    auto start = std::chrono::steady_clock::now();
    if (!advance(0, "initial forces")) return false;   // Host.cpp:132-138
    int t = 0, printout = c.OutputTimesteps, animt = c.animts;
    while (t < c.Timesteps) {
        int e = c.Timesteps - 1;
        if (printout >= t && printout < e) e = printout;
        if (animt >= t && animt < e) e = animt;
        if (!advance(e - t + 1, "step loop")) return false;
        t = e + 1;
        if (printout == e) {
            for (size_t k = 0; k < R; k++)
                if (!perform_output(sys[k], *ms[k], e, c.DT, c.Debug, periodic))
                    return report(sys[k], *ms[k], "performOutput");
            printout += c.OutputTimesteps;
        }
        if (animt == e) {
            for (size_t k = 0; k < R; k++) {
                Member& m = *ms[k];
                if (mdx_get_state(sys[k], m.pos.data(), m.vel.data(), nullptr, nullptr, nullptr, nullptr) != MDX_OK)
                    return report(sys[k], m, "mdx_get_state");
                pushAnim(m.posVec, m.velVec, m.pos.data(), m.vel.data(), n);
            }
            animt += c.animts;
        }
    }
    for (mdx_system* s : sys) mdx_synchronize(s);
    auto stop = std::chrono::steady_clock::now();
    const double elapsed = std::chrono::duration<double, std::milli>(stop - start).count();
    for (size_t k = 0; k < R; k++) {
        Member& m = *ms[k];
        m.elapsed = elapsed;
        m.out << "\nAfter Simulation [VXCM = " << m.vxcm << "]:" << std::endl;
        if (!perform_output(sys[k], m, t, c.DT, true, periodic)) return report(sys[k], m, "performOutput");
        m.out << "\nTime took: " << (long long)m.elapsed << "ms" << std::endl << std::endl;
        if (mdx_get_state(sys[k], m.pos.data(), m.vel.data(), nullptr, nullptr, nullptr, nullptr) != MDX_OK)
            return report(sys[k], m, "mdx_get_state");
        m.ok = true;
    }
    return true;
}

}  // namespace

// The sweep's vxcm values, by the reference's float loop (Host.cpp:75); a
// step that is not positive gives one member instead of looping forever.
std::vector<float> sweepValues(const Configuration& c) {
    std::vector<float> v;
    for (float vxcm = c.vxcmFrom; vxcm <= c.vxcmTo; vxcm += c.vxcmStep) {
        v.push_back(vxcm);
        if (!(c.vxcmStep > 0)) break;
    }
    return v;
}

int hostMain(Configuration* c) {
    const int n = c->LennardJonesParticles + c->ManyBodyParticles;
    const double boxSize = std::pow((double)n, 1.0 / 3.0) * 52.8;  // Host.cpp:31
    c->energyLoss = std::sqrt(c->energyLoss);  // Host.cpp:73 (float sqrt)
    const mdx_config mc = make_config(*c, n, boxSize);
    const std::vector<float> values = sweepValues(*c);
    const int count = (int)values.size();
    const bool periodic = c->BoxLength > 0;
    // walled members run as one mdx_sweep per batch; the periodic box (not a
    // configuration of the reference) one member at a time
    const int batch = periodic ? 1 : std::max(1, std::min(c->SweepBatch, count));
    const Configuration& cfg = *c;

    for (int b0 = 0; b0 < count; b0 += batch) {
        const int nb = std::min(batch, count - b0);
        std::vector<std::unique_ptr<Member>> members(nb);
        std::vector<Member*> ms(nb);
        for (int k = 0; k < nb; k++) {
            members[k].reset(new Member);
            members[k]->vxcm = values[b0 + k];
            ms[k] = members[k].get();
        }
        std::vector<mdx_system*> sys(nb, nullptr);
        mdx_sweep* w = nullptr;
        mdx_system* single = nullptr;
        std::string batch_err;
        bool ok = true;
        if (periodic) {
            if (mdx_create(&mc, &single) != MDX_OK) {
                batch_err = std::string("mdx_create: ") + mdx_last_error(nullptr);
                ok = false;
            }
            sys[0] = single;
        } else if (mdx_sweep_create(&mc, nb, &w) != MDX_OK) {
            batch_err = std::string("mdx_sweep_create: ") + mdx_sweep_last_error(nullptr);
            ok = false;
        } else {
            for (int k = 0; k < nb; k++) mdx_sweep_member(w, k, &sys[k]);
        }
        if (ok) {
            auto advance = [&](int steps, const char* what) {
                int rc = periodic ? mdx_run(single, steps) : mdx_sweep_run(w, steps);
                if (rc != MDX_OK) {
                    batch_err = std::string(what) + ": " + (periodic ? mdx_last_error(single) : mdx_sweep_last_error(w));
                    return false;
                }
                return true;
            };
            ok = run_batch(sys, ms, cfg, n, boxSize, advance);
        }
        if (w) mdx_sweep_destroy(w);
        if (single) mdx_destroy(single);
        // output in sweep order; the first failure ends the sweep, and a
        // member stopped by another's failure reports that one's error
        std::string first_err = batch_err;
        for (Member* m : ms)
            if (!m->err.empty()) {
                first_err = m->err;
                break;
            }
        for (int k = 0; k < nb; k++) {
            Member& m = *ms[k];
            std::cout << m.out.str() << std::flush;
            std::string err = m.err.empty() ? first_err : m.err;
            bool mok = m.ok;
            if (mok && (!writeOutput(cfg, m.pos.data(), m.vel.data(), &err) ||
                        !writeResults(m.rv, cfg, (float)m.elapsed, tailResults(m.pos.data(), cfg.ManyBodyParticles),
                                      m.vxcm, &err) ||
                        (cfg.animts > -1 && !writeAnimationBinaryData(m.posVec, m.velVec, cfg, m.vxcm, &err))))
                mok = false;
            if (!mok) {
                std::cerr << "Error: " << err << std::endl;
                return EXIT_FAILURE;
            }
        }
    }
    return EXIT_SUCCESS;
}

}  // namespace mdxh
