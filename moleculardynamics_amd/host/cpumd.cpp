// cpumd.cpp -- hostMain of cpu/cpumd/Host.cpp:14-240 driving the gfx950
// engine through include/mdx.h.  Each vxcm member keeps the reference's order
// of events (initial lists, initial output, timer, initial forces, steps,
// in-loop outputs at printout == t, animation frames, final output, files);
// steps between two events run device-resident in one mdx_run call.
//
// The reference runs the sweep's members one after another
// (Host.cpp:75-216).  They are independent systems, and one small system
// leaves most of the GPU idle, so here up to SweepStreams members run at
// once, each on a host thread with its own mdx_system (and so its own HIP
// stream).  Every member's text is buffered and printed, and its files are
// written, in sweep order by the main thread: stdout and the files are those
// of the sequential sweep, except that "Time took" is each member's own wall
// time while it shared the GPU.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <future>
#include <iomanip>
#include <iostream>
#include <memory>
#include <mutex>
#include <sstream>
#include <thread>

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

// The body of the sweep loop (Host.cpp:84-197) for one vxcm on system s.
bool run_member(mdx_system* s, const Configuration& c, int n, double boxSize, Member& m) {
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
    if (c.animts > -1) pushAnim(m.posVec, m.velVec, pos, vel, n);

    auto start = std::chrono::steady_clock::now();
    if (mdx_run(s, 0) != MDX_OK) return report(s, m, "initial forces");  // Host.cpp:132-138
    int t = 0, printout = c.OutputTimesteps, animt = c.animts;
    while (t < c.Timesteps) {
        int e = c.Timesteps - 1;
        if (printout >= t && printout < e) e = printout;
        if (animt >= t && animt < e) e = animt;
        if (mdx_run(s, e - t + 1) != MDX_OK) return report(s, m, "step loop");
        t = e + 1;
        if (printout == e) {
            if (!perform_output(s, m, e, c.DT, c.Debug, periodic)) return report(s, m, "performOutput");
            printout += c.OutputTimesteps;
        }
        if (animt == e) {
            if (mdx_get_state(s, pos, vel, nullptr, nullptr, nullptr, nullptr) != MDX_OK)
                return report(s, m, "mdx_get_state");
            pushAnim(m.posVec, m.velVec, pos, vel, n);
            animt += c.animts;
        }
    }
    mdx_synchronize(s);
    auto stop = std::chrono::steady_clock::now();
    m.elapsed = std::chrono::duration<double, std::milli>(stop - start).count();

    out << "\nAfter Simulation [VXCM = " << m.vxcm << "]:" << std::endl;
    if (!perform_output(s, m, t, c.DT, true, periodic)) return report(s, m, "performOutput");
    out << "\nTime took: " << (long long)m.elapsed << "ms" << std::endl << std::endl;
    if (mdx_get_state(s, pos, vel, nullptr, nullptr, nullptr, nullptr) != MDX_OK)
        return report(s, m, "mdx_get_state");
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
    const int workers = std::max(1, std::min(c->SweepStreams, count));
    const Configuration& cfg = *c;

    std::vector<std::unique_ptr<Member>> members(count);
    std::vector<std::promise<void>> done(count);
    std::vector<std::future<void>> ready;
    for (int i = 0; i < count; i++) {
        members[i].reset(new Member);
        members[i]->vxcm = values[i];
        ready.push_back(done[i].get_future());
    }
    // Workers take members in sweep order, at most 2 x workers ahead of the
    // last member written out, so finished members do not pile up in memory.
    std::mutex mu;
    std::condition_variable cv;
    int next = 0, written = 0;
    bool stop = false;
    auto worker = [&]() {
        mdx_system* s = nullptr;
        std::string create_err;
        if (mdx_create(&mc, &s) != MDX_OK) create_err = std::string("mdx_create: ") + mdx_last_error(nullptr);
        for (;;) {
            int i;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || next >= count || next < written + 2 * workers; });
                if (stop || next >= count) break;
                i = next++;
            }
            Member& m = *members[i];
            if (!create_err.empty()) m.err = create_err;
            else m.ok = run_member(s, cfg, n, boxSize, m);
            done[i].set_value();
        }
        if (s) mdx_destroy(s);
    };
    std::vector<std::thread> pool;
    for (int w = 0; w < workers; w++) pool.emplace_back(worker);

    int status = EXIT_SUCCESS;
    for (int i = 0; i < count; i++) {
        ready[i].wait();
        Member& m = *members[i];
        std::cout << m.out.str() << std::flush;
        std::string err = m.err;
        bool ok = m.ok;
        if (ok && (!writeOutput(cfg, m.pos.data(), m.vel.data(), &err) ||
                   !writeResults(m.rv, cfg, (float)m.elapsed, tailResults(m.pos.data(), cfg.ManyBodyParticles),
                                 m.vxcm, &err) ||
                   (cfg.animts > -1 && !writeAnimationBinaryData(m.posVec, m.velVec, cfg, m.vxcm, &err))))
            ok = false;
        members[i].reset();
        {
            std::lock_guard<std::mutex> lk(mu);
            written = i + 1;
            if (!ok) stop = true;
        }
        cv.notify_all();
        if (!ok) {
            std::cerr << "Error: " << err << std::endl;
            status = EXIT_FAILURE;
            break;
        }
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        stop = true;
    }
    cv.notify_all();
    for (auto& th : pool) th.join();
    return status;
}

}  // namespace mdxh
