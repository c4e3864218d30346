// cpumd.cpp -- hostMain of cpu/cpumd/Host.cpp:14-240 driving the gfx950
// engine through include/mdx.h.  The loop keeps the reference's order of
// events (initial lists, initial output, timer, initial forces, steps,
// in-loop outputs at printout == t, animation frames, final output, files);
// steps between two events run device-resident in one mdx_run call.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <iomanip>
#include <iostream>

#include "mdx.h"
#include "mdx_host.h"

namespace mdxh {

namespace {

constexpr double kBoltzmann = (double)8.314e-7f;  // common/DebugCalculations.cpp:122

int report(mdx_system* s, const char* what) {
    std::cerr << "Error: " << what << ": " << mdx_last_error(s) << std::endl;
    return EXIT_FAILURE;
}

// performOutput (common/DebugCalculations.cpp:14-41) on device reductions
bool perform_output(mdx_system* s, std::vector<Results>& rv, int t, float dt, bool print, bool periodic) {
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
    rv.push_back(r);
    if (print)
        std::cout << "t=" << r.time << "\tT=" << r.temperature << "\tE=" << r.e << "\tEK=" << r.ek
                  << "\tEU=" << r.eu << std::endl;
    return true;
}

}  // namespace

int hostMain(Configuration* c) {
    const int n = c->LennardJonesParticles + c->ManyBodyParticles;
    const double boxSize = std::pow((double)n, 1.0 / 3.0) * 52.8;  // Host.cpp:31
    const bool periodic = c->BoxLength > 0;
    mdx_config mc{};
    mc.n_atoms = n;
    mc.n_bond = c->ManyBodyParticles;
    mc.box = periodic ? MDX_BOX_PERIODIC : MDX_BOX_WALLED;
    mc.precision = c->Precision == 0 ? MDX_PREC_MIXED : MDX_PREC_FP64;
    mc.use_lj = c->useLennardJones;
    mc.use_bond = c->useManyBody;
    mc.lj_table = periodic ? MDX_LJ_UNIFORM : MDX_LJ_REFERENCE;
    mc.bond_form = periodic ? MDX_BOND_TERSOFF : MDX_BOND_REFERENCE;
    mc.dt = c->DT;
    mc.lj_rs = c->LennardJonesRS; mc.lj_rcut = c->LennardJonesRCUT;
    mc.mb_rs = c->ManyBodyRS; mc.mb_rcut = c->ManyBodyRCUT;
    c->energyLoss = std::sqrt(c->energyLoss);  // Host.cpp:73 (float sqrt)
    mc.energy_loss = c->energyLoss;
    mc.wall = boxSize;
    mc.box_length = c->BoxLength;
    mc.lj_sigma = c->LJSigma; mc.lj_eps = c->LJEpsilon; mc.mass = c->Mass;
    mc.accel_scale = 1.0;
    mdx_system* s = nullptr;
    if (mdx_create(&mc, &s) != MDX_OK) return report(nullptr, "mdx_create");

    std::vector<double> pos(4 * (size_t)n), vel(3 * (size_t)n);
    std::vector<Results> rv;
    std::vector<std::vector<float>> posVec, velVec;
    std::string err;
    for (float vxcm = c->vxcmFrom; vxcm <= c->vxcmTo; vxcm += c->vxcmStep) {
        if (!readInput(*c, pos.data(), vel.data(), &err)) {
            std::cerr << "Error: " << err << std::endl;
            mdx_destroy(s);
            return EXIT_FAILURE;
        }
        for (int i = 0; i < n; i++) {  // Host.cpp:94-106
            for (int d = 0; d < 3 && !periodic; d++) {
                double& x = pos[4 * i + d];
                if (std::fabs(x) > boxSize / 2) x = (boxSize / 2) * (x / std::fabs(x));
            }
            vel[3 * i] += vxcm;
        }
        std::cout << std::fixed << std::setprecision(6);
        std::cout.setf(std::ios::fixed, std::ios::floatfield);
        std::cout << "Before Simulation [VXCM = " << vxcm << "]:" << std::endl;
        if (mdx_set_state(s, pos.data(), vel.data(), nullptr, nullptr, nullptr) != MDX_OK)
            return report(s, "mdx_set_state");
        double vmax, bsl;
        if (mdx_ready_list(s, &vmax, &bsl) != MDX_OK) return report(s, "readyList");
        int32_t nls, nlg;
        if (c->useLennardJones && mdx_build_list(s, MDX_LIST_LJ, bsl, &nls, &nlg) != MDX_OK)
            return report(s, "buildList (Lennard-Jones)");
        if (c->useManyBody && !periodic && mdx_build_list(s, MDX_LIST_BOND, bsl, &nls, &nlg) != MDX_OK)
            return report(s, "buildList (many body)");
        if (!perform_output(s, rv, 0, c->DT, true, periodic)) return report(s, "performOutput");
        std::cout << std::endl;
        if (c->animts > -1) pushAnim(posVec, velVec, pos.data(), vel.data(), n);

        auto start = std::chrono::steady_clock::now();
        if (mdx_run(s, 0) != MDX_OK) return report(s, "initial forces");  // Host.cpp:132-138
        int t = 0, printout = c->OutputTimesteps, animt = c->animts;
        while (t < c->Timesteps) {
            int e = c->Timesteps - 1;
            if (printout >= t && printout < e) e = printout;
            if (animt >= t && animt < e) e = animt;
            if (mdx_run(s, e - t + 1) != MDX_OK) return report(s, "step loop");
            t = e + 1;
            if (printout == e) {
                if (!perform_output(s, rv, e, c->DT, c->Debug, periodic)) return report(s, "performOutput");
                printout += c->OutputTimesteps;
            }
            if (animt == e) {
                if (mdx_get_state(s, pos.data(), vel.data(), nullptr, nullptr, nullptr, nullptr) != MDX_OK)
                    return report(s, "mdx_get_state");
                pushAnim(posVec, velVec, pos.data(), vel.data(), n);
                animt += c->animts;
            }
        }
        mdx_synchronize(s);
        auto stop = std::chrono::steady_clock::now();
        double elapsed = std::chrono::duration<double, std::milli>(stop - start).count();

        std::cout << "\nAfter Simulation [VXCM = " << vxcm << "]:" << std::endl;
        if (!perform_output(s, rv, t, c->DT, true, periodic)) return report(s, "performOutput");
        std::cout << "\nTime took: " << (long long)elapsed << "ms" << std::endl << std::endl;
        if (mdx_get_state(s, pos.data(), vel.data(), nullptr, nullptr, nullptr, nullptr) != MDX_OK)
            return report(s, "mdx_get_state");
        if (!writeOutput(*c, pos.data(), vel.data(), &err) ||
            !writeResults(rv, *c, (float)elapsed, tailResults(pos.data(), c->ManyBodyParticles), vxcm, &err) ||
            (c->animts > -1 && !writeAnimationBinaryData(posVec, velVec, *c, vxcm, &err))) {
            std::cerr << "Error: " << err << std::endl;
            mdx_destroy(s);
            return EXIT_FAILURE;
        }
        rv.clear();
        posVec.clear();
        velVec.clear();
        if (!(c->vxcmStep > 0)) break;  // a zero step would repeat forever
    }
    mdx_destroy(s);
    return EXIT_SUCCESS;
}

}  // namespace mdxh
