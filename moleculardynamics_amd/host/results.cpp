// results.cpp -- result files of common/Results.cpp: the binary .dat record
// stream, the .txt log, the .anm animation frames and the N-O pair census.
#include <cmath>
#include <fstream>
#include <iomanip>
#include <sstream>

#include "mdx_host.h"

namespace mdxh {

std::string vxcmName(const std::string& stem, float vxcm, const char* ext) {
    std::ostringstream s;
    s << stem << "(vxcm-" << vxcm << ")." << ext;
    return s.str();
}

void pushAnim(std::vector<std::vector<float>>& posVec, std::vector<std::vector<float>>& velVec,
              const double* pos4, const double* vel3, int nop) {
    std::vector<float> p(4 * (size_t)nop), v(3 * (size_t)nop);
    for (size_t i = 0; i < 4 * (size_t)nop; i++) p[i] = (float)pos4[i];
    for (size_t i = 0; i < 3 * (size_t)nop; i++) v[i] = (float)vel3[i];
    posVec.push_back(std::move(p));
    velVec.push_back(std::move(v));
}

// Results.cpp:49-86: int nop, int animts, int count, all position frames
// (x, y, z, type as floats), then all velocity frames
bool writeAnimationBinaryData(const std::vector<std::vector<float>>& posVec,
                              const std::vector<std::vector<float>>& velVec, const Configuration& c,
                              float vxcm, std::string* err) {
    std::ofstream f(vxcmName(c.AnimFile, vxcm, "anm").c_str(),
                    std::ios::binary | std::ios::out | std::ios::trunc);
    if (f.fail()) {
        if (err) *err = "Error opening result file.";
        return false;
    }
    int nop = c.LennardJonesParticles + c.ManyBodyParticles, count = (int)posVec.size();
    f.write((const char*)&nop, sizeof(int));
    f.write((const char*)&c.animts, sizeof(int));
    f.write((const char*)&count, sizeof(int));
    for (auto& p : posVec) f.write((const char*)p.data(), sizeof(float) * p.size());
    for (auto& v : velVec) f.write((const char*)v.data(), sizeof(float) * v.size());
    if (f.fail()) {
        if (err) *err = "Error writing to result file.";
        return false;
    }
    return true;
}

// Results.cpp:88-118: unlike bond-order pairs closer than 2 reNO
std::string tailResults(const double* pos4, int mb) {
    int pairs = 0;
    std::ostringstream list, res;
    res << "ManyBody Pairs:\n";
    const double lim = (double)(2 * 1.15f);
    for (int i = 0; i < mb; i++)
        for (int j = i + 1; j < mb; j++) {
            if ((int)pos4[4 * i + 3] == (int)pos4[4 * j + 3]) continue;
            double dx = pos4[4 * i] - pos4[4 * j], dy = pos4[4 * i + 1] - pos4[4 * j + 1];
            double dz = pos4[4 * i + 2] - pos4[4 * j + 2];
            if (std::sqrt(dx * dx + dy * dy + dz * dz) < lim) {
                pairs++;
                list << "[" << i << ", " << j << "] ";
            }
        }
    res << "Amount: " << pairs << "\n" << list.str();
    return res.str();
}

// Results.cpp:120-229
bool writeResults(const std::vector<Results>& rv, const Configuration& c, float elapsed,
                  const std::string& tail, float vxcm, std::string* err) {
    std::ofstream result(vxcmName(c.Filename, vxcm, "dat").c_str(),
                         std::ios::binary | std::ios::out | std::ios::trunc);
    if (result.fail()) {
        if (err) *err = "Error opening result file.";
        return false;
    }
    std::ofstream log(vxcmName(c.Filename, vxcm, "txt").c_str(), std::ios::out | std::ios::trunc);
    if (log.fail()) {
        if (err) *err = "Error opening log file.";
        return false;
    }
    log << std::fixed << std::setprecision(6);
    log.setf(std::ios::fixed, std::ios::floatfield);
    log << "Number Of Particles = " << c.LennardJonesParticles + c.ManyBodyParticles
        << ", Timesteps = " << c.Timesteps << "\n";
    log << "DT = " << c.DT;
    if (c.useLennardJones)
        log << ", LennardJonesRS = " << c.LennardJonesRS << ", LennardJonesRCUT = " << c.LennardJonesRCUT;
    if (c.useManyBody) log << ", ManyBodyRS = " << c.ManyBodyRS << ", ManyBodyRCUT = " << c.ManyBodyRCUT;
    log << "\n\n";
    log << "\nVXCM [" << vxcm << "]: Time took: " << elapsed << " ms\n\n";
    log << "t\t\tT\t\tE\t\tEK\t\tEU\t\tR.x\t\tR.y\t\tR.z\t\tp.x\t\tp.y\t\tp.z\n";
    int nop = c.LennardJonesParticles + c.ManyBodyParticles;
    unsigned char use;
    result.write((const char*)&nop, sizeof(int));
    result.write((const char*)&c.Timesteps, sizeof(int));
    result.write((const char*)&c.DT, sizeof(float));
    use = c.useLennardJones ? 0xFF : 0x00;
    result.write((const char*)&use, 1);
    if (c.useLennardJones) {
        result.write((const char*)&c.LennardJonesRS, sizeof(float));
        result.write((const char*)&c.LennardJonesRCUT, sizeof(float));
    }
    use = c.useManyBody ? 0xFF : 0x00;
    result.write((const char*)&use, 1);
    if (c.useManyBody) {
        result.write((const char*)&c.ManyBodyRS, sizeof(float));
        result.write((const char*)&c.ManyBodyRCUT, sizeof(float));
    }
    result.write((const char*)&vxcm, sizeof(float));
    for (const Results& r : rv) {
        log << std::endl;
        const double rec[11] = {r.time, r.temperature, r.e, r.ek, r.eu, r.centerOfMassx, r.centerOfMassy,
                                r.centerOfMassz, r.momentumx, r.momentumy, r.momentumz};
        result.write((const char*)rec, sizeof rec);
        log << r.time << std::scientific;
        for (int k = 1; k < 11; k++) log << "\t" << rec[k];
        log << std::fixed;
    }
    log << "\n\n" << tail;
    if (result.fail() || log.fail()) {
        if (err) *err = "Error writing to result file.";
        return false;
    }
    return true;
}

}  // namespace mdxh
