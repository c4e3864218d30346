// host_capi.cpp -- C entry points into the host layer so the parser, loader
// and result writers can be exercised without a GPU (and from Python).
#include <cstring>
#include <sstream>

#include "mdx_host.h"

namespace {

int put(const std::string& s, char* out, int cap) {
    if (!out || cap <= 0) return -1;
    std::strncpy(out, s.c_str(), (size_t)cap - 1);
    out[cap - 1] = 0;
    return (int)s.size() < cap ? 0 : -1;
}

std::string to_json(const mdxh::Configuration& c) {
    std::ostringstream j;
    j.precision(9);
    j << "{\"LennardJonesParticles\":" << c.LennardJonesParticles << ",\"ManyBodyParticles\":" << c.ManyBodyParticles
      << ",\"DT\":" << c.DT << ",\"Timesteps\":" << c.Timesteps << ",\"Temperature\":" << c.Temperature
      << ",\"LennardJonesRS\":" << c.LennardJonesRS << ",\"LennardJonesRCUT\":" << c.LennardJonesRCUT
      << ",\"ManyBodyRS\":" << c.ManyBodyRS << ",\"ManyBodyRCUT\":" << c.ManyBodyRCUT
      << ",\"CudaBlocks\":" << c.CudaBlocks << ",\"UseCuda\":" << c.UseCuda
      << ",\"useLennardJones\":" << c.useLennardJones << ",\"useManyBody\":" << c.useManyBody
      << ",\"lennardJonesBPP\":" << c.lennardJonesBPP << ",\"manyBodyBPP\":" << c.manyBodyBPP
      << ",\"OutputTimesteps\":" << c.OutputTimesteps << ",\"Debug\":" << c.Debug << ",\"Fallback\":" << c.Fallback
      << ",\"vxcmFrom\":" << c.vxcmFrom << ",\"vxcmTo\":" << c.vxcmTo << ",\"vxcmStep\":" << c.vxcmStep
      << ",\"energyLoss\":" << c.energyLoss << ",\"animts\":" << c.animts << ",\"Input\":\"" << c.Input
      << "\",\"Output\":\"" << c.Output << "\",\"Filename\":\"" << c.Filename << "\",\"AnimFile\":\"" << c.AnimFile
      << "\",\"BoxLength\":" << c.BoxLength << ",\"Precision\":" << c.Precision
      << ",\"SweepBatch\":" << c.SweepBatch << "}";
    return j.str();
}

}  // namespace

extern "C" {

// parse a configuration file; writes JSON of the fields or the error text
int mdxh_read_configuration(const char* path, char* out, int cap) {
    mdxh::Configuration c;
    std::string err;
    if (!mdxh::readConfiguration(path, &c, &err)) {
        put(err, out, cap);
        return 1;
    }
    return put(to_json(c), out, cap);
}

// the vxcm sweep's member values for from/to/step; returns the count (values
// beyond cap are not written)
int mdxh_sweep_values(float from, float to, float step, float* out, int cap) {
    mdxh::Configuration c;
    c.vxcmFrom = from;
    c.vxcmTo = to;
    c.vxcmStep = step;
    const std::vector<float> v = mdxh::sweepValues(c);
    for (int i = 0; i < (int)v.size() && i < cap; i++) out[i] = v[i];
    return (int)v.size();
}

// ConfigParser on text: returns 0 and the value of `key` as a string, 1 when missing, 2 on parse error
int mdxh_config_lookup(const char* text, const char* key, char* out, int cap) {
    mdxh::ConfigParser cp;
    std::string err, v;
    if (!cp.ParseText(text, &err)) {
        put(err, out, cap);
        return 2;
    }
    if (!cp.GetString(key, v)) return 1;
    return put(v, out, cap);
}

int mdxh_write_output(const char* path, int n, const double* pos4, const double* vel3) {
    mdxh::Configuration c;
    c.LennardJonesParticles = n;
    c.Output = path;
    return mdxh::writeOutput(c, pos4, vel3, nullptr) ? 0 : 1;
}

int mdxh_read_input(const char* path, int n, double* pos4, double* vel3) {
    mdxh::Configuration c;
    c.LennardJonesParticles = n;
    c.Input = path;
    return mdxh::readInput(c, pos4, vel3, nullptr) ? 0 : 1;
}

// write .dat/.txt for `count` records of 11 doubles with the given header fields
int mdxh_write_results(const char* stem, int nop, int timesteps, float dt, int use_lj, float lj_rs, float lj_rc,
                       int use_mb, float mb_rs, float mb_rc, float vxcm, float elapsed, const double* rec,
                       int count, const double* pos4, int mb) {
    mdxh::Configuration c;
    c.LennardJonesParticles = nop;
    c.Timesteps = timesteps;
    c.DT = dt;
    c.useLennardJones = use_lj != 0;
    c.LennardJonesRS = lj_rs; c.LennardJonesRCUT = lj_rc;
    c.useManyBody = use_mb != 0;
    c.ManyBodyRS = mb_rs; c.ManyBodyRCUT = mb_rc;
    c.Filename = stem;
    std::vector<mdxh::Results> rv(count);
    for (int k = 0; k < count; k++) std::memcpy(&rv[k], rec + 11 * k, sizeof(mdxh::Results));
    return mdxh::writeResults(rv, c, elapsed, mdxh::tailResults(pos4, mb), vxcm, nullptr) ? 0 : 1;
}

}  // extern "C"
