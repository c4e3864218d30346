// mdx_host.h -- the cpumd host layer kept from the reference, re-implemented
// over the C-ABI of include/mdx.h:
//   ConfigParser   common/ConfigParser.{h,cpp}  KEY=VALUE parser
//   Loader         common/Loader.{h,cpp}        27 mandatory keys, text I/O
//   Results        common/Results.{h,cpp}       .dat / .txt / .anm writers
//   hostMain       cpu/cpumd/Host.cpp           the vxcm sweep and step loop
// File formats and parsing rules are those of the reference; errors are
// reported through return values (the driver turns them into the
// reference's "Error: ..." message and exit status).
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

namespace mdxh {

// common/Global.h:75-104 plus optional keys the reference does not have.
struct Configuration {
    int LennardJonesParticles = 0;
    int ManyBodyParticles = 0;
    float DT = 0;
    int Timesteps = 0;
    float Temperature = 0;
    float LennardJonesRS = 0, LennardJonesRCUT = 0;
    float ManyBodyRS = 0, ManyBodyRCUT = 0;
    int CudaBlocks = 0;
    bool UseCuda = false;
    bool useLennardJones = false, useManyBody = false;
    bool lennardJonesBPP = false, manyBodyBPP = false;
    int OutputTimesteps = 0;
    bool Debug = false, Fallback = false;
    float vxcmFrom = 0, vxcmTo = 0, vxcmStep = 0;
    float energyLoss = 0;
    int animts = 0;
    std::string Input, Output, Filename, AnimFile;
    // optional keys (absent -> reference behaviour)
    double BoxLength = 0;       // BOXLENGTH: > 0 selects the periodic box
    int Precision = 1;          // PRECISION: 0 mixed (FP32 pair), 1 FP64 (the reference's real = double)
    double LJSigma = 1, LJEpsilon = 1, Mass = 1;  // LJSIGMA / LJEPSILON / MASS for the periodic box
    int SweepBatch = 8;         // SWEEPBATCH: vxcm sweep members advanced together (one launch per step kernel)
};

// common/ConfigParser.cpp:22-98 semantics: "//" comment lines and trailing
// comments, KEY=VALUE split at the first '=', keys upper-cased and not
// trimmed, trailing blanks of a value trimmed (not when only its first
// character is non-blank), the first occurrence of a key wins.
class ConfigParser {
  public:
    bool Parse(const std::string& filename, std::string* err);
    bool ParseText(const std::string& text, std::string* err);
    bool GetString(std::string key, std::string& value) const;
    bool GetInt(std::string key, int& value) const;
    bool GetFloat(std::string key, float& value) const;
    bool GetDouble(std::string key, double& value) const;
    bool GetBoolean(std::string key, bool& value) const;
    bool GetFloats(std::string key, std::vector<float>& value) const;

  private:
    std::unordered_map<std::string, std::string> data_;
};

// common/Loader.cpp:34-83: all 27 keys are mandatory.
bool readConfiguration(const std::string& filename, Configuration* cfg, std::string* err);
// common/Loader.cpp:125-155 / 88-120: "x y z type vx vy vz" per atom; output in
// scientific notation with single spaces and no newlines.
bool readInput(const Configuration& cfg, double* pos4, double* vel3, std::string* err);
bool writeOutput(const Configuration& cfg, const double* pos4, const double* vel3, std::string* err);

// common/Results.h:15-28
struct Results {
    double time, temperature, e, ek, eu, centerOfMassx, centerOfMassy, centerOfMassz, momentumx,
        momentumy, momentumz;
};

// common/Results.cpp:12-229
void pushAnim(std::vector<std::vector<float>>& posVec, std::vector<std::vector<float>>& velVec,
              const double* pos4, const double* vel3, int nop);
bool writeAnimationBinaryData(const std::vector<std::vector<float>>& posVec,
                              const std::vector<std::vector<float>>& velVec, const Configuration& cfg,
                              float vxcm, std::string* err);
std::string tailResults(const double* pos4, int manyBodyParticles);
bool writeResults(const std::vector<Results>& res, const Configuration& cfg, float elapsed_ms,
                  const std::string& tail, float vxcm, std::string* err);
// "<stem>(vxcm-<v>).<ext>" as the reference's stringstream prints a float
std::string vxcmName(const std::string& stem, float vxcm, const char* ext);

// the vxcm values of the sweep (cpu/cpumd/Host.cpp:75, float accumulation);
// one value when the step is not positive
std::vector<float> sweepValues(const Configuration& cfg);
// cpu/cpumd/Host.cpp:14-240 on the engine: returns the process exit status.
int hostMain(Configuration* cfg);

}  // namespace mdxh
