// config.cpp -- ConfigParser and Loader (common/ConfigParser.cpp,
// common/Loader.cpp) with the reference's parsing rules and file formats.
#include <cctype>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <sstream>

#include "mdx_host.h"

namespace mdxh {

namespace {

std::string upper(std::string s) {
    for (auto& c : s) c = (char)std::toupper((unsigned char)c);
    return s;
}

// from_string of ConfigParser.cpp:200-211: istringstream >> dec >> value
template <class T> bool from_string(T& t, const std::string& s) {
    std::istringstream iss(s);
    return !(iss >> std::dec >> t).fail();
}

}  // namespace

bool ConfigParser::ParseText(const std::string& text, std::string* err) {
    data_.clear();
    std::istringstream in(text);
    std::string line;
    int lineid = 0;
    for (;;) {
        line.clear();
        std::getline(in, line);
        if (in.eof() && line.empty()) break;
        if (in.fail()) {
            if (err) *err = "Unable to read line from configuration file.";
            return false;
        }
        if (line.compare(0, 2, "//") == 0 || line.empty()) continue;
        size_t pos;
        while ((pos = line.rfind("//")) != std::string::npos) line = line.substr(0, pos);
        pos = line.find('=');
        if (pos == std::string::npos) {
            if (err) *err = "Error in configuration file!\nLine: " + std::to_string(lineid);
            return false;
        }
        std::string key = line.substr(0, pos), value = line.substr(pos + 1);
        if (key.empty()) {
            if (err) *err = "Error in configuration file!\nLine: " + std::to_string(lineid);
            return false;
        }
        if (!value.empty()) {
            // trailing blanks are cut unless only the first character is left
            // (ConfigParser.cpp:76-88)
            int ipos = (int)value.size() - 1;
            while (value[ipos] == ' ' || value[ipos] == '\t') {
                ipos--;
                if (ipos < 0) break;
            }
            if (ipos > 0 && ipos != (int)value.size() - 1) value = value.substr(0, ipos + 1);
        }
        data_.emplace(upper(key), value);  // first occurrence wins (unordered_map::insert)
        lineid++;
    }
    return true;
}

bool ConfigParser::Parse(const std::string& filename, std::string* err) {
    std::ifstream ifs(filename.c_str());
    if (ifs.fail()) {
        if (err) *err = "Error opening configuration file.";
        return false;
    }
    std::stringstream ss;
    ss << ifs.rdbuf();
    return ParseText(ss.str(), err);
}

bool ConfigParser::GetString(std::string key, std::string& value) const {
    auto it = data_.find(upper(key));
    if (it == data_.end()) return false;
    value = it->second;
    return true;
}
bool ConfigParser::GetInt(std::string key, int& value) const {
    auto it = data_.find(upper(key));
    return it != data_.end() && from_string(value, it->second);
}
bool ConfigParser::GetFloat(std::string key, float& value) const {
    auto it = data_.find(upper(key));
    return it != data_.end() && from_string(value, it->second);
}
bool ConfigParser::GetDouble(std::string key, double& value) const {
    auto it = data_.find(upper(key));
    return it != data_.end() && from_string(value, it->second);
}
// operator>>(bool) accepts 0 and 1 only (ConfigParser.cpp:147-156)
bool ConfigParser::GetBoolean(std::string key, bool& value) const {
    auto it = data_.find(upper(key));
    return it != data_.end() && from_string(value, it->second);
}
bool ConfigParser::GetFloats(std::string key, std::vector<float>& value) const {
    auto it = data_.find(upper(key));
    if (it == data_.end()) return false;
    value.clear();
    std::string line = it->second;
    float v;
    while (from_string(v, line)) {
        value.push_back(v);
        size_t pos = line.find(' ');
        if (pos == std::string::npos) break;
        line = line.substr(pos + 1);
    }
    return !value.empty();
}

// key names of common/Loader.h:25-51
bool readConfiguration(const std::string& filename, Configuration* c, std::string* err) {
    ConfigParser cp;
    std::string perr;
    if (!cp.Parse(filename, &perr)) {
        if (err) *err = perr + "\nError in configuration file!";
        return false;
    }
    *c = Configuration();
    bool ok = true;
    ok &= cp.GetInt("Lennard Jones Particles", c->LennardJonesParticles);
    ok &= cp.GetInt("Timesteps", c->Timesteps);
    ok &= cp.GetFloat("DT", c->DT);
    ok &= cp.GetFloat("Temperature", c->Temperature);
    ok &= cp.GetFloat("LennardJonesRS", c->LennardJonesRS);
    ok &= cp.GetFloat("LennardJonesRCUT", c->LennardJonesRCUT);
    ok &= cp.GetFloat("ManyBodyRS", c->ManyBodyRS);
    ok &= cp.GetFloat("ManyBodyRCUT", c->ManyBodyRCUT);
    ok &= cp.GetString("Input", c->Input);
    ok &= cp.GetString("Output", c->Output);
    ok &= cp.GetString("Filename", c->Filename);
    ok &= cp.GetInt("CUDA Blocks", c->CudaBlocks);
    ok &= cp.GetBoolean("Use Cuda", c->UseCuda);
    ok &= cp.GetInt("OutputTimesteps", c->OutputTimesteps);
    ok &= cp.GetBoolean("Debug", c->Debug);
    ok &= cp.GetBoolean("Fallback", c->Fallback);
    ok &= cp.GetBoolean("UseLennardJones", c->useLennardJones);
    ok &= cp.GetBoolean("UseManyBody", c->useManyBody);
    ok &= cp.GetBoolean("LennardJonesBPP", c->lennardJonesBPP);
    ok &= cp.GetBoolean("ManyBodyBPP", c->manyBodyBPP);
    ok &= cp.GetFloat("VXCMFROM", c->vxcmFrom);
    ok &= cp.GetFloat("VXCMTO", c->vxcmTo);
    ok &= cp.GetFloat("VXCMSTEP", c->vxcmStep);
    ok &= cp.GetFloat("EnergyLoss", c->energyLoss);
    ok &= cp.GetInt("Many Body Particles", c->ManyBodyParticles);
    ok &= cp.GetInt("AnimationTimesteps", c->animts);
    ok &= cp.GetString("AnimationData", c->AnimFile);
    if (!ok) {
        if (err) *err = "Unable to read configuration file!";
        return false;
    }
    // optional keys; unknown keys are ignored as in the reference
    cp.GetDouble("BoxLength", c->BoxLength);
    cp.GetInt("Precision", c->Precision);
    cp.GetDouble("LJSigma", c->LJSigma);
    cp.GetDouble("LJEpsilon", c->LJEpsilon);
    cp.GetDouble("Mass", c->Mass);
    cp.GetInt("SweepBatch", c->SweepBatch);
    return true;
}

bool readInput(const Configuration& c, double* pos4, double* vel3, std::string* err) {
    std::ifstream in(c.Input.c_str());
    if (!in.good()) {
        if (err) *err = "Error opening input file.";
        return false;
    }
    const int n = c.LennardJonesParticles + c.ManyBodyParticles;
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < 4; j++) in >> pos4[4 * i + j];
        for (int j = 0; j < 3; j++) in >> vel3[3 * i + j];
        if (!in.good()) {  // Loader.cpp:150-151
            if (err) *err = "Error reading from input file.";
            return false;
        }
    }
    return true;
}

bool writeOutput(const Configuration& c, const double* pos4, const double* vel3, std::string* err) {
    std::ofstream out(c.Output.c_str(), std::ofstream::out | std::ofstream::trunc);
    if (!out.good()) {
        if (err) *err = "Error opening output file.";
        return false;
    }
    out << std::scientific;
    const int n = c.LennardJonesParticles + c.ManyBodyParticles;
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < 4; j++) out << pos4[4 * i + j] << " ";
        for (int j = 0; j < 3; j++) out << vel3[3 * i + j] << " ";
        if (!out.good()) {
            if (err) *err = "Error writing to output file file.";
            return false;
        }
    }
    return true;
}

}  // namespace mdxh
