// cpumd_main.cpp -- the cpumd executable (cpu/cpumd/Main.cpp:18-27 with the
// Loader CLI of common/Loader.cpp:14-29): `mdx_cpumd [configuration-file]`,
// default.cfg when no argument is given.
#include <cstdlib>
#include <iostream>

#include "mdx_host.h"

int main(int argc, char** argv) {
    std::string configfile = "default.cfg";
    if (argc == 2)
        configfile = argv[1];
    else if (argc > 2) {
        std::cerr << "usage: mdx_cpumd [configuration-file].\n";
        return EXIT_FAILURE;
    }
    mdxh::Configuration cfg;
    std::string err;
    if (!mdxh::readConfiguration(configfile, &cfg, &err)) {
        std::cerr << err << std::endl;
        return EXIT_FAILURE;
    }
    std::cout << "Molecular::Dynamics\n";
    return mdxh::hostMain(&cfg);
}
