// main.cpp -- the cgppd command line (src/main.cpp:3-44) on the device path.
#include <unistd.h>

#include <cstdio>
#include <exception>

#include "cgppd_host.h"

int main(int argc, char **argv)
{
    try {
        cgppd::Simulation sim;
        sim.init(argc, argv, (int)getpid());
        if (!sim.parameters.skipsimulation) sim.run();
        std::printf("cgppd: %zu replicas, %d MC steps each, %d exchanges accepted of %d\n", sim.replica.size(),
                    sim.parameters.MCsteps, sim.totalExchanges, sim.totalTests);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "cgppd: %s\n", e.what());
        return 1;
    }
    return 0;
}
