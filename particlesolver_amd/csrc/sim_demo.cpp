// sim_demo.cpp — drives the C++ Simulation facade the way the reference's
// CPU app does (cpu/src/view.cpp:185-199: Simulation::init(scene), then
// tick(.01) per frame).  Usage: sim_demo <scene> <ticks> [trajectory]
// with scene one of fluid, gas, smoke_open, smoke_closed, volcano.  Prints
// the particle count and kinetic energy; with a path, writes the final
// positions and phases as a one-frame trajectory file.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "simulation.hpp"

using namespace ps;

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <fluid|gas|smoke_open|smoke_closed|volcano> <ticks> [out]\n",
                     argv[0]);
        return 2;
    }
    const std::string scene = argv[1];
    const int ticks = std::atoi(argv[2]);
    try {
        Simulation sim;
        if (scene == "fluid") sim.init(FLUID_TEST);
        else if (scene == "gas") sim.initGas();
        else if (scene == "smoke_open") sim.init(SMOKE_OPEN_TEST);
        else if (scene == "smoke_closed") sim.init(SMOKE_CLOSED_TEST);
        else if (scene == "volcano") sim.init(VOLCANO_TEST);
        else {
            std::fprintf(stderr, "unknown scene %s\n", scene.c_str());
            return 2;
        }
        for (int t = 0; t < ticks; t++) sim.tick(.01);
        std::printf("%s: ticks %d particles %d kinetic_energy %.9g\n", scene.c_str(), ticks,
                    sim.getNumParticles(), sim.getKineticEnergy());
        if (argc > 3) {
            ps_trajectory *tr = nullptr;
            int rc = ps_trajectory_open(argv[3], PS_FIELD_POSITIONS | PS_FIELD_PHASES, 2, &tr);
            if (rc == PS_OK) rc = ps_trajectory_write(tr, sim.handle());
            const int rc2 = ps_trajectory_close(tr);
            if (rc != PS_OK || rc2 != PS_OK) {
                std::fprintf(stderr, "trajectory: error %d %d\n", rc, rc2);
                return 1;
            }
        }
    } catch (const Error &e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
