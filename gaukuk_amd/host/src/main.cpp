// main.cpp -- `gaukuk.sim -i input.in`: the reference's driver loop (src/main.cpp:12-65):
// load the configuration, build the Sim, run the setup plug-in, fill boundaries, write output 0,
// then Advance by dtoutput and write one output per interval until tmax (or stepmax).
// Extra flags: --setup-only (stop after output 0), --quiet (no per-step lines), --restart cons_XXXXX
// (continue from that output file of the same grid: its state and time; outputs resume at XXXXX+1).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <ctime>
#include <cstring>
#include <iomanip>
#include <iostream>

#include "sim.hpp"
#include "utils/utils.hpp"

int main(int argc, char* argv[]) {
    using namespace Gaukuk;
    if (argc < 3 || std::string(argv[1]) != "-i") {
        std::cerr << "Usage: " << argv[0] << " -i input.in [--setup-only] [--quiet] [--restart cons_XXXXX]"
                  << std::endl;
        return 1;
    }
    bool setupOnly = false, quiet = false;
    const char* restart = nullptr;
    for (int a = 3; a < argc; a++) {
        if (!std::strcmp(argv[a], "--setup-only")) setupOnly = true;
        else if (!std::strcmp(argv[a], "--quiet")) quiet = true;
        else if (!std::strcmp(argv[a], "--restart") && a + 1 < argc) restart = argv[++a];
    }
    try {
        Config::getInstance().loadFromFile(argv[2]);
        Real tmax = Config::getInstance().get("tmax");
        Real dtoutput = Config::getInstance().get("dtoutput");

        Sim sim;
        sim.SetVerbose(!quiet);
        sim.Setup();
        DataType outputType = DataType::Cons;
        if ((int)Config::getInstance().get("dataType") == 1) outputType = DataType::Prim;
        int outputStep = 0;
        if (restart) {
            // the file holds the state as the run left it: no boundary update, no output 0
            outputStep = sim.Restart(restart);
        } else {
            sim.boundary.UpdateBD(sim.cons, sim.grid, sim.domain, sim.eos);
            sim.MarkHostModified();
            sim.WriteData(outputStep, outputType);
        }
        Real tnow = sim.GetTime();
        if (setupOnly) return 0;

        auto time0 = std::chrono::high_resolution_clock::now();
        clock_t cpu0 = clock();
        while (std::abs(tmax - tnow) > 1e-12 && sim.isContinue) {
            dtoutput = std::min(dtoutput, tmax - tnow);
            sim.Advance(dtoutput);
            tnow = sim.GetTime();
            outputStep++;
            sim.WriteData(outputStep, outputType);
            std::cout << "output " << outputStep << std::endl;
        }
        auto time1 = std::chrono::high_resolution_clock::now();
        clock_t cpu1 = clock();
        double wall = std::chrono::duration<double>(time1 - time0).count();
        double cpu = double(cpu1 - cpu0) / CLOCKS_PER_SEC;
        std::cout << std::right << "Finish! "
                  << "\n    total steps = " << std::setw(6) << sim.GetStep()
                  << " , \n    walltime = " << std::setw(8) << std::fixed << std::setprecision(3) << wall
                  << " s, \n    cputime = " << std::setw(8) << std::fixed << std::setprecision(3) << cpu << " s"
                  << "\n-----------------------------------------------------------" << std::endl;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 2;
    }
    return 0;
}
