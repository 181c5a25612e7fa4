// setup_shock_tube.cpp -- Sod shock tube plug-in (keys of the reference's
// src/setup/setup_shock_tube.cpp:12-16): rho/p left of x0 and right of it, at rest.
#include "../gaukuk.hpp"
#include "../sim.hpp"
#include "../utils/utils.hpp"

namespace Gaukuk {

void Sim::Setup() {
    Config& c = Config::getInstance();
    const Real x0 = c.get("x0"), rhoL = c.get("rhoLeft"), rhoR = c.get("rhoRight");
    const Real pL = c.get("pressureLeft"), pR = c.get("pressureRight");
    for (int k = grid.kb; k < grid.ke; k++)
        for (int j = grid.jb; j < grid.je; j++)
            for (int i = grid.ib; i < grid.ie; i++) {
                const bool left = domain.xc(i) < x0;
                prim(DEN, k, j, i) = left ? rhoL : rhoR;
                prim(VLX, k, j, i) = 0;
                prim(VLY, k, j, i) = 0;
                prim(VLZ, k, j, i) = 0;
                prim(PRE, k, j, i) = left ? pL : pR;
            }
    eos.PrimToCons(prim, cons, grid);
}

}  // namespace Gaukuk
