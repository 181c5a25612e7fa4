// setup_blast.cpp -- 2D circular overpressure: rho = rho0 everywhere, p = pressure_low*pRatio
// inside radius r0 of (x0, y0) (default: the domain centre), pressure_low outside, at rest.
#include <cmath>

#include "../gaukuk.hpp"
#include "../sim.hpp"
#include "../utils/utils.hpp"

namespace Gaukuk {

void Sim::Setup() {
    Config& c = Config::getInstance();
    const Real rho0 = c.get("rho0", 1.0), pLow = c.get("pressure_low", 0.1), ratio = c.get("pRatio", 50.0);
    const Real r0 = c.get("r0", 0.1);
    const Real x0 = c.get("x0", 0.5 * (domain.xmin + domain.xmax)), y0 = c.get("y0", 0.5 * (domain.ymin + domain.ymax));
    for (int k = grid.kb; k < grid.ke; k++)
        for (int j = grid.jb; j < grid.je; j++)
            for (int i = grid.ib; i < grid.ie; i++) {
                const Real dx = domain.xc(i) - x0, dy = domain.yc(j) - y0;
                const Real r = std::sqrt(dx * dx + dy * dy);
                prim(DEN, k, j, i) = rho0;
                prim(VLX, k, j, i) = 0;
                prim(VLY, k, j, i) = 0;
                prim(VLZ, k, j, i) = 0;
                prim(PRE, k, j, i) = r < r0 ? pLow * ratio : pLow;
            }
    eos.PrimToCons(prim, cons, grid);
}

}  // namespace Gaukuk
