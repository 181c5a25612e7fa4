// setup_rt.cpp -- Rayleigh-Taylor with constant gravity: rho = rhoUp above yInterface, rhoDown
// below; hydrostatic p = p0 + gravAccY*rho*(y - yInterface); vy = amp*(1 + cos(4 pi x))*
// (1 + cos(3 pi (y - yInterface)))/4; gravity (0, gravAccY, 0) enrolled as a source
// (the reference's setup_rt.cpp:11-70 with the interface moved to yInterface).
#include <cmath>

#include "../gaukuk.hpp"
#include "../sim.hpp"
#include "../utils/utils.hpp"

namespace Gaukuk {

void Sim::Setup() {
    Config& c = Config::getInstance();
    const Real rhoUp = c.get("rhoUp", 2.0), rhoDown = c.get("rhoDown", 1.0);
    const Real gy = c.get("gravAccY", -0.1), p0 = c.get("p0", 2.5), amp = c.get("amp", 0.01);
    const Real y0 = c.get("yInterface", 0.75);
    const double pi = std::acos(-1.0);
    for (int k = grid.kb; k < grid.ke; k++)
        for (int j = grid.jb; j < grid.je; j++)
            for (int i = grid.ib; i < grid.ie; i++) {
                const Real x = domain.xc(i), y = domain.yc(j);
                const Real rho = y > y0 ? rhoUp : rhoDown;
                prim(DEN, k, j, i) = rho;
                prim(VLX, k, j, i) = 0;
                prim(VLY, k, j, i) = amp * (1.0 + std::cos(4.0 * pi * x)) * (1.0 + std::cos(3.0 * pi * (y - y0))) / 4.0;
                prim(VLZ, k, j, i) = 0;
                prim(PRE, k, j, i) = p0 + gy * rho * (y - y0);
            }
    srcTerm.EnrollConstGravityVector(0, gy, 0);
    eos.PrimToCons(prim, cons, grid);
}

}  // namespace Gaukuk
