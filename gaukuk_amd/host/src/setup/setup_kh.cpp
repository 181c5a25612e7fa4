// setup_kh.cpp -- Kelvin-Helmholtz shear layer: rho = rhoIn, vx = vxIn for |y - yc| < width,
// else rhoOut, vxOut; p = pressure; vy = amp*sin(4 pi (x - xmin)/Lx) + noise*U(-1,1) with U the
// SplitMix64 stream of `seed` keyed by the active cell index j*NX + i.
#include <cmath>

#include "../gaukuk.hpp"
#include "../sim.hpp"
#include "../utils/utils.hpp"
#include "splitmix.hpp"

namespace Gaukuk {

void Sim::Setup() {
    Config& c = Config::getInstance();
    const Real rhoIn = c.get("rhoIn", 2.0), rhoOut = c.get("rhoOut", 1.0);
    const Real vxIn = c.get("vxIn", 0.5), vxOut = c.get("vxOut", -0.5);
    const Real pressure = c.get("pressure", 2.5), amp = c.get("amp", 0.01), noise = c.get("noise", 0.01);
    const Real width = c.get("width", 0.25);
    const uint64_t seed = (uint64_t)c.get("seed", 0);
    const Real yc = 0.5 * (domain.ymin + domain.ymax), Lx = domain.xmax - domain.xmin;
    const double pi = std::acos(-1.0);
    for (int k = grid.kb; k < grid.ke; k++)
        for (int j = grid.jb; j < grid.je; j++)
            for (int i = grid.ib; i < grid.ie; i++) {
                const Real x = domain.xc(i), y = domain.yc(j);
                const bool inner = std::fabs(y - yc) < width;
                const uint64_t idx = (uint64_t)(j - grid.jb) * grid.nx + (i - grid.ib);
                prim(DEN, k, j, i) = inner ? rhoIn : rhoOut;
                prim(VLX, k, j, i) = inner ? vxIn : vxOut;
                prim(VLY, k, j, i) = amp * std::sin(4.0 * pi * (x - domain.xmin) / Lx) + noise * UniformPm1(seed, idx);
                prim(VLZ, k, j, i) = 0;
                prim(PRE, k, j, i) = pressure;
            }
    eos.PrimToCons(prim, cons, grid);
}

}  // namespace Gaukuk
