// setup_solid_body.cpp -- supersonic flow past a reflective cylinder: a uniform state (rho, vxIn, p)
// fed through a self-defined x-left face, and an immersed solid of radius r_solid at (x0, y0, z0).
// Keys as the reference's solid-body problem (input/solid_body.in): rho, p, vxIn, x0, y0, z0,
// r_solid.  The inflow face writes the same state whatever the flow does, so it runs on the device
// as a fixed face.
#include "../gaukuk.hpp"
#include "../sim.hpp"
#include "../utils/utils.hpp"

namespace Gaukuk {

void Sim::Setup() {
    Config& c = Config::getInstance();
    const Real rho = c.get("rho"), pre = c.get("p"), vxIn = c.get("vxIn");
    const Real x0 = c.get("x0", 0.0), y0 = c.get("y0", 0.0), z0 = c.get("z0", 0.0), rSolid = c.get("r_solid");
    for (int k = grid.kb; k < grid.ke; k++)
        for (int j = grid.jb; j < grid.je; j++)
            for (int i = grid.ib; i < grid.ie; i++) {
                prim(DEN, k, j, i) = rho;
                prim(VLX, k, j, i) = vxIn;
                prim(VLY, k, j, i) = 0;
                prim(VLZ, k, j, i) = 0;
                prim(PRE, k, j, i) = pre;
            }
    eos.PrimToCons(prim, cons, grid);
    boundary.EnrollSelfDefineBDXL();
    boundary.SetupImmersedShell(grid, domain, x0, y0, z0, rSolid);
}

// The inflow: every x-left ghost cell of the active rows holds (rho, rho vxIn, 0, 0, E).
void Boundary::SelfDefineBDXL(TArray<Real>& cons, const Grid& grid, const EquationOfState& eos) {
    Config& c = Config::getInstance();
    const Real rho = c.get("rho"), pre = c.get("p"), vxIn = c.get("vxIn");
    const Real eng = 0.5 * rho * (vxIn * vxIn) + eos.EGas(rho, pre);
    for (int k = grid.kb; k < grid.ke; k++)
        for (int j = grid.jb; j < grid.je; j++)
            for (int i = 0; i < grid.ib; i++) {
                cons(DEN, k, j, i) = rho;
                cons(MTX, k, j, i) = rho * vxIn;
                cons(MTY, k, j, i) = 0;
                cons(MTZ, k, j, i) = 0;
                cons(ENG, k, j, i) = eng;
            }
}

}  // namespace Gaukuk
