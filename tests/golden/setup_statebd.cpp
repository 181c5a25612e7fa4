// setup_statebd.cpp -- a test problem with self-defined faces whose ghost values depend on the
// state, written as a setup plug-in of the reference's format (a Sim::Setup plus the
// Boundary::SelfDefineBD* functions it enrolls, src/boundary/boundary.hpp:28-45) so that the same
// file builds into the reference (the fixture) and into this project's host driver.
//
//   x left   copy outflow: every ghost cell takes the edge cell of its row (the behaviour of the
//            reference's documented template, src/setup/setup_example.cpp:128-351)
//   x right  linear extrapolation from the two nearest active columns, density and energy kept
//            at least half the edge cell's
//   y left   a moving reflective lid: mirror image with the normal momentum negated and the
//            tangential velocity reflected about lidU, energy adjusted for the new kinetic energy
//   y right  copy outflow
//
// Initial state: uniform rho0, p0 with an overpressure disk (p0 * p_ratio, radius r0 at x0, y0)
// and a shear flow vx = vshear * sin(2 pi y).
#include <cmath>

#include "../gaukuk.hpp"
#include "../sim.hpp"

namespace Gaukuk
{

void Sim::Setup()
{
    const Real rho0 = Config::getInstance().get("rho0");
    const Real p0 = Config::getInstance().get("p0");
    const Real pRatio = Config::getInstance().get("p_ratio");
    const Real r0 = Config::getInstance().get("r0");
    const Real x0 = Config::getInstance().get("x0");
    const Real y0 = Config::getInstance().get("y0");
    const Real vShear = Config::getInstance().get("vshear");
    const Real twoPi = 2 * 3.14159265358979323846;
    for (int k = grid.kb; k < grid.ke; k++)
        for (int j = grid.jb; j < grid.je; j++)
            for (int i = grid.ib; i < grid.ie; i++) {
                const Real dx = domain.xc(i) - x0, dy = domain.yc(j) - y0;
                const bool inside = dx * dx + dy * dy < r0 * r0;
                prim(DEN, k, j, i) = rho0;
                prim(VLX, k, j, i) = vShear * std::sin(twoPi * domain.yc(j));
                prim(VLY, k, j, i) = 0;
                prim(VLZ, k, j, i) = 0;
                prim(PRE, k, j, i) = inside ? p0 * pRatio : p0;
            }
    eos.PrimToCons(prim, cons, grid);
    boundary.EnrollSelfDefineBDXL();
    boundary.EnrollSelfDefineBDXR();
    boundary.EnrollSelfDefineBDYL();
    boundary.EnrollSelfDefineBDYR();
}

void Boundary::SelfDefineBDXL(TArray<Real>& cons, const Grid& grid, const EquationOfState&)
{
    for (int v = 0; v < NVar; v++)
        for (int k = grid.kb; k < grid.ke; k++)
            for (int j = grid.jb; j < grid.je; j++)
                for (int i = grid.igb; i < grid.ib; i++) cons(v, k, j, i) = cons(v, k, j, grid.ib);
}

void Boundary::SelfDefineBDXR(TArray<Real>& cons, const Grid& grid, const EquationOfState&)
{
    const int e = grid.ie - 1;   // edge column
    for (int k = grid.kb; k < grid.ke; k++)
        for (int j = grid.jb; j < grid.je; j++)
            for (int i = grid.ie; i < grid.ige; i++) {
                const Real d = (Real)(i - e);
                for (int v = 0; v < NVar; v++) {
                    const Real edge = cons(v, k, j, e);
                    Real u = edge + d * (edge - cons(v, k, j, e - 1));
                    if ((v == DEN || v == ENG) && u < (Real)0.5 * edge) u = (Real)0.5 * edge;
                    cons(v, k, j, i) = u;
                }
            }
}

void Boundary::SelfDefineBDYL(TArray<Real>& cons, const Grid& grid, const EquationOfState&)
{
    const Real lidU = Config::getInstance().get("lidU");
    for (int k = grid.kb; k < grid.ke; k++)
        for (int j = grid.jgb; j < grid.jb; j++) {
            const int src = 2 * grid.jb - 1 - j;   // mirror row
            for (int i = grid.ib; i < grid.ie; i++) {
                const Real rho = cons(DEN, k, src, i);
                const Real mx = cons(MTX, k, src, i), my = cons(MTY, k, src, i), mz = cons(MTZ, k, src, i);
                const Real mxNew = 2 * rho * lidU - mx;
                const Real keOld = (Real)0.5 * (mx * mx + my * my + mz * mz) / rho;
                const Real keNew = (Real)0.5 * (mxNew * mxNew + my * my + mz * mz) / rho;
                cons(DEN, k, j, i) = rho;
                cons(MTX, k, j, i) = mxNew;
                cons(MTY, k, j, i) = -my;
                cons(MTZ, k, j, i) = mz;
                cons(ENG, k, j, i) = cons(ENG, k, src, i) - keOld + keNew;
            }
        }
}

void Boundary::SelfDefineBDYR(TArray<Real>& cons, const Grid& grid, const EquationOfState&)
{
    const int e = grid.je - 1;
    for (int v = 0; v < NVar; v++)
        for (int k = grid.kb; k < grid.ke; k++)
            for (int j = grid.je; j < grid.jge; j++)
                for (int i = grid.ib; i < grid.ie; i++) cons(v, k, j, i) = cons(v, k, e, i);
}

} // namespace Gaukuk
