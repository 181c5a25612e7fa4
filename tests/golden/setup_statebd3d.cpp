// setup_statebd3d.cpp -- a test problem with self-defined faces whose ghost values depend on the
// state on all six sides, optionally under a point mass, written as a setup plug-in of the
// reference's format (a Sim::Setup plus the Boundary::SelfDefineBD* functions it enrolls,
// src/boundary/boundary.hpp:28-45) so that the same file builds into the reference (the fixture)
// and into this project's host driver.  With NZ = 1 the z faces are enrolled but never called
// (UpdateBD calls them only in 3D, src/boundary/boundary.cpp:127-135).
//
//   x left   copy outflow: every ghost cell takes the edge cell of its row (the behaviour of the
//            reference's documented template, src/setup/setup_example.cpp:128-351)
//   x right  linear extrapolation from the two nearest active columns, density and energy kept
//            at least half the edge cell's
//   y left   a moving reflective lid: mirror image with the normal momentum negated and the
//            tangential x velocity reflected about lidU, energy adjusted for the new kinetic energy
//   y right  copy outflow
//   z left   copy outflow from the first active plane (the template's SelfDefineBDZL)
//   z right  mirror image with the normal momentum negated and the tangential momenta halved,
//            energy adjusted for the kinetic energy removed
//
// Initial state: uniform rho0, p0 with an overpressure sphere (p0 * p_ratio, radius r0 at x0, y0,
// z0), a shear flow vx = vshear * sin(2 pi y) and vz = vz0 * sin(2 pi x); a point mass gm at
// (xg, yg, zg) with softening rs when gm is not zero.
#include <cmath>

#include "../gaukuk.hpp"
#include "../sim.hpp"

namespace Gaukuk
{

void Sim::Setup()
{
    Config& cfg = Config::getInstance();
    const Real rho0 = cfg.get("rho0"), p0 = cfg.get("p0"), pRatio = cfg.get("p_ratio");
    const Real r0 = cfg.get("r0"), x0 = cfg.get("x0"), y0 = cfg.get("y0"), z0 = cfg.get("z0");
    const Real vShear = cfg.get("vshear"), vz0 = cfg.get("vz0");
    const Real gm = cfg.get("gm");
    const Real twoPi = 2 * 3.14159265358979323846;
    for (int k = grid.kb; k < grid.ke; k++)
        for (int j = grid.jb; j < grid.je; j++)
            for (int i = grid.ib; i < grid.ie; i++) {
                const Real dx = domain.xc(i) - x0, dy = domain.yc(j) - y0, dz = domain.zc(k) - z0;
                const bool inside = dx * dx + dy * dy + dz * dz < r0 * r0;
                prim(DEN, k, j, i) = rho0;
                prim(VLX, k, j, i) = vShear * std::sin(twoPi * domain.yc(j));
                prim(VLY, k, j, i) = 0;
                prim(VLZ, k, j, i) = vz0 * std::sin(twoPi * domain.xc(i));
                prim(PRE, k, j, i) = inside ? p0 * pRatio : p0;
            }
    eos.PrimToCons(prim, cons, grid);
    if (gm != 0) srcTerm.EnrollPointGravity(gm, cfg.get("xg"), cfg.get("yg"), cfg.get("zg"), 0, 0, 0, cfg.get("rs"));
    boundary.EnrollSelfDefineBDXL();
    boundary.EnrollSelfDefineBDXR();
    boundary.EnrollSelfDefineBDYL();
    boundary.EnrollSelfDefineBDYR();
    boundary.EnrollSelfDefineBDZL();
    boundary.EnrollSelfDefineBDZR();
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

void Boundary::SelfDefineBDZL(TArray<Real>& cons, const Grid& grid, const EquationOfState&)
{
    for (int v = 0; v < NVar; v++)
        for (int k = grid.kgb; k < grid.kb; k++)
            for (int j = grid.jb; j < grid.je; j++)
                for (int i = grid.ib; i < grid.ie; i++) cons(v, k, j, i) = cons(v, grid.kb, j, i);
}

void Boundary::SelfDefineBDZR(TArray<Real>& cons, const Grid& grid, const EquationOfState&)
{
    for (int k = grid.ke; k < grid.kge; k++) {
        const int src = 2 * grid.ke - 1 - k;   // mirror plane
        for (int j = grid.jb; j < grid.je; j++)
            for (int i = grid.ib; i < grid.ie; i++) {
                const Real rho = cons(DEN, src, j, i);
                const Real mx = cons(MTX, src, j, i), my = cons(MTY, src, j, i), mz = cons(MTZ, src, j, i);
                const Real keRemoved = (Real)0.375 * (mx * mx + my * my) / rho;   // 1/2 (1 - 1/4) |m_t|^2 / rho
                cons(DEN, k, j, i) = rho;
                cons(MTX, k, j, i) = (Real)0.5 * mx;
                cons(MTY, k, j, i) = (Real)0.5 * my;
                cons(MTZ, k, j, i) = -mz;
                cons(ENG, k, j, i) = cons(ENG, src, j, i) - keRemoved;
            }
    }
}

} // namespace Gaukuk
