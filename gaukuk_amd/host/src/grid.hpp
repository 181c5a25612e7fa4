// grid.hpp -- mesh sizes and coordinates from the configuration (the reference's Grid and
// Domain, src/grid.hpp:11-56, src/sim.cpp:16-42): nx, ny, nz, nGhost; len* = n* + 2*nGhost;
// active index ranges [ib, ie) etc.; dx = (xmax - xmin)/nx; dyRec = 0 when ny <= 1 (and dzRec
// when nz <= 1); drmin = min(dx, dy, dz); cell centres xc(i) = xmin - nGhost*dx + 0.5*dx + i*dx.
#pragma once

#include <algorithm>
#include <stdexcept>

#include "gaukuk.hpp"
#include "template_array.hpp"
#include "utils/utils.hpp"

namespace Gaukuk {

class Grid {
public:
    int nx, ny, nz, nGhost, lenx, leny, lenz, lenArr;
    int igb, ige, jgb, jge, kgb, kge;
    int ib, ie, jb, je, kb, ke;

    Grid() {
        Config& c = Config::getInstance();
        nx = (int)c.get("NX");
        ny = (int)c.get("NY");
        nz = (int)c.get("NZ");
        nGhost = (int)c.get("NGhost");
        if (nx <= 0 || ny <= 0 || nz <= 0 || nGhost < 0) throw std::runtime_error("Invalid grid parameters from config");
        lenx = nx + 2 * nGhost; leny = ny + 2 * nGhost; lenz = nz + 2 * nGhost;
        lenArr = lenx * leny * lenz;
        igb = jgb = kgb = 0;
        ige = lenx; jge = leny; kge = lenz;
        ib = nGhost; ie = ib + nx;
        jb = nGhost; je = jb + ny;
        kb = nGhost; ke = kb + nz;
    }
};

class Domain {
public:
    explicit Domain(const Grid& g) {
        Config& c = Config::getInstance();
        xmin = c.get("xmin"); xmax = c.get("xmax");
        ymin = c.get("ymin"); ymax = c.get("ymax");
        zmin = c.get("zmin"); zmax = c.get("zmax");
        dx = (xmax - xmin) / g.nx;
        dy = (ymax - ymin) / g.ny;
        dz = (zmax - zmin) / g.nz;
        dxRec = 1. / dx;
        dyRec = g.ny <= 1 ? 0 : 1. / dy;
        dzRec = g.nz <= 1 ? 0 : 1. / dz;
        drmin = std::min(std::min(dx, dy), dz);
        xc.NewArray(g.lenx); yc.NewArray(g.leny); zc.NewArray(g.lenz);
        for (int i = 0; i < g.lenx; i++) xc(i) = xmin - g.nGhost * dx + 0.5 * dx + i * dx;
        for (int j = 0; j < g.leny; j++) yc(j) = ymin - g.nGhost * dy + 0.5 * dy + j * dy;
        for (int k = 0; k < g.lenz; k++) zc(k) = zmin - g.nGhost * dz + 0.5 * dz + k * dz;
    }
    Real xmin, xmax, ymin, ymax, zmin, zmax;
    Real dx, dy, dz, drmin;
    Real dxRec, dyRec, dzRec;
    TArray<Real> xc, yc, zc;
};

}  // namespace Gaukuk
