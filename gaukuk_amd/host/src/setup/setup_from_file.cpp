// setup_from_file.cpp -- active-region conserved state from the binary file named by
// $GK_IC_FILE: int32 nvar (5), int32 nx, int32 ny, float64 cons[5][ny][nx]; config keys
// gx, gy, gz enroll constant gravity; npoint and pg<q>_* enroll point masses; fixedbd = 1 fills
// every ghost cell with the conserved state fix_0..fix_4, fixedbd = 2 with the plane read from
// $GK_FIXED_FILE, and enrolls it as fixed-face data.  Lets
// any array (e.g. from gaukuk_amd.problems) start a run.
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "../gaukuk.hpp"
#include "../sim.hpp"
#include "../utils/utils.hpp"

namespace Gaukuk {

// fixedbd = 2: the ghost-inclusive plane [5][NY+2g][NX+2g] from the file named by $GK_FIXED_FILE
// (int32 nvar, lenx, leny; float64 data), copied into every z plane of cons
static void ReadFixedPlane(TArray<Real>& cons, const Grid& grid) {
    const char* path = std::getenv("GK_FIXED_FILE");
    if (!path) throw std::runtime_error("fixedbd = 2 needs GK_FIXED_FILE");
    FILE* f = std::fopen(path, "rb");
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    int hdr[3] = {0, 0, 0};
    if (std::fread(hdr, sizeof(int), 3, f) != 3 || hdr[0] != NVar || hdr[1] != grid.lenx || hdr[2] != grid.leny) {
        std::fclose(f);
        throw std::runtime_error("fixed-data file header does not match the ghost-inclusive grid");
    }
    const size_t n = (size_t)NVar * grid.lenx * grid.leny;
    std::vector<double> buf(n);
    const size_t got = std::fread(buf.data(), sizeof(double), n, f);
    std::fclose(f);
    if (got != n) throw std::runtime_error("short fixed data");
    for (int v = 0; v < NVar; v++)
        for (int k = 0; k < grid.lenz; k++)
            for (int j = 0; j < grid.leny; j++)
                for (int i = 0; i < grid.lenx; i++)
                    cons(v, k, j, i) = (Real)buf[((size_t)v * grid.leny + j) * grid.lenx + i];
}

void Sim::Setup() {
    const char* path = std::getenv("GK_IC_FILE");
    if (!path) throw std::runtime_error("GK_IC_FILE is not set");
    FILE* f = std::fopen(path, "rb");
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    int hdr[3] = {0, 0, 0};
    const bool ok = std::fread(hdr, sizeof(int), 3, f) == 3;
    if (!ok || hdr[0] != NVar || hdr[1] != grid.nx || hdr[2] != grid.ny) {
        std::fclose(f);
        throw std::runtime_error("IC file header does not match NX/NY");
    }
    const size_t n = (size_t)NVar * grid.nx * grid.ny;
    std::vector<double> buf(n);
    const size_t got = std::fread(buf.data(), sizeof(double), n, f);
    std::fclose(f);
    if (got != n) throw std::runtime_error("short IC data");
    Config& c = Config::getInstance();
    const int fixedbd = (int)c.get("fixedbd", 0.0);
    if (fixedbd == 2) ReadFixedPlane(cons, grid);
    if (fixedbd == 1)
        for (int v = 0; v < NVar; v++) {
            const Real fv = (Real)c.get("fix_" + std::to_string(v), 0.0);
            for (int k = 0; k < grid.lenz; k++)
                for (int j = 0; j < grid.leny; j++)
                    for (int i = 0; i < grid.lenx; i++) cons(v, k, j, i) = fv;
        }
    for (int v = 0; v < NVar; v++)
        for (int j = 0; j < grid.ny; j++)
            for (int i = 0; i < grid.nx; i++)
                cons(v, grid.kb, grid.jb + j, grid.ib + i) = (Real)buf[((size_t)v * grid.ny + j) * grid.nx + i];
    if (fixedbd == 1 || fixedbd == 2) boundary.EnrollFixedBoundaryData(cons, grid);
    srcTerm.EnrollConstGravityVector(c.get("gx", 0.0), c.get("gy", 0.0), c.get("gz", 0.0));
    // npoint (1 or 2) and pg<q>_gm, _x, _y, _z, _vx, _vy, _vz, _rs enroll point masses in order
    const int npoint = (int)c.get("npoint", 0.0);
    for (int q = 0; q < npoint; q++) {
        auto get = [&](const char* name) { return (Real)c.get("pg" + std::to_string(q) + "_" + name, 0.0); };
        srcTerm.EnrollPointGravity(get("gm"), get("x"), get("y"), get("z"), get("vx"), get("vy"), get("vz"),
                                   get("rs"));
    }
}

}  // namespace Gaukuk
