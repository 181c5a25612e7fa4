// ref_setup_from_file.cpp -- a problem setup in the reference's plug-in format
// (void Sim::Setup(), src/sim.hpp:38; one setup file per binary, Makefile:82).
//
// TEST INFRASTRUCTURE ONLY: compiled with the reference sources into oracle/_ref/ to
// produce golden fixtures.  It loads the active-region conserved state from the binary
// file named by $GK_IC_FILE, so every implementation starts from bit-identical initial
// conditions (the reference setups derive theirs from a truncated pi, src/gaukuk.hpp:19):
//   int32 nvar(=5), int32 nx, int32 ny, float64 cons[5][ny][nx]
// Optional config keys gx, gy, gz enroll constant gravity
// (SourceTerm::EnrollConstGravityVector, src/source_term/source.cpp:23-31); npoint (1 or 2) and
// pg<q>_gm, _x, _y, _z, _vx, _vy, _vz, _rs enroll point masses in order
// (SourceTerm::EnrollPointGravity, src/source_term/source.cpp:33-67; two masses = binary gravity);
// fixedbd = 1 fills every ghost cell with the conserved state fix_0..fix_4, fixedbd = 2 with the
// plane read from $GK_FIXED_FILE, and enrolls the fixed boundary data
// (Boundary::EnrollFixedBoundaryData, src/boundary/fixed.cpp:242-269).
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "gaukuk.hpp"
#include "sim.hpp"

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
    int hdr[3];
    if (std::fread(hdr, sizeof(int), 3, f) != 3) throw std::runtime_error("short IC header");
    if (hdr[0] != NVar || hdr[1] != grid.nx || hdr[2] != grid.ny)
        throw std::runtime_error("IC file does not match NX/NY");
    const size_t n = (size_t)NVar * grid.nx * grid.ny;
    std::vector<double> buf(n);
    if (std::fread(buf.data(), sizeof(double), n, f) != n) throw std::runtime_error("short IC data");
    std::fclose(f);
    const int k = grid.kb;
    const int fixedbd = (int)Config::getInstance().get("fixedbd", 0.0);
    if (fixedbd == 2) ReadFixedPlane(cons, grid);
    if (fixedbd == 1) {
        for (int v = 0; v < NVar; v++) {
            const Real fv = (Real)Config::getInstance().get("fix_" + std::to_string(v), 0.0);
            for (int kk = 0; kk < grid.lenz; kk++)
                for (int j = 0; j < grid.leny; j++)
                    for (int i = 0; i < grid.lenx; i++) cons(v, kk, j, i) = fv;
        }
    }
    for (int v = 0; v < NVar; v++)
        for (int j = 0; j < grid.ny; j++)
            for (int i = 0; i < grid.nx; i++)
                cons(v, k, grid.jb + j, grid.ib + i) = (Real)buf[((size_t)v * grid.ny + j) * grid.nx + i];
    Real gx = Config::getInstance().get("gx", 0.0);
    Real gy = Config::getInstance().get("gy", 0.0);
    Real gz = Config::getInstance().get("gz", 0.0);
    srcTerm.EnrollConstGravityVector(gx, gy, gz);
    if (fixedbd == 1 || fixedbd == 2) boundary.EnrollFixedBoundaryData(cons, grid);
    const int npoint = (int)Config::getInstance().get("npoint", 0.0);
    for (int q = 0; q < npoint; q++) {
        auto key = [&](const char* name) { return "pg" + std::to_string(q) + "_" + name; };
        auto get = [&](const char* name) { return (Real)Config::getInstance().get(key(name), 0.0); };
        srcTerm.EnrollPointGravity(get("gm"), get("x"), get("y"), get("z"), get("vx"), get("vy"), get("vz"),
                                   get("rs"));
    }
}

}  // namespace Gaukuk
