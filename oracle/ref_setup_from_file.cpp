// ref_setup_from_file.cpp -- a problem setup in the reference's plug-in format
// (void Sim::Setup(), src/sim.hpp:38; one setup file per binary, Makefile:82).
//
// TEST INFRASTRUCTURE ONLY: compiled with the reference sources into oracle/_ref/ to
// produce golden fixtures.  It loads the active-region conserved state from the binary
// file named by $GK_IC_FILE, so every implementation starts from bit-identical initial
// conditions (the reference setups derive theirs from a truncated pi, src/gaukuk.hpp:19):
//   int32 nvar(=5), int32 nx, int32 ny, float64 cons[5][nz][ny][nx] (nz = NZ, 1 in 2D)
// Optional config keys gx, gy, gz enroll constant gravity
// (SourceTerm::EnrollConstGravityVector, src/source_term/source.cpp:23-31); npoint (1 or 2) and
// pg<q>_gm, _x, _y, _z, _vx, _vy, _vz, _rs enroll point masses in order
// (SourceTerm::EnrollPointGravity, src/source_term/source.cpp:33-67; two masses = binary gravity);
// fixedbd = 1 fills every ghost cell with the conserved state fix_0..fix_4, fixedbd = 2 with the
// plane read from $GK_FIXED_FILE, fixedbd = 3 with the whole ghost-inclusive 3D array read from
// $GK_FIXED_FILE, and enrolls the fixed boundary data
// (Boundary::EnrollFixedBoundaryData, src/boundary/fixed.cpp:242-269); fixinner = 1 enrolls it with
// the inner region (fixed.cpp:181-207); selfbd_<f> = 1 enrolls a self-defined face f (0 xleft,
// 1 xright, 2 yleft, 3 yright) whose ghost cells take the conserved state sd<f>_0..sd<f>_4 (the
// form of the inflow face of src/setup/setup_solid_body.cpp); immersed = 1 sets up the immersed
// shell at shell_x, shell_y, shell_z with radius shell_r (Boundary::SetupImmersedShell,
// src/boundary/immersed.cpp:89-235).
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

// fixedbd = 3: the ghost-inclusive array [5][NZ+2g][NY+2g][NX+2g] from $GK_FIXED_FILE
// (int32 nvar, lenx, leny, lenz; float64 data)
static void ReadFixedVolume(TArray<Real>& cons, const Grid& grid) {
    const char* path = std::getenv("GK_FIXED_FILE");
    if (!path) throw std::runtime_error("fixedbd = 3 needs GK_FIXED_FILE");
    FILE* f = std::fopen(path, "rb");
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    int hdr[4] = {0, 0, 0, 0};
    if (std::fread(hdr, sizeof(int), 4, f) != 4 || hdr[0] != NVar || hdr[1] != grid.lenx || hdr[2] != grid.leny ||
        hdr[3] != grid.lenz) {
        std::fclose(f);
        throw std::runtime_error("fixed-data file header does not match the ghost-inclusive grid");
    }
    const size_t n = (size_t)NVar * grid.lenx * grid.leny * grid.lenz;
    std::vector<double> buf(n);
    const size_t got = std::fread(buf.data(), sizeof(double), n, f);
    std::fclose(f);
    if (got != n) throw std::runtime_error("short fixed data");
    for (int v = 0; v < NVar; v++)
        for (int k = 0; k < grid.lenz; k++)
            for (int j = 0; j < grid.leny; j++)
                for (int i = 0; i < grid.lenx; i++)
                    cons(v, k, j, i) = (Real)buf[(((size_t)v * grid.lenz + k) * grid.leny + j) * grid.lenx + i];
}

// a self-defined face writing one conserved state into its ghost cells (active tangential range)
static void ConstantFace(TArray<Real>& cons, const Grid& grid, int f) {
    Real v[NVar];
    for (int q = 0; q < NVar; q++)
        v[q] = (Real)Config::getInstance().get("sd" + std::to_string(f) + "_" + std::to_string(q), 0.0);
    const int il = f == 0 ? grid.igb : (f == 1 ? grid.ie : grid.ib);
    const int ir = f == 0 ? grid.ib : (f == 1 ? grid.ige : grid.ie);
    const int jl = f == 2 ? grid.jgb : (f == 3 ? grid.je : grid.jb);
    const int jr = f == 2 ? grid.jb : (f == 3 ? grid.jge : grid.je);
    for (int q = 0; q < NVar; q++)
        for (int k = grid.kb; k < grid.ke; k++)
            for (int j = jl; j < jr; j++)
                for (int i = il; i < ir; i++) cons(q, k, j, i) = v[q];
}

void Boundary::SelfDefineBDXL(TArray<Real>& cons, const Grid& grid, const EquationOfState&) { ConstantFace(cons, grid, 0); }
void Boundary::SelfDefineBDXR(TArray<Real>& cons, const Grid& grid, const EquationOfState&) { ConstantFace(cons, grid, 1); }
void Boundary::SelfDefineBDYL(TArray<Real>& cons, const Grid& grid, const EquationOfState&) { ConstantFace(cons, grid, 2); }
void Boundary::SelfDefineBDYR(TArray<Real>& cons, const Grid& grid, const EquationOfState&) { ConstantFace(cons, grid, 3); }

void Sim::Setup() {
    const char* path = std::getenv("GK_IC_FILE");
    if (!path) throw std::runtime_error("GK_IC_FILE is not set");
    FILE* f = std::fopen(path, "rb");
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    int hdr[3];
    if (std::fread(hdr, sizeof(int), 3, f) != 3) throw std::runtime_error("short IC header");
    if (hdr[0] != NVar || hdr[1] != grid.nx || hdr[2] != grid.ny)
        throw std::runtime_error("IC file does not match NX/NY");
    const size_t n = (size_t)NVar * grid.nx * grid.ny * grid.nz;
    std::vector<double> buf(n);
    if (std::fread(buf.data(), sizeof(double), n, f) != n) throw std::runtime_error("short IC data");
    std::fclose(f);
    const int fixedbd = (int)Config::getInstance().get("fixedbd", 0.0);
    if (fixedbd == 2) ReadFixedPlane(cons, grid);
    if (fixedbd == 3) ReadFixedVolume(cons, grid);
    if (fixedbd == 1) {
        for (int v = 0; v < NVar; v++) {
            const Real fv = (Real)Config::getInstance().get("fix_" + std::to_string(v), 0.0);
            for (int kk = 0; kk < grid.lenz; kk++)
                for (int j = 0; j < grid.leny; j++)
                    for (int i = 0; i < grid.lenx; i++) cons(v, kk, j, i) = fv;
        }
    }
    for (int v = 0; v < NVar; v++)
        for (int k = 0; k < grid.nz; k++)
            for (int j = 0; j < grid.ny; j++)
                for (int i = 0; i < grid.nx; i++)
                    cons(v, grid.kb + k, grid.jb + j, grid.ib + i) =
                        (Real)buf[(((size_t)v * grid.nz + k) * grid.ny + j) * grid.nx + i];
    Real gx = Config::getInstance().get("gx", 0.0);
    Real gy = Config::getInstance().get("gy", 0.0);
    Real gz = Config::getInstance().get("gz", 0.0);
    srcTerm.EnrollConstGravityVector(gx, gy, gz);
    Config& c = Config::getInstance();
    const bool fixinner = (int)c.get("fixinner", 0.0) == 1;
    if ((fixedbd >= 1 && fixedbd <= 3) || fixinner) boundary.EnrollFixedBoundaryData(cons, grid, fixinner);
    if ((int)c.get("selfbd_0", 0.0) == 1) boundary.EnrollSelfDefineBDXL();
    if ((int)c.get("selfbd_1", 0.0) == 1) boundary.EnrollSelfDefineBDXR();
    if ((int)c.get("selfbd_2", 0.0) == 1) boundary.EnrollSelfDefineBDYL();
    if ((int)c.get("selfbd_3", 0.0) == 1) boundary.EnrollSelfDefineBDYR();
    if ((int)c.get("immersed", 0.0) == 1)
        boundary.SetupImmersedShell(grid, domain, (Real)c.get("shell_x", 0.0), (Real)c.get("shell_y", 0.0),
                                    (Real)c.get("shell_z", 0.0), (Real)c.get("shell_r", 0.0));
    const int npoint = (int)Config::getInstance().get("npoint", 0.0);
    for (int q = 0; q < npoint; q++) {
        auto key = [&](const char* name) { return "pg" + std::to_string(q) + "_" + name; };
        auto get = [&](const char* name) { return (Real)Config::getInstance().get(key(name), 0.0); };
        srcTerm.EnrollPointGravity(get("gm"), get("x"), get("y"), get("z"), get("vx"), get("vy"), get("vz"),
                                   get("rs"));
    }
}

}  // namespace Gaukuk
