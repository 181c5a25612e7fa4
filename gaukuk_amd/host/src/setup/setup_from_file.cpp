// setup_from_file.cpp -- active-region conserved state from the binary file named by
// $GK_IC_FILE: int32 nvar (5), int32 nx, int32 ny, float64 cons[5][nz][ny][nx]; config keys
// gx, gy, gz enroll constant gravity; npoint and pg<q>_* enroll point masses; fixedbd = 1 fills
// every ghost cell with the conserved state fix_0..fix_4, fixedbd = 2 with the plane read from
// $GK_FIXED_FILE, fixedbd = 3 with the whole ghost-inclusive 3D array read from it, and enrolls it
// as fixed-face data (fixinner = 1: with the inner region);
// selfbd_<f> = 1 enrolls a self-defined face f (0 x left .. 3 y right) that writes the conserved
// state sd<f>_0..sd<f>_4 into its ghost cells (with sdcopy_<f> = 1 it copies the edge cell
// instead, which depends on the state); immersed = 1 sets up the immersed shell
// (shell_x, shell_y, shell_z, shell_r).  Lets
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

// fixedbd = 3: the ghost-inclusive array [5][NZ+2g][NY+2g][NX+2g] from $GK_FIXED_FILE (int32 nvar,
// lenx, leny, lenz; float64 data)
static void ReadFixedVolume(TArray<Real>& cons, const Grid& grid) {
    const char* path = std::getenv("GK_FIXED_FILE");
    if (!path) throw std::runtime_error("fixedbd = 3 needs GK_FIXED_FILE");
    FILE* f = std::fopen(path, "rb");
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    int hdr[4] = {0, 0, 0, 0};
    const bool ok = std::fread(hdr, sizeof(int), 4, f) == 4 && hdr[0] == NVar && hdr[1] == grid.lenx &&
                    hdr[2] == grid.leny && hdr[3] == grid.lenz;
    std::vector<double> buf;
    if (ok) {
        buf.resize(cons.GetSize());
        if (std::fread(buf.data(), sizeof(double), buf.size(), f) != buf.size()) buf.clear();
    }
    std::fclose(f);
    if (!ok) throw std::runtime_error("fixed-data file header does not match the ghost-inclusive grid");
    if (buf.empty()) throw std::runtime_error("short fixed data");
    for (size_t q = 0; q < buf.size(); q++) cons.data()[q] = (Real)buf[q];
}

// a self-defined face writing one conserved state into its ghost cells (active tangential range)
static void ConstantFace(TArray<Real>& cons, const Grid& grid, int f) {
    Real v[NVar];
    for (int q = 0; q < NVar; q++)
        v[q] = (Real)Config::getInstance().get("sd" + std::to_string(f) + "_" + std::to_string(q), 0.0);
    const int il = f == 0 ? 0 : (f == 1 ? grid.ie : grid.ib);
    const int ir = f == 0 ? grid.ib : (f == 1 ? grid.ie + grid.nGhost : grid.ie);
    const int jl = f == 2 ? 0 : (f == 3 ? grid.je : grid.jb);
    const int jr = f == 2 ? grid.jb : (f == 3 ? grid.je + grid.nGhost : grid.je);
    const bool copy = (int)Config::getInstance().get("sdcopy_" + std::to_string(f), 0.0) == 1;
    for (int q = 0; q < NVar; q++)
        for (int k = grid.kb; k < grid.ke; k++)
            for (int j = jl; j < jr; j++)
                for (int i = il; i < ir; i++) {
                    const int si = f == 0 ? grid.ib : (f == 1 ? grid.ie - 1 : i);
                    const int sj = f == 2 ? grid.jb : (f == 3 ? grid.je - 1 : j);
                    cons(q, k, j, i) = copy ? cons(q, k, sj, si) : v[q];
                }
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
    int hdr[3] = {0, 0, 0};
    const bool ok = std::fread(hdr, sizeof(int), 3, f) == 3;
    if (!ok || hdr[0] != NVar || hdr[1] != grid.nx || hdr[2] != grid.ny) {
        std::fclose(f);
        throw std::runtime_error("IC file header does not match NX/NY");
    }
    const size_t n = (size_t)NVar * grid.nx * grid.ny * grid.nz;
    std::vector<double> buf(n);
    const size_t got = std::fread(buf.data(), sizeof(double), n, f);
    std::fclose(f);
    if (got != n) throw std::runtime_error("short IC data");
    Config& c = Config::getInstance();
    const int fixedbd = (int)c.get("fixedbd", 0.0);
    if (fixedbd == 2) ReadFixedPlane(cons, grid);
    if (fixedbd == 3) ReadFixedVolume(cons, grid);
    if (fixedbd == 1)
        for (int v = 0; v < NVar; v++) {
            const Real fv = (Real)c.get("fix_" + std::to_string(v), 0.0);
            for (int k = 0; k < grid.lenz; k++)
                for (int j = 0; j < grid.leny; j++)
                    for (int i = 0; i < grid.lenx; i++) cons(v, k, j, i) = fv;
        }
    for (int v = 0; v < NVar; v++)
        for (int k = 0; k < grid.nz; k++)
            for (int j = 0; j < grid.ny; j++)
                for (int i = 0; i < grid.nx; i++)
                    cons(v, grid.kb + k, grid.jb + j, grid.ib + i) =
                        (Real)buf[(((size_t)v * grid.nz + k) * grid.ny + j) * grid.nx + i];
    const bool fixinner = (int)c.get("fixinner", 0.0) == 1;
    if ((fixedbd >= 1 && fixedbd <= 3) || fixinner) boundary.EnrollFixedBoundaryData(cons, grid, fixinner);
    if ((int)c.get("selfbd_0", 0.0) == 1) boundary.EnrollSelfDefineBDXL();
    if ((int)c.get("selfbd_1", 0.0) == 1) boundary.EnrollSelfDefineBDXR();
    if ((int)c.get("selfbd_2", 0.0) == 1) boundary.EnrollSelfDefineBDYL();
    if ((int)c.get("selfbd_3", 0.0) == 1) boundary.EnrollSelfDefineBDYR();
    if ((int)c.get("immersed", 0.0) == 1)
        boundary.SetupImmersedShell(grid, domain, (Real)c.get("shell_x", 0.0), (Real)c.get("shell_y", 0.0),
                                    (Real)c.get("shell_z", 0.0), (Real)c.get("shell_r", 0.0));
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
