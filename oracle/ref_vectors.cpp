// ref_vectors.cpp -- known-answer vector harness linked against the reference's own objects.
//
// TEST INFRASTRUCTURE ONLY: built into oracle/_ref/ from the reference sources, run here to
// write per-function golden vectors (tests/golden/).  It calls, unchanged:
//   Flux::RiemannSolver            src/flux/hllc.cpp:22-103
//   Reconstruction::ReconstructXPLM src/reconstruction/recon_piecewise_linear.cpp:8-32
//   EquationOfState::ConsToPrim    src/eos/adiabatic.cpp:18-61
//   EquationOfState::CalCmax       src/eos/adiabatic.cpp:115-160
//   EquationOfState::PrimToCons    src/eos/adiabatic.cpp:162-198
// Usage: ref_vectors <config.in> <op> <in.bin> <out.bin>
//   Every input starts with two int32 (the second is 0 unless stated).
//   op hllc:  in = int32 n, int32 dir, f64 ul[5][n], f64 ur[5][n]      out = f64 flux[5][n]
//   op plm:   in = int32 m, f64 prim[5][m]                             out = f64 ul[5][m+1], ur[5][m+1]
//   op c2p:   in = int32 n, f64 cons[5][n]                             out = f64 prim[5][n]
//   op p2c:   in = int32 n, f64 prim[5][n]                             out = f64 cons[5][n]
//   op cmax:  in = int32 n, f64 cons[5][n]                             out = f64 cmax
// The config file supplies gamma and a grid NX/NY/NZ/NGhost whose row holds n cells.
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "eos/eos.hpp"
#include "flux/flux.hpp"
#include "grid.hpp"
#include "reconstruction/reconstruction.hpp"
#include "template_array.hpp"
#include "utils/utils.hpp"

using namespace Gaukuk;

static std::vector<char> slurp(const char* p) {
    FILE* f = std::fopen(p, "rb");
    if (!f) throw std::runtime_error("cannot open input");
    std::vector<char> b;
    char tmp[65536];
    size_t r;
    while ((r = std::fread(tmp, 1, sizeof tmp, f)) > 0) b.insert(b.end(), tmp, tmp + r);
    std::fclose(f);
    return b;
}

static void spit(const char* p, const void* d, size_t n) {
    FILE* f = std::fopen(p, "wb");
    if (!f || std::fwrite(d, 1, n, f) != n) throw std::runtime_error("cannot write output");
    std::fclose(f);
}

int main(int argc, char** argv) {
    if (argc != 5) { std::fprintf(stderr, "usage: %s cfg op in out\n", argv[0]); return 2; }
    Config::getInstance().loadFromFile(argv[1]);
    std::string op = argv[2];
    std::vector<char> in = slurp(argv[3]);
    const int* hi = reinterpret_cast<const int*>(in.data());
    EquationOfState eos;
    if (op == "hllc") {
        int n = hi[0], dir = hi[1];
        const double* ul = reinterpret_cast<const double*>(in.data() + 8);
        const double* ur = ul + 5 * n;
        TArray<Real> L(NVar, n), R(NVar, n), F(NVar, 1, 1, n);
        for (int v = 0; v < NVar; v++)
            for (int i = 0; i < n; i++) { L(v, i) = ul[v * n + i]; R(v, i) = ur[v * n + i]; }
        Flux flux(n);
        flux.RiemannSolver(L, R, dir, eos, F, 0, 0, 0, n);
        spit(argv[4], F.data(), F.GetSizeInBytes());
    } else if (op == "plm") {
        int m = hi[0];
        const double* w = reinterpret_cast<const double*>(in.data() + 8);
        TArray<Real> W(NVar, 1, 1, m), L(NVar, m + 1), R(NVar, m + 1);
        for (int v = 0; v < NVar; v++)
            for (int i = 0; i < m; i++) W(v, 0, 0, i) = w[v * m + i];
        Reconstruction rc;
        rc.ReconstructXPLM(W, L, R, 0, 0, 0, m);
        std::vector<double> out(L.data(), L.data() + L.GetSize());
        out.insert(out.end(), R.data(), R.data() + R.GetSize());
        spit(argv[4], out.data(), out.size() * sizeof(double));
    } else if (op == "c2p" || op == "p2c" || op == "cmax") {
        int n = hi[0];
        const double* u = reinterpret_cast<const double*>(in.data() + 8);
        Grid grid;  // NX = n, NY = NZ = 1 from the config
        if (grid.nx != n) throw std::runtime_error("config NX must equal n");
        TArray<Real> A(NVar, grid.lenz, grid.leny, grid.lenx), B(NVar, grid.lenz, grid.leny, grid.lenx);
        for (int v = 0; v < NVar; v++)
            for (int i = 0; i < n; i++) A(v, grid.kb, grid.jb, grid.ib + i) = u[v * n + i];
        if (op == "cmax") {
            Real c = 1e-16;
            eos.CalCmax(A, grid, c);
            spit(argv[4], &c, sizeof c);
            return 0;
        }
        if (op == "c2p") eos.ConsToPrim(A, B, grid); else eos.PrimToCons(A, B, grid);
        std::vector<double> out((size_t)NVar * n);
        for (int v = 0; v < NVar; v++)
            for (int i = 0; i < n; i++) out[v * n + i] = B(v, grid.kb, grid.jb, grid.ib + i);
        spit(argv[4], out.data(), out.size() * sizeof(double));
    } else {
        std::fprintf(stderr, "unknown op %s\n", op.c_str());
        return 2;
    }
    return 0;
}
