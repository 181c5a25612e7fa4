// gk_device.hpp -- device-side data shared by the kernels and the host runtime.
#pragma once

#include <stdint.h>

#include "gk_physics.hpp"

// Compile-time policy choice (configure --eos/--flux).  The reference links one EOS and one flux
// source file per build (Makefile:64-90); here they are template arguments.
#ifndef GK_EOS
#define GK_EOS Adiabatic
#endif
#ifndef GK_FLUX
#define GK_FLUX Hllc
#endif

namespace gk {

template <class R> using EosPolicy = GK_EOS<R>;
using RiemannPolicy = GK_FLUX;

// The running max of the wave speed is spread over this many words per parity (each wave of a
// stage adds its max to one, chosen by its index): all waves of a launch can finish at once, and
// device-scope atomics on a single word from thousands of waves serialize.  The step's dt takes the
// max of the words, which equals the max over all waves whatever their order.
constexpr int kCmaxSlots = 32;

// Device-resident time state, so that whole steps (dt included) run without a host round
// trip and can be captured in a HIP graph.  Mirrors the scalars of Sim::Advance
// (src/sim.cpp:69-117): t, dt, dtUntilOutput, tNext, step, stepMax, isContinue.
struct DevClock {
    double t, dt, t_next, dt_until_output;
    unsigned long long cmax_bits[2][kCmaxSlots];  // running max of |v|+c_s as IEEE bits, by parity
    int step, stepmax, is_continue, active;
    int parity;      // which cmax slot the current step reads
    int cur;         // forward Euler: 1 when the current state lives in the second buffer
    double cfl, drmin;
};

struct StepRecord {
    double t, dt;
    int step, pad;
};

// Arguments of one fused stage launch.  All arrays are SoA planes of the active region only:
// element (n, j, i) at n*plane + j*pitch + i.
template <class R>
struct StageArgs {
    const R* in;        // stage input U_s: fluxes are computed from it (halo via BC remap)
    const R* base;      // U^n for the three-coefficient combine (unused for two coefficients)
    R* out;             // stage output (may alias base, never in)
    long long plane;
    int pitch, nx, ny;
    int seg_rows;       // rows per wave segment
    int strips, nwaves; // wave grid: strips x segments
    int bc[6];          // xleft, xright, yleft, yright, zleft, zright (2D, ny == 1: y forced to type 0)
    R c1, c2, c3;       // combine coefficients
    int kick_in;        // stage input is U^n: apply the first gravity half-kick on load
    int kick_base;      // three-coefficient base is U^n: apply the first half-kick on load
    int pingpong;       // forward Euler: swap in/out when clock->cur is set
    int kick_out;       // last stage: apply the second half-kick to the result
    int want_cmax;      // last stage: accumulate next step's max wave speed
    R gx, gy, gz;
    double dxRec, dyRec;
    EosPolicy<R> eos;
    DevClock* clock;
    // point masses (SourceTerm::EnrollPointGravity): 0 none, 1 point gravity, 2 binary gravity
    int npoint;
    double pm0[5], pm1[5];       // gm, x, y, z, rs of each mass
    double ab, omega, theta1;    // the binary's separation, angular rate and initial phase
    // fixed (type 4) faces: the two saved ghost layers nearest each face, by distance d from it,
    // per z plane k: x faces [5][nz][ny][2] (element (n, k, j, d)), y faces [5][nz][2][nx]
    // (element (n, k, d, i)) (nz = 1 in 2D); 3D z faces the nearest saved plane [5][ny][nx]
    const R* fxl;
    const R* fxr;
    const R* fyl;
    const R* fyr;
    const R* fzl;
    const R* fzr;
    // inner fixed region (EnrollFixedBoundaryData with fixInner): the saved state (state layout)
    // and per active row (k * ny + j) the column interval [x, y) of region cells (empty when
    // x >= y)
    const R* inner;
    const int2* inner_rows;
    // 3D (nz > 1): planes of ny rows follow each other in every state plane of `plane` elements.
    // zstale: the nearest z ghost plane on each side (rows -1..ny) of the array the reference's
    // Sim holds as `cons` at step stale_base, [2][5][ny + 2][nx]; the reference never writes those
    // cells (edge cells, and whole planes of type-0 z faces), so they keep that array's initial
    // contents, and the other array's are zero.  stage_q: the stage's index in the step.
    int nz;
    long long plane2;            // elements per z plane (ny * pitch)
    double dzRec;
    const R* zstale;
    int stale_base, stage_q;
    const double* zcs;           // z centres, ghost-inclusive index
    const double* xc;            // cell centres, ghost-inclusive index (Domain, src/sim.cpp:37-39)
    const double* yc;
    double zc;                   // centre of the active z plane
    int ng;
};

// Immersed reflective shell (Boundary::ImmersedReflective, src/boundary/immersed.cpp:237-300).
// Cells are element offsets j * pitch + i of the state planes.
constexpr int kShellTargets = 8;
template <class R>
struct ShellArgs {
    R* u;                   // the state whose shell cells are rewritten
    R* alt;                 // forward Euler: the other buffer (the clock's `cur` selects)
    int pingpong;
    int always;             // 1: run outside an advance (clock inactive)
    long long plane;
    int nshell;
    const long long* cell;  // [nshell]
    const int* ntar;        // [nshell]
    const long long* tar;   // [kShellTargets][nshell]
    const R* dir;           // [4][nshell]: cos phi, sin phi, cos theta, sin theta
    const R* gam;           // [kShellTargets + 1][nshell]: distance weights, 0 = the shell cell
    R gm1, gm1Rec;
    const DevClock* clock;
};

constexpr int kWave = 64;
constexpr int kHalo = 2;
constexpr int kStrip = kWave - 2 * kHalo;  // owned columns per wave
constexpr int kBlock = 256;                // threads per workgroup (4 waves)

}  // namespace gk
