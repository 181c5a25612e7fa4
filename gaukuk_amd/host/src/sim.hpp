// sim.hpp -- the simulation object of the drop-in host driver.
//
// Surface of the reference's Sim (src/sim.hpp:23-58): Setup() from a setup_XXX.cpp plug-in,
// Advance(dtoutput), WriteData(id, type), GetTime/Getdt/GetStep, isContinue, and the public
// grid, domain, cons, prim, eos, boundary and srcTerm members a setup file uses.  The integrator
// is chosen through a pointer-to-member from the `integrator` key (src/sim.cpp:60-66).  The time
// step runs on the GPU through the C ABI of include/gaukuk.h; host arrays are only synchronised
// with the device at output time.
#pragma once

#include <string>

#include "boundary/boundary.hpp"
#include "eos/eos.hpp"
#include "gaukuk.h"
#include "gaukuk.hpp"
#include "grid.hpp"
#include "source_term/source.hpp"
#include "template_array.hpp"

namespace Gaukuk {

enum class DataType { Prim, Cons };

class Sim {
public:
    Sim();
    ~Sim();
    Sim(const Sim&) = delete;
    Sim& operator=(const Sim&) = delete;

    bool isContinue;
    const Grid grid;
    const Domain domain;
    TArray<Real> cons, prim;
    EquationOfState eos;
    Boundary boundary;
    SourceTerm srcTerm;

    void Setup();  // provided by the linked setup_XXX.cpp
    void Advance(Real dtoutput);
    void WriteData(const int outputID, DataType dType);
    // Continue from a cons_XXXXX file of this grid (an extension: the reference always starts at
    // t = 0): loads its state, ghost cells included, and its time; returns XXXXX.
    int Restart(const std::string& path);

    Real GetTime() { return t; }
    Real Getdt() { return dt; }
    int GetStep() { return step; }
    void SetVerbose(bool v) { verbose_ = v; }
    // Host state changed (e.g. by a setup or a restart): upload before the next Advance.
    void MarkHostModified() { hostDirty_ = true; }

private:
    using IntegratorFunc = int (Sim::*)() const;  // the integrator the device runs
    int ForwardEuler_() const { return 1; }
    int RK2_() const { return 2; }
    int RK3_() const { return 3; }
    IntegratorFunc hydroIntegrator_;

    int step, stepMax;
    Real t, dt, CFL;
    int rcOrder, integratorType;

    gk_ctx* ctx_ = nullptr;
    bool hostDirty_ = true;   // host cons newer than the device copy
    bool hostStale_ = false;  // device newer than host cons
    int fixedSent_ = -1;      // version of the fixed-face data the device holds (-1: none)
    bool timePending_ = false;  // a restart time the device has not been given yet
    bool shellSent_ = false;  // the immersed shell is set up on the device
    bool selfFaces_[6] = {false, false, false, false, false, false};  // self-defined faces run as fixed
    TArray<Real> selfGhosts_;  // the ghost layers those faces write
    int hookFaces_ = 0;        // state-dependent self-defined faces, refreshed through the hook
    TArray<Real> hookMirror_;  // the hook's mirror array (NaN outside the mirrored bands)
    bool ProbeSelfDefined(int f, TArray<Real>& out) const;
    static void BdHook(void* user, void* mirror);
    bool verbose_ = true;
    void EnsureDevice();
    void Check(gk_status s, const char* what) const;
};

}  // namespace Gaukuk
