// boundary.hpp -- boundary-condition dispatch.  As in the reference (src/boundary/boundary.hpp:
// 15-117, boundary.cpp:13-125) each face gets a function chosen by its type code from the
// configuration: 0 copy the edge cell, 1 periodic, 2 reflective, 3 outflow (mirror, inflow
// momentum clamped, energy adjusted), 4 fixed (ghost layers restored from data the setup
// enrolled); unknown codes fall back to 0.  On the device the same codes
// select an index remap applied while the stage kernel loads its halo, so no ghost cells are
// stored; the host functions here fill ghosts of host arrays (initial state and output files).
// The in-place parts of a boundary update (immersed shell, inner fixed region) run on the device:
// UpdateBD marks them pending and the Sim applies them to the uploaded state.
#pragma once

#include <functional>
#include <memory>

#include "../eos/eos.hpp"
#include "../gaukuk.hpp"
#include "../grid.hpp"
#include "../template_array.hpp"

namespace Gaukuk {

class Boundary {
public:
    using BoundaryFunc = std::function<void(TArray<Real>&, const Grid&, const EquationOfState&)>;
    Boundary();
    BoundaryFunc Bdxl, Bdxr, Bdyl, Bdyr, Bdzl, Bdzr;
    // type codes per face (xleft, xright, yleft, yright, zleft, zright) as passed to the device
    int type[6];
    bool selfDefined[6] = {false, false, false, false, false, false};

    void UpdateBD(TArray<Real>& cons, const Grid& grid, const Domain& domain, const EquationOfState& eos);

    // Self-defined faces (setup-file hooks, src/boundary/boundary.hpp:28-45): the setup file defines
    // the static member for a face and enrolls it by name, or passes any function.  On the device a
    // self-defined face whose ghost values do not depend on the state (an inflow; the Sim checks
    // that on a perturbed state) runs as a fixed face; any other 2D face runs through the boundary
    // hook (gk_set_bd_hook), which calls the function before every stage on a mirror of the
    // state's edge bands.
    static void SelfDefineBDXL(TArray<Real>& cons, const Grid& grid, const EquationOfState& eos);
    static void SelfDefineBDXR(TArray<Real>& cons, const Grid& grid, const EquationOfState& eos);
    static void SelfDefineBDYL(TArray<Real>& cons, const Grid& grid, const EquationOfState& eos);
    static void SelfDefineBDYR(TArray<Real>& cons, const Grid& grid, const EquationOfState& eos);
    static void SelfDefineBDZL(TArray<Real>& cons, const Grid& grid, const EquationOfState& eos);
    static void SelfDefineBDZR(TArray<Real>& cons, const Grid& grid, const EquationOfState& eos);
    void EnrollSelfDefineBDXL() { EnrollSelfDefineBDXL(&SelfDefineBDXL); }
    void EnrollSelfDefineBDXR() { EnrollSelfDefineBDXR(&SelfDefineBDXR); }
    void EnrollSelfDefineBDYL() { EnrollSelfDefineBDYL(&SelfDefineBDYL); }
    void EnrollSelfDefineBDYR() { EnrollSelfDefineBDYR(&SelfDefineBDYR); }
    void EnrollSelfDefineBDZL() { EnrollSelfDefineBDZL(&SelfDefineBDZL); }
    void EnrollSelfDefineBDZR() { EnrollSelfDefineBDZR(&SelfDefineBDZR); }
    void EnrollSelfDefineBDXL(BoundaryFunc f) { Bdxl = std::move(f); selfDefined[0] = true; }
    void EnrollSelfDefineBDXR(BoundaryFunc f) { Bdxr = std::move(f); selfDefined[1] = true; }
    void EnrollSelfDefineBDYL(BoundaryFunc f) { Bdyl = std::move(f); selfDefined[2] = true; }
    void EnrollSelfDefineBDYR(BoundaryFunc f) { Bdyr = std::move(f); selfDefined[3] = true; }
    void EnrollSelfDefineBDZL(BoundaryFunc f) { Bdzl = std::move(f); selfDefined[4] = true; }
    void EnrollSelfDefineBDZR(BoundaryFunc f) { Bdzr = std::move(f); selfDefined[5] = true; }
    bool AnySelfDefined() const;
    const BoundaryFunc& Face(int f) const;

    // Immersed reflective sphere (src/boundary/immersed.cpp:89-235): recorded here, set up and
    // applied on the device (gk_set_immersed_shell).
    void SetupImmersedShell(const Grid& grid, const Domain& domain, Real xShell, Real yShell, Real zShell,
                            Real rShell);
    bool HasShell() const { return hasShell_; }
    const double* Shell() const { return shell_; }
    // an UpdateBD whose in-place part (shell, inner region) the device has not applied yet
    bool DeviceUpdatePending() const { return pending_; }
    void ClearDeviceUpdate() { pending_ = false; }

    // Fixed faces (src/boundary/fixed.cpp:242-269): keeps a copy of `cons`, whose ghost layers
    // type-4 faces restore; with fixInner the inner region's active cells are restored before every
    // stage as well (src/boundary/boundary.cpp:139-141, fixed.cpp:181-207).
    void EnrollFixedBoundaryData(const TArray<Real>& cons, const Grid& grid, bool fixInner = false);
    bool AnyFixed() const;
    bool FixedEnrolled() const { return fixed_->enrolled; }
    bool FixedInner() const { return fixed_->inner; }
    const TArray<Real>& FixedData() const { return fixed_->data; }
    int FixedVersion() const { return fixed_->version; }

    struct FixedState {
        TArray<Real> data;
        bool enrolled = false;
        bool inner = false;
        int version = 0;
    };
    static BoundaryFunc ForType(int axis, bool high, int code, std::shared_ptr<const FixedState> fixed = nullptr);

private:
    std::shared_ptr<FixedState> fixed_ = std::make_shared<FixedState>();
    bool hasShell_ = false;
    double shell_[4] = {0, 0, 0, 0};
    bool pending_ = false;
};

}  // namespace Gaukuk
