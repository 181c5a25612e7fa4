// eos.hpp -- host side of the equation of state: the conversions a problem setup and the
// output writer need (the reference's EquationOfState, src/eos/eos.hpp:13-36).  The time step
// itself runs on the device with the same policy compiled into the HIP library.
#pragma once

#include <cmath>

#include "../gaukuk.hpp"
#include "../grid.hpp"
#include "../template_array.hpp"

namespace Gaukuk {

class EquationOfState {
public:
    EquationOfState();
    // conversions over k-active (2D) or all k, and all j, i including ghosts
    void ConsToPrim(const TArray<Real>& cons, TArray<Real>& prim, const Grid& grid) const;
    void PrimToCons(const TArray<Real>& prim, TArray<Real>& cons, const Grid& grid) const;
    Real GetGamma() const { return gamma_; }
    Real GetGm1Rec() const { return gm1Rec_; }
    Real SoundSpeed(Real den, Real pre) const { return std::sqrt(gamma_ * pre / den); }
    Real EGas(Real, Real pre) const { return pre * gm1Rec_; }
    static const char* Name();

private:
    Real gamma_, gm1Rec_;
};

}  // namespace Gaukuk
