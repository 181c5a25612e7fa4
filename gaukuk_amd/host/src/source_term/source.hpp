// source.hpp -- source-term enrollment (the reference's SourceTerm, src/source_term/source.hpp:
// 23-50).  Constant gravity is applied on the device as Strang half-kicks fused into the first
// and last RK stage; point and binary gravity are not part of this build and are rejected.
#pragma once

#include <stdexcept>

#include "../gaukuk.hpp"

namespace Gaukuk {

class SourceTerm {
public:
    bool sourceEnrolled = false;
    Real gx_ = 0, gy_ = 0, gz_ = 0;
    // src/source_term/source.cpp:23-31: enrolled only for a non-zero vector
    void EnrollConstGravityVector(Real gx, Real gy, Real gz) {
        gx_ = gx; gy_ = gy; gz_ = gz;
        if (gx_ != 0 || gy_ != 0 || gz_ != 0) sourceEnrolled = true;
    }
    void EnrollPointGravity(Real, Real, Real, Real, Real, Real, Real, Real) {
        throw std::runtime_error("point/binary gravity is not available on the device path");
    }
};

}  // namespace Gaukuk
