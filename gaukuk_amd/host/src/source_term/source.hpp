// source.hpp -- source-term enrollment (the reference's SourceTerm, src/source_term/source.hpp:
// 23-50).  Constant gravity and point/binary gravity are applied on the device as Strang
// half-kicks fused into the first and last RK stage (gk_set_point_gravity for the masses).
#pragma once

#include <stdexcept>

#include "../gaukuk.hpp"

namespace Gaukuk {

class SourceTerm {
public:
    bool sourceEnrolled = false;
    Real gx_ = 0, gy_ = 0, gz_ = 0;
    // enrolled point masses: gm, x, y, z, vx, vy, vz, rs
    int nPoint = 0;
    Real mass[2][8] = {{0}};
    // src/source_term/source.cpp:23-31: enrolled only for a non-zero vector
    void EnrollConstGravityVector(Real gx, Real gy, Real gz) {
        gx_ = gx; gy_ = gy; gz_ = gz;
        if (gx_ != 0 || gy_ != 0 || gz_ != 0) sourceEnrolled = true;
    }
    // src/source_term/source.cpp:33-67: the first mass gives point gravity, a second binary
    // gravity, a third is an error
    void EnrollPointGravity(Real gm, Real x, Real y, Real z, Real vx, Real vy, Real vz, Real rs) {
        if (nPoint >= 2) throw std::runtime_error("Setup failed: too many point gravity.");
        const Real v[8] = {gm, x, y, z, vx, vy, vz, rs};
        for (int k = 0; k < 8; k++) mass[nPoint][k] = v[k];
        nPoint++;
        sourceEnrolled = true;
    }
};

}  // namespace Gaukuk
