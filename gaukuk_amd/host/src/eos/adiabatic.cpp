// adiabatic.cpp -- ideal-gas conversions on the host (operation order of the reference's
// src/eos/adiabatic.cpp:18-61 and :162-198, which the device kernels also follow).
#include <algorithm>

#include "../utils/utils.hpp"
#include "eos.hpp"

namespace Gaukuk {

EquationOfState::EquationOfState() {
    gamma_ = Config::getInstance().get("gamma");
    gm1Rec_ = 1.0 / (gamma_ - 1.0);
}

const char* EquationOfState::Name() { return "adiabatic"; }

void EquationOfState::ConsToPrim(const TArray<Real>& cons, TArray<Real>& prim, const Grid& g) const {
    const int kl = g.nz == 1 ? g.kb : g.kgb, kr = g.nz == 1 ? g.ke : g.kge;
    const Real gm1 = gamma_ - 1.0;
    for (int k = kl; k < kr; k++)
        for (int j = g.jgb; j < g.jge; j++)
            for (int i = g.igb; i < g.ige; i++) {
                Real d = std::max(cons(DEN, k, j, i), (Real)DENSITY_FLOOR);
                const Real m1 = cons(MTX, k, j, i), m2 = cons(MTY, k, j, i), m3 = cons(MTZ, k, j, i);
                const Real inv = 1.0 / d;
                prim(DEN, k, j, i) = d;
                prim(VLX, k, j, i) = m1 * inv;
                prim(VLY, k, j, i) = m2 * inv;
                prim(VLZ, k, j, i) = m3 * inv;
                const Real ek = 0.5 * inv * (m1 * m1 + m2 * m2 + m3 * m3);
                const Real p = gm1 * (cons(ENG, k, j, i) - ek);
                prim(PRE, k, j, i) = p > PRESSURE_FLOOR ? p : (Real)PRESSURE_FLOOR;
            }
}

void EquationOfState::PrimToCons(const TArray<Real>& prim, TArray<Real>& cons, const Grid& g) const {
    const int kl = g.nz == 1 ? g.kb : g.kgb, kr = g.nz == 1 ? g.ke : g.kge;
    for (int k = kl; k < kr; k++)
        for (int j = g.jgb; j < g.jge; j++)
            for (int i = g.igb; i < g.ige; i++) {
                const Real d = prim(DEN, k, j, i), v1 = prim(VLX, k, j, i), v2 = prim(VLY, k, j, i);
                const Real v3 = prim(VLZ, k, j, i), p = prim(PRE, k, j, i);
                cons(DEN, k, j, i) = d;
                cons(MTX, k, j, i) = d * v1;
                cons(MTY, k, j, i) = d * v2;
                cons(MTZ, k, j, i) = d * v3;
                cons(ENG, k, j, i) = p * gm1Rec_ + 0.5 * d * (v1 * v1 + v2 * v2 + v3 * v3);
            }
}

}  // namespace Gaukuk
