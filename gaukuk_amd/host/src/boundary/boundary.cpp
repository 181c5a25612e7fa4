// boundary.cpp -- host ghost fill by face type (semantics of the reference's
// src/boundary/{outflow_copy,periodic,reflective,outflow,fixed}.cpp): each face is filled over the
// active range of the other axes only, so corners stay as they are (zero in fresh arrays);
// z faces only when nz > 1 (src/boundary/boundary.cpp:127-142).  The immersed shell and the inner
// fixed region rewrite active cells; that part runs on the device and UpdateBD marks it pending.
#include <algorithm>
#include <stdexcept>

#include "../utils/utils.hpp"
#include "boundary.hpp"

namespace Gaukuk {

namespace {

// ghost index `g` on axis of length n (active range [b, b+n)) -> source index; kind as below
int SourceIndex(int g, int b, int n, bool high, int code, int& kind) {
    kind = 0;
    if (code == 1) return high ? g - n : g + n;
    if (code == 2 || code == 3) {
        kind = code == 2 ? 1 : (high ? 3 : 2);
        return high ? (b + n) + (b + n) - g - 1 : b + b - g - 1;
    }
    return high ? b + n - 1 : b;
}

void Transform(Real u[5], int comp, int kind) {
    if (kind == 1) {
        u[comp] = -u[comp];
    } else if (kind >= 2) {
        const Real m = u[comp];
        const Real mc = kind == 2 ? std::min(m, (Real)0.0) : std::max(m, (Real)0.0);
        const Real d = std::max(u[DEN], (Real)DENSITY_FLOOR);
        const Real inv = 1.0 / d;
        const Real keOld = 0.5 * inv * m * m;
        const Real keNew = 0.5 * inv * mc * mc;
        Real eint = u[ENG] - keOld;
        eint = std::max(eint, (Real)1e-16);
        u[DEN] = d;
        u[comp] = mc;
        u[ENG] = eint + keNew;
    }
}

}  // namespace

Boundary::BoundaryFunc Boundary::ForType(int axis, bool high, int code, std::shared_ptr<const FixedState> fixed) {
    return [axis, high, code, fixed](TArray<Real>& c, const Grid& g, const EquationOfState&) {
        if (code == 4 && !(fixed && fixed->enrolled))
            throw std::invalid_argument("fixed boundary is not initialized. call "
                                        "'Boundary::EnrollFixedBoundaryData(const TArray<Real>& cons, "
                                        "const Grid& grid)' in your setup");
        const int lo[3] = {g.ib, g.jb, g.kb}, hi[3] = {g.ie, g.je, g.ke}, n[3] = {g.nx, g.ny, g.nz};
        int from[3] = {g.ib, g.jb, g.kb}, to[3] = {g.ie, g.je, g.ke};
        // ghost layers along `axis`, active range along the other two
        from[axis] = high ? hi[axis] : 0;
        to[axis] = high ? hi[axis] + g.nGhost : lo[axis];
        for (int k = from[2]; k < to[2]; k++)
            for (int j = from[1]; j < to[1]; j++)
                for (int i = from[0]; i < to[0]; i++) {
                    if (code == 4) {
                        for (int v = 0; v < NVar; v++) c(v, k, j, i) = fixed->data(v, k, j, i);
                        continue;
                    }
                    int src[3] = {i, j, k};
                    const int here[3] = {i, j, k};
                    int kind;
                    src[axis] = SourceIndex(here[axis], lo[axis], n[axis], high, code, kind);
                    Real u[5];
                    for (int v = 0; v < NVar; v++) u[v] = c(v, src[2], src[1], src[0]);
                    Transform(u, MTX + axis, kind);
                    for (int v = 0; v < NVar; v++) c(v, k, j, i) = u[v];
                }
    };
}

Boundary::Boundary() {
    static const char* keys[6] = {"xleft", "xright", "yleft", "yright", "zleft", "zright"};
    Config& c = Config::getInstance();
    for (int f = 0; f < 6; f++) {
        int code = (int)c.get(keys[f]);
        if (code < 0 || code > 4) code = 0;
        type[f] = code;
    }
    Bdxl = ForType(0, false, type[0], fixed_);
    Bdxr = ForType(0, true, type[1], fixed_);
    Bdyl = ForType(1, false, type[2], fixed_);
    Bdyr = ForType(1, true, type[3], fixed_);
    // z faces of type 0 (and unknown codes) register the y copy functions, as the reference does
    // (src/boundary/boundary.cpp:90-122): their ghost planes are never written
    const bool zl = type[4] >= 1 && type[4] <= 4, zr = type[5] >= 1 && type[5] <= 4;
    Bdzl = zl ? ForType(2, false, type[4], fixed_) : ForType(1, false, 0, fixed_);
    Bdzr = zr ? ForType(2, true, type[5], fixed_) : ForType(1, true, 0, fixed_);
}

void Boundary::EnrollFixedBoundaryData(const TArray<Real>& cons, const Grid&, bool fixInner) {
    fixed_->data = cons;
    fixed_->enrolled = true;
    fixed_->inner = fixInner;
    fixed_->version++;
}

void Boundary::SetupImmersedShell(const Grid& grid, const Domain&, Real xShell, Real yShell, Real zShell,
                                  Real rShell) {
    if (grid.nx <= 1 || grid.ny <= 1) throw std::invalid_argument("immersed boundary has to be at least 2D. ");
    hasShell_ = true;
    shell_[0] = xShell; shell_[1] = yShell; shell_[2] = zShell; shell_[3] = rShell;
}

const Boundary::BoundaryFunc& Boundary::Face(int f) const {
    switch (f) {
        case 0: return Bdxl;
        case 1: return Bdxr;
        case 2: return Bdyl;
        case 3: return Bdyr;
        case 4: return Bdzl;
        default: return Bdzr;
    }
}

bool Boundary::AnyFixed() const {
    for (int f = 0; f < 6; f++)
        if (type[f] == 4 && !selfDefined[f]) return true;
    return false;
}

bool Boundary::AnySelfDefined() const {
    for (bool b : selfDefined)
        if (b) return true;
    return false;
}

void Boundary::UpdateBD(TArray<Real>& cons, const Grid& grid, const Domain&, const EquationOfState& eos) {
    Bdxl(cons, grid, eos);
    Bdxr(cons, grid, eos);
    Bdyl(cons, grid, eos);
    Bdyr(cons, grid, eos);
    if (grid.nz > 1) {
        Bdzl(cons, grid, eos);
        Bdzr(cons, grid, eos);
    }
    if (hasShell_ || fixed_->inner) pending_ = true;
}

}  // namespace Gaukuk
