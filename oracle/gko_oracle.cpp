// gko_oracle.cpp -- CPU restatement of Gaukuk's hydro step (TEST INFRASTRUCTURE ONLY).
//
// Every function below cites the reference file:line whose arithmetic it restates.  The
// order of every floating-point operation is kept, because the strict build of this file
// must be bitwise equal to the strict build of the reference.  Only tests, smoke() and the
// bench's CPU-baseline leg use this library; the HIP product path never does.
#include "gko_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace {

constexpr int NV = 5;
enum { DEN = 0, MX = 1, MY = 2, MZ = 3, ENG = 4 };
enum { VX = 1, VY = 2, VZ = 3, PRE = 4 };

// std::min / std::max semantics (first argument returned on ties / unordered operands).
template <class R> inline R smin(R a, R b) { return (b < a) ? b : a; }
template <class R> inline R smax(R a, R b) { return (a < b) ? b : a; }

template <class Real>
struct Oracle {
    gko_params P;
    int nx, ny, nz, ng, lenx, leny, lenz, rows, ib, ie, jb, je, kb, ke;
    Real gamma, gm1Rec, cfl;
    Real dx, dy, dz, dxRec, dyRec, dzRec, drmin;
    Real gx, gy, gz;
    bool src_enrolled, has_const, has_point, has_binary;
    // point masses (GravitySource, src/source_term/source.hpp:11-20) and the binary's orbit
    struct Mass { Real gm, x, y, z, vx, vy, vz, rs; } obj0{}, obj1{};
    Real ab_ = 0, omega_ = 0, theta1_ = 0;
    std::vector<Real> xc, yc;   // cell centres (Domain, src/sim.cpp:37-39), ghost-inclusive index
    Real zc0 = 0;               // z centre of the active plane (2D)
    std::vector<Real> zc;       // 3D: z centres, ghost-inclusive
    // state
    std::vector<Real> cons, prim, flx1, flx2, flx3, consTemp;
    Real t, dt, dtUntilOutput, cmax;
    int step;
    bool isContinue;
    std::vector<double> log_t, log_dt;

    explicit Oracle(const gko_params& p) : P(p) {
        nx = p.nx; ny = p.ny; nz = p.nz < 1 ? 1 : p.nz; ng = p.ng;
        if (nx <= 0 || ny <= 0 || ng < 0) throw std::runtime_error("Invalid grid parameters");
        lenx = nx + 2 * ng; leny = ny + 2 * ng;
        // 2D keeps only the active z plane; 3D keeps every plane (Grid: src/grid.hpp:23-35)
        lenz = nz > 1 ? nz + 2 * ng : 1;
        rows = lenz * leny;
        ib = ng; ie = ib + nx; jb = ng; je = jb + ny;
        kb = nz > 1 ? ng : 0; ke = kb + nz;
        // EquationOfState ctor: src/eos/adiabatic.cpp:12-16
        gamma = (Real)p.gamma;
        gm1Rec = Real(1.0) / (gamma - Real(1.0));
        cfl = (Real)p.cfl;
        // Domain ctor: src/sim.cpp:16-36 (nz == 1)
        dx = ((Real)p.xmax - (Real)p.xmin) / (Real)nx;
        dy = ((Real)p.ymax - (Real)p.ymin) / (Real)ny;
        dz = ((Real)p.zmax - (Real)p.zmin) / (Real)nz;
        dxRec = Real(1.) / dx;
        dyRec = Real(1.) / dy;
        if (ny <= 1) dyRec = 0;
        dzRec = nz > 1 ? Real(1.) / dz : Real(0);
        drmin = smin(smin(dx, dy), dz);
        // SourceTerm::EnrollConstGravityVector: src/source_term/source.cpp:23-31
        gx = (Real)p.g[0]; gy = (Real)p.g[1]; gz = (Real)p.g[2];
        has_const = (gx != 0 || gy != 0 || gz != 0);
        has_point = has_binary = false;
        src_enrolled = has_const;
        // cell centres: xmin - nGhost*dx + 0.5*dx + i*dx (src/sim.cpp:37-39); the active z plane is
        // k = nGhost of lenz = 1 + 2 nGhost
        xc.resize(lenx); yc.resize(leny);
        for (int i = 0; i < lenx; i++) xc[i] = (Real)p.xmin - (Real)ng * dx + (Real)0.5 * dx + (Real)i * dx;
        for (int j = 0; j < leny; j++) yc[j] = (Real)p.ymin - (Real)ng * dy + (Real)0.5 * dy + (Real)j * dy;
        zc0 = (Real)p.zmin - (Real)ng * dz + (Real)0.5 * dz + (Real)ng * dz;
        zc.resize(nz + 2 * ng);
        for (int k = 0; k < nz + 2 * ng; k++) zc[k] = (Real)p.zmin - (Real)ng * dz + (Real)0.5 * dz + (Real)k * dz;
        // SourceTerm::EnrollPointGravity, called npoint times: src/source_term/source.cpp:33-67
        for (int q = 0; q < p.npoint && q < 2; q++) {
            const double* g = p.pg[q];
            EnrollPointGravity((Real)g[0], (Real)g[1], (Real)g[2], (Real)g[3], (Real)g[4], (Real)g[5], (Real)g[6],
                               (Real)g[7]);
        }
        size_t n = (size_t)NV * rows * lenx;
        cons.assign(n, 0); prim.assign(n, 0); consTemp.assign(n, 0);
        flx1.assign((size_t)NV * rows * (lenx + 1), 0);
        flx2.assign((size_t)NV * lenz * (leny + 1) * lenx, 0);
        if (nz > 1) flx3.assign((size_t)NV * (lenz + 1) * leny * lenx, 0);
        // Sim ctor: src/sim.cpp:45-58
        t = 0; dt = (Real)1e10; dtUntilOutput = (Real)1e10; cmax = (Real)1e-16;
        step = 0; isContinue = true;
    }

    // Arrays are [n][k][j][i] (x fastest); `j` below is a row of plane k, k = 0 in 2D.
    inline size_t at(int n, int j, int i, int k = 0) const {
        return (size_t)i + (size_t)lenx * ((size_t)j + (size_t)leny * ((size_t)k + (size_t)lenz * n));
    }
    inline size_t at1(int n, int j, int i, int k = 0) const {
        return (size_t)i + (size_t)(lenx + 1) * ((size_t)j + (size_t)leny * ((size_t)k + (size_t)lenz * n));
    }
    inline size_t at2(int n, int j, int i, int k = 0) const {
        return (size_t)i + (size_t)lenx * ((size_t)j + (size_t)(leny + 1) * ((size_t)k + (size_t)lenz * n));
    }
    inline size_t at3(int n, int k, int j, int i) const {
        return (size_t)i + (size_t)lenx * ((size_t)j + (size_t)leny * ((size_t)k + (size_t)(lenz + 1) * n));
    }

    // ---- EOS -------------------------------------------------------------------------
    // ConsToPrim over all cells incl. ghosts: src/eos/adiabatic.cpp:18-61
    void ConsToPrim(const std::vector<Real>& U, std::vector<Real>& W) {
        const Real gm1 = gamma - Real(1.0);
        const Real dmin = (Real)1e-16, pmin = (Real)1e-16;
        for (int k = 0; k < lenz; k++) {
#pragma omp parallel for schedule(static)
            for (int j = 0; j < leny; j++) {
                for (int i = 0; i < lenx; i++) {
                    Real d = U[at(DEN, j, i, k)], m1 = U[at(MX, j, i, k)], m2 = U[at(MY, j, i, k)],
                         m3 = U[at(MZ, j, i, k)], e = U[at(ENG, j, i, k)];
                    d = smax(d, dmin);
                    W[at(DEN, j, i, k)] = d;
                    Real inv = Real(1.0) / d;
                    W[at(VX, j, i, k)] = m1 * inv;
                    W[at(VY, j, i, k)] = m2 * inv;
                    W[at(VZ, j, i, k)] = m3 * inv;
                    Real ek = Real(0.5) * inv * (m1 * m1 + m2 * m2 + m3 * m3);
                    Real p = gm1 * (e - ek);
                    W[at(PRE, j, i, k)] = (p > pmin) ? p : pmin;
                }
            }
        }
    }

    // CalCmax over active cells: src/eos/adiabatic.cpp:115-160 (its own kinetic-energy form, :148)
    void CalCmax(const std::vector<Real>& U, Real& cm) {
        const Real gm1 = gamma - Real(1.0);
        const Real dmin = (Real)1e-16, pmin = (Real)1e-16;
        Real best = cm;
        for (int k = kb; k < ke; k++) {
#pragma omp parallel for schedule(static) reduction(max : best)
        for (int j = jb; j < je; j++) {
            for (int i = ib; i < ie; i++) {
                Real d = U[at(DEN, j, i, k)], m1 = U[at(MX, j, i, k)], m2 = U[at(MY, j, i, k)],
                     m3 = U[at(MZ, j, i, k)], e = U[at(ENG, j, i, k)];
                d = smax(d, dmin);
                Real inv = Real(1.0) / d;
                Real v1 = m1 * inv, v2 = m2 * inv, v3 = m3 * inv;
                Real ek = Real(0.5) * (v1 * m1 + v2 * m2 + v3 * m3);
                Real p = gm1 * (e - ek);
                p = (p > pmin) ? p : pmin;
                Real cs = std::sqrt(gamma * p * inv);
                Real loc = std::abs(v1) + cs;
                loc = smax(loc, std::abs(v2) + cs);
                loc = smax(loc, std::abs(v3) + cs);
                best = smax(best, loc);
            }
        }
        }
        cm = best;
    }

    // ---- Boundaries (types 0..4), each face over the active tangential range ------------
    // Type 0: src/boundary/outflow_copy.cpp:16-124; 1: periodic.cpp:19-131;
    // 2: reflective.cpp:19-166; 3: outflow.cpp:20-213.  Dispatch order Bdxl, Bdxr, Bdyl,
    // Bdyr: src/boundary/boundary.cpp:127-142 (z faces only when nz > 1: not here).
    // Unknown codes fall back to type 0 (boundary.cpp:27-29 and siblings).
    // One ghost cell g of a face filled from source cell s (types 1-3; type 0 copies the edge cell):
    // comp is the normal momentum component, lowSide selects the outflow clamp's sign.
    void FillCell(std::vector<Real>& U, size_t g[NV], size_t s[NV], int type, int comp, bool lowSide) {
        if (type == 2 || type == 3) {
            Real d = U[s[DEN]], m1 = U[s[MX]], m2 = U[s[MY]], m3 = U[s[MZ]], e = U[s[ENG]];
            Real m[3] = {m1, m2, m3};
            Real& mn = m[comp - MX];
            if (type == 2) {
                mn = -mn;
                U[g[DEN]] = d; U[g[MX]] = m[0]; U[g[MY]] = m[1]; U[g[MZ]] = m[2]; U[g[ENG]] = e;
            } else {
                const Real mo = mn;
                Real mc = lowSide ? smin(mo, Real(0.0)) : smax(mo, Real(0.0));
                d = smax(d, (Real)1e-16);
                Real inv = Real(1.0) / d;
                Real keOld = Real(0.5) * inv * mo * mo;
                Real keNew = Real(0.5) * inv * mc * mc;
                Real eint = e - keOld;
                eint = smax(eint, (Real)1e-16);
                mn = mc;
                U[g[DEN]] = d; U[g[MX]] = m[0]; U[g[MY]] = m[1]; U[g[MZ]] = m[2]; U[g[ENG]] = eint + keNew;
            }
        } else {
            for (int n = 0; n < NV; n++) U[g[n]] = U[s[n]];
        }
    }
    void FillX(std::vector<Real>& U, bool left, int type) {
        const int gl = left ? 0 : ie, gr = left ? ib : lenx;
        for (int k = kb; k < ke; k++)
            for (int j = jb; j < je; j++)
                for (int i = gl; i < gr; i++) {
                    int src = left ? ib : ie - 1;
                    if (type == 1) src = left ? i + nx : i - nx;
                    else if (type == 2 || type == 3) src = left ? (ib + ib - i - 1) : (ie + ie - i - 1);
                    size_t g[NV], sv[NV];
                    for (int n = 0; n < NV; n++) { g[n] = at(n, j, i, k); sv[n] = at(n, j, src, k); }
                    FillCell(U, g, sv, type, MX, left);
                }
    }
    void FillY(std::vector<Real>& U, bool left, int type) {
        const int gl = left ? 0 : je, gr = left ? jb : leny;
        for (int k = kb; k < ke; k++)
            for (int j = gl; j < gr; j++)
                for (int i = ib; i < ie; i++) {
                    int src = left ? jb : je - 1;
                    if (type == 1) src = left ? j + ny : j - ny;
                    else if (type == 2 || type == 3) src = left ? (jb + jb - j - 1) : (je + je - j - 1);
                    size_t g[NV], sv[NV];
                    for (int n = 0; n < NV; n++) { g[n] = at(n, j, i, k); sv[n] = at(n, src, i, k); }
                    FillCell(U, g, sv, type, MY, left);
                }
    }
    // z faces (3D): types 1-3 over the active (j, i) range (periodic.cpp:135-190,
    // reflective.cpp:170-230, outflow.cpp:216-300).  Type 0 and unknown codes register the *y* copy
    // functions (boundary.cpp:90-122: Bdzl = &OutflowCopyYL, Bdzr = &OutflowCopyYR), so the z ghost
    // planes are left as they are and the y ghosts are copied once more.
    void FillZ(std::vector<Real>& U, bool left, int type) {
        if (type != 1 && type != 2 && type != 3) {
            FillY(U, left, 0);
            return;
        }
        const int gl = left ? 0 : ke, gr = left ? kb : lenz;
        for (int k = gl; k < gr; k++)
            for (int j = jb; j < je; j++)
                for (int i = ib; i < ie; i++) {
                    int src = type == 1 ? (left ? k + nz : k - nz) : (left ? (kb + kb - k - 1) : (ke + ke - k - 1));
                    size_t g[NV], sv[NV];
                    for (int n = 0; n < NV; n++) { g[n] = at(n, j, i, k); sv[n] = at(n, j, i, src); }
                    FillCell(U, g, sv, type, MZ, left);
                }
    }
    // Type 4, fixed: src/boundary/fixed.cpp:16-127 -- the ghost layers saved by
    // EnrollFixedBoundaryData (fixed.cpp:242-269) are copied back over the active tangential
    // range.  Here the saved layers are a snapshot of the whole ghosted array.
    std::vector<Real> fixedSnap;
    std::vector<Real> fixedInner;   // EnrollFixedBoundaryData(..., fixInner = true): consFixedInner = cons
    void EnrollFixedBoundaryData(bool fixInner = false) {
        fixedSnap = cons;
        if (fixInner) fixedInner = cons;
    }
    // every active plane k (fixed.cpp: k in [kb, ke))
    void FillFixedX(std::vector<Real>& U, bool left) {
        if (fixedSnap.empty())
            throw std::invalid_argument("fixed boundary is not initialized");
        const int gl = left ? 0 : ie, gr = left ? ib : lenx;
        for (int n = 0; n < NV; n++)
            for (int k = kb; k < ke; k++)
                for (int j = jb; j < je; j++)
                    for (int i = gl; i < gr; i++) U[at(n, j, i, k)] = fixedSnap[at(n, j, i, k)];
    }
    void FillFixedY(std::vector<Real>& U, bool left) {
        if (fixedSnap.empty())
            throw std::invalid_argument("fixed boundary is not initialized");
        const int gl = left ? 0 : je, gr = left ? jb : leny;
        for (int n = 0; n < NV; n++)
            for (int k = kb; k < ke; k++)
                for (int j = gl; j < gr; j++)
                    for (int i = ib; i < ie; i++) U[at(n, j, i, k)] = fixedSnap[at(n, j, i, k)];
    }
    // z faces (3D): the saved ghost planes over the active (j, i) range (fixed.cpp:126-179)
    void FillFixedZ(std::vector<Real>& U, bool left) {
        if (fixedSnap.empty())
            throw std::invalid_argument("fixed boundary is not initialized");
        const int gl = left ? 0 : ke, gr = left ? kb : lenz;
        for (int n = 0; n < NV; n++)
            for (int k = gl; k < gr; k++)
                for (int j = jb; j < je; j++)
                    for (int i = ib; i < ie; i++) U[at(n, j, i, k)] = fixedSnap[at(n, j, i, k)];
    }
    // Inner fixed region: src/boundary/fixed.cpp:181-207.  Cells of the central half-box in x and
    // y, on every active plane, whose centre lies within 0.5 of the origin take the saved state.
    void FixedBDInner(std::vector<Real>& U) {
        const int il = static_cast<int>(ib + 0.25 * (ie - ib));
        const int ir = static_cast<int>(ib + 0.75 * (ie - ib));
        const int jl = static_cast<int>(jb + 0.25 * (je - jb));
        const int jr = static_cast<int>(jb + 0.75 * (je - jb));
        for (int n = 0; n < NV; n++)
            for (int k = kb; k < ke; k++)
                for (int j = jl; j < jr; j++)
                    for (int i = il; i < ir; i++) {
                        const Real xnow = xc[i], ynow = yc[j], znow = nz > 1 ? zc[k] : zc0;
                        const Real r2 = xnow * xnow + ynow * ynow + znow * znow;
                        if (r2 < 0.5 * 0.5) U[at(n, j, i, k)] = fixedInner[at(n, j, i, k)];
                    }
    }

    // ---- Immersed reflective shell: src/boundary/immersed.cpp -----------------------------
    static constexpr int nTarMax = 8;
    struct Shell {
        int i, j, k, nTar;
        int ti[nTarMax + 1], tj[nTarMax + 1], tk[nTarMax + 1];
        Real cosp, sinp, cost, sint;
        Real gam[nTarMax + 1];
    };
    std::vector<Shell> shells;
    bool immersed = false;
    // FindLeftId / FindRightId: immersed.cpp:19-47
    static int FindLeftId(const std::vector<Real>& c, int b, int e, Real x0, Real r, Real lo) {
        const Real xb = x0 - 1.2 * r;
        if (xb < lo) throw std::invalid_argument("immersed boundary out of domain");
        for (int i = b; i < e; i++)
            if (c[i] >= xb) return i;
        return e;
    }
    static int FindRightId(const std::vector<Real>& c, int b, int e, Real x0, Real r, Real hi) {
        const Real xe = x0 + 1.2 * r;
        if (xe > hi) throw std::invalid_argument("immersed boundary out of domain");
        for (int i = e - 1; i >= b; i--)
            if (c[i] <= xe) return i;
        return b;
    }
    // IsShellCell: immersed.cpp:49-69 -- inside the body, with a neighbour nGhost cells away outside
    static bool IsShellCell(Real dx_, Real dy_, Real dz_, Real r, Real sx, Real sy, Real sz) {
        Real rn = std::sqrt(dx_ * dx_ + dy_ * dy_ + dz_ * dz_);
        if (rn < r) {
            rn = std::sqrt((dx_ + sx) * (dx_ + sx) + dy_ * dy_ + dz_ * dz_); if (rn > r) return true;
            rn = std::sqrt((dx_ - sx) * (dx_ - sx) + dy_ * dy_ + dz_ * dz_); if (rn > r) return true;
            rn = std::sqrt(dx_ * dx_ + (dy_ + sy) * (dy_ + sy) + dz_ * dz_); if (rn > r) return true;
            rn = std::sqrt(dx_ * dx_ + (dy_ - sy) * (dy_ - sy) + dz_ * dz_); if (rn > r) return true;
            rn = std::sqrt(dx_ * dx_ + dy_ * dy_ + (dz_ + sz) * (dz_ + sz)); if (rn > r) return true;
            rn = std::sqrt(dx_ * dx_ + dy_ * dy_ + (dz_ - sz) * (dz_ - sz)); if (rn > r) return true;
        }
        return false;
    }
    // Boundary::SetupImmersedShell: immersed.cpp:89-235 (2D: the active z plane only).  The shell
    // cells are listed in (k, j, i) order; the reference's order depends on its threads, and the
    // result does not (shell cells and their targets are disjoint).
    Real zcAt(int k) const { return nz > 1 ? zc[k] : zc0; }
    void SetupImmersedShell(Real xs, Real ys, Real zs, Real rs) {
        if (nx <= 1 || ny <= 1) throw std::invalid_argument("immersed boundary has to be at least 2D. ");
        immersed = true;
        const int il = FindLeftId(xc, ib, ie, xs, rs, (Real)P.xmin);
        const int ir = FindRightId(xc, ib, ie, xs, rs, (Real)P.xmax);
        const int jl = FindLeftId(yc, jb, je, ys, rs, (Real)P.ymin);
        const int jr = FindRightId(yc, jb, je, ys, rs, (Real)P.ymax);
        int kl = kb, kr = ke;
        if (nz > 1) {
            kl = FindLeftId(zc, kb, ke, zs, rs, (Real)P.zmin);
            kr = FindRightId(zc, kb, ke, zs, rs, (Real)P.zmax);
        }
        if (il > ir || jl > jr || kl > kr) throw std::invalid_argument("immersed boundary out of domain");
        shells.clear();
        const Real sx = ng * dx, sy = ng * dy, sz = ng * dz;
        for (int k = kl; k < kr; k++)
            for (int j = jl; j < jr; j++)
                for (int i = il; i < ir; i++)
                    if (IsShellCell(xc[i] - xs, yc[j] - ys, zcAt(k) - zs, rs, sx, sy, sz)) {
                        Shell s{};
                        s.i = i; s.j = j; s.k = k;
                        shells.push_back(s);
                    }
        const int nxs = 2 * ng, nys = 2 * ng, nzs = nz > 1 ? 2 * ng : 0;
        for (Shell& s : shells) {
            s.ti[0] = s.i; s.tj[0] = s.j; s.tk[0] = s.k;
            const Real delx = xc[s.i] - xs, dely = yc[s.j] - ys, delz = zcAt(s.k) - zs;
            const Real rnow = std::sqrt(delx * delx + dely * dely + delz * delz);
            Real cosp, sinp, cost, sint;
            const Real rxy = std::sqrt(delx * delx + dely * dely);
            if (rxy < 1e-15) { cosp = 1.0; sinp = 0.0; } else { cosp = delx / rxy; sinp = dely / rxy; }
            if (rnow < 1e-15) { cost = 1.0; sint = 0.0; } else { cost = delz / rnow; sint = rxy / rnow; }
            s.cosp = cosp; s.sinp = sinp; s.cost = cost; s.sint = sint;
            const Real xImg = sint * cosp * (2 * rs - rnow) + xs;
            const Real yImg = sint * sinp * (2 * rs - rnow) + ys;
            const Real zImg = cost * (2 * rs - rnow) + zs;
            int nt = 0;
            Real disDel[nTarMax + 1];
            disDel[0] = 2 * (rs - rnow);
            for (int ii = -nxs; ii < nxs + 1; ++ii)
                for (int jj = -nys; jj < nys + 1; ++jj)
                    for (int kk = -nzs; kk < nzs + 1; ++kk) {
                        const int iT = s.i + ii, jT = s.j + jj, kT = s.k + kk;
                        if (iT < ib || iT >= ie || jT < jb || jT >= je || kT < kb || kT >= ke)
                            throw std::invalid_argument("immersed body is too close to the domain boundary, "
                                                        "neighboor search out of range.");
                        const Real xT = xc[iT], yT = yc[jT], zT = zcAt(kT);
                        if (std::abs(xT - xImg) <= dx && std::abs(yT - yImg) <= dy && std::abs(zT - zImg) <= dz) {
                            const Real rT =
                                std::sqrt((xT - xs) * (xT - xs) + (yT - ys) * (yT - ys) + (zT - zs) * (zT - zs));
                            if (rT > rs) {
                                nt++;
                                if (nt > nTarMax) throw std::logic_error("n of imagine points is more than nTarMax");
                                s.ti[nt] = iT; s.tj[nt] = jT; s.tk[nt] = kT;
                                disDel[nt] = std::sqrt((xT - xImg) * (xT - xImg) + (yT - yImg) * (yT - yImg) +
                                                       (zT - zImg) * (zT - zImg));
                            }
                        }
                    }
            s.nTar = nt;
            Real sum = 0.;
            for (int n = 0; n < nt + 1; n++) sum = sum + 1. / (disDel[n] * disDel[n]);
            for (int n = 0; n < nt + 1; n++) s.gam[n] = 1. / (disDel[n] * disDel[n] * sum);
        }
    }
    // Boundary::ImmersedReflective: immersed.cpp:237-300 -- each shell cell takes the weighted image
    // of its targets with the radial velocity reflected
    void ImmersedReflective(std::vector<Real>& U) {
        const Real dmin = (Real)1e-16, pmin = (Real)1e-16;
        const Real gm1 = gamma - 1;
        for (const Shell& s : shells) {
            const Real cosp = s.cosp, sinp = s.sinp, cost = s.cost, sint = s.sint;
            Real dens = 0, vr = 0, vp = 0, vt = 0, pres = 0;
            const Real g0 = s.gam[0];
            for (int n = 1; n <= s.nTar; n++) {
                const int iT = s.ti[n], jT = s.tj[n], kT = s.tk[n];
                Real denTar = U[at(DEN, jT, iT, kT)];
                denTar = std::max(denTar, dmin);
                const Real denInv = 1.0 / denTar;
                dens += 1. / (1 - g0) * s.gam[n] * denTar;
                const Real vx = U[at(MX, jT, iT, kT)] * denInv;
                const Real vy = U[at(MY, jT, iT, kT)] * denInv;
                const Real vz = U[at(MZ, jT, iT, kT)] * denInv;
                // CarToLocal: immersed.cpp:73-79
                const Real vrl = vx * cosp * sint + vy * sinp * sint + vz * cost;
                const Real vpl = -vx * sinp + vy * cosp;
                const Real vtl = vx * cosp * cost + vy * sinp * cost - vz * sint;
                vr += 1. / (-1. - g0) * s.gam[n] * vrl;
                vp += 1. / (1. - g0) * s.gam[n] * vpl;
                vt += 1. / (1. - g0) * s.gam[n] * vtl;
                const Real engTar = U[at(ENG, jT, iT, kT)];
                const Real engKin = 0.5 * denTar * (vx * vx + vy * vy + vz * vz);
                Real preTar = gm1 * (engTar - engKin);
                preTar = (preTar > pmin) ? preTar : pmin;
                pres += 1. / (1. - g0) * s.gam[n] * preTar;
            }
            // LocalToCar: immersed.cpp:82-87
            const Real vx = vr * sint * cosp + vt * cost * cosp - vp * sinp;
            const Real vy = vr * sint * sinp + vt * cost * sinp + vp * cosp;
            const Real vz = vr * cost - vt * sint;
            U[at(DEN, s.j, s.i, s.k)] = dens;
            U[at(MX, s.j, s.i, s.k)] = dens * vx;
            U[at(MY, s.j, s.i, s.k)] = dens * vy;
            U[at(MZ, s.j, s.i, s.k)] = dens * vz;
            const Real engKin = 0.5 * dens * (vx * vx + vy * vy + vz * vz);
            U[at(ENG, s.j, s.i, s.k)] = engKin + pres * gm1Rec;   // EGas: eos.hpp:30-32
        }
    }

    // Boundary::UpdateBD: src/boundary/boundary.cpp:127-142 -- faces, then the immersed shell,
    // then the inner fixed region
    void UpdateBD(std::vector<Real>& U) {
        if (P.bc[0] == 4) FillFixedX(U, true); else FillX(U, true, P.bc[0]);
        if (P.bc[1] == 4) FillFixedX(U, false); else FillX(U, false, P.bc[1]);
        if (P.bc[2] == 4) FillFixedY(U, true); else FillY(U, true, P.bc[2]);
        if (P.bc[3] == 4) FillFixedY(U, false); else FillY(U, false, P.bc[3]);
        if (nz > 1) {
            if (P.bc[4] == 4) FillFixedZ(U, true); else FillZ(U, true, P.bc[4]);
            if (P.bc[5] == 4) FillFixedZ(U, false); else FillZ(U, false, P.bc[5]);
        }
        if (immersed) ImmersedReflective(U);
        if (!fixedInner.empty()) FixedBDInner(U);
    }

    // ---- Reconstruction + HLLC ---------------------------------------------------------
    // VanLeer: src/reconstruction/reconstruction.hpp:36-42 (division evaluated, then discarded)
    static inline Real VanLeer(Real a, Real b) {
        Real multi = a * b;
        Real sigma = Real(2.0) * multi / (a + b);
        return (multi <= Real(0.0)) ? Real(0.0) : sigma;
    }

    // HLLC at one face: src/flux/hllc.cpp:22-103.  wl/wr are primitive states (den, v_n,
    // v_t1, v_t2, p) already rotated so that index 1 is the face normal; f receives
    // (F_den, F_n, F_t1, F_t2, F_E) in the same rotated order.
    inline void Hllc(const Real* wl, const Real* wr, Real* f) const {
        const Real gmRec = Real(1.0) / gamma;
        Real denl = wl[0], vxl = wl[1], vyl = wl[2], vzl = wl[3], prel = wl[4];
        Real denr = wr[0], vxr = wr[1], vyr = wr[2], vzr = wr[3], prer = wr[4];
        Real al = std::sqrt(gamma * prel / denl);
        Real ar = std::sqrt(gamma * prer / denr);
        Real pv = Real(0.5) * (prel + prer) - Real(0.125) * (vxr - vxl) * (denl + denr) * (al + ar);
        pv = smax(pv, Real(0));
        Real tmp = smax(pv / prel - 1, Real(0));
        Real ql = std::sqrt(1 + Real(0.5) * (1 + gmRec) * tmp);
        tmp = smax(pv / prer - 1, Real(0));
        Real qr = std::sqrt(1 + Real(0.5) * (1 + gmRec) * tmp);
        Real sl = vxl - al * ql;
        Real sr = vxr + ar * qr;
        Real ss = (prer - prel + denl * vxl * (sl - vxl) - denr * vxr * (sr - vxr)) /
                  (denl * (sl - vxl) - denr * (sr - vxr));
        Real el = prel * gm1Rec + Real(0.5) * denl * (vxl * vxl + vyl * vyl + vzl * vzl);
        Real er = prer * gm1Rec + Real(0.5) * denr * (vxr * vxr + vyr * vyr + vzr * vzr);
        Real cl = smin(Real(0), sl) * (ss - vxl) / (sl - ss);
        Real cr = smax(Real(0), sr) * (ss - vxr) / (sr - ss);
        Real sel = (ss >= 0) ? Real(1.0) : Real(0.0);
        Real ser = Real(1.0) - sel;
        Real fl1 = denl * (vxl + cl);
        Real fl2 = denl * (vxl * vxl + cl * sl) + prel;
        Real fl3 = fl1 * vyl;
        Real fl4 = fl1 * vzl;
        Real fl5 = (el + prel) * (vxl + cl) + denl * cl * ss * (sl - vxl);
        Real fr1 = denr * (vxr + cr);
        Real fr2 = denr * (vxr * vxr + cr * sr) + prer;
        Real fr3 = fr1 * vyr;
        Real fr4 = fr1 * vzr;
        Real fr5 = (er + prer) * (vxr + cr) + denr * cr * ss * (sr - vxr);
        f[0] = sel * fl1 + ser * fr1;
        f[1] = sel * fl2 + ser * fr2;
        f[2] = sel * fl3 + ser * fr3;
        f[3] = sel * fl4 + ser * fr4;
        f[4] = sel * fl5 + ser * fr5;
    }

    // Left state at face f from cell f-1, right state from cell f (PLM: recon_piecewise_linear.cpp:8-55;
    // first order: recon_first_order.cpp:8-29).
    inline Real CellVal(int n, int j, int i, int k = 0) const { return prim[at(n, j, i, k)]; }

    // x sweep over active rows (and planes), faces ib..ie (the faces UpdateCons reads): cal_flux.cpp:90-101
    void SweepX() {
        const int rotate[NV] = {DEN, VX, VY, VZ, PRE};
        for (int k = kb; k < ke; k++) {
#pragma omp parallel for schedule(static)
            for (int j = jb; j < je; j++) {
                Real wl[NV], wr[NV], f[NV];
                for (int fi = ib; fi <= ie; fi++) {
                    for (int q = 0; q < NV; q++) {
                        int n = rotate[q];
                        if (P.rc_order == 1) {
                            wl[q] = CellVal(n, j, fi - 1, k);
                            wr[q] = CellVal(n, j, fi, k);
                        } else {
                            Real a0 = CellVal(n, j, fi - 2, k), a1 = CellVal(n, j, fi - 1, k),
                                 a2 = CellVal(n, j, fi, k), a3 = CellVal(n, j, fi + 1, k);
                            Real sl = VanLeer(a1 - a0, a2 - a1);
                            Real sr = VanLeer(a2 - a1, a3 - a2);
                            wl[q] = a1 + Real(0.5) * sl;
                            wr[q] = a2 - Real(0.5) * sr;
                        }
                    }
                    Hllc(wl, wr, f);
                    for (int q = 0; q < NV; q++) flx1[at1(rotate[q], j, fi, k)] = f[q];
                }
            }
        }
    }

    // y sweep, faces jb..je over active columns (and planes): cal_flux.cpp:103-121 (2D),
    // 123-138 (3D).  HLLC rotation for direction VLY (hllc.cpp:25-27): normal vy, tangents (vz, vx).
    void SweepY() {
        const int rotate[NV] = {DEN, VY, VZ, VX, PRE};
        for (int k = kb; k < ke; k++) {
#pragma omp parallel for schedule(static)
            for (int fj = jb; fj <= je; fj++) {
                Real wl[NV], wr[NV], f[NV];
                for (int i = ib; i < ie; i++) {
                    for (int q = 0; q < NV; q++) {
                        int n = rotate[q];
                        if (P.rc_order == 1) {
                            wl[q] = CellVal(n, fj - 1, i, k);
                            wr[q] = CellVal(n, fj, i, k);
                        } else {
                            Real a0 = CellVal(n, fj - 2, i, k), a1 = CellVal(n, fj - 1, i, k),
                                 a2 = CellVal(n, fj, i, k), a3 = CellVal(n, fj + 1, i, k);
                            Real sl = VanLeer(a1 - a0, a2 - a1);
                            Real sr = VanLeer(a2 - a1, a3 - a2);
                            wl[q] = a1 + Real(0.5) * sl;
                            wr[q] = a2 - Real(0.5) * sr;
                        }
                    }
                    Hllc(wl, wr, f);
                    for (int q = 0; q < NV; q++) flx2[at2(rotate[q], fj, i, k)] = f[q];
                }
            }
        }
    }

    // z sweep (3D), faces kb..ke over active rows and columns: cal_flux.cpp:140-157.  The PLM states
    // come from ReconstructYZPLM(prim, ..., k, j) (recon_piecewise_linear.cpp:34-55), whose slope is
    // taken along y: the face between planes k-1 and k takes cell (k-1, j) plus half its y slope on
    // the left and cell (k, j) minus half its y slope on the right.  First order copies the cells
    // (recon_first_order.cpp:18-29).  HLLC rotation for VLZ: normal vz, tangents (vx, vy).
    void SweepZ() {
        const int rotate[NV] = {DEN, VZ, VX, VY, PRE};
        for (int fk = kb; fk <= ke; fk++) {
#pragma omp parallel for schedule(static)
            for (int j = jb; j < je; j++) {
                Real wl[NV], wr[NV], f[NV];
                for (int i = ib; i < ie; i++) {
                    for (int q = 0; q < NV; q++) {
                        int n = rotate[q];
                        if (P.rc_order == 1) {
                            wl[q] = CellVal(n, j, i, fk - 1);
                            wr[q] = CellVal(n, j, i, fk);
                        } else {
                            const Real l0 = CellVal(n, j - 1, i, fk - 1), l1 = CellVal(n, j, i, fk - 1),
                                       l2 = CellVal(n, j + 1, i, fk - 1);
                            const Real r0 = CellVal(n, j - 1, i, fk), r1 = CellVal(n, j, i, fk),
                                       r2 = CellVal(n, j + 1, i, fk);
                            wl[q] = l1 + Real(0.5) * VanLeer(l1 - l0, l2 - l1);
                            wr[q] = r1 - Real(0.5) * VanLeer(r1 - r0, r2 - r1);
                        }
                    }
                    Hllc(wl, wr, f);
                    for (int q = 0; q < NV; q++) flx3[at3(rotate[q], fk, j, i)] = f[q];
                }
            }
        }
    }

    void CalFlux() {
        SweepX();
        SweepY();
        if (nz > 1) SweepZ();
    }

    // ---- UpdateCons: src/evolution/update_cons.cpp:11-75 (2 coef), 135-205 (3 coef); the 3D form
    //      adds the z flux difference last (update_cons.cpp:76-132) -----------------------------
    void UpdateCons(const std::vector<Real>& U0, std::vector<Real>& UT, Real c1, Real c2, Real c3, bool three) {
        const Real dtdx = dt * dxRec, dtdy = dt * dyRec, dtdz = dt * dzRec;
        const Real dmin = (Real)1e-16, pmin = (Real)1e-16;
        const Real gm1 = gamma - 1;
        for (int k = kb; k < ke; k++) {
#pragma omp parallel for schedule(static)
            for (int j = jb; j < je; j++) {
                for (int i = ib; i < ie; i++) {
                    Real r[NV];
                    for (int n = 0; n < NV; n++) {
                        Real dF = (flx1[at1(n, j, i + 1, k)] - flx1[at1(n, j, i, k)]) * dtdx +
                                  (flx2[at2(n, j + 1, i, k)] - flx2[at2(n, j, i, k)]) * dtdy;
                        if (nz > 1) dF = dF + (flx3[at3(n, k + 1, j, i)] - flx3[at3(n, k, j, i)]) * dtdz;
                        if (three)
                            r[n] = c1 * U0[at(n, j, i, k)] + c2 * UT[at(n, j, i, k)] - c3 * dF;
                        else
                            r[n] = c1 * U0[at(n, j, i, k)] - c2 * dF;
                    }
                    if (r[DEN] < dmin) {
                        r[DEN] = dmin; r[MX] = 0.0; r[MY] = 0.0; r[MZ] = 0.0;
                        r[ENG] = pmin / gm1 + Real(0.0);
                    } else {
                        Real inv = Real(1.0) / r[DEN];
                        Real ke_ = Real(0.5) * inv * (r[MX] * r[MX] + r[MY] * r[MY] + r[MZ] * r[MZ]);
                        Real press = gm1 * (r[ENG] - ke_);
                        if (press < pmin) r[ENG] = pmin / gm1 + ke_;
                    }
                    for (int n = 0; n < NV; n++) UT[at(n, j, i, k)] = r[n];
                }
            }
        }
    }

    // ---- Gravity half-kick: src/source_term/const_gravity.cpp:10-67 ---------------------------
    void ConstGravity(std::vector<Real>& U, Real h) {
        const Real g[3] = {gx, gy, gz};
        for (int a = 0; a < 3; a++) {
            if (g[a] == 0) continue;
            const int m = MX + a;
            for (int k = kb; k < ke; k++) {
#pragma omp parallel for schedule(static)
                for (int j = jb; j < je; j++) {
                    for (int i = ib; i < ie; i++) {
                        U[at(ENG, j, i, k)] += U[at(m, j, i, k)] * g[a] * h;
                        U[at(m, j, i, k)] += U[at(DEN, j, i, k)] * g[a] * h;
                    }
                }
            }
        }
    }

    void EnrollPointGravity(Real gm, Real x, Real y, Real z, Real vx, Real vy, Real vz, Real rs) {
        if (has_binary) throw std::runtime_error("Setup failed: too many point gravity.");
        if (has_point) {
            obj1 = {gm, x, y, z, vx, vy, vz, rs};
            ab_ = std::sqrt((x - obj0.x) * (x - obj0.x) + (y - obj0.y) * (y - obj0.y) + (z - obj0.z) * (z - obj0.z));
            omega_ = std::sqrt((obj0.gm + gm) / (ab_ * ab_ * ab_));
            theta1_ = std::atan2(y - obj0.y, x - obj0.x);
            has_binary = true;
            has_point = false;
        } else {
            obj0 = {gm, x, y, z, vx, vy, vz, rs};
            has_point = true;
        }
        src_enrolled = true;
    }

    // ---- Point-mass gravity kicks: src/source_term/point_gravity.cpp:11-93 and
    //      binary_gravity.cpp:11-101 (energy with the old momenta, then the momenta) -----------
    void PointGravity(std::vector<Real>& U, Real h) {
        const Real gm = obj0.gm, x0 = obj0.x, y0 = obj0.y, z0 = obj0.z, rs = obj0.rs;
        const Real gmdt = gm * h;
        for (int k = kb; k < ke; k++) {
        const Real zk = nz > 1 ? zc[k] : zc0;
#pragma omp parallel for schedule(static)
        for (int j = jb; j < je; j++) {
            for (int i = ib; i < ie; i++) {
                const Real delx = xc[i] - x0, dely = yc[j] - y0, delz = zk - z0;
                const Real r2 = delx * delx + dely * dely + delz * delz;
                const Real rSmoInv = Real(1.0) / std::sqrt(r2 + rs * rs);
                const Real gVecdtdr = -gmdt * rSmoInv * rSmoInv * rSmoInv;
                const Real gxdt = gVecdtdr * delx, gydt = gVecdtdr * dely, gzdt = gVecdtdr * delz;
                Real& d = U[at(DEN, j, i, k)];
                Real& mx = U[at(MX, j, i, k)];
                Real& my = U[at(MY, j, i, k)];
                Real& mz = U[at(MZ, j, i, k)];
                Real& e = U[at(ENG, j, i, k)];
                e += mx * gxdt + my * gydt + mz * gzdt;
                mx += d * gxdt;
                my += d * gydt;
                mz += d * gzdt;
            }
        }
        }
    }
    void BinaryGravity(std::vector<Real>& U, Real tsrc, Real h) {
        const Real gm0dt = obj0.gm * h, rs0 = obj0.rs, gm1dt = obj1.gm * h, rs1 = obj1.rs;
        const Real theta = omega_ * tsrc + theta1_;
        obj1.x = ab_ * std::cos(theta);
        obj1.y = ab_ * std::sin(theta);
        const Real x0 = obj0.x, y0 = obj0.y, z0 = obj0.z, x1 = obj1.x, y1 = obj1.y, z1 = obj1.z;
        const Real gIndirdtdr1 = 0;   // the indirect term is computed and then zeroed (binary_gravity.cpp:44-45)
        for (int k = kb; k < ke; k++) {
        const Real zk = nz > 1 ? zc[k] : zc0;
#pragma omp parallel for schedule(static)
        for (int j = jb; j < je; j++) {
            for (int i = ib; i < ie; i++) {
                const Real delx0 = xc[i] - x0, dely0 = yc[j] - y0, delz0 = zk - z0;
                const Real delx1 = xc[i] - x1, dely1 = yc[j] - y1, delz1 = zk - z1;
                const Real r0Sqr = delx0 * delx0 + dely0 * dely0 + delz0 * delz0;
                const Real r0SmoInv = Real(1.0) / std::sqrt(r0Sqr + rs0 * rs0);
                const Real g0 = -gm0dt * r0SmoInv * r0SmoInv * r0SmoInv;
                const Real r1Sqr = delx1 * delx1 + dely1 * dely1 + delz1 * delz1;
                const Real r1SmoInv = Real(1.0) / std::sqrt(r1Sqr + rs1 * rs1);
                const Real g1 = -gm1dt * r1SmoInv * r1SmoInv * r1SmoInv;
                const Real gxdt = g0 * delx0 + g1 * delx1 + gIndirdtdr1 * x1;
                const Real gydt = g0 * dely0 + g1 * dely1 + gIndirdtdr1 * y1;
                const Real gzdt = g0 * delz0 + g1 * delz1 + gIndirdtdr1 * z1;
                Real& d = U[at(DEN, j, i, k)];
                Real& mx = U[at(MX, j, i, k)];
                Real& my = U[at(MY, j, i, k)];
                Real& mz = U[at(MZ, j, i, k)];
                Real& e = U[at(ENG, j, i, k)];
                e += mx * gxdt + my * gydt + mz * gzdt;
                mx += d * gxdt;
                my += d * gydt;
                mz += d * gzdt;
            }
        }
        }
    }
    // SourceTerm::UpdateSource: src/source_term/source.cpp:12-21
    void UpdateSource(std::vector<Real>& U, Real tsrc, Real h) {
        if (has_const) ConstGravity(U, h);
        if (has_point) PointGravity(U, h);
        else if (has_binary) BinaryGravity(U, tsrc, h);
    }

    // ---- Integrators: src/evolution/{forward_euler,rk2,rk3}.cpp ------------------------------
    void Stage(std::vector<Real>& in) {
        UpdateBD(in);
        ConsToPrim(in, prim);
        CalFlux();
    }
    void Integrate() {
        if (P.integrator == 1) {
            Stage(cons); UpdateCons(cons, consTemp, 1.0, 1.0, 0, false);
            cons.swap(consTemp);
        } else if (P.integrator == 3) {
            Stage(cons); UpdateCons(cons, consTemp, 1.0, 1.0, 0, false);
            Stage(consTemp); UpdateCons(cons, consTemp, 0.75, 0.25, 0.25, true);
            Stage(consTemp);
            Real frac13 = Real(1.0) / Real(3.0);
            UpdateCons(cons, consTemp, frac13, 2 * frac13, 2 * frac13, true);
            cons.swap(consTemp);
        } else {
            Stage(cons); UpdateCons(cons, consTemp, 1.0, 1.0, 0, false);
            Stage(consTemp); UpdateCons(cons, consTemp, 0.5, 0.5, 0.5, true);
            cons.swap(consTemp);
        }
    }

    // ---- Sim::Advance: src/sim.cpp:69-117 ---------------------------------------------------
    int Advance(Real dtoutput) {
        Real tNext = t + dtoutput;
        dtUntilOutput = dtoutput;
        int taken = 0;
        log_t.clear(); log_dt.clear();
        while (std::abs(tNext - t) > (Real)1e-12 && isContinue) {
            cmax = (Real)1e-16;
            CalCmax(cons, cmax);
            dt = cfl * drmin / cmax;
            dt = smin(dt, dtUntilOutput);
            // src/sim.cpp:85-91: half-kicks at t + dt/2 and t + dt
            if (src_enrolled) UpdateSource(cons, t + Real(0.5) * dt, Real(0.5) * dt);
            Integrate();
            if (src_enrolled) UpdateSource(cons, t + dt, Real(0.5) * dt);
            t += dt;
            step++;
            taken++;
            dtUntilOutput = tNext - t;
            if (P.stepmax > 0 && step >= P.stepmax) isContinue = false;
            log_t.push_back((double)t); log_dt.push_back((double)dt);
        }
        return taken;
    }
};

struct Holder {
    int prec;
    Oracle<double>* d = nullptr;
    Oracle<float>* f = nullptr;
};

}  // namespace

struct gko_sim { Holder h; };

extern "C" {

gko_sim* gko_create(const gko_params* p, int precision) {
    try {
        gko_sim* s = new gko_sim();
        s->h.prec = precision;
        if (precision == 32) s->h.f = new Oracle<float>(*p);
        else s->h.d = new Oracle<double>(*p);
        return s;
    } catch (...) {
        return nullptr;
    }
}

void gko_destroy(gko_sim* s) {
    if (!s) return;
    delete s->h.d; delete s->h.f; delete s;
}

void gko_set_cons(gko_sim* s, const void* c) {
    if (s->h.f) std::memcpy(s->h.f->cons.data(), c, s->h.f->cons.size() * sizeof(float));
    else std::memcpy(s->h.d->cons.data(), c, s->h.d->cons.size() * sizeof(double));
}
void gko_get_cons(const gko_sim* s, void* c) {
    if (s->h.f) std::memcpy(c, s->h.f->cons.data(), s->h.f->cons.size() * sizeof(float));
    else std::memcpy(c, s->h.d->cons.data(), s->h.d->cons.size() * sizeof(double));
}
int gko_advance(gko_sim* s, double dtoutput) {
    return s->h.f ? s->h.f->Advance((float)dtoutput) : s->h.d->Advance(dtoutput);
}
double gko_time(const gko_sim* s) { return s->h.f ? (double)s->h.f->t : s->h.d->t; }
int gko_step(const gko_sim* s) { return s->h.f ? s->h.f->step : s->h.d->step; }
int gko_is_continue(const gko_sim* s) { return s->h.f ? s->h.f->isContinue : s->h.d->isContinue; }
double gko_last_dt(const gko_sim* s) { return s->h.f ? (double)s->h.f->dt : s->h.d->dt; }
void gko_enroll_fixed(gko_sim* s, int fix_inner) {
    if (s->h.f) s->h.f->EnrollFixedBoundaryData(fix_inner != 0);
    else s->h.d->EnrollFixedBoundaryData(fix_inner != 0);
}
int gko_setup_immersed(gko_sim* s, double x, double y, double z, double r) {
    try {
        if (s->h.f) s->h.f->SetupImmersedShell((float)x, (float)y, (float)z, (float)r);
        else s->h.d->SetupImmersedShell(x, y, z, r);
    } catch (...) {
        return -1;
    }
    return 0;
}
int gko_update_bd(gko_sim* s) {
    try {
        if (s->h.f) s->h.f->UpdateBD(s->h.f->cons); else s->h.d->UpdateBD(s->h.d->cons);
    } catch (...) {
        return -1;
    }
    return 0;
}
int gko_dt_log(const gko_sim* s, double* t_out, double* dt_out, int cap) {
    const std::vector<double>& lt = s->h.f ? s->h.f->log_t : s->h.d->log_t;
    const std::vector<double>& ld = s->h.f ? s->h.f->log_dt : s->h.d->log_dt;
    int n = (int)std::min<size_t>(lt.size(), (size_t)cap);
    for (int k = 0; k < n; k++) { t_out[k] = lt[k]; dt_out[k] = ld[k]; }
    return (int)lt.size();
}

// ---- per-function entry points ----------------------------------------------------------------
static gko_params unit_params(double gamma) {
    gko_params p;
    std::memset(&p, 0, sizeof(p));
    p.nx = 1; p.ny = 1; p.ng = 2;
    p.xmax = p.ymax = p.zmax = 1.0;
    p.gamma = gamma; p.cfl = 0.4; p.integrator = 2; p.rc_order = 2;
    return p;
}

void gko_hllc(int n, const double* ul, const double* ur, int dir, double gamma, double* flux) {
    Oracle<double> o(unit_params(gamma));
    // rotation of src/flux/hllc.cpp:25-27
    int rot[NV] = {DEN, 0, 0, 0, PRE};
    for (int k = 0; k < 3; k++) rot[1 + k] = VX + (dir - VX + k) % 3;
    for (int i = 0; i < n; i++) {
        double wl[NV], wr[NV], f[NV];
        for (int q = 0; q < NV; q++) { wl[q] = ul[rot[q] * n + i]; wr[q] = ur[rot[q] * n + i]; }
        o.Hllc(wl, wr, f);
        for (int q = 0; q < NV; q++) flux[rot[q] * n + i] = f[q];
    }
}

void gko_plm_row(int m, const double* prim, double* ul, double* ur) {
    for (int n = 0; n < NV; n++) {
        for (int i = 0; i < m - 2; i++) {
            double u1 = prim[n * m + i], u2 = prim[n * m + i + 1], u3 = prim[n * m + i + 2];
            double s = Oracle<double>::VanLeer(u2 - u1, u3 - u2);
            ul[n * (m + 1) + i + 2] = u2 + 0.5 * s;
            ur[n * (m + 1) + i + 1] = u2 - 0.5 * s;
        }
    }
}

void gko_cons_to_prim(int n, const double* cons, double gamma, double* prim) {
    gko_params p = unit_params(gamma);
    Oracle<double> o(p);
    // run ConsToPrim on a 1-row view: reuse the state arrays with lenx = n
    o.lenx = n; o.leny = 1;
    o.cons.assign(cons, cons + (size_t)NV * n);
    o.prim.assign((size_t)NV * n, 0.0);
    o.ConsToPrim(o.cons, o.prim);
    std::memcpy(prim, o.prim.data(), sizeof(double) * NV * n);
}

void gko_prim_to_cons(int n, const double* prim, double gamma, double* cons) {
    // src/eos/adiabatic.cpp:162-198
    double gm1Rec = 1.0 / (gamma - 1.0);
    for (int i = 0; i < n; i++) {
        double d = prim[i], v1 = prim[n + i], v2 = prim[2 * n + i], v3 = prim[3 * n + i], p = prim[4 * n + i];
        cons[i] = d;
        cons[n + i] = d * v1;
        cons[2 * n + i] = d * v2;
        cons[3 * n + i] = d * v3;
        cons[4 * n + i] = p * gm1Rec + 0.5 * d * (v1 * v1 + v2 * v2 + v3 * v3);
    }
}

double gko_cmax(int n, const double* cons, double gamma, double cmax0) {
    gko_params p = unit_params(gamma);
    Oracle<double> o(p);
    o.lenx = n; o.leny = 1; o.ib = 0; o.ie = n; o.jb = 0; o.je = 1;
    o.cons.assign(cons, cons + (size_t)NV * n);
    double c = cmax0;
    o.CalCmax(o.cons, c);
    return c;
}

}  // extern "C"
