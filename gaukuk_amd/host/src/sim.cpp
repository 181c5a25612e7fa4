// sim.cpp -- host driver: construction from the configuration, device set-up and Advance.
#include "sim.hpp"

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <vector>

#include "utils/utils.hpp"

namespace Gaukuk {

namespace {

// the ghost layers of face f (0 x left .. 5 z right) over the active tangential range
void FaceGhostRange(const Grid& g, int f, int from[3], int to[3]) {
    const int axis = f / 2;
    const bool high = f % 2;
    const int lo[3] = {g.ib, g.jb, g.kb}, hi[3] = {g.ie, g.je, g.ke};
    for (int d = 0; d < 3; d++) { from[d] = lo[d]; to[d] = hi[d]; }
    from[axis] = high ? hi[axis] : 0;
    to[axis] = high ? hi[axis] + g.nGhost : lo[axis];
}

void CopyFace(const Grid& g, int f, const TArray<Real>& src, TArray<Real>& dst) {
    int from[3], to[3];
    FaceGhostRange(g, f, from, to);
    for (int v = 0; v < NVar; v++)
        for (int k = from[2]; k < to[2]; k++)
            for (int j = from[1]; j < to[1]; j++)
                for (int i = from[0]; i < to[0]; i++) dst(v, k, j, i) = src(v, k, j, i);
}

}  // namespace

Sim::Sim() : isContinue(true), domain(grid), step(0), t(0), dt(1e10) {
    Config& c = Config::getInstance();
    CFL = c.get("CFL");
    cons.NewArray(NVar, grid.lenz, grid.leny, grid.lenx);
    prim.NewArray(NVar, grid.lenz, grid.leny, grid.lenx);
    integratorType = (int)c.get("integrator", 2);
    rcOrder = (int)c.get("RcOrder", 1);
    stepMax = (int)c.get("stepmax", -1);
    if (integratorType == 1) hydroIntegrator_ = &Sim::ForwardEuler_;
    else if (integratorType == 3) hydroIntegrator_ = &Sim::RK3_;
    else hydroIntegrator_ = &Sim::RK2_;
}

Sim::~Sim() {
    if (ctx_) gk_destroy(ctx_);
}

void Sim::Check(gk_status s, const char* what) const {
    if (s != GK_OK) throw std::runtime_error(std::string(what) + ": " + gk_last_error(ctx_));
}

// A self-defined face runs on the device as a fixed face when the ghost values it writes do not
// depend on the state: it is applied to the current state and to one with every active value
// changed, and the ghost layers it fills (the face's nGhost layers over the active tangential
// range, which are all the stage reads) must agree bit for bit.  Returns whether they do, with the
// layers in `out`.  A face that depends on the state runs through the boundary hook instead.
bool Sim::ProbeSelfDefined(int f, TArray<Real>& out) const {
    TArray<Real> a = cons, b = cons;
    for (int v = 0; v < NVar; v++)
        for (int k = grid.kb; k < grid.ke; k++)
            for (int j = grid.jb; j < grid.je; j++)
                for (int i = grid.ib; i < grid.ie; i++) b(v, k, j, i) = a(v, k, j, i) * (Real)1.25 + (Real)0.5;
    const Boundary::BoundaryFunc& fn = boundary.Face(f);
    fn(a, grid, eos);
    fn(b, grid, eos);
    int from[3], to[3];
    FaceGhostRange(grid, f, from, to);
    for (int v = 0; v < NVar; v++)
        for (int k = from[2]; k < to[2]; k++)
            for (int j = from[1]; j < to[1]; j++)
                for (int i = from[0]; i < to[0]; i++)
                    if (std::memcmp(&a(v, k, j, i), &b(v, k, j, i), sizeof(Real)) != 0) return false;
    out = a;
    return true;
}

// gk_set_bd_hook's function: the enrolled face functions of the state-dependent self-defined
// faces, in the reference's UpdateBD order (src/boundary/boundary.cpp:127-142: x, y, then z), on
// the mirror array, whose active cells outside the mirrored bands hold NaN.
void Sim::BdHook(void* user, void*) {
    Sim* s = static_cast<Sim*>(user);
    for (int f = 0; f < 6; f++)
        if (s->hookFaces_ & (1 << f)) s->boundary.Face(f)(s->hookMirror_, s->grid, s->eos);
}

void Sim::EnsureDevice() {
    if (!ctx_) {
        gk_grid g{grid.nx, grid.ny, grid.nz, grid.nGhost,
                  (double)domain.xmin, (double)domain.xmax, (double)domain.ymin,
                  (double)domain.ymax, (double)domain.zmin, (double)domain.zmax};
        gk_phys p{};
        p.gamma = eos.GetGamma();
        p.cfl = CFL;
        p.g[0] = srcTerm.sourceEnrolled ? srcTerm.gx_ : 0;
        p.g[1] = srcTerm.sourceEnrolled ? srcTerm.gy_ : 0;
        p.g[2] = srcTerm.sourceEnrolled ? srcTerm.gz_ : 0;
        p.integrator = (this->*hydroIntegrator_)();
        p.rc_order = rcOrder == 1 ? 1 : 2;
        p.stepmax = stepMax;
        for (int f = 0; f < 6; f++) p.bc[f] = boundary.type[f];
        // state-independent self-defined faces become fixed faces holding the ghost values their
        // function writes; state-dependent ones are refreshed before every stage through the
        // boundary hook
        const int nface = grid.nz > 1 ? 6 : 4;
        for (int f = 0; f < nface; f++) {
            if (!boundary.selfDefined[f]) continue;
            TArray<Real> ghosts;
            if (!ProbeSelfDefined(f, ghosts)) {
                hookFaces_ |= 1 << f;
                p.bc[f] = 4;
                continue;
            }
            if (selfGhosts_.GetSize() == 0) selfGhosts_ = ghosts;
            CopyFace(grid, f, ghosts, selfGhosts_);
            selfFaces_[f] = true;
            p.bc[f] = 4;
        }
        p.precision = sizeof(Real) == 4 ? 32 : 64;
        gk_ctx* c = nullptr;
        gk_status s = gk_create(&g, &p, &c);
        if (s != GK_OK) throw std::runtime_error(std::string("gk_create: ") + gk_last_error(nullptr));
        ctx_ = c;
        for (int q = 0; q < srcTerm.nPoint; q++) {
            const Real* m = srcTerm.mass[q];
            Check(gk_set_point_gravity(ctx_, m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7]), "gk_set_point_gravity");
        }
        if (hookFaces_) {
            hookMirror_.NewArray(NVar, grid.lenz, grid.leny, grid.lenx);
            for (size_t k = 0; k < hookMirror_.GetSize(); k++) hookMirror_.data()[k] = std::numeric_limits<Real>::quiet_NaN();
            Check(gk_set_bd_hook(ctx_, hookFaces_, grid.nGhost, &Sim::BdHook, this, hookMirror_.data()), "gk_set_bd_hook");
        }
    }
    // fixed faces and the inner region: the enrolled snapshot (re-sent when the setup enrolls new
    // data), with the layers of state-independent self-defined faces written in
    bool anySelf = false;
    for (bool s : selfFaces_) anySelf |= s;
    const bool wantFixed = anySelf || (boundary.FixedEnrolled() && (boundary.AnyFixed() || boundary.FixedInner()));
    if (wantFixed && fixedSent_ != boundary.FixedVersion()) {
        TArray<Real> snap = boundary.FixedEnrolled() ? boundary.FixedData() : cons;
        for (int f = 0; f < 6; f++)
            if (selfFaces_[f]) CopyFace(grid, f, selfGhosts_, snap);
        Check(gk_set_fixed_boundary(ctx_, snap.data(), boundary.FixedEnrolled() && boundary.FixedInner() ? 1 : 0),
              "gk_set_fixed_boundary");
        fixedSent_ = boundary.FixedVersion();
    }
    if (boundary.HasShell() && !shellSent_) {
        const double* s = boundary.Shell();
        Check(gk_set_immersed_shell(ctx_, s[0], s[1], s[2], s[3]), "gk_set_immersed_shell");
        shellSent_ = true;
    }
    if (hostDirty_) {
        Check(gk_upload_cons(ctx_, cons.data()), "gk_upload_cons");
        hostDirty_ = false;
    }
    if (timePending_) {
        Check(gk_set_time(ctx_, t, step), "gk_set_time");
        timePending_ = false;
    }
    // the in-place part of the last boundary update (immersed shell, inner region)
    if (boundary.DeviceUpdatePending()) {
        Check(gk_update_bd(ctx_), "gk_update_bd");
        boundary.ClearDeviceUpdate();
        hostStale_ = true;
    }
}

int Sim::Restart(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open restart file " + path);
    int realSize = 0, hdr[5] = {0, 0, 0, 0, 0};
    Real tFile = 0;
    in.read(reinterpret_cast<char*>(&realSize), sizeof(int));
    in.read(reinterpret_cast<char*>(&tFile), sizeof(Real));
    in.read(reinterpret_cast<char*>(hdr), sizeof hdr);
    if (!in || realSize != (int)sizeof(Real) || hdr[0] != NVar || hdr[1] != grid.nx || hdr[2] != grid.ny ||
        hdr[3] != grid.nz || hdr[4] != grid.nGhost)
        throw std::runtime_error("restart file " + path + " does not match this grid and precision");
    in.seekg((std::streamoff)(domain.xc.GetSizeInBytes() + domain.yc.GetSizeInBytes() + domain.zc.GetSizeInBytes()),
             std::ios::cur);
    int size = 0;
    in.read(reinterpret_cast<char*>(&size), sizeof(int));
    if (!in || size != (int)cons.GetSize()) throw std::runtime_error("restart file " + path + ": bad data size");
    in.read(reinterpret_cast<char*>(cons.data()), cons.GetSizeInBytes());
    if (!in) throw std::runtime_error("restart file " + path + ": short data");
    t = tFile;
    step = 0;
    hostDirty_ = true;
    hostStale_ = false;
    timePending_ = true;
    const size_t cut = path.find_last_of('_');
    return cut == std::string::npos ? 0 : std::atoi(path.c_str() + cut + 1);
}

// Sim::Advance (src/sim.cpp:69-117): the whole loop (dt from the CFL condition, source
// half-kicks, integrator, t/step update, stepmax) runs on the device; the per-step lines are
// printed afterwards from the device step log.
void Sim::Advance(Real dtoutput) {
    EnsureDevice();
    const int before = step;
    double tt = 0;
    int st = 0;
    Check(gk_advance(ctx_, dtoutput, &tt, &st), "gk_advance");
    const int n = st - before;
    if (verbose_ && n > 0) {
        std::vector<int> s(n);
        std::vector<double> ts(n), dts(n);
        int got = 0;
        Check(gk_step_log(ctx_, s.data(), ts.data(), dts.data(), n, &got), "gk_step_log");
        for (int k = 0; k < got; k++)
            std::cout << "step = " << s[k] << ", t = " << (Real)ts[k] << ", dt = " << (Real)dts[k] << "\n";
        std::cout.flush();
    }
    double dd = 0;
    int ic = 1;
    Check(gk_get_time(ctx_, nullptr, &dd, nullptr, &ic), "gk_get_time");
    t = tt;
    dt = dd;
    step = st;
    isContinue = ic != 0;
    hostStale_ = true;
}

}  // namespace Gaukuk
