// sim.cpp -- host driver: construction from the configuration, device set-up and Advance.
#include "sim.hpp"

#include <iostream>
#include <stdexcept>
#include <vector>

#include "utils/utils.hpp"

namespace Gaukuk {

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

void Sim::EnsureDevice() {
    if (boundary.AnySelfDefined())
        throw std::runtime_error("self-defined boundaries have no device implementation in this build");
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
        p.precision = sizeof(Real) == 4 ? 32 : 64;
        gk_ctx* c = nullptr;
        gk_status s = gk_create(&g, &p, &c);
        if (s != GK_OK) throw std::runtime_error(std::string("gk_create: ") + gk_last_error(nullptr));
        ctx_ = c;
        for (int q = 0; q < srcTerm.nPoint; q++) {
            const Real* m = srcTerm.mass[q];
            Check(gk_set_point_gravity(ctx_, m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7]), "gk_set_point_gravity");
        }
    }
    // fixed faces: the enrolled snapshot (re-sent when the setup enrolls new data)
    if (boundary.AnyFixed() && boundary.FixedEnrolled() && fixedSent_ != boundary.FixedVersion()) {
        Check(gk_set_fixed_boundary(ctx_, boundary.FixedData().data()), "gk_set_fixed_boundary");
        fixedSent_ = boundary.FixedVersion();
    }
    if (hostDirty_) {
        Check(gk_upload_cons(ctx_, cons.data()), "gk_upload_cons");
        hostDirty_ = false;
    }
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
