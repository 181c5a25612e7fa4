// gk_physics.hpp -- compile-time physics policies for the MI355X stage kernels.
//
// The reference selects its equation of state and Riemann solver by linking one source file
// each (Makefile:64-90, configure.py:8-19: EOS=adiabatic, FLUX=hllc).  Here the same choice is
// a template argument of the stage kernel: `--eos` / `--flux` at configure time set GK_EOS /
// GK_FLUX, which name the policy types below (gk_select.hpp).  Every policy is
// __host__ __device__ so host code and kernels share one definition.
//
// Arithmetic follows the reference's operation order so that the FP64 path differs from the
// reference only by FMA contraction and reassociation-free rounding (parity target 1e-12
// relative L1).  Comparisons use std::min/std::max semantics (second operand returned only when
// strictly better), not fmin/fmax, to keep the reference's signed-zero and NaN behaviour.
#pragma once

#include <hip/hip_runtime.h>

namespace gk {

template <class R> __host__ __device__ __forceinline__ R smin(R a, R b) { return (b < a) ? b : a; }
template <class R> __host__ __device__ __forceinline__ R smax(R a, R b) { return (a < b) ? b : a; }

template <class R> __host__ __device__ __forceinline__ R gsqrt(R x);
template <> __host__ __device__ __forceinline__ double gsqrt<double>(double x) { return __builtin_sqrt(x); }
template <> __host__ __device__ __forceinline__ float gsqrt<float>(float x) { return __builtin_sqrtf(x); }

template <class R> __host__ __device__ __forceinline__ R gabs(R x) { return x < R(0) ? -x : x; }

// Density and pressure floors: GAUKUK_DENSITY_FLOOR / GAUKUK_PRESSURE_FLOOR (src/gaukuk.hpp:11-21).
#ifndef GK_DENSITY_FLOOR
#define GK_DENSITY_FLOOR 1e-16
#endif
#ifndef GK_PRESSURE_FLOOR
#define GK_PRESSURE_FLOOR 1e-16
#endif

// ---------------------------------------------------------------------------------------------
// Adiabatic (ideal-gas) equation of state: src/eos/eos.hpp:13-36, src/eos/adiabatic.cpp:12-198.
template <class R>
struct Adiabatic {
    R gamma, gm1, gm1Rec, gmRec;

    static Adiabatic make(double g) {
        Adiabatic e;
        e.gamma = R(g);
        e.gm1 = e.gamma - R(1.0);
        e.gm1Rec = R(1.0) / (e.gamma - R(1.0));
        e.gmRec = R(1.0) / e.gamma;
        return e;
    }
    __host__ __device__ __forceinline__ R sound_speed(R d, R p) const { return gsqrt(gamma * p / d); }
    __host__ __device__ __forceinline__ R egas(R, R p) const { return p * gm1Rec; }

    // ConsToPrim with the floors clamped on local copies: adiabatic.cpp:18-61.
    __host__ __device__ __forceinline__ void cons_to_prim(const R u[5], R w[5]) const {
        R d = smax(u[0], R(GK_DENSITY_FLOOR));
        R inv = R(1.0) / d;
        w[0] = d;
        w[1] = u[1] * inv;
        w[2] = u[2] * inv;
        w[3] = u[3] * inv;
        R ek = R(0.5) * inv * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]);
        R p = gm1 * (u[4] - ek);
        w[4] = (p > R(GK_PRESSURE_FLOOR)) ? p : R(GK_PRESSURE_FLOOR);
    }

    // One cell's max_d |v_d| + c_s with CalCmax's own kinetic-energy form: adiabatic.cpp:115-160.
    __host__ __device__ __forceinline__ R wave_speed(const R u[5]) const {
        R d = smax(u[0], R(GK_DENSITY_FLOOR));
        R inv = R(1.0) / d;
        R v1 = u[1] * inv, v2 = u[2] * inv, v3 = u[3] * inv;
        R ek = R(0.5) * (v1 * u[1] + v2 * u[2] + v3 * u[3]);
        R p = gm1 * (u[4] - ek);
        p = (p > R(GK_PRESSURE_FLOOR)) ? p : R(GK_PRESSURE_FLOOR);
        R cs = gsqrt(gamma * p * inv);
        R loc = gabs(v1) + cs;
        loc = smax(loc, gabs(v2) + cs);
        loc = smax(loc, gabs(v3) + cs);
        return loc;
    }

    // Post-update floors of UpdateCons: src/evolution/update_cons.cpp:55-67.
    __host__ __device__ __forceinline__ void floors(R r[5]) const {
        const R dmin = R(GK_DENSITY_FLOOR), pmin = R(GK_PRESSURE_FLOOR);
        if (r[0] < dmin) {
            r[0] = dmin; r[1] = R(0.0); r[2] = R(0.0); r[3] = R(0.0);
            r[4] = pmin / gm1 + R(0.0);
        } else {
            R inv = R(1.0) / r[0];
            R ke = R(0.5) * inv * (r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
            R press = gm1 * (r[4] - ke);
            if (press < pmin) r[4] = pmin / gm1 + ke;
        }
    }
};

// ---------------------------------------------------------------------------------------------
// Slope limiters.  PLM with van Leer: src/reconstruction/reconstruction.hpp:36-42 (the quotient
// is evaluated and then discarded by the select when a*b <= 0).  First order = zero slope, which
// reproduces recon_first_order.cpp exactly (u +/- 0.5*0 == u).
struct VanLeer {
    static constexpr int order = 2;
    template <class R> __host__ __device__ __forceinline__ static R slope(R a, R b) {
        R multi = a * b;
        R sigma = R(2.0) * multi / (a + b);
        return (multi <= R(0.0)) ? R(0.0) : sigma;
    }
};
struct FirstOrder {
    static constexpr int order = 1;
    template <class R> __host__ __device__ __forceinline__ static R slope(R, R) { return R(0.0); }
};

// ---------------------------------------------------------------------------------------------
// HLLC Riemann solver (Toro ch. 10): src/flux/hllc.cpp:22-103.
// wl, wr: primitive states rotated so index 1 is the face-normal velocity and 2, 3 the
// tangential ones in the reference's cyclic order (hllc.cpp:25-27); f receives the flux in the
// same rotated order.  The left/right blend is a multiply, as in the reference, so a NaN on the
// unselected side propagates.
struct Hllc {
    template <class R, class EOS>
    __host__ __device__ __forceinline__ static void flux(const EOS& eos, const R wl[5], const R wr[5], R f[5]) {
        const R denl = wl[0], vxl = wl[1], vyl = wl[2], vzl = wl[3], prel = wl[4];
        const R denr = wr[0], vxr = wr[1], vyr = wr[2], vzr = wr[3], prer = wr[4];
        const R al = eos.sound_speed(denl, prel);
        const R ar = eos.sound_speed(denr, prer);
        R pv = R(0.5) * (prel + prer) - R(0.125) * (vxr - vxl) * (denl + denr) * (al + ar);
        pv = smax(pv, R(0));
        const R qk = R(0.5) * (R(1) + eos.gmRec);
        R tmp = smax(pv / prel - R(1), R(0));
        const R ql = gsqrt(R(1) + qk * tmp);
        tmp = smax(pv / prer - R(1), R(0));
        const R qr = gsqrt(R(1) + qk * tmp);
        const R sl = vxl - al * ql;
        const R sr = vxr + ar * qr;
        const R ss = (prer - prel + denl * vxl * (sl - vxl) - denr * vxr * (sr - vxr)) /
                     (denl * (sl - vxl) - denr * (sr - vxr));
        const R el = eos.egas(denl, prel) + R(0.5) * denl * (vxl * vxl + vyl * vyl + vzl * vzl);
        const R er = eos.egas(denr, prer) + R(0.5) * denr * (vxr * vxr + vyr * vyr + vzr * vzr);
        const R cl = smin(R(0), sl) * (ss - vxl) / (sl - ss);
        const R cr = smax(R(0), sr) * (ss - vxr) / (sr - ss);
        const R sel = (ss >= R(0)) ? R(1.0) : R(0.0);
        const R ser = R(1.0) - sel;
        const R fl1 = denl * (vxl + cl);
        const R fl2 = denl * (vxl * vxl + cl * sl) + prel;
        const R fl5 = (el + prel) * (vxl + cl) + denl * cl * ss * (sl - vxl);
        const R fr1 = denr * (vxr + cr);
        const R fr2 = denr * (vxr * vxr + cr * sr) + prer;
        const R fr5 = (er + prer) * (vxr + cr) + denr * cr * ss * (sr - vxr);
        f[0] = sel * fl1 + ser * fr1;
        f[1] = sel * fl2 + ser * fr2;
        f[2] = sel * (fl1 * vyl) + ser * (fr1 * vyr);
        f[3] = sel * (fl1 * vzl) + ser * (fr1 * vzr);
        f[4] = sel * fl5 + ser * fr5;
    }
};

// ---------------------------------------------------------------------------------------------
// Boundary transforms applied to a ghost cell's source value (types 2, 3; types 0 and 1 are
// pure index maps).  comp = 1 for x faces (normal momentum MTX), 2 for y faces (MTY).
//   kind 1: reflective, negate the normal momentum       src/boundary/reflective.cpp:29-44
//   kind 2: outflow low face, clamp m_n to <= 0, fix E    src/boundary/outflow.cpp:36-58
//   kind 3: outflow high face, clamp m_n to >= 0, fix E   src/boundary/outflow.cpp:85-107
template <class R>
__host__ __device__ __forceinline__ void bc_transform(R u[5], int comp, int kind) {
    if (kind == 1) {
        u[comp] = -u[comp];
    } else if (kind >= 2) {
        const R m = u[comp];
        const R mc = (kind == 2) ? smin(m, R(0.0)) : smax(m, R(0.0));
        const R d = smax(u[0], R(GK_DENSITY_FLOOR));
        const R inv = R(1.0) / d;
        const R keOld = R(0.5) * inv * m * m;
        const R keNew = R(0.5) * inv * mc * mc;
        R eint = u[4] - keOld;
        eint = smax(eint, R(1e-16));
        u[0] = d;
        u[comp] = mc;
        u[4] = eint + keNew;
    }
}

// Ghost index map for one axis of length n with face types (lo, hi): returns the source index
// and sets kind (0 none, 1 reflect, 2 outflow-low, 3 outflow-high).  Unknown codes behave as
// type 0, like the reference's registration fallback (src/boundary/boundary.cpp:27-29).
__host__ __device__ __forceinline__ int bc_remap(int k, int n, int lo, int hi, int& kind) {
    kind = 0;
    if (k < 0) {
        if (lo == 1) return k + n;
        if (lo == 2) { kind = 1; return -1 - k; }
        if (lo == 3) { kind = 2; return -1 - k; }
        return 0;
    }
    if (k >= n) {
        if (hi == 1) return k - n;
        if (hi == 2) { kind = 1; return 2 * n - 1 - k; }
        if (hi == 3) { kind = 3; return 2 * n - 1 - k; }
        return n - 1;
    }
    return k;
}

// Constant-gravity half kick on one cell, axis by axis, energy first with the old momentum:
// src/source_term/const_gravity.cpp:17-66.
template <class R>
__host__ __device__ __forceinline__ void gravity_kick(R u[5], R gx, R gy, R gz, R h) {
    if (gx != R(0)) { u[4] += u[1] * gx * h; u[1] += u[0] * gx * h; }
    if (gy != R(0)) { u[4] += u[2] * gy * h; u[2] += u[0] * gy * h; }
    if (gz != R(0)) { u[4] += u[3] * gz * h; u[3] += u[0] * gz * h; }
}

}  // namespace gk
