// gk_physics.hpp -- compile-time physics policies for the MI355X stage kernels.
//
// The reference selects its equation of state and Riemann solver by linking one source file
// each (Makefile:64-90, configure.py:8-19: EOS=adiabatic, FLUX=hllc).  Here the same choice is
// a template argument of the stage kernel: `--eos` / `--flux` at configure time set GK_EOS /
// GK_FLUX, which name the policy types below (gk_device.hpp).  Every policy is
// __host__ __device__ so host code and kernels share one definition.
//
// Arithmetic follows the reference's operation order so that the FP64 path differs from the
// reference only by FMA contraction and by the few-ulp quotient and root cores below (parity
// target 1e-12 relative L1).  Comparisons use std::min/std::max semantics (second operand returned only when
// strictly better), not fmin/fmax, to keep the reference's signed-zero and NaN behaviour, except
// for maxnum() below, used only where the two agree for every result that matters.
#pragma once

#include <hip/hip_runtime.h>

namespace gk {

template <class R> __host__ __device__ __forceinline__ R smin(R a, R b) { return (b < a) ? b : a; }
template <class R> __host__ __device__ __forceinline__ R smax(R a, R b) { return (a < b) ? b : a; }
// FP32 on the device (judged by L1 against FP64, not bit for bit): one v_min_f32 / v_max_f32
// instead of a compare and a select; the results differ from std::min/std::max only for a NaN
// operand and in the sign of a zero result.
template <> __host__ __device__ __forceinline__ float smin<float>(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_fminf(a, b);
#else
    return (b < a) ? b : a;
#endif
}
template <> __host__ __device__ __forceinline__ float smax<float>(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_fmaxf(a, b);
#else
    return (a < b) ? b : a;
#endif
}

// max(a, b) as the hardware's IEEE maxNum (v_max_f64 / v_max_f32, one instruction instead of a
// compare and two selects).  It equals smax(a, b) for all operands that are not NaN, except that
// the sign of a zero result may differ; used where a NaN operand makes the result NaN through
// another path anyway, and no zero sign reaches the result.  The host build uses smax.
template <class R> __host__ __device__ __forceinline__ R maxnum(R a, R b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_fmax(a, b);
#else
    return smax(a, b);
#endif
}
template <> __host__ __device__ __forceinline__ float maxnum<float>(float a, float b) { return smax(a, b); }

// Quotients and roots of the stencil arithmetic, for operands in the normal range (floored
// densities and pressures, sums of them).  The IEEE sequences the compiler emits wrap a core of
// hardware reciprocal (reciprocal square root) plus refinement steps in range-scaling and
// special-value steps (div_scale/div_fixup, ldexp and class tests), which are the identity for
// such operands.  The cores below keep one refinement: FP64 v_rcp_f64 / v_rsq_f64 are good to
// about 2^-25, and one residual correction (quotient), one Newton step (reciprocal) or one
// Goldschmidt step (roots) takes that to a few ulp.  Largest errors against the IEEE operations
// measured on the device over 3 x 64M draws (exponents within +-4, +-30 and +-500): FP64 quotient
// 19 ulp, reciprocal 11, root 37, sqrt(a / b) 38, 1 / sqrt 20; FP32 (the hardware v_rcp_f32,
// v_sqrt_f32, v_rsq_f32) 1-2 ulp.  gk_selftest() counts device results beyond twice those bounds.
// The FP64 path is held to 1e-12 relative L1 against the reference per conserved variable (KH
// 2048^2 after 200 RK3 steps: 1.2e-13 with these cores, 3.3e-14 with correctly rounded ones, for
// 6% more VALU instructions per stage); FP32 is judged by its L1 distance from the FP64 run.  The
// host build uses the plain operators.
__host__ __device__ __forceinline__ double srcp(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double r = __builtin_amdgcn_rcp(d);
    return __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
#else
    return 1.0 / d;
#endif
}
__host__ __device__ __forceinline__ double sdiv(double n, double d) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double r = __builtin_amdgcn_rcp(d);
    const double q = n * r;
    return __builtin_fma(__builtin_fma(-d, q, n), r, q);
#else
    return n / d;
#endif
}
// Goldschmidt on y = rsq(x): g = x y ~ sqrt(x), h = y / 2 ~ 1 / (2 sqrt(x)), r = 1/2 - g h;
// sqrt(x) ~ g + g r and 1 / sqrt(x) ~ 2 (h + h r) = y + y r.
__host__ __device__ __forceinline__ double ssqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    const double g = x * y, h = 0.5 * y;
    return __builtin_fma(g, __builtin_fma(-h, g, 0.5), g);
#else
    return __builtin_sqrt(x);
#endif
}
__host__ __device__ __forceinline__ float srcp(float d) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(d);
#else
    return 1.0f / d;
#endif
}
__host__ __device__ __forceinline__ float sdiv(float n, float d) { return n * srcp(d); }
__host__ __device__ __forceinline__ float ssqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return __builtin_sqrtf(x);
#endif
}
// sqrt(a / b) for positive normal a, b (the sound speed): a / sqrt(a b), one reciprocal square
// root instead of a quotient and a root.
__host__ __device__ __forceinline__ double sqrt_ratio(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double z = a * b, y = __builtin_amdgcn_rsq(z);
    const double g = z * y, h = 0.5 * y;
    return a * __builtin_fma(y, __builtin_fma(-h, g, 0.5), y);
#else
    return __builtin_sqrt(a / b);
#endif
}
__host__ __device__ __forceinline__ float sqrt_ratio(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return a * __builtin_amdgcn_rsqf(a * b);
#else
    return __builtin_sqrtf(a / b);
#endif
}
__host__ __device__ __forceinline__ double srsqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    const double g = x * y, h = 0.5 * y;
    return __builtin_fma(y, __builtin_fma(-h, g, 0.5), y);
#else
    return 1.0 / __builtin_sqrt(x);
#endif
}
__host__ __device__ __forceinline__ float srsqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / __builtin_sqrtf(x);
#endif
}

template <class R> __host__ __device__ __forceinline__ R gabs(R x) { return x < R(0) ? -x : x; }
template <> __host__ __device__ __forceinline__ float gabs<float>(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_fabsf(x);
#else
    return x < 0.0f ? -x : x;
#endif
}

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
    __host__ __device__ __forceinline__ R sound_speed(R d, R p) const { return sqrt_ratio(gamma * p, d); }
    __host__ __device__ __forceinline__ R egas(R, R p) const { return p * gm1Rec; }

    // ConsToPrim with the floors clamped on local copies: adiabatic.cpp:18-61.
    __host__ __device__ __forceinline__ void cons_to_prim(const R u[5], R w[5]) const {
        R d = smax(u[0], R(GK_DENSITY_FLOOR));
        R inv = srcp(d);
        w[0] = d;
        w[1] = u[1] * inv;
        w[2] = u[2] * inv;
        w[3] = u[3] * inv;
        R ek = R(0.5) * inv * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]);
        R p = gm1 * (u[4] - ek);
        // (p > floor) ? p : floor, which is maxNum(p, floor) for every p, NaN and -0 included
        w[4] = maxnum(p, R(GK_PRESSURE_FLOOR));
    }

    // One cell's max_d |v_d| + c_s with CalCmax's own kinetic-energy form: adiabatic.cpp:115-160.
    __host__ __device__ __forceinline__ R wave_speed(const R u[5]) const {
        return wave_speed(u, srcp(smax(u[0], R(GK_DENSITY_FLOOR))));
    }
    // The same with inv = 1 / max(u[0], floor) already formed (floors() returns it).
    __host__ __device__ __forceinline__ R wave_speed(const R u[5], R inv) const {
        R v1 = u[1] * inv, v2 = u[2] * inv, v3 = u[3] * inv;
        R ek = R(0.5) * (v1 * u[1] + v2 * u[2] + v3 * u[3]);
        R p = gm1 * (u[4] - ek);
        p = (p > R(GK_PRESSURE_FLOOR)) ? p : R(GK_PRESSURE_FLOOR);
        R cs = ssqrt(gamma * p * inv);
        R loc = gabs(v1) + cs;
        loc = smax(loc, gabs(v2) + cs);
        loc = smax(loc, gabs(v3) + cs);
        return loc;
    }

    // Post-update floors of UpdateCons: src/evolution/update_cons.cpp:55-67, written without
    // branches (both outcomes are computed and selected; the quotient 1/r[0] of the low-density
    // case is discarded).  pm_gm1 = pmin/gm1 is the same quotient the reference forms in place.
    // Returns 1 / max(r[0], floor) of the floored state, the quotient wave_speed forms (the
    // source kicks after the floors leave the density unchanged).
    __host__ __device__ __forceinline__ R floors(R r[5]) const {
        const R dmin = R(GK_DENSITY_FLOOR), pmin = R(GK_PRESSURE_FLOOR);
        const R pm_gm1 = pmin / gm1;
        const bool low_d = r[0] < dmin;
        const R inv = srcp(r[0]);
        const R ke = R(0.5) * inv * (r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
        const R press = gm1 * (r[4] - ke);
        const R e_fix = (press < pmin) ? pm_gm1 + ke : r[4];
        r[4] = low_d ? pm_gm1 + R(0.0) : e_fix;
        r[1] = low_d ? R(0.0) : r[1];
        r[2] = low_d ? R(0.0) : r[2];
        r[3] = low_d ? R(0.0) : r[3];
        r[0] = low_d ? dmin : r[0];
        return low_d ? R(1.0) / dmin : inv;
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
        R sigma = sdiv(R(2.0) * multi, a + b);
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
// same rotated order.
struct Hllc {
    template <class R, class EOS>
    __host__ __device__ __forceinline__ static void flux(const EOS& eos, const R wl[5], const R wr[5], R f[5]) {
        const R denl = wl[0], vxl = wl[1], vyl = wl[2], vzl = wl[3], prel = wl[4];
        const R denr = wr[0], vxr = wr[1], vyr = wr[2], vzr = wr[3], prer = wr[4];
        const R al = eos.sound_speed(denl, prel);
        const R ar = eos.sound_speed(denr, prer);
        // maxnum: pv and tmp are NaN only for a NaN state value, which also makes S* below NaN and
        // with it every flux component; the sign of a zero pv only reaches tmp = pv/p - 1 = -1
        R pv = R(0.5) * (prel + prer) - R(0.125) * (vxr - vxl) * (denl + denr) * (al + ar);
        pv = maxnum(pv, R(0));
        const R qk = R(0.5) * (R(1) + eos.gmRec);
        R tmp = maxnum(sdiv(pv, prel) - R(1), R(0));
        const R ql = ssqrt(R(1) + qk * tmp);
        tmp = maxnum(sdiv(pv, prer) - R(1), R(0));
        const R qr = ssqrt(R(1) + qk * tmp);
        const R sl = vxl - al * ql;
        const R sr = vxr + ar * qr;
        const R ss = sdiv(prer - prel + denl * vxl * (sl - vxl) - denr * vxr * (sr - vxr),
                          denl * (sl - vxl) - denr * (sr - vxr));
        // Only the upwind side of the contact enters the flux: the reference blends
        // F = sel*F_L + (1-sel)*F_R with sel in {0,1} (hllc.cpp:64-83), which equals the
        // selected side's flux whenever the other side is finite.
        const bool left = ss >= R(0);
        const R dk = left ? denl : denr, vk = left ? vxl : vxr, v2 = left ? vyl : vyr;
        const R v3 = left ? vzl : vzr, pk = left ? prel : prer, sk = left ? sl : sr;
        // left: smin(0, sl) = (sl < 0 ? sl : 0); right: smax(0, sr) = (0 < sr ? sr : 0)
        // (bitwise logic: a short-circuit form becomes a branch)
        const bool take = (left & (sk < R(0))) | ((!left) & (R(0) < sk));
        const R mk = take ? sk : R(0);
        const R ek = eos.egas(dk, pk) + R(0.5) * dk * (vk * vk + v2 * v2 + v3 * v3);
        const R ck = sdiv(mk * (ss - vk), sk - ss);
        const R f1 = dk * (vk + ck);
        f[0] = f1;
        f[1] = dk * (vk * vk + ck * sk) + pk;
        f[2] = f1 * v2;
        f[3] = f1 * v3;
        f[4] = (ek + pk) * (vk + ck) + dk * ck * ss * (sk - vk);
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
    } else if (kind == 2 || kind == 3) {
        const R m = u[comp];
        const R mc = (kind == 2) ? smin(m, R(0.0)) : smax(m, R(0.0));
        const R d = smax(u[0], R(GK_DENSITY_FLOOR));
        const R inv = srcp(d);
        const R keOld = R(0.5) * inv * m * m;
        const R keNew = R(0.5) * inv * mc * mc;
        R eint = u[4] - keOld;
        eint = smax(eint, R(1e-16));
        u[0] = d;
        u[comp] = mc;
        u[4] = eint + keNew;
    }
}

// The same transforms without branches, for a kind that differs between lanes of a wave
// (ghost columns): every outcome is computed and the right one selected.
template <class R>
__host__ __device__ __forceinline__ void bc_transform_lanes(R u[5], int comp, int kind) {
    const R m = u[comp];
    const R mc = (kind == 2) ? smin(m, R(0.0)) : smax(m, R(0.0));
    const R d = smax(u[0], R(GK_DENSITY_FLOOR));
    const R inv = srcp(d);
    const R keOld = R(0.5) * inv * m * m;
    const R keNew = R(0.5) * inv * mc * mc;
    R eint = u[4] - keOld;
    eint = smax(eint, R(1e-16));
    const bool outflow = (kind == 2) | (kind == 3);
    u[4] = outflow ? eint + keNew : u[4];
    u[comp] = outflow ? mc : ((kind == 1) ? -m : m);
    u[0] = outflow ? d : u[0];
}

// Ghost index map for one axis of length n with face types (lo, hi): returns the source index
// and sets kind (0 none, 1 reflect, 2 outflow-low, 3 outflow-high, 4 fixed).  For a fixed face
// (type 4, src/boundary/fixed.cpp) the return value is the ghost's distance from the face (0 for
// the nearest layer), an index into that face's saved layers.  Unknown codes behave as type 0,
// like the reference's registration fallback (src/boundary/boundary.cpp:27-29).
__host__ __device__ __forceinline__ int bc_remap(int k, int n, int lo, int hi, int& kind) {
    kind = 0;
    // active index: one unsigned compare, the face-type logic stays off the common path
    if (__builtin_expect((unsigned)k < (unsigned)n, 1)) return k;
    if (k < 0) {
        if (lo == 1) return k + n;
        if (lo == 2) { kind = 1; return -1 - k; }
        if (lo == 3) { kind = 2; return -1 - k; }
        if (lo == 4) { kind = 4; return -1 - k; }
        return 0;
    }
    if (k >= n) {
        if (hi == 1) return k - n;
        if (hi == 2) { kind = 1; return 2 * n - 1 - k; }
        if (hi == 3) { kind = 3; return 2 * n - 1 - k; }
        if (hi == 4) { kind = 4; return k - n; }
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

// Point-mass gravity half-kick on one cell at (x, y, z): src/source_term/point_gravity.cpp:20-74
// (one mass) and binary_gravity.cpp:29-96 (two masses, the second at (x1, y1, z1); the indirect
// term is computed and then zeroed there, so it contributes 0 * x1).  Energy takes the three old
// momenta, then the momenta are kicked.  m = {gm, x, y, z, rs}.
template <class R>
__host__ __device__ __forceinline__ void point_kick(R u[5], int npoint, const double m0[5], const double m1[5],
                                                    R x1, R y1, R h, R x, R y, R z) {
    R gxdt, gydt, gzdt;
    const R gm0dt = R(m0[0]) * h, rs0 = R(m0[4]);
    const R delx0 = x - R(m0[1]), dely0 = y - R(m0[2]), delz0 = z - R(m0[3]);
    const R r0Sqr = delx0 * delx0 + dely0 * dely0 + delz0 * delz0;
    const R r0SmoInv = srsqrt(r0Sqr + rs0 * rs0);
    const R g0 = -gm0dt * r0SmoInv * r0SmoInv * r0SmoInv;
    if (npoint == 1) {
        gxdt = g0 * delx0;
        gydt = g0 * dely0;
        gzdt = g0 * delz0;
    } else {
        const R gm1dt = R(m1[0]) * h, rs1 = R(m1[4]), z1 = R(m1[3]);
        const R delx1 = x - x1, dely1 = y - y1, delz1 = z - z1;
        const R r1Sqr = delx1 * delx1 + dely1 * dely1 + delz1 * delz1;
        const R r1SmoInv = srsqrt(r1Sqr + rs1 * rs1);
        const R g1 = -gm1dt * r1SmoInv * r1SmoInv * r1SmoInv;
        const R gIndir = R(0);
        gxdt = g0 * delx0 + g1 * delx1 + gIndir * x1;
        gydt = g0 * dely0 + g1 * dely1 + gIndir * y1;
        gzdt = g0 * delz0 + g1 * delz1 + gIndir * z1;
    }
    u[4] += u[1] * gxdt + u[2] * gydt + u[3] * gzdt;
    u[1] += u[0] * gxdt;
    u[2] += u[0] * gydt;
    u[3] += u[0] * gzdt;
}

}  // namespace gk
