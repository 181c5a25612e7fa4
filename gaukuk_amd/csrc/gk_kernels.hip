// gk_kernels.hip -- the fused RK stage kernel and its helpers, hand-written for gfx950.
//
// One launch of stage_kernel performs a whole RK stage of the reference
// (src/evolution/rk2.cpp:11-22, rk3.cpp:11-28, forward_euler.cpp:11-17):
//   boundary fill      -> an index remap + transform applied while loading (no ghost storage)
//   ConsToPrim         -> once per loaded cell, in registers
//   PLM + HLLC (x, y)  -> x faces with cross-lane exchange, y faces from a rolling row window
//   UpdateCons         -> flux divergence, RK combine, floors, store
//   gravity            -> first half-kick on U^n loads, second half-kick before the last store
//   CalCmax            -> max wave speed of the stored result, for the next step's dt
//
// Execution shape.  A wave of 64 lanes owns a strip of 60 columns (2 halo lanes each side) and
// marches up a segment of rows.  Each lane holds the primitive state of its column for three
// consecutive rows; the y-face flux between rows r-2 and r-1 and the x-face fluxes of row r-2
// are computed once per face, so every cell is converted once and every face solved once
// (plus the strip halo).  Loads of the next row are issued before the current row's flux math.
// Waves never exchange data, so there is no barrier; x neighbours move between lanes (DPP shifts
// in FP64, permutes in FP32), and LDS serves only as per-wave private storage (the 2D variants'
// ring of raw input rows, the 3D variant's ring of neighbour-plane rows).
#include "gk_device.hpp"

namespace gk {

__device__ __forceinline__ unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }
__device__ __forceinline__ double bitsd(unsigned long long b) { return __longlong_as_double((long long)b); }

// Neighbour-lane values: lane_below gives lane i the value of lane i-1, lane_above that of lane
// i+1.  Doubles move through DPP wavefront shifts (wave_shr:1 / wave_shl:1), two VALU moves
// instead of two LDS permutes and their wait: the FP64 stage kernels run 2-3% faster.  The end
// lane with no source (0, resp. 63) is not written by the wave shift and receives an unspecified
// value: each 32-bit half is then one v_mov_b32_dpp into a fresh register, where giving that lane
// its own value needs a copy into the destination first (30 fewer VALU instructions per row
// iteration, +1.7% per step).  Lanes 0 and 63 are halo lanes of every strip; no owner lane's
// result depends on a value they receive (the stencil of owner lane i reaches lanes i-2 .. i+2,
// owner lanes are 2 .. 61).  Floats keep the LDS permute: with DPP moves the FP32 stage kernels
// ran 7-10% slower.  Every lane of the wave must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_shift(double v) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_mov_dpp(hi, CTRL, 0xf, 0xf, false),
                            __builtin_amdgcn_mov_dpp(lo, CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ float dpp_shift(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, false));
}
__device__ __forceinline__ double lane_below(double v) { return dpp_shift<0x138>(v); }
__device__ __forceinline__ double lane_above(double v) { return dpp_shift<0x130>(v); }
__device__ __forceinline__ float lane_below(float v) { return __shfl_up(v, 1); }
__device__ __forceinline__ float lane_above(float v) { return __shfl_down(v, 1); }

// Max over the 64 lanes of a wave (all lanes get the result).
__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = smax(v, __shfl_xor(v, off));
    return v;
}

// Workgroup id -> logical id so that each of the 8 XCDs (which receive workgroups round-robin)
// processes one contiguous band of the grid and its halo rows stay in that XCD's L2.
__device__ __forceinline__ int xcd_logical_block(int b, int nb) {
    return (b % 8) * (nb / 8) + b / 8;
}

// Three waves per SIMD: the register budget (168 VGPRs) holds the three-row window, the
// prefetched row and both Riemann solves without spills.
#ifndef GK_WAVES_PER_EU
#define GK_WAVES_PER_EU 3
#endif
// Levels of the progress-based issue priority of the row march (0: off, at most 4).
#ifndef GK_PACE_LEVELS
#define GK_PACE_LEVELS 4
#endif

// GRAV: constant gravity is enrolled; LAST: the stage ends the step (second half-kick, next
// step's max wave speed).  THREE: the three-coefficient combine of RK stages 2 and 3.  VAR selects
// the boundary/geometry variant, so that plain 2D grids run code with no trace of the others:
// 0 plain 2D; 1 2D with fixed faces or the inner fixed region (saved data); 2 3D (nz > 1); 3 3D
// with saved data; 4 plain 2D whose states are too large for two of them to stay in the
// memory-side cache (cache-policy hints on the state streams, below).
// One wave advances the cells of strip `strip`, segment `seg` of one grid by one stage.
// `strip` and `seg` are wave-uniform (scalar registers), so row logic is scalar branches and
// scalar address arithmetic.
constexpr bool is3d(int var) { return var == 2 || var == 3; }

template <class R, class Lim, bool THREE, bool GRAV, bool LAST, int VAR, bool PACE = true>
__device__ __forceinline__ void march(const StageArgs<R>& a, int strip, int seg, int lane, int kp = 0,
                                      R* zring = nullptr, R* oring = nullptr) {
    constexpr bool FIXED = VAR == 1 || VAR == 3;
    constexpr bool THREED = is3d(VAR);
    constexpr bool KICK_IN = GRAV && !THREE;  // the stage input is U^n
    constexpr bool KICK_BASE = GRAV && THREE; // the combine base is U^n
    DevClock* clk = a.clock;
    const int c = strip * kStrip + lane - kHalo;
    const int j0 = seg * a.seg_rows;
    const int j1 = min(j0 + a.seg_rows, a.ny);
    const bool owner = lane >= kHalo && lane < kWave - kHalo && c < a.nx;

    const double dt = clk->dt;
    const R dtdx = R(dt * a.dxRec), dtdy = R(dt * a.dyRec), h = R(0.5 * dt);
    // sources of the two half-kicks (src/sim.cpp:85-91, src/source_term/source.cpp:12-21):
    // constant gravity, then point or binary gravity at the source cell's centre; the binary's
    // second mass sits at angle omega t + theta1 for t = t + dt/2 (first kick) and t + dt (second)
    R bx1[2] = {R(0), R(0)}, by1[2] = {R(0), R(0)};
    if (GRAV && a.npoint == 2) {
        const double ts[2] = {clk->t + 0.5 * dt, clk->t + dt};
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const double th = a.omega * ts[q] + a.theta1;
            bx1[q] = R(a.ab * cos(th));
            by1[q] = R(a.ab * sin(th));
        }
    }
    // the z centre of plane kz (3D) or of the single plane (2D); the 3D table exists only when
    // point masses are enrolled, the only source that reads it
    auto zcentre = [&](int kz) -> R { return (THREED && a.npoint) ? R(a.zcs[kz + a.ng]) : R(a.zc); };
    const R zc_here = zcentre(kp);
    auto kick_at = [&](R u[5], int q, int i, int j, R z) {
        gravity_kick(u, a.gx, a.gy, a.gz, h);
        if (a.npoint) point_kick(u, a.npoint, a.pm0, a.pm1, bx1[q], by1[q], h, R(a.xc[i + a.ng]),
                                 R(a.yc[j + a.ng]), z);
    };
    auto kick = [&](R u[5], int q, int i, int j) { kick_at(u, q, i, j, zc_here); };
    const long long P = a.plane;
    const int pitch = a.pitch;
    // forward-Euler ping-pong: the current state alternates between the two buffers
    const R* in = a.in;
    R* out = a.out;
    if (a.pingpong && clk->cur) { in = a.out; out = const_cast<R*>(a.in); }
    // 3D: the wave works in plane kp; its rows are addressed from the plane's first row
    const R* in3 = in;        // plane 0 of the stage input, for the z neighbours
    const R* base = a.base;
    if (THREED) {
        in += kp * a.plane2;
        out += kp * a.plane2;
        base += kp * a.plane2;
    }

    int kx;
    const int cs = bc_remap(min(c, a.nx + 1), a.nx, a.bc[0], a.bc[1], kx);
    // column of the combine operands: every lane loads them (no branch), so halo lanes clamp
    // into [0, nx-1]; only owner lanes store
    const int cown = min(max(c, 0), a.nx - 1);
    // ghost columns sit in the first strip and in every strip whose last lane reaches column nx
    // (the last strip, and also the one before it when nx = 1 mod 60); only reflective and outflow
    // faces transform them: the x transform is a uniform branch (periodic, copy and fixed faces
    // skip it, so their edge strips run no longer than the others)
    const bool xedge = strip == 0 || strip * kStrip + (kWave - 1 - kHalo) >= a.nx;
    const bool xtrans = xedge && ((a.bc[0] == 2) | (a.bc[0] == 3) | (a.bc[1] == 2) | (a.bc[1] == 3));

    // Addresses are a wave-uniform row pointer (scalar registers) plus a per-lane 32-bit byte
    // offset, which the compiler turns into scalar-base loads and stores.
    const long long Pb = P * (long long)sizeof(R);
    const unsigned off_in = (unsigned)cs * (unsigned)sizeof(R);
    const unsigned off_own = (unsigned)cown * (unsigned)sizeof(R);
    auto at = [&](const R* row, int n, unsigned off) -> const R& {
        return *reinterpret_cast<const R*>(reinterpret_cast<const char*>(row) + n * Pb + off);
    };
    // Cache policy of the state streams (VAR 4).  A step's states pass through the 256 MB
    // memory-side cache; when one state is larger than half of it (KH 2048^2 FP64: 168 MB) the
    // combine stages load their input (U1, U2), read for the last time, non-temporally, and the
    // first stage stores its output (U1 of a multi-stage integrator, read once by the next stage)
    // non-temporally, which leaves the cache to U^n, read again by every stage of the step.
    // Measured at KH 2048^2 FP64 RK3 against default loads and stores: stage 2 123 -> 102 us, step
    // +7.5%; hints on U^n's last read or on the other stores were no better.  Grids whose states
    // mostly fit lose by them (RT 1024x3072, 126 MB FP64 and 63 MB FP32: 2%), and so does 3D,
    // where waves also read their neighbour planes' rows: those run with the default policy.
    constexpr bool NT_IN = VAR == 4 && THREE;
    constexpr bool NT_OUT = VAR == 4 && !THREE && !LAST;
    auto in_at = [&](const R* row, int n) -> R {
        if (NT_IN) return __builtin_nontemporal_load(&at(row, n, off_in));
        return at(row, n, off_in);
    };
    // fixed (type 4) x faces: only the edge strips hold their ghost lanes
    const bool xfixed = FIXED && xedge && (a.bc[0] == 4 || a.bc[1] == 4);
    const int nzf = THREED ? a.nz : 1;   // planes of the saved x and y face layers
    // a ghost row of a fixed y face in plane kz: the saved layer rs (distance from the face)
    auto fixed_yrow = [&](int kz, int r, int rs, R u[5]) {
        const R* F = r < 0 ? a.fyl : a.fyr;
        const long long pl = 2LL * a.nx * nzf, o = ((long long)kz * 2 + rs) * a.nx + cs;
#pragma unroll
        for (int n = 0; n < 5; n++) u[n] = F[n * pl + o];
    };
    // inner fixed region (src/boundary/fixed.cpp:181-207): every read of a stage input or of U^n at
    // a region cell returns the saved state, unkicked (the reference resets those cells after the
    // face fills and after the first half-kick); results are stored as computed.  Region rows are
    // indexed kz * ny + r (one plane in 2D).
    auto inner_at_k = [&](int kz, int r) -> bool {
        if (!FIXED || !a.inner || r < 0 || r >= a.ny) return false;
        const int2 iv = a.inner_rows[kz * a.ny + r];
        return c >= iv.x && c < iv.y;
    };
    auto inner_load_k = [&](int kz, int r, R u[5]) {
        const long long o = ((long long)kz * a.ny + r) * pitch + c;
#pragma unroll
        for (int n = 0; n < 5; n++) u[n] = a.inner[n * P + o];
    };
    auto inner_at = [&](int r) -> bool { return inner_at_k(kp, r); };
    auto inner_load = [&](int r, R u[5]) { inner_load_k(kp, r, u); };
    // raw conserved values of row r (ghost rows remapped; the row index is wave-uniform)
    auto load_row = [&](int r, R u[5]) {
        int ky;
        const int rs = bc_remap(r, a.ny, a.bc[2], a.bc[3], ky);
        if (FIXED && ky == 4) {
            fixed_yrow(kp, r, rs, u);
        } else {
            const R* row = in + (long long)rs * pitch;
#pragma unroll
            for (int n = 0; n < 5; n++) u[n] = in_at(row, n);
        }
        if (xfixed && kx == 4) {
            // a ghost column of a fixed x face (cs is its distance from the face); corner cells
            // are never read, so ghost rows take the nearest active row
            const R* F = c < 0 ? a.fxl : a.fxr;
            const long long pl = 2LL * a.ny * nzf;
            const long long o = ((long long)kp * a.ny + min(max(r, 0), a.ny - 1)) * 2 + cs;
#pragma unroll
            for (int n = 0; n < 5; n++) u[n] = F[n * pl + o];
        }
        if (inner_at(r)) inner_load(r, u);
    };
    // gravity half-kick on U^n, boundary transforms of ghost cells, ConsToPrim
    auto finish_row = [&](int r, R u[5], R wout[5]) {
        if (KICK_IN) {
            // ghosts copy kicked active cells, except those of fixed faces, which hold saved data
            int kyy;
            const int rs = bc_remap(r, a.ny, a.bc[2], a.bc[3], kyy);
            if (!FIXED || (kyy != 4 && !(xfixed && kx == 4) && !inner_at(r))) kick(u, 0, cs, rs);
        }
        if (xtrans) bc_transform_lanes(u, 1, kx);
        if (r < 0 || r >= a.ny) {
            int ky;
            bc_remap(r, a.ny, a.bc[2], a.bc[3], ky);
            bc_transform(u, 2, ky);
        }
        a.eos.cons_to_prim(u, wout);
    };
    // y slope of row B (between A below and C above): right state of B's lower face, left state
    // of its upper face
    auto yslope = [&](const R A[5], const R B[5], const R C[5], R wR[5], R wL[5]) {
#pragma unroll
        for (int n = 0; n < 5; n++) {
            const R s = Lim::slope(B[n] - A[n], C[n] - B[n]);
            wR[n] = B[n] - R(0.5) * s;
            wL[n] = B[n] + R(0.5) * s;
        }
    };
    // y-face flux with the direction-VLY rotation (src/flux/hllc.cpp:25-27): (den, vy, vz, vx, p)
    auto yflux = [&](const R wl[5], const R wr[5], R G[5]) {
        const R l[5] = {wl[0], wl[2], wl[3], wl[1], wl[4]};
        const R rr[5] = {wr[0], wr[2], wr[3], wr[1], wr[4]};
        R f[5];
        RiemannPolicy::flux(a.eos, l, rr, f);
        G[0] = f[0]; G[2] = f[1]; G[3] = f[2]; G[1] = f[3]; G[4] = f[4];
    };
    // x-face fluxes of one row: face c (F) and face c+1 (Fn) through neighbour lanes
    auto xflux = [&](const R A[5], R F[5], R Fn[5]) {
        R wR[5], wLo[5], wL[5];
#pragma unroll
        for (int n = 0; n < 5; n++) {
            const R am = lane_below(A[n]), ap = lane_above(A[n]);
            const R s = Lim::slope(A[n] - am, ap - A[n]);
            wR[n] = A[n] - R(0.5) * s;   // right state of face c
            wLo[n] = A[n] + R(0.5) * s;  // left state of face c+1
        }
#pragma unroll
        for (int n = 0; n < 5; n++) wL[n] = lane_below(wLo[n]);
        RiemannPolicy::flux(a.eos, wL, wR, F);
#pragma unroll
        for (int n = 0; n < 5; n++) Fn[n] = lane_above(F[n]);
    };

    // ---- z faces (3D) ----------------------------------------------------------------------
    // The reference's z sweep reconstructs with ReconstructYZPLM at (k, j), whose slope runs along y
    // (recon_piecewise_linear.cpp:34-55, cal_flux.cpp:140-157): the face between planes k-1 and k
    // takes W(k-1, j) + sigma_y/2 on the left and W(k, j) - sigma_y/2 on the right (first order:
    // the cells themselves).  This wave has both states of its own plane from the y march; the
    // planes below and above are read directly (their rows come from L2: neighbouring planes are
    // marched by neighbouring waves).
    const R dtdz = R(dt * a.dzRec);
    R wRz[5];   // W - sigma_y/2 of the row the next update reads (plane kp)
    // The stage reads the array that held the state at step stale_base on even steps from there
    // when it is the first stage (the reference's `cons`), on odd steps otherwise (`consTemp`,
    // swapped with `cons` at every step end).
    const bool readsX = THREED && ((a.stage_q == 0) == (((clk->step - a.stale_base) & 1) == 0));
    // raw conserved values of (plane kz, row jz) in this lane's column, boundary rules applied
    auto zload = [&](int kz, int jz, R u[5]) {
        if (kz >= 0 && kz < a.nz) {
            int ky;
            const int rs = bc_remap(jz, a.ny, a.bc[2], a.bc[3], ky);
            // saved data (fixed y-face rows, inner-region cells) is read as it is: no kick
            if (FIXED && ky == 4) {
                fixed_yrow(kz, jz, rs, u);
                return;
            }
            if (inner_at_k(kz, jz)) {
                inner_load_k(kz, jz, u);
                return;
            }
            const R* row = in3 + kz * a.plane2 + (long long)rs * pitch;
#pragma unroll
            for (int n = 0; n < 5; n++) u[n] = at(row, n, off_in);
            if (KICK_IN) kick_at(u, 0, cs, rs, zcentre(kz));
            if (jz < 0 || jz >= a.ny) bc_transform(u, 2, ky);
            return;
        }
        // a z ghost plane: z faces of types 1-3 fill its active rows (periodic.cpp:135-190,
        // reflective.cpp:170-230, outflow.cpp:216-300) from the active planes as they stand before
        // the inner region's reset (boundary.cpp:127-142); a fixed face restores its saved plane
        // (fixed.cpp:126-179); type 0 registers the y copy instead (boundary.cpp:90-122), and no
        // face fills edge cells
        const int zt = kz < 0 ? a.bc[4] : a.bc[5];
        if (FIXED && zt == 4 && jz >= 0 && jz < a.ny) {
            const R* F = kz < 0 ? a.fzl : a.fzr;
            const long long pl = (long long)a.ny * a.nx, o = (long long)jz * a.nx + cown;
#pragma unroll
            for (int n = 0; n < 5; n++) u[n] = F[n * pl + o];
        } else if ((zt == 1 || zt == 2 || zt == 3) && jz >= 0 && jz < a.ny) {
            int kk;
            const int ks = bc_remap(kz, a.nz, a.bc[4], a.bc[5], kk);
            const R* row = in3 + ks * a.plane2 + (long long)jz * pitch;
#pragma unroll
            for (int n = 0; n < 5; n++) u[n] = at(row, n, off_in);
            if (KICK_IN) kick_at(u, 0, cs, jz, zcentre(ks));
            bc_transform(u, 3, kk);
        } else {
            const long long pl = (long long)(a.ny + 2) * a.nx;
            const long long o = (long long)(kz < 0 ? 0 : 5) * pl + (long long)(jz + 1) * a.nx + cown;
#pragma unroll
            for (int n = 0; n < 5; n++) u[n] = readsX ? a.zstale[o + n * pl] : R(0);
        }
    };
    // Primitive rows of the two neighbour planes (pl 0: kp-1, 1: kp+1), rows j0-1 .. j1, in a
    // three-row ring in this wave's LDS slice ([2][3][5][64], each lane its own column): every
    // row is loaded and converted once per segment instead of three times.
    auto zslot = [&](int pl, int row, int n) -> R& {
        return zring[((pl * 3 + (row - j0 + 1) % 3) * 5 + n) * kWave + lane];
    };
    auto zrow = [&](int pl, int row) {
        R u[5], w[5];
        zload(pl ? kp + 1 : kp - 1, row, u);
        a.eos.cons_to_prim(u, w);
#pragma unroll
        for (int n = 0; n < 5; n++) zslot(pl, row, n) = w[n];
    };
    // W -+ sigma_y/2 of (neighbour plane pl, row jz): rows jz-1, jz from the ring, row jz+1 added
    auto zstate = [&](int pl, int jz, R lo[5], R hi[5]) {
        zrow(pl, jz + 1);
        R w0[5], w1[5], w2[5];
#pragma unroll
        for (int n = 0; n < 5; n++) {
            w0[n] = zslot(pl, jz - 1, n);
            w1[n] = zslot(pl, jz, n);
            w2[n] = zslot(pl, jz + 1, n);
        }
        yslope(w0, w1, w2, lo, hi);
    };
    // z-face flux with the direction-VLZ rotation (hllc.cpp:25-27): (den, vz, vx, vy, p)
    auto zflux = [&](const R wl[5], const R wr[5], R H[5]) {
        const R l[5] = {wl[0], wl[3], wl[1], wl[2], wl[4]};
        const R rr[5] = {wr[0], wr[3], wr[1], wr[2], wr[4]};
        R f[5];
        RiemannPolicy::flux(a.eos, l, rr, f);
        H[0] = f[0]; H[3] = f[1]; H[1] = f[2]; H[2] = f[3]; H[4] = f[4];
    };
    // the z-flux difference H(k+1/2) - H(k-1/2) of row rr, from this plane's states (lo, hi)
    auto zfluxes = [&](int rr, const R lo_here[5], const R hi_here[5], R Hz[5]) {
        R lo[5], hi[5], Hd[5], Hu[5];
        zstate(0, rr, lo, hi);
        zflux(hi, lo_here, Hd);
        zstate(1, rr, lo, hi);
        zflux(hi_here, lo, Hu);
#pragma unroll
        for (int n = 0; n < 5; n++) Hz[n] = Hu[n] - Hd[n];
    };

    R pre[5];               // next input row, raw
    R wLy[5], Gp[5];        // left state of the next y face; y flux at the lower face of row r-2
    R cacc = R(0);
    const int rlast = j1 + 1;

    // the update of row r-2 from its x fluxes (F, Fn) and y fluxes (Gp below, G above)
    auto update = [&](int r, const R bs[5], const R ut[5], const R F[5], const R Fn[5], const R G[5],
                      const R Hz[5]) {
        const long long o = (long long)(r - 2) * pitch;
        R b2[5];
#pragma unroll
        for (int n = 0; n < 5; n++) b2[n] = bs[n];
        // the base of every stage is U^n after the first gravity half-kick
        if (KICK_IN || KICK_BASE)
            if (!inner_at(r - 2)) kick(b2, 0, cown, r - 2);
        R res[5];
#pragma unroll
        for (int n = 0; n < 5; n++) {
            R dF = (Fn[n] - F[n]) * dtdx + (G[n] - Gp[n]) * dtdy;
            if (THREED) dF = dF + Hz[n] * dtdz;   // update_cons.cpp:76-132: the z term is added last
            if (THREE) res[n] = a.c1 * b2[n] + a.c2 * ut[n] - a.c3 * dF;
            else res[n] = a.c1 * b2[n] - a.c2 * dF;
        }
        const R rinv = a.eos.floors(res);
        if (GRAV && LAST) kick(res, 1, cown, r - 2);
        if (LAST) {
            const R ws = a.eos.wave_speed(res, rinv);
            cacc = owner ? smax(cacc, ws) : cacc;
        }
        if (owner) {
            char* row = reinterpret_cast<char*>(out + o) + off_own;
#pragma unroll
            for (int n = 0; n < 5; n++) {
                R* dst = reinterpret_cast<R*>(row + n * Pb);
                if (NT_OUT) __builtin_nontemporal_store(res[n], dst);
                else *dst = res[n];
            }
        }
    };
    // 2D: the raw stage-input rows this wave loads, at its own column, stay in a 4-row ring in the
    // wave's slice of LDS until the update two rows later takes them as its combine operand.
    // Re-reading them from memory missed L2 (the rows were loaded three row iterations, tens of
    // megabytes of traffic, earlier), and cost up to one extra read of the state per stage.  Owner
    // lanes load their own column, so the ring holds exactly the operand (inner-region values
    // included, before the gravity kick).  3D keeps its LDS for the z-plane ring.
    constexpr bool RING = !THREED;
    auto ring_put = [&](int r, const R u[5]) {
        if (RING) {
#pragma unroll
            for (int n = 0; n < 5; n++) oring[((r & 3) * 5 + n) * kWave + lane] = u[n];
        }
    };
    auto ring_get = [&](int r, int n) -> R { return oring[((r & 3) * 5 + n) * kWave + lane]; };
    // operands of the update of row r-2, issued before the flux math
    auto load_operands = [&](int r, R bs[5], R ut[5]) {
        const long long o = (long long)(r - 2) * pitch;
        if (THREE) {
#pragma unroll
            for (int n = 0; n < 5; n++) {
                bs[n] = at(base + o, n, off_own);
                ut[n] = RING ? ring_get(r - 2, n) : at(in + o, n, off_own);
            }
        } else {
#pragma unroll
            for (int n = 0; n < 5; n++) { bs[n] = RING ? ring_get(r - 2, n) : at(in + o, n, off_own); ut[n] = R(0); }
        }
        if (inner_at(r - 2)) {
            inner_load(r - 2, bs);
            if (THREE) inner_load(r - 2, ut);
        }
    };

    R W0[5], W1[5], W2[5];  // primitive rows, rotated by the unrolled loop
    {
        R u[5];
        load_row(j0 - 2, pre);
#pragma unroll
        for (int n = 0; n < 5; n++) u[n] = pre[n];
        load_row(j0 - 1, pre);
        finish_row(j0 - 2, u, W0);
#pragma unroll
        for (int n = 0; n < 5; n++) u[n] = pre[n];
        load_row(j0, pre);
        finish_row(j0 - 1, u, W1);
#pragma unroll
        for (int n = 0; n < 5; n++) u[n] = pre[n];
        ring_put(j0, u);
        load_row(j0 + 1, pre);
        finish_row(j0, u, W2);
        R wR[5];
        yslope(W0, W1, W2, wR, wLy);   // wLy: left state of face j0
#pragma unroll
        for (int n = 0; n < 5; n++) u[n] = pre[n];
        ring_put(j0 + 1, u);
        load_row(j0 + 2, pre);
        finish_row(j0 + 1, u, W0);     // W0 now holds row j0+1
        R wLn[5];
        yslope(W1, W2, W0, wR, wLn);   // slope of row j0
        if (THREED) {
#pragma unroll
            for (int n = 0; n < 5; n++) wRz[n] = wR[n];
        }
        yflux(wLy, wR, Gp);            // face j0
        if (THREED) {
            // the neighbour planes' rows below the first update row
#pragma unroll
            for (int pl = 0; pl < 2; pl++) { zrow(pl, j0 - 1); zrow(pl, j0); }
        }
#pragma unroll
        for (int n = 0; n < 5; n++) wLy[n] = wLn[n];
    }
    // FP32: the combine operands of the next update row are loaded one row ahead, a whole row
    // iteration before their use (10 more registers, within what the FP32 kernels' occupancy
    // leaves free; FP64 would lose a resident wave per SIMD).  Past the segment the row index is
    // clamped, as for the prefetched input row.
    constexpr bool EARLY_OPS = sizeof(R) == 4;
    R nbs[5], nut[5];
    if (EARLY_OPS) load_operands(j0 + 2, nbs, nut);   // row j0, the first update
    // steady state at newest row r: y face r-1 and the update of row r-2 (primitives in A)
    auto body = [&](int r, R (&A)[5], R (&B)[5], R (&Cn)[5]) {
        R u[5], bs[5], ut[5];
#pragma unroll
        for (int n = 0; n < 5; n++) u[n] = pre[n];
        ring_put(r, u);
        load_row(min(r + 1, rlast), pre);
        if (EARLY_OPS) {
#pragma unroll
            for (int n = 0; n < 5; n++) { bs[n] = nbs[n]; ut[n] = nut[n]; }
            load_operands(min(r + 1, rlast), nbs, nut);
        } else {
            load_operands(r, bs, ut);
        }
        finish_row(r, u, Cn);
        R wR[5], wLn[5], G[5], F[5], Fn[5];
        yslope(A, B, Cn, wR, wLn);
        yflux(wLy, wR, G);
        xflux(A, F, Fn);
        R Hz[5] = {R(0), R(0), R(0), R(0), R(0)};
        if (THREED) zfluxes(r - 2, wRz, wLy, Hz);
        update(r, bs, ut, F, Fn, G, Hz);
#pragma unroll
        for (int n = 0; n < 5; n++) { Gp[n] = G[n]; wLy[n] = wLn[n]; wRz[n] = wR[n]; }
    };
    // Issue priority falls as the wave moves through its segment.  The SIMD's arbiter otherwise
    // serves the oldest wave first, so of the waves that start a one-round launch together on a
    // SIMD one finishes after about a third of the launch and the last runs the final third alone,
    // with nothing to hide its latencies; with the priority following progress the waves keep pace
    // and stay resident together (wave timestamps at KH 2048^2: 2.0 -> 2.5 of 3 resident waves on
    // average; 0.3-1.5% faster single-grid steps).  The persistent ensemble kernel, whose waves
    // take work items from a queue, runs without it (2% slower with it).  Wave-uniform scalar
    // branches between rows; no effect on results.
    const int total = rlast - (j0 + 2) + 1;
    auto pace = [&](int r) {
#if GK_PACE_LEVELS > 0
        if (!PACE) return;
        const int lvl = (GK_PACE_LEVELS * (r - (j0 + 2))) / total;
        if (lvl <= 0) __builtin_amdgcn_s_setprio(3);
        else if (lvl == 1) __builtin_amdgcn_s_setprio(2);
        else if (lvl == 2) __builtin_amdgcn_s_setprio(1);
        else __builtin_amdgcn_s_setprio(0);
#else
        (void)r;
#endif
    };
    for (int r = j0 + 2;;) {
        pace(r); body(r, W2, W0, W1); if (++r > rlast) break;
        pace(r); body(r, W0, W1, W2); if (++r > rlast) break;
        pace(r); body(r, W1, W2, W0); if (++r > rlast) break;
    }
    (void)total;

    if (LAST) {
        const double m = wave_max((double)cacc);
        if (lane == 0) atomicMax(&clk->cmax_bits[clk->parity ^ 1][(seg * a.strips + strip + kp) % kCmaxSlots], dbits(m));
    }
}

// One RK stage of one grid: wave w of the (XCD-remapped) grid takes strip w % strips, segment
// w / strips.
// 3D waves hold the z-face work on top of the 2D march: two waves per SIMD keep it out of scratch.
template <class R, class Lim, bool THREE, bool GRAV, bool LAST, int VAR>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(is3d(VAR) ? 2 : GK_WAVES_PER_EU)))
void stage_kernel(const StageArgs<R> a) {
    if (!a.clock->active) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int w = xcd_logical_block(blockIdx.x, gridDim.x) * (kBlock / kWave) + wid;
    if (w >= a.nwaves) return;
    if constexpr (is3d(VAR)) {
        // 3D: strip fastest, then segment, then plane (waves of neighbouring planes share an XCD);
        // each wave keeps its neighbour-plane row ring in its own LDS slice
        __shared__ R zring[kBlock / kWave][2 * 3 * 5 * kWave];
        const int per_plane = a.nwaves / a.nz;
        const int kp = w / per_plane, r = w - kp * per_plane;
        march<R, Lim, THREE, GRAV, LAST, VAR>(a, r % a.strips, r / a.strips, lane, kp, zring[wid]);
    } else {
        __shared__ R oring[kBlock / kWave][4 * 5 * kWave];
        march<R, Lim, THREE, GRAV, LAST, VAR>(a, w % a.strips, w / a.strips, lane, 0, nullptr, oring[wid]);
    }
}

// ---------------------------------------------------------------------------------------------
// Immersed reflective shell (Boundary::ImmersedReflective, src/boundary/immersed.cpp:237-300),
// run on a stage's input before the stage (Boundary::UpdateBD applies it after the face fills,
// src/boundary/boundary.cpp:136-138).  Each shell cell -- inside the body, within nGhost cells of
// its surface -- takes the distance-weighted image of up to 8 fluid cells around its mirror
// point, with the radial velocity reflected.  Shell cells lie inside the body and targets outside
// it, so one thread per shell cell rewrites the state in place with no ordering.
template <class R>
__global__ __launch_bounds__(256) void immersed_kernel(const ShellArgs<R> a) {
    if (!a.always && !a.clock->active) return;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nshell) return;
    R* u = (a.pingpong && a.clock->cur) ? a.alt : a.u;
    const long long P = a.plane;
    const int ns = a.nshell;
    const R cosp = a.dir[s], sinp = a.dir[ns + s], cost = a.dir[2 * ns + s], sint = a.dir[3 * ns + s];
    const R g0 = a.gam[s];
    const R wd = R(1.) / (R(1) - g0), wr = R(1.) / (R(-1.) - g0);
    R dens = 0, vr = 0, vp = 0, vt = 0, pres = 0;
    const int nt = a.ntar[s];
    for (int n = 1; n <= nt; n++) {
        const long long o = a.tar[(long long)(n - 1) * ns + s];
        const R gn = a.gam[(long long)n * ns + s];
        const R denTar = smax(u[o], R(GK_DENSITY_FLOOR));
        const R denInv = R(1.0) / denTar;
        dens += wd * gn * denTar;
        const R vx = u[P + o] * denInv, vy = u[2 * P + o] * denInv, vz = u[3 * P + o] * denInv;
        // Cartesian -> (radial, azimuthal, polar) about the body centre (CarToLocal, immersed.cpp:73-79)
        const R vrl = vx * cosp * sint + vy * sinp * sint + vz * cost;
        const R vpl = -vx * sinp + vy * cosp;
        const R vtl = vx * cosp * cost + vy * sinp * cost - vz * sint;
        vr += wr * gn * vrl;
        vp += wd * gn * vpl;
        vt += wd * gn * vtl;
        const R engKin = R(0.5) * denTar * (vx * vx + vy * vy + vz * vz);
        R preTar = a.gm1 * (u[4 * P + o] - engKin);
        preTar = (preTar > R(GK_PRESSURE_FLOOR)) ? preTar : R(GK_PRESSURE_FLOOR);
        pres += wd * gn * preTar;
    }
    // back to Cartesian (LocalToCar, immersed.cpp:82-87)
    const R vx = vr * sint * cosp + vt * cost * cosp - vp * sinp;
    const R vy = vr * sint * sinp + vt * cost * sinp + vp * cosp;
    const R vz = vr * cost - vt * sint;
    const long long c = a.cell[s];
    u[c] = dens;
    u[P + c] = dens * vx;
    u[2 * P + c] = dens * vy;
    u[3 * P + c] = dens * vz;
    u[4 * P + c] = R(0.5) * dens * (vx * vx + vy * vy + vz * vz) + pres * a.gm1Rec;
}

// The inner fixed region written into a state (Boundary::FixedBDInner, src/boundary/fixed.cpp:
// 181-207), for the boundary update outside the stages; rows [j0, j0 + nrows), one thread per
// (row, column) of the region's bounding columns [i0, i0 + width).
template <class R>
__global__ __launch_bounds__(256) void inner_reset_kernel(R* u, const R* saved, const int2* rows, long long plane,
                                                          int pitch, int j0, int nrows, int i0, int width) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long long)nrows * width) return;
    const int j = j0 + (int)(k / width), i = i0 + (int)(k % width);
    const int2 iv = rows[j];
    if (i < iv.x || i >= iv.y) return;
    const long long o = (long long)j * pitch + i;
#pragma unroll
    for (int n = 0; n < 5; n++) u[n * plane + o] = saved[n * plane + o];
}

// ---------------------------------------------------------------------------------------------
// Ensembles: many independent grids of uneven sizes advanced together.  Each problem's stage is
// split into work items (strip, segment); a persistent grid of waves claims items from an atomic
// counter until all are done, so large and small problems share the machine without a launch per
// problem.  Each problem has its own clock (its own dt).
// Work queues of the ensemble kernel: one per XCD (workgroups are dealt to the 8 XCDs round-robin,
// so workgroup b runs on XCD b % 8).  Each queue holds a contiguous eighth of the items, so
// neighbouring strips and segments of a problem -- which share halo rows and the cache lines at
// strip edges -- run on the same XCD and meet in its L2.
constexpr int kEnsQueues = 8;

struct EnsProblem {
    long long off[3];     // element offsets of the problem's state buffers A, B, C in the arena
    long long plane;
    int pitch, nx, ny, seg_rows, strips, items, item_base;
    int bc[4];
    double dxRec, dyRec;
};

template <class R>
struct EnsArgs {
    R* arena;
    const EnsProblem* probs;
    DevClock* clocks;
    int* counter;         // this stage's work-queue heads, one per XCD (reset by ens_step_begin_kernel)
    int nprob, total_items;
    int xcd_begin[kEnsQueues + 1];   // queue x hands out items [xcd_begin[x], xcd_begin[x + 1])
    int reverse;          // hand the queue's items out from its last one down
    int in_buf, base_buf, out_buf;
    StageArgs<R> proto;   // stage constants (coefficients, flags, gravity, EOS)
};

template <class R, class Lim, bool THREE, bool GRAV, bool LAST>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(GK_WAVES_PER_EU)))
void ens_stage_kernel(const EnsArgs<R> e) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    __shared__ R oring[kBlock / kWave][4 * 5 * kWave];
    const int q = blockIdx.x % kEnsQueues;
    const int first = e.xcd_begin[q], count = e.xcd_begin[q + 1] - first;
    int* head = e.counter + q;
    // the k-th claim of a queue takes its item first + k, or first + count - 1 - k in reverse
    int k = 0;
    if (lane == 0) k = atomicAdd(head, 1);
    k = __builtin_amdgcn_readfirstlane(k);
    while (k < count) {
        const int item = e.reverse ? first + count - 1 - k : first + k;
        // claim the next item before working on this one, to hide the atomic's latency
        int next = 0;
        if (lane == 0) next = atomicAdd(head, 1);
        next = __builtin_amdgcn_readfirstlane(next);
        // the problem owning `item`: last p with item_base <= item (binary search, scalar loads)
        int lo = 0, hi = e.nprob - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (e.probs[mid].item_base <= item) lo = mid; else hi = mid - 1;
        }
        const EnsProblem& d = e.probs[lo];
        DevClock* clk = e.clocks + lo;
        if (clk->active) {
            StageArgs<R> a = e.proto;
            a.in = e.arena + d.off[e.in_buf];
            a.base = e.arena + d.off[e.base_buf];
            a.out = e.arena + d.off[e.out_buf];
            a.plane = d.plane;
            a.pitch = d.pitch; a.nx = d.nx; a.ny = d.ny;
            a.seg_rows = d.seg_rows; a.strips = d.strips;
            a.bc[0] = d.bc[0]; a.bc[1] = d.bc[1]; a.bc[2] = d.bc[2]; a.bc[3] = d.bc[3];
            a.dxRec = d.dxRec; a.dyRec = d.dyRec;
            a.clock = clk;
            const int local = item - d.item_base;
            march<R, Lim, THREE, GRAV, LAST, 0, false>(a, local % d.strips, local / d.strips, lane, 0, nullptr,
                                                       oring[wid]);
        }
        k = next;
    }
}

// Max wave speed of a state (step-0 dt, or after an upload): CalCmax, adiabatic.cpp:115-160.
template <class R>
__global__ __launch_bounds__(kBlock) void cmax_kernel(const R* u, long long plane, int pitch, int nx, int ny,
                                                     EosPolicy<R> eos, DevClock* clk) {
    const R* src = u;
    R acc = R(0);
    const long long n = (long long)nx * ny;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
        const long long j = k / nx, i = k - j * nx;
        const long long o = j * pitch + i;
        R c[5];
#pragma unroll
        for (int q = 0; q < 5; q++) c[q] = src[o + q * plane];
        acc = smax(acc, eos.wave_speed(c));
    }
    const double m = wave_max((double)acc);
    if ((threadIdx.x & 63) == 0) atomicMax(&clk->cmax_bits[clk->parity][blockIdx.x % kCmaxSlots], dbits(m));
}

// Sim::Advance prologue (src/sim.cpp:70-71): tNext = t + dtoutput, dtUntilOutput = dtoutput.
__global__ void advance_begin_kernel(DevClock* c, double dtoutput) {
    c->t_next = c->t + dtoutput;
    c->dt_until_output = dtoutput;
    c->active = c->is_continue;
}

// Loop test and dt of one step (src/sim.cpp:72-81).  The words cmax_bits[parity][*] hold
// max(1e-16, cells) between them.
__device__ __forceinline__ void clock_begin(DevClock* c) {
    if (!c->active) return;
    const double gap = c->t_next - c->t;
    if (!((gap < 0 ? -gap : gap) > 1e-12 && c->is_continue)) { c->active = 0; return; }
    // positive doubles order like their bit patterns
    unsigned long long mb = c->cmax_bits[c->parity][0];
    for (int s = 1; s < kCmaxSlots; s++) mb = mb < c->cmax_bits[c->parity][s] ? c->cmax_bits[c->parity][s] : mb;
    const double cm = bitsd(mb);
    double dt = c->cfl * c->drmin / cm;
    dt = smin(dt, c->dt_until_output);
    c->dt = dt;
    for (int s = 0; s < kCmaxSlots; s++) c->cmax_bits[c->parity ^ 1][s] = dbits(1e-16);
}

// Loop tail (src/sim.cpp:93-98).
__device__ __forceinline__ bool clock_end(DevClock* c, int toggle_cur) {
    if (!c->active) return false;
    c->t += c->dt;
    c->step++;
    c->dt_until_output = c->t_next - c->t;
    if (c->stepmax > 0 && c->step >= c->stepmax) c->is_continue = 0;
    c->parity ^= 1;
    if (toggle_cur) c->cur ^= 1;
    return true;
}

__global__ void step_begin_kernel(DevClock* c) { clock_begin(c); }

// plus the step log entry (the reference prints it, src/sim.cpp:100-103)
__device__ __forceinline__ void clock_end_logged(DevClock* c, StepRecord* log, int cap, int toggle_cur) {
    if (!clock_end(c, toggle_cur)) return;
    StepRecord rec;
    rec.t = c->t; rec.dt = c->dt; rec.step = c->step; rec.pad = 0;
    log[(c->step - 1) % cap] = rec;
}
__global__ void step_end_kernel(DevClock* c, StepRecord* log, int cap, int toggle_cur) {
    clock_end_logged(c, log, cap, toggle_cur);
}
// Between two steps enqueued back to back: the tail of the first and the loop test and dt of the
// second in one launch (one dispatch less per step; the same operations in the same order).
__global__ void step_seam_kernel(DevClock* c, StepRecord* log, int cap, int toggle_cur) {
    clock_end_logged(c, log, cap, toggle_cur);
    clock_begin(c);
}

// Ensemble step prologue/epilogue: one thread per problem; the first thread also rewinds the
// work-queue heads of the step's stages (a kernel, not a memset, so the step can be captured).
__global__ void ens_step_begin_kernel(DevClock* clocks, int n, int* counters, int ncounters) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < ncounters) counters[p] = 0;
    if (p < n) clock_begin(clocks + p);
}
__global__ void ens_step_end_kernel(DevClock* clocks, int n, int toggle_cur) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) clock_end(clocks + p, toggle_cur);
}
__global__ void ens_advance_begin_kernel(DevClock* clocks, int n, double dtoutput) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) {
        DevClock* c = clocks + p;
        c->t_next = c->t + dtoutput;
        c->dt_until_output = dtoutput;
        c->active = c->is_continue;
    }
}

// Deterministic FP64 sums of the five conserved variables: fixed cell->thread assignment and
// a fixed tree order per block, then the partials are summed in block order.
template <class R>
__global__ __launch_bounds__(256) void sums_partial_kernel(const R* u, long long plane, int pitch, int nx, int ny,
                                                          double* partial) {
    __shared__ double sh[5][256];
    const long long n = (long long)nx * ny;
    double acc[5] = {0, 0, 0, 0, 0};
    for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long long)gridDim.x * 256) {
        const long long j = k / nx, i = k - j * nx;
        const long long o = j * pitch + i;
#pragma unroll
        for (int q = 0; q < 5; q++) acc[q] += (double)u[o + q * plane];
    }
#pragma unroll
    for (int q = 0; q < 5; q++) sh[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
#pragma unroll
            for (int q = 0; q < 5; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x < 5) partial[blockIdx.x * 5 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void sums_final_kernel(const double* partial, int nblocks, double* out) {
    if (threadIdx.x < 5) {
        double s = 0;
        for (int b = 0; b < nblocks; b++) s += partial[b * 5 + threadIdx.x];
        out[threadIdx.x] = s;
    }
}

// Self-check of the stencil's quotient and root cores (srcp, sdiv, ssqrt, sqrt_ratio, srsqrt in
// FP64 and FP32) against the compiler's IEEE operators on operands drawn from a SplitMix64 stream:
// mantissas uniform, exponents uniform in [-emax, emax] (quotient operands) and [-2*emax, 2*emax]
// (root operand), random signs for the quotient.  Counts every result farther from the IEEE one
// than its bound (about twice the largest error measured over 3 x 64M draws).
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double draw(unsigned long long bits, int emax, bool sign) {
    const unsigned long long mant = bits & 0xFFFFFFFFFFFFFull;
    const int e = (int)((bits >> 52) % (unsigned long long)(2 * emax + 1)) - emax;
    unsigned long long v = ((unsigned long long)(e + 1023) << 52) | mant;
    if (sign && ((bits >> 63) & 1ull)) v |= 1ull << 63;
    return __longlong_as_double((long long)v);
}
// FP32 operand: 23 mantissa bits, exponent in [-emax, emax]
__device__ __forceinline__ float drawf(unsigned mant, unsigned bits, int emax) {
    const int e = (int)(bits % (unsigned)(2 * emax + 1)) - emax;
    unsigned v = ((unsigned)(e + 127) << 23) | (mant & 0x7FFFFFu);
    if ((bits >> 31) & 1u) v |= 1u << 31;
    return __uint_as_float(v);
}
__device__ __forceinline__ unsigned long long ulp_dist(double a, double b) {
    const long long d = __double_as_longlong(a) - __double_as_longlong(b);
    return (unsigned long long)(d < 0 ? -d : d);
}
__device__ __forceinline__ unsigned ulp_dist(float a, float b) {
    const int d = __float_as_int(a) - __float_as_int(b);
    return (unsigned)(d < 0 ? -d : d);
}
// FP64 draws with exponents up to +-emax; FP32 draws with exponents up to +-min(emax, 60) whose
// divisor mantissa runs through all 2^23 values as k does (count >= 2^23 covers every divisor
// mantissa); each result beyond its bound counts once.  Bounds in ulp: FP64 quotient 32,
// reciprocal 16, root 64, sqrt(a / b) 64, 1 / sqrt 32 (measured 19, 11, 37, 38, 20); FP32 4 each
// (measured 1-2).
__global__ void selftest_kernel(long long n, unsigned long long seed, int emax, unsigned long long* bad) {
    unsigned long long local = 0;
    const int emax32 = emax < 60 ? emax : 60;
    {
        // the DPP lane shifts against the cross-lane shuffles (every lane active) on every lane
        // that has a source lane (the end lane's value is unspecified)
        const unsigned long long h = mix64(~seed ^ (blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x));
        const double v = draw(h, emax, true);
        const float f = drawf((unsigned)h, (unsigned)(h >> 32), emax32);
        const int ln = threadIdx.x & (kWave - 1);
        const double vb = lane_below(v), va = lane_above(v);
        const float fb = dpp_shift<0x138>(f), fa = dpp_shift<0x130>(f);
        const double sb = __shfl_up(v, 1), sa = __shfl_down(v, 1);
        const float tb = __shfl_up(f, 1), ta = __shfl_down(f, 1);
        local += ln > 0 && dbits(vb) != dbits(sb);
        local += ln < kWave - 1 && dbits(va) != dbits(sa);
        local += ln > 0 && __float_as_uint(fb) != __float_as_uint(tb);
        local += ln < kWave - 1 && __float_as_uint(fa) != __float_as_uint(ta);
    }
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
        const double a = draw(mix64(seed ^ (3ull * k)), emax, true);
        const double b = draw(mix64(seed ^ (3ull * k + 1)), emax, true);
        const double x = fabs(draw(mix64(seed ^ (3ull * k + 2)), 2 * emax, false));
        const double pa = fabs(a), pb = fabs(b);
        local += ulp_dist(sdiv(a, b), a / b) > 32;
        local += ulp_dist(srcp(b), 1.0 / b) > 16;
        local += ulp_dist(ssqrt(x), __builtin_sqrt(x)) > 64;
        local += ulp_dist(sqrt_ratio(pa, pb), __builtin_sqrt(pa / pb)) > 64;
        local += ulp_dist(srsqrt(x), 1.0 / __builtin_sqrt(x)) > 32;
        const unsigned long long h = mix64(seed ^ (3ull * k + 5));
        const float a32 = drawf((unsigned)h, (unsigned)(h >> 32), emax32);
        const float b32 = drawf((unsigned)k, (unsigned)(h >> 40) * 2654435761u, emax32);
        const float x32 = fabsf(drawf((unsigned)(h >> 7), (unsigned)(h >> 21), emax32));
        const float pa32 = fabsf(a32), pb32 = fabsf(b32);
        local += ulp_dist(sdiv(a32, b32), a32 / b32) > 4;
        local += ulp_dist(srcp(b32), 1.0f / b32) > 4;
        local += ulp_dist(ssqrt(x32), __builtin_sqrtf(x32)) > 4;
        local += ulp_dist(sqrt_ratio(pa32, pb32), __builtin_sqrtf(pa32 / pb32)) > 4;
        local += ulp_dist(srsqrt(x32), 1.0f / __builtin_sqrtf(x32)) > 4;
    }
    if (local) atomicAdd(bad, local);
}

}  // namespace gk
