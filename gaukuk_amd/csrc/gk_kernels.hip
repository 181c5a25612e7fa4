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
// Waves never exchange data, so there is no LDS and no barrier; x neighbours move through
// cross-lane permutes.
#include "gk_device.hpp"

namespace gk {

__device__ __forceinline__ unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }
__device__ __forceinline__ double bitsd(unsigned long long b) { return __longlong_as_double((long long)b); }

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

template <class R, class Lim, bool THREE>
__global__ __launch_bounds__(kBlock) void stage_kernel(const StageArgs<R> a) {
    DevClock* clk = a.clock;
    if (!clk->active) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = xcd_logical_block(blockIdx.x, gridDim.x) * (kBlock / kWave) + (threadIdx.x >> 6);
    if (w >= a.nwaves) return;
    const int strip = w % a.strips, seg = w / a.strips;
    const int c = strip * kStrip + lane - kHalo;
    const int j0 = seg * a.seg_rows;
    const int j1 = min(j0 + a.seg_rows, a.ny);
    const bool owner = lane >= kHalo && lane < kWave - kHalo && c < a.nx;

    const double dt = clk->dt;
    const R dtdx = R(dt * a.dxRec), dtdy = R(dt * a.dyRec), h = R(0.5 * dt);
    const long long P = a.plane;
    // forward-Euler ping-pong: the current state alternates between the two buffers
    const R* in = a.in;
    R* out = a.out;
    if (a.pingpong && clk->cur) { in = a.out; out = const_cast<R*>(a.in); }

    int kx;
    const int cs = bc_remap(min(c, a.nx + 1), a.nx, a.bc[0], a.bc[1], kx);
    const int cown = min(c, a.nx - 1);

    R W0[5], W1[5], W2[5];  // primitive rows r-2, r-1, r (ring, renamed by the unrolled loop)
    R pre[5];               // next input row, raw
    R wLy[5], Gp[5];        // left state of the next y face; lower y-face flux of the update row
#pragma unroll
    for (int n = 0; n < 5; n++) { W0[n] = W1[n] = W2[n] = wLy[n] = Gp[n] = R(0); }
    R cacc = R(0);

    auto raw_load = [&](int r, R u[5]) {
        int ky;
        const int rs = bc_remap(r, a.ny, a.bc[2], a.bc[3], ky);
        const R* p = in + (long long)rs * a.pitch + cs;
#pragma unroll
        for (int n = 0; n < 5; n++) u[n] = p[n * P];
    };

    const int rlast = j1 + 1;
    raw_load(j0 - 2, pre);

    auto body = [&](int r, R (&A)[5], R (&B)[5], R (&Cn)[5]) {
        R u[5];
#pragma unroll
        for (int n = 0; n < 5; n++) u[n] = pre[n];
        if (r < rlast) raw_load(r + 1, pre);

        // operands of the cell update of row r-2 (owner lanes), issued before the flux math
        const int ur = r - 2;
        const bool upd = ur >= j0;
        R bs[5], ut[5];
        if (upd && owner) {
            const long long o = (long long)ur * a.pitch + cown;
            if (THREE) {
#pragma unroll
                for (int n = 0; n < 5; n++) { bs[n] = a.base[o + n * P]; ut[n] = in[o + n * P]; }
            } else {
#pragma unroll
                for (int n = 0; n < 5; n++) bs[n] = in[o + n * P];
            }
        }

        // row r: boundary transform, gravity half-kick on U^n, ConsToPrim
        if (a.kick_in) gravity_kick(u, a.gx, a.gy, a.gz, h);
        {
            int ky;
            bc_remap(r, a.ny, a.bc[2], a.bc[3], ky);
            bc_transform(u, 1, kx);
            bc_transform(u, 2, ky);
        }
        a.eos.cons_to_prim(u, Cn);

        // y face r-1 (between rows r-2 and r-1): left state from the slope of row r-2 (kept in
        // wLy), right state from the slope of row r-1.
        R G[5];
        if (r >= j0) {
            R wR[5], wLn[5];
#pragma unroll
            for (int n = 0; n < 5; n++) {
                const R s = Lim::slope(B[n] - A[n], Cn[n] - B[n]);
                wR[n] = B[n] - R(0.5) * s;
                wLn[n] = B[n] + R(0.5) * s;
            }
            if (r >= j0 + 1) {
                // direction VLY rotation (src/flux/hllc.cpp:25-27): (den, vy, vz, vx, p)
                const R l[5] = {wLy[0], wLy[2], wLy[3], wLy[1], wLy[4]};
                const R rr[5] = {wR[0], wR[2], wR[3], wR[1], wR[4]};
                R f[5];
                RiemannPolicy::flux(a.eos, l, rr, f);
                G[0] = f[0]; G[2] = f[1]; G[3] = f[2]; G[1] = f[3]; G[4] = f[4];
            }
#pragma unroll
            for (int n = 0; n < 5; n++) wLy[n] = wLn[n];
        }

        // update of row r-2 (primitives in A): x faces c and c+1 through neighbour lanes
        if (upd) {
            R F[5], Fn[5];
            {
                R wR[5], wLo[5], wL[5];
#pragma unroll
                for (int n = 0; n < 5; n++) {
                    const R am = __shfl_up(A[n], 1), ap = __shfl_down(A[n], 1);
                    const R s = Lim::slope(A[n] - am, ap - A[n]);
                    wR[n] = A[n] - R(0.5) * s;   // right state of face c
                    wLo[n] = A[n] + R(0.5) * s;  // left state of face c+1
                }
#pragma unroll
                for (int n = 0; n < 5; n++) wL[n] = __shfl_up(wLo[n], 1);
                RiemannPolicy::flux(a.eos, wL, wR, F);
#pragma unroll
                for (int n = 0; n < 5; n++) Fn[n] = __shfl_down(F[n], 1);
            }
            if (owner) {
                // the base of every stage is U^n after the first gravity half-kick
                if (THREE ? a.kick_base : a.kick_in) gravity_kick(bs, a.gx, a.gy, a.gz, h);
                R res[5];
#pragma unroll
                for (int n = 0; n < 5; n++) {
                    const R dF = (Fn[n] - F[n]) * dtdx + (G[n] - Gp[n]) * dtdy;
                    if (THREE) res[n] = a.c1 * bs[n] + a.c2 * ut[n] - a.c3 * dF;
                    else res[n] = a.c1 * bs[n] - a.c2 * dF;
                }
                a.eos.floors(res);
                if (a.kick_out) gravity_kick(res, a.gx, a.gy, a.gz, h);
                const long long o = (long long)ur * a.pitch + c;
#pragma unroll
                for (int n = 0; n < 5; n++) out[o + n * P] = res[n];
                if (a.want_cmax) cacc = smax(cacc, a.eos.wave_speed(res));
            }
        }
        if (r >= j0 + 1) {
#pragma unroll
            for (int n = 0; n < 5; n++) Gp[n] = G[n];
        }
    };

    for (int r = j0 - 2;;) {
        body(r, W0, W1, W2); if (++r > rlast) break;
        body(r, W1, W2, W0); if (++r > rlast) break;
        body(r, W2, W0, W1); if (++r > rlast) break;
    }

    if (a.want_cmax) {
        const double m = wave_max((double)cacc);
        if (lane == 0) atomicMax(&clk->cmax_bits[clk->parity ^ 1], dbits(m));
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
    if ((threadIdx.x & 63) == 0) atomicMax(&clk->cmax_bits[clk->parity], dbits(m));
}

// Sim::Advance prologue (src/sim.cpp:70-71): tNext = t + dtoutput, dtUntilOutput = dtoutput.
__global__ void advance_begin_kernel(DevClock* c, double dtoutput) {
    c->t_next = c->t + dtoutput;
    c->dt_until_output = dtoutput;
    c->active = c->is_continue;
}

// Loop test and dt of one step (src/sim.cpp:72-81).  cmax_bits[parity] holds max(1e-16, cells).
__global__ void step_begin_kernel(DevClock* c) {
    if (!c->active) return;
    const double gap = c->t_next - c->t;
    if (!((gap < 0 ? -gap : gap) > 1e-12 && c->is_continue)) { c->active = 0; return; }
    const double cm = bitsd(c->cmax_bits[c->parity]);
    double dt = c->cfl * c->drmin / cm;
    dt = smin(dt, c->dt_until_output);
    c->dt = dt;
    c->cmax_bits[c->parity ^ 1] = dbits(1e-16);
}

// Loop tail (src/sim.cpp:93-98) plus the step log entry (the reference prints it, :100-103).
__global__ void step_end_kernel(DevClock* c, StepRecord* log, int cap, int toggle_cur) {
    if (!c->active) return;
    c->t += c->dt;
    c->step++;
    c->dt_until_output = c->t_next - c->t;
    if (c->stepmax > 0 && c->step >= c->stepmax) c->is_continue = 0;
    StepRecord rec;
    rec.t = c->t; rec.dt = c->dt; rec.step = c->step; rec.pad = 0;
    log[(c->step - 1) % cap] = rec;
    c->parity ^= 1;
    if (toggle_cur) c->cur ^= 1;
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

}  // namespace gk
