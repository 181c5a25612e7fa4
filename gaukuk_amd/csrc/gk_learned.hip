// gk_learned.hip -- learned flux correction: the MLP's inference, gradient and Adam kernels, the
// training-data pipeline (flux recording, restriction, face features and targets) and the
// corrected FP32 stage kernel.  Included at the end of gk_capi.hip (one shared library).
//
// None of this has a counterpart in the reference, whose coarse flux is plain HLLC
// (src/flux/hllc.cpp:22-103).  The correction is defined as follows.
//   target   Over one coarse step of length D from the restricted fine state, the fine run's
//            time- and face-averaged flux through a coarse face minus the coarse scheme's
//            RK-weighted flux through it.  Adding the target to the coarse fluxes reproduces the
//            restricted fine state exactly (the conservative-average identity).
//   frame    Per face: normal n, tangent t; densities rho, velocities v_n, v_t, pressure p of the
//            four cells across the face.  Features and targets live in the frame moving
//            tangentially with vt_bar (the scheme is Galilean invariant along the face) and are
//            scaled by the face means (rho_bar, c_bar), so they are dimensionless.
//   features f[4k + 0..3] = rho_k / rho_bar - 1, v_n,k / c_bar, (v_t,k - vt_bar) / c_bar,
//            p_k / p_bar - 1, k = 0..3 for cells i-2..i+1; means over cells i-1, i.
//   outputs  (mass, m_n, m_t', E') flux / (rho c, rho c^2, rho c^2, rho c^3).
//   bound    each output is clipped to +-kappa * D_q, D_q the upwind (Roe-wave) dissipation of the
//            jump between the two PLM face states, so the correction can remove at most the
//            fraction kappa of the scheme's dissipation (unbounded, an anti-diffusive correction
//            makes the coarse run blow up).  Training targets are clipped the same way.
//   fallback plain HLLC where max_k |p_k+1 - p_k| / min(p_k, p_k+1) > 0.1 over the stencil
//            (shock sensor), or where a one-face share U_L - 4 nu dF, U_R + 4 nu dF of either
//            neighbour (nu = dt / dx) loses positive density or pressure.
#include "gk_mlp.hpp"

namespace gk {

// ---------------------------------------------------------------------------------------------
// Tiles through LDS: a D-layout tile (lane (g, s) holds rows 4g+i of column s) is written as a
// 16 x 16 BF16 matrix and read back so that lane (g, x) holds row x, columns 4g..4g+3, the operand
// layout of a product that contracts over the tile's columns.
constexpr int kTp = 20;  // row pitch of a transpose tile (BF16 elements)
__device__ __forceinline__ void tile_put(unsigned short* t, mlp::bf16x4 v, int lane) {
    const int g = lane >> 4, s = lane & 15;
#pragma unroll
    for (int i = 0; i < 4; i++) t[(4 * g + i) * kTp + s] = (unsigned short)v[i];
}
__device__ __forceinline__ mlp::bf16x4 tile_get(const unsigned short* t, int lane) {
    return mlp::lds4(&t[(lane & 15) * kTp + 4 * (lane >> 4)]);
}

__device__ __forceinline__ long long sample_index(unsigned long long seed, long long iter, long long slot, long long n) {
    const unsigned long long h = mix64(seed ^ mix64((unsigned long long)iter * 0x9E3779B97F4A7C15ull + (unsigned long long)slot));
    return (long long)(h % (unsigned long long)n);
}

// Scaled BF16 inputs of lane (g, s): features 4g..4g+3 of sample `idx`.
__device__ __forceinline__ mlp::bf16x4 load_inputs(const float* X, long long idx, const float* insc, int g) {
    const float4 xv = *reinterpret_cast<const float4*>(X + idx * mlp::kIn + 4 * g);
    mlp::f32x4 xs = {xv.x * insc[0], xv.y * insc[1], xv.z * insc[2], xv.w * insc[3]};
    return mlp::pack4(xs);
}

// ---------------------------------------------------------------------------------------------
// Inference on a sample array: Y[k][0..3] = network output (before out_scale) of X[k][0..15].
__global__ __launch_bounds__(256) void mlp_forward_kernel(const float* params, const float* scales, const float* X,
                                                          long long n, float* Y) {
    __shared__ mlp::Smem s;
    const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, r = lane & 15;
    mlp::stage(s, params, scales, scales + mlp::kIn, tid, 256);
    __syncthreads();
    mlp::Regs w;
    mlp::load_regs(w, s, lane);
    float insc[4];
#pragma unroll
    for (int i = 0; i < 4; i++) insc[i] = s.in_scale[4 * g + i];
    const long long ntiles = (n + 15) / 16;
    for (long long t = (long long)blockIdx.x * 4 + (tid >> 6); t < ntiles; t += (long long)gridDim.x * 4) {
        const long long slot = t * 16 + r;
        const long long idx = slot < n ? slot : n - 1;
        const mlp::f32x4 y = mlp::forward_tile(w, load_inputs(X, idx, insc, g));
        if (g == 0 && slot < n) *reinterpret_cast<float4*>(Y + slot * 4) = make_float4(y[0], y[1], y[2], y[3]);
    }
}

// ---------------------------------------------------------------------------------------------
// Gradient of L = sum_k wt_k sum_o (y_ko - t_ko / out_scale_o)^2 * inv_norm over one batch.
// Each workgroup (4 waves) handles 4 * tiles_per_wave tiles of 16 samples and writes its partial
// gradient (fixed summation order, so the result is deterministic) and partial loss.
struct TrainArgs {
    const float* params;
    const float* scales;   // in_scale[16], out_scale[4]
    const float* X;
    const float* T;
    const float* Wt;
    long long n;
    int sequential;        // 1: slot k is sample k (slots >= n weigh 0); 0: hashed draw
    int tiles_per_wave;
    unsigned long long seed;
    long long iter;
    float inv_norm;        // 1 / (4 * batch)
    float* partial;        // [gridDim.x][kParams]
    float* loss_partial;   // [gridDim.x]
};

__global__ __launch_bounds__(256) void mlp_grad_kernel(const TrainArgs a) {
    __shared__ mlp::Smem s;
    __shared__ mlp::SmemT st;
    __shared__ unsigned short tb[4][18][16 * kTp];
    __shared__ float gsum[mlp::kParams];
    __shared__ float lsum[4];
    using mlp::bf16x4;
    using mlp::f32x4;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, g = lane >> 4, r = lane & 15;
    mlp::stage(s, a.params, a.scales, a.scales + mlp::kIn, tid, 256);
    mlp::stage_t(st, a.params, tid, 256);
    __syncthreads();
    mlp::Regs w;
    mlp::load_regs(w, s, lane);
    bf16x4 w2t[4][4], w3t[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
#pragma unroll
        for (int kb = 0; kb < 4; kb++) w2t[t][kb] = mlp::lds4(&st.w2t[(16 * t + r) * mlp::kP64 + 16 * kb + 4 * g]);
        w3t[t] = mlp::lds4(&st.w3t[(16 * t + r) * mlp::kP16 + 4 * g]);
    }
    float insc[4], outinv[4];
#pragma unroll
    for (int i = 0; i < 4; i++) { insc[i] = s.in_scale[4 * g + i]; outinv[i] = 1.0f / s.out_scale[i]; }
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    f32x4 gw2[4][4], gw1[4], gw3[4], gb1[4], gb2[4], gb3 = zero;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        gw1[t] = zero; gw3[t] = zero; gb1[t] = zero; gb2[t] = zero;
#pragma unroll
        for (int u = 0; u < 4; u++) gw2[t][u] = zero;
    }
    float lacc = 0.f;
    unsigned short* T = &tb[wv][0][0];
    constexpr int TS = 16 * kTp;

    for (int q = 0; q < a.tiles_per_wave; q++) {
        const long long slot = (((long long)blockIdx.x * 4 + wv) * a.tiles_per_wave + q) * 16 + r;
        bool live = true;
        long long idx;
        if (a.sequential) { live = slot < a.n; idx = live ? slot : 0; }
        else idx = sample_index(a.seed, a.iter, slot, a.n);
        const bf16x4 x = load_inputs(a.X, idx, insc, g);
        // forward, keeping the pre-activations for the ReLU masks
        f32x4 z1[4], z2[4];
        bf16x4 h1[4], h2[4];
        const f32x4 y = mlp::forward_tile_keep(w, x, z1, z2, h1, h2);
        // output error (lanes 0..15 hold the 4 outputs of sample r)
        f32x4 dy = zero;
        if (g == 0) {
            const float4 tv = *reinterpret_cast<const float4*>(a.T + idx * 4);
            const float wt = live ? a.Wt[idx] : 0.f;
            const float tt[4] = {tv.x, tv.y, tv.z, tv.w};
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float e = y[i] - tt[i] * outinv[i];
                lacc += wt * e * e;
                dy[i] = 2.f * wt * e * a.inv_norm;
            }
        }
        const bf16x4 dyq = mlp::pack4(dy);
        // back through the layers: dH2 = W3^T dY, dZ2 = dH2 (Z2 > 0), dH1 = W2^T dZ2, dZ1 = ...
        f32x4 dz2[4], dz1[4];
        bf16x4 dz2q[4], dz1q[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            f32x4 d = mlp::mfma(w3t[t], dyq, zero);
#pragma unroll
            for (int i = 0; i < 4; i++) d[i] = z2[t][i] > 0.f ? d[i] : 0.f;
            dz2[t] = d;
            dz2q[t] = mlp::pack4(d);
        }
#pragma unroll
        for (int t = 0; t < 4; t++) {
            f32x4 d = zero;
#pragma unroll
            for (int kb = 0; kb < 4; kb++) d = mlp::mfma(w2t[t][kb], dz2q[kb], d);
#pragma unroll
            for (int i = 0; i < 4; i++) d[i] = z1[t][i] > 0.f ? d[i] : 0.f;
            dz1[t] = d;
            dz1q[t] = mlp::pack4(d);
        }
#pragma unroll
        for (int t = 0; t < 4; t++) { gb1[t] += dz1[t]; gb2[t] += dz2[t]; }
        gb3 += dy;
        // weight gradients contract over the samples: transpose the operands through LDS
        tile_put(T, x, lane);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            tile_put(T + (1 + t) * TS, h1[t], lane);
            tile_put(T + (5 + t) * TS, h2[t], lane);
            tile_put(T + (9 + t) * TS, dz2q[t], lane);
            tile_put(T + (13 + t) * TS, dz1q[t], lane);
        }
        tile_put(T + 17 * TS, dyq, lane);
        __syncthreads();
        const bf16x4 tx = tile_get(T, lane), tdy = tile_get(T + 17 * TS, lane);
        bf16x4 th1[4], th2[4], tdz2[4], tdz1[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            th1[t] = tile_get(T + (1 + t) * TS, lane);
            th2[t] = tile_get(T + (5 + t) * TS, lane);
            tdz2[t] = tile_get(T + (9 + t) * TS, lane);
            tdz1[t] = tile_get(T + (13 + t) * TS, lane);
        }
        __syncthreads();
#pragma unroll
        for (int ot = 0; ot < 4; ot++) {
#pragma unroll
            for (int ht = 0; ht < 4; ht++) gw2[ot][ht] = mlp::mfma(tdz2[ot], th1[ht], gw2[ot][ht]);
            gw1[ot] = mlp::mfma(tdz1[ot], tx, gw1[ot]);
            gw3[ot] = mlp::mfma(tdy, th2[ot], gw3[ot]);
        }
    }
    // bias sums over the 16 samples (lanes) of each row group; loss over the wave
#pragma unroll
    for (int m = 1; m < 16; m <<= 1) {
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                gb1[t][i] += __shfl_xor(gb1[t][i], m);
                gb2[t][i] += __shfl_xor(gb2[t][i], m);
            }
#pragma unroll
        for (int i = 0; i < 4; i++) gb3[i] += __shfl_xor(gb3[i], m);
    }
    for (int m = 1; m < 64; m <<= 1) lacc += __shfl_xor(lacc, m);
    // the four waves add into gsum in a fixed order
    for (int ws = 0; ws < 4; ws++) {
        if (wv == ws) {
            auto put = [&](int k, float v) { gsum[k] = (ws == 0 ? 0.f : gsum[k]) + v; };
#pragma unroll
            for (int ot = 0; ot < 4; ot++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int row = 16 * ot + 4 * g + i;
#pragma unroll
                    for (int ht = 0; ht < 4; ht++) put(mlp::oW2 + row * mlp::kHid + 16 * ht + r, gw2[ot][ht][i]);
                    put(mlp::oW1 + row * mlp::kIn + r, gw1[ot][i]);
                    if (g == 0) put(mlp::oW3 + i * mlp::kHid + 16 * ot + r, gw3[ot][i]);
                    if (r == 0) { put(mlp::ob1 + row, gb1[ot][i]); put(mlp::ob2 + row, gb2[ot][i]); }
                }
            if (r == 0 && g == 0)
#pragma unroll
                for (int i = 0; i < 4; i++) put(mlp::ob3 + i, gb3[i]);
            if (lane == 0) lsum[ws] = lacc;
        }
        __syncthreads();
    }
    for (int k = tid; k < mlp::kParams; k += 256) a.partial[(long long)blockIdx.x * mlp::kParams + k] = gsum[k];
    if (tid == 0) a.loss_partial[blockIdx.x] = ((lsum[0] + lsum[1]) + lsum[2]) + lsum[3];
}

// Sum of the block partials (block order) and, if `apply`, one Adam step (Kingma & Ba 2015,
// Algorithm 1, bias-corrected): bc1 = 1 - b1^t, bc2 = 1 - b2^t.
__global__ __launch_bounds__(256) void adam_kernel(float* p, float* m, float* v, const float* partial, int nblk,
                                                   const float* loss_partial, float inv_norm, float* loss_out,
                                                   float* grad_out, int apply, float lr, float b1, float b2, float eps,
                                                   float bc1, float bc2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < mlp::kParams) {
        float gsum = 0.f;
        for (int b = 0; b < nblk; b++) gsum += partial[(long long)b * mlp::kParams + k];
        if (grad_out) grad_out[k] = gsum;
        if (apply) {
            const float mk = b1 * m[k] + (1.f - b1) * gsum;
            const float vk = b2 * v[k] + (1.f - b2) * gsum * gsum;
            m[k] = mk;
            v[k] = vk;
            p[k] -= lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
        }
    }
    if (k == 0 && loss_out) {
        float l = 0.f;
        for (int b = 0; b < nblk; b++) l += loss_partial[b];
        *loss_out = l * inv_norm;
    }
}

// Weighted mean squares of the inputs and targets of a data set (per-block FP64 partials).
__global__ __launch_bounds__(256) void dataset_stats_kernel(const float* X, const float* T, const float* Wt, long long n,
                                                            double* partial) {
    __shared__ double sh[21][256];
    double acc[21];
#pragma unroll
    for (int q = 0; q < 21; q++) acc[q] = 0.0;
    for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long long)gridDim.x * 256) {
        const double wt = Wt[k];
        if (wt == 0.0) continue;
#pragma unroll
        for (int q = 0; q < 16; q++) { const double x = X[k * 16 + q]; acc[q] += wt * x * x; }
#pragma unroll
        for (int q = 0; q < 4; q++) { const double t = T[k * 4 + q]; acc[16 + q] += wt * t * t; }
        acc[20] += wt;
    }
#pragma unroll
    for (int q = 0; q < 21; q++) sh[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
#pragma unroll
            for (int q = 0; q < 21; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 21) partial[blockIdx.x * 21 + threadIdx.x] = sh[threadIdx.x][0];
}

// ---------------------------------------------------------------------------------------------
// Faces of a grid: x faces i = 0..nx at rows j (index j*(nx+1) + i), then y faces j = 0..ny at
// columns i (index (nx+1)*ny + j*nx + i).  Face i of an axis lies between cells i-1 and i.
struct FaceGeom {
    int nx, ny;
    __device__ long long nxf() const { return (long long)(nx + 1) * ny; }
    __device__ long long total() const { return nxf() + (long long)nx * (ny + 1); }
    __device__ void decode(long long k, bool& yface, int& i, int& j) const {
        yface = k >= nxf();
        if (!yface) { j = (int)(k / (nx + 1)); i = (int)(k - (long long)j * (nx + 1)); }
        else { const long long k2 = k - nxf(); j = (int)(k2 / nx); i = (int)(k2 - (long long)j * nx); }
    }
};

// Conserved state of cell (i, j) of a stage input, ghost cells by the boundary rules: the index
// remap and transforms of the stage kernels (x rule, then y rule).
template <class R>
__device__ __forceinline__ void load_cell(const R* u, long long plane, int pitch, int nx, int ny, const int bc[4],
                                          int i, int j, R out[5]) {
    int kx, ky;
    const int si = bc_remap(min(max(i, -2), nx + 1), nx, bc[0], bc[1], kx);
    const int sj = bc_remap(min(max(j, -2), ny + 1), ny, bc[2], bc[3], ky);
    const long long o = (long long)sj * pitch + si;
#pragma unroll
    for (int n = 0; n < 5; n++) out[n] = u[o + n * plane];
    bc_transform(out, 1, kx);
    bc_transform(out, 2, ky);
}

// PLM (van Leer) + Riemann flux at a face from the primitive states of its four cells
// (w[0..3] = cells i-2..i+1 along the face normal), in the lab frame.
template <class R>
__device__ __forceinline__ void face_flux(const EosPolicy<R>& eos, const R w[4][5], bool yface, R F[5]) {
    R wl[5], wr[5];
#pragma unroll
    for (int n = 0; n < 5; n++) {
        const R sl = VanLeer::slope(w[1][n] - w[0][n], w[2][n] - w[1][n]);
        const R sr = VanLeer::slope(w[2][n] - w[1][n], w[3][n] - w[2][n]);
        wl[n] = w[1][n] + R(0.5) * sl;
        wr[n] = w[2][n] - R(0.5) * sr;
    }
    if (!yface) {
        RiemannPolicy::flux(eos, wl, wr, F);
    } else {  // direction VLY: (den, vy, vz, vx, p), src/flux/hllc.cpp:25-27
        const R l[5] = {wl[0], wl[2], wl[3], wl[1], wl[4]};
        const R rr[5] = {wr[0], wr[2], wr[3], wr[1], wr[4]};
        R f[5];
        RiemannPolicy::flux(eos, l, rr, f);
        F[0] = f[0]; F[2] = f[1]; F[3] = f[2]; F[1] = f[3]; F[4] = f[4];
    }
}

template <class R>
__device__ __forceinline__ void face_stencil(const R* u, long long plane, int pitch, int nx, int ny, const int bc[4],
                                             const EosPolicy<R>& eos, bool yface, int i, int j, R w[4][5]) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
        R c[5];
        if (!yface) load_cell(u, plane, pitch, nx, ny, bc, i - 2 + k, j, c);
        else load_cell(u, plane, pitch, nx, ny, bc, i, j - 2 + k, c);
        eos.cons_to_prim(c, w[k]);
    }
}

// Recording: acc[c][face] += weight * dt * F_c for c = mass, m_x, m_y, E, at every face, from the
// stage input (weight = the stage's share of the step's flux).  Inactive steps add nothing.
template <class R>
__global__ __launch_bounds__(256) void flux_record_kernel(const R* in0, const R* in1, int pingpong, long long plane,
                                                          int pitch, int nx, int ny, int4 bcv, EosPolicy<R> eos,
                                                          const DevClock* clk, double weight, double* acc) {
    if (!clk->active) return;
    const FaceGeom fg{nx, ny};
    const long long nf = fg.total();
    const R* in = (pingpong && clk->cur) ? in1 : in0;
    const int bc[4] = {bcv.x, bcv.y, bcv.z, bcv.w};
    const double wdt = weight * clk->dt;
    for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < nf; k += (long long)gridDim.x * 256) {
        bool yface;
        int i, j;
        fg.decode(k, yface, i, j);
        R w[4][5], F[5];
        face_stencil(in, plane, pitch, nx, ny, bc, eos, yface, i, j, w);
        face_flux(eos, w, yface, F);
        acc[k] += wdt * (double)F[0];
        acc[nf + k] += wdt * (double)F[1];
        acc[2 * nf + k] += wdt * (double)F[2];
        acc[3 * nf + k] += wdt * (double)F[4];
    }
}

// Conservative average of f x f blocks of a fine state into a coarse state (fixed order, FP64).
template <class RF, class RC>
__global__ __launch_bounds__(256) void restrict_kernel(const RF* fine, long long fplane, int fpitch, RC* coarse,
                                                       long long cplane, int cpitch, int cnx, int cny, int f) {
    const long long n = (long long)cnx * cny;
    const double inv = 1.0 / ((double)f * f);
    for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < n; k += (long long)gridDim.x * 256) {
        const int J = (int)(k / cnx), I = (int)(k - (long long)J * cnx);
        for (int q = 0; q < 5; q++) {
            double s = 0.0;
            for (int b = 0; b < f; b++)
                for (int a = 0; a < f; a++) s += (double)fine[q * fplane + (long long)(f * J + b) * fpitch + f * I + a];
            coarse[q * cplane + (long long)J * cpitch + I] = (RC)(s * inv);
        }
    }
}

// Reciprocal and root for the learned path's own arithmetic (features, bound): the hardware
// approximations in FP32 (v_rcp_f32, v_sqrt_f32, within 1 ulp), exact operators in FP64.
__device__ __forceinline__ float lrcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double lrcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float lsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double lsqrt(double x) { return __builtin_sqrt(x); }

// Face means, features, sensor: shared by the data pipeline and the corrected stage kernel.
// st[k] = (rho, v_n, v_t, p) of cell k = 0..3 across the face.
template <class R>
struct FaceFrame {
    R rb, pb, cb, vtb;
    R irb, ipb, icb;   // reciprocals of the means
    __device__ __forceinline__ void init(const R st[4][4], R gamma) {
        rb = R(0.5) * (st[1][0] + st[2][0]);
        pb = R(0.5) * (st[1][3] + st[2][3]);
        irb = lrcp(rb);
        ipb = lrcp(pb);
        cb = lsqrt(gamma * pb * irb);
        icb = lrcp(cb);
        vtb = R(0.5) * (st[1][2] + st[2][2]);
    }
    __device__ __forceinline__ void features(const R c[4], R f[4]) const {
        f[0] = c[0] * irb - R(1);
        f[1] = c[1] * icb;
        f[2] = (c[2] - vtb) * icb;
        f[3] = c[3] * ipb - R(1);
    }
    // network output (times out_scale) -> lab-frame (mass, m_n, m_t, E) correction
    __device__ __forceinline__ void to_lab(const R y[4], R d[4]) const {
        const R a = y[0] * rb * cb, b = y[1] * rb * cb * cb, c = y[2] * rb * cb * cb, e = y[3] * rb * cb * cb * cb;
        d[0] = a;
        d[1] = b;
        d[2] = c + vtb * a;
        d[3] = e + vtb * c + R(0.5) * vtb * vtb * a;
    }
    // lab-frame (mass, m_n, m_t, E) flux difference -> network units (before out_scale)
    __device__ __forceinline__ void to_net(const R d[4], R y[4]) const {
        y[0] = d[0] / (rb * cb);
        y[1] = d[1] / (rb * cb * cb);
        y[2] = (d[2] - vtb * d[0]) / (rb * cb * cb);
        y[3] = (d[3] - vtb * d[2] + R(0.5) * vtb * vtb * d[0]) / (rb * cb * cb * cb);
    }
};
// max_k |p_k+1 - p_k| / min(p_k, p_k+1) > 0.1, written without the quotient (p > 0)
template <class R>
__device__ __forceinline__ bool sensor_fires(const R p[4]) {
    bool fire = false;
#pragma unroll
    for (int k = 0; k < 3; k++) fire |= gabs(p[k + 1] - p[k]) > R(0.1) * smin(p[k], p[k + 1]);
    return fire;
}

// The dissipation bound D_q in the units of the network's output (face frame moving with vt_bar,
// fluxes over rho_bar c_bar^k), from the 16 features alone: cell k has rho = 1 + f[4k],
// v_n = f[4k+1], v_t = f[4k+2], p = (1 + f[4k+3]) / gamma in units of (rho_bar, c_bar).  D is the
// upwind dissipation 1/2 sum_k |lambda_k| |alpha_k| |r_k| of the jump between the two PLM face
// states split into the Euler waves (Toro ch. 3: acoustic v_n -+ c, contact and shear v_n), taken
// component by component with arithmetic face averages.  A contact or shear layer at rest gets
// D = 0: HLLC keeps it sharp and there is no dissipation to remove.
template <class R>
__device__ __forceinline__ void dissipation_bound(const R f[16], R gamma, R D[4]) {
    R w[4][4];
    const R ig = lrcp(gamma);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        w[k][0] = R(1) + f[4 * k];
        w[k][1] = f[4 * k + 1];
        w[k][2] = f[4 * k + 2];
        w[k][3] = (R(1) + f[4 * k + 3]) * ig;
    }
    R l[4], r[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        l[q] = w[1][q] + R(0.5) * VanLeer::slope(w[1][q] - w[0][q], w[2][q] - w[1][q]);
        r[q] = w[2][q] - R(0.5) * VanLeer::slope(w[2][q] - w[1][q], w[3][q] - w[2][q]);
    }
    const R rho = R(0.5) * (l[0] + r[0]), vn = R(0.5) * (l[1] + r[1]);
    const R vt = R(0.5) * (l[2] + r[2]), p = R(0.5) * (l[3] + r[3]);
    const R c2 = gamma * p * lrcp(rho), c = lsqrt(c2), ic2 = lrcp(c2);
    const R dr = r[0] - l[0], dvn = r[1] - l[1], dvt = r[2] - l[2], dp = r[3] - l[3];
    const R a1 = R(0.5) * (dp - rho * c * dvn) * ic2, a4 = R(0.5) * (dp + rho * c * dvn) * ic2;
    const R a2 = dr - dp * ic2, a3 = rho * dvt;
    const R H = c2 * lrcp(gamma - R(1)) + R(0.5) * (vn * vn + vt * vt);
    const R w1 = gabs(vn - c) * gabs(a1), w2 = gabs(vn) * gabs(a2);
    const R w3 = gabs(vn) * gabs(a3), w4 = gabs(vn + c) * gabs(a4);
    D[0] = R(0.5) * (w1 + w2 + w4);
    D[1] = R(0.5) * (w1 * gabs(vn - c) + w2 * gabs(vn) + w4 * gabs(vn + c));
    D[2] = R(0.5) * ((w1 + w2 + w4) * gabs(vt) + w3);
    D[3] = R(0.5) * (w1 * gabs(H - vn * c) + w2 * R(0.5) * (vn * vn + vt * vt) + w3 * gabs(vt) +
                     w4 * gabs(H + vn * c));
}
template <class R>
__device__ __forceinline__ R clip(R y, R b) { return smin(smax(y, -b), b); }

// Features, face means and sensor flag of every face of a (coarse) state.
template <class R>
__global__ __launch_bounds__(256) void face_features_kernel(const R* u, long long plane, int pitch, int nx, int ny,
                                                            int4 bcv, EosPolicy<R> eos, float* feat, float* meta) {
    const FaceGeom fg{nx, ny};
    const long long nf = fg.total();
    const int bc[4] = {bcv.x, bcv.y, bcv.z, bcv.w};
    for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < nf; k += (long long)gridDim.x * 256) {
        bool yface;
        int i, j;
        fg.decode(k, yface, i, j);
        R w[4][5], st[4][4];
        face_stencil(u, plane, pitch, nx, ny, bc, eos, yface, i, j, w);
#pragma unroll
        for (int c = 0; c < 4; c++) {
            st[c][0] = w[c][0];
            st[c][1] = yface ? w[c][2] : w[c][1];
            st[c][2] = yface ? w[c][1] : w[c][2];
            st[c][3] = w[c][4];
        }
        FaceFrame<R> fr;
        fr.init(st, eos.gamma);
#pragma unroll
        for (int c = 0; c < 4; c++) {
            R f[4];
            fr.features(st[c], f);
#pragma unroll
            for (int q = 0; q < 4; q++) feat[k * 16 + 4 * c + q] = (float)f[q];
        }
        const R p[4] = {st[0][3], st[1][3], st[2][3], st[3][3]};
        meta[k * 4 + 0] = (float)fr.rb;
        meta[k * 4 + 1] = (float)fr.pb;
        meta[k * 4 + 2] = (float)fr.vtb;
        // periodic grids list the wrap face twice (i = 0 and i = nx): keep one
        const bool dup = (!yface && i == nx && bc[1] == 1) || (yface && j == ny && bc[3] == 1);
        meta[k * 4 + 3] = (sensor_fires(p) || dup) ? 1.f : 0.f;
    }
}

// Targets from the recorded fluxes: the fine accumulators of the f fine faces on each coarse
// face, over f * D, minus the coarse accumulator over D, in network units; appended with the
// features at dataset position `base`.
__global__ __launch_bounds__(256) void face_targets_kernel(const double* facc, int fnx, int fny, const double* cacc,
                                                           int cnx, int cny, int f, double D, double gamma,
                                                           double kappa, const float* feat, const float* meta,
                                                           float* X, float* T, float* Wt, long long base) {
    const FaceGeom cg{cnx, cny}, fg{fnx, fny};
    const long long nc = cg.total(), nfine = fg.total();
    for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < nc; k += (long long)gridDim.x * 256) {
        bool yface;
        int I, J;
        cg.decode(k, yface, I, J);
        double fl[4] = {0, 0, 0, 0};
        for (int r = 0; r < f; r++) {
            const long long fk = !yface ? (long long)(f * J + r) * (fnx + 1) + (long long)f * I
                                        : fg.nxf() + (long long)(f * J) * fnx + (long long)f * I + r;
#pragma unroll
            for (int c = 0; c < 4; c++) fl[c] += facc[c * nfine + fk];
        }
        double lab[4];
#pragma unroll
        for (int c = 0; c < 4; c++) lab[c] = fl[c] / ((double)f * D) - cacc[c * nc + k] / D;
        // lab (mass, m_x, m_y, E) -> face frame (mass, m_n, m_t, E)
        const double d[4] = {lab[0], yface ? lab[2] : lab[1], yface ? lab[1] : lab[2], lab[3]};
        FaceFrame<double> fr;
        fr.rb = meta[k * 4 + 0];
        fr.pb = meta[k * 4 + 1];
        fr.vtb = meta[k * 4 + 2];
        fr.cb = sqrt(gamma * fr.pb / fr.rb);
        double y[4], fe[16], bd[4];
        fr.to_net(d, y);
#pragma unroll
        for (int q = 0; q < 16; q++) fe[q] = feat[k * 16 + q];
        dissipation_bound(fe, gamma, bd);
#pragma unroll
        for (int c = 0; c < 4; c++) y[c] = clip(y[c], kappa * bd[c]);
        bool ok = meta[k * 4 + 3] == 0.f;
#pragma unroll
        for (int c = 0; c < 4; c++) ok = ok && isfinite(y[c]);
        const long long o = base + k;
#pragma unroll
        for (int q = 0; q < 16; q++) X[o * 16 + q] = feat[k * 16 + q];
#pragma unroll
        for (int c = 0; c < 4; c++) T[o * 4 + c] = ok ? (float)y[c] : 0.f;
        Wt[o] = ok ? 1.f : 0.f;
    }
}

// ---------------------------------------------------------------------------------------------
// The corrected FP32 stage: one 512-thread workgroup per 16 x 8 tile of cells.
//   1. the tile and a 2-cell halo -> LDS (conserved after the boundary rules, and primitives)
//   2. PLM + HLLC at the tile's 136 x faces and 144 y faces -> LDS, one face per thread, and from
//      the same four cells each face's network inputs, dissipation bound and sensor
//   3. the MLP on 18 tiles of 16 faces (BF16 MFMA), positivity tests, correction
//   4. flux divergence, RK combine, floors, store, max wave speed
// A tile this size needs 53 KB of LDS and 128 registers per lane, so two workgroups (16 waves)
// share a CU.  (A 16 x 16 tile needed 88 KB, one workgroup per CU, and 544 faces for 512 threads:
// two rounds of every face phase.  At 256^2 the two shapes measured the same per launch:
// the phases are latency chains separated by barriers, not issue-bound.)  Every face is solved
// from the same inputs whatever tile holds it, and the network treats each face (one MFMA column)
// independently, so results do not depend on the tile shape.
// Weights image: BF16 matrices and FP32 biases/scales in the LDS layout, rebuilt when the
// network's parameters or scales change.  kImageCopies replicas (one per block of this kernel):
// every workgroup of a stage reads the image at the same moment, and one copy read by all of
// them queues on the few L2 channels that hold it.
constexpr int kImageCopies = 32;
__global__ __launch_bounds__(256) void mlp_image_kernel(const float* params, const float* scales, mlp::Smem* img) {
    mlp::stage(img[blockIdx.x], params, scales, scales + mlp::kIn, threadIdx.x, 256);
}

struct NetArgs {
    const mlp::Smem* image;           // prebuilt LDS images of the weights (kImageCopies replicas)
    unsigned long long* stats;        // per tile: corrected, sensor fallbacks, positivity fallbacks
    float kappa;                      // dissipation bound factor
    int enabled;
};

constexpr int kLTX = 16, kLTY = 8;         // tile width and height (cells)
constexpr int kLEX = kLTX + 2 * kHalo;     // tile width and height with halo
constexpr int kLEY = kLTY + 2 * kHalo;
constexpr int kLFx = (kLTX + 1) * kLTY;    // x faces per tile (136)
constexpr int kLFy = kLTX * (kLTY + 1);    // y faces per tile (144)
constexpr int kLFaces = kLFx + kLFy;       // 280
constexpr int kLFacesPad = (kLFaces + 15) / 16 * 16;   // whole MLP tiles of 16 faces (288)
constexpr int kLMinSide = 16;              // smallest grid side the corrected stage takes

constexpr int kLThreads = 512;             // 8 waves per tile: one face per thread in every face phase
constexpr int kLWaves = kLThreads / kWave;

// Two workgroups (16 waves) per CU: the register budget is 128 per lane, VGPRs and AGPRs together
// (without the bound the compiler assumes fewer waves and spends more on MFMA accumulators).
template <bool THREE, bool LAST>
__global__ __launch_bounds__(kLThreads) __attribute__((amdgpu_waves_per_eu(4)))
void learned_stage_kernel(const StageArgs<float> a, const NetArgs net) {
    DevClock* clk = a.clock;
    if (!clk->active) return;
    __shared__ mlp::Smem s;
    __shared__ float sw[5][kLEY][kLEX + 1];
    __shared__ float su[5][kLEY][kLEX + 1];
    __shared__ float sf[5][kLFaces];
    __shared__ unsigned long long scount[kLWaves][3];
    __shared__ unsigned short sx[kLFacesPad * mlp::kP16];   // scaled BF16 network inputs per face
    __shared__ float smeta[kLFaces][8];                     // bound D[4], rho_bar, c_bar, vt_bar, sensor
    __shared__ float sy[kLFacesPad][4];                     // network outputs (rows past kLFaces unused)
    __shared__ double swmax[kLWaves];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int tiles_x = (a.nx + kLTX - 1) / kLTX;
    const int i0 = (blockIdx.x % tiles_x) * kLTX, j0 = (blockIdx.x / tiles_x) * kLTY;
    const double dt = clk->dt;
    const float dtdx = float(dt * a.dxRec), dtdy = float(dt * a.dyRec);
    const float* in = a.in;
    float* out = a.out;
    if (a.pingpong && clk->cur) { in = a.out; out = const_cast<float*>(a.in); }
    const float gamma = a.eos.gamma;
    // Every global load of the kernel is issued here, before any result is used, so the
    // workgroup waits out one memory latency: the tile and halo cells (240, one per thread),
    // the combine operand of this thread's cell, and the weights image.
    constexpr int kCells = kLEX * kLEY;
    static_assert(kCells <= kLThreads, "one cell per thread");
    const int kc = min(tid, kCells - 1);
    const int ly = kc / kLEX, lx = kc - ly * kLEX;
    int kx, ky;
    const int si = bc_remap(min(max(i0 + lx - kHalo, -2), a.nx + 1), a.nx, a.bc[0], a.bc[1], kx);
    const int sj = bc_remap(min(max(j0 + ly - kHalo, -2), a.ny + 1), a.ny, a.bc[2], a.bc[3], ky);
    const long long oc = (long long)sj * a.pitch + si;
    float u[5];
#pragma unroll
    for (int n = 0; n < 5; n++) u[n] = in[oc + n * a.plane];
    const int cx = tid % kLTX, cy = min(tid / kLTX, kLTY - 1);
    const int i = i0 + cx, j = j0 + cy;
    const bool mine = tid < kLTX * kLTY && i < a.nx && j < a.ny;
    const long long o = (long long)min(j, a.ny - 1) * a.pitch + min(i, a.nx - 1);
    float bse[5];
#pragma unroll
    for (int n = 0; n < 5; n++) bse[n] = THREE ? a.base[o + n * a.plane] : 0.f;
    mlp::ImageCopy<kLThreads> wimg;
    if (net.enabled) wimg.load(net.image + blockIdx.x % kImageCopies, tid);
    __builtin_amdgcn_sched_barrier(0);   // keep the loads above from being split by waits

    bc_transform(u, 1, kx);   // x rule, then y rule (corner cells are never read)
    bc_transform(u, 2, ky);
    if (tid < kCells) {
        float w[5];
        a.eos.cons_to_prim(u, w);
#pragma unroll
        for (int n = 0; n < 5; n++) { su[n][ly][lx] = u[n]; sw[n][ly][lx] = w[n]; }
    }
    if (net.enabled) wimg.store(s, tid);
    __syncthreads();

    // face f: x faces fx = 0..16 of row fy (between cells fx-1 and fx), then y faces fy = 0..8 of
    // column fx (between rows fy-1 and fy)
    auto face_cell = [&](int f, int k, int& cy, int& cx) {  // cell k = 0..3 across face f (LDS coords)
        if (f < kLFx) { const int fy = f / (kLTX + 1), fx = f - fy * (kLTX + 1); cy = fy + kHalo; cx = fx + k; }
        else { const int f2 = f - kLFx, fy = f2 / kLTX, fx = f2 - fy * kLTX; cy = fy + k; cx = fx + kHalo; }
    };
    // One face per thread: PLM + HLLC flux and, with a network, from the same four cells the face
    // frame, features, dissipation bound and sensor (M1)
    for (int f = tid; f < kLFaces; f += kLThreads) {
        const bool yface = f >= kLFx;
        float w[4][5], F[5];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            int cy, cx;
            face_cell(f, k, cy, cx);
#pragma unroll
            for (int n = 0; n < 5; n++) w[k][n] = sw[n][cy][cx];
        }
        face_flux(a.eos, w, yface, F);
#pragma unroll
        for (int n = 0; n < 5; n++) sf[n][f] = F[n];
        if (net.enabled) {
            float st[4][4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                st[k][0] = w[k][0];
                st[k][1] = yface ? w[k][2] : w[k][1];
                st[k][2] = yface ? w[k][1] : w[k][2];
                st[k][3] = w[k][4];
            }
            FaceFrame<float> fr;
            fr.init(st, gamma);
            float fa[16], bd[4];
#pragma unroll
            for (int k = 0; k < 4; k++) fr.features(st[k], fa + 4 * k);
            dissipation_bound(fa, gamma, bd);
            const float p[4] = {st[0][3], st[1][3], st[2][3], st[3][3]};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const mlp::f32x4 v = {fa[4 * q] * s.in_scale[4 * q], fa[4 * q + 1] * s.in_scale[4 * q + 1],
                                      fa[4 * q + 2] * s.in_scale[4 * q + 2], fa[4 * q + 3] * s.in_scale[4 * q + 3]};
                *reinterpret_cast<mlp::bf16x4*>(&sx[f * mlp::kP16 + 4 * q]) = mlp::pack4(v);
            }
            *reinterpret_cast<float4*>(&smeta[f][0]) = make_float4(bd[0], bd[1], bd[2], bd[3]);
            *reinterpret_cast<float4*>(&smeta[f][4]) = make_float4(fr.rb, fr.cb, fr.vtb, sensor_fires(p) ? 1.f : 0.f);
        }
    }
    __syncthreads();

    if (net.enabled) {
        // M2, per wave: the network on tiles of 16 faces (lane (g, s) feeds inputs 4g..4g+3 of face s)
        {
            mlp::Regs wr;
            mlp::load_regs(wr, s, lane);
            const int g = lane >> 4, r = lane & 15;
            for (int q = wv; q * 16 < kLFaces; q += kLWaves) {
                const int f = q * 16 + r;
                const mlp::f32x4 y = mlp::forward_tile(wr, mlp::lds4(&sx[f * mlp::kP16 + 4 * g]));
                if (g == 0) *reinterpret_cast<float4*>(&sy[f][0]) = make_float4(y[0], y[1], y[2], y[3]);
            }
        }
        __syncthreads();
        // M3, one face per thread: bound, lab frame, fallbacks, corrected flux
        unsigned long long n_corr = 0, n_sensor = 0, n_pos = 0;
        for (int f = tid; f < kLFaces; f += kLThreads) {
            const bool yface = f >= kLFx;
            const float4 bd = *reinterpret_cast<const float4*>(&smeta[f][0]);
            const float4 mt = *reinterpret_cast<const float4*>(&smeta[f][4]);
            const float4 yv = *reinterpret_cast<const float4*>(&sy[f][0]);
            FaceFrame<float> fr;
            fr.rb = mt.x;
            fr.cb = mt.y;
            fr.vtb = mt.z;
            const float yh[4] = {clip(yv.x * s.out_scale[0], net.kappa * bd.x), clip(yv.y * s.out_scale[1], net.kappa * bd.y),
                                 clip(yv.z * s.out_scale[2], net.kappa * bd.z), clip(yv.w * s.out_scale[3], net.kappa * bd.w)};
            float d[4];
            fr.to_lab(yh, d);
            float d5[5];
            d5[0] = d[0];
            d5[1] = yface ? d[2] : d[1];
            d5[2] = yface ? d[1] : d[2];
            d5[3] = 0.f;
            d5[4] = d[3];
            const bool fired = mt.w != 0.f;
            // one-face shares of the two neighbours (cells 1 and 2)
            const float nu4 = 4.f * (yface ? dtdy : dtdx);
            int ly, lx, hy, hx;
            face_cell(f, 1, ly, lx);
            face_cell(f, 2, hy, hx);
            float uL[5], uR[5];
#pragma unroll
            for (int n = 0; n < 5; n++) {
                uL[n] = su[n][ly][lx] - nu4 * d5[n];
                uR[n] = su[n][hy][hx] + nu4 * d5[n];
            }
            // p > 0 <=> rho E - |m|^2 / 2 > 0 for rho > 0
            const float qL = uL[0] * uL[4] - 0.5f * (uL[1] * uL[1] + uL[2] * uL[2] + uL[3] * uL[3]);
            const float qR = uR[0] * uR[4] - 0.5f * (uR[1] * uR[1] + uR[2] * uR[2] + uR[3] * uR[3]);
            bool ok = uL[0] > 0.f && uR[0] > 0.f && qL > 0.f && qR > 0.f;
#pragma unroll
            for (int n = 0; n < 5; n++) ok = ok && __builtin_isfinite(d5[n]);
            if (!fired && ok) {
#pragma unroll
                for (int n = 0; n < 5; n++) sf[n][f] += d5[n];
            }
            // counts: each grid face once (faces on a tile's high edge belong to the next tile)
            const int fl = yface ? (f - kLFx) / kLTX : f % (kLTX + 1);
            const bool own = fl < (yface ? kLTY : kLTX) && (yface ? j0 + fl < a.ny : i0 + fl < a.nx);
            if (own) {
                if (fired) n_sensor++;
                else if (ok) n_corr++;
                else n_pos++;
            }
        }
        // face counts: wave sums to LDS (a thread holds at most one face, so a wave's count is
        // the population of a ballot); after the barrier one thread adds them to this tile's slot
        // (plain read-modify-write: a tile belongs to one workgroup, launches are ordered)
        static_assert(kLFaces <= kLThreads, "one face per thread");
        n_corr = __popcll(__ballot(n_corr != 0));
        n_sensor = __popcll(__ballot(n_sensor != 0));
        n_pos = __popcll(__ballot(n_pos != 0));
        if (lane == 0) { scount[wv][0] = n_corr; scount[wv][1] = n_sensor; scount[wv][2] = n_pos; }
        __syncthreads();
        if (net.stats && tid == 0) {
            unsigned long long* slot = net.stats + 3 * (long long)blockIdx.x;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                unsigned long long t = 0;
                for (int w = 0; w < kLWaves; w++) t += scount[w][k];
                slot[k] += t;
            }
        }
    }

    float cacc = 0.f;
    if (mine) {
        const int fxl = cy * (kLTX + 1) + cx, fyl = kLFx + cy * kLTX + cx;
        float res[5];
#pragma unroll
        for (int n = 0; n < 5; n++) {
            const float dF = (sf[n][fxl + 1] - sf[n][fxl]) * dtdx + (sf[n][fyl + kLTX] - sf[n][fyl]) * dtdy;
            const float ut = su[n][cy + kHalo][cx + kHalo];
            if (THREE) res[n] = a.c1 * bse[n] + a.c2 * ut - a.c3 * dF;
            else res[n] = a.c1 * ut - a.c2 * dF;
        }
        const float rinv = a.eos.floors(res);
        if (LAST) cacc = a.eos.wave_speed(res, rinv);
#pragma unroll
        for (int n = 0; n < 5; n++) out[o + n * a.plane] = res[n];
    }
    if (LAST) {
        // one atomic per workgroup: simultaneous atomics on one address serialise
        const double m = wave_max((double)cacc);
        if (lane == 0) swmax[wv] = m;
        __syncthreads();
        if (tid == 0) {
            double mm = swmax[0];
            for (int w = 1; w < kLWaves; w++) mm = smax(mm, swmax[w]);
            atomicMax(&clk->cmax_bits[clk->parity ^ 1][blockIdx.x % kCmaxSlots], dbits(mm));
        }
    }
}

}  // namespace gk

// =============================================================================================
// Host side
struct gk_net {
    float* params = nullptr;   // [kParams] FP32 master weights
    float* m = nullptr;        // Adam moments
    float* v = nullptr;
    float* scales = nullptr;   // in_scale[16], out_scale[4]
    float* grad = nullptr;     // [kParams]
    float* loss = nullptr;     // [1]
    float* partial = nullptr;  // [partial_blocks][kParams]
    float* loss_partial = nullptr;
    int partial_blocks = 0;
    long long adam_t = 0;
    long long train_iter = 0;
    float* X = nullptr;        // data set: features [cap][16], targets [cap][4], weights [cap]
    float* T = nullptr;
    float* Wt = nullptr;
    long long cap = 0, n = 0;
    float* feat = nullptr;     // per-face staging of gk_net_collect
    float* meta = nullptr;
    long long stage_cap = 0;
    double* stats_partial = nullptr;
    float kappa = 0.5f;        // dissipation bound of corrections and training targets
    mlp::Smem* image = nullptr;  // LDS image of the weights for the corrected stage kernel
    long long version = 0;       // bumped whenever parameters or scales change
    long long image_version = -1;
    hipStream_t stream = nullptr;
    std::string err;
};

namespace {

gk_status net_fail(gk_net* n, gk_status s, const std::string& m) {
    n->err = m;
    return s;
}

#define GK_NET_HIP(net, call)                                                              \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            (net)->err = std::string(#call) + ": " + hipGetErrorString(e_);                \
            return GK_EHIP;                                                                \
        }                                                                                  \
    } while (0)

gk_status net_partials(gk_net* n, int blocks) {
    if (blocks <= n->partial_blocks) return GK_OK;
    if (n->partial) (void)hipFree(n->partial);
    if (n->loss_partial) (void)hipFree(n->loss_partial);
    n->partial = nullptr;
    n->loss_partial = nullptr;
    n->partial_blocks = 0;
    if (hipMalloc(&n->partial, sizeof(float) * (size_t)mlp::kParams * blocks) != hipSuccess ||
        hipMalloc(&n->loss_partial, sizeof(float) * blocks) != hipSuccess)
        return net_fail(n, GK_ENOMEM, "partial-gradient allocation failed");
    n->partial_blocks = blocks;
    return GK_OK;
}

// One gradient pass over `blocks` workgroups; the Adam step if `apply`.
gk_status net_step(gk_net* n, const float* X, const float* T, const float* Wt, long long count, int sequential,
                   int blocks, int tpw, unsigned long long seed, float lr, bool apply, float* grad_out, float* loss_out) {
    gk_status st = net_partials(n, blocks);
    if (st != GK_OK) return st;
    const long long batch = (long long)blocks * 64 * tpw;
    TrainArgs a;
    a.params = n->params; a.scales = n->scales;
    a.X = X; a.T = T; a.Wt = Wt; a.n = count;
    a.sequential = sequential; a.tiles_per_wave = tpw;
    a.seed = seed; a.iter = sequential ? 0 : n->train_iter++;   // draws depend on (seed, training step)
    a.inv_norm = (float)(1.0 / (4.0 * (double)(sequential ? count : batch)));
    a.partial = n->partial; a.loss_partial = n->loss_partial;
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(256), 0, n->stream, a);
    if (apply) n->version++;
    float bc1 = 1.f, bc2 = 1.f;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    if (apply) {
        n->adam_t++;
        bc1 = (float)(1.0 - std::pow(0.9, (double)n->adam_t));
        bc2 = (float)(1.0 - std::pow(0.999, (double)n->adam_t));
    }
    hipLaunchKernelGGL(adam_kernel, dim3((mlp::kParams + 255) / 256), dim3(256), 0, n->stream, n->params, n->m, n->v,
                       n->partial, blocks, n->loss_partial, a.inv_norm, loss_out, grad_out, apply ? 1 : 0, lr, b1, b2,
                       eps, bc1, bc2);
    GK_NET_HIP(n, hipGetLastError());
    return GK_OK;
}

template <class T>
gk_status upload_tmp(gk_net* n, T** dst, const T* src, size_t count) {
    GK_NET_HIP(n, hipMalloc(dst, sizeof(T) * (count ? count : 1)));
    if (count) GK_NET_HIP(n, hipMemcpy(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice));
    return GK_OK;
}

}  // namespace

// Stage weights of the step's flux: the stage's own flux coefficient times the coefficients with
// which later stages carry its output (RK3: 1/6, 1/6, 2/3; RK2: 1/2, 1/2; forward Euler: 1).
static std::vector<double> stage_flux_weights(const gk_ctx* c) {
    const size_t S = c->stages.size();
    std::vector<double> w(S);
    for (size_t q = 0; q < S; q++) {
        double v = c->stages[q].three ? c->stages[q].c3 : c->stages[q].c2;
        for (size_t p = q + 1; p < S; p++) v *= c->stages[p].c2;
        w[q] = v;
    }
    return w;
}

static long long face_count(const gk_ctx* c) { return (long long)(c->nx + 1) * c->ny + (long long)c->nx * (c->ny + 1); }

static void record_launch(gk_ctx* c, size_t q) {
    const int4 bcv = make_int4(c->bc[0], c->bc[1], c->bc[2], c->bc[3]);
    const StageSpec& s = c->stages[q];
    const long long nf = face_count(c);
    const int blocks = (int)std::min<long long>((nf + 255) / 256, 4096);
    const bool pp = c->phys.integrator == 1;
    if (c->prec == 32)
        hipLaunchKernelGGL(flux_record_kernel<float>, dim3(blocks), dim3(256), 0, c->stream, (const float*)c->buf[s.in],
                           (const float*)c->buf[1], pp ? 1 : 0, c->plane, c->pitch, c->nx, c->ny, bcv,
                           EosPolicy<float>::make(c->phys.gamma), c->clock, c->rec_w[q], c->acc);
    else
        hipLaunchKernelGGL(flux_record_kernel<double>, dim3(blocks), dim3(256), 0, c->stream, (const double*)c->buf[s.in],
                           (const double*)c->buf[1], pp ? 1 : 0, c->plane, c->pitch, c->nx, c->ny, bcv,
                           EosPolicy<double>::make(c->phys.gamma), c->clock, c->rec_w[q], c->acc);
}

static int learned_tiles(const gk_ctx* c) { return ((c->nx + kLTX - 1) / kLTX) * ((c->ny + kLTY - 1) / kLTY); }

static bool learned_launch(gk_ctx* c, const StageSpec& s) {
    if (!c->net || c->prec != 32) return false;
    StageArgs<float> a = make_args<float>(c, s);
    gk_net* net = c->net;
    if (net->image_version != net->version) {
        // the network's own stream has finished its updates (every net call that changes
        // parameters synchronises it); the image is rebuilt in this grid's stream order
        hipLaunchKernelGGL(mlp_image_kernel, dim3(kImageCopies), dim3(256), 0, c->stream, net->params, net->scales,
                           net->image);
        net->image_version = net->version;
    }
    NetArgs na;
    na.image = net->image;
    na.stats = c->lstats;
    na.kappa = c->net->kappa;
    na.enabled = 1;
    dim3 grid(learned_tiles(c)), block(kLThreads);
    if (s.three) {
        if (s.last) hipLaunchKernelGGL((learned_stage_kernel<true, true>), grid, block, 0, c->stream, a, na);
        else hipLaunchKernelGGL((learned_stage_kernel<true, false>), grid, block, 0, c->stream, a, na);
    } else {
        if (s.last) hipLaunchKernelGGL((learned_stage_kernel<false, true>), grid, block, 0, c->stream, a, na);
        else hipLaunchKernelGGL((learned_stage_kernel<false, false>), grid, block, 0, c->stream, a, na);
    }
    return true;
}

extern "C" {

gk_status gk_net_create(gk_net** out) {
    if (!out) return GK_EINVAL;
    *out = nullptr;
    gk_net* n = new gk_net();
    auto bail = [&](gk_status s) { gk_net_destroy(n); return s; };
    if (hipStreamCreate(&n->stream) != hipSuccess) return bail(GK_EHIP);
    if (hipMalloc(&n->params, sizeof(float) * mlp::kParams) != hipSuccess ||
        hipMalloc(&n->m, sizeof(float) * mlp::kParams) != hipSuccess ||
        hipMalloc(&n->v, sizeof(float) * mlp::kParams) != hipSuccess ||
        hipMalloc(&n->grad, sizeof(float) * mlp::kParams) != hipSuccess ||
        hipMalloc(&n->scales, sizeof(float) * (mlp::kIn + mlp::kOut)) != hipSuccess ||
        hipMalloc(&n->loss, sizeof(float)) != hipSuccess ||
        hipMalloc(&n->stats_partial, sizeof(double) * 21 * 1024) != hipSuccess ||
        hipMalloc(&n->image, sizeof(mlp::Smem) * kImageCopies) != hipSuccess)
        return bail(GK_ENOMEM);
    std::vector<float> zero(mlp::kParams, 0.f), ones(mlp::kIn + mlp::kOut, 1.f);
    if (hipMemcpy(n->params, zero.data(), sizeof(float) * mlp::kParams, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(n->m, zero.data(), sizeof(float) * mlp::kParams, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(n->v, zero.data(), sizeof(float) * mlp::kParams, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(n->scales, ones.data(), sizeof(float) * ones.size(), hipMemcpyHostToDevice) != hipSuccess)
        return bail(GK_EHIP);
    *out = n;
    return GK_OK;
}

void gk_net_destroy(gk_net* n) {
    if (!n) return;
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    void* bufs[] = {n->params, n->m, n->v, n->scales, n->grad, n->loss, n->partial, n->loss_partial,
                    n->X, n->T, n->Wt, n->feat, n->meta, n->stats_partial, n->image};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
}

gk_status gk_net_set_params(gk_net* n, const float* params) {
    if (!n || !params) return GK_EINVAL;
    GK_NET_HIP(n, hipMemcpy(n->params, params, sizeof(float) * mlp::kParams, hipMemcpyHostToDevice));
    // a new starting point restarts the optimiser
    std::vector<float> zero(mlp::kParams, 0.f);
    GK_NET_HIP(n, hipMemcpy(n->m, zero.data(), sizeof(float) * mlp::kParams, hipMemcpyHostToDevice));
    GK_NET_HIP(n, hipMemcpy(n->v, zero.data(), sizeof(float) * mlp::kParams, hipMemcpyHostToDevice));
    n->adam_t = 0;
    n->train_iter = 0;
    n->version++;
    return GK_OK;
}

gk_status gk_net_get_params(gk_net* n, float* params) {
    if (!n || !params) return GK_EINVAL;
    GK_NET_HIP(n, hipStreamSynchronize(n->stream));
    GK_NET_HIP(n, hipMemcpy(params, n->params, sizeof(float) * mlp::kParams, hipMemcpyDeviceToHost));
    return GK_OK;
}

gk_status gk_net_set_scales(gk_net* n, const float in_scale[16], const float out_scale[4]) {
    if (!n || !in_scale || !out_scale) return GK_EINVAL;
    for (int k = 0; k < 4; k++)
        if (!(out_scale[k] > 0.f) || !std::isfinite(out_scale[k])) return net_fail(n, GK_EINVAL, "out_scale must be positive");
    float h[mlp::kIn + mlp::kOut];
    std::memcpy(h, in_scale, sizeof(float) * mlp::kIn);
    std::memcpy(h + mlp::kIn, out_scale, sizeof(float) * mlp::kOut);
    GK_NET_HIP(n, hipMemcpy(n->scales, h, sizeof h, hipMemcpyHostToDevice));
    n->version++;
    return GK_OK;
}

gk_status gk_net_get_scales(gk_net* n, float in_scale[16], float out_scale[4]) {
    if (!n || !in_scale || !out_scale) return GK_EINVAL;
    float h[mlp::kIn + mlp::kOut];
    GK_NET_HIP(n, hipMemcpy(h, n->scales, sizeof h, hipMemcpyDeviceToHost));
    std::memcpy(in_scale, h, sizeof(float) * mlp::kIn);
    std::memcpy(out_scale, h + mlp::kIn, sizeof(float) * mlp::kOut);
    return GK_OK;
}

gk_status gk_net_set_kappa(gk_net* n, float kappa) {
    if (!n || !(kappa >= 0.f)) return GK_EINVAL;
    n->kappa = kappa;
    return GK_OK;
}

gk_status gk_net_get_kappa(gk_net* n, float* kappa) {
    if (!n || !kappa) return GK_EINVAL;
    *kappa = n->kappa;
    return GK_OK;
}

gk_status gk_net_forward(gk_net* n, long long count, const float* x, float* y) {
    if (!n || count < 0 || (count && (!x || !y))) return GK_EINVAL;
    if (count == 0) return GK_OK;
    float *dx = nullptr, *dy = nullptr;
    gk_status st = upload_tmp(n, &dx, x, (size_t)count * mlp::kIn);
    if (st == GK_OK && hipMalloc(&dy, sizeof(float) * (size_t)count * mlp::kOut) != hipSuccess) st = net_fail(n, GK_ENOMEM, "alloc");
    if (st == GK_OK) {
        const long long tiles = (count + 15) / 16;
        const int blocks = (int)std::min<long long>((tiles + 3) / 4, 2048);
        hipLaunchKernelGGL(mlp_forward_kernel, dim3(blocks), dim3(256), 0, n->stream, n->params, n->scales, dx, count, dy);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(n->stream) != hipSuccess ||
            hipMemcpy(y, dy, sizeof(float) * (size_t)count * mlp::kOut, hipMemcpyDeviceToHost) != hipSuccess)
            st = net_fail(n, GK_EHIP, "forward failed");
    }
    if (dx) (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    return st;
}

gk_status gk_net_gradient(gk_net* n, long long count, const float* x, const float* t, const float* wt, float* grad,
                          double* loss) {
    if (!n || count <= 0 || !x || !t || !wt || !grad) return GK_EINVAL;
    float *dx = nullptr, *dt = nullptr, *dw = nullptr;
    gk_status st = upload_tmp(n, &dx, x, (size_t)count * mlp::kIn);
    if (st == GK_OK) st = upload_tmp(n, &dt, t, (size_t)count * mlp::kOut);
    if (st == GK_OK) st = upload_tmp(n, &dw, wt, (size_t)count);
    if (st == GK_OK) {
        const int tpw = 4;
        const int blocks = (int)((count + 64 * tpw - 1) / (64 * tpw));
        st = net_step(n, dx, dt, dw, count, 1, blocks, tpw, 0, 0.f, false, n->grad, n->loss);
    }
    float hl = 0.f;
    if (st == GK_OK && (hipStreamSynchronize(n->stream) != hipSuccess ||
                        hipMemcpy(grad, n->grad, sizeof(float) * mlp::kParams, hipMemcpyDeviceToHost) != hipSuccess ||
                        hipMemcpy(&hl, n->loss, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess))
        st = net_fail(n, GK_EHIP, "gradient failed");
    if (st == GK_OK && loss) *loss = hl;
    for (float* b : {dx, dt, dw})
        if (b) (void)hipFree(b);
    return st;
}

gk_status gk_net_adam(gk_net* n, const float* grad, float lr) {
    if (!n || !grad) return GK_EINVAL;
    gk_status st = net_partials(n, 1);
    if (st != GK_OK) return st;
    GK_NET_HIP(n, hipMemcpy(n->partial, grad, sizeof(float) * mlp::kParams, hipMemcpyHostToDevice));
    n->adam_t++;
    n->version++;
    const float bc1 = (float)(1.0 - std::pow(0.9, (double)n->adam_t));
    const float bc2 = (float)(1.0 - std::pow(0.999, (double)n->adam_t));
    hipLaunchKernelGGL(adam_kernel, dim3((mlp::kParams + 255) / 256), dim3(256), 0, n->stream, n->params, n->m, n->v,
                       n->partial, 1, n->loss_partial, 1.f, (float*)nullptr, (float*)nullptr, 1, lr, 0.9f, 0.999f,
                       1e-8f, bc1, bc2);
    GK_NET_HIP(n, hipGetLastError());
    GK_NET_HIP(n, hipStreamSynchronize(n->stream));
    return GK_OK;
}

gk_status gk_net_reserve(gk_net* n, long long cap) {
    if (!n || cap < 0) return GK_EINVAL;
    if (cap <= n->cap) return GK_OK;
    float *X = nullptr, *T = nullptr, *W = nullptr;
    if (hipMalloc(&X, sizeof(float) * 16 * (size_t)cap) != hipSuccess ||
        hipMalloc(&T, sizeof(float) * 4 * (size_t)cap) != hipSuccess ||
        hipMalloc(&W, sizeof(float) * (size_t)cap) != hipSuccess) {
        for (float* b : {X, T, W})
            if (b) (void)hipFree(b);
        return net_fail(n, GK_ENOMEM, "data-set allocation failed");
    }
    if (n->n) {
        GK_NET_HIP(n, hipMemcpy(X, n->X, sizeof(float) * 16 * (size_t)n->n, hipMemcpyDeviceToDevice));
        GK_NET_HIP(n, hipMemcpy(T, n->T, sizeof(float) * 4 * (size_t)n->n, hipMemcpyDeviceToDevice));
        GK_NET_HIP(n, hipMemcpy(W, n->Wt, sizeof(float) * (size_t)n->n, hipMemcpyDeviceToDevice));
    }
    for (float* b : {n->X, n->T, n->Wt})
        if (b) (void)hipFree(b);
    n->X = X; n->T = T; n->Wt = W;
    n->cap = cap;
    return GK_OK;
}

gk_status gk_net_add_samples(gk_net* n, long long count, const float* x, const float* t, const float* wt) {
    if (!n || count < 0 || (count && (!x || !t || !wt))) return GK_EINVAL;
    if (n->n + count > n->cap) return net_fail(n, GK_EINVAL, "data set full (gk_net_reserve)");
    GK_NET_HIP(n, hipMemcpy(n->X + 16 * n->n, x, sizeof(float) * 16 * (size_t)count, hipMemcpyHostToDevice));
    GK_NET_HIP(n, hipMemcpy(n->T + 4 * n->n, t, sizeof(float) * 4 * (size_t)count, hipMemcpyHostToDevice));
    GK_NET_HIP(n, hipMemcpy(n->Wt + n->n, wt, sizeof(float) * (size_t)count, hipMemcpyHostToDevice));
    n->n += count;
    return GK_OK;
}

gk_status gk_net_dataset(gk_net* n, long long start, long long count, float* x, float* t, float* wt, long long* size) {
    if (!n || start < 0 || count < 0) return GK_EINVAL;
    if (size) *size = n->n;
    if (count == 0) return GK_OK;
    if (start + count > n->n) return net_fail(n, GK_EINVAL, "range beyond the data set");
    GK_NET_HIP(n, hipStreamSynchronize(n->stream));
    if (x) GK_NET_HIP(n, hipMemcpy(x, n->X + 16 * start, sizeof(float) * 16 * (size_t)count, hipMemcpyDeviceToHost));
    if (t) GK_NET_HIP(n, hipMemcpy(t, n->T + 4 * start, sizeof(float) * 4 * (size_t)count, hipMemcpyDeviceToHost));
    if (wt) GK_NET_HIP(n, hipMemcpy(wt, n->Wt + start, sizeof(float) * (size_t)count, hipMemcpyDeviceToHost));
    return GK_OK;
}

gk_status gk_net_dataset_stats(gk_net* n, double ms_x[16], double ms_t[4], double* weight) {
    if (!n || !ms_x || !ms_t) return GK_EINVAL;
    const int blocks = 1024;
    std::vector<double> h(21 * blocks);
    if (n->n > 0) {
        hipLaunchKernelGGL(dataset_stats_kernel, dim3(blocks), dim3(256), 0, n->stream, n->X, n->T, n->Wt, n->n,
                           n->stats_partial);
        GK_NET_HIP(n, hipGetLastError());
        GK_NET_HIP(n, hipStreamSynchronize(n->stream));
        GK_NET_HIP(n, hipMemcpy(h.data(), n->stats_partial, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    }
    double s[21] = {0};
    for (int b = 0; b < blocks; b++)
        for (int q = 0; q < 21; q++) s[q] += h[b * 21 + q];
    const double w = s[20] > 0 ? s[20] : 1.0;
    for (int q = 0; q < 16; q++) ms_x[q] = s[q] / w;
    for (int q = 0; q < 4; q++) ms_t[q] = s[16 + q] / w;
    if (weight) *weight = s[20];
    return GK_OK;
}

gk_status gk_net_train(gk_net* n, int iters, int blocks, int tiles_per_wave, float lr, unsigned long long seed,
                       float* loss_log) {
    if (!n || iters < 0 || blocks <= 0 || tiles_per_wave <= 0 || !(lr >= 0.f)) return GK_EINVAL;
    if (n->n <= 0) return net_fail(n, GK_ESTATE, "empty data set");
    std::vector<float*> dlog;
    float* dl = nullptr;
    if (loss_log && iters > 0) GK_NET_HIP(n, hipMalloc(&dl, sizeof(float) * iters));
    gk_status st = GK_OK;
    for (int it = 0; it < iters && st == GK_OK; it++)
        st = net_step(n, n->X, n->T, n->Wt, n->n, 0, blocks, tiles_per_wave, seed, lr, true, nullptr,
                      dl ? dl + it : nullptr);
    if (st == GK_OK && hipStreamSynchronize(n->stream) != hipSuccess) st = net_fail(n, GK_EHIP, "training failed");
    if (st == GK_OK && dl && hipMemcpy(loss_log, dl, sizeof(float) * iters, hipMemcpyDeviceToHost) != hipSuccess)
        st = net_fail(n, GK_EHIP, "loss copy failed");
    if (dl) (void)hipFree(dl);
    return st;
}

const char* gk_net_last_error(const gk_net* n) { return n ? n->err.c_str() : ""; }

// ---- flux recording, restriction, data collection
gk_status gk_flux_record(gk_ctx* c, int on) {
    if (!c) return GK_EINVAL;
    if (!on) { c->recording = false; return GK_OK; }
    if (c->gravity) return fail(c, GK_EINVAL, "flux recording needs a grid without gravity");
    if (has_extra_boundaries(c))
        return fail(c, GK_EINVAL, "flux recording takes face types 0-3 only (no fixed, inner or immersed boundaries)");
    if (c->nz > 1) return fail(c, GK_EINVAL, "flux recording is 2D only");
    if (c->ny < 2) return fail(c, GK_EINVAL, "flux recording needs a 2D grid");
    const long long nf = face_count(c);
    if (!c->acc) GK_TRY_HIP(c, hipMalloc(&c->acc, sizeof(double) * 4 * (size_t)nf));
    GK_TRY_HIP(c, hipMemsetAsync(c->acc, 0, sizeof(double) * 4 * (size_t)nf, c->stream));
    c->rec_w = stage_flux_weights(c);
    c->recording = true;
    return GK_OK;
}

gk_status gk_flux_download(gk_ctx* c, double* acc) {
    if (!c || !acc) return GK_EINVAL;
    if (!c->acc) return fail(c, GK_ESTATE, "nothing recorded");
    GK_TRY_HIP(c, hipStreamSynchronize(c->stream));
    GK_TRY_HIP(c, hipMemcpy(acc, c->acc, sizeof(double) * 4 * (size_t)face_count(c), hipMemcpyDeviceToHost));
    return GK_OK;
}

gk_status gk_restrict(gk_ctx* fine, gk_ctx* coarse) {
    if (!fine || !coarse) return GK_EINVAL;
    if (!fine->uploaded) return fail(coarse, GK_ESTATE, "fine grid has no state");
    const int f = fine->nx / coarse->nx;
    if (f < 1 || fine->nx != f * coarse->nx || fine->ny != f * coarse->ny)
        return fail(coarse, GK_EINVAL, "fine grid is not an integer refinement of the coarse grid");
    gk_status s = read_clock(fine);
    if (s != GK_OK) return s;
    if ((s = read_clock(coarse)) != GK_OK) return s;
    GK_TRY_HIP(fine, hipStreamSynchronize(fine->stream));
    const void* src = fine->buf[current_buffer(fine)];
    const long long n = (long long)coarse->nx * coarse->ny;
    const int blocks = (int)std::min<long long>((n + 255) / 256, 4096);
    if (fine->prec == 64 && coarse->prec == 64)
        hipLaunchKernelGGL((restrict_kernel<double, double>), dim3(blocks), dim3(256), 0, coarse->stream,
                           (const double*)src, fine->plane, fine->pitch, (double*)coarse->buf[0], coarse->plane,
                           coarse->pitch, coarse->nx, coarse->ny, f);
    else if (fine->prec == 64)
        hipLaunchKernelGGL((restrict_kernel<double, float>), dim3(blocks), dim3(256), 0, coarse->stream,
                           (const double*)src, fine->plane, fine->pitch, (float*)coarse->buf[0], coarse->plane,
                           coarse->pitch, coarse->nx, coarse->ny, f);
    else if (coarse->prec == 64)
        hipLaunchKernelGGL((restrict_kernel<float, double>), dim3(blocks), dim3(256), 0, coarse->stream,
                           (const float*)src, fine->plane, fine->pitch, (double*)coarse->buf[0], coarse->plane,
                           coarse->pitch, coarse->nx, coarse->ny, f);
    else
        hipLaunchKernelGGL((restrict_kernel<float, float>), dim3(blocks), dim3(256), 0, coarse->stream,
                           (const float*)src, fine->plane, fine->pitch, (float*)coarse->buf[0], coarse->plane,
                           coarse->pitch, coarse->nx, coarse->ny, f);
    GK_TRY_HIP(coarse, hipGetLastError());
    // the coarse run continues from the fine run's time
    coarse->hclock->cur = 0;
    coarse->hclock->t = fine->hclock->t;
    coarse->hclock->step = 0;
    coarse->hclock->is_continue = 1;
    if ((s = write_clock(coarse)) != GK_OK) return s;
    coarse->uploaded = true;
    return refresh_cmax(coarse);
}

gk_status gk_net_collect(gk_net* n, gk_ctx* fine, gk_ctx* coarse, int steps, double* span, long long* added) {
    if (!n || !fine || !coarse || steps < 1) return GK_EINVAL;
    if (fine->prec != 64 || coarse->prec != 64) return net_fail(n, GK_EINVAL, "collection runs both grids in FP64");
    if (fine->phys.integrator != coarse->phys.integrator || fine->phys.gamma != coarse->phys.gamma)
        return net_fail(n, GK_EINVAL, "fine and coarse grids must share integrator and gamma");
    const int f = fine->nx / coarse->nx;
    if (f < 1 || fine->nx != f * coarse->nx || fine->ny != f * coarse->ny)
        return net_fail(n, GK_EINVAL, "fine grid is not an integer refinement of the coarse grid");
    const long long nf = face_count(coarse);
    if (n->n + nf > n->cap) return net_fail(n, GK_EINVAL, "data set full (gk_net_reserve)");
    if (nf > n->stage_cap) {
        if (n->feat) (void)hipFree(n->feat);
        if (n->meta) (void)hipFree(n->meta);
        n->feat = n->meta = nullptr;
        n->stage_cap = 0;
        if (hipMalloc(&n->feat, sizeof(float) * 16 * (size_t)nf) != hipSuccess ||
            hipMalloc(&n->meta, sizeof(float) * 4 * (size_t)nf) != hipSuccess)
            return net_fail(n, GK_ENOMEM, "staging allocation failed");
        n->stage_cap = nf;
    }
    auto sub = [&](gk_ctx* c, gk_status s) {
        if (s != GK_OK) n->err = std::string("grid: ") + c->err;
        return s;
    };
    gk_status s;
    if ((s = sub(coarse, gk_restrict(fine, coarse))) != GK_OK) return s;
    const int4 bcv = make_int4(coarse->bc[0], coarse->bc[1], coarse->bc[2], coarse->bc[3]);
    const int blocks = (int)std::min<long long>((nf + 255) / 256, 4096);
    hipLaunchKernelGGL(face_features_kernel<double>, dim3(blocks), dim3(256), 0, coarse->stream,
                       (const double*)coarse->buf[0], coarse->plane, coarse->pitch, coarse->nx, coarse->ny, bcv,
                       EosPolicy<double>::make(coarse->phys.gamma), n->feat, n->meta);
    GK_NET_HIP(n, hipGetLastError());
    // `steps` coarse steps with recording: their flux and their span D
    const double tc0 = coarse->hclock->t;
    if ((s = sub(coarse, gk_flux_record(coarse, 1))) != GK_OK) return s;
    if ((s = sub(coarse, gk_run_steps(coarse, steps))) != GK_OK) return s;
    if ((s = sub(coarse, read_clock(coarse))) != GK_OK) return s;
    coarse->recording = false;
    if (coarse->hclock->step != steps) return net_fail(n, GK_ESTATE, "coarse steps did not run (stepmax?)");
    const double D = coarse->hclock->t - tc0;
    // the fine run over exactly D, recorded
    if ((s = sub(fine, gk_flux_record(fine, 1))) != GK_OK) return s;
    double t0 = fine->hclock->t;
    double t1 = 0;
    s = sub(fine, gk_advance(fine, D, &t1, nullptr));
    fine->recording = false;
    if (s != GK_OK) return s;
    if (!(std::fabs(t1 - (t0 + D)) <= 1e-12 * std::max(1.0, std::fabs(t1)))) {
        char msg[256];
        std::snprintf(msg, sizeof msg, "fine run did not end at the coarse step's end: t0 %.17g D %.17g t1 %.17g step %d",
                      t0, D, t1, fine->hclock->step);
        return net_fail(n, GK_ESTATE, msg);
    }
    GK_NET_HIP(n, hipStreamSynchronize(fine->stream));
    hipLaunchKernelGGL(face_targets_kernel, dim3(blocks), dim3(256), 0, coarse->stream, fine->acc, fine->nx, fine->ny,
                       coarse->acc, coarse->nx, coarse->ny, f, D, coarse->phys.gamma, (double)n->kappa, n->feat,
                       n->meta, n->X, n->T, n->Wt, n->n);
    GK_NET_HIP(n, hipGetLastError());
    GK_NET_HIP(n, hipStreamSynchronize(coarse->stream));
    n->n += nf;
    if (span) *span = D;
    if (added) *added = nf;
    return GK_OK;
}

// ---- the corrected solver
gk_status gk_set_net(gk_ctx* c, gk_net* n) {
    if (!c) return GK_EINVAL;
    if (!n) { c->net = nullptr; return GK_OK; }
    if (const char* m = net_conflict(c)) return fail(c, GK_EINVAL, m);
    if (c->nx < kLMinSide || c->ny < kLMinSide) return fail(c, GK_EINVAL, "the corrected stage needs NX, NY >= 16");
    if (!c->lstats) {
        const size_t bytes = sizeof(unsigned long long) * 3 * (size_t)learned_tiles(c);
        GK_TRY_HIP(c, hipMalloc(&c->lstats, bytes));
        GK_TRY_HIP(c, hipMemset(c->lstats, 0, bytes));
    }
    c->net = n;
    return GK_OK;
}

gk_status gk_net_stats(gk_ctx* c, long long out[3], int reset) {
    if (!c || !out) return GK_EINVAL;
    long long sum[3] = {0, 0, 0};
    if (c->lstats) {
        const int tiles = learned_tiles(c);
        std::vector<unsigned long long> h(3 * (size_t)tiles);
        GK_TRY_HIP(c, hipStreamSynchronize(c->stream));
        GK_TRY_HIP(c, hipMemcpy(h.data(), c->lstats, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost));
        if (reset) GK_TRY_HIP(c, hipMemset(c->lstats, 0, sizeof(unsigned long long) * h.size()));
        for (int t = 0; t < tiles; t++)
            for (int k = 0; k < 3; k++) sum[k] += (long long)h[3 * (size_t)t + k];
    }
    for (int k = 0; k < 3; k++) out[k] = sum[k];
    return GK_OK;
}

}  // extern "C"
