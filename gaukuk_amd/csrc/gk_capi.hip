// gk_capi.hip -- host runtime behind include/gaukuk.h.
//
// A context owns three state buffers (SoA planes of the active region), the device clock, a
// step log and reduction scratch.  Integrators are tables of stage launches selected once at
// creation, as the reference selects its pointer-to-member integrator (src/sim.cpp:60-66).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "gaukuk.h"
#include "gk_kernels.hip"

using namespace gk;

#define GK_STR2(x) #x
#define GK_STR(x) GK_STR2(x)

namespace {

thread_local std::string g_create_error;

struct StageSpec {
    int q;              // index of the stage in the step
    int in, base, out;  // buffer indices 0 (A = U^n), 1 (B), 2 (C)
    bool three;
    double c1, c2, c3;
    bool first, last;
};

// Blocks of the conservation-sum reduction: each writes five partial sums, added in block order.
constexpr int kSumBlocks = 512;

}  // namespace

// Partial sums of gk_cons_sums_noalloc: one array per GPU in the library's device memory (20 KiB),
// held until the process ends.  Host code never uses the name itself (a host-side stand-in, not a
// device address): it uses the address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double scratch[5 * kSumBlocks];

// Per-device address of `scratch`, looked up once per device on the first call made with that
// device current.
constexpr int kMaxDevices = 64;
static double* g_scratch_dev[kMaxDevices] = {nullptr};
static std::mutex g_scratch_mu;

// How one stage launch divides the grid: segments of seg_rows rows, strips x segs (x planes) waves.
struct WavePlan {
    int seg_rows = 0, segs = 0, nwaves = 0, nblocks = 0;
};

struct gk_ctx {
    gk_grid grid;
    gk_phys phys;
    int prec = 64;
    size_t esz = 8;
    int nx = 0, ny = 0, nz = 1, pitch = 0;
    long long plane = 0;     // elements per state variable (all planes)
    long long plane2 = 0;    // elements per z plane
    void* buf[3] = {nullptr, nullptr, nullptr};
    DevClock* clock = nullptr;
    StepRecord* log = nullptr;
    int log_cap = 4096;
    double* partial = nullptr;
    double* sums = nullptr;
    int sum_blocks = kSumBlocks;
    DevClock* hclock = nullptr;  // pinned mirror
    hipStream_t stream = nullptr;
    int seg_rows = 0, strips = 0, segs = 0, nwaves = 0, nblocks = 0;
    bool tile_fixed = false;   // gk_set_tile chose the segment height for every stage kernel
    int cus = 0;               // compute units of the device
    // per stage kernel (keyed by its entry point): segments sized to one resident round at that
    // kernel's occupancy
    std::vector<std::pair<const void*, WavePlan>> plans;
    bool uploaded = false;
    bool gravity = false;
    double dxRec = 0, dyRec = 0, drmin = 0;
    int bc[6] = {0, 0, 0, 0, 0, 0};
    std::vector<StageSpec> stages;
    std::string err;
    // point masses (gk_set_point_gravity)
    int npoint = 0;
    double pm[2][8] = {{0}};           // gm, x, y, z, vx, vy, vz, rs
    double ab = 0, omega = 0, theta1 = 0;
    double* xc = nullptr;              // device cell centres, ghost-inclusive
    double* yc = nullptr;
    double zc = 0;
    double* zcs = nullptr;             // 3D: device z centres, ghost-inclusive
    // 3D: the nearest z ghost plane on each side as the uploaded state holds it (rows -1..ny),
    // [2][5][ny + 2][nx]; zero when the upload carried no ghost cells
    void* zstale = nullptr;
    int stale_base = 0;
    // fixed (type 4) faces (gk_set_fixed_boundary): the saved state in the reference layout, and
    // on the device the two ghost layers nearest each face (xl, xr, yl, yr) in the state precision
    std::vector<double> fixed_ref;
    void* fix[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    // inner fixed region: saved state (state layout) and per-row column intervals; bounding box
    void* inner = nullptr;
    int2* inner_rows = nullptr;
    int inner_j0 = 0, inner_j1 = 0, inner_i0 = 0, inner_i1 = 0;
    // immersed shell (gk_set_immersed_shell): the shell tables in device memory
    int nshell = 0;
    bool immersed = false;
    long long* sh_cell = nullptr;
    int* sh_ntar = nullptr;
    long long* sh_tar = nullptr;
    void* sh_dir = nullptr;
    void* sh_gam = nullptr;
    // boundary hook (gk_set_bd_hook): faces whose ghost layers a host function refreshes before
    // every stage, from a mirror of the stage input's edge bands
    gk_bd_hook hook = nullptr;
    void* hook_user = nullptr;
    void* hook_mirror = nullptr;   // caller's reference-layout host array, state precision
    int hook_mask = 0, hook_band = 0;
    std::vector<char> hook_stage;  // staging of one face's band or ghost layers
    // learned flux correction (gk_learned.hip)
    gk_net* net = nullptr;                 // attached network: stages use the corrected kernel
    unsigned long long* lstats = nullptr;  // corrected / sensor / positivity face counts
    double* acc = nullptr;                 // recorded RK-weighted dt * flux per face
    bool recording = false;
    std::vector<double> rec_w;             // each stage's share of the step's flux
};

static bool learned_launch(gk_ctx* c, const StageSpec& s);
static void record_launch(gk_ctx* c, size_t q);

#define GK_TRY_HIP(ctx, call)                                                              \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                \
            return GK_EHIP;                                                                \
        }                                                                                  \
    } while (0)

static gk_status fail(gk_ctx* c, gk_status s, const std::string& m) {
    c->err = m;
    return s;
}

// faces of the grid: x and y, and z in 3D
static int nfaces(const gk_ctx* c) { return c->nz > 1 ? 6 : 4; }
static bool has_fixed_face(const gk_ctx* c) {
    for (int f = 0; f < nfaces(c); f++)
        if (c->bc[f] == 4) return true;
    return false;
}
static bool hooked(const gk_ctx* c, int f) { return c->hook && ((c->hook_mask >> f) & 1); }
// type-4 faces whose data must come from gk_set_fixed_boundary (not refreshed by a hook)
static bool has_enrolled_face(const gk_ctx* c) {
    for (int f = 0; f < nfaces(c); f++)
        if (c->bc[f] == 4 && !hooked(c, f)) return true;
    return false;
}
// boundaries beyond the face remaps: fixed faces, the inner fixed region, an immersed shell
static bool has_extra_boundaries(const gk_ctx* c) { return has_fixed_face(c) || c->inner || c->immersed; }

// What the corrected (learned) stage kernel cannot run: it has no gravity kick, no saved-data
// boundaries and no z faces, works in FP32 and reconstructs with PLM.  nullptr when it can.
static const char* net_conflict(const gk_ctx* c) {
    if (c->prec != 32) return "the corrected stage runs in precision 32";
    if (c->gravity) return "the corrected stage does not take gravity";
    if (has_extra_boundaries(c))
        return "the corrected stage takes face types 0-3 only (no fixed, inner or immersed boundaries)";
    if (c->nz > 1) return "the corrected stage is 2D only";
    if (c->phys.rc_order == 1) return "the corrected stage uses PLM (RcOrder 2)";
    return nullptr;
}

// Boundary::FixedBDXL..FixedBDYR throw when no data was enrolled (src/boundary/fixed.cpp:18-19)
static gk_status check_fixed(gk_ctx* c) {
    // an attached network runs every stage: the configuration must still be one it can run
    if (c->net)
        if (const char* m = net_conflict(c)) return fail(c, GK_EINVAL, m);
    if (has_enrolled_face(c) && c->fixed_ref.empty())
        return fail(c, GK_ESTATE,
                    "fixed boundary is not initialized: call gk_set_fixed_boundary (EnrollFixedBoundaryData)");
    // the shell is rewritten from the stage input in memory, while the first gravity half-kick is
    // applied as the stage kernel loads U^n
    if (c->immersed && c->gravity)
        return fail(c, GK_EINVAL, "an immersed shell together with gravity is not supported");
    return GK_OK;
}

// Domain cell centres (src/sim.cpp:37-39), ghost-inclusive index, as the reference computes them
static void domain_centres(const gk_ctx* c, std::vector<double>& hx, std::vector<double>& hy, double& zc) {
    const int g = c->grid.nghost, lx = c->nx + 2 * g, ly = c->ny + 2 * g;
    const double dx = (c->grid.xmax - c->grid.xmin) / c->grid.nx;
    const double dy = (c->grid.ymax - c->grid.ymin) / c->grid.ny;
    const double dz = (c->grid.zmax - c->grid.zmin) / c->grid.nz;
    hx.resize(lx); hy.resize(ly);
    for (int i = 0; i < lx; i++) hx[i] = c->grid.xmin - g * dx + 0.5 * dx + i * dx;
    for (int j = 0; j < ly; j++) hy[j] = c->grid.ymin - g * dy + 0.5 * dy + j * dy;
    zc = c->grid.zmin - g * dz + 0.5 * dz + g * dz;   // the active plane k = nGhost
}

template <class R>
static void launch_immersed_t(gk_ctx* c, void* u, void* alt, int pingpong, int always) {
    ShellArgs<R> a;
    a.u = (R*)u; a.alt = (R*)alt; a.pingpong = pingpong; a.always = always;
    a.plane = c->plane; a.nshell = c->nshell;
    a.cell = c->sh_cell; a.ntar = c->sh_ntar; a.tar = c->sh_tar;
    a.dir = (const R*)c->sh_dir; a.gam = (const R*)c->sh_gam;
    const EosPolicy<R> eos = EosPolicy<R>::make(c->phys.gamma);
    a.gm1 = eos.gm1; a.gm1Rec = eos.gm1Rec;
    a.clock = c->clock;
    hipLaunchKernelGGL(immersed_kernel<R>, dim3((c->nshell + 255) / 256), dim3(256), 0, c->stream, a);
}
// the shell update of a stage's input (Boundary::UpdateBD, src/boundary/boundary.cpp:136-138)
static void launch_immersed(gk_ctx* c, const StageSpec& s) {
    if (!c->nshell) return;
    const int pingpong = c->phys.integrator == 1;
    if (c->prec == 32) launch_immersed_t<float>(c, c->buf[s.in], c->buf[s.out], pingpong, 0);
    else launch_immersed_t<double>(c, c->buf[s.in], c->buf[s.out], pingpong, 0);
}

// Segments for a target number of waves (seg_rows <= 0), or of the given height.  Segments of at
// least 8 rows (2 on small grids, below); with a target, at most `target` waves, so that they are resident at once, unless
// that leaves more than a fifth of the target idle: then about five rounds of shorter segments,
// whose waves refill the slots as others finish (a 3D grid of 256 planes x 5 strips at a target
// of 2048: one round of 1280 waves 3.3, 2560 waves 3.3, 10240 waves of 32 rows 4.2 G zone-updates/s).
static WavePlan make_plan(const gk_ctx* c, int target, int seg_rows) {
    WavePlan p;
    if (seg_rows <= 0) {
        const int per_seg = c->strips * c->nz;
        long long segs = std::max(1, target / per_seg);
        if (5LL * segs * per_seg < 4LL * target) segs = (5LL * target + per_seg - 1) / per_seg;
        seg_rows = std::max((int)((c->ny + segs - 1) / segs), 8);
        // a grid whose 8-row segments fill less than an eighth of the slots: 2-row segments (KH
        // 256^2 FP64: 160 waves 55.9, 640 waves 32.9 us per RK3 step; at 512^2, 576 waves of 8
        // rows are already faster than 2304 of 2)
        const long long waves8 = (long long)per_seg * ((c->ny + 7) / 8);
        if (8 * waves8 < target) seg_rows = 2;
    }
    p.seg_rows = std::min(seg_rows, c->ny);
    p.segs = (c->ny + p.seg_rows - 1) / p.seg_rows;
    p.nwaves = c->strips * p.segs * c->nz;
    const int nb = (p.nwaves + (kBlock / kWave) - 1) / (kBlock / kWave);
    p.nblocks = (nb + 7) / 8 * 8;  // a multiple of 8 keeps the XCD remap a bijection
    return p;
}

// The default plan (reported by gk_get_tile, used where no kernel-specific plan applies): one
// resident round of 3 waves per SIMD on 256 CUs x 4 SIMDs (2 for the 3D variant).
static void plan_waves(gk_ctx* c, int seg_rows) {
    c->strips = (c->nx + kStrip - 1) / kStrip;
    const WavePlan p = make_plan(c, c->nz > 1 ? 2048 : 3072, seg_rows);
    c->seg_rows = p.seg_rows;
    c->segs = p.segs;
    c->nwaves = p.nwaves;
    c->nblocks = p.nblocks;
    c->tile_fixed = seg_rows > 0;
    c->plans.clear();
}

// The plan of one stage kernel: as many waves as the device holds at once at that kernel's
// occupancy (its register count decides it: 3 waves per SIMD for the FP64 2D kernels, 4-6 for
// FP32), so that every wave starts in the first round; the default plan if the runtime cannot
// say, or the fixed height of gk_set_tile.
static WavePlan stage_plan(gk_ctx* c, const void* kernel) {
    for (const auto& e : c->plans)
        if (e.first == kernel) return e.second;
    WavePlan p;
    int blocks = 0;
    if (!c->tile_fixed && c->cus > 0 &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, kBlock, 0) == hipSuccess && blocks > 0) {
        p = make_plan(c, blocks * (kBlock / kWave) * c->cus, 0);
    } else {
        p.seg_rows = c->seg_rows; p.segs = c->segs; p.nwaves = c->nwaves; p.nblocks = c->nblocks;
    }
    c->plans.emplace_back(kernel, p);
    return p;
}

static void build_integrator(gk_ctx* c) {
    c->stages.clear();
    const int it = c->phys.integrator;
    if (it == 1) {  // ForwardEuler_: src/evolution/forward_euler.cpp:11-17
        c->stages.push_back({0, 0, 0, 1, false, 1.0, 1.0, 0.0, true, true});
    } else if (it == 3) {  // RK3_: src/evolution/rk3.cpp:11-28
        const double f13 = 1.0 / 3.0;
        c->stages.push_back({0, 0, 0, 1, false, 1.0, 1.0, 0.0, true, false});
        c->stages.push_back({1, 1, 0, 2, true, 0.75, 0.25, 0.25, false, false});
        c->stages.push_back({2, 2, 0, 0, true, f13, 2 * f13, 2 * f13, false, true});
    } else {  // RK2_ (also the reference's default branch): src/evolution/rk2.cpp:11-22
        c->stages.push_back({0, 0, 0, 1, false, 1.0, 1.0, 0.0, true, false});
        c->stages.push_back({1, 1, 0, 0, true, 0.5, 0.5, 0.5, false, true});
    }
}

template <class R>
static StageArgs<R> make_args(gk_ctx* c, const StageSpec& s) {
    StageArgs<R> a;
    std::memset(&a, 0, sizeof(a));
    a.in = (const R*)c->buf[s.in];
    a.base = (const R*)c->buf[s.base];
    a.out = (R*)c->buf[s.out];
    a.plane = c->plane;
    a.pitch = c->pitch; a.nx = c->nx; a.ny = c->ny;
    a.nz = c->nz; a.plane2 = c->plane2;
    a.dzRec = c->nz > 1 ? 1. / ((c->grid.zmax - c->grid.zmin) / c->grid.nz) : 0.0;   // src/sim.cpp:26-33
    a.zstale = (const R*)c->zstale; a.stale_base = c->stale_base; a.stage_q = s.q;
    a.zcs = c->zcs;
    a.seg_rows = c->seg_rows; a.strips = c->strips; a.nwaves = c->nwaves;
    for (int k = 0; k < 6; k++) a.bc[k] = c->bc[k];
    // RK3's last coefficients are frac13 = 1.0/3.0 and 2*frac13 (src/evolution/rk3.cpp:25-26);
    // rounding the double values to float equals evaluating them in float.
    a.c1 = R(s.c1); a.c2 = R(s.c2); a.c3 = R(s.c3);
    a.kick_in = c->gravity && s.first;
    a.kick_base = c->gravity;
    a.pingpong = c->phys.integrator == 1;
    a.kick_out = c->gravity && s.last;
    a.want_cmax = s.last;
    a.gx = R(c->phys.g[0]); a.gy = R(c->phys.g[1]); a.gz = R(c->phys.g[2]);
    a.dxRec = c->dxRec; a.dyRec = c->dyRec;
    a.eos = EosPolicy<R>::make(c->phys.gamma);
    a.clock = c->clock;
    a.npoint = c->npoint;
    for (int q = 0; q < 2; q++) {
        double* d = q == 0 ? a.pm0 : a.pm1;
        d[0] = c->pm[q][0]; d[1] = c->pm[q][1]; d[2] = c->pm[q][2]; d[3] = c->pm[q][3]; d[4] = c->pm[q][7];
    }
    a.ab = c->ab; a.omega = c->omega; a.theta1 = c->theta1;
    a.xc = c->xc; a.yc = c->yc; a.zc = c->zc; a.ng = c->grid.nghost;
    a.fxl = (const R*)c->fix[0]; a.fxr = (const R*)c->fix[1];
    a.fyl = (const R*)c->fix[2]; a.fyr = (const R*)c->fix[3];
    a.fzl = (const R*)c->fix[4]; a.fzr = (const R*)c->fix[5];
    a.inner = (const R*)c->inner; a.inner_rows = c->inner_rows;
    return a;
}

using PlmLimiter = VanLeer;

template <class R, class L, bool THREE, int VAR>
static void launch_stage_f(gk_ctx* c, const StageArgs<R>& a, bool last) {
    void (*k)(StageArgs<R>) =
        c->gravity ? (last ? stage_kernel<R, L, THREE, true, true, VAR> : stage_kernel<R, L, THREE, true, false, VAR>)
                   : (last ? stage_kernel<R, L, THREE, false, true, VAR> : stage_kernel<R, L, THREE, false, false, VAR>);
    const WavePlan p = stage_plan(c, reinterpret_cast<const void*>(k));
    StageArgs<R> b = a;
    b.seg_rows = p.seg_rows;
    b.nwaves = p.nwaves;
    hipLaunchKernelGGL(k, dim3(p.nblocks), dim3(kBlock), 0, c->stream, b);
}

// One state larger than half of the 256 MB memory-side cache: two of a step's states no longer fit
// together, and plain 2D PLM grids run the stage variant with cache-policy hints (VAR 4).
static bool states_exceed_cache(const gk_ctx* c) { return (size_t)c->plane * 5 * c->esz > (size_t(128) << 20); }

template <class R, class L, bool THREE>
static void launch_stage_t(gk_ctx* c, const StageArgs<R>& a, bool last) {
    const bool saved = has_fixed_face(c) || c->inner;
    if (c->nz > 1) {
        if (saved) launch_stage_f<R, L, THREE, 3>(c, a, last);
        else launch_stage_f<R, L, THREE, 2>(c, a, last);
    } else if (saved) {
        launch_stage_f<R, L, THREE, 1>(c, a, last);
    } else {
        if constexpr (L::order == 2)
            if (states_exceed_cache(c)) return launch_stage_f<R, L, THREE, 4>(c, a, last);
        launch_stage_f<R, L, THREE, 0>(c, a, last);
    }
}

template <class R>
static void launch_stage(gk_ctx* c, const StageSpec& s) {
    StageArgs<R> a = make_args<R>(c, s);
    const bool plm = c->phys.rc_order != 1;
    if (s.three) {
        if (plm) launch_stage_t<R, PlmLimiter, true>(c, a, s.last);
        else launch_stage_t<R, FirstOrder, true>(c, a, s.last);
    } else {
        if (plm) launch_stage_t<R, PlmLimiter, false>(c, a, s.last);
        else launch_stage_t<R, FirstOrder, false>(c, a, s.last);
    }
}

static void launch_stage_any(gk_ctx* c, const StageSpec& s) {
    if (c->net && learned_launch(c, s)) return;
    if (c->prec == 32) launch_stage<float>(c, s);
    else launch_stage<double>(c, s);
}

static void enqueue_stages(gk_ctx* c) {
    for (size_t q = 0; q < c->stages.size(); q++) {
        if (c->recording) record_launch(c, q);
        launch_immersed(c, c->stages[q]);
        launch_stage_any(c, c->stages[q]);
    }
}

// n steps back to back: the loop test and dt of the first step, each later step's joined to the
// previous step's tail in one launch, and the last step's tail
static void enqueue_steps(gk_ctx* c, int n) {
    if (n <= 0) return;
    const int toggle = c->phys.integrator == 1 ? 1 : 0;
    hipLaunchKernelGGL(step_begin_kernel, dim3(1), dim3(1), 0, c->stream, c->clock);
    for (int k = 0; k < n; k++) {
        if (k > 0)
            hipLaunchKernelGGL(step_seam_kernel, dim3(1), dim3(1), 0, c->stream, c->clock, c->log, c->log_cap, toggle);
        enqueue_stages(c);
    }
    hipLaunchKernelGGL(step_end_kernel, dim3(1), dim3(1), 0, c->stream, c->clock, c->log, c->log_cap, toggle);
}

// Steps to enqueue before the host looks at the clock again: the steps left to the output time at
// the next step's dt (from the wave speed the last stage left for it), one more for a dt that
// shrinks on the way and for the loop test that ends the advance, at most the steps stepmax still
// allows plus that test, and at most 256 per host round trip.  Steps past the end cost one early
// return per launch.
static int advance_chunk(const DevClock& h) {
    unsigned long long mb = 0;
    for (int s = 0; s < kCmaxSlots; s++) mb = std::max(mb, h.cmax_bits[h.parity][s]);
    double cm;
    std::memcpy(&cm, &mb, sizeof cm);
    const double dt = h.cfl * h.drmin / cm;
    const double rem = h.t_next - h.t;
    double est = (dt > 0 && std::isfinite(dt)) ? std::ceil(std::max(rem, 0.0) / dt) + 1 : 256;
    if (h.stepmax > 0) est = std::min(est, (double)(h.stepmax - h.step) + 1);
    return (int)std::max(1.0, std::min(est, 256.0));
}

static gk_status read_clock(gk_ctx* c) {
    GK_TRY_HIP(c, hipMemcpyAsync(c->hclock, c->clock, sizeof(DevClock), hipMemcpyDeviceToHost, c->stream));
    GK_TRY_HIP(c, hipStreamSynchronize(c->stream));
    return GK_OK;
}

static gk_status write_clock(gk_ctx* c) {
    GK_TRY_HIP(c, hipMemcpyAsync(c->clock, c->hclock, sizeof(DevClock), hipMemcpyHostToDevice, c->stream));
    GK_TRY_HIP(c, hipStreamSynchronize(c->stream));
    return GK_OK;
}

// Recompute the max wave speed of the current state into the slot the next step reads.
static gk_status refresh_cmax(gk_ctx* c) {
    gk_status s = read_clock(c);
    if (s != GK_OK) return s;
    unsigned long long one = 0;
    double floor_c = 1e-16;
    std::memcpy(&one, &floor_c, sizeof one);
    for (int k = 0; k < kCmaxSlots; k++) c->hclock->cmax_bits[c->hclock->parity][k] = one;
    if ((s = write_clock(c)) != GK_OK) return s;
    const int cur = c->hclock->cur;
    const void* u = c->buf[cur ? 1 : 0];
    dim3 grid(1024), block(kBlock);
    if (c->prec == 32)
        hipLaunchKernelGGL(cmax_kernel<float>, grid, block, 0, c->stream, (const float*)u, c->plane, c->pitch, c->nx,
                           c->ny * c->nz, EosPolicy<float>::make(c->phys.gamma), c->clock);
    else
        hipLaunchKernelGGL(cmax_kernel<double>, grid, block, 0, c->stream, (const double*)u, c->plane, c->pitch, c->nx,
                           c->ny * c->nz, EosPolicy<double>::make(c->phys.gamma), c->clock);
    GK_TRY_HIP(c, hipGetLastError());
    return GK_OK;
}

static int current_buffer(gk_ctx* c) { return c->hclock->cur ? 1 : 0; }

// Reference layout helpers: Real[5][lenz][leny][lenx]; active planes k = nghost .. nghost + nz - 1
// (one plane in 2D).  Active states here are [5][nz][ny][nx].
template <class R>
static void ref_to_active(const gk_ctx* c, const R* ref, double* act) {
    const int g = c->grid.nghost, nx = c->nx, ny = c->ny, nz = c->nz;
    const int lx = nx + 2 * g, ly = ny + 2 * g, lz = nz + 2 * g;
    for (int n = 0; n < 5; n++)
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++)
                    act[(((size_t)n * nz + k) * ny + j) * nx + i] =
                        (double)ref[(((size_t)n * lz + g + k) * ly + (j + g)) * lx + (i + g)];
}

template <class R>
static void active_to_ref(const gk_ctx* c, const double* act, R* ref) {
    const int g = c->grid.nghost, nx = c->nx, ny = c->ny, nz = c->nz;
    const int lx = nx + 2 * g, ly = ny + 2 * g, lz = nz + 2 * g;
    std::memset(ref, 0, sizeof(R) * 5 * (size_t)lx * ly * lz);
    auto at = [&](int n, int k, int j, int i) -> R& { return ref[(((size_t)n * lz + k) * ly + j) * lx + i]; };
    for (int n = 0; n < 5; n++)
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) at(n, k + g, j + g, i + g) = (R)act[(((size_t)n * nz + k) * ny + j) * nx + i];
    // ghost fill over the active tangential range, x faces, y faces, then z faces
    // (src/boundary/boundary.cpp:127-142); fixed faces copy the saved state at the same position
    // (src/boundary/fixed.cpp:16-179); type-0 z faces leave their planes zero (the reference's z copy
    // refills the y ghosts instead, boundary.cpp:90-122)
    auto saved = [&](int n, int k, int j, int i) {
        return (R)c->fixed_ref[(((size_t)n * lz + k) * ly + j) * lx + i];
    };
    for (int k = g; k < g + nz; k++) {
        for (int j = g; j < g + ny; j++)
            for (int i = 0; i < lx; i++) {
                if (i >= g && i < g + nx) continue;
                if (hooked(c, i < g ? 0 : 1)) continue;   // filled by the hook (gk_download_cons)
                int kind;
                const int si = bc_remap(i - g, nx, c->bc[0], c->bc[1], kind) + g;
                R u[5];
                for (int n = 0; n < 5; n++) u[n] = kind == 4 ? saved(n, k, j, i) : at(n, k, j, si);
                bc_transform(u, 1, kind);
                for (int n = 0; n < 5; n++) at(n, k, j, i) = u[n];
            }
        if (ny > 1 || nz > 1)
            for (int j = 0; j < ly; j++) {
                if (j >= g && j < g + ny) continue;
                if (hooked(c, j < g ? 2 : 3)) continue;
                for (int i = g; i < g + nx; i++) {
                    int kind;
                    const int sj = bc_remap(j - g, ny, c->bc[2], c->bc[3], kind) + g;
                    R u[5];
                    for (int n = 0; n < 5; n++) u[n] = kind == 4 ? saved(n, k, j, i) : at(n, k, sj, i);
                    bc_transform(u, 2, kind);
                    for (int n = 0; n < 5; n++) at(n, k, j, i) = u[n];
                }
            }
    }
    if (nz > 1)
        for (int k = 0; k < lz; k++) {
            if (k >= g && k < g + nz) continue;
            if (hooked(c, k < g ? 4 : 5)) continue;   // filled by the hook (gk_download_cons)
            const int zt = k < g ? c->bc[4] : c->bc[5];
            if (zt < 1 || zt > 4) continue;
            int kind;
            const int sk = bc_remap(k - g, nz, c->bc[4], c->bc[5], kind) + g;
            for (int j = g; j < g + ny; j++)
                for (int i = g; i < g + nx; i++) {
                    R u[5];
                    for (int n = 0; n < 5; n++) u[n] = kind == 4 ? saved(n, k, j, i) : at(n, sk, j, i);
                    bc_transform(u, 3, kind);
                    for (int n = 0; n < 5; n++) at(n, k, j, i) = u[n];
                }
        }
}

// ---- boundary hook (gk_set_bd_hook) ------------------------------------------------------
// Indices below are active-relative: cell (i, j, k) of the active box, ghosts at -1, -2, ... and
// nx, nx + 1, ...  The mirror is the reference layout [5][nz + 2g][ny + 2g][nx + 2g] (2D: one
// active plane between 2g ghost planes).
template <class R>
static R& mirror_at(const gk_ctx* c, int n, int k, int j, int i) {
    const int g = c->grid.nghost;
    const size_t lx = c->nx + 2 * g, ly = c->ny + 2 * g, lz = c->nz + 2 * g;
    return static_cast<R*>(c->hook_mirror)[(((size_t)n * lz + (k + g)) * ly + (j + g)) * lx + (i + g)];
}

// Face f's band: the `band` active layers nearest the face across the whole active extent of the
// other axes, [lo, hi) per axis (x, y, z).
static void hook_band_box(const gk_ctx* c, int f, int lo[3], int hi[3]) {
    const int n[3] = {c->nx, c->ny, c->nz}, a = f / 2, b = c->hook_band;
    for (int d = 0; d < 3; d++) { lo[d] = 0; hi[d] = n[d]; }
    if (f & 1) lo[a] = n[a] - b;
    else hi[a] = b;
}

// Write the bands of the hooked faces of `state` (device, [5][nz][ny][nx] of R, row pitch `pitch`)
// into the mirror; with `kick`, after the first half-kick of the step's sources (constant, then
// point or binary gravity at the cell centre, the binary's second mass at t + dt/2), as the stage
// kernel applies it when it loads U^n.
template <class R>
static gk_status hook_write_bands(gk_ctx* c, const void* state, bool kick) {
    const int ny = c->ny, nz = c->nz, b = c->hook_band, pitch = c->pitch, g = c->grid.nghost;
    const R* dev = static_cast<const R*>(state);
    const double dt = c->hclock->dt;
    const R gx = R(c->phys.g[0]), gy = R(c->phys.g[1]), gz = R(c->phys.g[2]), h = R(0.5 * dt);
    // point masses as the stage kernel takes them ({gm, x, y, z, rs}) and the cell centres
    double m0[5], m1[5];
    const double* pm[2] = {c->pm[0], c->pm[1]};
    for (int q = 0; q < 2; q++) {
        double* d = q ? m1 : m0;
        d[0] = pm[q][0]; d[1] = pm[q][1]; d[2] = pm[q][2]; d[3] = pm[q][3]; d[4] = pm[q][7];
    }
    R bx1 = R(0), by1 = R(0);
    std::vector<double> xc, yc, zcs;
    double zc = 0;
    if (kick && c->npoint) {
        domain_centres(c, xc, yc, zc);
        const double dz = (c->grid.zmax - c->grid.zmin) / c->grid.nz;
        zcs.resize((size_t)nz + 2 * g);
        for (int k = 0; k < (int)zcs.size(); k++) zcs[k] = nz > 1 ? c->grid.zmin - g * dz + 0.5 * dz + k * dz : zc;
        if (c->npoint == 2) {
            const double th = c->omega * (c->hclock->t + 0.5 * dt) + c->theta1;
            bx1 = R(c->ab * std::cos(th));
            by1 = R(c->ab * std::sin(th));
        }
    }
    for (int f = 0; f < nfaces(c); f++) {
        if (!hooked(c, f)) continue;
        int lo[3], hi[3];
        hook_band_box(c, f, lo, hi);
        // one strided copy per variable into a staging box; st(n, k, j, i) is its element
        size_t width, rows, spitch, off;
        if (f < 2) {        // b columns of all nz * ny rows (planes of ny rows follow each other)
            width = b; rows = (size_t)nz * ny; spitch = pitch; off = lo[0];
        } else if (f < 4) { // b consecutive rows of every plane
            width = (size_t)b * pitch; rows = nz; spitch = (size_t)c->plane2; off = (size_t)lo[1] * pitch;
        } else {            // b consecutive planes
            width = (size_t)b * c->plane2; rows = 1; spitch = width; off = (size_t)lo[2] * c->plane2;
        }
        c->hook_stage.resize(sizeof(R) * 5 * width * rows);
        R* st = reinterpret_cast<R*>(c->hook_stage.data());
        for (int n = 0; n < 5; n++)
            GK_TRY_HIP(c, hipMemcpy2DAsync(st + (size_t)n * width * rows, sizeof(R) * width,
                                           dev + (size_t)n * c->plane + off, sizeof(R) * spitch, sizeof(R) * width,
                                           rows, hipMemcpyDeviceToHost, c->stream));
        GK_TRY_HIP(c, hipStreamSynchronize(c->stream));
        auto at = [&](int n, int k, int j, int i) -> R {
            size_t o;
            if (f < 2) o = (((size_t)n * nz + k) * ny + j) * b + (i - lo[0]);
            else if (f < 4) o = (((size_t)n * nz + k) * b + (j - lo[1])) * pitch + i;
            else o = (((size_t)n * b + (k - lo[2])) * ny + j) * pitch + i;
            return st[o];
        };
        for (int k = lo[2]; k < hi[2]; k++)
            for (int j = lo[1]; j < hi[1]; j++)
                for (int i = lo[0]; i < hi[0]; i++) {
                    R u[5];
                    for (int n = 0; n < 5; n++) u[n] = at(n, k, j, i);
                    if (kick) {
                        gravity_kick(u, gx, gy, gz, h);
                        if (c->npoint)
                            point_kick(u, c->npoint, m0, m1, bx1, by1, h, R(xc[i + g]), R(yc[j + g]), R(zcs[k + g]));
                    }
                    for (int n = 0; n < 5; n++) mirror_at<R>(c, n, k, j, i) = u[n];
                }
    }
    return GK_OK;
}

// Before stage q: mirror the stage input's bands, call the hook, upload the ghost layers the stage
// reads into each hooked face's fixed-face array (layouts of gk_set_fixed_boundary: x faces
// [5][nz][ny][2], y faces [5][nz][2][nx], both the two layers nearest the face; z faces
// [5][ny][nx], the nearest plane).
template <class R>
static gk_status hook_refresh_t(gk_ctx* c, size_t q) {
    const StageSpec& s = c->stages[q];
    // forward Euler alternates its two buffers with the clock's `cur`
    const bool swap = c->phys.integrator == 1 && c->hclock->cur;
    const void* in = c->buf[swap ? s.out : s.in];
    if (gk_status st = hook_write_bands<R>(c, in, s.first && c->gravity)) return st;
    c->hook(c->hook_user, c->hook_mirror);
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    for (int f = 0; f < nfaces(c); f++) {
        if (!hooked(c, f)) continue;
        const size_t len = f < 2 ? (size_t)10 * nz * ny : f < 4 ? (size_t)10 * nz * nx : (size_t)5 * ny * nx;
        c->hook_stage.resize(sizeof(R) * len);
        R* st = reinterpret_cast<R*>(c->hook_stage.data());
        bool finite = true;
        auto put = [&](size_t o, R v) {
            finite &= std::isfinite((double)v);
            st[o] = v;
        };
        for (int n = 0; n < 5; n++)
            for (int k = 0; k < nz; k++) {
                if (f < 4) {
                    for (int d = 0; d < 2; d++) {
                        if (f < 2)
                            for (int j = 0; j < ny; j++)
                                put((((size_t)n * nz + k) * ny + j) * 2 + d,
                                    mirror_at<R>(c, n, k, j, f == 0 ? -1 - d : nx + d));
                        else
                            for (int i = 0; i < nx; i++)
                                put((((size_t)n * nz + k) * 2 + d) * nx + i,
                                    mirror_at<R>(c, n, k, f == 2 ? -1 - d : ny + d, i));
                    }
                } else if (k == 0) {
                    for (int j = 0; j < ny; j++)
                        for (int i = 0; i < nx; i++)
                            put(((size_t)n * ny + j) * nx + i, mirror_at<R>(c, n, f == 4 ? -1 : nz, j, i));
                }
            }
        if (!finite)
            return fail(c, GK_EINVAL, "the boundary hook left a non-finite ghost value on face " + std::to_string(f) +
                                          " (does the face function read cells beyond the mirrored band?)");
        GK_TRY_HIP(c, hipMemcpyAsync(c->fix[f], st, sizeof(R) * len, hipMemcpyHostToDevice, c->stream));
        GK_TRY_HIP(c, hipStreamSynchronize(c->stream));
    }
    return GK_OK;
}
static gk_status hook_refresh(gk_ctx* c, size_t q) {
    return c->prec == 32 ? hook_refresh_t<float>(c, q) : hook_refresh_t<double>(c, q);
}

// Output ghost cells of the hooked faces: the hook applied to the current state; all nGhost
// layers over the active tangential range, as the mirror holds them after the call.
template <class R>
static gk_status hook_fill_output(gk_ctx* c, const void* state, R* ref) {
    if (gk_status st = hook_write_bands<R>(c, state, false)) return st;
    c->hook(c->hook_user, c->hook_mirror);
    const int g = c->grid.nghost, nx = c->nx, ny = c->ny, nz = c->nz;
    const R* mir = static_cast<const R*>(c->hook_mirror);
    for (int f = 0; f < nfaces(c); f++) {
        if (!hooked(c, f)) continue;
        const int a = f / 2, n[3] = {nx, ny, nz};
        int lo[3] = {0, 0, 0}, hi[3] = {nx, ny, nz};
        lo[a] = f & 1 ? n[a] : -g;
        hi[a] = f & 1 ? n[a] + g : 0;
        for (int v = 0; v < 5; v++)
            for (int k = lo[2]; k < hi[2]; k++)
                for (int j = lo[1]; j < hi[1]; j++)
                    for (int i = lo[0]; i < hi[0]; i++) {
                        const size_t o = &mirror_at<R>(c, v, k, j, i) - mir;
                        ref[o] = mir[o];
                    }
    }
    return GK_OK;
}

// Whole steps with the host back in the loop before every stage (contexts with a boundary hook):
// the same kernels as enqueue_steps, one step at a time; stops early when the clock goes inactive.
static gk_status hooked_steps(gk_ctx* c, int n) {
    const int toggle = c->phys.integrator == 1 ? 1 : 0;
    for (int k = 0; k < n; k++) {
        hipLaunchKernelGGL(step_begin_kernel, dim3(1), dim3(1), 0, c->stream, c->clock);
        if (gk_status s = read_clock(c)) return s;
        if (!c->hclock->active) break;
        for (size_t q = 0; q < c->stages.size(); q++) {
            if (gk_status s = hook_refresh(c, q)) return s;
            launch_immersed(c, c->stages[q]);
            launch_stage_any(c, c->stages[q]);
        }
        hipLaunchKernelGGL(step_end_kernel, dim3(1), dim3(1), 0, c->stream, c->clock, c->log, c->log_cap, toggle);
        GK_TRY_HIP(c, hipGetLastError());
    }
    return GK_OK;
}

template <class T>
static gk_status upload_vec(gk_ctx* c, T** dst, const std::vector<T>& v) {
    if (!*dst) GK_TRY_HIP(c, hipMalloc((void**)dst, sizeof(T) * (v.empty() ? 1 : v.size())));
    if (!v.empty()) GK_TRY_HIP(c, hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return GK_OK;
}
// a vector of doubles in the state precision
static gk_status upload_real(gk_ctx* c, void** dst, const std::vector<double>& v) {
    if (c->prec == 32) {
        std::vector<float> f(v.begin(), v.end());
        return upload_vec(c, (float**)dst, f);
    }
    return upload_vec(c, (double**)dst, v);
}

// Boundary::FixedBDInner's region (src/boundary/fixed.cpp:181-207): the central half-box of the
// active cells in x and y, on every active plane, restricted to cell centres within 0.5 of the
// origin; per active row the region is one run of columns (x^2 grows with |x| and rounding keeps
// that order).  Rows are numbered k * ny + j over the planes, which is also their order in the
// state (planes of ny rows follow each other), so the bounding rows [inner_j0, inner_j1) span
// planes in 3D.
static gk_status set_inner(gk_ctx* c) {
    const int g = c->grid.nghost, nx = c->nx, ny = c->ny, nz = c->nz;
    const int ib = g, ie = g + nx, jb = g, je = g + ny;
    const int il = static_cast<int>(ib + 0.25 * (ie - ib)), ir = static_cast<int>(ib + 0.75 * (ie - ib));
    const int jl = static_cast<int>(jb + 0.25 * (je - jb)), jr = static_cast<int>(jb + 0.75 * (je - jb));
    std::vector<double> xc, yc;
    double zc;
    domain_centres(c, xc, yc, zc);
    const double dz = (c->grid.zmax - c->grid.zmin) / c->grid.nz;
    std::vector<int2> rows((size_t)ny * nz, make_int2(0, 0));
    c->inner_j0 = ny * nz; c->inner_j1 = 0; c->inner_i0 = nx; c->inner_i1 = 0;
    for (int k = 0; k < nz; k++) {
        // Domain::zc of active plane k (src/sim.cpp:39); the single plane's centre in 2D
        const double z = nz > 1 ? c->grid.zmin - g * dz + 0.5 * dz + (k + g) * dz : zc;
        for (int j = jl; j < jr; j++) {
            int lo = -1, hi = -1;
            for (int i = il; i < ir; i++) {
                const double r2 = xc[i] * xc[i] + yc[j] * yc[j] + z * z;
                if (r2 < 0.5 * 0.5) {
                    if (lo < 0) lo = i;
                    else if (hi != i) return fail(c, GK_EINVAL, "inner fixed region is not one run of columns per row");
                    hi = i + 1;
                }
            }
            if (lo < 0) continue;
            const int row = k * ny + j - g;
            rows[row] = make_int2(lo - g, hi - g);
            c->inner_j0 = std::min(c->inner_j0, row); c->inner_j1 = std::max(c->inner_j1, row + 1);
            c->inner_i0 = std::min(c->inner_i0, lo - g); c->inner_i1 = std::max(c->inner_i1, hi - g);
        }
    }
    std::vector<double> saved((size_t)c->plane * 5);
    const int lx = nx + 2 * g, ly = ny + 2 * g, lz = nz + 2 * g;
    for (int n = 0; n < 5; n++)
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++)
                    saved[(size_t)n * c->plane + (size_t)k * c->plane2 + (size_t)j * c->pitch + i] =
                        c->fixed_ref[(((size_t)n * lz + k + g) * ly + j + g) * lx + i + g];
    if (gk_status s = upload_vec(c, &c->inner_rows, rows)) return s;
    return upload_real(c, &c->inner, saved);
}


extern "C" {

gk_status gk_create(const gk_grid* grid, const gk_phys* phys, gk_ctx** out) {
    if (!grid || !phys || !out) { g_create_error = "null argument"; return GK_EINVAL; }
    *out = nullptr;
    // src/grid.hpp:23-25: non-positive sizes are a configuration error
    if (grid->nx <= 0 || grid->ny <= 0 || grid->nz <= 0 || grid->nghost < 0) {
        g_create_error = "Invalid grid parameters from config"; return GK_EINVAL;
    }
    if (grid->nz > 1) {
        // 3D: ensembles and the learned correction are 2D only
        if (grid->ny < 2) { g_create_error = "3D grids need NY >= 2"; return GK_EINVAL; }
    }
    if (grid->nghost < 2) { g_create_error = "NGhost must be >= 2 for the PLM stencil"; return GK_EINVAL; }
    if (grid->nx < 2) { g_create_error = "NX must be >= 2"; return GK_EINVAL; }
    if (phys->precision != 64 && phys->precision != 32) { g_create_error = "precision must be 64 or 32"; return GK_EINVAL; }
    if (!(phys->gamma > 1.0)) { g_create_error = "gamma must exceed 1"; return GK_EINVAL; }
    if ((long long)grid->nx * grid->ny * grid->nz * 5 >= (1LL << 40)) { g_create_error = "grid too large"; return GK_EINVAL; }

    gk_ctx* c = new gk_ctx();
    c->grid = *grid;
    c->phys = *phys;
    c->prec = phys->precision;
    c->esz = c->prec == 32 ? 4 : 8;
    c->nx = grid->nx; c->ny = grid->ny; c->nz = grid->nz > 1 ? grid->nz : 1; c->pitch = grid->nx;
    c->plane2 = (long long)c->pitch * c->ny;
    c->plane = c->plane2 * c->nz;
    for (int f = 0; f < 6; f++) c->bc[f] = (phys->bc[f] >= 0 && phys->bc[f] <= 4) ? phys->bc[f] : 0;
    if (c->ny == 1) { c->bc[2] = 0; c->bc[3] = 0; }  // y fluxes are multiplied by dyRec = 0
    // 3D: a z face of type 0 (or an unknown code) registers the y copy (boundary.cpp:90-114), which
    // UpdateBD runs after both y faces (boundary.cpp:127-135) over the active x and z range, so the
    // y ghosts on that side are copies of the nearest row whatever the y face's own type
    if (c->nz > 1) {
        if (c->bc[4] == 0) c->bc[2] = 0;
        if (c->bc[5] == 0) c->bc[3] = 0;
    }
    c->gravity = phys->g[0] != 0 || phys->g[1] != 0 || phys->g[2] != 0;
    // Domain (src/sim.cpp:16-36)
    const double dx = (grid->xmax - grid->xmin) / grid->nx;
    const double dy = (grid->ymax - grid->ymin) / grid->ny;
    const double dz = (grid->zmax - grid->zmin) / grid->nz;
    c->dxRec = 1. / dx;
    c->dyRec = grid->ny <= 1 ? 0.0 : 1. / dy;
    c->drmin = smin(smin(dx, dy), dz);
    plan_waves(c, 0);
    build_integrator(c);

    auto bail = [&](gk_status s, const std::string& m) {
        g_create_error = m;
        gk_destroy(c);
        return s;
    };
    if (hipStreamCreate(&c->stream) != hipSuccess) return bail(GK_EHIP, "hipStreamCreate failed");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        c->cus = 0;  // stage kernels then use the default plan
    const size_t bytes = (size_t)c->plane * 5 * c->esz;
    const int nbuf = phys->integrator == 3 ? 3 : 2;
    for (int k = 0; k < nbuf; k++)
        if (hipMalloc(&c->buf[k], bytes) != hipSuccess) return bail(GK_ENOMEM, "state allocation failed");
    if (nbuf == 2) c->buf[2] = c->buf[1];
    if (hipMalloc(&c->clock, sizeof(DevClock)) != hipSuccess ||
        hipMalloc(&c->log, sizeof(StepRecord) * c->log_cap) != hipSuccess ||
        hipMalloc(&c->partial, sizeof(double) * 5 * c->sum_blocks) != hipSuccess ||
        hipMalloc(&c->sums, sizeof(double) * 5) != hipSuccess ||
        hipHostMalloc((void**)&c->hclock, sizeof(DevClock)) != hipSuccess)
        return bail(GK_ENOMEM, "auxiliary allocation failed");
    // Sim::Sim initial scalars (src/sim.cpp:45-47)
    DevClock h;
    std::memset(&h, 0, sizeof h);
    h.t = 0; h.dt = 1e10; h.t_next = 0; h.dt_until_output = 1e10;
    double floor_c = 1e-16;
    for (int k = 0; k < kCmaxSlots; k++) {
        std::memcpy(&h.cmax_bits[0][k], &floor_c, 8);
        std::memcpy(&h.cmax_bits[1][k], &floor_c, 8);
    }
    h.step = 0; h.stepmax = phys->stepmax; h.is_continue = 1; h.active = 0; h.parity = 0; h.cur = 0;
    h.cfl = phys->cfl; h.drmin = c->drmin;
    *c->hclock = h;
    if (hipMemcpy(c->clock, &h, sizeof h, hipMemcpyHostToDevice) != hipSuccess) return bail(GK_EHIP, "clock init failed");
    *out = c;
    return GK_OK;
}

void gk_destroy(gk_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void* seen[3] = {nullptr, nullptr, nullptr};
    for (int k = 0; k < 3; k++) {
        bool dup = false;
        for (int q = 0; q < k; q++) dup |= seen[q] == c->buf[k];
        if (c->buf[k] && !dup) (void)hipFree(c->buf[k]);
        seen[k] = c->buf[k];
    }
    if (c->clock) (void)hipFree(c->clock);
    if (c->log) (void)hipFree(c->log);
    if (c->partial) (void)hipFree(c->partial);
    if (c->sums) (void)hipFree(c->sums);
    if (c->acc) (void)hipFree(c->acc);
    if (c->xc) (void)hipFree(c->xc);
    if (c->yc) (void)hipFree(c->yc);
    if (c->lstats) (void)hipFree(c->lstats);
    for (int k = 0; k < 6; k++)
        if (c->fix[k]) (void)hipFree(c->fix[k]);
    if (c->inner) (void)hipFree(c->inner);
    if (c->inner_rows) (void)hipFree(c->inner_rows);
    if (c->sh_cell) (void)hipFree(c->sh_cell);
    if (c->sh_ntar) (void)hipFree(c->sh_ntar);
    if (c->sh_tar) (void)hipFree(c->sh_tar);
    if (c->sh_dir) (void)hipFree(c->sh_dir);
    if (c->sh_gam) (void)hipFree(c->sh_gam);
    if (c->zcs) (void)hipFree(c->zcs);
    if (c->zstale) (void)hipFree(c->zstale);
    if (c->hclock) (void)hipHostFree(c->hclock);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// 3D: the cells of the nearest z ghost planes the reference never writes keep the uploaded
// state's values in the array that holds it (zero in the other): [2][5][ny + 2][nx], rows -1..ny.
static gk_status set_stale(gk_ctx* c, const std::vector<double>& planes) {
    if (c->nz <= 1) return GK_OK;
    c->stale_base = c->hclock->step;
    return upload_real(c, &c->zstale, planes);
}

gk_status gk_upload_active(gk_ctx* c, const double* host) {
    if (!c || !host) return GK_EINVAL;
    gk_status s = read_clock(c);
    if (s != GK_OK) return s;
    c->hclock->cur = 0;
    if ((s = write_clock(c)) != GK_OK) return s;
    const size_t n = (size_t)c->plane * 5;
    if (c->prec == 32) {
        std::vector<float> tmp(n);
        for (size_t k = 0; k < n; k++) tmp[k] = (float)host[k];
        GK_TRY_HIP(c, hipMemcpy(c->buf[0], tmp.data(), n * 4, hipMemcpyHostToDevice));
    } else {
        GK_TRY_HIP(c, hipMemcpy(c->buf[0], host, n * 8, hipMemcpyHostToDevice));
    }
    c->uploaded = true;
    if ((s = set_stale(c, std::vector<double>((size_t)10 * (c->ny + 2) * c->nx, 0.0))) != GK_OK) return s;
    return refresh_cmax(c);
}

gk_status gk_download_active(gk_ctx* c, double* host) {
    if (!c || !host) return GK_EINVAL;
    gk_status s = read_clock(c);
    if (s != GK_OK) return s;
    const size_t n = (size_t)c->plane * 5;
    const void* src = c->buf[current_buffer(c)];
    if (c->prec == 32) {
        std::vector<float> tmp(n);
        GK_TRY_HIP(c, hipMemcpy(tmp.data(), src, n * 4, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < n; k++) host[k] = tmp[k];
    } else {
        GK_TRY_HIP(c, hipMemcpy(host, src, n * 8, hipMemcpyDeviceToHost));
    }
    return GK_OK;
}

gk_status gk_upload_cons(gk_ctx* c, const void* host) {
    if (!c || !host) return GK_EINVAL;
    std::vector<double> act((size_t)c->plane * 5);
    if (c->prec == 32) ref_to_active(c, (const float*)host, act.data());
    else ref_to_active(c, (const double*)host, act.data());
    gk_status s = gk_upload_active(c, act.data());
    if (s != GK_OK || c->nz <= 1) return s;
    const int g = c->grid.nghost, nx = c->nx, ny = c->ny, nz = c->nz;
    const int lx = nx + 2 * g, ly = ny + 2 * g, lz = nz + 2 * g;
    std::vector<double> planes((size_t)10 * (ny + 2) * nx);
    for (int side = 0; side < 2; side++)
        for (int n = 0; n < 5; n++)
            for (int jj = -1; jj <= ny; jj++)
                for (int i = 0; i < nx; i++) {
                    const size_t src = (((size_t)n * lz + (side ? g + nz : g - 1)) * ly + (g + jj)) * lx + (g + i);
                    const double v = c->prec == 32 ? (double)((const float*)host)[src] : ((const double*)host)[src];
                    planes[(((size_t)side * 5 + n) * (ny + 2) + (jj + 1)) * nx + i] = v;
                }
    return set_stale(c, planes);
}

gk_status gk_download_cons(gk_ctx* c, void* host) {
    if (!c || !host) return GK_EINVAL;
    std::vector<double> act((size_t)c->plane * 5);
    gk_status s = gk_download_active(c, act.data());
    if (s != GK_OK) return s;
    if (c->prec == 32) active_to_ref(c, act.data(), (float*)host);
    else active_to_ref(c, act.data(), (double*)host);
    if (c->hook) {
        const void* cur = c->buf[current_buffer(c)];
        return c->prec == 32 ? hook_fill_output<float>(c, cur, (float*)host)
                             : hook_fill_output<double>(c, cur, (double*)host);
    }
    return GK_OK;
}

gk_status gk_advance(gk_ctx* c, double dtoutput, double* t, int* step) {
    if (!c) return GK_EINVAL;
    if (!c->uploaded) return fail(c, GK_ESTATE, "no state uploaded");
    if (gk_status fs = check_fixed(c)) return fs;
    hipLaunchKernelGGL(advance_begin_kernel, dim3(1), dim3(1), 0, c->stream, c->clock, dtoutput);
    GK_TRY_HIP(c, hipGetLastError());
    gk_status s = read_clock(c);
    if (s != GK_OK) return s;
    int last_step = -1;
    while (c->hclock->active) {
        if (c->hook) {
            if ((s = hooked_steps(c, advance_chunk(*c->hclock))) != GK_OK) return s;
        } else {
            enqueue_steps(c, advance_chunk(*c->hclock));
        }
        GK_TRY_HIP(c, hipGetLastError());
        if ((s = read_clock(c)) != GK_OK) return s;
        if (!c->hclock->active) break;
        // an active clock that did not move over a whole chunk would loop forever
        if (c->hclock->step == last_step) return fail(c, GK_ESTATE, "device clock stopped advancing");
        last_step = c->hclock->step;
    }
    if (t) *t = c->hclock->t;
    if (step) *step = c->hclock->step;
    return GK_OK;
}

gk_status gk_get_time(gk_ctx* c, double* t, double* dt, int* step, int* is_continue) {
    if (!c) return GK_EINVAL;
    gk_status s = read_clock(c);
    if (s != GK_OK) return s;
    if (t) *t = c->hclock->t;
    if (dt) *dt = c->hclock->dt;
    if (step) *step = c->hclock->step;
    if (is_continue) *is_continue = c->hclock->is_continue;
    return GK_OK;
}

gk_status gk_set_time(gk_ctx* c, double t, int step) {
    if (!c || step < 0) return GK_EINVAL;
    gk_status s = read_clock(c);
    if (s != GK_OK) return s;
    c->hclock->t = t;
    c->hclock->step = step;
    c->hclock->is_continue = !(c->hclock->stepmax > 0 && step >= c->hclock->stepmax);
    c->stale_base = step;
    return write_clock(c);
}

gk_status gk_step_log(gk_ctx* c, int* steps, double* t, double* dt, int cap, int* count) {
    if (!c || cap < 0) return GK_EINVAL;
    gk_status s = read_clock(c);
    if (s != GK_OK) return s;
    std::vector<StepRecord> all(c->log_cap);
    GK_TRY_HIP(c, hipMemcpy(all.data(), c->log, sizeof(StepRecord) * c->log_cap, hipMemcpyDeviceToHost));
    const int total = c->hclock->step;
    int n = total < c->log_cap ? total : c->log_cap;
    if (n > cap) n = cap;
    for (int k = 0; k < n; k++) {
        const StepRecord& r = all[(total - n + k) % c->log_cap];
        if (steps) steps[k] = r.step;
        if (t) t[k] = r.t;
        if (dt) dt[k] = r.dt;
    }
    if (count) *count = n;
    return GK_OK;
}

gk_status gk_set_point_gravity(gk_ctx* c, double gm, double x, double y, double z, double vx, double vy, double vz,
                                double rs) {
    if (!c) return GK_EINVAL;
    if (c->net) return fail(c, GK_EINVAL, "a network is attached (gk_set_net): detach it first");
    // SourceTerm::EnrollPointGravity (src/source_term/source.cpp:33-67): the first mass gives point
    // gravity, a second turns it into binary gravity, a third is an error
    if (c->npoint >= 2) return fail(c, GK_EINVAL, "Setup failed: too many point gravity.");
    if (!c->xc) {
        std::vector<double> hx, hy;
        domain_centres(c, hx, hy, c->zc);
        const int lx = (int)hx.size(), ly = (int)hy.size();
        if (c->nz > 1) {
            // every plane's z centre (Domain, src/sim.cpp:39)
            const int g = c->grid.nghost;
            const double dz = (c->grid.zmax - c->grid.zmin) / c->grid.nz;
            std::vector<double> hz((size_t)c->nz + 2 * g);
            for (int k = 0; k < (int)hz.size(); k++) hz[k] = c->grid.zmin - g * dz + 0.5 * dz + k * dz;
            if (gk_status s = upload_vec(c, &c->zcs, hz)) return s;
        }
        GK_TRY_HIP(c, hipMalloc(&c->xc, sizeof(double) * lx));
        GK_TRY_HIP(c, hipMalloc(&c->yc, sizeof(double) * ly));
        GK_TRY_HIP(c, hipMemcpy(c->xc, hx.data(), sizeof(double) * lx, hipMemcpyHostToDevice));
        GK_TRY_HIP(c, hipMemcpy(c->yc, hy.data(), sizeof(double) * ly, hipMemcpyHostToDevice));
    }
    double* m = c->pm[c->npoint];
    m[0] = gm; m[1] = x; m[2] = y; m[3] = z; m[4] = vx; m[5] = vy; m[6] = vz; m[7] = rs;
    if (c->npoint == 1) {
        const double* m0 = c->pm[0];
        c->ab = std::sqrt((x - m0[1]) * (x - m0[1]) + (y - m0[2]) * (y - m0[2]) + (z - m0[3]) * (z - m0[3]));
        c->omega = std::sqrt((m0[0] + gm) / (c->ab * c->ab * c->ab));
        c->theta1 = std::atan2(y - m0[2], x - m0[1]);
    }
    c->npoint++;
    c->gravity = true;
    return GK_OK;
}

gk_status gk_set_fixed_boundary(gk_ctx* c, const void* host, int fix_inner) {
    if (!c || !host) return GK_EINVAL;
    if (c->net) return fail(c, GK_EINVAL, "a network is attached (gk_set_net): detach it first");
    // Boundary::EnrollFixedBoundaryData (src/boundary/fixed.cpp:242-269): the ghost layers of the
    // given state become the fixed faces' values; the stencil reads the two layers nearest each x
    // and y face on every active plane, and the nearest plane of each z face
    const int g = c->grid.nghost, nx = c->nx, ny = c->ny, nz = c->nz;
    const int lx = nx + 2 * g, ly = ny + 2 * g, lz = nz + 2 * g;
    const size_t total = 5 * (size_t)lx * ly * lz;
    c->fixed_ref.resize(total);
    for (size_t k = 0; k < total; k++)
        c->fixed_ref[k] = c->prec == 32 ? (double)((const float*)host)[k] : ((const double*)host)[k];
    auto ref = [&](int n, int k, int j, int i) { return c->fixed_ref[(((size_t)n * lz + k) * ly + j) * lx + i]; };
    const int nf = nfaces(c);
    std::vector<double> face[6];
    face[0].resize(10 * (size_t)ny * nz); face[1].resize(10 * (size_t)ny * nz);
    face[2].resize(10 * (size_t)nx * nz); face[3].resize(10 * (size_t)nx * nz);
    if (nf == 6) { face[4].resize(5 * (size_t)nx * ny); face[5].resize(5 * (size_t)nx * ny); }
    for (int n = 0; n < 5; n++)
        for (int k = 0; k < nz; k++)
            for (int d = 0; d < 2; d++) {
                for (int j = 0; j < ny; j++) {
                    face[0][(((size_t)n * nz + k) * ny + j) * 2 + d] = ref(n, k + g, j + g, g - 1 - d);
                    face[1][(((size_t)n * nz + k) * ny + j) * 2 + d] = ref(n, k + g, j + g, g + nx + d);
                }
                for (int i = 0; i < nx; i++) {
                    face[2][(((size_t)n * nz + k) * 2 + d) * nx + i] = ref(n, k + g, g - 1 - d, i + g);
                    face[3][(((size_t)n * nz + k) * 2 + d) * nx + i] = ref(n, k + g, g + ny + d, i + g);
                }
            }
    if (nf == 6)
        for (int n = 0; n < 5; n++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    face[4][((size_t)n * ny + j) * nx + i] = ref(n, g - 1, j + g, i + g);
                    face[5][((size_t)n * ny + j) * nx + i] = ref(n, g + nz, j + g, i + g);
                }
    GK_TRY_HIP(c, hipStreamSynchronize(c->stream));
    for (int f = 0; f < nf; f++) {
        const size_t cnt = face[f].size();
        if (!c->fix[f]) GK_TRY_HIP(c, hipMalloc(&c->fix[f], cnt * c->esz));
        if (c->prec == 32) {
            std::vector<float> tmp(face[f].begin(), face[f].end());
            GK_TRY_HIP(c, hipMemcpy(c->fix[f], tmp.data(), cnt * 4, hipMemcpyHostToDevice));
        } else {
            GK_TRY_HIP(c, hipMemcpy(c->fix[f], face[f].data(), cnt * 8, hipMemcpyHostToDevice));
        }
    }
    return fix_inner ? set_inner(c) : GK_OK;
}

gk_status gk_set_immersed_shell(gk_ctx* c, double xs, double ys, double zs, double rs) {
    if (!c) return GK_EINVAL;
    if (c->net) return fail(c, GK_EINVAL, "a network is attached (gk_set_net): detach it first");
    // Boundary::SetupImmersedShell (src/boundary/immersed.cpp:89-235); in 2D the single plane
    if (c->nx <= 1 || c->ny <= 1) return fail(c, GK_EINVAL, "immersed boundary has to be at least 2D. ");
    const bool three = c->nz > 1;
    const int g = c->grid.nghost, ib = g, ie = g + c->nx, jb = g, je = g + c->ny, kb = g, ke = g + c->nz;
    std::vector<double> xc, yc;
    double zc0;
    domain_centres(c, xc, yc, zc0);
    const double dx = (c->grid.xmax - c->grid.xmin) / c->grid.nx;
    const double dy = (c->grid.ymax - c->grid.ymin) / c->grid.ny;
    const double dz = (c->grid.zmax - c->grid.zmin) / c->grid.nz;
    // z centres, ghost-inclusive index (Domain, src/sim.cpp:39); 2D uses the active plane's
    std::vector<double> zc((size_t)c->nz + 2 * g);
    for (int k = 0; k < (int)zc.size(); k++) zc[k] = three ? c->grid.zmin - g * dz + 0.5 * dz + k * dz : zc0;
    // first cell at or past the body's extent from the low end / last at or before it from the
    // high end (FindLeftId, FindRightId: immersed.cpp:19-47)
    bool outside = false;
    auto first_from = [&](const std::vector<double>& cc, int b, int e, double centre, double lo) {
        const double edge = centre - 1.2 * rs;
        if (edge < lo) outside = true;
        for (int i = b; i < e; i++)
            if (cc[i] >= edge) return i;
        return e;
    };
    auto last_to = [&](const std::vector<double>& cc, int b, int e, double centre, double hi) {
        const double edge = centre + 1.2 * rs;
        if (edge > hi) outside = true;
        for (int i = e - 1; i >= b; i--)
            if (cc[i] <= edge) return i;
        return b;
    };
    const int il = first_from(xc, ib, ie, xs, c->grid.xmin), ir = last_to(xc, ib, ie, xs, c->grid.xmax);
    const int jl = first_from(yc, jb, je, ys, c->grid.ymin), jr = last_to(yc, jb, je, ys, c->grid.ymax);
    int kl = kb, kr = ke;
    if (three) {
        kl = first_from(zc, kb, ke, zs, c->grid.zmin);
        kr = last_to(zc, kb, ke, zs, c->grid.zmax);
    }
    if (outside || il > ir || jl > jr || kl > kr) return fail(c, GK_EINVAL, "immersed boundary out of domain");
    // state offset of ghost-inclusive cell (i, j, k): planes of ny rows follow each other
    auto offset = [&](int i, int j, int k) {
        return (long long)(k - g) * c->plane2 + (long long)(j - g) * c->pitch + (i - g);
    };
    // shell cells: inside the body with a neighbour nGhost cells away (in one axis) outside it
    // (IsShellCell: immersed.cpp:49-69)
    const double sx = g * dx, sy = g * dy, sz = g * dz;
    auto outside_at = [&](double ax, double ay, double az) { return std::sqrt(ax * ax + ay * ay + az * az) > rs; };
    std::vector<int> si, sj, sk;
    for (int k = kl; k < kr; k++)
        for (int j = jl; j < jr; j++)
            for (int i = il; i < ir; i++) {
                const double ax = xc[i] - xs, ay = yc[j] - ys, az = zc[k] - zs;
                if (!(std::sqrt(ax * ax + ay * ay + az * az) < rs)) continue;
                if (outside_at(ax + sx, ay, az) || outside_at(ax - sx, ay, az) || outside_at(ax, ay + sy, az) ||
                    outside_at(ax, ay - sy, az) || outside_at(ax, ay, az + sz) || outside_at(ax, ay, az - sz)) {
                    si.push_back(i); sj.push_back(j); sk.push_back(k);
                }
            }
    const int ns = (int)si.size();
    std::vector<long long> cell(ns), tar((size_t)kShellTargets * ns, 0);
    std::vector<int> ntar(ns);
    std::vector<double> dir((size_t)4 * ns), gam((size_t)(kShellTargets + 1) * ns, 0.0);
    const int reach = 2 * g, zreach = three ? 2 * g : 0;
    for (int s = 0; s < ns; s++) {
        const int i = si[s], j = sj[s], k = sk[s];
        cell[s] = offset(i, j, k);
        const double ax = xc[i] - xs, ay = yc[j] - ys, az = zc[k] - zs;
        const double rnow = std::sqrt(ax * ax + ay * ay + az * az);
        const double rxy = std::sqrt(ax * ax + ay * ay);
        double cosp = 1.0, sinp = 0.0, cost = 1.0, sint = 0.0;
        if (!(rxy < 1e-15)) { cosp = ax / rxy; sinp = ay / rxy; }
        if (!(rnow < 1e-15)) { cost = az / rnow; sint = rxy / rnow; }
        dir[s] = cosp; dir[ns + s] = sinp; dir[2 * ns + s] = cost; dir[3 * ns + s] = sint;
        // mirror point of the cell through the surface, and the fluid cells within one cell of it
        const double xi = sint * cosp * (2 * rs - rnow) + xs;
        const double yi = sint * sinp * (2 * rs - rnow) + ys;
        const double zi = cost * (2 * rs - rnow) + zs;
        double dist[kShellTargets + 1];
        dist[0] = 2 * (rs - rnow);
        int nt = 0;
        for (int di = -reach; di <= reach; di++)
            for (int dj = -reach; dj <= reach; dj++)
                for (int dk = -zreach; dk <= zreach; dk++) {
                    const int ti = i + di, tj = j + dj, tk = k + dk;
                    if (ti < ib || ti >= ie || tj < jb || tj >= je || tk < kb || tk >= ke)
                        return fail(c, GK_EINVAL, "immersed body is too close to the domain boundary, "
                                                  "neighboor search out of range.");
                    const double tx = xc[ti], ty = yc[tj], tz = zc[tk];
                    if (!(std::abs(tx - xi) <= dx && std::abs(ty - yi) <= dy && std::abs(tz - zi) <= dz)) continue;
                    const double rt =
                        std::sqrt((tx - xs) * (tx - xs) + (ty - ys) * (ty - ys) + (tz - zs) * (tz - zs));
                    if (!(rt > rs)) continue;
                    if (++nt > kShellTargets) return fail(c, GK_EINVAL, "n of imagine points is more than nTarMax");
                    tar[(size_t)(nt - 1) * ns + s] = offset(ti, tj, tk);
                    dist[nt] = std::sqrt((tx - xi) * (tx - xi) + (ty - yi) * (ty - yi) + (tz - zi) * (tz - zi));
                }
        ntar[s] = nt;
        // inverse-square distance weights normalised over the cell and its targets
        double sum = 0.;
        for (int n = 0; n <= nt; n++) sum = sum + 1. / (dist[n] * dist[n]);
        for (int n = 0; n <= nt; n++) gam[(size_t)n * ns + s] = 1. / (dist[n] * dist[n] * sum);
    }
    GK_TRY_HIP(c, hipStreamSynchronize(c->stream));
    for (void** p : {(void**)&c->sh_cell, (void**)&c->sh_ntar, (void**)&c->sh_tar, &c->sh_dir, &c->sh_gam})
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    gk_status s;
    if ((s = upload_vec(c, &c->sh_cell, cell)) || (s = upload_vec(c, &c->sh_ntar, ntar)) ||
        (s = upload_vec(c, &c->sh_tar, tar)) || (s = upload_real(c, &c->sh_dir, dir)) ||
        (s = upload_real(c, &c->sh_gam, gam)))
        return s;
    c->nshell = ns;
    c->immersed = true;
    return GK_OK;
}

gk_status gk_set_bd_hook(gk_ctx* c, int face_mask, int band, gk_bd_hook hook, void* user, void* mirror) {
    if (!c) return GK_EINVAL;
    if (!hook || !mirror) return fail(c, GK_EINVAL, "gk_set_bd_hook: NULL hook or mirror");
    const int nf = nfaces(c);
    if (face_mask <= 0 || face_mask >= (1 << nf))
        return fail(c, GK_EINVAL, nf == 6 ? "gk_set_bd_hook: face_mask selects faces 0-5"
                                          : "gk_set_bd_hook: face_mask selects faces 0-3 (z faces need NZ > 1)");
    if ((face_mask & 12) && c->ny < 2) return fail(c, GK_EINVAL, "gk_set_bd_hook: y faces of a one-row grid");
    if (band < 1 || band > c->nx || band > c->ny || (nf == 6 && band > c->nz))
        return fail(c, GK_EINVAL, "gk_set_bd_hook: band must lie in [1, min(nx, ny, nz)]");
    if (c->net) return fail(c, GK_EINVAL, "a network is attached (gk_set_net): detach it first");
    // a hooked face became a type-4 face; without its hook it would keep its last ghost layers
    if (c->hook && (c->hook_mask & ~face_mask))
        return fail(c, GK_EINVAL, "gk_set_bd_hook: a later call must keep every face an earlier call hooked");
    // 3D: a z face of type 0 registers the y copy, which UpdateBD runs after both y faces
    // (boundary.cpp:90-135): it overwrites a hooked y face's ghost layers on its side, so that y
    // face stays the copy gk_create made it.  A hooked z face replaces the copy, and the y face on
    // its side gets its own type back.
    int mask = face_mask;
    if (nf == 6)
        for (int s = 0; s < 2; s++) {
            if ((face_mask >> (4 + s)) & 1) {
                if (c->bc[4 + s] == 0) {
                    const int t = c->phys.bc[2 + s];
                    c->bc[2 + s] = t >= 0 && t <= 4 ? t : 0;
                }
            } else if (c->bc[4 + s] == 0) {
                mask &= ~(4 << s);
            }
        }
    GK_TRY_HIP(c, hipStreamSynchronize(c->stream));
    for (int f = 0; f < nf; f++) {
        if (!((mask >> f) & 1)) continue;
        const size_t len = f < 2 ? (size_t)10 * c->nz * c->ny : f < 4 ? (size_t)10 * c->nz * c->nx
                                                                     : (size_t)5 * c->ny * c->nx;
        if (!c->fix[f]) GK_TRY_HIP(c, hipMalloc(&c->fix[f], c->esz * len));
        c->bc[f] = 4;
    }
    c->hook = hook;
    c->hook_user = user;
    c->hook_mirror = mirror;
    c->hook_mask = mask;
    c->hook_band = band;
    c->plans.clear();   // the fixed-face kernel variant runs from now on
    return GK_OK;
}

gk_status gk_update_bd(gk_ctx* c) {
    if (!c) return GK_EINVAL;
    if (!c->uploaded) return fail(c, GK_ESTATE, "no state uploaded");
    if (gk_status fs = check_fixed(c)) return fs;
    // Boundary::UpdateBD's in-place parts on the current state (src/boundary/boundary.cpp:127-142):
    // the immersed shell, then the inner fixed region; ghost cells live only in the stage loads
    gk_status s = read_clock(c);
    if (s != GK_OK) return s;
    void* u = c->buf[current_buffer(c)];
    if (c->nshell) {
        if (c->prec == 32) launch_immersed_t<float>(c, u, u, 0, 1);
        else launch_immersed_t<double>(c, u, u, 0, 1);
    }
    const int rows = c->inner_j1 - c->inner_j0, width = c->inner_i1 - c->inner_i0;
    if (c->inner && rows > 0 && width > 0) {
        const long long n = (long long)rows * width;
        const dim3 grid((unsigned)((n + 255) / 256));
        if (c->prec == 32)
            hipLaunchKernelGGL(inner_reset_kernel<float>, grid, dim3(256), 0, c->stream, (float*)u,
                               (const float*)c->inner, c->inner_rows, c->plane, c->pitch, c->inner_j0, rows,
                               c->inner_i0, width);
        else
            hipLaunchKernelGGL(inner_reset_kernel<double>, grid, dim3(256), 0, c->stream, (double*)u,
                               (const double*)c->inner, c->inner_rows, c->plane, c->pitch, c->inner_j0, rows,
                               c->inner_i0, width);
    }
    GK_TRY_HIP(c, hipGetLastError());
    return refresh_cmax(c);
}

gk_status gk_run_steps(gk_ctx* c, int nsteps) {
    if (!c || nsteps < 0) return GK_EINVAL;
    if (!c->uploaded) return fail(c, GK_ESTATE, "no state uploaded");
    if (gk_status fs = check_fixed(c)) return fs;
    hipLaunchKernelGGL(advance_begin_kernel, dim3(1), dim3(1), 0, c->stream, c->clock, 1e300);
    if (c->hook) {
        if (gk_status s = hooked_steps(c, nsteps)) return s;
    } else {
        enqueue_steps(c, nsteps);
    }
    GK_TRY_HIP(c, hipGetLastError());
    return GK_OK;
}

gk_status gk_synchronize(gk_ctx* c) {
    if (!c) return GK_EINVAL;
    GK_TRY_HIP(c, hipStreamSynchronize(c->stream));
    return GK_OK;
}

gk_status gk_profile_steps(gk_ctx* c, int nsteps, double* total_ms, double kernel_ms[4], int launches[4]) {
    if (!c || nsteps < 0) return GK_EINVAL;
    if (!c->uploaded) return fail(c, GK_ESTATE, "no state uploaded");
    if (c->hook) return fail(c, GK_EINVAL, "per-launch profiling is not available with a boundary hook");
    if (gk_status fs = check_fixed(c)) return fs;
    hipLaunchKernelGGL(advance_begin_kernel, dim3(1), dim3(1), 0, c->stream, c->clock, 1e300);
    GK_TRY_HIP(c, hipStreamSynchronize(c->stream));
    // an event after every launch; the interval ending at a launch's event is its duration.
    // classes: 0..2 the stage kernels by position in the step, 3 the clock kernels.  Every event is
    // created before the first launch, so the timed region holds only launches and records.
    const size_t per_step = 1 + c->stages.size() * (c->nshell ? 2 : 1);
    std::vector<hipEvent_t> ev((size_t)nsteps * per_step + 2, nullptr);
    for (hipEvent_t& e : ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            for (hipEvent_t& d : ev)
                if (d) (void)hipEventDestroy(d);
            return fail(c, GK_EHIP, "hipEventCreate failed");
        }
    }
    std::vector<int> kind;
    kind.reserve(ev.size());
    size_t nrec = 0;
    auto rec = [&]() { (void)hipEventRecord(ev[nrec++], c->stream); };
    rec();
    const int toggle = c->phys.integrator == 1 ? 1 : 0;
    for (int k = 0; k < nsteps; k++) {
        // the same launches as enqueue_steps
        if (k == 0) hipLaunchKernelGGL(step_begin_kernel, dim3(1), dim3(1), 0, c->stream, c->clock);
        else hipLaunchKernelGGL(step_seam_kernel, dim3(1), dim3(1), 0, c->stream, c->clock, c->log, c->log_cap, toggle);
        rec(); kind.push_back(3);
        for (size_t q = 0; q < c->stages.size(); q++) {
            if (c->nshell) {
                launch_immersed(c, c->stages[q]);
                rec(); kind.push_back(3);
            }
            launch_stage_any(c, c->stages[q]);
            rec(); kind.push_back((int)q);
        }
    }
    if (nsteps > 0) {
        hipLaunchKernelGGL(step_end_kernel, dim3(1), dim3(1), 0, c->stream, c->clock, c->log, c->log_cap, toggle);
        rec(); kind.push_back(3);
    }
    GK_TRY_HIP(c, hipGetLastError());
    GK_TRY_HIP(c, hipStreamSynchronize(c->stream));
    double kms[4] = {0, 0, 0, 0};
    int cnt[4] = {0, 0, 0, 0};
    for (size_t q = 0; q + 1 < nrec; q++) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ev[q], ev[q + 1]);
        kms[kind[q]] += ms;
        cnt[kind[q]] += 1;
    }
    float tot = 0;
    if (nrec > 1) (void)hipEventElapsedTime(&tot, ev[0], ev[nrec - 1]);
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    if (total_ms) *total_ms = tot;
    for (int k = 0; k < 4; k++) {
        if (kernel_ms) kernel_ms[k] = kms[k];
        if (launches) launches[k] = cnt[k];
    }
    return GK_OK;
}

gk_status gk_cons_sums(gk_ctx* c, double out[5]) {
    if (!c || !out) return GK_EINVAL;
    gk_status s = read_clock(c);
    if (s != GK_OK) return s;
    const void* u = c->buf[current_buffer(c)];
    if (c->prec == 32)
        hipLaunchKernelGGL(sums_partial_kernel<float>, dim3(c->sum_blocks), dim3(256), 0, c->stream, (const float*)u,
                           c->plane, c->pitch, c->nx, c->ny * c->nz, c->partial);
    else
        hipLaunchKernelGGL(sums_partial_kernel<double>, dim3(c->sum_blocks), dim3(256), 0, c->stream, (const double*)u,
                           c->plane, c->pitch, c->nx, c->ny * c->nz, c->partial);
    hipLaunchKernelGGL(sums_final_kernel, dim3(1), dim3(64), 0, c->stream, c->partial, c->sum_blocks, c->sums);
    GK_TRY_HIP(c, hipGetLastError());
    GK_TRY_HIP(c, hipMemcpyAsync(out, c->sums, sizeof(double) * 5, hipMemcpyDeviceToHost, c->stream));
    GK_TRY_HIP(c, hipStreamSynchronize(c->stream));
    return GK_OK;
}

gk_status gk_cons_sums_noalloc(const void* state, int nx, int ny, int nz, int precision, double* out, void* stream) {
    if (!state || !out || nx <= 0 || ny <= 0 || nz <= 0 || (precision != 64 && precision != 32)) {
        g_create_error = "gk_cons_sums_noalloc: null pointer, non-positive size or precision not 64/32";
        return GK_EINVAL;
    }
    // the 2560 partials do not depend on the input's size; the element count bounds the indices
    if ((long long)nx * ny * nz >= (1LL << 40)) {
        g_create_error = "gk_cons_sums_noalloc: state larger than 2^40 cells";
        return GK_EINVAL;
    }
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        g_create_error = "gk_cons_sums_noalloc: no current device (or its index exceeds 63)";
        return GK_EHIP;
    }
    double* partial = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_dev[dev]) {
            void* p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                g_create_error = std::string("gk_cons_sums_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e);
                return GK_EHIP;
            }
            g_scratch_dev[dev] = static_cast<double*>(p);
        }
        partial = g_scratch_dev[dev];
    }
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const long long plane = (long long)nx * ny * nz;
    // the kernels of gk_cons_sums: every block writes its five partials, so nothing needs clearing
    if (precision == 32)
        hipLaunchKernelGGL(sums_partial_kernel<float>, dim3(kSumBlocks), dim3(256), 0, st, (const float*)state, plane,
                           nx, nx, ny * nz, partial);
    else
        hipLaunchKernelGGL(sums_partial_kernel<double>, dim3(kSumBlocks), dim3(256), 0, st, (const double*)state,
                           plane, nx, nx, ny * nz, partial);
    hipLaunchKernelGGL(sums_final_kernel, dim3(1), dim3(64), 0, st, partial, kSumBlocks, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_create_error = std::string("gk_cons_sums_noalloc: launch: ") + hipGetErrorString(e);
        return GK_EHIP;
    }
    return GK_OK;
}

gk_status gk_check_finite(gk_ctx* c) {
    if (!c) return GK_EINVAL;
    std::vector<double> act((size_t)c->plane * 5);
    gk_status s = gk_download_active(c, act.data());
    if (s != GK_OK) return s;
    for (double v : act)
        if (!std::isfinite(v)) return fail(c, GK_ENUMERIC, "non-finite conserved value");
    return GK_OK;
}

gk_status gk_set_tile(gk_ctx* c, int seg_rows) {
    if (!c || seg_rows < 0) return GK_EINVAL;
    plan_waves(c, seg_rows);
    return GK_OK;
}

gk_status gk_get_tile(gk_ctx* c, int* seg_rows, int* strips, int* waves) {
    if (!c) return GK_EINVAL;
    if (seg_rows) *seg_rows = c->seg_rows;
    if (strips) *strips = c->strips;
    if (waves) *waves = c->nwaves;
    return GK_OK;
}

gk_status gk_selftest(long long count, unsigned long long seed, int emax, long long* mismatches) {
    if (count < 0 || emax < 1 || emax > 500 || !mismatches) {
        g_create_error = "gk_selftest: bad argument";
        return GK_EINVAL;
    }
    unsigned long long* d = nullptr;
    if (hipMalloc(&d, sizeof(unsigned long long)) != hipSuccess) { g_create_error = "selftest alloc"; return GK_ENOMEM; }
    unsigned long long h = 0;
    gk_status st = GK_OK;
    if (hipMemcpy(d, &h, sizeof h, hipMemcpyHostToDevice) != hipSuccess) st = GK_EHIP;
    if (st == GK_OK) {
        hipLaunchKernelGGL(selftest_kernel, dim3(1024), dim3(256), 0, 0, count, seed, emax, d);
        if (hipGetLastError() != hipSuccess || hipMemcpy(&h, d, sizeof h, hipMemcpyDeviceToHost) != hipSuccess)
            st = GK_EHIP;
    }
    (void)hipFree(d);
    if (st != GK_OK) { g_create_error = "gk_selftest: HIP failure"; return st; }
    *mismatches = (long long)h;
    return GK_OK;
}

const char* gk_last_error(const gk_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

const char* gk_version(void) {
    return "gaukuk_amd 0.1 gfx950 eos=" GK_STR(GK_EOS) " flux=" GK_STR(GK_FLUX);
}

}  // extern "C"

#include "gk_ensemble.hip"
#include "gk_learned.hip"
