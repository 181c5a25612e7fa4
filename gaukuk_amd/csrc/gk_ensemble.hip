// gk_ensemble.hip -- host runtime of the ensemble entry points (gk_ens_* in include/gaukuk.h).
// Included by gk_capi.hip after the single-grid runtime, whose helpers it reuses.
//
// All problems share one state arena (per problem: two or three buffers of 5 planes of its
// active cells), one descriptor table and one clock per problem.  A step is: one clock kernel
// for every problem (dt from each problem's own max wave speed; rewinds the work-queue heads),
// one persistent launch per RK stage over all problems' work items, one clock kernel for the t
// update.  The same step can be run with one single-grid launch per problem per stage (the
// baseline the batching is measured against) or captured once as a HIP graph and replayed.

struct gk_ens {
    int P = 0;
    int prec = 64;
    size_t esz = 8;
    gk_phys common;
    std::vector<gk_grid> grids;
    std::vector<gk_phys> phys;
    std::vector<EnsProblem> hprobs;
    EnsProblem* dprobs = nullptr;
    void* arena = nullptr;
    size_t arena_elems = 0;
    DevClock* clocks = nullptr;
    DevClock* hclocks = nullptr;  // pinned mirror
    int* counters = nullptr;      // work-queue heads: kEnsQueues per stage of a step
    unsigned long long launches = 0;   // stage launches enqueued so far (queue direction)
    hipStream_t stream = nullptr;
    int total_items = 0;
    int xcd_begin[kEnsQueues + 1] = {0};
    int seg_rows = 0;
    int waves_per_cu = 12;
    int blocks = 0;
    bool gravity = false;
    std::vector<StageSpec> stages;
    std::vector<gk_ctx> views;    // per-problem single-grid views (no allocations) for the looped mode
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_steps = 0;
    std::string err;
};

static thread_local std::string g_ens_error;

#define GK_ENS_TRY(e, call)                                                                \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            (e)->err = std::string(#call) + ": " + hipGetErrorString(e_);                  \
            return GK_EHIP;                                                                \
        }                                                                                  \
    } while (0)

static constexpr int kNumCUs = 256;  // MI355X

// Work items and descriptor offsets for the current tile choice.
static void ens_plan(gk_ens* e) {
    int base = 0;
    for (int p = 0; p < e->P; p++) {
        EnsProblem& d = e->hprobs[p];
        d.strips = (d.nx + kStrip - 1) / kStrip;
        int s = e->seg_rows > 0 ? e->seg_rows : 16;
        if (s > d.ny) s = d.ny;
        d.seg_rows = s;
        const int segs = (d.ny + s - 1) / s;
        d.items = d.strips * segs;
        d.item_base = base;
        base += d.items;
    }
    e->total_items = base;
    // kEnsQueues contiguous, equal shares of the items, one per XCD queue
    for (int x = 0; x <= kEnsQueues; x++) e->xcd_begin[x] = (int)((long long)base * x / kEnsQueues);
    e->blocks = kNumCUs * e->waves_per_cu / (kBlock / kWave);
    // the looped baseline runs each problem as a single grid would run alone: every stage kernel
    // with the automatic plan for that grid's size at that kernel's occupancy (stage_plan)
    for (int p = 0; p < e->P; p++) plan_waves(&e->views[p], 0);
}

static gk_status ens_upload_table(gk_ens* e) {
    GK_ENS_TRY(e, hipMemcpyAsync(e->dprobs, e->hprobs.data(), sizeof(EnsProblem) * e->P, hipMemcpyHostToDevice,
                                 e->stream));
    GK_ENS_TRY(e, hipStreamSynchronize(e->stream));
    return GK_OK;
}

template <class R, class L, bool THREE>
static void ens_launch_t(gk_ens* e, const EnsArgs<R>& a, bool last) {
    dim3 grid(e->blocks), block(kBlock);
    if (e->gravity) {
        if (last) hipLaunchKernelGGL((ens_stage_kernel<R, L, THREE, true, true>), grid, block, 0, e->stream, a);
        else hipLaunchKernelGGL((ens_stage_kernel<R, L, THREE, true, false>), grid, block, 0, e->stream, a);
    } else {
        if (last) hipLaunchKernelGGL((ens_stage_kernel<R, L, THREE, false, true>), grid, block, 0, e->stream, a);
        else hipLaunchKernelGGL((ens_stage_kernel<R, L, THREE, false, false>), grid, block, 0, e->stream, a);
    }
}

template <class R>
static void ens_launch_stage(gk_ens* e, size_t q) {
    const StageSpec& s = e->stages[q];
    EnsArgs<R> a;
    std::memset(&a, 0, sizeof a);
    a.arena = (R*)e->arena;
    a.probs = e->dprobs;
    a.clocks = e->clocks;
    a.counter = e->counters + kEnsQueues * q;
    for (int x = 0; x <= kEnsQueues; x++) a.xcd_begin[x] = e->xcd_begin[x];
    a.nprob = e->P;
    a.total_items = e->total_items;
    // Alternate the queues' direction from one stage launch to the next, so that each stage starts
    // with the items the previous one touched last, whose states are the likeliest to be still in
    // the 256 MB memory-side cache (a 512-problem ensemble streams about 0.8 GB per state).
    // Items of a stage are independent: the order does not change results.
    a.reverse = (int)(e->launches++ & 1);
    a.in_buf = s.in; a.base_buf = s.base; a.out_buf = s.out;
    a.proto = make_args<R>(&e->views[0], s);  // stage constants; per-problem fields set on device
    const bool plm = e->common.rc_order != 1;
    if (s.three) {
        if (plm) ens_launch_t<R, PlmLimiter, true>(e, a, s.last);
        else ens_launch_t<R, FirstOrder, true>(e, a, s.last);
    } else {
        if (plm) ens_launch_t<R, PlmLimiter, false>(e, a, s.last);
        else ens_launch_t<R, FirstOrder, false>(e, a, s.last);
    }
}

static void ens_enqueue_step(gk_ens* e) {
    const int nthreads = 256, nb = (e->P + nthreads - 1) / nthreads;
    hipLaunchKernelGGL(ens_step_begin_kernel, dim3(nb), dim3(nthreads), 0, e->stream, e->clocks, e->P, e->counters,
                       (int)e->stages.size() * kEnsQueues);
    for (size_t q = 0; q < e->stages.size(); q++) {
        if (e->prec == 32) ens_launch_stage<float>(e, q);
        else ens_launch_stage<double>(e, q);
    }
    hipLaunchKernelGGL(ens_step_end_kernel, dim3(nb), dim3(nthreads), 0, e->stream, e->clocks, e->P,
                       e->common.integrator == 1 ? 1 : 0);
}

// The baseline: every problem as its own single-grid run, one launch per stage per problem.
static void ens_enqueue_step_looped(gk_ens* e) {
    for (int p = 0; p < e->P; p++) {
        gk_ctx* v = &e->views[p];
        hipLaunchKernelGGL(step_begin_kernel, dim3(1), dim3(1), 0, e->stream, v->clock);
        for (const StageSpec& s : v->stages) launch_stage_any(v, s);
        hipLaunchKernelGGL(ens_step_end_kernel, dim3(1), dim3(1), 0, e->stream, v->clock, 1,
                           v->phys.integrator == 1 ? 1 : 0);
    }
}

static void ens_release_graph(gk_ens* e) {
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->gexec = nullptr;
    e->graph = nullptr;
    e->graph_steps = 0;
}

extern "C" {

gk_status gk_ens_create(int count, const gk_grid* grids, const gk_phys* phys, gk_ens** out) {
    if (count <= 0 || !grids || !phys || !out) { g_ens_error = "bad argument"; return GK_EINVAL; }
    *out = nullptr;
    const gk_phys& c0 = phys[0];
    for (int p = 0; p < count; p++) {
        const gk_grid& g = grids[p];
        const gk_phys& f = phys[p];
        if (g.nx < 2 || g.ny <= 0 || g.nz != 1 || g.nghost < 2) { g_ens_error = "problem " + std::to_string(p) + ": bad grid"; return GK_EINVAL; }
        if (f.precision != c0.precision || f.integrator != c0.integrator || f.rc_order != c0.rc_order ||
            f.gamma != c0.gamma || f.g[0] != c0.g[0] || f.g[1] != c0.g[1] || f.g[2] != c0.g[2]) {
            g_ens_error = "all problems of an ensemble share precision, integrator, RcOrder, gamma and gravity";
            return GK_EINVAL;
        }
        for (int q = 0; q < 4; q++)
            if (f.bc[q] == 4) { g_ens_error = "fixed boundaries (type 4) are not supported"; return GK_EINVAL; }
    }
    if (c0.precision != 64 && c0.precision != 32) { g_ens_error = "precision must be 64 or 32"; return GK_EINVAL; }
    gk_ens* e = new gk_ens();
    e->P = count;
    e->prec = c0.precision;
    e->esz = e->prec == 32 ? 4 : 8;
    e->common = c0;
    e->gravity = c0.g[0] != 0 || c0.g[1] != 0 || c0.g[2] != 0;
    e->grids.assign(grids, grids + count);
    e->phys.assign(phys, phys + count);
    e->hprobs.resize(count);
    e->views.resize(count);
    const int nbuf = c0.integrator == 3 ? 3 : 2;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 0;  // looped launches then use the default plan
    size_t off = 0;
    for (int p = 0; p < count; p++) {
        EnsProblem& d = e->hprobs[p];
        std::memset(&d, 0, sizeof d);
        d.nx = grids[p].nx; d.ny = grids[p].ny; d.pitch = d.nx;
        d.plane = (long long)d.nx * d.ny;
        for (int k = 0; k < 3; k++) {
            d.off[k] = (long long)off;
            if (k < nbuf) off += (size_t)d.plane * 5;
        }
        if (nbuf == 2) d.off[2] = d.off[1];
        for (int q = 0; q < 4; q++) d.bc[q] = (phys[p].bc[q] >= 0 && phys[p].bc[q] <= 3) ? phys[p].bc[q] : 0;
        if (d.ny == 1) { d.bc[2] = 0; d.bc[3] = 0; }
        const double dx = (grids[p].xmax - grids[p].xmin) / grids[p].nx;
        const double dy = (grids[p].ymax - grids[p].ymin) / grids[p].ny;
        d.dxRec = 1. / dx;
        d.dyRec = grids[p].ny <= 1 ? 0.0 : 1. / dy;
        // a single-grid view of the problem for the looped mode (shares the ensemble's memory)
        gk_ctx& v = e->views[p];
        v.grid = grids[p]; v.phys = phys[p]; v.prec = e->prec; v.esz = e->esz;
        v.nx = d.nx; v.ny = d.ny; v.pitch = d.pitch; v.plane = d.plane; v.plane2 = d.plane;
        v.cus = cus;
        for (int q = 0; q < 4; q++) v.bc[q] = d.bc[q];
        v.gravity = e->gravity;
        v.dxRec = d.dxRec; v.dyRec = d.dyRec;
        build_integrator(&v);
    }
    e->arena_elems = off;
    build_integrator(&e->views[0]);
    e->stages = e->views[0].stages;
    auto bail = [&](gk_status s, const std::string& m) {
        g_ens_error = m;
        gk_ens_destroy(e);
        return s;
    };
    if (hipStreamCreate(&e->stream) != hipSuccess) return bail(GK_EHIP, "hipStreamCreate failed");
    if (hipMalloc(&e->arena, e->arena_elems * e->esz) != hipSuccess) return bail(GK_ENOMEM, "arena allocation failed");
    if (hipMalloc(&e->dprobs, sizeof(EnsProblem) * count) != hipSuccess ||
        hipMalloc(&e->clocks, sizeof(DevClock) * count) != hipSuccess ||
        hipMalloc(&e->counters, sizeof(int) * 4 * kEnsQueues) != hipSuccess ||
        hipHostMalloc((void**)&e->hclocks, sizeof(DevClock) * count) != hipSuccess)
        return bail(GK_ENOMEM, "auxiliary allocation failed");
    double floor_c = 1e-16;
    for (int p = 0; p < count; p++) {
        DevClock& h = e->hclocks[p];
        std::memset(&h, 0, sizeof h);
        h.dt = 1e10; h.dt_until_output = 1e10;
        for (int k = 0; k < kCmaxSlots; k++) {
            std::memcpy(&h.cmax_bits[0][k], &floor_c, 8);
            std::memcpy(&h.cmax_bits[1][k], &floor_c, 8);
        }
        h.stepmax = phys[p].stepmax; h.is_continue = 1;
        h.cfl = phys[p].cfl;
        const double dx = (grids[p].xmax - grids[p].xmin) / grids[p].nx;
        const double dy = (grids[p].ymax - grids[p].ymin) / grids[p].ny;
        const double dz = (grids[p].zmax - grids[p].zmin) / grids[p].nz;
        h.drmin = smin(smin(dx, dy), dz);
        gk_ctx& v = e->views[p];
        v.clock = e->clocks + p;
        for (int k = 0; k < 3; k++) v.buf[k] = (char*)e->arena + e->hprobs[p].off[k] * e->esz;
        v.stream = e->stream;
        v.uploaded = true;
    }
    if (hipMemcpy(e->clocks, e->hclocks, sizeof(DevClock) * count, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(e->counters, 0, sizeof(int) * 4 * kEnsQueues) != hipSuccess)
        return bail(GK_EHIP, "clock init failed");
    ens_plan(e);
    if (ens_upload_table(e) != GK_OK) return bail(GK_EHIP, "descriptor upload failed");
    *out = e;
    return GK_OK;
}

void gk_ens_destroy(gk_ens* e) {
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    ens_release_graph(e);
    for (gk_ctx& v : e->views) {  // views borrow the ensemble's memory and stream
        v.clock = nullptr; v.stream = nullptr;
        for (int k = 0; k < 3; k++) v.buf[k] = nullptr;
    }
    if (e->arena) (void)hipFree(e->arena);
    if (e->dprobs) (void)hipFree(e->dprobs);
    if (e->clocks) (void)hipFree(e->clocks);
    if (e->counters) (void)hipFree(e->counters);
    if (e->hclocks) (void)hipHostFree(e->hclocks);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

gk_status gk_ens_upload(gk_ens* e, int p, const double* host) {
    if (!e || !host || p < 0 || p >= e->P) return GK_EINVAL;
    gk_ctx* v = &e->views[p];
    GK_ENS_TRY(e, hipMemcpyAsync(e->hclocks + p, e->clocks + p, sizeof(DevClock), hipMemcpyDeviceToHost, e->stream));
    GK_ENS_TRY(e, hipStreamSynchronize(e->stream));
    DevClock& h = e->hclocks[p];
    h.cur = 0;
    double floor_c = 1e-16;
    for (int k = 0; k < kCmaxSlots; k++) std::memcpy(&h.cmax_bits[h.parity][k], &floor_c, 8);
    GK_ENS_TRY(e, hipMemcpyAsync(e->clocks + p, &h, sizeof(DevClock), hipMemcpyHostToDevice, e->stream));
    const size_t n = (size_t)v->plane * 5;
    if (e->prec == 32) {
        std::vector<float> tmp(n);
        for (size_t k = 0; k < n; k++) tmp[k] = (float)host[k];
        GK_ENS_TRY(e, hipMemcpyAsync(v->buf[0], tmp.data(), n * 4, hipMemcpyHostToDevice, e->stream));
        GK_ENS_TRY(e, hipStreamSynchronize(e->stream));
        hipLaunchKernelGGL(cmax_kernel<float>, dim3(256), dim3(kBlock), 0, e->stream, (const float*)v->buf[0],
                           v->plane, v->pitch, v->nx, v->ny, EosPolicy<float>::make(v->phys.gamma), v->clock);
    } else {
        GK_ENS_TRY(e, hipMemcpyAsync(v->buf[0], host, n * 8, hipMemcpyHostToDevice, e->stream));
        hipLaunchKernelGGL(cmax_kernel<double>, dim3(256), dim3(kBlock), 0, e->stream, (const double*)v->buf[0],
                           v->plane, v->pitch, v->nx, v->ny, EosPolicy<double>::make(v->phys.gamma), v->clock);
    }
    GK_ENS_TRY(e, hipGetLastError());
    GK_ENS_TRY(e, hipStreamSynchronize(e->stream));
    return GK_OK;
}

gk_status gk_ens_download(gk_ens* e, int p, double* host) {
    if (!e || !host || p < 0 || p >= e->P) return GK_EINVAL;
    gk_ctx* v = &e->views[p];
    GK_ENS_TRY(e, hipMemcpyAsync(e->hclocks + p, e->clocks + p, sizeof(DevClock), hipMemcpyDeviceToHost, e->stream));
    GK_ENS_TRY(e, hipStreamSynchronize(e->stream));
    const void* src = v->buf[e->hclocks[p].cur ? 1 : 0];
    const size_t n = (size_t)v->plane * 5;
    if (e->prec == 32) {
        std::vector<float> tmp(n);
        GK_ENS_TRY(e, hipMemcpy(tmp.data(), src, n * 4, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < n; k++) host[k] = tmp[k];
    } else {
        GK_ENS_TRY(e, hipMemcpy(host, src, n * 8, hipMemcpyDeviceToHost));
    }
    return GK_OK;
}

gk_status gk_ens_set_tile(gk_ens* e, int seg_rows, int waves_per_cu) {
    if (!e || seg_rows < 0 || waves_per_cu < 0 || waves_per_cu > 32) return GK_EINVAL;
    GK_ENS_TRY(e, hipStreamSynchronize(e->stream));
    e->seg_rows = seg_rows;
    if (waves_per_cu > 0) e->waves_per_cu = waves_per_cu;
    ens_plan(e);
    ens_release_graph(e);  // a captured graph holds the old item counts
    return ens_upload_table(e);
}

gk_status gk_ens_get_tile(gk_ens* e, int* seg_rows, int* waves_per_cu, int* total_items) {
    if (!e) return GK_EINVAL;
    if (seg_rows) *seg_rows = e->seg_rows > 0 ? e->seg_rows : 16;
    if (waves_per_cu) *waves_per_cu = e->waves_per_cu;
    if (total_items) *total_items = e->total_items;
    return GK_OK;
}

gk_status gk_ens_run_steps(gk_ens* e, int nsteps, int mode) {
    if (!e || nsteps < 0 || mode < 0 || mode > 2) return GK_EINVAL;
    const int nb = (e->P + 255) / 256;
    hipLaunchKernelGGL(ens_advance_begin_kernel, dim3(nb), dim3(256), 0, e->stream, e->clocks, e->P, 1e300);
    if (mode == 2) {  // replay a graph of graph_steps steps
        if (!e->gexec || e->graph_steps <= 0) { e->err = "no captured graph"; return GK_ESTATE; }
        if (nsteps % e->graph_steps) { e->err = "nsteps must be a multiple of the captured steps"; return GK_EINVAL; }
        for (int k = 0; k < nsteps / e->graph_steps; k++) GK_ENS_TRY(e, hipGraphLaunch(e->gexec, e->stream));
    } else {
        for (int k = 0; k < nsteps; k++) {
            if (mode == 0) ens_enqueue_step(e);
            else ens_enqueue_step_looped(e);
        }
    }
    GK_ENS_TRY(e, hipGetLastError());
    return GK_OK;
}

gk_status gk_ens_capture(gk_ens* e, int nsteps) {
    if (!e || nsteps <= 0) return GK_EINVAL;
    GK_ENS_TRY(e, hipStreamSynchronize(e->stream));
    ens_release_graph(e);
    GK_ENS_TRY(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < nsteps; k++) ens_enqueue_step(e);
    hipGraph_t g = nullptr;
    hipError_t ec = hipStreamEndCapture(e->stream, &g);
    if (ec != hipSuccess) { e->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(ec); return GK_EHIP; }
    e->graph = g;
    GK_ENS_TRY(e, hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0));
    e->graph_steps = nsteps;
    return GK_OK;
}

gk_status gk_ens_synchronize(gk_ens* e) {
    if (!e) return GK_EINVAL;
    GK_ENS_TRY(e, hipStreamSynchronize(e->stream));
    return GK_OK;
}

gk_status gk_ens_time(gk_ens* e, int nsteps, int mode, double* ms) {
    if (!e || !ms) return GK_EINVAL;
    hipEvent_t a, b;
    GK_ENS_TRY(e, hipEventCreate(&a));
    GK_ENS_TRY(e, hipEventCreate(&b));
    GK_ENS_TRY(e, hipStreamSynchronize(e->stream));
    GK_ENS_TRY(e, hipEventRecord(a, e->stream));
    gk_status s = gk_ens_run_steps(e, nsteps, mode);
    if (s != GK_OK) return s;
    GK_ENS_TRY(e, hipEventRecord(b, e->stream));
    GK_ENS_TRY(e, hipEventSynchronize(b));
    float f = 0;
    GK_ENS_TRY(e, hipEventElapsedTime(&f, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = f;
    return GK_OK;
}

gk_status gk_ens_get_time(gk_ens* e, int p, double* t, double* dt, int* step) {
    if (!e || p < 0 || p >= e->P) return GK_EINVAL;
    GK_ENS_TRY(e, hipMemcpyAsync(e->hclocks + p, e->clocks + p, sizeof(DevClock), hipMemcpyDeviceToHost, e->stream));
    GK_ENS_TRY(e, hipStreamSynchronize(e->stream));
    if (t) *t = e->hclocks[p].t;
    if (dt) *dt = e->hclocks[p].dt;
    if (step) *step = e->hclocks[p].step;
    return GK_OK;
}

const char* gk_ens_last_error(const gk_ens* e) { return e ? e->err.c_str() : g_ens_error.c_str(); }

}  // extern "C"
