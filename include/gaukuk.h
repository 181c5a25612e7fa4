/*
 * gaukuk.h -- C ABI of the MI355X hydro library (libgaukuk_hip.so).
 *
 * The reference has no FFI: its hot path is the C++ class Sim (src/sim.hpp:23-58) whose
 * Advance() (src/sim.cpp:69-117) runs the integrator selected by a pointer-to-member
 * (src/sim.hpp:47,51; src/sim.cpp:60-66) over TArray state (src/template_array.hpp).  This ABI
 * is the boundary that replaces that class's compute members; the C++ host driver in
 * gaukuk_amd/host/ keeps the reference's Sim surface (Setup, Advance, WriteData, the .in keys)
 * and calls these entry points.  Plain C types only: pointers, sizes, doubles, ints.
 *
 * Conventions
 *   ownership  the caller owns every host buffer; a context owns its device memory.
 *   errors     every call returns gk_status; gk_last_error() describes the last failure.  No
 *              exception crosses the ABI (the C++ wrapper rethrows, as the reference throws
 *              std::runtime_error from its constructors, src/grid.hpp:23-25).
 *   threading  one context per host thread; work is ordered on the context's HIP stream.
 *   numerics   floors repair states silently, as in the reference (update_cons.cpp:55-67);
 *              a floor hit is not an error.
 */
#ifndef GAUKUK_H
#define GAUKUK_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gk_ctx gk_ctx;

typedef enum {
    GK_OK = 0,
    GK_EINVAL = 1,   /* bad argument or unsupported configuration */
    GK_EHIP = 2,     /* a HIP runtime call failed (text in gk_last_error) */
    GK_ENOMEM = 3,   /* device allocation failed */
    GK_ESTATE = 4,   /* call out of order (e.g. advance before any state was uploaded) */
    GK_ENUMERIC = 5  /* non-finite state detected by gk_check_finite */
} gk_status;

/* Grid: the keys NX, NY, NZ, NGhost, xmin..zmax of the .in file (src/grid.hpp:18-41,
 * src/sim.cpp:16-42).  nz = 1 runs the 1D/2D path, nz > 1 the 3D path (ny >= 2; every face type,
 * constant and point gravity, the inner fixed region, the immersed sphere and boundary hooks);
 * nghost must be at least 2 (PLM stencil). */
typedef struct {
    int nx, ny, nz, nghost;
    double xmin, xmax, ymin, ymax, zmin, zmax;
} gk_grid;

/* Physics and integrator: gamma (src/eos/adiabatic.cpp:12-16), CFL, integrator 1/2/3 and
 * RcOrder 1/2 (src/sim.cpp:47,56-66), stepmax, boundary codes xleft..zright: 0 copy, 1 periodic,
 * 2 reflective, 3 outflow, 4 fixed (its data comes from gk_set_fixed_boundary)
 * (src/boundary/boundary.cpp:13-125), constant gravity (src/source_term/source.cpp:23-31),
 * precision 64 or 32 (the reference's Real, src/gaukuk.hpp:5-9). */
typedef struct {
    double gamma, cfl;
    double g[3];
    int integrator;
    int rc_order;
    int stepmax;
    int bc[6];
    int precision;
} gk_phys;

/* SourceTerm::EnrollPointGravity (src/source_term/source.cpp:33-67): enroll a point mass (gm, position,
 * velocity, smoothing length rs).  The first call enables point gravity
 * (src/source_term/point_gravity.cpp), a second call turns it into binary gravity (the second mass
 * orbits the first, src/source_term/binary_gravity.cpp), a third call is an error.  Applied at both
 * Strang half-kicks after constant gravity (src/source_term/source.cpp:12-21). */
gk_status gk_set_point_gravity(gk_ctx* ctx, double gm, double x, double y, double z, double vx, double vy, double vz,
                               double rs);

/* Boundary::EnrollFixedBoundaryData(cons, grid, fixInner) (src/boundary/fixed.cpp:242-269; faces
 * FixedBDXL..FixedBDZR, fixed.cpp:16-179): `cons` is a state in the reference layout and the
 * context's precision; its ghost layers become the values that type-4 faces restore before every
 * stage (x and y faces on every active plane, z faces over the active rows).  With fix_inner != 0
 * its active cells inside the inner region (the central half-box in x and y on every active plane,
 * cell centres within 0.5 of the origin: Boundary::FixedBDInner, fixed.cpp:181-207) are restored
 * there before every stage as well.  Advancing a context with a type-4 face fails with GK_ESTATE
 * until this is called, as the reference throws. */
gk_status gk_set_fixed_boundary(gk_ctx* ctx, const void* cons, int fix_inner);

/* Boundary::SetupImmersedShell (src/boundary/immersed.cpp:89-235): a reflective solid sphere of
 * radius r at (x, y, z).  Before every stage the shell cells (inside the body, within nGhost cells
 * of its surface) take the weighted image of the fluid around their mirror points
 * (Boundary::ImmersedReflective, immersed.cpp:237-300).  A later call replaces the shell.  Fails
 * with GK_EINVAL where the reference throws (body outside the domain, too close to a face, more
 * than 8 image cells); advancing with gravity enrolled fails with GK_EINVAL. */
gk_status gk_set_immersed_shell(gk_ctx* ctx, double x, double y, double z, double r);

/* Self-defined faces whose ghost values depend on the state (Boundary::EnrollSelfDefineBD*,
 * src/boundary/boundary.hpp:28-45; the reference calls them in every UpdateBD, i.e. before every
 * stage: src/evolution/rk2.cpp:12,17, boundary.cpp:127-142).  The faces in `face_mask` (bit f for
 * face f: 0 x left, 1 x right, 2 y left, 3 y right, and on 3D grids 4 z left, 5 z right) become
 * fixed (type-4) faces whose ghost layers the stage reads (the two nearest an x or y face, the
 * nearest plane of a z face, over the active tangential range) are refreshed before every stage:
 * the library writes the `band` active layers nearest each of those faces of the stage's input
 * (after the first half-kick of constant, point or binary gravity in the first stage, as the
 * reference's face functions see it) into `mirror` at their reference-layout positions, calls
 * hook(user, mirror), and reads the ghost layers back from `mirror`.  `mirror` is the caller's host
 * array Real[5][lenz][leny][lenx] in the context's precision; the library writes only those bands,
 * so cells the caller poisoned (e.g. with NaN) stay poisoned, and a face function that reads beyond
 * the band yields non-finite ghosts, which fail the step with GK_EINVAL.  Ghost cells outside the
 * active tangential range (edges and corners) are not taken from the mirror: they keep what the
 * reference's own faces leave there.  In 3D a y face on the side of a type-0 z face stays the copy
 * that z face registers (the reference runs it after the y faces), hooked or not.  Stepping
 * returns to the host once per stage; gk_download_cons fills the hooked faces' ghost cells by
 * calling the hook on the current state.  Fails with GK_EINVAL for a NULL hook or mirror, a mask
 * naming z faces of a 2D grid, band outside [1, min(nx, ny, nz)] (2D: [1, min(nx, ny)]), y faces
 * of a one-row grid or an attached network.  A later call replaces the hook; its mask must include
 * every face an earlier call hooked (GK_EINVAL otherwise). */
typedef void (*gk_bd_hook)(void* user, void* mirror);
gk_status gk_set_bd_hook(gk_ctx* ctx, int face_mask, int band, gk_bd_hook hook, void* user, void* mirror);

/* The in-place part of Boundary::UpdateBD on the current state (src/boundary/boundary.cpp:127-142):
 * the immersed shell, then the inner fixed region.  The reference's main calls UpdateBD once after
 * Setup (src/main.cpp:30); callers that upload a state and advance it call this in between. */
gk_status gk_update_bd(gk_ctx* ctx);

/* Sim::Sim() (src/sim.cpp:45-67): validate, allocate device state, select the integrator. */
gk_status gk_create(const gk_grid* grid, const gk_phys* phys, gk_ctx** out);
/* Sim::~Sim(): free everything owned by the context. */
void gk_destroy(gk_ctx* ctx);

/* State transfer in the reference TArray layout Real[5][lenz][leny][lenx] with ghosts
 * (src/template_array.hpp:67-72; Real per gk_phys.precision).  Upload reads active cells only;
 * download writes active cells, fills ghost cells by the boundary rules and zeroes corners and
 * z ghost planes.  Replaces the host arrays Sim::cons is read from / written to. */
gk_status gk_upload_cons(gk_ctx* ctx, const void* host);
gk_status gk_download_cons(gk_ctx* ctx, void* host);
/* Active cells only, always float64 [5][ny][nx] (converted for precision 32). */
gk_status gk_upload_active(gk_ctx* ctx, const double* host);
gk_status gk_download_active(gk_ctx* ctx, double* host);

/* Sim::Advance(dtoutput) (src/sim.cpp:69-117): step until t reaches t + dtoutput or stepmax.
 * On return *t and *step hold the new time and step count (either may be NULL).  Steps are
 * enqueued in chunks sized from the steps left at the current dt (at most 256 per host check). */
gk_status gk_advance(gk_ctx* ctx, double dtoutput, double* t, int* step);
/* Sim::GetTime / Getdt / GetStep and isContinue (src/sim.hpp:41-43, :26). */
gk_status gk_get_time(gk_ctx* ctx, double* t, double* dt, int* step, int* is_continue);
/* Set the simulation time and step count (a restart from an output file; the reference has no
 * restart, its Sim starts at t = 0, src/sim.cpp:45-47).  The next step's dt follows from the current
 * state as usual. */
gk_status gk_set_time(gk_ctx* ctx, double t, int step);
/* The per-step lines the reference prints (src/sim.cpp:100-103): the last n (<= cap) steps'
 * (step, t, dt), oldest first.  *count receives the number written. */
gk_status gk_step_log(gk_ctx* ctx, int* steps, double* t, double* dt, int cap, int* count);

/* Enqueue exactly nsteps steps with no output deadline (for benchmarking); asynchronous. */
gk_status gk_run_steps(gk_ctx* ctx, int nsteps);
gk_status gk_synchronize(gk_ctx* ctx);
/* Time nsteps steps with HIP events on the context's stream (one event after every launch; all
 * events are created before the first launch).
 * total_ms: the whole region; kernel_ms[k] / launches[k]: summed duration and count of the
 * stage kernel at position k = 0, 1, 2 within the step, and k = 3 the one-thread clock kernels. */
gk_status gk_profile_steps(gk_ctx* ctx, int nsteps, double* total_ms, double kernel_ms[4], int launches[4]);

/* FP64 sums of the five conserved variables over active cells, deterministic. */
gk_status gk_cons_sums(gk_ctx* ctx, double out[5]);
/* gk_cons_sums without allocating (no counterpart in the reference: for callers that cannot
 * allocate device memory, such as code under HIP graph capture or a tight serving loop).  The same
 * kernels and summation order as gk_cons_sums, so the same bits, on a state the caller holds in
 * device memory: `state` is [5][nz][ny][nx] (x fastest) of double (precision 64) or float (32);
 * `out` receives 5 doubles in device memory; `stream` is a hipStream_t (NULL: the null stream);
 * the call only enqueues work.  The partial sums live in a 20 KiB array of the library's device
 * memory, one per GPU, held until the process ends (its size does not depend on the state's);
 * every call overwrites all of it, and calls that overlap on one GPU race on it.  Calls no
 * hipMalloc or hipFree.  Errors: GK_EINVAL (null pointer, non-positive size, precision not 64/32,
 * more than 2^40 cells), GK_EHIP (no current device, symbol lookup or launch failed); the text is
 * in gk_last_error(NULL). */
gk_status gk_cons_sums_noalloc(const void* state, int nx, int ny, int nz, int precision, double* out, void* stream);
/* GK_ENUMERIC if any active conserved value is NaN or infinite. */
gk_status gk_check_finite(gk_ctx* ctx);

/* Tuning: rows per wave segment of every stage kernel (0 = automatic: each stage kernel gets the
 * segment height that fills the device once at that kernel's occupancy).  Results do not depend
 * on the choice.  gk_get_tile reports the fixed height, or in automatic mode the plan for three
 * waves per SIMD (the FP64 2D kernels' occupancy; two for 3D). */
gk_status gk_set_tile(gk_ctx* ctx, int seg_rows);
gk_status gk_get_tile(gk_ctx* ctx, int* seg_rows, int* strips, int* waves);

/* Device self-check, no counterpart in the reference: the stage kernels' quotient, reciprocal,
 * square-root, sqrt(a / b) and reciprocal-square-root cores (FP64 and FP32) against the IEEE
 * operators on `count` operand draws with binary exponents in [-emax, emax] (root operands
 * [-2*emax, 2*emax]; FP32 within +-60); *mismatches counts results farther from the IEEE result
 * than the core's bound (FP64 16-64 ulp, FP32 4 ulp: about twice the largest error measured), and
 * the stage kernels' neighbour-lane moves that differ from the cross-lane shuffles on a lane that
 * has a source lane (one draw per launched thread).  Expected 0 for emax <= 500. */
gk_status gk_selftest(long long count, unsigned long long seed, int emax, long long* mismatches);

/* ---- Ensembles: many independent grids advanced together (no counterpart in the reference,
 * which runs one problem per process; each problem follows Sim::Advance semantics with its own
 * dt).  Problems share precision, integrator, RcOrder, gamma and gravity; sizes, extents, CFL,
 * stepmax and boundary codes are per problem.  A step runs one clock kernel for all problems, one
 * persistent launch per RK stage over every problem's (strip, segment) work items, and one clock
 * kernel for the t update. */
typedef struct gk_ens gk_ens;
gk_status gk_ens_create(int count, const gk_grid* grids, const gk_phys* phys, gk_ens** out);
void gk_ens_destroy(gk_ens* ens);
/* Active cells of problem p, float64 [5][ny][nx] (converted for precision 32). */
gk_status gk_ens_upload(gk_ens* ens, int p, const double* host);
gk_status gk_ens_download(gk_ens* ens, int p, double* host);
/* Tile: rows per segment (0 = default 16) and resident waves per CU of the persistent grid. */
gk_status gk_ens_set_tile(gk_ens* ens, int seg_rows, int waves_per_cu);
gk_status gk_ens_get_tile(gk_ens* ens, int* seg_rows, int* waves_per_cu, int* total_items);
/* Enqueue nsteps steps without an output deadline.  mode 0: batched persistent launches;
 * mode 1: one single-grid launch per problem per stage (the looped baseline); mode 2: replay the
 * captured graph (nsteps a multiple of the captured step count). */
gk_status gk_ens_run_steps(gk_ens* ens, int nsteps, int mode);
/* Capture nsteps batched steps into a HIP graph for mode 2. */
gk_status gk_ens_capture(gk_ens* ens, int nsteps);
gk_status gk_ens_synchronize(gk_ens* ens);
/* Run nsteps in `mode` between two HIP events on the ensemble's stream; *ms = elapsed time. */
gk_status gk_ens_time(gk_ens* ens, int nsteps, int mode, double* ms);
gk_status gk_ens_get_time(gk_ens* ens, int p, double* t, double* dt, int* step);
const char* gk_ens_last_error(const gk_ens* ens);

/* ---- Learned flux correction (no counterpart in the reference, whose coarse flux is plain HLLC,
 * src/flux/hllc.cpp:22-103).  A network (gk_net) is an MLP 16 -> 64 -> 64 -> 4 with ReLU, run on
 * BF16 matrix cores with FP32 accumulation, plus its Adam state and a device-resident data set.
 * Parameters: 5508 floats, W1[64][16], b1[64], W2[64][64], b2[64], W3[4][64], b3[4] (row-major,
 * W[out][in]).  Network input = feature * in_scale; flux correction = output * out_scale, in the
 * units described in gk_learned.hip (tangential face frame over rho c^k of the face means). */
#define GK_NET_PARAMS 5508
typedef struct gk_net gk_net;
gk_status gk_net_create(gk_net** out);
void gk_net_destroy(gk_net* net);
/* Setting parameters also clears the Adam moments and the minibatch draw counter. */
gk_status gk_net_set_params(gk_net* net, const float* params);
gk_status gk_net_get_params(gk_net* net, float* params);
gk_status gk_net_set_scales(gk_net* net, const float in_scale[16], const float out_scale[4]);
gk_status gk_net_get_scales(gk_net* net, float in_scale[16], float out_scale[4]);
/* Dissipation bound (default 0.5): each correction component, and each collected training target,
 * is clipped to +-kappa * 1/2 S |U_R - U_L| of the PLM face states (kappa = 0: no correction). */
gk_status gk_net_set_kappa(gk_net* net, float kappa);
gk_status gk_net_get_kappa(gk_net* net, float* kappa);
/* y[k][0..3] = network output (before out_scale) for x[k][0..15] (raw features); host buffers. */
gk_status gk_net_forward(gk_net* net, long long count, const float* x, float* y);
/* Gradient of L = sum_k wt_k sum_o (y_ko - t_ko / out_scale_o)^2 / (4 count); host buffers. */
gk_status gk_net_gradient(gk_net* net, long long count, const float* x, const float* t, const float* wt,
                          float* grad, double* loss);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8, bias-corrected) with a host gradient. */
gk_status gk_net_adam(gk_net* net, const float* grad, float lr);
/* Data set on the device: reserve room, append host samples, read back a range (x, t, wt may be
 * NULL) and its size, weighted mean squares of inputs and targets. */
gk_status gk_net_reserve(gk_net* net, long long capacity);
gk_status gk_net_add_samples(gk_net* net, long long count, const float* x, const float* t, const float* wt);
gk_status gk_net_dataset(gk_net* net, long long start, long long count, float* x, float* t, float* wt,
                         long long* size);
gk_status gk_net_dataset_stats(gk_net* net, double ms_x[16], double ms_t[4], double* weight);
/* iters Adam steps on minibatches of blocks * 64 * tiles_per_wave samples drawn (with replacement,
 * from `seed`) from the data set; loss_log (may be NULL) receives each step's loss. */
gk_status gk_net_train(gk_net* net, int iters, int blocks, int tiles_per_wave, float lr, unsigned long long seed,
                       float* loss_log);
const char* gk_net_last_error(const gk_net* net);

/* Training data from a fine FP64 run (both grids FP64, same integrator and gamma, fine = f x coarse
 * in both axes): restrict the fine state onto the coarse grid, take the coarse grid's features at
 * every face, run `steps` recorded coarse steps (span D) and the fine run over exactly D, recorded;
 * append one sample per coarse face (target: restricted fine flux minus coarse flux, averaged over
 * D; weight 0 where the shock sensor fires).  The fine run keeps going from t + D. */
gk_status gk_net_collect(gk_net* net, gk_ctx* fine, gk_ctx* coarse, int steps, double* span, long long* added);
/* Flux recording: while on, every stage adds (stage weight) * dt * flux at every face: mass, m_x,
 * m_y, E planes of (nx+1)*ny x faces then nx*(ny+1) y faces (face i between cells i-1 and i).
 * Turning it on zeroes the sums. */
gk_status gk_flux_record(gk_ctx* ctx, int on);
gk_status gk_flux_download(gk_ctx* ctx, double* acc);
/* Conservative average of the fine grid's state onto the coarse grid (the coarse grid takes the
 * fine grid's time and restarts its step count). */
gk_status gk_restrict(gk_ctx* fine, gk_ctx* coarse);
/* Attach a network to a precision-32 PLM grid without gravity (NULL detaches): its stages then
 * run the corrected kernel.  The network must outlive the attachment.  While one is attached,
 * gk_set_point_gravity, gk_set_fixed_boundary and gk_set_immersed_shell fail with GK_EINVAL, and
 * stepping re-checks the configuration. */
gk_status gk_set_net(gk_ctx* ctx, gk_net* net);
/* Face counts since the last reset: corrected, shock-sensor fallbacks, positivity fallbacks. */
gk_status gk_net_stats(gk_ctx* ctx, long long out[3], int reset);

/* Text of the last error on this context (or of gk_create when ctx is NULL). */
const char* gk_last_error(const gk_ctx* ctx);
/* Library build identification (gfx target, policies). */
const char* gk_version(void);

#ifdef __cplusplus
}
#endif
#endif /* GAUKUK_H */
