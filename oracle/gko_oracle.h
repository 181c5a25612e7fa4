/*
 * gko_oracle.h -- CPU restatement of Gaukuk's per-timestep hydro update.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle for the HIP
 * path and the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product path (gaukuk_amd/)
 * never links or calls it.
 *
 * It restates, loop for loop and operation for operation, the reference's
 * 1D/2D path (nz == 1):
 *   Sim::Advance                 src/sim.cpp:69-117
 *   Sim::RK2_/RK3_/ForwardEuler_ src/evolution/rk2.cpp:11-22, rk3.cpp:11-28,
 *                                forward_euler.cpp:11-17
 *   Sim::UpdateCons (2/3 coef)   src/evolution/update_cons.cpp:11-75, 135-205
 *   Flux::CalFlux                src/flux/cal_flux.cpp:13-159 (nz == 1 part)
 *   Flux::RiemannSolver (HLLC)   src/flux/hllc.cpp:22-103
 *   Reconstruction PLM/1st order src/reconstruction/*.cpp, reconstruction.hpp:36-42
 *   EquationOfState              src/eos/adiabatic.cpp:12-198, eos.hpp:27-32
 *   Boundary types 0..3          src/boundary/{outflow_copy,periodic,reflective,outflow}.cpp
 *   Boundary fixed (type 4)      src/boundary/fixed.cpp:16-127, 242-269
 *   Boundary inner fixed region  src/boundary/fixed.cpp:181-207
 *   Immersed reflective shell    src/boundary/immersed.cpp:19-300
 *   SourceTerm::ConstGravity     src/source_term/const_gravity.cpp:10-67
 *   SourceTerm::PointGravity     src/source_term/point_gravity.cpp:11-93
 *   SourceTerm::BinaryGravity    src/source_term/binary_gravity.cpp:11-101
 *   SourceTerm::UpdateSource / EnrollPointGravity   src/source_term/source.cpp:12-67
 *
 * Built with strict IEEE flags (-O2 -fno-fast-math -ffp-contract=off) it is
 * bitwise equal to the reference built the same way (golden fixtures under
 * tests/golden/).  A second build with the reference's own fast flags serves
 * as the timed CPU baseline.
 *
 * State layout: Real[5][lenz][leny][lenx] with x fastest -- the reference TArray
 * layout (src/template_array.hpp:67-72); in 2D (nz == 1) only the single active
 * z plane is kept (lenz = 1; the reference keeps 2*nGhost empty z planes).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nx, ny, ng;            /* active cells and ghost width (ng >= 2 for PLM) */
    double xmin, xmax, ymin, ymax, zmin, zmax;
    double gamma, cfl;
    int integrator;            /* 1 forward Euler, 2 RK2, 3 RK3 (src/sim.cpp:60-66) */
    int rc_order;              /* 1 first order, 2 PLM */
    int stepmax;               /* <= 0: unlimited */
    int bc[6];                 /* xleft, xright, yleft, yright, zleft, zright: 0 copy, 1 periodic, 2 reflective,
                                  3 outflow, 4 fixed (z faces only when nz > 1) */
    double g[3];               /* constant gravity vector; zero vector = not enrolled */
    int npoint;                /* point masses enrolled (0, 1: point gravity, 2: binary) */
    double pg[2][8];           /* per mass: gm, x, y, z, vx, vy, vz, rs (EnrollPointGravity order) */
    int nz;                    /* active z planes (<= 1: 2D, a single plane) */
} gko_params;

typedef struct gko_sim gko_sim;

/* precision: 64 or 32 (32 has no reference counterpart: every operation in float). */
gko_sim* gko_create(const gko_params* p, int precision);
void     gko_destroy(gko_sim* s);
/* Copy the full ghosted state in/out (Real[5][lenz][leny][lenx], Real per precision). */
void     gko_set_cons(gko_sim* s, const void* cons);
void     gko_get_cons(const gko_sim* s, void* cons);
/* Sim::Advance(dtoutput).  Returns the number of steps taken. */
int      gko_advance(gko_sim* s, double dtoutput);
double   gko_time(const gko_sim* s);
int      gko_step(const gko_sim* s);
int      gko_is_continue(const gko_sim* s);
double   gko_last_dt(const gko_sim* s);
/* Boundary::EnrollFixedBoundaryData (src/boundary/fixed.cpp:242-269): save the ghost layers of
 * the current (set) state for type-4 faces, and with fix_inner the state of the inner region
 * (fixed.cpp:181-207). */
void     gko_enroll_fixed(gko_sim* s, int fix_inner);
/* Boundary::SetupImmersedShell (src/boundary/immersed.cpp:89-235): a reflective solid sphere of
 * radius r at (x, y, z), applied at every boundary update.  Returns -1 where the reference throws. */
int      gko_setup_immersed(gko_sim* s, double x, double y, double z, double r);
/* Boundary::UpdateBD on the current state (main calls it once after Setup, src/main.cpp:30).
 * Returns -1 where the reference throws. */
int      gko_update_bd(gko_sim* s);
/* Current step log: up to cap entries of (t, dt) for the last gko_advance. */
int      gko_dt_log(const gko_sim* s, double* t_out, double* dt_out, int cap);

/* ---- per-function entry points (FP64) for known-answer vectors ---- */
/* Flux::RiemannSolver over n faces; ul/ur/flux are [5][n]; dir = 1 (x), 2 (y), 3 (z). */
void   gko_hllc(int n, const double* ul, const double* ur, int dir, double gamma, double* flux);
/* ReconstructXPLM over one row of m cells: prim [5][m] -> ul, ur [5][m+1]
   (faces 2..m-1 written for ul, 1..m-2 for ur; other entries left untouched). */
void   gko_plm_row(int m, const double* prim, double* ul, double* ur);
/* ConsToPrim over n cells ([5][n] each). */
void   gko_cons_to_prim(int n, const double* cons, double gamma, double* prim);
/* PrimToCons over n cells. */
void   gko_prim_to_cons(int n, const double* prim, double gamma, double* cons);
/* CalCmax over n cells, starting from cmax0. */
double gko_cmax(int n, const double* cons, double gamma, double cmax0);

#ifdef __cplusplus
}
#endif
