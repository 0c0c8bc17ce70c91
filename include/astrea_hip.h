/*
 * astrea_hip.h -- C ABI of libastrea_hip.so, the MI355X (gfx950) finite-volume
 * update for astrea (mervyzr/astrea 1.3.2).
 *
 * The reference has no FFI: its operator interface for this path is the pair
 *   evolve_space(grid, sim_variables) -> fluxes   (num_methods/evolvers.py:12)
 *   evolve_time(grid, fluxes, dt, sim_variables) -> grid (num_methods/evolvers.py:38)
 * driven by core_run's time loop (astrea.py:31-85).  The entry points below
 * replace those calls; the comment on each names the reference code it stands
 * in for.  Only plain C types cross the ABI; device pointers are HIP device
 * addresses (e.g. torch CUDA-tensor data pointers), ``stream`` is a hipStream_t
 * passed as void*.
 *
 * State layout ("packed SoA"): for a plan over P problems with C_p cells each
 * (C_p = n0*n1, cell (i0, i1) at i0*n1 + i1, i.e. the reference's C order of
 * grid[i0, i1]), the state is nvars planes of sum(C_p) doubles; plane v holds
 * conservative component varidx[v] of every problem back to back.  The
 * components of a plan are the live ones of its variable mask (bit c set =
 * component c of rho, rho*vx, rho*vy, rho*vz, E, Bx, By, Bz is carried);
 * components outside the mask are identically zero in the reference for the
 * problem class the mask names and are not stored.
 *
 * All functions return AH_OK (0) or a negative status; none throws.
 * A plan is single-stream and not re-entrant; distinct plans are independent.
 */
#ifndef ASTREA_HIP_H
#define ASTREA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AH_VERSION_MAJOR 1
#define AH_VERSION_MINOR 0

enum ah_status {
    AH_OK = 0,
    AH_ERR_INVALID = -1,     /* bad argument or option combination */
    AH_ERR_UNSUPPORTED = -2, /* option the device path does not implement */
    AH_ERR_HIP = -3,         /* a HIP runtime call failed */
    AH_ERR_NOMEM = -4,       /* workspace too small / allocation failed */
    AH_ERR_STATE = -5,       /* plan not bound, or used out of order */
    AH_ERR_TOO_LARGE = -6    /* problem larger than the static scratch array (ah_evolve_space_static) */
};

/* options -- same meaning as astrea's option strings (static/.db.json:220-347) */
enum ah_boundary { AH_EDGE = 0, AH_WRAP = 1 };                /* "edge" (outlet) / "wrap" (periodic) */
enum ah_recon { AH_PCM = 0, AH_PLM = 1, AH_PPM = 2, AH_WENO3 = 3, AH_WENO5 = 4, AH_WENO7 = 5 };
enum ah_limiter { AH_MINMOD = 0, AH_VANLEER = 1, AH_OSPRE = 2, AH_VANALBADA1 = 3, AH_KOREN = 4, AH_SUPERBEE = 5 };
enum ah_flattener { AH_FLAT_OFF = 0, AH_FLAT_ASTREA = 1, AH_FLAT_COLELLA90 = 2 };
/* PPM authors (schemes/ppm.py:43-71): McCorquodale & Colella 2011 ('mc', the
   reference's only reachable one), Colella et al. 2011 ('c': interface limiter
   limiters.py:53-79, then the extrapolant limiter limiters.py:144-201) and
   Peterson & Hammett 2008 ('ph') */
enum ah_ppm_author { AH_PPM_MC = 0, AH_PPM_C = 1, AH_PPM_PH = 2 };
/* solvers.py:69-75 (LLF), 92-138 (HLLC), 142-232 (HLLD), 236-284 (Osher-Solomon
   with the Dumbser-Toro path integral; FP64, 1D hydro/MHD and 2D hydro),
   288-391 (entropy-stable as the reference writes it; FP64, 1D, all 8 variables),
   79-88 (Lax-Wendroff, FP64, 1D/2D hydrodynamics and MHD: the coefficient ss^2/max|lambda|
   with ss the sound speed sqrt(gamma p/rho); the reference takes ss from a
   LAPACK-order-dependent eigenvalue column instead, so its results are not
   reproduced) */
enum ah_solver { AH_LLF = 0, AH_HLLC = 1, AH_HLLD = 2, AH_DOTS = 3, AH_ES = 4, AH_LW = 5 };
enum ah_rk {
    AH_EULER = 0, AH_RK4 = 1, AH_SSPRK22 = 2, AH_SSPRK33 = 3, AH_SSPRK43 = 4,
    AH_SSPRK53 = 5, AH_SSPRK54 = 6, AH_SSPRK104 = 7
};
/* AH_FP64: every operation in FP64, in the reference's order (bitwise the numpy
 * reference).  AH_SPLIT: the HLLC solver's coefficient algebra (wave-speed estimates,
 * star-region factors), the spectral radius and the reciprocal of the primitive
 * recovery in FP32; states, interface states, reconstruction, fluxes, HLLD,
 * constrained transport, stage combinations and the dt reduction in FP64.  Not
 * bitwise. */
enum ah_precision { AH_FP64 = 0, AH_SPLIT = 1 };
/* AH_COMPAT_ASTREA reproduces the reference's Riemann-solver axis argument:
 * the enumeration index of the sweep, which swaps the normal velocity of
 * HLLC/HLLD between the axes on every other step (solvers.py:34-36,63,
 * astrea.py:85), with the HLLC star momentum always at index 1
 * (solvers.py:127-128) and F(w+) as the flux when no HLLC region mask applies,
 * i.e. when all waves move right (solvers.py:133).  AH_COMPAT_PHYSICAL passes
 * the physical sweep axis, writes the star momentum at that axis and returns
 * the upwind F(w-) there. */
enum ah_compat { AH_COMPAT_ASTREA = 0, AH_COMPAT_PHYSICAL = 1 };
enum ah_varmask {
    AH_VARS_H1 = 0x13,  /* rho, rho*vx, E          : 1D hydrodynamics */
    AH_VARS_H2 = 0x17,  /* rho, rho*vx, rho*vy, E  : 2D hydrodynamics */
    AH_VARS_M6 = 0x77,  /* + Bx, By                : 2D MHD with vz = Bz = 0 */
    AH_VARS_ALL = 0xFF
};

typedef struct ah_scheme {
    int dims;        /* 1 or 2 (sim_variables.dimension) */
    int recon;       /* ah_recon   (subgrid, evolvers.py:14-21) */
    int limiter;     /* ah_limiter (PLM slope limiter; reference: minmod, plm.py:66) */
    int flattener;   /* ah_flattener (PPM only) */
    int solver;      /* ah_solver  (solvers.py:13-31) */
    int rk;          /* ah_rk      (evolvers.py:79-206) */
    int precision;   /* ah_precision */
    int varmask;     /* ah_varmask */
    int magnetic_2d; /* constrained transport (evolvers.py:26-32, 51-58, 63-67, 73-76) */
    int tile_s;      /* 1D sweep tile length in cells: 0 (the default, 243 for hydrodynamics) or
                        115 (128-thread workgroups; 1D hydrodynamic PLM/PPM with LLF/HLLC) */
    int tile_t;      /* reserved (0) */
    int compat;      /* ah_compat */
    int av;          /* 1: artificial viscosity on the fluxes (1D PPM only; schemes/ppm.py:138-170,
                        whose output the reference computes but never applies: 0 is the
                        reference-equivalent setting) */
    int ppm_author;  /* ah_ppm_author (ppm.py:13; the reference's evolve_space always passes 'mc',
                        evolvers.py:17); c / ph: 1D and 2D hydrodynamics in FP64 without
                        flattener or artificial viscosity, on the default tiles */
    int fuse2d;      /* 2D hydrodynamics (PLM/PPM with LLF/HLLC): 1 = one launch per sweep, the
                        interface states reconstructed inside each flux tile instead of written to
                        and read back from HBM by a reconstruction launch; the same results bit for
                        bit, and no interface-state planes in the workspace.  0: two launches */
    int reserved[1];
} ah_scheme;

typedef struct ah_problem {
    int n0, n1;       /* cells on axis 0 and axis 1 (n1 = 1 in 1D) */
    int boundary;     /* ah_boundary */
    int checkpoints;  /* sim_variables.checkpoints (astrea.py:32) */
    double dx;        /* sim_variables.dx (used for both axes, evolvers.py:48) */
    double gamma;
    double cfl;
    double t_end;
    long long max_steps; /* < 0: no step limit (the reference runs to t_end only) */
} ah_problem;

typedef struct ah_problem_status {
    double t;          /* simulated time (astrea.py:84) */
    double dt;         /* last dt taken */
    long long steps;   /* steps taken */
    int active;        /* 1 while t < t_end and steps < max_steps */
    int failed;        /* 1 once a non-finite Jacobian state appeared (the reference's LAPACK call raises there) */
} ah_problem_status;

/* Full per-problem time state (what a problem carries between steps). */
typedef struct ah_time_state {
    double t, dt;
    long long steps;
    int idx;      /* next checkpoint index (astrea.py:44,78) */
    int active;
    int failed;
    int parity;   /* sweep-order parity (astrea.py:85) */
} ah_time_state;

typedef struct ah_plan ah_plan;

/* Library identity. */
int ah_version(int* major, int* minor);
const char* ah_status_str(int status);

/* Build a plan for `nprob` independent problems sharing one scheme (one problem
 * = the reference's single run; several = the ragged batch).  Allocates the
 * plan's small device tables (problem descriptors, tile lists, time state). */
int ah_plan_create(const ah_scheme* scheme, int nprob, const ah_problem* problems, ah_plan** out);
int ah_plan_destroy(ah_plan* plan);

/* Packed state layout: total cells, number of stored components and their
 * component indices (varidx[v] in 0..7), first cell of each problem. */
int ah_plan_layout(const ah_plan* plan, long long* total_cells, int* nvars, int* varidx,
                   long long* problem_offsets);

/* Caller-provided device workspace (stage registers, interface states,
 * reductions).  No allocation happens in ah_plan_step.  Binding again (new
 * state or workspace) drops a step graph captured with the old buffers. */
int ah_plan_workspace_size(const ah_plan* plan, size_t* bytes);
int ah_plan_bind(ah_plan* plan, double* state, void* workspace, size_t workspace_bytes);

/* t = 0, step counters 0, every problem active (astrea.py:44). */
int ah_plan_reset(ah_plan* plan, void* stream);

/* Advance every active problem by up to `nsteps` full time steps of the
 * reference loop (astrea.py:45-85): stage-0 evolve_space, dt = cfl*min(dx/eigmax)
 * with the checkpoint clamp, evolve_time, t += dt, sweep-order flip.  Time,
 * dt and termination stay on the device; a problem that reaches t_end or
 * max_steps (or fails) is skipped by the remaining launches. */
int ah_plan_step(ah_plan* plan, int nsteps, void* stream);

/* The same steps captured once into a hipGraph of `nsteps` steps and
 * launched `nlaunch` times (reference loop astrea.py:45-85, replayed without
 * host work).  Needs a non-default stream; the graph stays valid across
 * ah_plan_compact.  No reference counterpart beyond the loop itself. */
int ah_plan_step_graph(ah_plan* plan, int nsteps, int nlaunch, void* stream);

/* Drop the tiles of problems that are no longer active from every kernel's
 * work list (synchronises the stream); `nactive` receives the number of
 * active problems.  Batched runs call it between chunks of steps. */
int ah_plan_compact(ah_plan* plan, int* nactive, void* stream);

/* Persistent grid cap: workgroups per launch (each loops over its tiles). */
int ah_plan_set_blocks(ah_plan* plan, int max_blocks);

/* Initial state of problem `prob` built on the device straight into the
 * bound state planes, replacing constructor.initialise
 * (functions/constructor.py:11-109) for piecewise-constant problems:
 * primitive states at the cell centres of [0,1]^d (constructor.py:13-16),
 * then the conservative cell averages of constructor.py:105-109 (4th-order
 * conversion when `high_order`, i.e. for PPM and WENO, then the h^2/24
 * Laplacian), equal to that pipeline to the last bit.
 *   AH_IC_RIEMANN_1D:   w[0] for x <= x0, w[1] otherwise;
 *   AH_IC_QUADRANTS_2D: w[0] for x <= x0 and y < y0, w[1] for x > x0 and
 *                       y < y0, w[2] for x <= x0 and y >= y0, w[3] otherwise
 *                       (the Lax-Liu layout of constructor.py:57-60).
 * w[k] are primitive 8-vectors (rho, vx, vy, vz, p, Bx, By, Bz); components
 * the plan does not store must be zero.  The serving loop admits requests
 * without shipping their grids. */
enum ah_ic_kind { AH_IC_RIEMANN_1D = 0, AH_IC_QUADRANTS_2D = 1 };
typedef struct ah_ic {
    int kind;        /* ah_ic_kind */
    int high_order;  /* 1: 4th-order averaging (PPM, WENO); 0: pointwise conversion */
    double x0, y0;   /* region split positions */
    double w[4][8];  /* primitive state per region */
} ah_ic;
int ah_plan_init_ic(ah_plan* plan, int prob, const ah_ic* ic, void* stream);

/* The primitive snapshot of problem `prob` as the reference saves it every
 * step (astrea.py:47): sim_variables.convert_conservative of the state (the
 * 4th-order cell conversion for PPM/WENO, the pointwise one otherwise,
 * generic.py:250-255) transposed by ortho_axis (generic.py:264), all 8
 * components, FP64: `out` (device memory, cells * 8 doubles) receives the
 * array of shape (n0, 8) in 1D and (n1, n0, 8) in 2D, equal to the
 * reference's to the last bit.  The HDF5 writing around it (astrea.py:48-50)
 * stays with the caller. */
int ah_plan_snapshot(ah_plan* plan, int prob, double* out, void* stream);

/* Snapshots taken by the step program itself (astrea.py:46-50 writes one per
 * pass of its loop; the host need not wait on the device between steps).
 * After every step of problem `prob` (AH_CAPTURE_EVERY_STEP), or only after
 * the steps whose dt was clamped to a checkpoint t_end*idx/checkpoints
 * (AH_CAPTURE_CHECKPOINTS, astrea.py:74-78), the snapshot of
 * ah_plan_snapshot goes to slot ctl[0] % slots of `ring` (device memory,
 * slots * cells * 8 doubles), its t to times[slot] (device, `slots` doubles),
 * and ctl[0] (device, 2 long longs; ctl[1] is internal) counts the snapshots
 * taken. Steps that fail or do not happen take none, nor does a step that
 * ends past t_end (the reference's loop stops without saving there). The caller drains the
 * ring before ctl[0] runs `slots` ahead of what it has read. The launches are
 * part of ah_plan_step and of captured step graphs; a call re-captures graphs
 * and restarts the count. AH_CAPTURE_OFF (buffers may be null) stops it. */
enum ah_capture { AH_CAPTURE_OFF = 0, AH_CAPTURE_EVERY_STEP = 1, AH_CAPTURE_CHECKPOINTS = 2 };
int ah_plan_set_capture(ah_plan* plan, int prob, int mode, double* ring, int slots, double* times,
                        long long* ctl, void* stream);

/* Export / import the time state of all problems of a plan (synchronous),
 * so a problem can move between plans with its state planes: the serving
 * loop re-plans its batch as requests arrive and finish.  No reference
 * counterpart (the reference runs one problem per process). */
int ah_plan_time_export(ah_plan* plan, ah_time_state* out, void* stream);
int ah_plan_time_import(ah_plan* plan, const ah_time_state* in, void* stream);

/* Copy the per-problem time state to the host (synchronises the stream). */
int ah_plan_status(ah_plan* plan, ah_problem_status* out, void* stream);

/* The same statuses written by a kernel, in stream order, into `out` (device
 * memory, nprob entries) without synchronising: a caller keeps stepping and
 * reads the failure flags and times of that point of the stream when its copy
 * arrives (the reference has no asynchronous path; its loop checks nothing
 * until an exception). */
int ah_plan_status_async(ah_plan* plan, ah_problem_status* out, void* stream);

/* Parity entry points mirroring the reference's operator interface, for one
 * problem (index `prob`) of a bound plan, at step parity `parity` (0: axis-0
 * sweep first, as on even steps; 1: reversed, astrea.py:85).
 *
 * ah_evolve_space replaces evolvers.evolve_space + solvers.calculate_Riemann_flux:
 * writes, for each physical axis a, the interface fluxes in the sweep frame
 * (flux[a]: nvars planes of (n_a + 1) * n_other doubles, interface-major) and
 * eigmax[a] (functions/fv.py:157-169, closed-form spectral radius).
 * `grid` is the packed state of the whole plan (only problem `prob` is read). */
int ah_evolve_space(ah_plan* plan, int prob, int parity, const double* grid,
                    double* flux0, double* flux1, double* eigmax, void* stream);

/* ah_compute_L replaces compute_L (evolvers.py:41-60): L = -sum_a dF_a/dx for
 * problem `prob`, from fluxes laid out as ah_evolve_space writes them (`L`:
 * the packed layout of the state).  Plans with constrained transport return
 * AH_ERR_UNSUPPORTED: their field rows need the corner EMF, ah_compute_L_ct. */
int ah_compute_L(ah_plan* plan, int prob, int parity, const double* flux0, const double* flux1,
                 double* L, void* stream);

/* Constrained transport (magnetic_2d plans).  ah_evolve_space_ct is
 * ah_evolve_space plus the EMF attachment of evolvers.py:26-32: `corner`
 * receives the corner EMF e3U (mag_field.py:125-187; n0*n1 doubles in the
 * state's (x, y) order, the reference's fluxes[(1,0,2)]['mag_corner'];
 * fluxes[(0,1,2)]['mag_corner'] is its transpose), `face0` / `face1` the
 * components of fluxes[...]['face_avg'] that evolve_time reads: B_x of the
 * x sweep's face values and B_y of the y sweep's, in the state's (x, y)
 * order.  ah_compute_L_ct is compute_L with the EMF-difference rows of B_x,
 * B_y (evolvers.py:51-58) from that corner EMF. */
int ah_evolve_space_ct(ah_plan* plan, int prob, int parity, const double* grid, double* flux0, double* flux1,
                       double* eigmax, double* corner, double* face0, double* face1, void* stream);
int ah_compute_L_ct(ah_plan* plan, int prob, int parity, const double* flux0, const double* flux1,
                    const double* corner, double* L, void* stream);

/* ah_evolve_time replaces evolve_time(grid, fluxes, dt, sim_variables)
 * (evolvers.py:38-206) for problem `prob` of a single-problem plan at step
 * parity `parity`: L0 from the given fluxes (and, with constrained
 * transport, `corner`), the in-place reassignment of B_x, B_y to the face
 * values `face0`, `face1` (evolvers.py:73-76), the plan's Runge-Kutta stages
 * with evolve_space at each later stage and, with constrained transport, the
 * inverse reconstruction of the field after every stage (evolvers.py:63-67).
 * `grid` (the packed state layout) is overwritten with the result; the
 * plan's own state and time state are left as they were.  `corner`, `face0`
 * and `face1` may be null for plans without constrained transport. */
int ah_evolve_time(ah_plan* plan, int prob, int parity, double* grid, const double* flux0, const double* flux1,
                   const double* corner, const double* face0, const double* face1, double dt, void* stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under stream capture, serving loops, harnesses that forbid
 * hipMalloc).  ah_evolve_space for one problem without a plan: the same
 * sweeps, fluxes and eigmax as ah_evolve_space on a single-problem plan of
 * (`scheme`, `problem`), with `grid` the packed state of that plan (nvars
 * planes of n0*n1 doubles).  The problem tables, tile lists and interface-
 * state workspace live in a fixed device array of AH_STATIC_SCRATCH_BYTES
 * per device (held until the process ends); nothing is allocated.  A problem
 * whose workspace exceeds it returns AH_ERR_TOO_LARGE (2D hydrodynamics in
 * FP64 needs 136 bytes per cell: about 490 000 cells).  Calls on one device
 * share the array, so calls that overlap on one device race.  Constrained
 * transport is not supported (AH_ERR_UNSUPPORTED).  Synchronises `stream`. */
#define AH_STATIC_SCRATCH_BYTES (64ll << 20)
int ah_evolve_space_static(const ah_scheme* scheme, const ah_problem* problem, int parity, const double* grid,
                           double* flux0, double* flux1, double* eigmax, void* stream);

/* Kernel timing with HIP events inside ah_plan_step (no reference
 * counterpart: the reference's only timing is perf_counter around a whole run,
 * astrea.py:178-180).  ah_plan_profile(plan, capacity) enables recording of up
 * to `capacity` launches (0 disables); ah_plan_profile_read synchronises,
 * returns the summed milliseconds and launch counts per kernel kind and
 * clears the record.  Without constrained transport a step has no dt launch
 * (its kernels evaluate dt) and AH_K_ADVANCE times the end-of-step launch
 * that stores dt and advances t, which runs only on plans with more than 64
 * combination tiles (1024 cells each in 2D, 256 in 1D) per problem; on the
 * others the step's last combination launch does that work.  The middle
 * SSPRK(3,3) stage has no
 * combination launch (the launch that finishes its L writes K2). */
enum ah_kernel_kind {
    AH_K_SWEEP1D = 0, AH_K_RECON0 = 1, AH_K_FLUX0 = 2, AH_K_RECON1 = 3, AH_K_FLUX1 = 4,
    AH_K_COMBINE = 5, AH_K_DT = 6, AH_K_ADVANCE = 7, AH_NKIND = 8
};
int ah_plan_profile(ah_plan* plan, int capacity);
int ah_plan_profile_read(ah_plan* plan, double* ms, int* counts);

#ifdef __cplusplus
}
#endif
#endif /* ASTREA_HIP_H */
