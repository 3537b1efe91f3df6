/* mh_euler.h -- C ABI of the MI355X (gfx950) MUSCL-Hancock Euler solver.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes; no
 * exceptions and no exit() cross this boundary.  Every entry point takes an
 * explicit HIP stream (pass 0 for the null stream) and, unless it says it
 * synchronises, only enqueues work, so sequences of calls are
 * graph-capturable.
 *
 * The reference (HarryKKKKK/large_sys_eq_opt) has no C ABI; each entry point
 * replaces the C++ free function or method cited beside it (paths relative
 * to the reference repository).
 *
 * Host state arrays use the reference layout: (ny+4)*(nx+4) cells with two
 * ghost layers, flat index j*(nx+4)+i (head/cpu/grid_cpu.hpp:86-88); "AoS"
 * means Conserved{rho,rhou,rhov,E} per cell (head/types.hpp:13-23), "SoA"
 * means four planes rho, rhou, rhov, E (head/gpu/grid_gpu.cuh:43-46).
 * On the device a grid is four FP64 planes with a padded row pitch.
 *
 * Concurrency: a grid carries its own device scratch (a max-speed slot and
 * the y sweep's list of tiles to redo with IEEE division), so calls that
 * write the same grid, or compute dt from it, must not overlap (order them
 * on one stream or with events); calls on different grids may run on
 * different streams at once.  The sweeps also write the column left of the
 * two left ghost columns (row padding that no call reads).
 */
#ifndef MH_EULER_H
#define MH_EULER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mh_grid_s* mh_grid;
typedef struct mh_stepper_s* mh_stepper;
typedef struct mh_batch_s* mh_batch;
typedef struct mh_adjoint_s* mh_adjoint;
typedef struct mh_snapshot_s* mh_snapshot;
typedef void* mh_stream; /* hipStream_t */

/* Status codes (the reference prints and calls std::exit instead:
 * src/gpu/solver_gpu.cu:41-49). */
enum {
    MH_OK = 0,
    MH_ERR_INVALID = 1,   /* bad argument, mismatched grids, unknown solver */
    MH_ERR_OOM = 2,       /* device allocation failed */
    MH_ERR_LAUNCH = 3,    /* kernel launch or HIP runtime failure */
    MH_ERR_WAVESPEED = 4, /* max signal speed <= 0: no finite dt (the CPU
                             reference throws, src/cpu/solver_cpu.cpp:304-306) */
    MH_ERR_NODEVICE = 5   /* no HIP device visible */
};

/* Riemann solvers, numbered as head/riemann.hpp:13-18. */
enum { MH_HLL = 0, MH_HLLC = 1, MH_EXACT = 2, MH_FORCE = 3 };

/* Boundary kinds per edge.  TRANSMISSIVE = zero-gradient copy of the nearest
 * interior cell into both ghost layers (head/cpu/boundary_cpu.hpp:5-36);
 * INFLOW = fixed (Dirichlet) conserved state in the ghost layers. */
enum { MH_BC_TRANSMISSIVE = 0, MH_BC_INFLOW = 1 };
enum { MH_EDGE_LEFT = 0, MH_EDGE_RIGHT = 1, MH_EDGE_BOTTOM = 2, MH_EDGE_TOP = 3 };

typedef struct {
    int kind[4];          /* per edge, MH_EDGE_* order */
    double state[4][4];   /* conserved (rho, rhou, rhov, E) for INFLOW edges */
} mh_bc;

typedef struct {
    int nx, ny;                        /* interior cells */
    double x_min, x_max, y_min, y_max; /* domain; dx = (x_max-x_min)/nx */
} mh_grid_desc;

/* Last error message of the calling thread ("" when none). */
const char* mh_last_error(void);
/* Library version string. */
const char* mh_version(void);
/* Selects the device for the calling thread (hipSetDevice). */
int mh_set_device(int device);

/* ---- grids: replaces Grid2DGPU (head/gpu/grid_gpu.cuh:11-154) ---------- */

/* Grid2DGPU::Grid2DGPU / allocate (head/gpu/grid_gpu.cuh:15-47); ng is 2. */
int mh_grid_create(const mh_grid_desc* desc, mh_grid* out);
/* Grid2DGPU::~Grid2DGPU / release (head/gpu/grid_gpu.cuh:31, :49-59). */
int mh_grid_destroy(mh_grid g);
/* Grid2DGPU::nx() ... dy() accessors (head/gpu/grid_gpu.cuh:61-75); pitch is
 * the device row pitch in doubles, plane_stride the distance between planes. */
int mh_grid_info(mh_grid g, mh_grid_desc* desc, double* dx, double* dy, int64_t* pitch, int64_t* plane_stride);
/* Device address of cell (i, j) of plane c, i, j in [-2, n+2) (interior
 * origin at 0).  For zero-copy interop (e.g. torch tensors over the planes). */
int mh_grid_plane_ptr(mh_grid g, int c, double** ptr);
/* Grid2DGPU::upload_from_aos / download_to_aos (head/gpu/grid_gpu.cuh:90-121):
 * host AoS in the reference layout, ghosts included.  Synchronise. */
int mh_grid_upload_aos(mh_grid g, const double* host_aos);
int mh_grid_download_aos(mh_grid g, double* host_aos);
/* The same with four host SoA planes of (ny+4)*(nx+4) doubles each. */
int mh_grid_upload_soa(mh_grid g, const double* const planes[4]);
int mh_grid_download_soa(mh_grid g, double* const planes[4]);
/* Device-to-device copy of the interior (nx*ny per plane) from/to a dense
 * device buffer laid out [4][ny][nx].  Enqueued on `stream`. */
int mh_grid_copy_interior_from_device(mh_grid g, const double* src, mh_stream stream);
int mh_grid_copy_interior_to_device(mh_grid g, double* dst, mh_stream stream);
/* Grid2DGPU::swap (head/gpu/grid_gpu.cuh:123-128): exchanges storage. */
int mh_grid_swap(mh_grid a, mh_grid b);

/* ---- boundary: replaces apply_transmissive_boundary_gpu
 *      (head/gpu/boundary_gpu.cuh:5, src/gpu/boundary_gpu.cu:71-84) -------- */
int mh_fill_ghosts(mh_grid g, const mh_bc* bc, mh_stream stream);

/* ---- time step: replaces compute_dt_gpu (head/gpu/solver_gpu.cuh:25-29,
 *      src/gpu/solver_gpu.cu:741-782).  Synchronises `stream`.
 *      dt = cfl*min(dx,dy)/max over interior cells of max(|u|+a, |v|+a);
 *      MH_ERR_WAVESPEED when that max is <= 0. ---------------------------- */
int mh_compute_dt(mh_grid g, double cfl, double* dt, double* max_speed, mh_stream stream);

/* ---- advance: replaces advance_second_order_gpu / advance_first_order_gpu
 *      (head/gpu/solver_gpu.cuh:31-44, src/gpu/solver_gpu.cu:788-885) and
 *      the CPU twins (head/cpu/solver_cpu.hpp:46-79).
 *      Second order: x sweep Uold -> Utmp, ghosts of Utmp, y sweep
 *      Utmp -> Unew, ghosts of Unew (src/cpu/solver_cpu.cpp:426-498).
 *      Unew may alias Uold (two-buffer stepping); Utmp must be distinct.
 *      bc NULL = transmissive on every edge. --------------------------------- */
int mh_advance_second_order(mh_grid Uold, mh_grid Utmp, mh_grid Unew, double dt, int solver, const mh_bc* bc,
                            mh_stream stream);
int mh_advance_first_order(mh_grid Uold, mh_grid Unew, double dt, int solver, const mh_bc* bc, mh_stream stream);
/* One sweep alone (dir 0 = x, 1 = y) over interior cells of `in` -> `out`. */
int mh_sweep(mh_grid in, mh_grid out, double dt, int dir, int solver, mh_stream stream);

/* ---- device-resident time loop (no host round trip per step).  Replaces the
 *      loop of scripts/gpu/main_gpu.cu:243-299: each step computes dt on the
 *      device from the previous step's fused max-signal-speed reduction,
 *      clamps it to t_end - t, advances, and accumulates t.  --------------- */
/* Creates a stepper over state U (owned by the caller) with its own scratch
 * grid.  t_end <= 0: no clamp, steps run unconditionally.  bc NULL =
 * transmissive.  Fills the ghosts of U and reduces its max signal speed.
 * With the launch shape's `fused` set (by default for grids of 2^20 cells
 * or more, 2^24 for the exact solver), each step is one pass and the state
 * alternates between U and the scratch grid: it is in U again after
 * mh_stepper_clock, mh_stepper_finish or mh_stepper_destroy, which is when
 * U may be read.  An exact-solver stepper under the default shape changes
 * between the fused step and the two sweeps at its clock reads, by the
 * share of tiles its fast passes left to the IEEE redo pass (the same
 * results either way). */
int mh_stepper_create(mh_grid U, int solver, double cfl, double t_end, const mh_bc* bc, int dt_log_capacity,
                      mh_stepper* out, mh_stream stream);
int mh_stepper_destroy(mh_stepper s);
/* Enqueues n steps (steps past t_end are no-ops).  No synchronisation. */
int mh_stepper_advance(mh_stepper s, int nsteps, mh_stream stream);
/* Synchronises and reads the clock: time, completed steps, dt log (up to
 * capacity entries; may be NULL).  Returns MH_ERR_WAVESPEED if a step met a
 * non-positive max signal speed. */
int mh_stepper_clock(mh_stepper s, double* t, int* steps, double* dt_log, mh_stream stream);
/* Re-fills the ghosts of the state (after the last step); synchronises
 * first if steps were enqueued since the last mh_stepper_clock. */
int mh_stepper_finish(mh_stepper s, mh_stream stream);
/* Holds the loop after the first step with t >= t_stop (0: no stop) -- the
 * reference writes a snapshot there (scripts/gpu/main_gpu.cu:272-298); steps
 * enqueued past it are no-ops until the next call moves the stop and
 * releases the clock.  Enqueued on `stream`. */
int mh_stepper_set_stop(mh_stepper s, double t_stop, mh_stream stream);
/* mh_stepper_advance with a HIP event pair around every launch; synchronises.
 * ms_total: elapsed time of the n steps; ms_kernels[3]: average duration per
 * step of the clock kernel, the x sweep and the y sweep -- or, for a stepper
 * on the fused step, of the clock kernel, the fused step (with its redo
 * pass) and 0.  (No reference counterpart: the reference times whole runs
 * externally.) */
int mh_stepper_advance_profiled(mh_stepper s, int nsteps, mh_stream stream, double* ms_total, double* ms_kernels);
/* Launch shape of the sweeps (process-wide; chosen by the autotuner).  No
 * reference counterpart: src/gpu/solver_gpu.cu:830-831 hard-codes 16x16
 * blocks. */
typedef struct {
    int seg_x;    /* x tile: cells per row segment (0 = whole row) */
    int seg_y;    /* y tile: rows per chunk (0 = whole column) */
    int y_waves;  /* y sweep register target: 4 wavefronts/SIMD, or 1 = compiler's choice */
    int x_layout; /* x tile LDS exchange: 0 compact (6.3 KB, selects), 1 66-slot (10.3 KB), 2 compact with the update deferred one iteration (two LDS round trips instead of three) */
    int x_wpb;    /* wavefronts (tiles) per workgroup, x sweep: 1 or 4 */
    int y_wpb;    /* wavefronts (tiles) per workgroup, y sweep: 1 or 4 */
    int min_tiles; /* a sweep whose segments give fewer tiles than this gets
                      shorter ones (x: multiples of 60 cells, at least 60;
                      y: at least 4 rows) and, while still below it, one
                      tile per workgroup; 0 keeps the shape as set
                      (default 8192) */
    int fused;    /* steppers created from now on run each step as one fused
                     x-then-y pass, with seg_y rows per chunk: 1 (default)
                     on grids of at least 2^20 cells (exact solver: 2^24),
                     2 on every grid; 0: always the x sweep and the y
                     sweep */
    int fused_tail; /* the fused step's last rows of every strip in chunks
                       of this many rows, dispatched last: about
                       min_tiles / 2 such tiles, at most half the rows, and
                       only when shorter than the seg_y chunks; 0: none
                       (default 16) */
} mh_launch_shape;
int mh_set_launch_shape(const mh_launch_shape* shape);
int mh_get_launch_shape(mh_launch_shape* shape);

/* ---- snapshot output: replaces the download + write_aos_csv of the
 *      reference's loop (scripts/gpu/main_gpu.cu:272-298, :146-171;
 *      Grid2DGPU::download_to_aos, head/gpu/grid_gpu.cuh:106-121).  A capture
 *      copies the state on the device (on the caller's stream) and then to
 *      pinned host memory on the snapshot's own copy stream, so stepping
 *      continues while the data travels; the CSV writer formats rows on
 *      several host threads. ---------------------------------------------- */
/* nbuf slots (1..8), each a device copy and pinned host planes of a grid
 * shaped like `like`. */
int mh_snapshot_create(mh_grid like, int nbuf, mh_snapshot* out);
int mh_snapshot_destroy(mh_snapshot s);
/* Enqueues the capture of U into `slot` (waits on the device for that
 * slot's previous host copy).  The caller must be done with the slot's host
 * planes from an earlier capture. */
int mh_snapshot_capture(mh_snapshot s, mh_grid U, int slot, mh_stream stream);
/* Waits for the slot's host copy; planes (may be NULL) receives its four
 * host planes of (ny+4)*(nx+4) doubles in the reference's cell order
 * (index j*(nx+4)+i, ghosts included). */
int mh_snapshot_wait(mh_snapshot s, int slot, double** planes);
/* Waits for the slot, then writes it as the reference's snapshot CSV
 * (header i,j,x,y,rho,rhou,rhov,E; one row per interior cell; i, j with the
 * ghost offset; values as %.{precision}g, the reference's ostream default
 * being 6; scripts/gpu/main_gpu.cu:146-171) on nthreads host threads
 * (0: up to 16). */
int mh_snapshot_write_csv(mh_snapshot s, int slot, const char* path, int precision, int nthreads);
/* The same CSV from a host state in the reference AoS layout
 * ((ny+4)*(nx+4) Conserved cells) of the grid `desc` describes. */
int mh_write_csv_aos(const double* host_aos, const mh_grid_desc* desc, int precision, int nthreads, const char* path);

/* ---- initial conditions on the device (point sampling at cell centres,
 *      ghosts included: src/init.cpp:29-38).  kind 0: reference shock_bubble
 *      (src/test_cases.cpp:24-69); 1: reference blast_wave (:78-101);
 *      2: blast with p[0..6] = x0,y0,r0,rho_in,p_in,rho_out,p_out;
 *      3: shock-bubble with p[0..7] = rho0,p0,mach,shock_x,bubble_rho,
 *      bubble_x,bubble_y,bubble_r. ------------------------------------------ */
int mh_init_state(mh_grid g, int kind, const double* params, mh_stream stream);

/* ---- batched ensemble (no reference counterpart): n independent grids of
 *      uneven size, each with its own CFL dt, advanced one split step per
 *      persistent launch; step state lives on the device, so a captured
 *      graph of k launches replays k steps. -------------------------------- */
/* Allocates n instance grids (descs[k]) and their scratch in one arena.
 * bcs: n boundary specs or NULL (transmissive). */
int mh_batch_create(int n, const mh_grid_desc* descs, int solver, double cfl, const mh_bc* bcs, mh_batch* out);
int mh_batch_destroy(mh_batch b);
/* Non-owning handle of instance k: upload/download/init/fill through the
 * grid entry points (mh_grid_destroy refuses it). */
int mh_batch_grid(mh_batch b, int k, mh_grid* g);
/* After the instances are initialised: fills ghosts, resets the clocks and
 * reduces every instance's initial max signal speed.  Synchronises. */
int mh_batch_prepare(mh_batch b, mh_stream stream);
/* Launch schedule: x tile row segment (0 = whole row) and rows per x tile
 * (>= 1), y tile chunk rows (0 = whole column), mode (0: one persistent
 * launch per step; 1: two persistent launches, x then y; 2: two dispatched
 * launches with one tile per wavefront; 3: the fused step, a wavefront pair
 * per y tile sweeping x into LDS and y out of it, plus its redo launch),
 * persistent
 * workgroups per CU (0 = occupancy limit).  Under mode 3 an instance's state
 * alternates between its grid and its scratch grid; mh_batch_sync,
 * mh_batch_clock, mh_batch_set_schedule and mh_batch_set_size_class
 * synchronise and put it back in the grid, so read instance grids after one
 * of them.  Drops a captured graph. */
int mh_batch_set_schedule(mh_batch b, int xseg, int xrows, int ychunk, int mode, int blocks_per_cu);
/* Size class of the schedule: instances with fewer than side_below cells a
 * side (nx) take y chunks of ychunk rows and x tiles of xrows rows instead
 * (0 keeps the schedule's value; side_below 0 turns the class off).  Drops
 * a captured graph. */
int mh_batch_set_size_class(mh_batch b, int side_below, int ychunk, int xrows);
/* Enqueues n steps of every instance (graph replays where captured). */
int mh_batch_step(mh_batch b, int nsteps, mh_stream stream);
/* Captures steps_per_graph steps into a hipGraph used by mh_batch_step. */
int mh_batch_capture(mh_batch b, int steps_per_graph, mh_stream stream);
/* Synchronises (and puts each instance's state back into its grid after
 * mode-3 steps); per-instance time, last dt, steps, status (arrays of n or
 * NULL). */
int mh_batch_clock(mh_batch b, double* t, double* dt, int* steps, int* status, mh_stream stream);
/* Synchronises and puts each instance's state back into its grid after
 * mode-3 steps (no status check): call before reading instance grids. */
int mh_batch_sync(mh_batch b, mh_stream stream);
/* Persistent grid size and tile counts of the current schedule. */
int mh_batch_info(mh_batch b, int* nblocks, int* nxtiles, int* nytiles);

/* ---- discrete adjoint (no reference counterpart): reverse mode of the
 *      second-order split step at a fixed dt (all four Riemann solvers;
 *      the exact solver's p* differentiated implicitly), with checkpointed
 *      forward/backward over N steps. ------------------------------------ */
/* Workspace for grids of shape desc: checkpoints every checkpoint_every
 * steps for up to max_steps steps. */
int mh_adjoint_create(const mh_grid_desc* desc, int solver, const mh_bc* bc, int checkpoint_every, int max_steps,
                      mh_adjoint* out);
int mh_adjoint_destroy(mh_adjoint a);
/* lam_out = (dU^{n+1}/dU^n)^T lam_in for one split step at state U and time
 * step dt (interior adjoints; U's ghosts are filled by the call). */
int mh_adjoint_step_vjp(mh_adjoint a, mh_grid U, double dt, mh_grid lam_in, mh_grid lam_out, mh_stream stream);
/* The same for one sweep alone (dir 0 = x, 1 = y): lam_out covers every
 * cell the sweep reads, ghost layers included (no boundary folding). */
int mh_adjoint_sweep_vjp(mh_adjoint a, mh_grid U, double dt, int dir, mh_grid lam_in, mh_grid lam_out,
                         mh_stream stream);
/* nsteps steps from U0 (unchanged) with CFL dt (or the sequence dt_in),
 * checkpointing; dt_out (nsteps, may be NULL) receives the dt sequence, UN
 * (may be NULL) the final state.  Synchronises. */
int mh_adjoint_forward(mh_adjoint a, mh_grid U0, int nsteps, double cfl, const double* dt_in, double* dt_out,
                       mh_grid UN, mh_stream stream);
/* lam0 = dJ/dU^0 from lamN = dJ/dU^N over the steps of the last forward call
 * (each checkpoint segment recomputed once).  Synchronises. */
int mh_adjoint_backward(mh_adjoint a, mh_grid lamN, mh_grid lam0, mh_stream stream);
/* dt sequence and step count of the last forward call. */
int mh_adjoint_dt(mh_adjoint a, double* dt, int* nsteps);
/* Zero every cell of a grid / copy every cell between equal-shape grids. */
int mh_grid_zero(mh_grid g, mh_stream stream);
int mh_grid_copy(mh_grid src, mh_grid dst, mh_stream stream);

/* ---- utilities (no reference counterpart) -------------------------------- */
/* Plain streaming copy of n doubles (n even) between device buffers: the
 * measured copy bandwidth and the calibration of HBM traffic counters. */
int mh_copy(const double* src, double* dst, int64_t n, mh_stream stream);
int mh_device_alloc(int64_t bytes, void** ptr);
int mh_device_free(void* ptr);
int mh_synchronize(mh_stream stream);

/* ---- pointwise entry points over host arrays of n states (synchronous) ---- */
/* riemann_flux (head/riemann.hpp:650-672): F = flux(UL, UR) along dir for n
 * faces of AoS conserved states.  iters (may be NULL) receives the exact
 * solver's Newton iteration count per face (0 for the other solvers). */
int mh_riemann_flux(int n, const double* UL, const double* UR, int dir, int solver, double* F, int* iters);
/* No reference counterpart.  mh_riemann_flux for callers that cannot
 * allocate device memory (code in tight serving loops, harnesses that forbid
 * hipMalloc): the same results, staged through a fixed device array of the
 * library (6.5 MB per device, held until the process ends) instead of
 * per-call allocations; n <= MH_NOALLOC_MAX_FACES, larger n is refused with
 * MH_ERR_INVALID.  Calls on one device share that array: calls that overlap
 * on one GPU (several host threads) race. */
#define MH_NOALLOC_MAX_FACES 65536
int mh_riemann_flux_noalloc(int n, const double* UL, const double* UR, int dir, int solver, double* F, int* iters);
/* No reference counterpart: the sweeps' fast arithmetic pointwise, for tests
 * that pin it to the compiler's (host buffers of n values; U is n AoS
 * states).  fast = 1: q = a/b and r = sqrt(a) by the unscaled Newton
 * sequences, s = the signal speed max(|u|+c, |v|+c) as the y sweep's fused
 * dt reduction forms it; fast = 2: the same with the division's final
 * special-case fixup left out (as the sweeps do where operands are normal);
 * fast = 0: the compiler's division and square root. */
int mh_fast_ops(int n, const double* a, const double* b, const double* U, int fast, double* q, double* r, double* s);
/* reconstruct_cell_muscl_hancock (src/cpu/solver_cpu.cpp:68-99) for n cells:
 * evolved left/right edge states from (Um, Uc, Up) and dt/d. */
int mh_reconstruct(int n, const double* Um, const double* Uc, const double* Up, double dt_over_d, int dir, double* UL,
                   double* UR);

#ifdef __cplusplus
}
#endif

#endif /* MH_EULER_H */
