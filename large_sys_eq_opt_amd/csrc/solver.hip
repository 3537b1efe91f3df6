// solver.hip -- gfx950 kernels of the plain MUSCL-Hancock path and the C ABI
// over them (include/mh_euler.h).
//
// Kernels
//   k_sweep_y<S>   column-streaming y sweep: each lane owns one column, the
//                  wavefront covers 64 adjacent columns (row-coalesced loads)
//                  and marches up a chunk of rows with a register window.
//                  Each cell is reconstructed once and each face solved once.
//                  Epilogue: left/right ghost columns of the output and the
//                  wavefront max of the output's signal speed (one atomic max
//                  per wavefront) -- the next step's dt needs no extra pass.
//   k_sweep_x<S>   row-streaming x sweep: the wavefront marches along a row
//                  segment 64 cells at a time; neighbour values cross lanes
//                  through LDS with a two-cell lag, again one reconstruction
//                  per cell and one Riemann solve per face.  Epilogue: bottom/
//                  top ghost rows of the output (read by the y sweep).
//   k_max_speed    wavefront-shuffle max of the signal speed (first step).
//   k_clock        one thread: commits the previous step (t += dt), forms dt
//                  from the reduced max speed and clamps it to t_end.
//   k_fill_ghosts  all ghost cells from the boundary spec.
//   k_first_order<S>, k_init   first-order step, initial conditions.
// The reference launches 16x16 blocks that reconstruct 18/16 cells and solve
// 17/16 faces per output cell (src/gpu/solver_gpu.cu:301-683).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "euler_device.h"
#include "grid.h"
#include "sweep_tiles.h"
#include "clock.h"

namespace mh {

// ---------------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------------
static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }
int fail(int code, const std::string& m) {
    g_err = m;
    return code;
}
int check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return MH_OK;
    return fail(e == hipErrorOutOfMemory ? MH_ERR_OOM : MH_ERR_LAUNCH,
                std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// device clock and sweep kernels
// ---------------------------------------------------------------------------
struct SweepArgs {
    DtSrc dts;
    double d;          // dx or dy
    int seg;           // rows (y) or cells (x) per wavefront
    int ngroups;       // column groups (y) or segments per row (x)
    int ghosts;        // write the fused ghost layers of the output
    int ntiles;        // wavefront tiles in the launch
    Bc bc;
    // y sweep only: where the wavefront max of the output's signal speed
    // goes -- a fixed slot, or (clock_speed) the clock's inactive slot.
    unsigned long long* speed;
    int clock_speed;
    int* redo;         // y sweep: the output grid's redo list (see k_sweep_y)
};

// Sweep kernels: WPB independent wavefronts per workgroup, one tile each
// (wavefront-level synchronisation only).  For the y sweep, W is the
// wavefronts per SIMD the register allocation aims for (1 = compiler's
// choice; four fit the HLL/HLLC/FORCE y tile in 128 VGPRs).  For the x
// sweep, L selects the LDS exchange layout.
//
// The y sweep runs each tile's fast pass (y_tile_pass<S, true>).  A tile
// whose states leave the fast-path range is, with RL = 1, appended to the
// output grid's redo list and left to k_sweep_y_redo, which the host
// launches right after on the same stream (it adds the tile's speed
// maximum); with RL = 0 the IEEE pass runs in place.  Keeping the IEEE pass
// out of the kernel keeps it within 128 VGPRs (four wavefronts per SIMD);
// small grids, where the second launch would cost more than occupancy
// gains, use RL = 0.
__device__ __forceinline__ unsigned long long* speed_slot(const SweepArgs& a) {
    return a.clock_speed ? &a.dts.clock->smax[a.dts.clock->slot ^ 1] : a.speed;
}

template <int S, int W, int WPB, int RL>
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(W, 8))) void k_sweep_y(View in, View out,
                                                                                               SweepArgs a) {
    double dt;
    if (!dt_of(a.dts, dt)) return;
    const double dtd = dt / a.d;
    const double half = 0.5 * dtd;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int tile = blockIdx.x * WPB + w;
    if (tile >= a.ntiles) return;
    const int cg = tile % a.ngroups;
    const int ch = tile / a.ngroups;
    const int j0 = ch * a.seg;
    const int j1 = min(j0 + a.seg, in.ny);
    const int want = (a.speed || a.clock_speed) ? 1 : 0;
    __shared__ double ring[WPB][kRing];
    double m = 0.0;
    if (!(dt_in_fast_range(dtd) &&
          y_tile_pass<S, true>(in, out, cg, j0, j1, dtd, half, a.bc, a.ghosts, want, ring[w], lane, m))) {
        if constexpr (RL) {
            if (lane == 0) a.redo[2 + atomicAdd(a.redo, 1)] = tile;
            return;
        } else {
            y_tile_pass<S, false>(in, out, cg, j0, j1, dtd, half, a.bc, a.ghosts, want, ring[w], lane, m);
        }
    }
    if (want) {
        m = wave_max(m);
        if (lane == 0) atomic_max_pos(speed_slot(a), m);
    }
}

// IEEE pass over the tiles the preceding k_sweep_y put on the redo list.  A
// list too short to fill the chip splits each tile into pieces of at least
// kRedoRowsY rows (with their own warm-up rows) that run side by side, so a
// few redone tiles cost a few pieces' chains of rows rather than a whole
// tile's each; a long list keeps whole tiles.  The last workgroup to finish
// empties the list for the next launch.
constexpr int kRedoRowsY = 16, kRedoPiecesY = 4096;
template <int S>
__global__ __launch_bounds__(256) void k_sweep_y_redo(View in, View out, SweepArgs a) {
    __shared__ double ring[4][kRing];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int n = __hip_atomic_load(a.redo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    double dt;
    if (n > 0 && dt_of(a.dts, dt)) {
        const double dtd = dt / a.d;
        const double half = 0.5 * dtd;
        const int want = (a.speed || a.clock_speed) ? 1 : 0;
        // pieces per tile and rows per piece (a tile's last piece may be
        // shorter, or empty)
        const int pieces = max(1, min((a.seg + kRedoRowsY - 1) / kRedoRowsY, (kRedoPiecesY + n - 1) / n));
        const int rows = (a.seg + pieces - 1) / pieces;
        for (int q = blockIdx.x * 4 + w; q < n * pieces; q += gridDim.x * 4) {
            const int tile = a.redo[2 + q / pieces];
            const int cg = tile % a.ngroups;
            const int t0 = (tile / a.ngroups) * a.seg;
            const int j0 = t0 + (q % pieces) * rows;
            const int j1 = min(min(j0 + rows, t0 + a.seg), in.ny);
            if (j0 >= j1) continue;
            double m = 0.0;
            y_tile_pass<S, false>(in, out, cg, j0, j1, dtd, half, a.bc, a.ghosts, want, ring[w], lane, m);
            if (want) {
                m = wave_max(m);
                if (lane == 0) atomic_max_pos(speed_slot(a), m);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && atomicAdd(a.redo + 1, 1) == (int)gridDim.x - 1) {
        if (a.dts.clock) a.dts.clock->redo_tiles += n;
        a.redo[0] = 0;
        a.redo[1] = 0;
    }
}

template <int S, int L, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_sweep_x(View in, View out, SweepArgs a) {
    double dt;
    if (!dt_of(a.dts, dt)) return;
    const double dtd = dt / a.d;
    const double half = 0.5 * dtd;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int tile = blockIdx.x * WPB + w;
    if (tile >= a.ntiles) return;
    const int j = tile / a.ngroups;
    const int sg = tile - j * a.ngroups;
    const int i0 = sg * a.seg;
    const int i1 = min(i0 + a.seg, in.nx);
    __shared__ typename XSlice<L>::type X[WPB];
    x_tile<S, L>(in, out, j, i0, i1, dtd, half, a.bc, a.ghosts, X[w], lane);
}

// ---------------------------------------------------------------------------
// first-order unsplit Godunov step (src/cpu/solver_cpu.cpp:315-369)
// ---------------------------------------------------------------------------
template <int S>
__global__ __launch_bounds__(256) void k_first_order(View in, View out, double dtdx, double dtdy) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= in.nx || j >= in.ny) return;
    int it = 0;
    const Cons Uc = ld(in, i, j), Uw = ld(in, i - 1, j), Ue = ld(in, i + 1, j);
    const Cons Us = ld(in, i, j - 1), Un = ld(in, i, j + 1);
    const Prim Vc = c2p(Uc), Vw = c2p(Uw), Ve = c2p(Ue), Vs = c2p(Us), Vn = c2p(Un);
    const Cons Fxp = riemann<S, kX>(Uc, Vc, Ue, Ve, it);
    const Cons Fxm = riemann<S, kX>(Uw, Vw, Uc, Vc, it);
    const Cons Fyp = riemann<S, kY>(Uc, Vc, Un, Vn, it);
    const Cons Fym = riemann<S, kY>(Us, Vs, Uc, Vc, it);
    const Cons Ux = update(Uc, dtdx, Fxm, Fxp);
    st(out, i, j, update(Ux, dtdy, Fym, Fyp));
}

// ---------------------------------------------------------------------------
// max signal speed over interior cells -> slot (bit pattern of a double)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_max_speed(View v, unsigned long long* slot) {
    const int64_t n = (int64_t)v.nx * v.ny;
    double m = 0.0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(c / v.nx);
        const int i = (int)(c - (int64_t)j * v.nx);
        m = smax(m, signal_speed(ld(v, i, j)));
    }
    m = wave_max(m);
    __shared__ double part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) atomic_max_pos(slot, smax(smax(part[0], part[1]), smax(part[2], part[3])));
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0ull;
}

// ---------------------------------------------------------------------------
// device clock: commit the pending step, then (unless commit_only) form the
// next dt = cfl*h/smax clamped to t_end - t (scripts/cpu/main_cpu.cpp:221-228).
// ---------------------------------------------------------------------------
// `redo` (a fused stepper's redo list, or null) is emptied for the step.
__global__ void k_clock(Clock* c, double* dt_log, int commit_only, int* redo) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (redo) {
        c->redo_tiles += redo[0];
        redo[0] = 0;
    }
    if (c->pending) {
        c->t += c->dt;
        c->step += 1;
        c->pending = 0;
        c->slot ^= 1;
        c->cur ^= c->pp;
    }
    if (commit_only || c->done) return;
    if (c->t_end > 0.0 && !(c->t < c->t_end)) {
        c->done = 1;
        return;
    }
    if (c->t_stop > 0.0 && !(c->t < c->t_stop)) {
        c->done = 1;
        c->stopped = 1;
        return;
    }
    const double m = __longlong_as_double((long long)c->smax[c->slot]);
    if (!(m > 0.0)) {
        c->done = 1;
        c->status = MH_ERR_WAVESPEED;
        return;
    }
    double dt = c->cfl * c->h / m;
    if (c->t_end > 0.0 && c->t + dt > c->t_end) dt = c->t_end - c->t;
    c->dt = dt;
    c->smax[c->slot ^ 1] = 0ull;
    if (dt_log && c->step < c->log_cap) dt_log[c->step] = dt;
    c->pending = 1;
}

// Starts a clock at t = 0 with both speed slots clear.
__global__ void k_clock_init(Clock* c, double t_end, double cfl, double h, int log_cap, int pp) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    c->pp = pp;
    c->cur = 0;
    c->t = 0.0;
    c->dt = 0.0;
    c->t_end = t_end;
    c->cfl = cfl;
    c->h = h;
    c->t_stop = 0.0;
    c->smax[0] = 0ull;
    c->smax[1] = 0ull;
    c->step = 0;
    c->slot = 0;
    c->pending = 0;
    c->done = 0;
    c->status = 0;
    c->log_cap = log_cap;
    c->stopped = 0;
    c->redo_tiles = 0;
}

// The host has swapped the stepper's two grids: the state is in the first.
__global__ void k_clock_cur0(Clock* c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) c->cur = 0;
}

// The stepper changes between the fused step (pp = 1) and the two sweeps;
// the state is in its grid.
__global__ void k_clock_pp(Clock* c, int pp) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        c->pp = pp;
        c->cur = 0;
    }
}

// Moves the clock's stop time; releases a clock held at the previous one.
__global__ void k_clock_stop(Clock* c, double t_stop) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    c->t_stop = t_stop;
    if (c->stopped) {
        c->stopped = 0;
        c->done = 0;
    }
}

// ---------------------------------------------------------------------------
// ghost fill: every ghost cell from the interior and the boundary spec, with
// the reference's order (x ghosts of interior rows first, then y ghosts of
// all columns copying row 0 / ny-1 including its x ghosts:
// head/cpu/boundary_cpu.hpp:5-36).
// ---------------------------------------------------------------------------
__device__ __forceinline__ Cons ghost_value(const View& v, const Bc& bc, int i, int j) {
    if (j < 0 && bc.kind[MH_EDGE_BOTTOM] == MH_BC_INFLOW) return bc.s[MH_EDGE_BOTTOM];
    if (j >= v.ny && bc.kind[MH_EDGE_TOP] == MH_BC_INFLOW) return bc.s[MH_EDGE_TOP];
    const int jj = j < 0 ? 0 : (j >= v.ny ? v.ny - 1 : j);
    if (i < 0 && bc.kind[MH_EDGE_LEFT] == MH_BC_INFLOW) return bc.s[MH_EDGE_LEFT];
    if (i >= v.nx && bc.kind[MH_EDGE_RIGHT] == MH_BC_INFLOW) return bc.s[MH_EDGE_RIGHT];
    const int ii = i < 0 ? 0 : (i >= v.nx ? v.nx - 1 : i);
    return ld(v, ii, jj);
}

// Thread k < 4*(ny): x ghosts of interior rows; then 4*(nx+4): y ghosts.
__global__ __launch_bounds__(256) void k_fill_ghosts(View v, Bc bc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int nxg = 4 * v.ny;
    const int nyg = 4 * (v.nx + 4);
    if (k < nxg) {
        const int j = k >> 2, g = k & 3;
        const int i = g < 2 ? -1 - g : v.nx + (g - 2);
        st(v, i, j, ghost_value(v, bc, i, j));
    } else if (k < nxg + nyg) {
        const int q = k - nxg;
        const int i = q / 4 - 2, g = q & 3;
        const int j = g < 2 ? -1 - g : v.ny + (g - 2);
        st(v, i, j, ghost_value(v, bc, i, j));
    }
}

// ---------------------------------------------------------------------------
// initial conditions (point sampling at every cell centre, src/init.cpp:29-38)
// ---------------------------------------------------------------------------
struct InitArgs {
    int kind;         // 0 blast, 1 shock-bubble
    double xmin, ymin, dx, dy;
    double x0, y0, r0;
    Prim in, out;     // blast: inside / outside; shock-bubble: quiescent / post-shock
    double shock_x, bubble_rho;
};

__device__ __forceinline__ Cons make_cons(double rho, double u, double v, double p) {  // src/test_cases.cpp:11-18
    return {rho, rho * u, rho * v, p / kGm1 + 0.5 * rho * (u * u + v * v)};
}

__global__ __launch_bounds__(256) void k_init(View v, InitArgs a) {
    const int64_t tnx = v.nx + 4;
    const int64_t n = tnx * (v.ny + 4);
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(c / tnx) - 2;
        const int i = (int)(c % tnx) - 2;
        const double x = a.xmin + ((double)i + 0.5) * a.dx;
        const double y = a.ymin + ((double)j + 0.5) * a.dy;
        const double ddx = x - a.x0, ddy = y - a.y0;
        Cons U;
        if (a.kind == 0) {
            const double r2 = ddx * ddx + ddy * ddy;
            U = (r2 <= a.r0 * a.r0) ? make_cons(a.in.r, 0.0, 0.0, a.in.p) : make_cons(a.out.r, 0.0, 0.0, a.out.p);
        } else {
            double rho = a.in.r, u = 0.0, w = 0.0, p = a.in.p;
            if (x < a.shock_x) { rho = a.out.r; u = a.out.u; w = 0.0; p = a.out.p; }
            if (ddx * ddx + ddy * ddy <= a.r0 * a.r0) rho = a.bubble_rho;
            U = make_cons(rho, u, w, p);
        }
        st(v, i, j, U);
    }
}

// ---------------------------------------------------------------------------
// pointwise evaluation (riemann_flux of head/riemann.hpp:650-672 and the
// reconstruction of src/cpu/solver_cpu.cpp:68-99) over arrays of states
// ---------------------------------------------------------------------------
template <int S, int D>
__global__ __launch_bounds__(256) void k_face_flux(int n, const double* UL, const double* UR, double* F, int* iters) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const Cons L = {UL[4 * k], UL[4 * k + 1], UL[4 * k + 2], UL[4 * k + 3]};
    const Cons R = {UR[4 * k], UR[4 * k + 1], UR[4 * k + 2], UR[4 * k + 3]};
    int it = 0;
    const Cons f = riemann<S, D>(L, c2p(L), R, c2p(R), it);
    F[4 * k] = f.r; F[4 * k + 1] = f.mu; F[4 * k + 2] = f.mv; F[4 * k + 3] = f.e;
    if (iters) iters[k] = it;
}

template <int D>
__global__ __launch_bounds__(256) void k_recon(int n, const double* Um, const double* Uc, const double* Up, double half,
                                               double* L, double* R) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const Cons a = {Um[4 * k], Um[4 * k + 1], Um[4 * k + 2], Um[4 * k + 3]};
    const Cons b = {Uc[4 * k], Uc[4 * k + 1], Uc[4 * k + 2], Uc[4 * k + 3]};
    const Cons c = {Up[4 * k], Up[4 * k + 1], Up[4 * k + 2], Up[4 * k + 3]};
    Edge l, r;
    reconstruct<D>(c2p(a), c2p(b), c2p(c), half, l, r);
    L[4 * k] = l.U.r; L[4 * k + 1] = l.U.mu; L[4 * k + 2] = l.U.mv; L[4 * k + 3] = l.U.e;
    R[4 * k] = r.U.r; R[4 * k + 1] = r.U.mu; R[4 * k + 2] = r.U.mv; R[4 * k + 3] = r.U.e;
}

// Pointwise fast-path arithmetic for the tests: q = a / b, r = sqrt(a) and
// s = signal speed of U with the unscaled sequences (fast = 1; fast = 2
// divides without the fixup) or the compiler's (fast = 0).
__global__ __launch_bounds__(256) void k_fast_ops(int n, const double* a, const double* b, const double* U, int fast,
                                                 double* q, double* r, double* s) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const Cons u = {U[4 * k], U[4 * k + 1], U[4 * k + 2], U[4 * k + 3]};
    if (fast) {
        q[k] = fast == 2 ? div_nr_nf(a[k], b[k]) : div_nr(a[k], b[k]);
        r[k] = sqrt_nr(a[k]);
        s[k] = signal_speed<true>(u);
    } else {
        q[k] = a[k] / b[k];
        r[k] = sqrt(a[k]);
        s[k] = signal_speed<false>(u);
    }
}

// Plain streaming copy (calibrates traffic counters; measured copy peak).
// Launched with one 16-byte element per thread: the resident wavefronts then
// cover one contiguous window of the buffers, which HBM serves at 6.2-6.3
// TB/s on MI355X; the same kernel on 8192 workgroups striding over the
// whole buffer reached 4.7-4.8 TB/s.
__global__ __launch_bounds__(256) void k_copy(const double2* __restrict__ src, double2* __restrict__ dst, int64_t n2) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n2; k += (int64_t)gridDim.x * blockDim.x)
        dst[k] = src[k];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
Bc to_bc(const mh_bc* b) {
    Bc o;
    for (int e = 0; e < 4; ++e) {
        o.kind[e] = b ? b->kind[e] : MH_BC_TRANSMISSIVE;
        o.s[e] = b ? Cons{b->state[e][0], b->state[e][1], b->state[e][2], b->state[e][3]} : Cons{0, 0, 0, 0};
    }
    return o;
}

int valid_bc(const mh_bc* b) {
    if (!b) return 1;
    for (int e = 0; e < 4; ++e)
        if (b->kind[e] != MH_BC_TRANSMISSIVE && b->kind[e] != MH_BC_INFLOW) return 0;
    return 1;
}

// Sweep launch shapes (the autotuner may override them).
// The defaults keep the number of wavefronts well above the resident slots
// for large grids (whole rows and 256-row chunks leave a second, partly
// filled round of wavefronts at 8192^2); 1020 = 16 x 64 - 4 cells fill the
// x tile's 64-lane iterations exactly.  Compact x exchange with the update
// deferred one iteration (two LDS round trips per iteration instead of
// three: x sweep 2.7 % faster for HLLC, 5 % for HLL and FORCE), one
// wavefront per workgroup, and four y tiles per workgroup at the compiler's
// register choice measured fastest for HLLC at 8192^2 on MI355X; the
// autotuner can override every field.
// The shape is process-wide: mh_set_launch_shape replaces it under a lock
// and every launch reads one consistent copy (shape()), so a host thread
// that sets it while another launches races on nothing.
static mh_launch_shape g_shape = {1020, 128, 4, 2, 1, 4, 8192, 1, 16};
static std::mutex g_shape_lock;

mh_launch_shape shape() {
    std::lock_guard<std::mutex> lk(g_shape_lock);
    return g_shape;
}

template <int S, int L, int WPB>
void go_x(dim3 g, hipStream_t st, const View& vi, const View& vo, const SweepArgs& a) {
    hipLaunchKernelGGL((k_sweep_x<S, L, WPB>), g, dim3(64 * WPB), 0, st, vi, vo, a);
}
// The in-place variant (RL = 0: small grids, the IEEE pass inside the kernel)
// keeps the compiler's register choice: at the four-wave target it spills.
// The redo pass: 64 workgroups of four wavefronts walk the list (a few
// tiles outside the fast range); for the exact solver, whose fast pass
// leaves every tile with a non-uniform face to it, one wavefront per two
// tiles.
template <int S, int W, int WPB>
void go_y(dim3 g, hipStream_t st, const View& vi, const View& vo, const SweepArgs& a) {
    if (a.redo) {
        hipLaunchKernelGGL((k_sweep_y<S, W, WPB, 1>), g, dim3(64 * WPB), 0, st, vi, vo, a);
        const unsigned nredo = S == kExact ? (unsigned)std::max(64, (a.ntiles + 7) / 8) : 64u;
        hipLaunchKernelGGL(k_sweep_y_redo<S>, dim3(nredo), dim3(256), 0, st, vi, vo, a);
    } else {
        hipLaunchKernelGGL((k_sweep_y<S, 1, WPB, 0>), g, dim3(64 * WPB), 0, st, vi, vo, a);
    }
}
template <int S>
void go_x_s(int layout, int wpb, dim3 g, hipStream_t st, const View& vi, const View& vo, const SweepArgs& a) {
    if (layout == 0) (wpb == 4 ? go_x<S, 0, 4> : go_x<S, 0, 1>)(g, st, vi, vo, a);
    else if (layout == 2) (wpb == 4 ? go_x<S, 2, 4> : go_x<S, 2, 1>)(g, st, vi, vo, a);
    else (wpb == 4 ? go_x<S, 1, 4> : go_x<S, 1, 1>)(g, st, vi, vo, a);
}
template <int S>
void go_y_s(int waves, int wpb, dim3 g, hipStream_t st, const View& vi, const View& vo, const SweepArgs& a) {
    if (S == kExact || S == kExactFull || waves != 4) (wpb == 4 ? go_y<S, 1, 4> : go_y<S, 1, 1>)(g, st, vi, vo, a);
    else (wpb == 4 ? go_y<S, 4, 4> : go_y<S, 4, 1>)(g, st, vi, vo, a);
}

// A grid whose configured segments give fewer than min_tiles tiles (default
// 8192: eight wavefronts for each SIMD of 256 CUs, two rounds of the four
// each holds) gets shorter segments: x segments stay a multiple of 60 cells
// (one 64-lane pass with its 4-cell halo), y chunks at least 4 rows (each
// chunk reads 4 warm-up rows).  More, shorter tiles finish sooner although
// they reread more: at 4096^2 the step fell from 0.65 to 0.58 ms (2048
// tiles -> 8192), at 8192^2 nothing changes (the default shape gives 8192 y
// tiles).  A grid left below min_tiles takes one tile per workgroup.
int seg_x_for(const mh_grid_s* g, const mh_launch_shape& sh) {
    const int nx = g->d.nx, ny = g->d.ny, want_tiles = sh.min_tiles;
    int seg = sh.seg_x > 0 ? sh.seg_x : nx;
    if ((int64_t)((nx + seg - 1) / seg) * ny < want_tiles && seg > 60) {
        const int per_row = (want_tiles + ny - 1) / ny;
        const int want = (nx + per_row - 1) / per_row;
        seg = std::min(seg, std::max(60, (want + 59) / 60 * 60));
    }
    return seg;
}

int seg_y_for(const mh_grid_s* g, const mh_launch_shape& sh) {
    const int nx = g->d.nx, ny = g->d.ny, want_tiles = sh.min_tiles;
    int seg = sh.seg_y > 0 ? sh.seg_y : ny;
    const int ngroups = (nx + 63) / 64;
    if ((int64_t)ngroups * ((ny + seg - 1) / seg) < want_tiles && seg > 4) {
        const int chunks = (want_tiles + ngroups - 1) / ngroups;
        seg = std::min(seg, std::max(4, (ny + chunks - 1) / chunks));
    }
    return seg;
}

// full: the exact solver's fast pass solves every face with the full
// solver (kExactFull) instead of leaving the tiles with a non-uniform
// wavefront to the IEEE pass (an exact-solver stepper whose fast passes
// left most tiles to it).
int launch_sweep(const mh_grid_s* in, const mh_grid_s* out, int dir, int solver, DtSrc dts, const Bc& bc, int ghosts,
                 unsigned long long* speed, int clock_speed, hipStream_t st, int full = 0) {
    SweepArgs a;
    a.dts = dts;
    a.bc = bc;
    a.ghosts = ghosts;
    a.speed = speed;
    a.clock_speed = clock_speed;
    a.redo = nullptr;
    const View vi = in->view(), vo = out->view();
    const mh_launch_shape sh = shape();
    if (dir == kX) {
        a.d = in->dx;
        a.seg = seg_x_for(in, sh);
        a.ngroups = (in->d.nx + a.seg - 1) / a.seg;
        a.ntiles = a.ngroups * in->d.ny;
        const int wpb = a.ntiles < sh.min_tiles ? 1 : sh.x_wpb;
        const dim3 g((unsigned)((a.ntiles + wpb - 1) / wpb));
        const int L = sh.x_layout;
        switch (solver) {
            case kHLL: go_x_s<kHLL>(L, wpb, g, st, vi, vo, a); break;
            case kHLLC: go_x_s<kHLLC>(L, wpb, g, st, vi, vo, a); break;
            case kExact: (full ? go_x_s<kExactFull> : go_x_s<kExact>)(L, wpb, g, st, vi, vo, a); break;
            default: go_x_s<kForce>(L, wpb, g, st, vi, vo, a); break;
        }
    } else {
        a.d = in->dy;
        a.seg = seg_y_for(in, sh);
        a.ngroups = (in->d.nx + 63) / 64;
        a.ntiles = a.ngroups * ((in->d.ny + a.seg - 1) / a.seg);
        // The redo list costs a second launch: worth it only where the sweep
        // is throughput-bound (at least min_tiles tiles); small, latency-bound
        // grids take the IEEE pass in place.
        if (out->redo && a.ntiles >= sh.min_tiles && a.ntiles <= out->redo_cap) a.redo = out->redo;
        const int wpb = a.ntiles < sh.min_tiles ? 1 : sh.y_wpb;
        const dim3 g((unsigned)((a.ntiles + wpb - 1) / wpb));
        const int W = sh.y_waves;
        switch (solver) {
            case kHLL: go_y_s<kHLL>(W, wpb, g, st, vi, vo, a); break;
            case kHLLC: go_y_s<kHLLC>(W, wpb, g, st, vi, vo, a); break;
            case kExact: (full ? go_y_s<kExactFull> : go_y_s<kExact>)(W, wpb, g, st, vi, vo, a); break;
            default: go_y_s<kForce>(W, wpb, g, st, vi, vo, a); break;
        }
    }
    return check_hip(hipGetLastError(), "sweep launch");
}

// fused.hip
bool fused_applies(const mh_grid_s* g, int solver);
int fused_tiles(const mh_grid_s* g);
int launch_step_fused(const mh_grid_s* g0, const mh_grid_s* g1, int solver, DtSrc dts, const Bc& bc, int ghosts,
                      unsigned long long* speed, int clock_speed, int pp, int* redo, hipStream_t st);

int launch_fill(const mh_grid_s* g, const Bc& bc, hipStream_t st) {
    const int n = 4 * g->d.ny + 4 * (g->d.nx + 4);
    hipLaunchKernelGGL(k_fill_ghosts, dim3((n + 255) / 256), dim3(256), 0, st, g->view(), bc);
    return check_hip(hipGetLastError(), "ghost fill launch");
}

int launch_max_speed(const mh_grid_s* g, unsigned long long* slot, hipStream_t st) {
    const int64_t n = (int64_t)g->d.nx * g->d.ny;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_max_speed, dim3(blocks), dim3(256), 0, st, g->view(), slot);
    return check_hip(hipGetLastError(), "max speed launch");
}

int launch_sweep_host(const mh_grid_s* in, const mh_grid_s* out, int dir, int solver, double dt, const Bc& bc,
                      int ghosts, hipStream_t st) {
    return launch_sweep(in, out, dir, solver, DtSrc{nullptr, dt}, bc, ghosts, nullptr, 0, st);
}

// Max signal speed of a grid's interior, synchronising `st`.
int max_speed_host(const mh_grid_s* g, double* smax, hipStream_t st) {
    if (!g->aux) return fail(MH_ERR_INVALID, "max speed: grid of a batch");
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, g->aux, 1);
    int rc = launch_max_speed(g, g->aux, st);
    unsigned long long bits = 0;
    if (!rc) rc = check_hip(hipMemcpyAsync(&bits, g->aux, sizeof(bits), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    if (!rc) rc = check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (!rc) std::memcpy(smax, &bits, sizeof(double));
    return rc;
}

bool same_shape(const mh_grid_s* a, const mh_grid_s* b) {
    return a->d.nx == b->d.nx && a->d.ny == b->d.ny && a->dx == b->dx && a->dy == b->dy;
}

}  // namespace mh

using namespace mh;

struct mh_stepper_s {
    mh_grid U;
    mh_grid_s* tmp;
    int solver;
    Bc bc;
    double cfl, t_end;
    Clock* clock;      // device
    double* dt_log;    // device
    int log_cap;
    // fused: steps run as k_step_fused, alternating the state between U and
    // tmp (the clock's cur); dirty: steps were enqueued since the host last
    // moved the state back into U (stepper_settle)
    int fused, dirty;
    // adaptive (the exact solver under the default shape): at every clock
    // read the stepper takes the fused step or the two sweeps by the share
    // of tiles the fast passes left to the redo lists since the last read
    // (from mark_step, mark_redo); full: its two sweeps solve every face in
    // the fast pass (launch_sweep)
    int adaptive, mark_step, mark_redo, full;
};

namespace mh {
// (Re)starts a stepper at t = 0 from the current state of its grid: fills
// the ghosts and reduces the state's max signal speed into the clock.
// Enqueues only.  Used by mh_stepper_create and the adjoint's forward run.
int stepper_start(mh_stepper s, double cfl, hipStream_t st) {
    s->cfl = cfl;
    s->mark_step = 0;
    s->mark_redo = 0;
    hipLaunchKernelGGL(k_clock_init, dim3(1), dim3(64), 0, st, s->clock, s->t_end, cfl,
                       std::min(s->U->dx, s->U->dy), s->log_cap, s->fused);
    s->dirty = 0;
    int rc = check_hip(hipGetLastError(), "clock launch");
    if (!rc) rc = launch_fill(s->U, s->bc, st);
    if (!rc) rc = launch_max_speed(s->U, &s->clock->smax[0], st);
    return rc;
}

// Steps of a stepper: the fused step where it applies, else the x sweep
// U -> tmp and the y sweep tmp -> U.
int stepper_step(mh_stepper s, hipStream_t st) {
    const DtSrc dts{s->clock, 0.0};
    s->dirty = 1;
    if (s->fused) return launch_step_fused(s->U, s->tmp, s->solver, dts, s->bc, 1, nullptr, 1, 1, s->tmp->redo, st);
    int rc = launch_sweep(s->U, s->tmp, kX, s->solver, dts, s->bc, 1, nullptr, 0, st, s->full);
    // The y sweep reduces the new state's max signal speed into the
    // clock's inactive slot, which k_clock has just cleared.
    if (!rc) rc = launch_sweep(s->tmp, s->U, kY, s->solver, dts, s->bc, 1, nullptr, 1, st, s->full);
    return rc;
}

// After a fused run whose committed steps left the state in tmp (clock
// cur = 1, read by the caller from a synchronised copy c), swaps the two
// grids' storage so that the stepper's grid holds it again.
int stepper_settle(mh_stepper s, const Clock& c, hipStream_t st) {
    s->dirty = 0;
    if (!s->fused || !c.cur) return MH_OK;
    std::swap(s->U->base, s->tmp->base);
    hipLaunchKernelGGL(k_clock_cur0, dim3(1), dim3(64), 0, st, s->clock);
    return check_hip(hipGetLastError(), "clock launch");
}

// The exact solver's fast passes solve the wavefronts whose faces join
// equal states and leave every tile with another face to an IEEE redo
// pass: the fused step's (a pair of 217 VGPRs) or the y sweep's.  Where few
// tiles are left (the cylindrical blast at 8192^2: about 2 % of them) the
// fused step is faster (2.00 against 2.46 ms per step), where all are (a
// state with no two equal neighbours) the two sweeps are (6.9 against 7.9
// ms, with the full solver in their fast pass: no redo pass, whose IEEE
// division everywhere else costs more).  At a clock read after at least 4
// steps, a fused stepper that redid more than 40 % of its tiles since the
// last read changes to the two sweeps with the full solver, and stays
// there (they leave no redo count to measure the share by).  The state is
// in the stepper's grid (stepper_settle ran).
int stepper_adapt(mh_stepper s, const Clock& c, hipStream_t st) {
    const int steps = c.step - s->mark_step;
    if (!s->fused || steps < 4) return MH_OK;
    const double share = (double)(c.redo_tiles - s->mark_redo) / ((double)steps * fused_tiles(s->U));
    s->mark_step = c.step;
    s->mark_redo = c.redo_tiles;
    if (share <= 0.4) return MH_OK;
    s->fused = 0;
    s->full = 1;
    hipLaunchKernelGGL(k_clock_pp, dim3(1), dim3(64), 0, st, s->clock, 0);
    return check_hip(hipGetLastError(), "clock launch");
}

// Reads the clock (synchronising `st`), settles the state and, for an
// adaptive stepper, chooses the step's form for the next steps.
int stepper_sync(mh_stepper s, hipStream_t st, Clock* out) {
    Clock c;
    hipLaunchKernelGGL(k_clock, dim3(1), dim3(64), 0, st, s->clock, s->dt_log, 1, nullptr);
    int rc = check_hip(hipGetLastError(), "clock launch");
    if (!rc) rc = check_hip(hipMemcpyAsync(&c, s->clock, sizeof(c), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    if (!rc) rc = check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (!rc) rc = stepper_settle(s, c, st);
    if (!rc && s->adaptive) rc = stepper_adapt(s, c, st);
    if (!rc && out) *out = c;
    return rc;
}

// The adjoint's forward run copies checkpoints out of the stepper's grid
// between steps, so its stepper keeps the two sweeps.
void stepper_unfuse(mh_stepper s) {
    s->fused = 0;
    s->adaptive = 0;
}
}  // namespace mh

extern "C" {

const char* mh_last_error(void) { return g_err.c_str(); }
const char* mh_version(void) { return "mh_euler 1.0 gfx950"; }

int mh_set_device(int device) { return check_hip(hipSetDevice(device), "hipSetDevice"); }

int mh_grid_create(const mh_grid_desc* desc, mh_grid* out) {
    if (!desc || !out) return fail(MH_ERR_INVALID, "mh_grid_create: null argument");
    if (desc->nx < 1 || desc->ny < 1 || !(desc->x_max > desc->x_min) || !(desc->y_max > desc->y_min))
        return fail(MH_ERR_INVALID, "mh_grid_create: need nx, ny >= 1 and a non-empty domain");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(MH_ERR_NODEVICE, "no HIP device");
    mh_grid g = new (std::nothrow) mh_grid_s();
    if (!g) return fail(MH_ERR_OOM, "host allocation failed");
    g->d = *desc;
    g->owns = 1;
    g->dx = (desc->x_max - desc->x_min) / (double)desc->nx;   // head/gpu/grid_gpu.cuh:40-41
    g->dy = (desc->y_max - desc->y_min) / (double)desc->ny;
    g->rp = round_up((int64_t)kLead + desc->nx + kNg, 16);
    g->ps = round_up(g->rp * (desc->ny + 2 * kNg), 256);
    (void)hipGetDevice(&g->device);
    const hipError_t e = hipMalloc(&g->base, (size_t)(4 * g->ps) * sizeof(double));
    if (e != hipSuccess) {
        delete g;
        return check_hip(e, "hipMalloc(grid)");
    }
    const hipError_t e2 = hipMalloc((void**)&g->aux, 8 * sizeof(unsigned long long));
    if (e2 != hipSuccess) {
        (void)hipFree(g->base);
        delete g;
        return check_hip(e2, "hipMalloc(grid aux)");
    }
    // y-sweep redo list: room for every tile of the finest chunking (1 row)
    g->redo_cap = (int64_t)((desc->nx + 63) / 64) * desc->ny;
    hipError_t e3 = hipMalloc((void**)&g->redo, sizeof(int) * (size_t)(2 + g->redo_cap));
    if (e3 == hipSuccess) e3 = hipMemset(g->redo, 0, 2 * sizeof(int));
    if (e3 != hipSuccess) {
        (void)hipFree(g->base);
        (void)hipFree(g->aux);
        if (g->redo) (void)hipFree(g->redo);
        delete g;
        return check_hip(e3, "hipMalloc(grid redo list)");
    }
    *out = g;
    return MH_OK;
}

int mh_grid_destroy(mh_grid g) {
    if (!g) return MH_OK;
    if (!g->owns) return fail(MH_ERR_INVALID, "mh_grid_destroy: grid belongs to a batch");
    hipError_t e = hipFree(g->base);
    const hipError_t e2 = hipFree(g->aux);
    const hipError_t e3 = g->redo ? hipFree(g->redo) : hipSuccess;
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = e3;
    delete g;
    return check_hip(e, "hipFree(grid)");
}

int mh_grid_info(mh_grid g, mh_grid_desc* desc, double* dx, double* dy, int64_t* pitch, int64_t* plane_stride) {
    if (!g) return fail(MH_ERR_INVALID, "mh_grid_info: null grid");
    if (desc) *desc = g->d;
    if (dx) *dx = g->dx;
    if (dy) *dy = g->dy;
    if (pitch) *pitch = g->rp;
    if (plane_stride) *plane_stride = g->ps;
    return MH_OK;
}

int mh_grid_plane_ptr(mh_grid g, int c, double** ptr) {
    if (!g || c < 0 || c > 3 || !ptr) return fail(MH_ERR_INVALID, "mh_grid_plane_ptr: bad argument");
    *ptr = g->view().o + (int64_t)c * g->ps;
    return MH_OK;
}

static int copy_planes(mh_grid g, double* const host[4], bool upload) {
    const size_t w = (size_t)(g->d.nx + 4), h = (size_t)(g->d.ny + 4);
    for (int c = 0; c < 4; ++c) {
        double* dev = g->base + (int64_t)c * g->ps + (kLead - kNg);
        const hipError_t e = upload ? hipMemcpy2D(dev, g->rp * sizeof(double), host[c], w * sizeof(double),
                                                  w * sizeof(double), h, hipMemcpyHostToDevice)
                                    : hipMemcpy2D(host[c], w * sizeof(double), dev, g->rp * sizeof(double),
                                                  w * sizeof(double), h, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return check_hip(e, "hipMemcpy2D");
    }
    return MH_OK;
}

int mh_grid_upload_soa(mh_grid g, const double* const planes[4]) {
    if (!g || !planes) return fail(MH_ERR_INVALID, "mh_grid_upload_soa: null argument");
    double* p[4] = {const_cast<double*>(planes[0]), const_cast<double*>(planes[1]), const_cast<double*>(planes[2]),
                    const_cast<double*>(planes[3])};
    return copy_planes(g, p, true);
}

int mh_grid_download_soa(mh_grid g, double* const planes[4]) {
    if (!g || !planes) return fail(MH_ERR_INVALID, "mh_grid_download_soa: null argument");
    double* p[4] = {planes[0], planes[1], planes[2], planes[3]};
    return copy_planes(g, p, false);
}

int mh_grid_upload_aos(mh_grid g, const double* aos) {
    if (!g || !aos) return fail(MH_ERR_INVALID, "mh_grid_upload_aos: null argument");
    const size_t n = (size_t)(g->d.nx + 4) * (g->d.ny + 4);
    std::vector<double> soa(4 * n);
    for (size_t k = 0; k < n; ++k)
        for (int c = 0; c < 4; ++c) soa[c * n + k] = aos[4 * k + c];
    double* p[4] = {soa.data(), soa.data() + n, soa.data() + 2 * n, soa.data() + 3 * n};
    return copy_planes(g, p, true);
}

int mh_grid_download_aos(mh_grid g, double* aos) {
    if (!g || !aos) return fail(MH_ERR_INVALID, "mh_grid_download_aos: null argument");
    const size_t n = (size_t)(g->d.nx + 4) * (g->d.ny + 4);
    std::vector<double> soa(4 * n);
    double* p[4] = {soa.data(), soa.data() + n, soa.data() + 2 * n, soa.data() + 3 * n};
    const int rc = copy_planes(g, p, false);
    if (rc) return rc;
    for (size_t k = 0; k < n; ++k)
        for (int c = 0; c < 4; ++c) aos[4 * k + c] = soa[c * n + k];
    return MH_OK;
}

static int copy_interior_dev(mh_grid g, double* dense, bool to_grid, hipStream_t st) {
    const size_t nx = (size_t)g->d.nx, ny = (size_t)g->d.ny;
    for (int c = 0; c < 4; ++c) {
        double* dev = g->view().o + (int64_t)c * g->ps;
        double* d = dense + c * nx * ny;
        const hipError_t e = to_grid ? hipMemcpy2DAsync(dev, g->rp * sizeof(double), d, nx * sizeof(double),
                                                        nx * sizeof(double), ny, hipMemcpyDeviceToDevice, st)
                                     : hipMemcpy2DAsync(d, nx * sizeof(double), dev, g->rp * sizeof(double),
                                                        nx * sizeof(double), ny, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return check_hip(e, "hipMemcpy2DAsync");
    }
    return MH_OK;
}

int mh_grid_copy_interior_from_device(mh_grid g, const double* src, mh_stream stream) {
    if (!g || !src) return fail(MH_ERR_INVALID, "mh_grid_copy_interior_from_device: null argument");
    return copy_interior_dev(g, const_cast<double*>(src), true, (hipStream_t)stream);
}

int mh_grid_copy_interior_to_device(mh_grid g, double* dst, mh_stream stream) {
    if (!g || !dst) return fail(MH_ERR_INVALID, "mh_grid_copy_interior_to_device: null argument");
    return copy_interior_dev(g, dst, false, (hipStream_t)stream);
}

int mh_grid_swap(mh_grid a, mh_grid b) {
    if (!a || !b) return fail(MH_ERR_INVALID, "mh_grid_swap: null grid");
    if (a->d.nx != b->d.nx || a->d.ny != b->d.ny) return fail(MH_ERR_INVALID, "mh_grid_swap: shape mismatch");
    std::swap(a->base, b->base);
    return MH_OK;
}

int mh_fill_ghosts(mh_grid g, const mh_bc* bc, mh_stream stream) {
    if (!g) return fail(MH_ERR_INVALID, "mh_fill_ghosts: null grid");
    if (!valid_bc(bc)) return fail(MH_ERR_INVALID, "mh_fill_ghosts: unknown boundary kind");
    return launch_fill(g, to_bc(bc), (hipStream_t)stream);
}

int mh_compute_dt(mh_grid g, double cfl, double* dt, double* max_speed, mh_stream stream) {
    if (!g) return fail(MH_ERR_INVALID, "mh_compute_dt: null grid");
    if (!g->aux) return fail(MH_ERR_INVALID, "mh_compute_dt: grid of a batch (use mh_batch_clock)");
    hipStream_t st = (hipStream_t)stream;
    unsigned long long* slot = g->aux;
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, slot, 1);
    int rc = launch_max_speed(g, slot, st);
    unsigned long long bits = 0;
    if (!rc) rc = check_hip(hipMemcpyAsync(&bits, slot, sizeof(bits), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    if (!rc) rc = check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (rc) return rc;
    double m;
    std::memcpy(&m, &bits, sizeof(m));
    if (max_speed) *max_speed = m;
    if (!(m > 0.0)) return fail(MH_ERR_WAVESPEED, "compute_dt: non-positive maximum wave speed");
    if (dt) *dt = cfl * std::min(g->dx, g->dy) / m;   // src/cpu/solver_cpu.cpp:308
    return MH_OK;
}

int mh_sweep(mh_grid in, mh_grid out, double dt, int dir, int solver, mh_stream stream) {
    if (!in || !out || in == out) return fail(MH_ERR_INVALID, "mh_sweep: need distinct grids");
    if (!same_shape(in, out)) return fail(MH_ERR_INVALID, "mh_sweep: shape mismatch");
    if (dir != 0 && dir != 1) return fail(MH_ERR_INVALID, "mh_sweep: dir must be 0 or 1");
    if (solver < 0 || solver > 3) return fail(MH_ERR_INVALID, "mh_sweep: unknown solver");
    return launch_sweep(in, out, dir, solver, DtSrc{nullptr, dt}, to_bc(nullptr), 0, nullptr, 0, (hipStream_t)stream);
}

int mh_advance_second_order(mh_grid Uold, mh_grid Utmp, mh_grid Unew, double dt, int solver, const mh_bc* bc,
                            mh_stream stream) {
    if (!Uold || !Utmp || !Unew) return fail(MH_ERR_INVALID, "mh_advance_second_order: null grid");
    if (Utmp == Uold || Utmp == Unew) return fail(MH_ERR_INVALID, "mh_advance_second_order: Utmp must be distinct");
    if (!same_shape(Uold, Utmp) || !same_shape(Uold, Unew))
        return fail(MH_ERR_INVALID, "mh_advance_second_order: shape mismatch");
    if (solver < 0 || solver > 3) return fail(MH_ERR_INVALID, "mh_advance_second_order: unknown solver");
    if (!valid_bc(bc)) return fail(MH_ERR_INVALID, "mh_advance_second_order: unknown boundary kind");
    hipStream_t st = (hipStream_t)stream;
    const Bc b = to_bc(bc);
    int rc = launch_sweep(Uold, Utmp, kX, solver, DtSrc{nullptr, dt}, b, 1, nullptr, 0, st);
    if (!rc) rc = launch_sweep(Utmp, Unew, kY, solver, DtSrc{nullptr, dt}, b, 1, nullptr, 0, st);
    if (!rc) rc = launch_fill(Utmp, b, st);
    if (!rc) rc = launch_fill(Unew, b, st);
    return rc;
}

int mh_advance_first_order(mh_grid Uold, mh_grid Unew, double dt, int solver, const mh_bc* bc, mh_stream stream) {
    if (!Uold || !Unew || Uold == Unew) return fail(MH_ERR_INVALID, "mh_advance_first_order: need distinct grids");
    if (!same_shape(Uold, Unew)) return fail(MH_ERR_INVALID, "mh_advance_first_order: shape mismatch");
    if (solver < 0 || solver > 3) return fail(MH_ERR_INVALID, "mh_advance_first_order: unknown solver");
    if (!valid_bc(bc)) return fail(MH_ERR_INVALID, "mh_advance_first_order: unknown boundary kind");
    hipStream_t st = (hipStream_t)stream;
    const double dtdx = dt / Uold->dx, dtdy = dt / Uold->dy;
    const dim3 grid((Uold->d.nx + 63) / 64, (Uold->d.ny + 3) / 4);
    const View vi = Uold->view(), vo = Unew->view();
    switch (solver) {
        case kHLL: hipLaunchKernelGGL(k_first_order<kHLL>, grid, dim3(256), 0, st, vi, vo, dtdx, dtdy); break;
        case kHLLC: hipLaunchKernelGGL(k_first_order<kHLLC>, grid, dim3(256), 0, st, vi, vo, dtdx, dtdy); break;
        case kExact: hipLaunchKernelGGL(k_first_order<kExact>, grid, dim3(256), 0, st, vi, vo, dtdx, dtdy); break;
        default: hipLaunchKernelGGL(k_first_order<kForce>, grid, dim3(256), 0, st, vi, vo, dtdx, dtdy); break;
    }
    int rc = check_hip(hipGetLastError(), "first-order launch");
    if (!rc) rc = launch_fill(Unew, to_bc(bc), st);
    return rc;
}

int mh_stepper_create(mh_grid U, int solver, double cfl, double t_end, const mh_bc* bc, int dt_log_capacity,
                      mh_stepper* out, mh_stream stream) {
    if (!U || !out) return fail(MH_ERR_INVALID, "mh_stepper_create: null argument");
    if (solver < 0 || solver > 3) return fail(MH_ERR_INVALID, "mh_stepper_create: unknown solver");
    if (!valid_bc(bc)) return fail(MH_ERR_INVALID, "mh_stepper_create: unknown boundary kind");
    if (!(cfl > 0.0)) return fail(MH_ERR_INVALID, "mh_stepper_create: cfl must be positive");
    hipStream_t st = (hipStream_t)stream;
    mh_stepper s = new (std::nothrow) mh_stepper_s();
    if (!s) return fail(MH_ERR_OOM, "host allocation failed");
    s->U = U;
    s->solver = solver;
    s->bc = to_bc(bc);
    s->t_end = t_end;
    s->log_cap = dt_log_capacity > 0 ? dt_log_capacity : 0;
    int rc = mh_grid_create(&U->d, &s->tmp);
    if (rc) { delete s; return rc; }
    // fused = 1 takes the fused step from kFusedMinCells cells on: smaller
    // grids are latency-bound, and each wavefront pair's chain of rows is
    // longer than the two sweeps' (256^2: 57 against 38 us per step; the
    // fused step is ahead from 1024^2 on: 73 against 81 us).  The exact
    // solver's fast pass solves the uniform wavefronts and its redo pass
    // (a pair of 217 VGPRs, two wavefronts per SIMD) every tile with a
    // non-uniform face; on the cylindrical blast that is ahead of the two
    // sweeps from 4096^2 on (0.749 against 0.791 ms per step; 8192^2: 2.00
    // against 2.46; 16384^2: 6.71 against 8.79) and behind below (2048^2:
    // 0.404 against 0.342), so it starts at 2^24 cells.  On a state with no
    // uniform region every tile is redone, and the two sweeps stay ahead
    // (8192^2: 7.93 against 6.90 ms).
    constexpr int64_t kFusedMinCells = 1 << 20, kFusedMinCellsExact = int64_t(1) << 24;
    const int want = shape().fused;
    const int64_t min_cells = solver == kExact ? kFusedMinCellsExact : kFusedMinCells;
    s->fused = (want == 2 || (want == 1 && (int64_t)U->d.nx * U->d.ny >= min_cells)) &&
               fused_applies(U, solver) && fused_tiles(U) <= s->tmp->redo_cap;
    s->adaptive = want == 1 && solver == kExact && s->fused;
    s->dirty = 0;
    rc = check_hip(hipMalloc((void**)&s->clock, sizeof(Clock)), "hipMalloc(clock)");
    s->dt_log = nullptr;
    if (!rc && s->log_cap) rc = check_hip(hipMalloc((void**)&s->dt_log, sizeof(double) * s->log_cap), "hipMalloc(dt log)");
    if (!rc) rc = stepper_start(s, cfl, st);
    if (!rc) rc = check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (rc) { mh_stepper_destroy(s); return rc; }
    *out = s;
    return MH_OK;
}

int mh_stepper_destroy(mh_stepper s) {
    if (!s) return MH_OK;
    // the state goes back into the caller's grid before tmp is freed
    if (s->dirty && s->clock && s->tmp) (void)stepper_sync(s, nullptr, nullptr);
    if (s->tmp) mh_grid_destroy(s->tmp);
    if (s->clock) (void)hipFree(s->clock);
    if (s->dt_log) (void)hipFree(s->dt_log);
    delete s;
    return MH_OK;
}

int mh_stepper_advance(mh_stepper s, int nsteps, mh_stream stream) {
    if (!s || nsteps < 0) return fail(MH_ERR_INVALID, "mh_stepper_advance: bad argument");
    hipStream_t st = (hipStream_t)stream;
    for (int k = 0; k < nsteps; ++k) {
        hipLaunchKernelGGL(k_clock, dim3(1), dim3(64), 0, st, s->clock, s->dt_log, 0, s->fused ? s->tmp->redo : nullptr);
        int rc = check_hip(hipGetLastError(), "clock launch");
        if (!rc) rc = stepper_step(s, st);
        if (rc) return rc;
    }
    return MH_OK;
}

int mh_stepper_clock(mh_stepper s, double* t, int* steps, double* dt_log, mh_stream stream) {
    if (!s) return fail(MH_ERR_INVALID, "mh_stepper_clock: null stepper");
    hipStream_t st = (hipStream_t)stream;
    Clock c;
    // c.step sizes the log copy: the clock is read (and the state settled
    // into the stepper's grid) first
    int rc = stepper_sync(s, st, &c);
    if (!rc && dt_log && s->log_cap && c.step > 0) {
        rc = check_hip(hipMemcpyAsync(dt_log, s->dt_log, sizeof(double) * std::min(s->log_cap, c.step),
                                      hipMemcpyDeviceToHost, st),
                       "hipMemcpyAsync");
        if (!rc) rc = check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    }
    if (rc) return rc;
    if (t) *t = c.t;
    if (steps) *steps = c.step;
    if (c.status) return fail(c.status, "stepper: non-positive maximum wave speed");
    return MH_OK;
}

int mh_stepper_set_stop(mh_stepper s, double t_stop, mh_stream stream) {
    if (!s) return fail(MH_ERR_INVALID, "mh_stepper_set_stop: null stepper");
    if (!(t_stop >= 0.0)) return fail(MH_ERR_INVALID, "mh_stepper_set_stop: t_stop must be >= 0");
    hipLaunchKernelGGL(k_clock_stop, dim3(1), dim3(64), 0, (hipStream_t)stream, s->clock, t_stop);
    return check_hip(hipGetLastError(), "clock launch");
}

int mh_stepper_finish(mh_stepper s, mh_stream stream) {
    if (!s) return fail(MH_ERR_INVALID, "mh_stepper_finish: null stepper");
    if (s->dirty) {
        const int rc = stepper_sync(s, (hipStream_t)stream, nullptr);
        if (rc) return rc;
    }
    return launch_fill(s->U, s->bc, (hipStream_t)stream);
}

int mh_init_state(mh_grid g, int kind, const double* p, mh_stream stream) {
    if (!g) return fail(MH_ERR_INVALID, "mh_init_state: null grid");
    InitArgs a{};
    a.xmin = g->d.x_min; a.ymin = g->d.y_min; a.dx = g->dx; a.dy = g->dy;
    double q[8];
    if (kind == 0) {            // reference shock_bubble (src/test_cases.cpp:24-69)
        const double r[8] = {1.29, 1.01325e5, 1.22, 0.005, 0.214, 0.035, 0.0445, 0.025};
        std::memcpy(q, r, sizeof q);
        kind = 3;
    } else if (kind == 1) {     // reference blast_wave (src/test_cases.cpp:78-101)
        const double r[7] = {0.5, 0.5, 0.1, 1.0, 100.0, 1.0, 1.0};
        std::memcpy(q, r, 7 * sizeof(double));
        kind = 2;
    } else if (kind == 2 || kind == 3) {
        if (!p) return fail(MH_ERR_INVALID, "mh_init_state: params required");
        std::memcpy(q, p, (kind == 2 ? 7 : 8) * sizeof(double));
    } else {
        return fail(MH_ERR_INVALID, "mh_init_state: unknown kind");
    }
    if (kind == 2) {
        a.kind = 0;
        a.x0 = q[0]; a.y0 = q[1]; a.r0 = q[2];
        a.in = {q[3], 0.0, 0.0, q[4]};
        a.out = {q[5], 0.0, 0.0, q[6]};
    } else {
        a.kind = 1;
        // post-shock state from Rankine-Hugoniot (src/test_cases.cpp:35-47)
        const double rho0 = q[0], p0 = q[1], mach = q[2];
        const double a0 = std::sqrt(kGamma * p0 / rho0);
        const double M2 = mach * mach;
        const double rho2 = rho0 * ((kGamma + 1.0) * M2) / ((kGamma - 1.0) * M2 + 2.0);
        const double p2 = p0 * (1.0 + 2.0 * kGamma / (kGamma + 1.0) * (M2 - 1.0));
        const double u2 = 2.0 * a0 / (kGamma + 1.0) * (mach - 1.0 / mach);
        a.in = {rho0, 0.0, 0.0, p0};
        a.out = {rho2, u2, 0.0, p2};
        a.shock_x = q[3];
        a.bubble_rho = q[4];
        a.x0 = q[5]; a.y0 = q[6]; a.r0 = q[7];
    }
    const int64_t n = (int64_t)(g->d.nx + 4) * (g->d.ny + 4);
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_init, dim3(blocks), dim3(256), 0, (hipStream_t)stream, g->view(), a);
    return check_hip(hipGetLastError(), "init launch");
}

// Device scratch for the pointwise entry points: freed before returning.
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

int mh_fast_ops(int n, const double* a, const double* b, const double* U, int fast, double* q, double* r, double* s) {
    if (n < 0 || (n > 0 && (!a || !b || !U || !q || !r || !s)) || fast < 0 || fast > 2)
        return fail(MH_ERR_INVALID, "mh_fast_ops: bad argument");
    if (n == 0) return MH_OK;
    const size_t nb = sizeof(double) * (size_t)n;
    DevBuf buf;
    int rc = check_hip(hipMalloc(&buf.p, 10 * nb), "hipMalloc");
    if (rc) return rc;
    double* da = (double*)buf.p;
    double* db = da + n;
    double* dU = db + n;
    double* dq = dU + 4 * (size_t)n;
    double* dr = dq + n;
    double* ds = dr + n;
    rc = check_hip(hipMemcpy(da, a, nb, hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc) rc = check_hip(hipMemcpy(db, b, nb, hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc) rc = check_hip(hipMemcpy(dU, U, 4 * nb, hipMemcpyHostToDevice), "hipMemcpy");
    if (rc) return rc;
    hipLaunchKernelGGL(k_fast_ops, dim3((n + 255) / 256), dim3(256), 0, 0, n, da, db, dU, fast, dq, dr, ds);
    rc = check_hip(hipGetLastError(), "fast ops launch");
    if (!rc) rc = check_hip(hipMemcpy(q, dq, nb, hipMemcpyDeviceToHost), "hipMemcpy");
    if (!rc) rc = check_hip(hipMemcpy(r, dr, nb, hipMemcpyDeviceToHost), "hipMemcpy");
    if (!rc) rc = check_hip(hipMemcpy(s, ds, nb, hipMemcpyDeviceToHost), "hipMemcpy");
    return rc;
}

int mh_riemann_flux(int n, const double* UL, const double* UR, int dir, int solver, double* F, int* iters) {
    if (n < 0 || (n > 0 && (!UL || !UR || !F))) return fail(MH_ERR_INVALID, "mh_riemann_flux: bad argument");
    if (dir != 0 && dir != 1) return fail(MH_ERR_INVALID, "mh_riemann_flux: dir must be 0 or 1");
    if (solver < 0 || solver > 3) return fail(MH_ERR_INVALID, "mh_riemann_flux: unknown solver");
    if (n == 0) return MH_OK;
    const size_t bytes = sizeof(double) * 4 * (size_t)n;
    DevBuf b;
    int rc = check_hip(hipMalloc(&b.p, 3 * bytes + sizeof(int) * (size_t)n), "hipMalloc");
    if (rc) return rc;
    double* dL = (double*)b.p;
    double* dR = dL + 4 * (size_t)n;
    double* dF = dR + 4 * (size_t)n;
    int* dI = (int*)(dF + 4 * (size_t)n);
    rc = check_hip(hipMemcpy(dL, UL, bytes, hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc) rc = check_hip(hipMemcpy(dR, UR, bytes, hipMemcpyHostToDevice), "hipMemcpy");
    if (rc) return rc;
    const dim3 g((n + 255) / 256), blk(256);
#define MH_FF(S_, D_) hipLaunchKernelGGL((k_face_flux<S_, D_>), g, blk, 0, 0, n, dL, dR, dF, dI)
    if (dir == 0) {
        switch (solver) { case 0: MH_FF(kHLL, kX); break; case 1: MH_FF(kHLLC, kX); break;
                          case 2: MH_FF(kExact, kX); break; default: MH_FF(kForce, kX); }
    } else {
        switch (solver) { case 0: MH_FF(kHLL, kY); break; case 1: MH_FF(kHLLC, kY); break;
                          case 2: MH_FF(kExact, kY); break; default: MH_FF(kForce, kY); }
    }
#undef MH_FF
    rc = check_hip(hipGetLastError(), "face flux launch");
    if (!rc) rc = check_hip(hipMemcpy(F, dF, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
    if (!rc && iters) rc = check_hip(hipMemcpy(iters, dI, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost), "hipMemcpy");
    return rc;
}

// ---- mh_riemann_flux_noalloc: the same evaluation staged through a fixed
// device array instead of per-call allocations ------------------------------
}  // extern "C"

namespace mh {
namespace {
constexpr int kNoAllocMaxFaces = MH_NOALLOC_MAX_FACES;
// Staging for mh_riemann_flux_noalloc: UL, UR, F (4 doubles per face each),
// then the Newton iteration counts (ints), for at most kNoAllocMaxFaces
// faces: 12 * 2^16 doubles + 2^16 ints = 6.5 MB per device, held until the
// process ends.  In host code `scratch` names a host-side stand-in; its
// device address comes from hipGetSymbolAddress, per device.
static __device__ __align__(256) double scratch[12 * kNoAllocMaxFaces + kNoAllocMaxFaces / 2];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_lock;
void* g_scratch_ptr[kMaxDevices] = {};

// The calling thread's device's copy of `scratch`, looked up once per device.
int scratch_ptr(double** out) {
    int dev = 0;
    int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= kMaxDevices) return fail(MH_ERR_INVALID, "mh_riemann_flux_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_lock);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (rc) return rc;
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<double*>(g_scratch_ptr[dev]);
    return MH_OK;
}
}  // namespace
}  // namespace mh

extern "C" {

int mh_riemann_flux_noalloc(int n, const double* UL, const double* UR, int dir, int solver, double* F, int* iters) {
    if (n < 0 || (n > 0 && (!UL || !UR || !F))) return fail(MH_ERR_INVALID, "mh_riemann_flux_noalloc: bad argument");
    if (n > kNoAllocMaxFaces)
        return fail(MH_ERR_INVALID, "mh_riemann_flux_noalloc: n exceeds MH_NOALLOC_MAX_FACES (" +
                                        std::to_string(kNoAllocMaxFaces) + "); use mh_riemann_flux");
    if (dir != 0 && dir != 1) return fail(MH_ERR_INVALID, "mh_riemann_flux_noalloc: dir must be 0 or 1");
    if (solver < 0 || solver > 3) return fail(MH_ERR_INVALID, "mh_riemann_flux_noalloc: unknown solver");
    if (n == 0) return MH_OK;
    double* base = nullptr;
    int rc = scratch_ptr(&base);
    if (rc) return rc;
    // Every face slot below n is written by the kernel before it is read
    // back, so nothing needs clearing between calls.
    double* dL = base;
    double* dR = base + 4 * (size_t)kNoAllocMaxFaces;
    double* dF = base + 8 * (size_t)kNoAllocMaxFaces;
    int* dI = reinterpret_cast<int*>(base + 12 * (size_t)kNoAllocMaxFaces);
    const size_t bytes = sizeof(double) * 4 * (size_t)n;
    rc = check_hip(hipMemcpy(dL, UL, bytes, hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc) rc = check_hip(hipMemcpy(dR, UR, bytes, hipMemcpyHostToDevice), "hipMemcpy");
    if (rc) return rc;
    const dim3 g((n + 255) / 256), blk(256);
#define MH_FF(S_, D_) hipLaunchKernelGGL((k_face_flux<S_, D_>), g, blk, 0, 0, n, dL, dR, dF, dI)
    if (dir == 0) {
        switch (solver) { case 0: MH_FF(kHLL, kX); break; case 1: MH_FF(kHLLC, kX); break;
                          case 2: MH_FF(kExact, kX); break; default: MH_FF(kForce, kX); }
    } else {
        switch (solver) { case 0: MH_FF(kHLL, kY); break; case 1: MH_FF(kHLLC, kY); break;
                          case 2: MH_FF(kExact, kY); break; default: MH_FF(kForce, kY); }
    }
#undef MH_FF
    rc = check_hip(hipGetLastError(), "face flux launch");
    if (!rc) rc = check_hip(hipMemcpy(F, dF, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
    if (!rc && iters) rc = check_hip(hipMemcpy(iters, dI, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost), "hipMemcpy");
    return rc;
}

int mh_reconstruct(int n, const double* Um, const double* Uc, const double* Up, double dt_over_d, int dir, double* UL,
                   double* UR) {
    if (n < 0 || (n > 0 && (!Um || !Uc || !Up || !UL || !UR))) return fail(MH_ERR_INVALID, "mh_reconstruct: bad argument");
    if (dir != 0 && dir != 1) return fail(MH_ERR_INVALID, "mh_reconstruct: dir must be 0 or 1");
    if (n == 0) return MH_OK;
    const size_t bytes = sizeof(double) * 4 * (size_t)n;
    DevBuf b;
    int rc = check_hip(hipMalloc(&b.p, 5 * bytes), "hipMalloc");
    if (rc) return rc;
    double* d = (double*)b.p;
    const size_t q = 4 * (size_t)n;
    rc = check_hip(hipMemcpy(d, Um, bytes, hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc) rc = check_hip(hipMemcpy(d + q, Uc, bytes, hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc) rc = check_hip(hipMemcpy(d + 2 * q, Up, bytes, hipMemcpyHostToDevice), "hipMemcpy");
    if (rc) return rc;
    const double half = 0.5 * dt_over_d;
    const dim3 g((n + 255) / 256), blk(256);
    if (dir == 0) hipLaunchKernelGGL(k_recon<kX>, g, blk, 0, 0, n, d, d + q, d + 2 * q, half, d + 3 * q, d + 4 * q);
    else hipLaunchKernelGGL(k_recon<kY>, g, blk, 0, 0, n, d, d + q, d + 2 * q, half, d + 3 * q, d + 4 * q);
    rc = check_hip(hipGetLastError(), "reconstruct launch");
    if (!rc) rc = check_hip(hipMemcpy(UL, d + 3 * q, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
    if (!rc) rc = check_hip(hipMemcpy(UR, d + 4 * q, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
    return rc;
}

int mh_stepper_advance_profiled(mh_stepper s, int nsteps, mh_stream stream, double* ms_total, double* ms_kernels) {
    if (!s || nsteps < 1) return fail(MH_ERR_INVALID, "mh_stepper_advance_profiled: bad argument");
    hipStream_t st = (hipStream_t)stream;
    const DtSrc dts{s->clock, 0.0};
    // events: [k][0..3] = before clock, before x, before y, after y (fused:
    // before the step, after it, after it)
    std::vector<hipEvent_t> ev((size_t)nsteps * 4);
    int rc = MH_OK;
    for (auto& e : ev)
        if (!rc) rc = check_hip(hipEventCreate(&e), "hipEventCreate");
    for (int k = 0; k < nsteps && !rc; ++k) {
        hipEvent_t* e = &ev[(size_t)k * 4];
        rc = check_hip(hipEventRecord(e[0], st), "hipEventRecord");
        hipLaunchKernelGGL(k_clock, dim3(1), dim3(64), 0, st, s->clock, s->dt_log, 0, s->fused ? s->tmp->redo : nullptr);
        if (!rc) rc = check_hip(hipGetLastError(), "clock launch");
        if (!rc) rc = check_hip(hipEventRecord(e[1], st), "hipEventRecord");
        if (s->fused) {
            if (!rc) rc = stepper_step(s, st);
            if (!rc) rc = check_hip(hipEventRecord(e[2], st), "hipEventRecord");
        } else {
            s->dirty = 1;
            if (!rc) rc = launch_sweep(s->U, s->tmp, kX, s->solver, dts, s->bc, 1, nullptr, 0, st, s->full);
            if (!rc) rc = check_hip(hipEventRecord(e[2], st), "hipEventRecord");
            if (!rc) rc = launch_sweep(s->tmp, s->U, kY, s->solver, dts, s->bc, 1, nullptr, 1, st, s->full);
        }
        if (!rc) rc = check_hip(hipEventRecord(e[3], st), "hipEventRecord");
    }
    if (!rc) rc = check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    double sum[3] = {0, 0, 0};
    for (int k = 0; k < nsteps && !rc; ++k) {
        hipEvent_t* e = &ev[(size_t)k * 4];
        for (int q = 0; q < 3 && !rc; ++q) {
            float ms = 0.f;
            rc = check_hip(hipEventElapsedTime(&ms, e[q], e[q + 1]), "hipEventElapsedTime");
            sum[q] += ms;
        }
    }
    float tot = 0.f;
    if (!rc) rc = check_hip(hipEventElapsedTime(&tot, ev[0], ev[(size_t)nsteps * 4 - 1]), "hipEventElapsedTime");
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc) return rc;
    if (ms_total) *ms_total = tot;
    if (s->fused) sum[2] = 0.0;   // the step was one launch (and its redo pass)
    if (ms_kernels)
        for (int q = 0; q < 3; ++q) ms_kernels[q] = sum[q] / nsteps;
    return MH_OK;
}

int mh_set_launch_shape(const mh_launch_shape* s) {
    if (!s) return fail(MH_ERR_INVALID, "mh_set_launch_shape: null shape");
    if (s->seg_x < 0 || s->seg_y < 0) return fail(MH_ERR_INVALID, "mh_set_launch_shape: negative segment");
    if (s->min_tiles < 0) return fail(MH_ERR_INVALID, "mh_set_launch_shape: negative min_tiles");
    if (s->y_waves != 1 && s->y_waves != 4) return fail(MH_ERR_INVALID, "mh_set_launch_shape: y_waves must be 1 or 4");
    if (s->fused < 0 || s->fused > 2) return fail(MH_ERR_INVALID, "mh_set_launch_shape: fused must be 0, 1 or 2");
    if (s->fused_tail < 0) return fail(MH_ERR_INVALID, "mh_set_launch_shape: negative fused_tail");
    if (s->x_layout < 0 || s->x_layout > 2) return fail(MH_ERR_INVALID, "mh_set_launch_shape: x_layout must be 0, 1 or 2");
    if ((s->x_wpb != 1 && s->x_wpb != 4) || (s->y_wpb != 1 && s->y_wpb != 4))
        return fail(MH_ERR_INVALID, "mh_set_launch_shape: wavefronts per workgroup must be 1 or 4");
    std::lock_guard<std::mutex> lk(g_shape_lock);
    g_shape = *s;
    return MH_OK;
}

int mh_get_launch_shape(mh_launch_shape* s) {
    if (!s) return fail(MH_ERR_INVALID, "mh_get_launch_shape: null shape");
    *s = shape();
    return MH_OK;
}

int mh_copy(const double* src, double* dst, int64_t n, mh_stream stream) {
    if (!src || !dst || n < 0 || (n & 1)) return fail(MH_ERR_INVALID, "mh_copy: need device buffers and an even n");
    const int64_t n2 = n / 2;
    const int blocks = (int)std::min<int64_t>((n2 + 255) / 256, (int64_t)1 << 30);
    hipLaunchKernelGGL(k_copy, dim3(std::max(blocks, 1)), dim3(256), 0, (hipStream_t)stream, (const double2*)src,
                       (double2*)dst, n2);
    return check_hip(hipGetLastError(), "copy launch");
}

int mh_device_alloc(int64_t bytes, void** ptr) {
    if (!ptr || bytes <= 0) return fail(MH_ERR_INVALID, "mh_device_alloc: bad argument");
    return check_hip(hipMalloc(ptr, (size_t)bytes), "hipMalloc");
}

int mh_device_free(void* ptr) { return check_hip(hipFree(ptr), "hipFree"); }

int mh_synchronize(mh_stream stream) { return check_hip(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize"); }

}  // extern "C"



