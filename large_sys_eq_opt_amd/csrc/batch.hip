// batch.hip -- a persistent kernel that advances an ensemble of independent
// grids of uneven size, each with its own CFL dt, by one split step per
// launch (no reference counterpart: the reference runs one grid per process).
//
// Layout: every instance is an ordinary grid (four pitched FP64 planes with
// ghosts) carved out of one arena, plus a scratch grid for the x-sweep
// output.  Work is cut into the same wavefront tiles as the single-grid
// sweeps (sweep_tiles.h): x tiles are row segments, y tiles are 64-column x
// chunk-row blocks.  The host lays the tiles out instance by instance and
// splits each phase's list into contiguous ranges of equal estimated cost,
// one range per wavefront of the persistent grid, so no tile queue (and no
// cross-XCD atomic per tile) is needed.
//
// One launch = one step of every instance:
//   phase x   each wavefront runs its x tiles with its instance's dt, formed
//             from the max signal speed reduced in the previous step;
//             block 0 commits the clocks (t += dt) and clears the slots the
//             y phase will fill.
//   barrier   grid-wide (all workgroups are resident: the grid is sized from
//             the occupancy of this kernel); bounded spin, flags an error
//             instead of hanging.
//   phase y   y tiles; each wavefront max of the new signal speed goes to its
//             instance's slot with one atomic max.
// Step parity, counters and clocks live in device memory, so the launch is
// identical every step and a captured graph of k launches replays k steps.
// The schedule can instead use two launches per step (x, then y), either of
// the same persistent kernel or of dispatched kernels (k_batch_x, k_batch_y,
// k_batch_y_redo) with one tile per wavefront: there the hardware dispatcher
// balances the load as tiles finish, tiles are ordered largest instance
// first, and each phase gets its own register budget.  Or (mode 3) it uses
// the fused split step (fused_tile.h): one launch per step in which a pair
// of wavefronts per y tile sweeps x into an LDS ring and y out of it, plus
// the launch that redoes the tiles whose fast pass gave up and commits the
// clocks; an instance's state then alternates between its grid and its
// scratch grid (BatchClock::cur) and mh_batch_clock puts it back.  The
// autotuner times the modes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "euler_device.h"
#include "grid.h"
#include "sweep_tiles.h"
#include "fused_tile.h"

namespace mh {

struct BatchInst {
    View U, T;
    double dx, dy, cfl, h;
    Bc bc;
};

struct BatchClock {
    unsigned long long smax[2];
    double t, dt;
    int step, status;
    int cur;   // fused steps: the state is in the instance's grid (0) or its scratch grid (1)
};

struct XTile { int inst, j0, j1, i0, i1; };   // rows j0..j1-1, cells i0..i1-1
struct YTile { int inst, cg, j0, j1; };

struct Ctrl {
    int step;
    unsigned bar[2];
    unsigned done[2];
    int err;
};

struct BatchArgs {
    const BatchInst* inst;
    BatchClock* clk;
    int ninst;
    const XTile* xt;
    const int* xstart;   // nwaves + 1
    const YTile* yt;
    const int* ystart;   // nwaves + 1
    Ctrl* ctrl;
    int phases;          // 1 = x only, 2 = y only, 3 = x, barrier, y
    // dispatched y phase: y tiles whose fast pass gave up, left to
    // k_batch_y_redo ([0] count, [1] finished workgroups, [2..] tile indices)
    int* redo;
};

__device__ __forceinline__ bool inst_dt(const BatchInst& I, const BatchClock& c, int par, double& dt) {
    const double m = __longlong_as_double((long long)c.smax[par]);
    if (!(m > 0.0)) return false;
    dt = I.cfl * I.h / m;   // src/cpu/solver_cpu.cpp:308
    return true;
}

// Returns false (in every thread of the workgroup) when the bounded spin
// gave up; the flag in *err then stops this and every later launch.
__device__ bool grid_barrier(unsigned* bar, unsigned nblocks, int* err) {
    __shared__ int ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(bar, 1u);
        long spins = 0;
        int good = 1;
        while (__hip_atomic_load(bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < nblocks) {
            __builtin_amdgcn_s_sleep(8);
            if (++spins > (1L << 24)) {   // ~ seconds: not all workgroups resident
                atomicExch(err, 1);
                good = 0;
                break;
            }
        }
        __threadfence();
        ok = good && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    }
    __syncthreads();
    return ok != 0;
}

// Clock commit of phase x, shared by the persistent and dispatched kernels:
// t += dt for every instance, clear the speed slots the y phase fills and
// the next step's barrier counters.
__device__ __forceinline__ void commit_clocks(const BatchArgs& a, int par) {
    if (threadIdx.x == 0) {
        a.ctrl->bar[par ^ 1] = 0u;
        a.ctrl->done[par ^ 1] = 0u;
    }
    for (int k = threadIdx.x; k < a.ninst; k += blockDim.x) {
        BatchClock& c = a.clk[k];
        double dt;
        if (inst_dt(a.inst[k], c, par, dt)) {
            c.dt = dt;
            c.t += dt;
            c.step += 1;
        } else {
            c.status = MH_ERR_WAVESPEED;
        }
        c.smax[par ^ 1] = 0ull;
    }
}

// One x tile: whole rows as one stream of passes (x_stream), segments of a
// row one row at a time.
template <int S>
__device__ __forceinline__ void x_phase_tile(const BatchInst& I, const XTile& t, double dtd, XLdsC& X, int lane) {
    if (t.i0 == 0 && t.i1 == I.U.nx) {
        x_stream<S>(I.U, I.T, t.j0, t.j1, dtd, 0.5 * dtd, I.bc, 1, X, lane);
        return;
    }
    for (int j = t.j0; j < t.j1; ++j) x_tile<S, 2, false>(I.U, I.T, j, t.i0, t.i1, dtd, 0.5 * dtd, I.bc, 1, X, lane);
}

union BatchSlice {   // a wavefront's LDS: x-tile exchange or y-tile ring
    XLdsC x;
    double ring[kRing];
};

template <int S>
__global__ __launch_bounds__(256) void k_batch_step(BatchArgs a) {
    __shared__ BatchSlice X[4];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int gw = blockIdx.x * 4 + w;
    // A barrier that gave up in an earlier launch leaves the error flag set:
    // this launch then touches nothing, so the instances keep the state of
    // the last complete step (the clocks of the failed step stay committed;
    // mh_batch_clock reports the error and mh_batch_prepare clears it).
    if (__hip_atomic_load(&a.ctrl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const int step = __hip_atomic_load(&a.ctrl->step, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int par = step & 1;
    if (a.phases & 1) {
        if (blockIdx.x == 0) commit_clocks(a, par);
        for (int q = a.xstart[gw]; q < a.xstart[gw + 1]; ++q) {
            const XTile t = a.xt[q];
            const BatchInst& I = a.inst[t.inst];
            double dt;
            if (!inst_dt(I, a.clk[t.inst], par, dt)) continue;
            x_phase_tile<S>(I, t, dt / I.dx, X[w].x, lane);
        }
    }
    // phase y rewrites U: never on top of an incomplete x phase
    if (a.phases == 3 && !grid_barrier(&a.ctrl->bar[par], gridDim.x, &a.ctrl->err)) return;
    if (a.phases & 2) {
        for (int q = a.ystart[gw]; q < a.ystart[gw + 1]; ++q) {
            const YTile t = a.yt[q];
            const BatchInst& I = a.inst[t.inst];
            double dt;
            if (!inst_dt(I, a.clk[t.inst], par, dt)) continue;
            const double dtd = dt / I.dy;
            double m = y_tile<S, false>(I.T, I.U, t.cg, t.j0, t.j1, dtd, 0.5 * dtd, I.bc, 1, 1, X[w].ring, lane);
            m = wave_max(m);
            if (lane == 0 && m > 0.0)
                atomicMax(&a.clk[t.inst].smax[par ^ 1], (unsigned long long)__double_as_longlong(m));
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();
            if (atomicAdd(&a.ctrl->done[par], 1u) == gridDim.x - 1)
                __hip_atomic_store(&a.ctrl->step, step + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Dispatched x phase: wavefront q of the launch runs x tile q (one
// wavefront per workgroup, as the single-grid x sweep runs best).  Block 0
// also commits the clocks; x tiles read only smax[par], which it leaves
// alone.
template <int S>
__global__ __launch_bounds__(64) void k_batch_x(BatchArgs a, int ntiles) {
    __shared__ XLdsC X[1];
    const int w = 0, lane = threadIdx.x;
    const int par = __hip_atomic_load(&a.ctrl->step, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 1;
    if (blockIdx.x == 0) commit_clocks(a, par);
    const int q = blockIdx.x;
    if (q >= ntiles) return;
    const XTile t = a.xt[q];
    // the instance's fields are read once here, before any store (the x
    // tile's buffer stores could alias them and force reloads every pass)
    const BatchInst& I = a.inst[t.inst];
    double dt;
    if (!inst_dt(I, a.clk[t.inst], par, dt)) return;
    x_phase_tile<S>(I, t, dt / I.dx, X[w], lane);
}

// Dispatched y phase: wavefront q runs y tile q with the fast division
// (y_tile_pass<S, true>); a tile whose states leave the fast range goes on
// the redo list for k_batch_y_redo, launched right after, which also
// advances the step counter.  Keeping the IEEE pass out of this kernel keeps
// it within the register budget of four wavefronts per SIMD.
template <int S>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 8))) void k_batch_y(BatchArgs a, int ntiles) {
    __shared__ double ring[4][kRing];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int par = __hip_atomic_load(&a.ctrl->step, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 1;
    const int q = blockIdx.x * 4 + w;
    if (q >= ntiles) return;
    const YTile t = a.yt[q];
    // the instance's fields are read once here, before any store, so they
    // are scalar loads kept in scalar registers
    const BatchInst I = a.inst[t.inst];
    double dt;
    if (!inst_dt(I, a.clk[t.inst], par, dt)) return;
    const double dtd = dt / I.dy;
    double m = 0.0;
    if (dt_in_fast_range(dtd) &&
        y_tile_pass<S, true, false>(I.T, I.U, t.cg, t.j0, t.j1, dtd, 0.5 * dtd, I.bc, 1, 1, ring[w], lane, m)) {
        m = wave_max(m);
        if (lane == 0 && m > 0.0) atomicMax(&a.clk[t.inst].smax[par ^ 1], (unsigned long long)__double_as_longlong(m));
    } else if (lane == 0) {
        a.redo[2 + atomicAdd(a.redo, 1)] = q;
    }
}

// IEEE pass over the y tiles k_batch_y listed; the last workgroup to finish
// empties the list and advances the step counter.
template <int S>
__global__ __launch_bounds__(256) void k_batch_y_redo(BatchArgs a) {
    __shared__ double ring[4][kRing];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int step = __hip_atomic_load(&a.ctrl->step, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int par = step & 1;
    const int n = __hip_atomic_load(a.redo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int r = blockIdx.x * 4 + w; r < n; r += gridDim.x * 4) {
        const YTile t = a.yt[a.redo[2 + r]];
        const BatchInst& I = a.inst[t.inst];
        double dt;
        if (!inst_dt(I, a.clk[t.inst], par, dt)) continue;
        const double dtd = dt / I.dy;
        double m = 0.0;
        y_tile_pass<S, false, false>(I.T, I.U, t.cg, t.j0, t.j1, dtd, 0.5 * dtd, I.bc, 1, 1, ring[w], lane, m);
        m = wave_max(m);
        if (lane == 0 && m > 0.0) atomicMax(&a.clk[t.inst].smax[par ^ 1], (unsigned long long)__double_as_longlong(m));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(a.redo + 1, 1) == (int)gridDim.x - 1) {
            a.redo[0] = 0;
            a.redo[1] = 0;
            __hip_atomic_store(&a.ctrl->step, step + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Values read from an instance's record after the kernel may have stored
// (the compiler then loads them into vector registers, which the fused
// tile's HLLC pass cannot spare): the same value in every lane, moved into
// scalar registers.
__device__ __forceinline__ int64_t uniform64(int64_t x) {
    return (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(x >> 32)) << 32) |
                     (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x));
}
__device__ __forceinline__ double uniform(double x) {
    return __longlong_as_double(uniform64(__double_as_longlong(x)));
}
__device__ __forceinline__ View uniform(const View& v) {
    View u;
    u.o = reinterpret_cast<double*>(uniform64(reinterpret_cast<int64_t>(v.o)));
    u.ps = uniform64(v.ps);
    u.rp = uniform64(v.rp);
    u.nx = __builtin_amdgcn_readfirstlane(v.nx);
    u.ny = __builtin_amdgcn_readfirstlane(v.ny);
    return u;
}

// Fused step (mode 3): workgroup q is the x/y wavefront pair of y tile q
// (tiles of the largest instances first); each reads its instance's state
// where the instance's clock says it is and writes the other grid.  A tile
// whose fast pass gives up goes on the redo list.
template <int S>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_batch_fused(BatchArgs a,
                                                                                              int ntiles) {
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int q = blockIdx.x;
    if (q >= ntiles) return;
    const int par = __hip_atomic_load(&a.ctrl->step, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 1;
    const YTile t = a.yt[q];
    // the views and step factors in scalar registers; the boundary spec,
    // read only at the grid's edges, stays in memory
    const BatchInst& I = a.inst[t.inst];
    const BatchClock& c = a.clk[t.inst];
    double dt;
    if (!inst_dt(I, c, par, dt)) return;
    const bool swap = __builtin_amdgcn_readfirstlane(c.cur) != 0;
    const View in = uniform(swap ? I.T : I.U), out = uniform(swap ? I.U : I.T);
    const double dtdx = uniform(dt / I.dx), dtdy = uniform(dt / I.dy);
    const Bc& bc = I.bc;
    __shared__ PairLds X;
    if (!(dt_in_fast_range(dtdx) && dt_in_fast_range(dtdy))) {
        if (w == 1 && lane == 0) a.redo[2 + atomicAdd(a.redo, 1)] = q;
        return;
    }
    double m = 0.0;
    const bool ok = pair_side<S, true>(w, in, out, t.cg * 64, t.j0, t.j1, dtdx, dtdy, bc, 1, 1, X, lane, m);
    if (w == 1) {
        if (!(ok && X.xok != 0u)) {
            if (lane == 0) a.redo[2 + atomicAdd(a.redo, 1)] = q;
        } else {
            m = wave_max(m);
            if (lane == 0 && m > 0.0)
                atomicMax(&a.clk[t.inst].smax[par ^ 1], (unsigned long long)__double_as_longlong(m));
        }
    }
}

// IEEE pass over the tiles k_batch_fused listed (both wavefronts of a
// workgroup walk the same tiles, a barrier between tiles); then the last
// workgroup to finish commits the step: every instance that stepped adds its
// dt, flips its state location and clears the speed slot the step read; the
// list is emptied and the step counter advanced.
template <int S>
__global__ __launch_bounds__(128) void k_batch_fused_redo(BatchArgs a) {
    __shared__ PairLds X;
    __shared__ int last;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int step = __hip_atomic_load(&a.ctrl->step, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int par = step & 1;
    const int n = __hip_atomic_load(a.redo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // a list too short to fill the chip: each tile in up to 8 pieces of at
    // least 16 rows (each with its own warm-up rows) that run side by side
    const int pieces = max(1, min(8, (2048 + n - 1) / max(n, 1)));
    for (int q = blockIdx.x; q < n * pieces; q += gridDim.x) {
        const YTile t = a.yt[a.redo[2 + q / pieces]];
        const int rows = max(16, (t.j1 - t.j0 + pieces - 1) / pieces);
        const int p0 = t.j0 + (q % pieces) * rows;
        const BatchInst& I = a.inst[t.inst];
        const BatchClock& c = a.clk[t.inst];
        double dt;
        if (p0 < t.j1 && inst_dt(I, c, par, dt)) {
            const bool swap = __builtin_amdgcn_readfirstlane(c.cur) != 0;
            const View in = uniform(swap ? I.T : I.U), out = uniform(swap ? I.U : I.T);
            const double dtdx = uniform(dt / I.dx), dtdy = uniform(dt / I.dy);
            double m = 0.0;
            pair_side<S, false>(w, in, out, t.cg * 64, p0, min(p0 + rows, t.j1), dtdx, dtdy, I.bc, 1, 1, X, lane, m);
            if (w == 1) {
                m = wave_max(m);
                if (lane == 0 && m > 0.0)
                    atomicMax(&a.clk[t.inst].smax[par ^ 1], (unsigned long long)__double_as_longlong(m));
            }
        }
        pair_barrier();
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(a.redo + 1, 1) == (int)gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    for (int k = threadIdx.x; k < a.ninst; k += blockDim.x) {
        BatchClock& c = a.clk[k];
        double dt;
        if (inst_dt(a.inst[k], c, par, dt)) {
            c.dt = dt;
            c.t += dt;
            c.step += 1;
            c.cur ^= 1;
        } else {
            c.status = MH_ERR_WAVESPEED;
        }
        c.smax[par] = 0ull;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a.redo[0] = 0;
        a.redo[1] = 0;
        __hip_atomic_store(&a.ctrl->step, step + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Initial max signal speed of every instance into slot 0.
__global__ __launch_bounds__(256) void k_batch_speed(const BatchInst* inst, BatchClock* clk, int ninst) {
    const int k = blockIdx.y;
    if (k >= ninst) return;
    const View v = inst[k].U;
    const int64_t n = (int64_t)v.nx * v.ny;
    double m = 0.0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(c / v.nx);
        const int i = (int)(c - (int64_t)j * v.nx);
        m = smax(m, signal_speed(ld(v, i, j)));
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0 && m > 0.0)
        atomicMax(&clk[k].smax[0], (unsigned long long)__double_as_longlong(m));
}

}  // namespace mh

using namespace mh;

struct mh_batch_s {
    int n = 0;
    int solver = 1;
    double cfl = 0.8;
    double* arena = nullptr;
    std::vector<mh_grid_s> grids;       // non-owning views of the instances
    std::vector<BatchInst> hinst;
    BatchInst* dinst = nullptr;
    BatchClock* dclk = nullptr;
    Ctrl* dctrl = nullptr;
    XTile* dxt = nullptr;
    YTile* dyt = nullptr;
    int* dxs = nullptr;
    int* dys = nullptr;
    int* dredo = nullptr;   // dispatched y phase's redo list (2 + nyt ints)
    int nblocks = 0;
    int nxt = 0, nyt = 0;
    // schedule; mode 0: one persistent launch per step, 1: two persistent
    // launches, 2: two dispatched launches (one tile per wavefront), 3: the
    // fused step (a wavefront pair per y tile)
    int xseg = 0, xrows = 1, ychunk = 256, mode = 0, blocks_per_cu = 0;
    // size class: instances with fewer than small_side cells a side take
    // their own y chunk rows and x tile rows (0: the schedule's)
    int small_side = 0, small_ychunk = 0, small_xrows = 0;
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_def = nullptr;
    int graph_steps = 0;
};

namespace {

int64_t grid_rp(int nx) { return round_up((int64_t)kLead + nx + kNg, 16); }
int64_t grid_ps(int nx, int ny) { return round_up(grid_rp(nx) * (ny + 2 * kNg), 256); }

void free_dev(void* p) {
    if (p) (void)hipFree(p);
}

void drop_graph(mh_batch b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->graph_def) (void)hipGraphDestroy(b->graph_def);
    b->graph = nullptr;
    b->graph_def = nullptr;
    b->graph_steps = 0;
}

const void* batch_kernel(int solver) {
    switch (solver) {
        case kHLL: return (const void*)k_batch_step<kHLL>;
        case kHLLC: return (const void*)k_batch_step<kHLLC>;
        case kExact: return (const void*)k_batch_step<kExact>;
        default: return (const void*)k_batch_step<kForce>;
    }
}

// Contiguous ranges of equal cost over `cost` for nw wavefronts.
std::vector<int> partition(const std::vector<double>& cost, int nw) {
    std::vector<int> start(nw + 1, (int)cost.size());
    double total = 0.0;
    for (double c : cost) total += c;
    double acc = 0.0;
    int q = 0;
    for (int w = 0; w < nw; ++w) {
        start[w] = q;
        const double target = total * (double)(w + 1) / (double)nw;
        while (q < (int)cost.size() && acc + 0.5 * cost[q] <= target) acc += cost[q++];
    }
    start[nw] = (int)cost.size();
    return start;
}

int build_schedule(mh_batch b) {
    drop_graph(b);
    int dev = 0;
    int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    hipDeviceProp_t prop;
    rc = check_hip(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    if (rc) return rc;
    int occ = 0;
    rc = check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, batch_kernel(b->solver), 256, 0),
                   "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (rc) return rc;
    occ = std::max(occ, 1);
    if (b->blocks_per_cu > 0) occ = std::min(occ, b->blocks_per_cu);
    b->nblocks = prop.multiProcessorCount * occ;
    const int nw = b->nblocks * 4;
    std::vector<XTile> xt;
    std::vector<YTile> yt;
    std::vector<double> xc, yc;
    // Largest instances first: under dispatched launches the long tiles
    // start early and the short ones fill the tail.
    std::vector<int> order(b->n);
    for (int k = 0; k < b->n; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int p, int q) {
        return (int64_t)b->hinst[p].U.nx * b->hinst[p].U.ny > (int64_t)b->hinst[q].U.nx * b->hinst[q].U.ny;
    });
    for (int k : order) {
        const int nx = b->hinst[k].U.nx, ny = b->hinst[k].U.ny;
        const bool small = b->small_side > 0 && nx < b->small_side;
        const int seg = b->xseg > 0 ? b->xseg : nx;
        // whole-row tiles are streamed (x_stream): their rows must span less
        // than 2^31 bytes of a plane
        const int64_t rp = b->hinst[k].U.rp;
        const int rows = (int)std::min<int64_t>(std::max(1, small && b->small_xrows > 0 ? b->small_xrows : b->xrows),
                                                std::max<int64_t>(1, ((int64_t)1 << 28) / rp - 1));
        const int ychunk = small && b->small_ychunk > 0 ? b->small_ychunk : b->ychunk;
        for (int j0 = 0; j0 < ny; j0 += rows)
            for (int i0 = 0; i0 < nx; i0 += seg) {
                const int i1 = std::min(nx, i0 + seg), j1 = std::min(ny, j0 + rows);
                xt.push_back({k, j0, j1, i0, i1});
                xc.push_back(seg == nx ? 64.0 * (((int64_t)(j1 - j0) * (nx + 4) + 63) / 64)
                                       : 64.0 * (j1 - j0) * ((i1 - i0 + 4 + 63) / 64));
            }
        const int ch = ychunk > 0 ? ychunk : ny;
        for (int cg = 0; cg * 64 < nx; ++cg)
            for (int j0 = 0; j0 < ny; j0 += ch) {
                const int j1 = std::min(ny, j0 + ch);
                yt.push_back({k, cg, j0, j1});
                yc.push_back(64.0 * (j1 - j0 + 4));
            }
    }
    if (b->mode >= 2) {
        // Dispatched launches: longest tiles first (each instance's short
        // remainder tiles end up in the tail, where they fill idle SIMDs).
        auto lpt = [](auto& tiles, std::vector<double>& cost) {
            std::vector<int> idx(tiles.size());
            for (size_t q = 0; q < idx.size(); ++q) idx[q] = (int)q;
            std::stable_sort(idx.begin(), idx.end(), [&](int p, int q) { return cost[p] > cost[q]; });
            auto t2 = tiles;
            std::vector<double> c2(cost.size());
            for (size_t q = 0; q < idx.size(); ++q) { t2[q] = tiles[idx[q]]; c2[q] = cost[idx[q]]; }
            tiles.swap(t2);
            cost.swap(c2);
        };
        lpt(xt, xc);
        lpt(yt, yc);
    }
    const std::vector<int> xs = partition(xc, nw), ys = partition(yc, nw);
    free_dev(b->dxt); free_dev(b->dyt); free_dev(b->dxs); free_dev(b->dys); free_dev(b->dredo);
    b->dxt = nullptr; b->dyt = nullptr; b->dxs = nullptr; b->dys = nullptr; b->dredo = nullptr;
    rc = check_hip(hipMalloc((void**)&b->dxt, sizeof(XTile) * xt.size()), "hipMalloc(tiles)");
    if (!rc) rc = check_hip(hipMalloc((void**)&b->dyt, sizeof(YTile) * yt.size()), "hipMalloc(tiles)");
    if (!rc) rc = check_hip(hipMalloc((void**)&b->dxs, sizeof(int) * xs.size()), "hipMalloc(tiles)");
    if (!rc) rc = check_hip(hipMalloc((void**)&b->dys, sizeof(int) * ys.size()), "hipMalloc(tiles)");
    if (!rc) rc = check_hip(hipMalloc((void**)&b->dredo, sizeof(int) * (2 + yt.size())), "hipMalloc(tiles)");
    if (!rc) rc = check_hip(hipMemset(b->dredo, 0, 2 * sizeof(int)), "hipMemset");
    if (!rc) rc = check_hip(hipMemcpy(b->dxt, xt.data(), sizeof(XTile) * xt.size(), hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc) rc = check_hip(hipMemcpy(b->dyt, yt.data(), sizeof(YTile) * yt.size(), hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc) rc = check_hip(hipMemcpy(b->dxs, xs.data(), sizeof(int) * xs.size(), hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc) rc = check_hip(hipMemcpy(b->dys, ys.data(), sizeof(int) * ys.size(), hipMemcpyHostToDevice), "hipMemcpy");
    b->nxt = (int)xt.size();
    b->nyt = (int)yt.size();
    return rc;
}

int launch_step(mh_batch b, hipStream_t st) {
    BatchArgs a;
    a.inst = b->dinst;
    a.clk = b->dclk;
    a.ninst = b->n;
    a.xt = b->dxt;
    a.xstart = b->dxs;
    a.yt = b->dyt;
    a.ystart = b->dys;
    a.ctrl = b->dctrl;
    a.redo = b->dredo;
    const dim3 g(b->nblocks), blk(256);
    if (b->mode == 3) {
        // the redo pass: 256 workgroups, for the exact solver (whose fast
        // pass leaves every tile with a non-uniform face to it) one per four
        // tiles
        const int nr = b->solver == kExact ? std::max(256, b->nyt / 4) : 256;
        const dim3 gf((unsigned)std::max(1, b->nyt)), gr((unsigned)std::max(1, std::min(nr, b->nyt))), bf(128);
        switch (b->solver) {
            case kHLL: hipLaunchKernelGGL(k_batch_fused<kHLL>, gf, bf, 0, st, a, b->nyt);
                       hipLaunchKernelGGL(k_batch_fused_redo<kHLL>, gr, bf, 0, st, a); break;
            case kHLLC: hipLaunchKernelGGL(k_batch_fused<kHLLC>, gf, bf, 0, st, a, b->nyt);
                        hipLaunchKernelGGL(k_batch_fused_redo<kHLLC>, gr, bf, 0, st, a); break;
            case kExact: hipLaunchKernelGGL(k_batch_fused<kExact>, gf, bf, 0, st, a, b->nyt);
                         hipLaunchKernelGGL(k_batch_fused_redo<kExact>, gr, bf, 0, st, a); break;
            default: hipLaunchKernelGGL(k_batch_fused<kForce>, gf, bf, 0, st, a, b->nyt);
                     hipLaunchKernelGGL(k_batch_fused_redo<kForce>, gr, bf, 0, st, a); break;
        }
        return check_hip(hipGetLastError(), "batch step launch");
    }
    if (b->mode >= 2) {
        const dim3 gx((unsigned)std::max(1, b->nxt)), gy((unsigned)std::max(1, (b->nyt + 3) / 4)), bx(64);
        const dim3 gr((unsigned)std::max(1, std::min(64, (b->nyt + 3) / 4)));
        switch (b->solver) {
            case kHLL: hipLaunchKernelGGL(k_batch_x<kHLL>, gx, bx, 0, st, a, b->nxt);
                       hipLaunchKernelGGL(k_batch_y<kHLL>, gy, blk, 0, st, a, b->nyt);
                       hipLaunchKernelGGL(k_batch_y_redo<kHLL>, gr, blk, 0, st, a); break;
            case kHLLC: hipLaunchKernelGGL(k_batch_x<kHLLC>, gx, bx, 0, st, a, b->nxt);
                        hipLaunchKernelGGL(k_batch_y<kHLLC>, gy, blk, 0, st, a, b->nyt);
                        hipLaunchKernelGGL(k_batch_y_redo<kHLLC>, gr, blk, 0, st, a); break;
            case kExact: hipLaunchKernelGGL(k_batch_x<kExact>, gx, bx, 0, st, a, b->nxt);
                         hipLaunchKernelGGL(k_batch_y<kExact>, gy, blk, 0, st, a, b->nyt);
                         hipLaunchKernelGGL(k_batch_y_redo<kExact>, gr, blk, 0, st, a); break;
            default: hipLaunchKernelGGL(k_batch_x<kForce>, gx, bx, 0, st, a, b->nxt);
                     hipLaunchKernelGGL(k_batch_y<kForce>, gy, blk, 0, st, a, b->nyt);
                     hipLaunchKernelGGL(k_batch_y_redo<kForce>, gr, blk, 0, st, a); break;
        }
        return check_hip(hipGetLastError(), "batch step launch");
    }
    const int np = b->mode == 1 ? 2 : 1;
    for (int k = 0; k < np; ++k) {
        a.phases = b->mode == 1 ? (k == 0 ? 1 : 2) : 3;
        switch (b->solver) {
            case kHLL: hipLaunchKernelGGL(k_batch_step<kHLL>, g, blk, 0, st, a); break;
            case kHLLC: hipLaunchKernelGGL(k_batch_step<kHLLC>, g, blk, 0, st, a); break;
            case kExact: hipLaunchKernelGGL(k_batch_step<kExact>, g, blk, 0, st, a); break;
            default: hipLaunchKernelGGL(k_batch_step<kForce>, g, blk, 0, st, a); break;
        }
    }
    return check_hip(hipGetLastError(), "batch step launch");
}

// Puts every instance's state back into its grid after fused steps (an
// instance whose clock says cur = 1 has it in its scratch grid: the two
// grids swap roles) and leaves the clocks as the other modes expect them
// between steps: cur = 0 and the speed slot the next step fills cleared.
// Synchronises `st`.
int batch_settle(mh_batch b, hipStream_t st, std::vector<BatchClock>* out_clk, Ctrl* out_ctrl) {
    std::vector<BatchClock> clk(b->n);
    Ctrl c;
    int rc = check_hip(hipMemcpyAsync(clk.data(), b->dclk, sizeof(BatchClock) * b->n, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    if (!rc) rc = check_hip(hipMemcpyAsync(&c, b->dctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    if (!rc) rc = check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (rc) return rc;
    bool moved = false;
    const int nxt = (c.step & 1) ^ 1;
    for (int k = 0; k < b->n; ++k) {
        if (clk[k].smax[nxt] != 0ull) {
            clk[k].smax[nxt] = 0ull;
            moved = true;
        }
        if (!clk[k].cur) continue;
        BatchInst& I = b->hinst[k];
        std::swap(I.U, I.T);
        b->grids[k].base = I.U.o - ((int64_t)kNg * I.U.rp + kLead);
        clk[k].cur = 0;
        moved = true;
    }
    if (moved) {
        rc = check_hip(hipMemcpyAsync(b->dinst, b->hinst.data(), sizeof(BatchInst) * b->n, hipMemcpyHostToDevice, st),
                       "hipMemcpyAsync");
        if (!rc) rc = check_hip(hipMemcpyAsync(b->dclk, clk.data(), sizeof(BatchClock) * b->n, hipMemcpyHostToDevice, st),
                                "hipMemcpyAsync");
        if (!rc) rc = check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    }
    if (out_clk) *out_clk = clk;
    if (out_ctrl) *out_ctrl = c;
    return rc;
}

}  // namespace

extern "C" {

int mh_batch_create(int n, const mh_grid_desc* descs, int solver, double cfl, const mh_bc* bcs, mh_batch* out) {
    if (n < 1 || !descs || !out) return fail(MH_ERR_INVALID, "mh_batch_create: bad argument");
    if (solver < 0 || solver > 3) return fail(MH_ERR_INVALID, "mh_batch_create: unknown solver");
    if (!(cfl > 0.0)) return fail(MH_ERR_INVALID, "mh_batch_create: cfl must be positive");
    for (int k = 0; k < n; ++k)
        if (descs[k].nx < 1 || descs[k].ny < 1 || !(descs[k].x_max > descs[k].x_min) ||
            !(descs[k].y_max > descs[k].y_min))
            return fail(MH_ERR_INVALID, "mh_batch_create: bad instance descriptor");
    if (bcs)
        for (int k = 0; k < n; ++k)
            for (int e = 0; e < 4; ++e)
                if (bcs[k].kind[e] != MH_BC_TRANSMISSIVE && bcs[k].kind[e] != MH_BC_INFLOW)
                    return fail(MH_ERR_INVALID, "mh_batch_create: unknown boundary kind");
    mh_batch b = new (std::nothrow) mh_batch_s();
    if (!b) return fail(MH_ERR_OOM, "host allocation failed");
    b->n = n;
    b->solver = solver;
    b->cfl = cfl;
    std::vector<int64_t> off(n);
    int64_t total = 0;
    for (int k = 0; k < n; ++k) {
        off[k] = total;
        total += 2 * 4 * grid_ps(descs[k].nx, descs[k].ny);
    }
    int rc = check_hip(hipMalloc((void**)&b->arena, sizeof(double) * (size_t)total), "hipMalloc(batch arena)");
    if (rc) { delete b; return rc; }
    b->grids.resize(n);
    b->hinst.resize(n);
    std::vector<BatchClock> clk(n);
    for (int k = 0; k < n; ++k) {
        mh_grid_s& g = b->grids[k];
        g.d = descs[k];
        g.dx = (descs[k].x_max - descs[k].x_min) / (double)descs[k].nx;
        g.dy = (descs[k].y_max - descs[k].y_min) / (double)descs[k].ny;
        g.rp = grid_rp(descs[k].nx);
        g.ps = grid_ps(descs[k].nx, descs[k].ny);
        g.base = b->arena + off[k];
        g.aux = nullptr;
        g.owns = 0;
        (void)hipGetDevice(&g.device);
        mh_grid_s t = g;
        t.base = g.base + 4 * g.ps;
        BatchInst& I = b->hinst[k];
        I.U = g.view();
        I.T = t.view();
        I.dx = g.dx;
        I.dy = g.dy;
        I.cfl = cfl;
        I.h = std::min(g.dx, g.dy);
        I.bc = to_bc(bcs ? &bcs[k] : nullptr);
        std::memset(&clk[k], 0, sizeof(BatchClock));
    }
    rc = check_hip(hipMalloc((void**)&b->dinst, sizeof(BatchInst) * n), "hipMalloc(batch)");
    if (!rc) rc = check_hip(hipMalloc((void**)&b->dclk, sizeof(BatchClock) * n), "hipMalloc(batch)");
    if (!rc) rc = check_hip(hipMalloc((void**)&b->dctrl, sizeof(Ctrl)), "hipMalloc(batch)");
    if (!rc) rc = check_hip(hipMemcpy(b->dinst, b->hinst.data(), sizeof(BatchInst) * n, hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc) rc = check_hip(hipMemcpy(b->dclk, clk.data(), sizeof(BatchClock) * n, hipMemcpyHostToDevice), "hipMemcpy");
    Ctrl c0{};
    if (!rc) rc = check_hip(hipMemcpy(b->dctrl, &c0, sizeof(Ctrl), hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc) rc = build_schedule(b);
    if (rc) { mh_batch_destroy(b); return rc; }
    *out = b;
    return MH_OK;
}

int mh_batch_destroy(mh_batch b) {
    if (!b) return MH_OK;
    drop_graph(b);
    free_dev(b->arena); free_dev(b->dinst); free_dev(b->dclk); free_dev(b->dctrl);
    free_dev(b->dxt); free_dev(b->dyt); free_dev(b->dxs); free_dev(b->dys); free_dev(b->dredo);
    delete b;
    return MH_OK;
}

int mh_batch_grid(mh_batch b, int k, mh_grid* g) {
    if (!b || k < 0 || k >= b->n || !g) return fail(MH_ERR_INVALID, "mh_batch_grid: bad argument");
    *g = &b->grids[k];
    return MH_OK;
}

int mh_batch_prepare(mh_batch b, mh_stream stream) {
    if (!b) return fail(MH_ERR_INVALID, "mh_batch_prepare: null batch");
    hipStream_t st = (hipStream_t)stream;
    int rc = MH_OK;
    for (int k = 0; k < b->n && !rc; ++k) rc = launch_fill(&b->grids[k], b->hinst[k].bc, st);
    std::vector<BatchClock> clk(b->n);
    for (auto& c : clk) std::memset(&c, 0, sizeof(c));
    Ctrl c0{};
    if (!rc) rc = check_hip(hipMemcpyAsync(b->dclk, clk.data(), sizeof(BatchClock) * b->n, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    if (!rc) rc = check_hip(hipMemcpyAsync(b->dctrl, &c0, sizeof(Ctrl), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    if (!rc) rc = check_hip(hipMemsetAsync(b->dredo, 0, 2 * sizeof(int), st), "hipMemsetAsync");
    int maxcells = 0;
    for (auto& I : b->hinst) maxcells = std::max(maxcells, I.U.nx * I.U.ny);
    if (!rc) {
        hipLaunchKernelGGL(k_batch_speed, dim3(std::max(1, std::min(64, (maxcells + 255) / 256)), b->n), dim3(256), 0, st,
                           b->dinst, b->dclk, b->n);
        rc = check_hip(hipGetLastError(), "batch speed launch");
    }
    if (!rc) rc = check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    return rc;
}

int mh_batch_set_schedule(mh_batch b, int xseg, int xrows, int ychunk, int mode, int blocks_per_cu) {
    if (!b || xseg < 0 || xrows < 1 || ychunk < 0 || blocks_per_cu < 0 || mode < 0 || mode > 3)
        return fail(MH_ERR_INVALID, "mh_batch_set_schedule: bad argument");
    const int rc = batch_settle(b, nullptr, nullptr, nullptr);
    if (rc) return rc;
    b->xseg = xseg;
    b->xrows = xrows;
    b->ychunk = ychunk;
    b->mode = mode;
    b->blocks_per_cu = blocks_per_cu;
    return build_schedule(b);
}

int mh_batch_set_size_class(mh_batch b, int side_below, int ychunk, int xrows) {
    if (!b || side_below < 0 || ychunk < 0 || xrows < 0) return fail(MH_ERR_INVALID, "mh_batch_set_size_class: bad argument");
    const int rc = batch_settle(b, nullptr, nullptr, nullptr);
    if (rc) return rc;
    b->small_side = side_below;
    b->small_ychunk = ychunk;
    b->small_xrows = xrows;
    return build_schedule(b);
}

int mh_batch_step(mh_batch b, int nsteps, mh_stream stream) {
    if (!b || nsteps < 0) return fail(MH_ERR_INVALID, "mh_batch_step: bad argument");
    hipStream_t st = (hipStream_t)stream;
    int k = 0;
    if (b->graph && b->graph_steps > 0)
        for (; k + b->graph_steps <= nsteps; k += b->graph_steps) {
            const int rc = check_hip(hipGraphLaunch(b->graph, st), "hipGraphLaunch");
            if (rc) return rc;
        }
    for (; k < nsteps; ++k) {
        const int rc = launch_step(b, st);
        if (rc) return rc;
    }
    return MH_OK;
}

int mh_batch_capture(mh_batch b, int steps_per_graph, mh_stream stream) {
    if (!b || steps_per_graph < 1) return fail(MH_ERR_INVALID, "mh_batch_capture: bad argument");
    drop_graph(b);
    hipStream_t st = (hipStream_t)stream;
    hipStream_t cap = nullptr;
    int rc = check_hip(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate");
    if (rc) return rc;
    rc = check_hip(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    for (int k = 0; k < steps_per_graph && !rc; ++k) rc = launch_step(b, cap);
    hipGraph_t g = nullptr;
    const int rc2 = check_hip(hipStreamEndCapture(cap, &g), "hipStreamEndCapture");
    if (!rc) rc = rc2;
    if (!rc) rc = check_hip(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (!rc) {
        b->graph_def = g;
        b->graph_steps = steps_per_graph;
    } else if (g) {
        (void)hipGraphDestroy(g);
    }
    (void)hipStreamDestroy(cap);
    (void)st;
    return rc;
}

int mh_batch_clock(mh_batch b, double* t, double* dt, int* steps, int* status, mh_stream stream) {
    if (!b) return fail(MH_ERR_INVALID, "mh_batch_clock: null batch");
    hipStream_t st = (hipStream_t)stream;
    std::vector<BatchClock> clk;
    Ctrl c;
    const int rc = batch_settle(b, st, &clk, &c);
    if (rc) return rc;
    int bad = 0;
    for (int k = 0; k < b->n; ++k) {
        if (t) t[k] = clk[k].t;
        if (dt) dt[k] = clk[k].dt;
        if (steps) steps[k] = clk[k].step;
        if (status) status[k] = clk[k].status;
        bad |= clk[k].status;
    }
    if (c.err) return fail(MH_ERR_LAUNCH, "batch: grid barrier timed out (workgroups not co-resident)");
    if (bad) return fail(MH_ERR_WAVESPEED, "batch: an instance met a non-positive max wave speed");
    return MH_OK;
}

int mh_batch_sync(mh_batch b, mh_stream stream) {
    if (!b) return fail(MH_ERR_INVALID, "mh_batch_sync: null batch");
    return batch_settle(b, (hipStream_t)stream, nullptr, nullptr);
}

int mh_batch_info(mh_batch b, int* nblocks, int* nxtiles, int* nytiles) {
    if (!b) return fail(MH_ERR_INVALID, "mh_batch_info: null batch");
    if (nblocks) *nblocks = b->nblocks;
    if (nxtiles) *nxtiles = b->nxt;
    if (nytiles) *nytiles = b->nyt;
    return MH_OK;
}

}  // extern "C"
