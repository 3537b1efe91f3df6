// adjoint.hip -- discrete adjoint of the split MUSCL-Hancock step (no
// reference counterpart).
//
// Vector-Jacobian product of one sweep: k_vjp_y streams up a column chunk
// like the forward y tile, with a three-row lag: at row p it converts row p,
// reconstructs row p-1 (primal), back-propagates the face between rows p-2
// and p-1 (its adjoint is dt/dy times the difference of the output adjoints
// of the two rows), back-propagates the reconstruction of row p-2 (which now
// has both edge adjoints) and finalises row p-3, whose primitive adjoint has
// then received all three reconstruction contributions.  Each stage runs
// only where it feeds an output row of the chunk.
//
// The x sweep's product reuses the same kernel on transposed grids: with
// rho*u and rho*v exchanged, the solver's x sweep is its y sweep operation
// for operation (the flux functions are mirror images and the conversions
// add u*u and v*v, which commutes exactly), so transposing state and
// adjoint, running k_vjp_y and transposing back is exact.
//
// One split step U -> T -> U' (x sweep, ghosts of T, y sweep) at a given dt:
//   T  = xsweep(U)
//   bT = vjp_y(T, bU')  with the y-ghost adjoints of T folded by its edges,
//                       written transposed
//   bU = vjp_x(U, bT)   k_vjp_y on U transposed and bT (already transposed),
//                       x-ghost adjoints folded, written back untransposed
// (the folds and the output transposes happen inside k_vjp_y).  The time step
// is held fixed (the adjoint of the step at the primal dt sequence).
// mh_adjoint_forward / mh_adjoint_backward run N steps with a checkpoint
// every K steps.  The backward pass recomputes each K-step segment once from
// its checkpoint -- keeping its states' x-sweep outputs and transposes -- on
// a second stream, into one of two buffers, while the products of the
// previous segment run on the caller's stream from the other buffer; each
// step's product is then two k_vjp_y launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "adjoint_device.h"
#include "euler_device.h"
#include "grid.h"
#include "sweep_tiles.h"

namespace mh {

// Host helpers defined in solver.hip.
int launch_sweep_host(const mh_grid_s* in, const mh_grid_s* out, int dir, int solver, double dt, const Bc& bc,
                      int ghosts, hipStream_t st);
int max_speed_host(const mh_grid_s* g, double* smax, hipStream_t st);
int stepper_start(mh_stepper s, double cfl, hipStream_t st);
void stepper_unfuse(mh_stepper s);

// Output rows -2..ny+1 are cut into nch chunks of seg rows, the last one
// running to row ny+1 (seg..2*seg-1 rows), so that rows -2..0 share the
// first chunk and rows ny-1..ny+1 the last (seg >= 3).  fold < 0 writes the
// ghost-row adjoints as they are; otherwise they are folded into rows 0 and
// ny-1 by the edge kinds (transmissive: the ghosts copy those rows; inflow:
// the ghosts are constants) and the ghost rows are written as zero.
// TO = 1 writes the result transposed (cell (i, c) to out(c, i), rho*u and
// rho*v exchanged), i.e. in the layout the other direction's product reads,
// so no separate transpose is needed; the lanes of a store then hit 64 rows
// of one column, and consecutive rows fill those lines.
template <int TO>
__device__ __forceinline__ void vjp_store(const View& out, int i, int c, const Cons& v) {
    if (TO) st(out, c, i, Cons{v.r, v.mv, v.mu, v.e});
    else st(out, i, c, v);
}

// One pass of k_vjp_y over its chunk.  F (fast pass): the sweeps' fast
// division and square root (adjoint_device.h); the range key of every state
// read is folded into a lane maximum and one vote at the end says whether
// the pass's results stand (returns false if not: the IEEE pass then redoes
// the chunk, rewriting every cell the fast one stored).
template <int S, int TO, bool F>
__device__ __forceinline__ bool vjp_pass(const View& U, const View& lam, const View& out, double dtd, double half,
                                         int seg, int nch, int ngroups, int fold, int lo_kind, int hi_kind,
                                         double (&ringU)[4][4][64], double (&ringL)[4][4][64]) {
    const int lane = threadIdx.x;
    unsigned key = 0;
    const int cg = blockIdx.x % ngroups;
    const int ch = blockIdx.x / ngroups;
    const int i = cg * 64 + lane;
    const bool active = i < U.nx;
    const int ic = active ? i : U.nx - 1;
    const int ny = U.ny;
    const int c0 = -2 + ch * seg;
    const int c1 = ch == nch - 1 ? ny + 2 : c0 + seg;
    const bool lo_tr = lo_kind == MH_BC_TRANSMISSIVE, hi_tr = hi_kind == MH_BC_TRANSMISSIVE;
    Cons acc_lo = czero(), acc_hi = czero(), held = czero();
    auto putU = [&](int s, const Cons& v) { ringU[s][0][lane] = v.r; ringU[s][1][lane] = v.mu; ringU[s][2][lane] = v.mv; ringU[s][3][lane] = v.e; };
    auto getU = [&](int s) { return Cons{ringU[s][0][lane], ringU[s][1][lane], ringU[s][2][lane], ringU[s][3][lane]}; };
    auto putL = [&](int s, const Cons& v) { ringL[s][0][lane] = v.r; ringL[s][1][lane] = v.mu; ringL[s][2][lane] = v.mv; ringL[s][3][lane] = v.e; };
    auto getL = [&](int s) { return Cons{ringL[s][0][lane], ringL[s][1][lane], ringL[s][2][lane], ringL[s][3][lane]}; };

    Prim W3 = pzero(), W2 = pzero(), W1 = pzero();   // rows p-3, p-2, p-1
    Edge Rprev;                                      // right edge state of row p-2
    Rprev.U = czero();
    Rprev.V = pzero();
    Cons bLprev = czero();                           // adjoint of the left edge state of row p-2
    Prim a0 = pzero(), a1 = pzero(), a2 = pzero();   // primitive adjoints of rows p-3, p-2, p-1
    for (int p = c0 - 3; p <= c1 + 2; ++p) {
        const int pr = min(max(p, -2), ny + 1);
        const Cons Up = ld(U, ic, pr);
        const Prim W0 = c2p<F>(Up);
        if constexpr (F) key = max(key, range_key(Up.r, Up.e, W0.p));
        const int s0 = (p + 8) & 3;
        putU(s0, Up);
        putL(s0, (p >= 0 && p < ny) ? ld(lam, ic, p) : czero());
        // primal reconstruction of row q = p-1 (feeds the faces below and above it)
        Edge Lq, Rq;
        Lq.U = czero(); Lq.V = pzero(); Rq = Lq;
        if (p >= c0 - 1 && p - 1 >= -1 && p - 1 <= ny) reconstruct<kY, F>(W2, W1, W0, half, Lq, Rq);
        // face between rows p-2 and p-1
        Cons bRf = czero(), bLf = czero();
        if (p >= c0 && p - 2 >= -1 && p - 2 <= ny - 1) {
            const Cons l1 = getL((p + 7) & 3), l2 = getL((p + 6) & 3);
            const Cons mu = {dtd * (l1.r - l2.r), dtd * (l1.mu - l2.mu), dtd * (l1.mv - l2.mv), dtd * (l1.e - l2.e)};
            Prim bRV = pzero(), bLV = pzero();
            riemann_vjp<S, kY, F>(Rprev.U, Rprev.V, Lq.U, Lq.V, mu, bRf, bRV, bLf, bLV);
            c2p_vjp<F>(Rprev.U, bRV, bRf);
            c2p_vjp<F>(Lq.U, bLV, bLf);
        }
        // reconstruction of row p-2
        if (p >= c0 + 1 && p - 2 >= -1 && p - 2 <= ny) reconstruct_vjp<kY, F>(W3, W2, W1, half, bLprev, bRf, a0, a1, a2);
        // row c = p-3 is complete
        const int c = p - 3;
        if (c >= c0 && c < c1) {
            Cons bU = (c >= 0 && c < ny) ? getL((p + 5) & 3) : czero();
            c2p_vjp<F>(getU((p + 5) & 3), a0, bU);
            if (fold < 0) {
                if (active) vjp_store<TO>(out, i, c, bU);
            } else if (c < 0) {
                if (lo_tr) acc_lo = {acc_lo.r + bU.r, acc_lo.mu + bU.mu, acc_lo.mv + bU.mv, acc_lo.e + bU.e};
                if (active) vjp_store<TO>(out, i, c, czero());
            } else if (c >= ny) {
                if (hi_tr) acc_hi = {acc_hi.r + bU.r, acc_hi.mu + bU.mu, acc_hi.mv + bU.mv, acc_hi.e + bU.e};
                if (active) vjp_store<TO>(out, i, c, czero());
                if (c == ny + 1) {
                    const Cons h = {held.r + acc_hi.r, held.mu + acc_hi.mu, held.mv + acc_hi.mv, held.e + acc_hi.e};
                    if (active) vjp_store<TO>(out, i, ny - 1, h);
                }
            } else {
                if (c == 0) bU = {bU.r + acc_lo.r, bU.mu + acc_lo.mu, bU.mv + acc_lo.mv, bU.e + acc_lo.e};
                if (c == ny - 1) held = bU;
                else if (active) vjp_store<TO>(out, i, c, bU);
            }
        }
        W3 = W2; W2 = W1; W1 = W0;
        Rprev = Rq;
        bLprev = bLf;
        a0 = a1; a1 = a2; a2 = pzero();
    }
    if constexpr (F) return !__any(key >= kRangeKeyEnd);
    return true;
}

// The chunk's fast pass where the solver has one (HLL, HLLC, FORCE) and
// dt/d is in the fast range, the IEEE pass where it does not or gave up.
template <int S, int TO>
__global__ __launch_bounds__(64) void k_vjp_y(View U, View lam, View out, double dtd, double half, int seg,
                                              int nch, int ngroups, int fold, int lo_kind, int hi_kind) {
    __shared__ double ringU[4][4][64];
    __shared__ double ringL[4][4][64];
    if constexpr (S != kExact) {
        if (dt_in_fast_range(dtd) &&
            vjp_pass<S, TO, true>(U, lam, out, dtd, half, seg, nch, ngroups, fold, lo_kind, hi_kind, ringU, ringL))
            return;
    }
    vjp_pass<S, TO, false>(U, lam, out, dtd, half, seg, nch, ngroups, fold, lo_kind, hi_kind, ringU, ringL);
}

// dst(j, i) = src(i, j) over every cell including ghosts, rho*u <-> rho*v.
// All four planes are loaded before the one barrier, so a block makes a
// single round trip to memory.
__global__ __launch_bounds__(256) void k_transpose(View src, View dst) {
    __shared__ double tile[4][32][33];
    const int tnx = src.nx + 4, tny = src.ny + 4;
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int c = 0; c < 4; ++c) {
        const int cs = c == 1 ? 2 : (c == 2 ? 1 : c);
        for (int k = ty; k < 32; k += 8) {
            const int i = bx + tx, j = by + k;
            if (i < tnx && j < tny) tile[c][k][tx] = src.o[cs * src.ps + (int64_t)(j - 2) * src.rp + (i - 2)];
        }
    }
    __syncthreads();
    for (int c = 0; c < 4; ++c)
        for (int k = ty; k < 32; k += 8) {
            const int i = by + tx, j = bx + k;   // dst column i = src row, dst row j = src column
            if (j < tnx && i < tny) dst.o[c * dst.ps + (int64_t)(j - 2) * dst.rp + (i - 2)] = tile[c][tx][k];
        }
}

__global__ __launch_bounds__(256) void k_zero_grid(View a) {
    const int64_t tnx = a.nx + 4, n = tnx * (a.ny + 4);
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(k / tnx) - 2, i = (int)(k % tnx) - 2;
        for (int c = 0; c < 4; ++c) a.o[c * a.ps + (int64_t)j * a.rp + i] = 0.0;
    }
}

__global__ __launch_bounds__(256) void k_copy_grid(View src, View dst) {
    const int64_t tnx = src.nx + 4, n = tnx * (src.ny + 4);
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(k / tnx) - 2, i = (int)(k % tnx) - 2;
        for (int c = 0; c < 4; ++c) dst.o[c * dst.ps + (int64_t)j * dst.rp + i] = src.o[c * src.ps + (int64_t)j * src.rp + i];
    }
}

}  // namespace mh

using namespace mh;

struct mh_adjoint_s {
    mh_grid_desc d{};
    int solver = 1;
    Bc bc{};
    mh_bc bc_c{};
    int K = 20;
    int max_steps = 0;
    mh_grid_s* T = nullptr;      // x-sweep output (single-step products)
    mh_grid_s* Ut = nullptr;     // transposed state
    mh_grid_s* bTt = nullptr;    // transposed adjoint of T
    mh_grid_s* work = nullptr;   // forward working state
    mh_grid_s* A0 = nullptr;     // adjoint ping-pong
    mh_grid_s* A1 = nullptr;
    std::vector<mh_grid_s*> ckpt;
    // Two buffers of segment storage: the states of a segment, their
    // x-sweep outputs and their transposes.  The backward pass recomputes
    // segment s-1 into one buffer on `aux` while the products of segment s
    // read the other on the caller's stream.
    std::vector<mh_grid_s*> segs[2], segT[2], segUt[2];
    hipStream_t aux = nullptr;
    hipEvent_t rec_done[2] = {nullptr, nullptr}, vjp_done[2] = {nullptr, nullptr};
    std::vector<double> dts;
    int nsteps = 0;
    // device-resident loop over `work` for forward runs with CFL time steps
    // (dt from the speed maximum fused into the y sweep: no host round trip
    // per step)
    mh_stepper stepper = nullptr;
};

namespace {

int new_grid(const mh_grid_desc& d, mh_grid_s** g) { return mh_grid_create(&d, g); }

mh_grid_desc transposed(const mh_grid_desc& d) {
    mh_grid_desc t = d;
    t.nx = d.ny; t.ny = d.nx;
    t.x_min = d.y_min; t.x_max = d.y_max; t.y_min = d.x_min; t.y_max = d.x_max;
    return t;
}

int launch_copy(const mh_grid_s* s, mh_grid_s* d, hipStream_t st) {
    hipLaunchKernelGGL(k_copy_grid, dim3(1024), dim3(256), 0, st, s->view(), d->view());
    return check_hip(hipGetLastError(), "copy launch");
}

int launch_transpose(const mh_grid_s* s, mh_grid_s* d, hipStream_t st) {
    const dim3 g((s->d.nx + 4 + 31) / 32, (s->d.ny + 4 + 31) / 32);
    hipLaunchKernelGGL(k_transpose, g, dim3(256), 0, st, s->view(), d->view());
    return check_hip(hipGetLastError(), "transpose launch");
}

// Rows per chunk of k_vjp_y: enough chunks for 2048 wavefronts (two per
// SIMD of 256 CUs), 3..256 rows each.  Three is the fewest the folds allow
// (ghost rows -2, -1 fold into row 0 and ny, ny+1 into ny-1 within one
// chunk: the first chunk holds rows -2..0, the last at least 3 rows); small
// grids are latency-bound, so shorter chunks (3 rows of output for 9 row
// iterations, against 4 for 10) finish sooner.
int vjp_seg(const mh_grid_desc& d) {
    const int ngroups = (d.nx + 63) / 64, rows = d.ny + 4;
    const int chunks = (2048 + ngroups - 1) / ngroups;
    return std::min(256, std::max(3, (rows + chunks - 1) / chunks));
}

// fold < 0: ghost-row adjoints written as they are; else folded by the
// edge kinds lo (bottom of the swept direction) and hi.  to: write `out`
// transposed (out has the transposed shape).
int launch_vjp_y(const mh_grid_s* U, const mh_grid_s* lam, mh_grid_s* out, double dtd, int solver, int fold, int lo,
                 int hi, int to, hipStream_t st) {
    const int ngroups = (U->d.nx + 63) / 64;
    const int seg = vjp_seg(U->d);
    const int nch = std::max(1, (U->d.ny + 4) / seg);
    const dim3 g((unsigned)(ngroups * nch));
    const double half = 0.5 * dtd;
    const View u = U->view(), l = lam->view(), o = out->view();
    if (solver == kHLLC) {
        if (to) hipLaunchKernelGGL((k_vjp_y<kHLLC, 1>), g, dim3(64), 0, st, u, l, o, dtd, half, seg, nch, ngroups, fold, lo, hi);
        else hipLaunchKernelGGL((k_vjp_y<kHLLC, 0>), g, dim3(64), 0, st, u, l, o, dtd, half, seg, nch, ngroups, fold, lo, hi);
    } else if (solver == kExact) {
        if (to) hipLaunchKernelGGL((k_vjp_y<kExact, 1>), g, dim3(64), 0, st, u, l, o, dtd, half, seg, nch, ngroups, fold, lo, hi);
        else hipLaunchKernelGGL((k_vjp_y<kExact, 0>), g, dim3(64), 0, st, u, l, o, dtd, half, seg, nch, ngroups, fold, lo, hi);
    } else if (solver == kForce) {
        if (to) hipLaunchKernelGGL((k_vjp_y<kForce, 1>), g, dim3(64), 0, st, u, l, o, dtd, half, seg, nch, ngroups, fold, lo, hi);
        else hipLaunchKernelGGL((k_vjp_y<kForce, 0>), g, dim3(64), 0, st, u, l, o, dtd, half, seg, nch, ngroups, fold, lo, hi);
    } else {
        if (to) hipLaunchKernelGGL((k_vjp_y<kHLL, 1>), g, dim3(64), 0, st, u, l, o, dtd, half, seg, nch, ngroups, fold, lo, hi);
        else hipLaunchKernelGGL((k_vjp_y<kHLL, 0>), g, dim3(64), 0, st, u, l, o, dtd, half, seg, nch, ngroups, fold, lo, hi);
    }
    return check_hip(hipGetLastError(), "vjp launch");
}

// lam_out = (dU'/dU)^T lam_in for one split step U -> T -> U' at time step
// dt, given Ut = U transposed (with the x-ghost columns of U) and
// T = xsweep(U) with its y-ghost rows (the cells the two sweeps read).  The
// y product writes the adjoint of T transposed, the x product reads it there
// and writes lam_out back in the grid's own layout.  lam_out receives its
// interior rows (x-ghost columns zero); its y-ghost rows are not written.
int step_vjp_t(mh_adjoint a, const mh_grid_s* Ut, const mh_grid_s* T, double dt, const mh_grid_s* lam_in,
               mh_grid_s* lam_out, hipStream_t st) {
    int rc = launch_vjp_y(T, lam_in, a->bTt, dt / T->dy, a->solver, 1, a->bc.kind[MH_EDGE_BOTTOM],
                          a->bc.kind[MH_EDGE_TOP], 1, st);
    if (!rc) rc = launch_vjp_y(Ut, a->bTt, lam_out, dt / T->dx, a->solver, 1, a->bc.kind[MH_EDGE_LEFT],
                               a->bc.kind[MH_EDGE_RIGHT], 1, st);
    return rc;
}

// The same from U alone: fills U's ghosts, recomputes T and transposes U.
int step_vjp(mh_adjoint a, mh_grid_s* U, double dt, const mh_grid_s* lam_in, mh_grid_s* lam_out, hipStream_t st) {
    int rc = launch_fill(U, a->bc, st);
    if (!rc) rc = launch_sweep_host(U, a->T, kX, a->solver, dt, a->bc, 1, st);
    if (!rc) rc = launch_transpose(U, a->Ut, st);
    if (!rc) rc = step_vjp_t(a, a->Ut, a->T, dt, lam_in, lam_out, st);
    return rc;
}

// Recomputes segment s from its checkpoint (ghosts filled) into buffer b on
// stream st: every state's x-sweep output and transpose, and the next state.
// The sweeps' fused ghost writes give T its y-ghost rows and the next state
// its x-ghost columns, the cells the products read.
int recompute_segment(mh_adjoint a, int s, int b, hipStream_t st) {
    const int k0 = s * a->K, k1 = std::min(a->nsteps, k0 + a->K);
    int rc = launch_copy(a->ckpt[s], a->segs[b][0], st);
    for (int n = k0; n < k1 && !rc; ++n) {
        const int i = n - k0;
        rc = launch_sweep_host(a->segs[b][i], a->segT[b][i], kX, a->solver, a->dts[n], a->bc, 1, st);
        if (!rc) rc = launch_transpose(a->segs[b][i], a->segUt[b][i], st);
        if (!rc && n + 1 < k1)
            rc = launch_sweep_host(a->segT[b][i], a->segs[b][i + 1], kY, a->solver, a->dts[n], a->bc, 1, st);
    }
    if (!rc) rc = check_hip(hipEventRecord(a->rec_done[b], st), "hipEventRecord");
    return rc;
}

bool same(const mh_grid_s* g, const mh_grid_desc& d) { return g && g->d.nx == d.nx && g->d.ny == d.ny; }

}  // namespace

extern "C" {

int mh_adjoint_create(const mh_grid_desc* desc, int solver, const mh_bc* bc, int checkpoint_every, int max_steps,
                      mh_adjoint* out) {
    if (!desc || !out || checkpoint_every < 1 || max_steps < 0) return fail(MH_ERR_INVALID, "mh_adjoint_create: bad argument");
    if (solver < kHLL || solver > kForce) return fail(MH_ERR_INVALID, "mh_adjoint_create: unknown solver");
    if (bc)
        for (int e = 0; e < 4; ++e)
            if (bc->kind[e] != MH_BC_TRANSMISSIVE && bc->kind[e] != MH_BC_INFLOW)
                return fail(MH_ERR_INVALID, "mh_adjoint_create: unknown boundary kind");
    mh_adjoint a = new (std::nothrow) mh_adjoint_s();
    if (!a) return fail(MH_ERR_OOM, "host allocation failed");
    a->d = *desc;
    a->solver = solver;
    a->bc = to_bc(bc);
    if (bc) a->bc_c = *bc;
    a->K = checkpoint_every;
    a->max_steps = max_steps;
    const mh_grid_desc t = transposed(*desc);
    int rc = new_grid(*desc, &a->T);
    if (!rc) rc = new_grid(t, &a->Ut);
    if (!rc) rc = new_grid(t, &a->bTt);
    if (!rc) rc = new_grid(*desc, &a->work);
    if (!rc) rc = new_grid(*desc, &a->A0);
    if (!rc) rc = new_grid(*desc, &a->A1);
    const int nck = max_steps > 0 ? (max_steps + a->K - 1) / a->K : 0;
    for (int k = 0; k < nck && !rc; ++k) {
        mh_grid_s* g = nullptr;
        rc = new_grid(*desc, &g);
        a->ckpt.push_back(g);
    }
    const int nseg_bufs = max_steps > a->K ? 2 : (max_steps > 0 ? 1 : 0);
    for (int b = 0; b < nseg_bufs && !rc; ++b)
        for (int k = 0; k < std::min(a->K, max_steps) && !rc; ++k) {
            mh_grid_s* g = nullptr;
            rc = new_grid(*desc, &g);
            a->segs[b].push_back(g);
            g = nullptr;
            if (!rc) rc = new_grid(*desc, &g);
            a->segT[b].push_back(g);
            g = nullptr;
            if (!rc) rc = new_grid(t, &g);
            a->segUt[b].push_back(g);
        }
    if (!rc) rc = check_hip(hipStreamCreateWithFlags(&a->aux, hipStreamNonBlocking), "hipStreamCreate");
    for (int b = 0; b < 2 && !rc; ++b) {
        rc = check_hip(hipEventCreateWithFlags(&a->rec_done[b], hipEventDisableTiming), "hipEventCreate");
        if (!rc) rc = check_hip(hipEventCreateWithFlags(&a->vjp_done[b], hipEventDisableTiming), "hipEventCreate");
    }
    if (!rc) rc = mh_stepper_create(a->work, solver, 1.0, 0.0, bc, std::max(max_steps, 1), &a->stepper, nullptr);
    if (!rc) stepper_unfuse(a->stepper);
    // Cells no kernel writes (the ghost rows of the ping-pong grids) read
    // as zero.
    for (mh_grid_s* g : {a->bTt, a->A0, a->A1})
        if (!rc) rc = mh_grid_zero(g, nullptr);
    if (!rc) rc = check_hip(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
    if (rc) { mh_adjoint_destroy(a); return rc; }
    *out = a;
    return MH_OK;
}

int mh_adjoint_destroy(mh_adjoint a) {
    if (!a) return MH_OK;
    if (a->aux) (void)hipStreamSynchronize(a->aux);
    if (a->stepper) mh_stepper_destroy(a->stepper);
    for (mh_grid_s* g : {a->T, a->Ut, a->bTt, a->work, a->A0, a->A1})
        if (g) mh_grid_destroy(g);
    for (mh_grid_s* g : a->ckpt) if (g) mh_grid_destroy(g);
    for (int b = 0; b < 2; ++b) {
        for (mh_grid_s* g : a->segs[b]) if (g) mh_grid_destroy(g);
        for (mh_grid_s* g : a->segT[b]) if (g) mh_grid_destroy(g);
        for (mh_grid_s* g : a->segUt[b]) if (g) mh_grid_destroy(g);
        if (a->rec_done[b]) (void)hipEventDestroy(a->rec_done[b]);
        if (a->vjp_done[b]) (void)hipEventDestroy(a->vjp_done[b]);
    }
    if (a->aux) (void)hipStreamDestroy(a->aux);
    delete a;
    return MH_OK;
}

int mh_adjoint_step_vjp(mh_adjoint a, mh_grid U, double dt, mh_grid lam_in, mh_grid lam_out, mh_stream stream) {
    if (!a || !U || !lam_in || !lam_out || lam_in == lam_out) return fail(MH_ERR_INVALID, "mh_adjoint_step_vjp: bad argument");
    if (!same(U, a->d) || !same(lam_in, a->d) || !same(lam_out, a->d))
        return fail(MH_ERR_INVALID, "mh_adjoint_step_vjp: shape mismatch");
    return step_vjp(a, U, dt, lam_in, lam_out, (hipStream_t)stream);
}

int mh_adjoint_sweep_vjp(mh_adjoint a, mh_grid U, double dt, int dir, mh_grid lam_in, mh_grid lam_out,
                         mh_stream stream) {
    if (!a || !U || !lam_in || !lam_out || lam_in == lam_out || (dir != 0 && dir != 1))
        return fail(MH_ERR_INVALID, "mh_adjoint_sweep_vjp: bad argument");
    if (!same(U, a->d) || !same(lam_in, a->d) || !same(lam_out, a->d))
        return fail(MH_ERR_INVALID, "mh_adjoint_sweep_vjp: shape mismatch");
    hipStream_t st = (hipStream_t)stream;
    // k_vjp_y writes only the cells the sweep reads; the rest of lam_out is zero.
    int rc = mh_grid_zero(lam_out, stream);
    if (dir == kY) {
        if (!rc) rc = launch_vjp_y(U, lam_in, lam_out, dt / U->dy, a->solver, -1, 0, 0, 0, st);
        return rc;
    }
    if (!rc) rc = launch_transpose(U, a->Ut, st);
    if (!rc) rc = launch_transpose(lam_in, a->bTt, st);
    if (!rc) rc = launch_vjp_y(a->Ut, a->bTt, lam_out, dt / U->dx, a->solver, -1, 0, 0, 1, st);
    return rc;
}

int mh_adjoint_forward(mh_adjoint a, mh_grid U0, int nsteps, double cfl, const double* dt_in, double* dt_out,
                       mh_grid UN, mh_stream stream) {
    if (!a || !U0 || nsteps < 0 || nsteps > a->max_steps) return fail(MH_ERR_INVALID, "mh_adjoint_forward: bad argument");
    if (!same(U0, a->d) || (UN && !same(UN, a->d))) return fail(MH_ERR_INVALID, "mh_adjoint_forward: shape mismatch");
    if (!dt_in && !(cfl > 0.0)) return fail(MH_ERR_INVALID, "mh_adjoint_forward: cfl must be positive");
    hipStream_t st = (hipStream_t)stream;
    std::vector<double> dts(nsteps, 0.0);
    // Each step is the x sweep work -> T and the y sweep T -> work; their
    // fused ghost writes supply every ghost cell the next sweep reads (T's
    // y-ghost rows, work's x-ghost columns), which are also all the
    // checkpoints' recompute and the products read.
    int rc = launch_copy(U0, a->work, st);
    if (dt_in) {
        if (!rc) rc = launch_fill(a->work, a->bc, st);
        for (int n = 0; n < nsteps && !rc; ++n) {
            if (n % a->K == 0) rc = launch_copy(a->work, a->ckpt[n / a->K], st);
            if (!rc) rc = launch_sweep_host(a->work, a->T, kX, a->solver, dt_in[n], a->bc, 1, st);
            if (!rc) rc = launch_sweep_host(a->T, a->work, kY, a->solver, dt_in[n], a->bc, 1, st);
            dts[n] = dt_in[n];
        }
    } else {
        // dt = cfl*min(dx,dy)/max speed on the device (the arithmetic of
        // mh_compute_dt), logged per step and read back once at the end
        if (!rc) rc = stepper_start(a->stepper, cfl, st);
        for (int n = 0; n < nsteps && !rc; ++n) {
            if (n % a->K == 0) rc = launch_copy(a->work, a->ckpt[n / a->K], st);
            if (!rc) rc = mh_stepper_advance(a->stepper, 1, stream);
        }
        double t = 0.0;
        int done = 0;
        if (!rc) rc = mh_stepper_clock(a->stepper, &t, &done, dts.data(), stream);
        if (!rc && done != nsteps) rc = fail(MH_ERR_LAUNCH, "mh_adjoint_forward: the device loop stopped early");
    }
    if (!rc && UN) rc = launch_fill(a->work, a->bc, st);
    if (!rc && UN) rc = launch_copy(a->work, UN, st);
    if (!rc) rc = check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (rc) return rc;
    a->dts = dts;
    a->nsteps = nsteps;
    if (dt_out) std::memcpy(dt_out, a->dts.data(), sizeof(double) * nsteps);
    return MH_OK;
}

int mh_adjoint_backward(mh_adjoint a, mh_grid lamN, mh_grid lam0, mh_stream stream) {
    if (!a || !lamN || !lam0) return fail(MH_ERR_INVALID, "mh_adjoint_backward: bad argument");
    if (!same(lamN, a->d) || !same(lam0, a->d)) return fail(MH_ERR_INVALID, "mh_adjoint_backward: shape mismatch");
    hipStream_t st = (hipStream_t)stream;
    const int N = a->nsteps;
    int rc = launch_copy(lamN, a->A0, st);
    mh_grid_s* cur = a->A0;
    mh_grid_s* nxt = a->A1;
    const int nseg = (N + a->K - 1) / a->K;
    const int nbuf = a->segs[1].empty() ? 1 : 2;
    bool used[2] = {false, false};
    // The aux stream must not start before the forward pass's checkpoints
    // and lamN are in place: it waits for the caller's stream once.
    if (!rc && nseg > 0) rc = check_hip(hipEventRecord(a->vjp_done[0], st), "hipEventRecord");
    if (!rc && nseg > 0) rc = check_hip(hipStreamWaitEvent(a->aux, a->vjp_done[0], 0), "hipStreamWaitEvent");
    if (!rc && nseg > 0) rc = recompute_segment(a, nseg - 1, (nseg - 1) % nbuf, a->aux);
    for (int s = nseg - 1; s >= 0 && !rc; --s) {
        const int b = s % nbuf;
        const int k0 = s * a->K, k1 = std::min(N, k0 + a->K);
        rc = check_hip(hipStreamWaitEvent(st, a->rec_done[b], 0), "hipStreamWaitEvent");
        if (!rc && s > 0 && nbuf == 2) {
            // segment s-1 into the other buffer, once segment s+1's products
            // (the buffer's last readers) are done
            const int pb = (s - 1) % 2;
            if (used[pb]) rc = check_hip(hipStreamWaitEvent(a->aux, a->vjp_done[pb], 0), "hipStreamWaitEvent");
            if (!rc) rc = recompute_segment(a, s - 1, pb, a->aux);
        }
        for (int n = k1 - 1; n >= k0 && !rc; --n) {
            rc = step_vjp_t(a, a->segUt[b][n - k0], a->segT[b][n - k0], a->dts[n], cur, nxt, st);
            std::swap(cur, nxt);
        }
        if (!rc) rc = check_hip(hipEventRecord(a->vjp_done[b], st), "hipEventRecord");
        used[b] = true;
        if (!rc && s > 0 && nbuf == 1) {
            // one buffer: recompute the next segment after these products
            rc = check_hip(hipStreamWaitEvent(a->aux, a->vjp_done[b], 0), "hipStreamWaitEvent");
            if (!rc) rc = recompute_segment(a, s - 1, 0, a->aux);
        }
    }
    if (!rc) rc = launch_copy(cur, lam0, st);
    if (!rc) rc = check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (!rc) rc = check_hip(hipStreamSynchronize(a->aux), "hipStreamSynchronize");
    return rc;
}

int mh_adjoint_dt(mh_adjoint a, double* dt, int* nsteps) {
    if (!a) return fail(MH_ERR_INVALID, "mh_adjoint_dt: null adjoint");
    if (nsteps) *nsteps = a->nsteps;
    if (dt) std::memcpy(dt, a->dts.data(), sizeof(double) * a->nsteps);
    return MH_OK;
}

int mh_grid_zero(mh_grid g, mh_stream stream) {
    if (!g) return fail(MH_ERR_INVALID, "mh_grid_zero: null grid");
    hipLaunchKernelGGL(k_zero_grid, dim3(1024), dim3(256), 0, (hipStream_t)stream, g->view());
    return check_hip(hipGetLastError(), "zero launch");
}

int mh_grid_copy(mh_grid src, mh_grid dst, mh_stream stream) {
    if (!src || !dst || src->d.nx != dst->d.nx || src->d.ny != dst->d.ny) return fail(MH_ERR_INVALID, "mh_grid_copy: bad argument");
    return launch_copy(src, dst, (hipStream_t)stream);
}

}  // extern "C"
