// sweep_tiles.h -- one wavefront's share of an x or y sweep, shared by the
// single-grid kernels (solver.hip) and the persistent batch kernel
// (batch.hip).
//
// y tile: 64 adjacent columns x a chunk of rows.  Each lane owns a column and
//   marches up the chunk with a register window (rows p-2, p-1, p): at row p
//   it converts row p to primitive variables, reconstructs row p-1, solves the
//   face between rows p-2 and p-1 and updates row p-2.  Loads and stores are
//   64 consecutive doubles per plane (one 512-byte segment).  Two warm-up rows
//   below the chunk and two above are read, nothing is recomputed inside it.
// x tile: one row segment.  The wavefront marches along the segment 64 cells
//   at a time with the same two-cell lag; values of the left neighbours cross
//   lanes through a per-wavefront LDS slice (2 carry slots + 64 lanes).
//
// Both tiles synchronise only their own wavefront, so several independent
// tiles can share a workgroup.
#pragma once

#include "euler_device.h"
#include "grid.h"

namespace mh {

__device__ __forceinline__ Cons ld(const View& v, int i, int j) {
    const double* q = v.o + (int64_t)j * v.rp + i;
    return {q[0], q[v.ps], q[2 * v.ps], q[3 * v.ps]};
}
__device__ __forceinline__ void st(const View& v, int i, int j, const Cons& U) {
    double* q = v.o + (int64_t)j * v.rp + i;
    q[0] = U.r;
    q[v.ps] = U.mu;
    q[2 * v.ps] = U.mv;
    q[3 * v.ps] = U.e;
}

// Row access through buffer descriptors: a wavefront-uniform descriptor per
// plane whose base is the row's cell i0 (scalar arithmetic, rebuilt per
// row) and a per-lane 32-bit byte offset that stays in one register for the
// whole tile, so no vector instruction forms an address.  The descriptors
// span 2^31 bytes from the row (no range check is ever hit: offsets stay
// within the row and its padding).
typedef unsigned int mh_u32x2 __attribute__((ext_vector_type(2)));
constexpr int kRsrcWord3 = 0x00020000;   // raw buffer, 32-bit data format (gfx9)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc_at(const double* p) {
    // every address passed here is the same in all lanes of the wavefront
    // (a row of the tile); saying so keeps the descriptor in scalar
    // registers where the compiler cannot prove it
    const uint64_t a = reinterpret_cast<uint64_t>(p);
    const uint64_t u = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(a >> 32)) << 32) |
                       (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)a);
    return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>(u), (short)0, 0x7fffffff, kRsrcWord3);
}
__device__ __forceinline__ double bld(const double* base, uint32_t off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc_at(base), off, 0, 0));
}
__device__ __forceinline__ void bst(const double* base, uint32_t off, double x) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(mh_u32x2, x), rsrc_at(base), off, 0, 0);
}
__device__ __forceinline__ const double* row_at(const View& v, int i0, int j) {
    return v.o + (int64_t)j * v.rp + i0;
}
__device__ __forceinline__ Cons ld_off(const View& v, const double* row, uint32_t off) {
    return {bld(row, off), bld(row + v.ps, off), bld(row + 2 * v.ps, off), bld(row + 3 * v.ps, off)};
}
__device__ __forceinline__ void st_off(const View& v, const double* row, uint32_t off, const Cons& U) {
    bst(row, off, U.r);
    bst(row + v.ps, off, U.mu);
    bst(row + 2 * v.ps, off, U.mv);
    bst(row + 3 * v.ps, off, U.e);
}

// Every lane of a tile issues its stores, also a lane with no cell to write:
// that lane writes into column kSinkCol of the row, which lies in the row's
// lead padding (left of the two ghost columns) and is never read.  With no
// branch around the stores, the loads prefetched before them can be waited
// for without waiting for the stores (the vector memory counter counts both
// and completes in order: a skipped store block would leave the load the
// newest event and force a full wait every row).
constexpr int kSinkCol = -kNg - 1;
static_assert(kLead >= kNg + 1, "the sink column must lie in the lead padding");

// Boundary spec in kernel-argument form.
struct Bc {
    int kind[4];
    Cons s[4];
};

// Orders this wavefront's LDS accesses across its lanes.  No other wavefront
// touches its slice and a wavefront's LDS instructions execute in issue
// order, so only compiler reordering has to be prevented: wavefront-scope
// fences around a wave barrier do that without waiting on global loads in
// flight (a workgroup-scope fence could add a vmcnt wait).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The exact solver's fast pass gives up on the tile at the first row or
// pass with a face it leaves to the IEEE pass (solver_gave_up), which
// redoes the whole tile: what the fast pass would compute after it is
// wasted.
template <int S, bool F> __device__ __forceinline__ bool give_up_now(int it) {
    if constexpr (F && S == kExact) return __any(solver_gave_up(it));
    return false;
}

__device__ __forceinline__ double wave_max(double m) {
    for (int off = 32; off >= 1; off >>= 1) m = smax(m, __shfl_xor(m, off, 64));
    return m;
}

// One wavefront's LDS exchange slice (10560 bytes): per exchanged quantity
// 66 slots, slots 0-1 carrying the values of lanes 62-63 of the previous
// iteration and slot 2+l holding lane l's, so lane l reads its left
// neighbours at slots l+1 and l without any lane-dependent select.
constexpr int kSlots = 66;
struct XLds66 {
    double W[4][kSlots];   // primitive state of the loaded cell
    double R[8][kSlots];   // right edge state (U 0-3, V 4-7)
    double U[4][kSlots];   // conserved state of the loaded cell
    double F[4][kSlots];   // face flux
};

// Compact slice (6464 bytes): `M` holds the current 64 lanes' values of the
// quantity being exchanged (the phases of an iteration are separated by
// wave_sync, so they share it); each quantity keeps the values the next
// iteration needs from lanes 62-63 in its own carry slots, and lanes 0-1
// select between neighbour and carry.  Fewer bytes, more selects: which
// layout is faster depends on occupancy, so both exist.  The deferred-update
// x tile (x_tiled_pass, the default) uses this slice too.
struct XLdsC {
    double M[8][64];
    double U[4][64];
    double cW[4][2];
    double cU[4][2];
    double cR[8];
    double cF[4];
};

// y tile: three rows of four planes, 64 lanes (768 doubles), then each
// lane's range key (low word of slot 768 + lane)
constexpr int kRing = 832;

// Fast-path selection shared by the tiles: a tile's results computed with
// div_nr / sqrt_nr equal the IEEE ones when every state the tile reads lies
// in the fast range (range_key < kRangeKeyEnd) and dt/d is within
// [2^-20, 2^20] (euler_device.h gives the bound).  dt/d is tested before the
// tile; the states' keys are accumulated as they are read and voted on once,
// after the tile.
__device__ __forceinline__ bool dt_in_fast_range(double dtd) { return dtd >= 0x1p-20 && dtd <= 0x1p20; }

// One pass of a y tile with the fast (F = true) or IEEE (F = false)
// division and square root.  The fast pass runs the whole tile
// speculatively: each row's range key is folded into a lane-local maximum
// (no vote, no exit inside the loop) and one wavefront vote at the end says
// whether the tile's results stand (returns false if not).  Then the IEEE
// pass must redo the tile: it reads only the input grid and rewrites every
// cell and ghost the fast pass stored, and the fast pass's speed maximum is
// dropped.  The IEEE pass stays out of the fast pass's loop (and, for the
// single-grid sweep, out of its kernel), which keeps the hot loop lean.
// B: rows through buffer descriptors and the Riemann solve on the edge
// states' shared reciprocals (the single-grid sweeps); without it, 64-bit
// lane addresses and the solver's own divisions (the batch kernels, whose
// register budget the shared reciprocals exceed).  The results are the same.
template <int S, bool F, bool B = true>
__device__ __forceinline__ bool y_tile_pass(const View& in, const View& out, int cg, int j0, int j1, double dtd,
                                            double half, const Bc& bc, int ghosts, int speed, double* ring, int lane,
                                            double& m) {
    const int i = cg * 64 + lane;
    const bool active = i < in.nx;
    const int ic = active ? i : in.nx - 1;
    // loads: cell ic of the row from column 0; stores: the lane's cell, or
    // the sink column for a lane past the row's end, from the sink column
    const uint32_t lo = (uint32_t)ic * (uint32_t)sizeof(double);
    const uint32_t so = (uint32_t)(active ? i - kSinkCol : 0) * (uint32_t)sizeof(double);
    auto load_row = [&](int r) {
        if constexpr (B) return ld_off(in, row_at(in, 0, r), lo);
        else return ld(in, ic, r);
    };
    int it = 0;
    auto put = [&](int slot, const Cons& U) {
        double* q = ring + slot * 256 + lane;
        q[0] = U.r; q[64] = U.mu; q[128] = U.mv; q[192] = U.e;
    };
    auto get = [&](int slot) {
        const double* q = ring + slot * 256 + lane;
        return Cons{q[0], q[64], q[128], q[192]};
    };
    // converts row state U to W, folding its range key into the lane's key
    // slot -- in LDS: one more register would take the tile past 128 VGPRs
    unsigned* key = reinterpret_cast<unsigned*>(ring + 768 + lane);
    if constexpr (F) __hip_atomic_store(key, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    auto conv = [&](const Cons& U, Prim& W) {
        W = c2p<F>(U);
        if constexpr (F)
            __hip_atomic_fetch_max(key, range_key(U.r, U.e, W.p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    };
    m = 0.0;
    // rows j0-2 and j0-1: convert only.  The window carries W of row p-1
    // and the difference dm = W(p-1) - W(p-2), which is all the limiter
    // takes of row p-2.
    Cons U0 = load_row(j0 - 2);
    Prim dm, W1, W0;
    conv(U0, W0);
    put(0, U0);
    U0 = load_row(j0 - 1);
    conv(U0, W1);
    put(1, U0);
    dm = diff(W1, W0);
    Cons Un = load_row(j0);
    Edge L, R, Rp;
    {   // row p = j0: reconstruct j0-1
        U0 = Un;
        Un = load_row(j0 + 1);
        conv(U0, W0);
        put(2, U0);
        dm = reconstruct_d<kY, F>(dm, W1, W0, half, L, Rp);
        W1 = W0;
    }
    Cons Fp;
    {   // row p = j0+1: reconstruct j0, face j0-1/2
        U0 = Un;
        Un = load_row(j0 + 2);
        conv(U0, W0);
        put(0, U0);   // row j0-2 is no longer needed
        dm = reconstruct_d<kY, F>(dm, W1, W0, half, L, R);
        Fp = riemann<S, kY, F, B>(Rp.U, Rp.V, L.U, L.V, it, Rp.rr, L.rr);
        if (give_up_now<S, F>(it)) return false;
        Rp = R;
        W1 = W0;
    }
    // rows p = j0+2 .. j1+1; the ring slot of row p is (p - j0 + 2) % 3
    int slot_new = 1, slot_old = 2;   // slots of rows p and p-2
    for (int p = j0 + 2; p <= j1 + 1; ++p) {
        // Row p goes into the ring slot of row p-3 (updated already) before
        // the next row's load, so that load may reuse Un's registers.
        conv(Un, W0);
        put(slot_new, Un);
        Un = load_row(min(p + 1, j1 + 1));
        dm = reconstruct_d<kY, F>(dm, W1, W0, half, L, R);
        const Cons Fc = riemann<S, kY, F, B>(Rp.U, Rp.V, L.U, L.V, it, Rp.rr, L.rr);
        if (give_up_now<S, F>(it)) return false;
        const int r = p - 2;
        const Cons Unew = update(get(slot_old), dtd, Fp, Fc);
        if constexpr (B) st_off(out, row_at(out, kSinkCol, r), so, Unew);
        else st(out, active ? i : kSinkCol, r, Unew);
        if (active) {
            // From rows in range, the new state's rho and p are zero or within
            // about 2^-150..2^400 (differences of in-range values cannot fall
            // below their unit in the last place), so its signal speed's
            // 1/rho, gamma*p/rho and square root stay exact with the fast
            // sequences (signal_speed<true> takes the square root of zero
            // apart); a NaN speed never enters the maximum.
            if (speed) m = smax(m, signal_speed<F>(Unew));
            if (ghosts) {
                if (i == 0) {
                    const Cons G = bc.kind[MH_EDGE_LEFT] == MH_BC_INFLOW ? bc.s[MH_EDGE_LEFT] : Unew;
                    st(out, -1, r, G);
                    st(out, -2, r, G);
                }
                if (i == in.nx - 1) {
                    const Cons G = bc.kind[MH_EDGE_RIGHT] == MH_BC_INFLOW ? bc.s[MH_EDGE_RIGHT] : Unew;
                    st(out, in.nx, r, G);
                    st(out, in.nx + 1, r, G);
                }
            }
        }
        Fp = Fc;
        Rp = R;
        W1 = W0;
        slot_old = slot_old == 2 ? 0 : slot_old + 1;
        slot_new = slot_new == 2 ? 0 : slot_new + 1;
    }
    if constexpr (F)
        return !__any(__hip_atomic_load(key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) >= kRangeKeyEnd ||
                      solver_gave_up(it));
    return true;
}

// y tile over columns [64*cg, 64*cg+64) and rows [j0, j1).  Returns the
// lane's running max of the output's signal speed when `speed` is set
// (std::max semantics: a NaN speed never enters), 0 otherwise.  Each lane
// marches up its column; the conserved states of rows p-1 and p-2 wait for
// the update in a three-row LDS ring (each lane touches only its own
// column, so program order is the only ordering needed); `ring` is the
// wavefront's kRing-double LDS slice.  The fast pass runs first when dt/d is
// in range (see y_tile_pass).
template <int S, bool B = true>
__device__ __forceinline__ double y_tile(const View& in, const View& out, int cg, int j0, int j1, double dtd,
                                         double half, const Bc& bc, int ghosts, int speed, double* ring, int lane) {
    double m = 0.0;
    if (dt_in_fast_range(dtd) &&
        y_tile_pass<S, true, B>(in, out, cg, j0, j1, dtd, half, bc, ghosts, speed, ring, lane, m))
        return m;
    y_tile_pass<S, false, B>(in, out, cg, j0, j1, dtd, half, bc, ghosts, speed, ring, lane, m);
    return m;
}

// x tile over row j, cells [i0, i1).  The lane holding cell p at iteration
// k reconstructs cell p-1, solves face (p-2)+1/2 and updates cell p-2; the
// W and U of cells p-1, p-2 and the right edge state and flux of the left
// neighbour come from lanes l-1, l-2 through LDS (the carry slots for lanes
// 0 and 1).
template <int S, bool F>
__device__ __forceinline__ bool x_tile66_pass(const View& in, const View& out, int j, int i0, int i1, double dtd,
                                       double half, const Bc& bc, int ghosts, XLds66& X, int lane) {
    int it = 0;
    const int niter = (i1 - i0 + 4 + 63) / 64;
    unsigned key = 0;   // running range key of the states read (fast pass)
    Cons Un = ld(in, min(i0 - 2 + lane, in.nx + 1), j);
    for (int k = 0; k < niter; ++k) {
        const int p = i0 - 2 + 64 * k + lane;
        const Cons U0 = Un;
        Un = ld(in, min(p + 64, in.nx + 1), j);
        const Prim W0 = c2p<F>(U0);
        if constexpr (F) key = max(key, range_key(U0.r, U0.e, W0.p));
        X.W[0][2 + lane] = W0.r; X.W[1][2 + lane] = W0.u; X.W[2][2 + lane] = W0.v; X.W[3][2 + lane] = W0.p;
        X.U[0][2 + lane] = U0.r; X.U[1][2 + lane] = U0.mu; X.U[2][2 + lane] = U0.mv; X.U[3][2 + lane] = U0.e;
        wave_sync();
        const Prim W1 = {X.W[0][1 + lane], X.W[1][1 + lane], X.W[2][1 + lane], X.W[3][1 + lane]};
        const Prim W2 = {X.W[0][lane], X.W[1][lane], X.W[2][lane], X.W[3][lane]};
        wave_sync();
        if (lane >= 62) {
            X.W[0][lane - 62] = W0.r; X.W[1][lane - 62] = W0.u; X.W[2][lane - 62] = W0.v; X.W[3][lane - 62] = W0.p;
        }
        Edge L, R;   // reconstruct q = p-1
        reconstruct<kX, F>(W2, W1, W0, half, L, R);
        X.R[0][2 + lane] = R.U.r; X.R[1][2 + lane] = R.U.mu; X.R[2][2 + lane] = R.U.mv; X.R[3][2 + lane] = R.U.e;
        X.R[4][2 + lane] = R.V.r; X.R[5][2 + lane] = R.V.u; X.R[6][2 + lane] = R.V.v; X.R[7][2 + lane] = R.V.p;
        wave_sync();
        const Cons RU = {X.R[0][1 + lane], X.R[1][1 + lane], X.R[2][1 + lane], X.R[3][1 + lane]};
        const Prim RV = {X.R[4][1 + lane], X.R[5][1 + lane], X.R[6][1 + lane], X.R[7][1 + lane]};
        wave_sync();
        if (lane == 63) {
            X.R[0][1] = R.U.r; X.R[1][1] = R.U.mu; X.R[2][1] = R.U.mv; X.R[3][1] = R.U.e;
            X.R[4][1] = R.V.r; X.R[5][1] = R.V.u; X.R[6][1] = R.V.v; X.R[7][1] = R.V.p;
        }
        const Cons Fc = riemann<S, kX, F>(RU, RV, L.U, L.V, it);   // face q-1/2 = r+1/2, r = p-2
        if (give_up_now<S, F>(it)) {
            wave_sync();
            return false;
        }
        X.F[0][2 + lane] = Fc.r; X.F[1][2 + lane] = Fc.mu; X.F[2][2 + lane] = Fc.mv; X.F[3][2 + lane] = Fc.e;
        wave_sync();
        const Cons Fm = {X.F[0][1 + lane], X.F[1][1 + lane], X.F[2][1 + lane], X.F[3][1 + lane]};
        const Cons U2 = {X.U[0][lane], X.U[1][lane], X.U[2][lane], X.U[3][lane]};
        wave_sync();
        if (lane == 63) { X.F[0][1] = Fc.r; X.F[1][1] = Fc.mu; X.F[2][1] = Fc.mv; X.F[3][1] = Fc.e; }
        if (lane >= 62)
            for (int c = 0; c < 4; ++c) X.U[c][lane - 62] = X.U[c][lane + 2];
        const int r = p - 2;
        const bool mine = r >= i0 && r < i1;
        const Cons Unew = update(U2, dtd, Fm, Fc);
        st(out, mine ? r : kSinkCol, j, Unew);
        if (mine) {
            if (ghosts) {
                if (j == 0) {
                    const Cons G = bc.kind[MH_EDGE_BOTTOM] == MH_BC_INFLOW ? bc.s[MH_EDGE_BOTTOM] : Unew;
                    st(out, r, -1, G);
                    st(out, r, -2, G);
                }
                if (j == in.ny - 1) {
                    const Cons G = bc.kind[MH_EDGE_TOP] == MH_BC_INFLOW ? bc.s[MH_EDGE_TOP] : Unew;
                    st(out, r, in.ny, G);
                    st(out, r, in.ny + 1, G);
                }
            }
        }
    }
    if constexpr (F) return !__any(key >= kRangeKeyEnd || solver_gave_up(it));
    return true;
}

// The same x tile over the compact slice.
template <int S, bool F>
__device__ __forceinline__ bool x_tilec_pass(const View& in, const View& out, int j, int i0, int i1, double dtd,
                                        double half, const Bc& bc, int ghosts, XLdsC& X, int lane) {
    int it = 0;
    const int niter = (i1 - i0 + 4 + 63) / 64;
    const bool l1 = lane >= 1, l2 = lane >= 2;
    unsigned key = 0;   // running range key of the states read (fast pass)
    Cons Un = ld(in, min(i0 - 2 + lane, in.nx + 1), j);
    for (int k = 0; k < niter; ++k) {
        const int p = i0 - 2 + 64 * k + lane;
        const Cons U0 = Un;
        Un = ld(in, min(p + 64, in.nx + 1), j);
        const Prim W0 = c2p<F>(U0);
        if constexpr (F) key = max(key, range_key(U0.r, U0.e, W0.p));
        X.M[0][lane] = W0.r; X.M[1][lane] = W0.u; X.M[2][lane] = W0.v; X.M[3][lane] = W0.p;
        X.U[0][lane] = U0.r; X.U[1][lane] = U0.mu; X.U[2][lane] = U0.mv; X.U[3][lane] = U0.e;
        wave_sync();
        double q1[4], q2[4];
        for (int c = 0; c < 4; ++c) {
            q1[c] = l1 ? X.M[c][lane - 1] : X.cW[c][1];
            q2[c] = l2 ? X.M[c][lane - 2] : X.cW[c][lane];
        }
        wave_sync();
        if (lane >= 62) {
            X.cW[0][lane - 62] = W0.r; X.cW[1][lane - 62] = W0.u; X.cW[2][lane - 62] = W0.v;
            X.cW[3][lane - 62] = W0.p;
        }
        const Prim W1 = {q1[0], q1[1], q1[2], q1[3]};
        const Prim W2 = {q2[0], q2[1], q2[2], q2[3]};
        Edge L, R;
        reconstruct<kX, F>(W2, W1, W0, half, L, R);
        X.M[0][lane] = R.U.r; X.M[1][lane] = R.U.mu; X.M[2][lane] = R.U.mv; X.M[3][lane] = R.U.e;
        X.M[4][lane] = R.V.r; X.M[5][lane] = R.V.u; X.M[6][lane] = R.V.v; X.M[7][lane] = R.V.p;
        wave_sync();
        double r1[8];
        for (int c = 0; c < 8; ++c) r1[c] = l1 ? X.M[c][lane - 1] : X.cR[c];
        wave_sync();
        if (lane == 63) {
            X.cR[0] = R.U.r; X.cR[1] = R.U.mu; X.cR[2] = R.U.mv; X.cR[3] = R.U.e;
            X.cR[4] = R.V.r; X.cR[5] = R.V.u; X.cR[6] = R.V.v; X.cR[7] = R.V.p;
        }
        const Cons RU = {r1[0], r1[1], r1[2], r1[3]};
        const Prim RV = {r1[4], r1[5], r1[6], r1[7]};
        const Cons Fc = riemann<S, kX, F>(RU, RV, L.U, L.V, it);
        if (give_up_now<S, F>(it)) {
            wave_sync();
            return false;
        }
        X.M[0][lane] = Fc.r; X.M[1][lane] = Fc.mu; X.M[2][lane] = Fc.mv; X.M[3][lane] = Fc.e;
        wave_sync();
        double f1[4], u2[4];
        for (int c = 0; c < 4; ++c) {
            f1[c] = l1 ? X.M[c][lane - 1] : X.cF[c];
            u2[c] = l2 ? X.U[c][lane - 2] : X.cU[c][lane];
        }
        wave_sync();
        if (lane == 63) { X.cF[0] = Fc.r; X.cF[1] = Fc.mu; X.cF[2] = Fc.mv; X.cF[3] = Fc.e; }
        if (lane >= 62)
            for (int c = 0; c < 4; ++c) X.cU[c][lane - 62] = X.U[c][lane];
        const Cons Fm = {f1[0], f1[1], f1[2], f1[3]};
        const Cons U2 = {u2[0], u2[1], u2[2], u2[3]};
        const int r = p - 2;
        const bool mine = r >= i0 && r < i1;
        const Cons Unew = update(U2, dtd, Fm, Fc);
        st(out, mine ? r : kSinkCol, j, Unew);
        if (mine) {
            if (ghosts) {
                if (j == 0) {
                    const Cons G = bc.kind[MH_EDGE_BOTTOM] == MH_BC_INFLOW ? bc.s[MH_EDGE_BOTTOM] : Unew;
                    st(out, r, -1, G);
                    st(out, r, -2, G);
                }
                if (j == in.ny - 1) {
                    const Cons G = bc.kind[MH_EDGE_TOP] == MH_BC_INFLOW ? bc.s[MH_EDGE_TOP] : Unew;
                    st(out, r, in.ny, G);
                    st(out, r, in.ny + 1, G);
                }
            }
        }
    }
    if constexpr (F) return !__any(key >= kRangeKeyEnd || solver_gave_up(it));
    return true;
}

// The x tile with the update deferred by one iteration, over the compact
// slice: the face flux of iteration k goes to LDS (M[4..7], which the next
// iteration's primitive exchange leaves alone) and is read, with the
// conserved state of the cell two lanes left, in iteration k+1's primitive
// exchange; the cell update of iteration k runs there.  Two LDS round trips
// per iteration instead of three; the own face flux stays in registers
// across the next iteration's conversion only.  The U slots hold the
// previous iteration's states until that read is done, so lanes 62-63 copy
// theirs into the carry before every lane writes its new one.
// B as for y_tile_pass.
template <int S, bool F, bool B = true>
__device__ __forceinline__ bool x_tiled_pass(const View& in, const View& out, int j, int i0, int i1, double dtd,
                                             double half, const Bc& bc, int ghosts, XLdsC& X, int lane) {
    int it = 0;
    const int niter = (i1 - i0 + 4 + 63) / 64;
    const bool l1 = lane >= 1, l2 = lane >= 2;
    unsigned key = 0;   // running range key of the states read (fast pass)
    Cons Fprev = {0.0, 0.0, 0.0, 0.0};
    const uint32_t lane8 = (uint32_t)lane * (uint32_t)sizeof(double);
    // the update of the cell two left of lane `lane`'s cell of iteration kk:
    // stores address cell i0-4+64kk of the row plus the lane's offset; a
    // lane with no cell of the segment gets an offset past the descriptor's
    // range, and the hardware drops its store (every lane still issues it)
    auto update_prev = [&](int kk, const Cons& Fm, const Cons& U2) {
        const int c0 = i0 - 4 + 64 * kk;
        const int r = c0 + lane;
        const bool mine = (unsigned)(lane - (i0 - c0)) < (unsigned)(i1 - i0);
        const Cons Unew = update(U2, dtd, Fm, Fprev);
        if constexpr (B) st_off(out, row_at(out, c0, j), mine ? lane8 : 0x80000000u, Unew);
        else st(out, mine ? r : kSinkCol, j, Unew);
        if (mine && ghosts) {
            if (j == 0) {
                const Cons G = bc.kind[MH_EDGE_BOTTOM] == MH_BC_INFLOW ? bc.s[MH_EDGE_BOTTOM] : Unew;
                st(out, r, -1, G);
                st(out, r, -2, G);
            }
            if (j == in.ny - 1) {
                const Cons G = bc.kind[MH_EDGE_TOP] == MH_BC_INFLOW ? bc.s[MH_EDGE_TOP] : Unew;
                st(out, r, in.ny, G);
                st(out, r, in.ny + 1, G);
            }
        }
    };
    // reads the previous iteration's left face flux and the conserved state
    // two lanes left (carries for lanes 0-1)
    auto read_prev = [&](Cons& Fm, Cons& U2) {
        double f1[4], u2[4];
        for (int c = 0; c < 4; ++c) {
            f1[c] = l1 ? X.M[4 + c][lane - 1] : X.cF[c];
            u2[c] = l2 ? X.U[c][lane - 2] : X.cU[c][lane];
        }
        Fm = {f1[0], f1[1], f1[2], f1[3]};
        U2 = {u2[0], u2[1], u2[2], u2[3]};
    };
    // loads: the pass's cells i0-2+64k+lane, clamped to the last ghost cell
    // nx+1 -- the row address of cell min(c0, nx+1) and the lane's offset up
    // to that ghost
    auto load_pass = [&](int c0) {
        if constexpr (B) {
            const int cb = min(c0, in.nx + 1);
            const uint32_t lim = (uint32_t)(in.nx + 1 - cb) * (uint32_t)sizeof(double);
            return ld_off(in, row_at(in, cb, j), min(lane8, lim));
        } else {
            return ld(in, min(c0 + lane, in.nx + 1), j);
        }
    };
    Cons Un = load_pass(i0 - 2);
    for (int k = 0; k < niter; ++k) {
        const Cons U0 = Un;
        Un = load_pass(i0 - 2 + 64 * (k + 1));
        const Prim W0 = c2p<F>(U0);
        if constexpr (F) key = max(key, range_key(U0.r, U0.e, W0.p));
        X.M[0][lane] = W0.r; X.M[1][lane] = W0.u; X.M[2][lane] = W0.v; X.M[3][lane] = W0.p;
        wave_sync();
        double q1[4], q2[4];
        for (int c = 0; c < 4; ++c) {
            q1[c] = l1 ? X.M[c][lane - 1] : X.cW[c][1];
            q2[c] = l2 ? X.M[c][lane - 2] : X.cW[c][lane];
        }
        Cons Fm, U2;
        read_prev(Fm, U2);
        wave_sync();
        if (lane >= 62) {
            X.cW[0][lane - 62] = W0.r; X.cW[1][lane - 62] = W0.u; X.cW[2][lane - 62] = W0.v;
            X.cW[3][lane - 62] = W0.p;
            for (int c = 0; c < 4; ++c) X.cU[c][lane - 62] = X.U[c][lane];
        }
        X.U[0][lane] = U0.r; X.U[1][lane] = U0.mu; X.U[2][lane] = U0.mv; X.U[3][lane] = U0.e;
        if (lane == 63) { X.cF[0] = Fprev.r; X.cF[1] = Fprev.mu; X.cF[2] = Fprev.mv; X.cF[3] = Fprev.e; }
        if (k > 0) update_prev(k - 1, Fm, U2);
        const Prim W1 = {q1[0], q1[1], q1[2], q1[3]};
        const Prim W2 = {q2[0], q2[1], q2[2], q2[3]};
        Edge L, R;
        reconstruct<kX, F>(W2, W1, W0, half, L, R);
        // the right edge state to the lane on the right: conserved state,
        // u, v, p, and in the slot of its density (which equals the
        // conserved density) rcp_nr of the density for the fast tile's
        // Riemann solve
        const double r4 = (F && B) ? R.rr : R.V.r;
        X.M[0][lane] = R.U.r; X.M[1][lane] = R.U.mu; X.M[2][lane] = R.U.mv; X.M[3][lane] = R.U.e;
        X.M[4][lane] = r4; X.M[5][lane] = R.V.u; X.M[6][lane] = R.V.v; X.M[7][lane] = R.V.p;
        wave_sync();
        double r1[8];
        for (int c = 0; c < 8; ++c) r1[c] = l1 ? X.M[c][lane - 1] : X.cR[c];
        wave_sync();
        if (lane == 63) {
            X.cR[0] = R.U.r; X.cR[1] = R.U.mu; X.cR[2] = R.U.mv; X.cR[3] = R.U.e;
            X.cR[4] = r4; X.cR[5] = R.V.u; X.cR[6] = R.V.v; X.cR[7] = R.V.p;
        }
        const Cons RU = {r1[0], r1[1], r1[2], r1[3]};
        const Prim RV = {r1[0], r1[5], r1[6], r1[7]};
        Fprev = riemann<S, kX, F, B>(RU, RV, L.U, L.V, it, r1[4], L.rr);
        if (give_up_now<S, F>(it)) {
            wave_sync();
            return false;
        }
        X.M[4][lane] = Fprev.r; X.M[5][lane] = Fprev.mu; X.M[6][lane] = Fprev.mv; X.M[7][lane] = Fprev.e;
    }
    wave_sync();
    Cons Fm, U2;
    read_prev(Fm, U2);
    update_prev(niter - 1, Fm, U2);
    wave_sync();   // the slice is free for the next tile of this wavefront
    if constexpr (F) return !__any(key >= kRangeKeyEnd || solver_gave_up(it));
    return true;
}

// XStream: whole rows j0..j1-1 as one stream of rows of nx+4 cells (each
// row with its two ghost cells on either side), 64 stream positions per
// pass, so one pass may hold the end of a row and the start of the next.  A
// cell's update reads only cells c-2..c+2 of its own row, so every update
// stored is the one the row alone gives; the positions whose reconstruction
// or faces mix two rows are ghost cells, whose updates are not stored.
// Against one segment per row, short rows leave no lanes idle (a row of 64
// cells takes 68 lanes, not 128) and a row's first loads are in flight
// while the row before it ends.  Rows through buffer descriptors at cell
// (-2, j0) of each grid; in and out have the same pitch and plane size, and
// (j1-j0) rows of them span less than 2^31 bytes (the host keeps to this).
struct XStream {
    // by value: the buffer stores below could alias fields read through a
    // reference, which would then be reloaded every pass
    View in, out;
    const Bc& bc;
    int j0, j1, ghosts, lane;
    double dtd;
    int w, len;              // stream row length nx+4, positions (j1-j0)*w
    int wraps;               // most row ends one pass can cross
    uint32_t olast, wrapb;   // byte offsets: position len-1; row change
    const double* inb;
    const double* outb;
    // load and store cursors: cell index and byte offset of the lane's
    // position in the next pass to load / store
    int lc, sc;
    uint32_t lo, so;
    __device__ XStream(const View& in_, const View& out_, const Bc& bc_, int j0_, int j1_, int ghosts_, int lane_,
                       double dtd_)
        : in(in_), out(out_), bc(bc_), j0(j0_), j1(j1_), ghosts(ghosts_), lane(lane_), dtd(dtd_) {
        w = in.nx + 4;
        len = (j1 - j0) * w;
        wraps = w >= 64 ? 1 : 64 / w + 1;
        olast = (uint32_t)((int64_t)(j1 - j0 - 1) * in.rp + (w - 1)) * (uint32_t)sizeof(double);
        wrapb = (uint32_t)(in.rp - w) * (uint32_t)sizeof(double);
        inb = row_at(in, -2, j0);
        outb = row_at(out, -2, j0);
        lc = lane - 2;
        lo = (uint32_t)lane * (uint32_t)sizeof(double);
        for (int m = 0; m < wraps; ++m) wrap(lc, lo);
        sc = lc;
        so = lo;
    }
    __device__ void wrap(int& c, uint32_t& o) const {
        const bool e = c > in.nx + 1;
        c = e ? c - w : c;
        o = e ? o + wrapb : o;
    }
    __device__ void advance(int& c, uint32_t& o) const {
        c += 64;
        o += 64u * (uint32_t)sizeof(double);
        for (int m = 0; m < wraps; ++m) wrap(c, o);
    }
    __device__ int passes() const { return (len + 63) / 64; }
    // positions past the stream's end load its last cell
    __device__ Cons load(int) {
        const Cons U = ld_off(in, inb, min(lo, olast));
        advance(lc, lo);
        return U;
    }
    __device__ void store(int kk, const Cons& U2, const Cons& Fm, const Cons& Fp) {
        const int c = sc - 2;   // the updated cell, in the lane's row
        const bool mine = (unsigned)c < (unsigned)in.nx && 64 * kk + lane < len;
        const Cons Unew = update(U2, dtd, Fm, Fp);
        st_off(out, outb, mine ? so - 2u * (uint32_t)sizeof(double) : 0x80000000u, Unew);
        if (mine && ghosts) {
            const int s = 64 * kk + lane;
            if (j0 == 0 && s < w) {
                const Cons G = bc.kind[MH_EDGE_BOTTOM] == MH_BC_INFLOW ? bc.s[MH_EDGE_BOTTOM] : Unew;
                st(out, c, -1, G);
                st(out, c, -2, G);
            }
            if (j1 == in.ny && s >= len - w) {
                const Cons G = bc.kind[MH_EDGE_TOP] == MH_BC_INFLOW ? bc.s[MH_EDGE_TOP] : Unew;
                st(out, c, in.ny, G);
                st(out, c, in.ny + 1, G);
            }
        }
        advance(sc, so);
    }
};

// The deferred-update x tile of x_tiled_pass over an XStream: the same
// passes, exchanges and face solves, with the stream's loads and stores (the
// single-grid sweeps keep their own copy of the pass, which the compiler
// schedules 0.5 % faster than this form with the addressing passed in).
template <int S, bool F>
__device__ __forceinline__ bool x_stream_pass(XStream& a, double half, XLdsC& X, int lane) {
    int it = 0;
    const int niter = a.passes();
    const bool l1 = lane >= 1, l2 = lane >= 2;
    unsigned key = 0;   // running range key of the states read (fast pass)
    Cons Fprev = {0.0, 0.0, 0.0, 0.0};
    // reads the previous iteration's left face flux and the conserved state
    // two lanes left (carries for lanes 0-1)
    auto read_prev = [&](Cons& Fm, Cons& U2) {
        double f1[4], u2[4];
        for (int c = 0; c < 4; ++c) {
            f1[c] = l1 ? X.M[4 + c][lane - 1] : X.cF[c];
            u2[c] = l2 ? X.U[c][lane - 2] : X.cU[c][lane];
        }
        Fm = {f1[0], f1[1], f1[2], f1[3]};
        U2 = {u2[0], u2[1], u2[2], u2[3]};
    };
    Cons Un = a.load(0);
    for (int k = 0; k < niter; ++k) {
        const Cons U0 = Un;
        Un = a.load(k + 1);
        const Prim W0 = c2p<F>(U0);
        if constexpr (F) key = max(key, range_key(U0.r, U0.e, W0.p));
        X.M[0][lane] = W0.r; X.M[1][lane] = W0.u; X.M[2][lane] = W0.v; X.M[3][lane] = W0.p;
        wave_sync();
        double q1[4], q2[4];
        for (int c = 0; c < 4; ++c) {
            q1[c] = l1 ? X.M[c][lane - 1] : X.cW[c][1];
            q2[c] = l2 ? X.M[c][lane - 2] : X.cW[c][lane];
        }
        Cons Fm, U2;
        read_prev(Fm, U2);
        wave_sync();
        if (lane >= 62) {
            X.cW[0][lane - 62] = W0.r; X.cW[1][lane - 62] = W0.u; X.cW[2][lane - 62] = W0.v;
            X.cW[3][lane - 62] = W0.p;
            for (int c = 0; c < 4; ++c) X.cU[c][lane - 62] = X.U[c][lane];
        }
        X.U[0][lane] = U0.r; X.U[1][lane] = U0.mu; X.U[2][lane] = U0.mv; X.U[3][lane] = U0.e;
        if (lane == 63) { X.cF[0] = Fprev.r; X.cF[1] = Fprev.mu; X.cF[2] = Fprev.mv; X.cF[3] = Fprev.e; }
        if (k > 0) a.store(k - 1, U2, Fm, Fprev);
        const Prim W1 = {q1[0], q1[1], q1[2], q1[3]};
        const Prim W2 = {q2[0], q2[1], q2[2], q2[3]};
        Edge L, R;
        reconstruct<kX, F>(W2, W1, W0, half, L, R);
        // the right edge state to the lane on the right: conserved state and
        // u, v, p (its primitive density is the conserved one)
        X.M[0][lane] = R.U.r; X.M[1][lane] = R.U.mu; X.M[2][lane] = R.U.mv; X.M[3][lane] = R.U.e;
        X.M[5][lane] = R.V.u; X.M[6][lane] = R.V.v; X.M[7][lane] = R.V.p;
        wave_sync();
        double r1[8];
        for (int c = 0; c < 8; ++c)
            if (c != 4) r1[c] = l1 ? X.M[c][lane - 1] : X.cR[c];
        wave_sync();
        if (lane == 63) {
            X.cR[0] = R.U.r; X.cR[1] = R.U.mu; X.cR[2] = R.U.mv; X.cR[3] = R.U.e;
            X.cR[5] = R.V.u; X.cR[6] = R.V.v; X.cR[7] = R.V.p;
        }
        const Cons RU = {r1[0], r1[1], r1[2], r1[3]};
        const Prim RV = {r1[0], r1[5], r1[6], r1[7]};
        Fprev = riemann<S, kX, F>(RU, RV, L.U, L.V, it);
        if (give_up_now<S, F>(it)) {
            wave_sync();
            return false;
        }
        X.M[4][lane] = Fprev.r; X.M[5][lane] = Fprev.mu; X.M[6][lane] = Fprev.mv; X.M[7][lane] = Fprev.e;
    }
    wave_sync();
    Cons Fm, U2;
    read_prev(Fm, U2);
    a.store(niter - 1, U2, Fm, Fprev);
    wave_sync();   // the slice is free for the next tile of this wavefront
    if constexpr (F) return !__any(key >= kRangeKeyEnd || solver_gave_up(it));
    return true;
}

// Rows [j0, j1) of a grid as one x tile (XStream), fast pass first as in
// x_tile; the batch kernels' x tiles of whole rows.
template <int S>
__device__ __forceinline__ void x_stream(const View& in, const View& out, int j0, int j1, double dtd, double half,
                                         const Bc& bc, int ghosts, XLdsC& X, int lane) {
    if (dt_in_fast_range(dtd)) {
        XStream a(in, out, bc, j0, j1, ghosts, lane, dtd);
        if (x_stream_pass<S, true>(a, half, X, lane)) return;
    }
    XStream a(in, out, bc, j0, j1, ghosts, lane, dtd);
    x_stream_pass<S, false>(a, half, X, lane);
}

// Layout selector: L = 0 compact, 1 = 66-slot, 2 = compact with the update
// deferred by one iteration.
template <int L> struct XSlice { using type = XLdsC; };
template <> struct XSlice<1> { using type = XLds66; };

template <int S, int L, bool B = true>
__device__ __forceinline__ void x_tile(const View& in, const View& out, int j, int i0, int i1, double dtd,
                                       double half, const Bc& bc, int ghosts, typename XSlice<L>::type& X,
                                       int lane) {
    // fast pass first (see y_tile_pass); the IEEE pass redoes the segment if
    // the fast one gives up
    const bool fast = dt_in_fast_range(dtd);
    if constexpr (L == 1) {
        if (fast && x_tile66_pass<S, true>(in, out, j, i0, i1, dtd, half, bc, ghosts, X, lane)) return;
        x_tile66_pass<S, false>(in, out, j, i0, i1, dtd, half, bc, ghosts, X, lane);
    } else if constexpr (L == 2) {
        if (fast && x_tiled_pass<S, true, B>(in, out, j, i0, i1, dtd, half, bc, ghosts, X, lane)) return;
        x_tiled_pass<S, false, B>(in, out, j, i0, i1, dtd, half, bc, ghosts, X, lane);
    } else {
        if (fast && x_tilec_pass<S, true>(in, out, j, i0, i1, dtd, half, bc, ghosts, X, lane)) return;
        x_tilec_pass<S, false>(in, out, j, i0, i1, dtd, half, bc, ghosts, X, lane);
    }
}

// Host helpers defined in solver.hip.
Bc to_bc(const mh_bc* b);
int launch_fill(const mh_grid_s* g, const Bc& bc, hipStream_t st);

}  // namespace mh
