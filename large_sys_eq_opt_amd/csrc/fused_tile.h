// fused_tile.h -- the x-then-y split step as one pass over the grid: a pair
// of wavefronts per 64-column strip and chunk of rows, one workgroup.
//
// The x wavefront sweeps the strip's rows j0-2 .. j1+1 of the input state
// one row at a time (lane l owns column c0+l) and writes each x-updated row
// into a four-row LDS ring; the y wavefront marches up the same columns as
// the y tile does (y_tile_pass), reading its rows from that ring instead of
// HBM.  A workgroup barrier after every row hands one row over: the x
// wavefront is one row ahead, and the ring holds the row being written and
// the three the y wavefront still reads (rows p, p-1 and, for the update,
// p-2).  The intermediate state never leaves the CU: a step reads 32 and
// writes 32 bytes per cell instead of 64 and 64.
//
// The x update of a row's 64 columns needs the primitive states of the two
// cells left and right of the strip, the face flux at its left edge and the
// left edge state of the cell right of it.  A lane cannot compute those for
// its own row without a second pass over the whole wavefront, so the x
// wavefront computes them for 32 rows at a time in one pass (window pass):
// lanes 0-31 the left side of rows k..k+31, lanes 32-63 the right side, from
// four cells each, and keeps them in LDS (B).  Every value is computed by the
// same functions on the same operands as in the x sweep, so the step is the
// x sweep followed by the y sweep bit for bit.
//
// Ghost rows of the intermediate state (rows -2, -1, ny, ny+1 that the y
// update reads) are the x update of row 0 / ny-1 (transmissive) or the
// inflow state, as the x sweep's epilogue writes them.
//
// The last strip of a grid whose nx is not a multiple of 64 is partial (P):
// its W = nx - c0 < 64 columns are lanes 0..W-1.  Row loads clamp at the
// last ghost cell nx+1, so lanes W and W+1 hold the ghost cells nx and nx+1
// and give lane W-1 its right neighbour's values in the row pass (for W = 63
// the window's right side supplies W(nx+1)); lanes past W compute on copies
// of ghost states and the y wavefront neither stores nor counts them.  Full
// strips take the code without these clamps and masks (P = false): it keeps
// the register schedule of the hot path.
#pragma once

#include "sweep_tiles.h"

namespace mh {

// One workgroup's LDS (19 KB): eight fit the 160 KB of a CU, so sixteen
// wavefronts (four per SIMD) can be resident.
//
// The exact solver's fast pass (exact_fast) leaves every tile with a
// non-uniform face to the IEEE pass, so its pair stops at the first row
// that gives up: the wavefront that finds it sets stop[b & 1] before
// barrier b, and both read stop[b & 1] after barrier b and leave the tile
// there, after the same number of barriers.  A flag for barrier b + 1 goes
// to the other slot, and one for barrier b + 2 can only be written once the
// reader of slot b & 1 has passed barrier b + 1.
struct PairLds {
    double ring[4][4][64];   // intermediate rows: slot r & 3, plane, lane
    double M[8][66];         // x exchange: lane l's value at slot l + 1
    double B[12][66];        // window values: left side at slot i, right at 32 + i (row k0 + i)
    unsigned key[64];        // y wavefront: lane's running range key
    unsigned xok;            // x wavefront's fast-range verdict for the tile
    unsigned stop[2];        // exact solver's fast pass: stop after barrier b if stop[b & 1]
    unsigned verdict;        // redo pass: the y wavefront's verdict on a piece, for both
};
constexpr int kPairWindow = 32;

// Both wavefronts of the pair meet here once per row; only LDS needs
// ordering across the barrier (each wavefront's global loads and stores are
// its own), so the fences are restricted to the local address space and
// leave loads in flight.
__device__ __forceinline__ void pair_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Window pass of the x wavefront for rows k0 .. k0+31 of the tile (row
// index k: grid row j0-2+k, clamped to the grid for ghost rows).  Lane i <
// 32 loads cells c0-2 .. c0+1 of row k0+i: W(c0-1), the left edge state of
// cell c0 and the right edge state of cell c0-1, and the flux between them;
// lane 32+i loads cells c0+62 .. c0+65: W(c0+64) and the left edge state of
// cell c0+64.  B rows: 0-3 W, 4-7 flux (left) / conserved edge state
// (right), 8-11 the edge state's rcp_nr of its density (fast pass) or its
// density, then u, v, p (right).  In a partial strip the right side loads
// from column min(c0 + 62, nx - 1) and its fourth column clamps at nx+1:
// every load stays in the row, and for W = 63 the values of B are the full
// strip's (for fewer columns no lane that counts reads them).
template <int S, bool F, bool P>
__device__ __forceinline__ void pair_x_window(const View& in, int c0, int j0, int k0, double half, PairLds& X,
                                              int lane, unsigned& key) {
    int it = 0;
    const int i = lane & 31;
    const bool right = lane >= 32;
    const int src = min(max(j0 - 2 + k0 + i, 0), in.ny - 1);
    const int cs = right ? (P ? min(c0 + 62, in.nx - 1) : c0 + 62) : c0 - 2;
    Prim Wq[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const Cons U = ld(in, P && t == 3 ? min(cs + t, in.nx + 1) : cs + t, src);
        Wq[t] = c2p<F>(U);
        if constexpr (F) key = max(key, range_key(U.r, U.e, Wq[t].p));
    }
    Edge La, Ra, Lb, Rb;
    reconstruct<kX, F>(Wq[1], Wq[2], Wq[3], half, Lb, Rb);   // cell c0 (left) / c0+64 (right)
    reconstruct<kX, F>(Wq[0], Wq[1], Wq[2], half, La, Ra);   // cell c0-1 (left)
    const Cons Fl = riemann<S, kX, F, true>(Ra.U, Ra.V, Lb.U, Lb.V, it, Ra.rr, Lb.rr);
    const Prim Wb = right ? Wq[2] : Wq[1];
    X.B[0][lane] = Wb.r; X.B[1][lane] = Wb.u; X.B[2][lane] = Wb.v; X.B[3][lane] = Wb.p;
    X.B[4][lane] = right ? Lb.U.r : Fl.r;
    X.B[5][lane] = right ? Lb.U.mu : Fl.mu;
    X.B[6][lane] = right ? Lb.U.mv : Fl.mv;
    X.B[7][lane] = right ? Lb.U.e : Fl.e;
    X.B[8][lane] = F ? Lb.rr : Lb.V.r;
    X.B[9][lane] = Lb.V.u; X.B[10][lane] = Lb.V.v; X.B[11][lane] = Lb.V.p;
    if constexpr (F) key = solver_gave_up(it) ? kRangeKeyEnd : key;
    wave_sync();
}

// The x wavefront of a pair: rows k = 0 .. nrows-1 (grid rows j0-2 .. j1+1)
// of columns c0 .. c0+63 into the ring, a barrier after each.  Per row:
// conversion, exchange of W with both neighbours, reconstruction, exchange
// of the left edge state to the left neighbour, the face flux at the right
// of each cell (the x sweep's riemann call on the same states), exchange of
// the flux to the right neighbour, update.  Lane 0 takes its left
// neighbour's values and lane 63 its right neighbour's from the window
// (B); the LDS address is selected once per exchange, the components follow
// at the same stride (66 doubles) in M and B.  Writes the tile's fast-range
// verdict (F) into X.xok before the last barrier.
template <int S, bool F, bool P>
__device__ __forceinline__ void pair_x(const View& in, int c0, int j0, int j1, double dtd, double half, const Bc& bc,
                                       PairLds& X, int lane) {
    constexpr bool A = F && S == kExact;   // stops at the first row the solver gives up
    int it = 0;
    unsigned key = 0;
    const int nrows = j1 - j0 + 4;
    const uint32_t lane8 = (uint32_t)(P ? min(c0 + lane, in.nx + 1) - c0 : lane) * (uint32_t)sizeof(double);
    if constexpr (A) {
        if (lane == 0) {
            X.stop[0] = 0u;
            X.stop[1] = 0u;
        }
    }
    auto src_row = [&](int k) { return min(max(j0 - 2 + k, 0), in.ny - 1); };
    Cons Ua = ld_off(in, row_at(in, c0, src_row(0)), lane8), Ub;
    const double* mL = &X.M[0][lane];
    const double* mR = &X.M[0][lane + 2];
    // Row k from its state U0 (loaded one row ahead), loading row k+1 into
    // Un; false: the pair stops here (A).  The loaded rows alternate between
    // two sets of registers, two rows per loop pass, as in the y wavefront.
    auto step = [&](int k, const Cons& U0, Cons& Un) {
        const int w = k & (kPairWindow - 1);
        if (w == 0) pair_x_window<S, F, P>(in, c0, j0, k, half, X, lane, key);
        Un = ld_off(in, row_at(in, c0, src_row(min(k + 1, nrows - 1))), lane8);
        const Prim W0 = c2p<F>(U0);
        if constexpr (F) key = max(key, range_key(U0.r, U0.e, W0.p));
        X.M[0][lane + 1] = W0.r; X.M[1][lane + 1] = W0.u; X.M[2][lane + 1] = W0.v; X.M[3][lane + 1] = W0.p;
        wave_sync();
        const double* pl = lane == 0 ? &X.B[0][w] : mL;
        const double* pr = lane == 63 ? &X.B[0][32 + w] : mR;
        const Prim Wm = {pl[0], pl[66], pl[132], pl[198]};
        const Prim Wp = {pr[0], pr[66], pr[132], pr[198]};
        wave_sync();
        Edge L, R;
        reconstruct<kX, F>(Wm, W0, Wp, half, L, R);
        X.M[0][lane + 1] = L.U.r; X.M[1][lane + 1] = L.U.mu; X.M[2][lane + 1] = L.U.mv; X.M[3][lane + 1] = L.U.e;
        X.M[4][lane + 1] = F ? L.rr : L.V.r;
        X.M[5][lane + 1] = L.V.u; X.M[6][lane + 1] = L.V.v; X.M[7][lane + 1] = L.V.p;
        wave_sync();
        const double* pe = lane == 63 ? &X.B[4][32 + w] : mR;
        double q[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) q[c] = pe[66 * c];
        wave_sync();
        // face c0+l+1/2: this lane's right edge state and its right
        // neighbour's left edge state (its primitive density is the
        // conserved one)
        const Cons Fc = riemann<S, kX, F, true>(R.U, R.V, Cons{q[0], q[1], q[2], q[3]}, Prim{q[0], q[5], q[6], q[7]},
                                                it, R.rr, q[4]);
        X.M[0][lane + 1] = Fc.r; X.M[1][lane + 1] = Fc.mu; X.M[2][lane + 1] = Fc.mv; X.M[3][lane + 1] = Fc.e;
        wave_sync();
        const double* pf = lane == 0 ? &X.B[4][w] : mL;
        const Cons Fm = {pf[0], pf[66], pf[132], pf[198]};
        wave_sync();
        const int r = j0 - 2 + k;
        Cons Us = update(U0, dtd, Fm, Fc);
        if (r < 0 && bc.kind[MH_EDGE_BOTTOM] == MH_BC_INFLOW) Us = bc.s[MH_EDGE_BOTTOM];
        if (r >= in.ny && bc.kind[MH_EDGE_TOP] == MH_BC_INFLOW) Us = bc.s[MH_EDGE_TOP];
        double* s = &X.ring[r & 3][0][lane];
        s[0] = Us.r; s[64] = Us.mu; s[128] = Us.mv; s[192] = Us.e;
        if (k == nrows - 1) {
            const bool ok = F ? !__any(key >= kRangeKeyEnd || solver_gave_up(it)) : true;
            if (lane == 0) X.xok = ok ? 1u : 0u;
        }
        if constexpr (A) {
            if (__any(key >= kRangeKeyEnd || solver_gave_up(it)) && lane == 0) X.stop[k & 1] = 1u;
        }
        pair_barrier();
        if constexpr (A) {
            if (X.stop[k & 1]) return false;
        }
        return true;
    };
    int k = 0;
    if constexpr (S != kExact && S != kExactFull) {
        // the window pass falls on even rows only (kPairWindow is even)
        for (; k + 1 < nrows; k += 2) {
            step(k, Ua, Ub);
            step(k + 1, Ub, Ua);
        }
    }
    for (; k < nrows; ++k) {
        if (!step(k, Ua, Ub)) return;
        Ua = Ub;
    }
}

// The y wavefront of a pair: y_tile_pass over columns c0 .. c0+63 and rows
// [j0, j1), each row read from the ring after the barrier that follows its x
// update; the update of row p-2 reads that row's intermediate state from
// the ring too.  Returns the y rows' fast-range verdict (F); m receives the
// lane's maximum signal speed of the new states when `speed` is set.
template <int S, bool F, bool P>
__device__ __forceinline__ bool pair_y(const View& out, int c0, int j0, int j1, double dtd, double half, const Bc& bc,
                                       int ghosts, int speed, PairLds& X, int lane, double& m) {
    constexpr bool A = F && S == kExact;   // stops at the first row the solver gives up
    int it = 0;
    const int i = c0 + lane;
    // stores: the lane's cell, or the sink column for a lane past the row's end
    const uint32_t so = (uint32_t)(!P || i < out.nx ? i - kSinkCol : 0) * (uint32_t)sizeof(double);
    auto held = [&](int r) {
        const double* q = &X.ring[r & 3][0][lane];
        return Cons{q[0], q[64], q[128], q[192]};
    };
    auto row = [&](int r) {
        pair_barrier();
        return held(r);
    };
    // (A) the pair stops after the barrier of row r (barrier r - j0 + 2 of
    // the tile); a give-up before the barrier of row r goes to its slot
    auto stopped = [&](int r) { return A && X.stop[(r - j0 + 2) & 1] != 0u; };
    auto give_up_before = [&](int r) {
        if (A && __any(solver_gave_up(it)) && lane == 0) X.stop[(r - j0 + 2) & 1] = 1u;
    };
    unsigned* key = &X.key[lane];
    if constexpr (F) *key = 0u;
    auto conv = [&](const Cons& U, Prim& W) {
        W = c2p<F>(U);
        if constexpr (F) *key = max(*key, range_key(U.r, U.e, W.p));
    };
    m = 0.0;
    Prim dm, W1, W0;
    conv(row(j0 - 2), W0);
    if (stopped(j0 - 2)) return false;
    conv(row(j0 - 1), W1);
    if (stopped(j0 - 1)) return false;
    dm = diff(W1, W0);
    Edge L, R, Rp;
    conv(row(j0), W0);
    if (stopped(j0)) return false;
    dm = reconstruct_d<kY, F>(dm, W1, W0, half, L, Rp);
    W1 = W0;
    conv(row(j0 + 1), W0);
    if (stopped(j0 + 1)) return false;
    dm = reconstruct_d<kY, F>(dm, W1, W0, half, L, R);
    Cons Fp = riemann<S, kY, F, true>(Rp.U, Rp.V, L.U, L.V, it, Rp.rr, L.rr);
    give_up_before(j0 + 2);
    // Row p of the march: its primitive state Wn, the face below it from the
    // previous row's right edge state Ro and this row's left edge state, and
    // the update of row p-2 from the face fluxes Fo (below) and Fn (above).
    // The carried values (W, R, F) alternate between two sets of registers,
    // two rows per loop pass, so no copies close the loop (not for the exact
    // solver, whose wavefront has no registers to spare for the second set:
    // it copies them after every row).
    auto step = [&](int p, const Prim& Wo, Prim& Wn, const Edge& Ro, Edge& Rn, const Cons& Fo, Cons& Fn) {
        conv(row(p), Wn);
        if (stopped(p)) return false;
        Edge Lp;
        dm = reconstruct_d<kY, F>(dm, Wo, Wn, half, Lp, Rn);
        Fn = riemann<S, kY, F, true>(Ro.U, Ro.V, Lp.U, Lp.V, it, Ro.rr, Lp.rr);
        if (p < j1 + 1) give_up_before(p + 1);
        const int r = p - 2;
        const Cons Unew = update(held(r), dtd, Fo, Fn);
        st_off(out, row_at(out, kSinkCol, r), so, Unew);
        if (speed) m = smax(m, signal_speed<F>(Unew));
        if (ghosts) {
            if (i == 0) {
                const Cons G = bc.kind[MH_EDGE_LEFT] == MH_BC_INFLOW ? bc.s[MH_EDGE_LEFT] : Unew;
                st(out, -1, r, G);
                st(out, -2, r, G);
            }
            if (i == out.nx - 1) {
                const Cons G = bc.kind[MH_EDGE_RIGHT] == MH_BC_INFLOW ? bc.s[MH_EDGE_RIGHT] : Unew;
                st(out, out.nx, r, G);
                st(out, out.nx + 1, r, G);
            }
        }
        return true;
    };
    Cons Fq;
    int p = j0 + 2;
    if constexpr (S != kExact && S != kExactFull) {
        for (; p + 1 <= j1 + 1; p += 2) {
            if (!step(p, W0, W1, R, Rp, Fp, Fq)) return false;
            if (!step(p + 1, W1, W0, Rp, R, Fq, Fp)) return false;
        }
    }
    for (; p <= j1 + 1; ++p) {
        if (!step(p, W0, W1, R, Rp, Fp, Fq)) return false;
        W0 = W1;
        R = Rp;
        Fp = Fq;
    }
    // lanes past a partial strip's last column computed on ghost copies:
    // their speeds and keys do not count
    const bool active = !P || i < out.nx;
    if (!active) m = 0.0;
    if constexpr (F) return !__any(active && (*key >= kRangeKeyEnd || solver_gave_up(it)));
    return true;
}

// Both wavefronts' side of a pair, with the strip's kind (full or partial)
// chosen per tile: w = 0 sweeps x into the ring, w = 1 sweeps y out of it
// and returns its verdict (the x wavefront returns true).
template <int S, bool F>
__device__ __forceinline__ bool pair_side(int w, const View& in, const View& out, int c0, int j0, int j1, double dtdx,
                                          double dtdy, const Bc& bc, int ghosts, int speed, PairLds& X, int lane,
                                          double& m) {
    const bool partial = c0 + 64 > in.nx;
    if (w == 0) {
        if (partial) pair_x<S, F, true>(in, c0, j0, j1, dtdx, 0.5 * dtdx, bc, X, lane);
        else pair_x<S, F, false>(in, c0, j0, j1, dtdx, 0.5 * dtdx, bc, X, lane);
        return true;
    }
    if (partial) return pair_y<S, F, true>(out, c0, j0, j1, dtdy, 0.5 * dtdy, bc, ghosts, speed, X, lane, m);
    return pair_y<S, F, false>(out, c0, j0, j1, dtdy, 0.5 * dtdy, bc, ghosts, speed, X, lane, m);
}

}  // namespace mh
