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
// layout is faster depends on occupancy, so both exist.
struct XLdsC {
    double M[8][64];
    double U[4][64];
    double cW[4][2];
    double cU[4][2];
    double cR[8];
    double cF[4];
};

constexpr int kRing = 768;   // y tile: three rows of four planes, 64 lanes

// y tile over columns [64*cg, 64*cg+64) and rows [j0, j1).  Returns the
// lane's running max of the output's signal speed when `speed` is set
// (std::max semantics: a NaN speed never enters), 0 otherwise.  The
// conserved states of rows p-1 and p-2 wait for the update in a three-row
// LDS ring (each lane touches only its own column, so program order is the
// only ordering needed); `ring` is the wavefront's kRing-double LDS slice.
template <int S>
__device__ __forceinline__ double y_tile(const View& in, const View& out, int cg, int j0, int j1, double dtd,
                                         double half, const Bc& bc, int ghosts, int speed, double* ring, int lane) {
    const int i = cg * 64 + lane;
    const bool active = i < in.nx;
    const int ic = active ? i : in.nx - 1;
    int it = 0;
    auto put = [&](int slot, const Cons& U) {
        double* q = ring + slot * 256 + lane;
        q[0] = U.r; q[64] = U.mu; q[128] = U.mv; q[192] = U.e;
    };
    auto get = [&](int slot) {
        const double* q = ring + slot * 256 + lane;
        return Cons{q[0], q[64], q[128], q[192]};
    };
    // rows j0-2 and j0-1: convert only
    Cons U0 = ld(in, ic, j0 - 2);
    Prim W2 = c2p(U0);
    put(0, U0);
    U0 = ld(in, ic, j0 - 1);
    Prim W1 = c2p(U0);
    put(1, U0);
    Cons Un = ld(in, ic, j0);
    Edge L, R, Rp;
    {   // row p = j0: reconstruct j0-1
        U0 = Un;
        Un = ld(in, ic, j0 + 1);
        const Prim W0 = c2p(U0);
        put(2, U0);
        reconstruct<kY>(W2, W1, W0, half, L, Rp);
        W2 = W1; W1 = W0;
    }
    Cons Fp;
    {   // row p = j0+1: reconstruct j0, face j0-1/2
        U0 = Un;
        Un = ld(in, ic, j0 + 2);
        const Prim W0 = c2p(U0);
        put(0, U0);   // row j0-2 is no longer needed
        reconstruct<kY>(W2, W1, W0, half, L, R);
        Fp = riemann<S, kY>(Rp.U, Rp.V, L.U, L.V, it);
        Rp = R;
        W2 = W1; W1 = W0;
    }
    double m = 0.0;
    // rows p = j0+2 .. j1+1; the ring slot of row p is (p - j0 + 2) % 3
    int slot_new = 1, slot_old = 2;   // slots of rows p and p-2
    for (int p = j0 + 2; p <= j1 + 1; ++p) {
        U0 = Un;
        Un = ld(in, ic, min(p + 1, j1 + 1));
        const Prim W0 = c2p(U0);
        reconstruct<kY>(W2, W1, W0, half, L, R);
        const Cons F = riemann<S, kY>(Rp.U, Rp.V, L.U, L.V, it);
        const int r = p - 2;
        const Cons Unew = update(get(slot_old), dtd, Fp, F);
        put(slot_new, U0);
        if (active) {
            st(out, i, r, Unew);
            if (speed) m = smax(m, signal_speed(Unew));
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
        Fp = F;
        Rp = R;
        W2 = W1; W1 = W0;
        slot_old = slot_old == 2 ? 0 : slot_old + 1;
        slot_new = slot_new == 2 ? 0 : slot_new + 1;
    }
    return m;
}

// x tile over row j, cells [i0, i1).  The lane holding cell p at iteration
// k reconstructs cell p-1, solves face (p-2)+1/2 and updates cell p-2; the
// W and U of cells p-1, p-2 and the right edge state and flux of the left
// neighbour come from lanes l-1, l-2 through LDS (the carry slots for lanes
// 0 and 1).
template <int S>
__device__ __forceinline__ void x_tile66(const View& in, const View& out, int j, int i0, int i1, double dtd,
                                       double half, const Bc& bc, int ghosts, XLds66& X, int lane) {
    int it = 0;
    const int niter = (i1 - i0 + 4 + 63) / 64;
    Cons Un = ld(in, min(i0 - 2 + lane, in.nx + 1), j);
    for (int k = 0; k < niter; ++k) {
        const int p = i0 - 2 + 64 * k + lane;   // cell loaded by this lane
        const Cons U0 = Un;
        Un = ld(in, min(p + 64, in.nx + 1), j);
        const Prim W0 = c2p(U0);
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
        reconstruct<kX>(W2, W1, W0, half, L, R);
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
        const Cons F = riemann<S, kX>(RU, RV, L.U, L.V, it);   // face q-1/2 = r+1/2, r = p-2
        X.F[0][2 + lane] = F.r; X.F[1][2 + lane] = F.mu; X.F[2][2 + lane] = F.mv; X.F[3][2 + lane] = F.e;
        wave_sync();
        const Cons Fm = {X.F[0][1 + lane], X.F[1][1 + lane], X.F[2][1 + lane], X.F[3][1 + lane]};
        const Cons U2 = {X.U[0][lane], X.U[1][lane], X.U[2][lane], X.U[3][lane]};
        wave_sync();
        if (lane == 63) { X.F[0][1] = F.r; X.F[1][1] = F.mu; X.F[2][1] = F.mv; X.F[3][1] = F.e; }
        if (lane >= 62)
            for (int c = 0; c < 4; ++c) X.U[c][lane - 62] = X.U[c][lane + 2];
        const int r = p - 2;
        if (r >= i0 && r < i1) {
            const Cons Unew = update(U2, dtd, Fm, F);
            st(out, r, j, Unew);
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
}

// The same x tile over the compact slice.
template <int S>
__device__ __forceinline__ void x_tilec(const View& in, const View& out, int j, int i0, int i1, double dtd,
                                        double half, const Bc& bc, int ghosts, XLdsC& X, int lane) {
    int it = 0;
    const int niter = (i1 - i0 + 4 + 63) / 64;
    const bool l1 = lane >= 1, l2 = lane >= 2;
    Cons Un = ld(in, min(i0 - 2 + lane, in.nx + 1), j);
    for (int k = 0; k < niter; ++k) {
        const int p = i0 - 2 + 64 * k + lane;
        const Cons U0 = Un;
        Un = ld(in, min(p + 64, in.nx + 1), j);
        const Prim W0 = c2p(U0);
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
        reconstruct<kX>(W2, W1, W0, half, L, R);
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
        const Cons F = riemann<S, kX>(RU, RV, L.U, L.V, it);
        X.M[0][lane] = F.r; X.M[1][lane] = F.mu; X.M[2][lane] = F.mv; X.M[3][lane] = F.e;
        wave_sync();
        double f1[4], u2[4];
        for (int c = 0; c < 4; ++c) {
            f1[c] = l1 ? X.M[c][lane - 1] : X.cF[c];
            u2[c] = l2 ? X.U[c][lane - 2] : X.cU[c][lane];
        }
        wave_sync();
        if (lane == 63) { X.cF[0] = F.r; X.cF[1] = F.mu; X.cF[2] = F.mv; X.cF[3] = F.e; }
        if (lane >= 62)
            for (int c = 0; c < 4; ++c) X.cU[c][lane - 62] = X.U[c][lane];
        const Cons Fm = {f1[0], f1[1], f1[2], f1[3]};
        const Cons U2 = {u2[0], u2[1], u2[2], u2[3]};
        const int r = p - 2;
        if (r >= i0 && r < i1) {
            const Cons Unew = update(U2, dtd, Fm, F);
            st(out, r, j, Unew);
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
}

// Layout selector: L = 0 compact, 1 = 66-slot.
template <int L> struct XSlice { using type = XLdsC; };
template <> struct XSlice<1> { using type = XLds66; };

template <int S, int L>
__device__ __forceinline__ void x_tile(const View& in, const View& out, int j, int i0, int i1, double dtd,
                                       double half, const Bc& bc, int ghosts, typename XSlice<L>::type& X,
                                       int lane) {
    if constexpr (L == 1) x_tile66<S>(in, out, j, i0, i1, dtd, half, bc, ghosts, X, lane);
    else x_tilec<S>(in, out, j, i0, i1, dtd, half, bc, ghosts, X, lane);
}

// Host helpers defined in solver.hip.
Bc to_bc(const mh_bc* b);
int launch_fill(const mh_grid_s* g, const Bc& bc, hipStream_t st);

}  // namespace mh
