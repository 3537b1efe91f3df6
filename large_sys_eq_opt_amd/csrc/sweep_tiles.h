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

constexpr int kSlots = 66;
struct XLds {            // one wavefront's exchange slice (8448 bytes)
    double W[4][kSlots];
    double R[8][kSlots];
    double F[4][kSlots];
};

// y tile over columns [64*cg, 64*cg+64) and rows [j0, j1).  Returns the
// lane's running max of the output's signal speed when `speed` is set
// (std::max semantics: a NaN speed never enters), 0 otherwise.
template <int S>
__device__ __forceinline__ double y_tile(const View& in, const View& out, int cg, int j0, int j1, double dtd,
                                         double half, const Bc& bc, int ghosts, int speed, int lane) {
    const int i = cg * 64 + lane;
    const bool active = i < in.nx;
    const int ic = active ? i : in.nx - 1;
    int it = 0;
    Cons U2 = ld(in, ic, j0 - 2), U1 = ld(in, ic, j0 - 1);
    Prim W2 = c2p(U2), W1 = c2p(U1);
    Cons Un = ld(in, ic, j0);
    Edge L, R, Rp;
    {   // row p = j0: reconstruct j0-1
        const Cons U0 = Un;
        Un = ld(in, ic, j0 + 1);
        const Prim W0 = c2p(U0);
        reconstruct<kY>(W2, W1, W0, half, L, Rp);
        U2 = U1; U1 = U0; W2 = W1; W1 = W0;
    }
    Cons Fp;
    {   // row p = j0+1: reconstruct j0, face j0-1/2
        const Cons U0 = Un;
        Un = ld(in, ic, j0 + 2);
        const Prim W0 = c2p(U0);
        reconstruct<kY>(W2, W1, W0, half, L, R);
        Fp = riemann<S, kY>(Rp.U, Rp.V, L.U, L.V, it);
        Rp = R;
        U2 = U1; U1 = U0; W2 = W1; W1 = W0;
    }
    double m = 0.0;
    for (int p = j0 + 2; p <= j1 + 1; ++p) {
        const Cons U0 = Un;
        Un = ld(in, ic, min(p + 1, j1 + 1));
        const Prim W0 = c2p(U0);
        reconstruct<kY>(W2, W1, W0, half, L, R);
        const Cons F = riemann<S, kY>(Rp.U, Rp.V, L.U, L.V, it);
        const int r = p - 2;
        const Cons Unew = update(U2, dtd, Fp, F);
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
        U2 = U1; U1 = U0; W2 = W1; W1 = W0;
    }
    return m;
}

// x tile over row j, cells [i0, i1).
template <int S>
__device__ __forceinline__ void x_tile(const View& in, const View& out, int j, int i0, int i1, double dtd,
                                       double half, const Bc& bc, int ghosts, XLds& X, int lane) {
    int it = 0;
    const int niter = (i1 - i0 + 4 + 63) / 64;
    Cons Un = ld(in, min(i0 - 2 + lane, in.nx + 1), j);
    for (int k = 0; k < niter; ++k) {
        const int p = i0 - 2 + 64 * k + lane;   // cell loaded by this lane
        const Cons U0 = Un;
        Un = ld(in, min(p + 64, in.nx + 1), j);
        const Prim W0 = c2p(U0);
        X.W[0][2 + lane] = W0.r; X.W[1][2 + lane] = W0.u; X.W[2][2 + lane] = W0.v; X.W[3][2 + lane] = W0.p;
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
        wave_sync();
        if (lane == 63) { X.F[0][1] = F.r; X.F[1][1] = F.mu; X.F[2][1] = F.mv; X.F[3][1] = F.e; }
        const int r = p - 2;
        if (r >= i0 && r < i1) {
            const Cons Unew = update(ld(in, r, j), dtd, Fm, F);
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

// Host helpers defined in solver.hip.
Bc to_bc(const mh_bc* b);
int launch_fill(const mh_grid_s* g, const Bc& bc, hipStream_t st);

}  // namespace mh
