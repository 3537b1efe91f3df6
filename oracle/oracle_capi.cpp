// oracle_capi.cpp -- extern "C" surface of the oracle for the Python test
// harness (ctypes).  TEST INFRASTRUCTURE ONLY: tests/, the smoke check and the
// CPU-baseline leg of the benchmark load it; the product never does.
//
// State arrays are the reference's AoS layout: Conserved[(ny+2ng)*(nx+2ng)],
// flat index j*(nx+2ng)+i (head/cpu/grid_cpu.hpp:86-88), ng = 2.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "cases_oracle.hpp"
#include "dual.hpp"
#include "euler_oracle.hpp"

using namespace orc;

namespace {

BcSpec make_bc(const int* kind, const double* state) {
    BcSpec bc;
    for (int e = 0; e < 4; ++e) {
        bc.kind[e] = kind ? kind[e] : TRANSMISSIVE;
        for (int c = 0; c < 4; ++c) bc.state[e][c] = state ? state[4 * e + c] : 0.0;
    }
    return bc;
}

template <class T> Cons<T> ld(const double* p) { return {T(p[0]), T(p[1]), T(p[2]), T(p[3])}; }
inline void st(double* p, const Cons<double>& U) { p[0] = U.rho; p[1] = U.rhou; p[2] = U.rhov; p[3] = U.E; }

std::size_t cells(int nx, int ny) { return (std::size_t)(nx + 4) * (ny + 4); }

}  // namespace

extern "C" {

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// Riemann flux for n independent faces; UL/UR/F are n x 4 conserved.
void orc_riemann_flux(int n, const double* UL, const double* UR, int dir, int solver, double* F) {
    for (int k = 0; k < n; ++k) st(F + 4 * k, riemann_flux(ld<double>(UL + 4 * k), ld<double>(UR + 4 * k), dir, solver));
}

// Exact-solver star state for n faces from primitive (rho,u,v,p) pairs.
void orc_exact_star(int n, const double* VL, const double* VR, int dir, double* p_star, double* u_star, int* iters,
                    int* ok) {
    for (int k = 0; k < n; ++k) {
        const Prim<double> L{VL[4 * k], VL[4 * k + 1], VL[4 * k + 2], VL[4 * k + 3]};
        const Prim<double> R{VR[4 * k], VR[4 * k + 1], VR[4 * k + 2], VR[4 * k + 3]};
        double ps = 0.0, us = 0.0;
        int it = 0;
        ok[k] = exact_star_pressure_velocity(L, R, dir, ps, us, &it) ? 1 : 0;
        p_star[k] = ps;
        u_star[k] = us;
        iters[k] = it;
    }
}

// MUSCL-Hancock reconstruction of n cells from (Um, Uc, Up).
void orc_reconstruct(int n, const double* Um, const double* Uc, const double* Up, double dt_over_d, int dir,
                     double* UL, double* UR) {
    for (int k = 0; k < n; ++k) {
        Cons<double> a, b;
        reconstruct(ld<double>(Um + 4 * k), ld<double>(Uc + 4 * k), ld<double>(Up + 4 * k), dt_over_d, dir, a, b);
        st(UL + 4 * k, a);
        st(UR + 4 * k, b);
    }
}

void orc_prim_to_cons(int n, const double* V, double* U) {
    for (int k = 0; k < n; ++k) st(U + 4 * k, prim_to_cons(Prim<double>{V[4 * k], V[4 * k + 1], V[4 * k + 2], V[4 * k + 3]}));
}

void orc_cons_to_prim(int n, const double* U, double* V) {
    for (int k = 0; k < n; ++k) {
        const Prim<double> W = cons_to_prim(ld<double>(U + 4 * k));
        V[4 * k] = W.rho; V[4 * k + 1] = W.u; V[4 * k + 2] = W.v; V[4 * k + 3] = W.p;
    }
}

// Initial conditions (point sampling at every cell centre, ghosts included).
// kind 0: reference shock_bubble; 1: reference blast_wave; 2: blast with
// params p[0..6] = x0,y0,r0,rho_in,p_in,rho_out,p_out; 3: shock-bubble with
// params p[0..7] = rho0,p0,mach,shock_x,bubble_rho,bubble_x,bubble_y,bubble_r.
int orc_init(int kind, const double* p, int nx, int ny, double xmin, double xmax, double ymin, double ymax,
             double* U) {
    const double dx = (xmax - xmin) / (double)nx, dy = (ymax - ymin) / (double)ny;
    const int tnx = nx + 4, tny = ny + 4;
    for (int j = 0; j < tny; ++j)
        for (int i = 0; i < tnx; ++i) {
            const double x = cell_center(xmin, i, 2, dx), y = cell_center(ymin, j, 2, dy);
            Cons<double> s;
            if (kind == 0) s = shock_bubble_state(reference_shock_bubble(), x, y);
            else if (kind == 1) s = blast_state(reference_blast(), x, y);
            else if (kind == 2) s = blast_state(BlastSpec{p[0], p[1], p[2], p[3], p[4], p[5], p[6]}, x, y);
            else if (kind == 3) s = shock_bubble_state(ShockBubbleSpec{p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7]}, x, y);
            else return -1;
            st(U + 4 * ((std::size_t)j * tnx + i), s);
        }
    return 0;
}

// Smoothed shock-bubble (params as kind 3 of orc_init, bubble_rho unused):
// per-cell bubble density rho_b[(ny+4)*(nx+4)] and bubble radius R, each with
// a tangent (d_rho_b may be null, d_R a scalar).  Writes state and tangent.
int orc_init_smooth(const double* p, int nx, int ny, double xmin, double xmax, double ymin, double ymax,
                    const double* rho_b, const double* d_rho_b, double R, double d_R, double* U, double* dU) {
    const double dx = (xmax - xmin) / (double)nx, dy = (ymax - ymin) / (double)ny;
    const double h = 0.5 * (dx + dy);
    const ShockBubbleSpec s{p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7]};
    const int tnx = nx + 4, tny = ny + 4;
    for (int j = 0; j < tny; ++j)
        for (int i = 0; i < tnx; ++i) {
            const std::size_t c = (std::size_t)j * tnx + i;
            const double x = cell_center(xmin, i, 2, dx), y = cell_center(ymin, j, 2, dy);
            const Dual rb(rho_b[c], d_rho_b ? d_rho_b[c] : 0.0);
            const Cons<Dual> q = shock_bubble_smooth_state(s, x, y, h, Dual(R, d_R), rb);
            U[4 * c] = q.rho.v; U[4 * c + 1] = q.rhou.v; U[4 * c + 2] = q.rhov.v; U[4 * c + 3] = q.E.v;
            if (dU) { dU[4 * c] = q.rho.d; dU[4 * c + 1] = q.rhou.d; dU[4 * c + 2] = q.rhov.d; dU[4 * c + 3] = q.E.d; }
        }
    return 0;
}

void orc_apply_boundary(int nx, int ny, double* U, const int* bc_kind, const double* bc_state) {
    Grid<double> g;
    g.nx = nx; g.ny = ny; g.U = reinterpret_cast<Cons<double>*>(U);
    apply_boundary(g, make_bc(bc_kind, bc_state));
}

// Max signal speed over interior cells and dt = cfl*min(dx,dy)/smax
// (src/cpu/solver_cpu.cpp:287-309).  Returns -1 when smax <= 0 (the
// reference throws there).
double orc_compute_dt(int nx, int ny, double dx, double dy, const double* U, double cfl, double* smax_out) {
    Grid<double> g;
    g.nx = nx; g.ny = ny; g.dx = dx; g.dy = dy;
    g.U = const_cast<Cons<double>*>(reinterpret_cast<const Cons<double>*>(U));
    const double m = max_speed(g);
    if (smax_out) *smax_out = m;
    if (m <= 0.0) return -1.0;
    return cfl * std::min(dx, dy) / m;
}

// One split step Uold -> Unew (Utmp scratch), order 2 = MUSCL-Hancock x-then-y,
// order 1 = unsplit first-order Godunov.
int orc_advance(int nx, int ny, double dx, double dy, const double* Uold, double* Utmp, double* Unew, double dt,
                int solver, int order, const int* bc_kind, const double* bc_state) {
    Grid<double> a, b, c;
    for (Grid<double>* g : {&a, &b, &c}) { g->nx = nx; g->ny = ny; g->dx = dx; g->dy = dy; }
    a.U = const_cast<Cons<double>*>(reinterpret_cast<const Cons<double>*>(Uold));
    b.U = reinterpret_cast<Cons<double>*>(Utmp);
    c.U = reinterpret_cast<Cons<double>*>(Unew);
    const BcSpec bc = make_bc(bc_kind, bc_state);
    std::vector<Cons<double>> ws;
    if (order == 1) advance_first_order(a, c, dt, solver, bc);
    else advance_second_order(a, b, c, dt, solver, bc, ws);
    return 0;
}

// Time loop of scripts/cpu/main_cpu.cpp:206-262 on state U (in/out):
//   while (t < t_end): BC(U); dt = compute_dt(U); clamp to t_end; advance; t += dt.
// t_end <= 0 runs exactly max_steps steps without a clamp.  With dt_in non-null
// the dt sequence is taken from it instead of being computed (frozen dt).
// dt_log (max_steps) receives each dt; snapshots follow the reference rule
// (first step with t >= k*t_end/nsnap) and are copied to snap_states.
// Returns the number of steps, or -1 on a non-positive wave speed.
int orc_run(int nx, int ny, double dx, double dy, double* U, int solver, int order, double cfl, double t_end,
            int max_steps, const int* bc_kind, const double* bc_state, const double* dt_in, double* dt_log,
            double* t_out, int nsnap, double* snap_t, double* snap_states, double* seconds) {
    const std::size_t n = cells(nx, ny);
    std::vector<double> buf1(4 * n), buf2(4 * n);
    std::memcpy(buf1.data(), U, 4 * n * sizeof(double));
    Grid<double> cur, tmp, nxt;
    for (Grid<double>* g : {&cur, &tmp, &nxt}) { g->nx = nx; g->ny = ny; g->dx = dx; g->dy = dy; }
    cur.U = reinterpret_cast<Cons<double>*>(U);
    tmp.U = reinterpret_cast<Cons<double>*>(buf1.data());
    nxt.U = reinterpret_cast<Cons<double>*>(buf2.data());
    const BcSpec bc = make_bc(bc_kind, bc_state);
    std::vector<Cons<double>> ws;
    double t = 0.0;
    int step = 0, next_snap = 0;
    const auto t0 = std::chrono::steady_clock::now();
    while ((t_end > 0.0 ? t < t_end : true) && step < max_steps) {
        apply_boundary(cur, bc);
        double dt;
        if (dt_in) {
            dt = dt_in[step];
        } else {
            const double m = max_speed(cur);
            if (m <= 0.0) return -1;
            dt = cfl * std::min(dx, dy) / m;
        }
        if (t_end > 0.0 && t + dt > t_end) dt = t_end - t;
        if (order == 1) advance_first_order(cur, nxt, dt, solver, bc);
        else advance_second_order(cur, tmp, nxt, dt, solver, bc, ws);
        std::swap(cur.U, nxt.U);
        t += dt;
        if (dt_log) dt_log[step] = dt;
        ++step;
        while (t_end > 0.0 && next_snap < nsnap && t >= t_end * (double)(next_snap + 1) / (double)nsnap) {
            if (snap_t) snap_t[next_snap] = t;
            if (snap_states) std::memcpy(snap_states + 4 * n * next_snap, cur.U, 4 * n * sizeof(double));
            ++next_snap;
        }
    }
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (cur.U != reinterpret_cast<Cons<double>*>(U)) std::memcpy(U, cur.U, 4 * n * sizeof(double));
    if (t_out) *t_out = t;
    return step;
}

// Forward-mode AD (frozen dt sequence): propagates tangent dU through
// `nsteps` second-order steps from state U (both in/out).
int orc_run_tangent(int nx, int ny, double dx, double dy, double* U, double* dU, int solver, int nsteps,
                    const double* dt_seq, const int* bc_kind, const double* bc_state) {
    const std::size_t n = cells(nx, ny);
    std::vector<Cons<Dual>> a(n), b(n), c(n);
    for (std::size_t k = 0; k < n; ++k)
        for (int q = 0; q < 4; ++q) (&a[k].rho)[q] = Dual(U[4 * k + q], dU[4 * k + q]);
    Grid<Dual> cur, tmp, nxt;
    for (Grid<Dual>* g : {&cur, &tmp, &nxt}) { g->nx = nx; g->ny = ny; g->dx = dx; g->dy = dy; }
    cur.U = a.data(); tmp.U = b.data(); nxt.U = c.data();
    const BcSpec bc = make_bc(bc_kind, bc_state);
    std::vector<Cons<Dual>> ws;
    for (int s = 0; s < nsteps; ++s) {
        apply_boundary(cur, bc);
        advance_second_order(cur, tmp, nxt, Dual(dt_seq[s]), solver, bc, ws);
        std::swap(cur.U, nxt.U);
    }
    for (std::size_t k = 0; k < n; ++k)
        for (int q = 0; q < 4; ++q) { U[4 * k + q] = (&cur.U[k].rho)[q].v; dU[4 * k + q] = (&cur.U[k].rho)[q].d; }
    return 0;
}

// Tangent of one sweep (frozen dt): inputs U, dU over every cell (ghosts
// included), outputs Uout, dUout over interior cells (ghosts of Uout copy U).
int orc_sweep_tangent(int nx, int ny, double dx, double dy, const double* U, const double* dU, double dt, int dir,
                      int solver, double* Uout, double* dUout) {
    const std::size_t n = cells(nx, ny);
    std::vector<Cons<Dual>> a(n), b(n);
    for (std::size_t k = 0; k < n; ++k)
        for (int q = 0; q < 4; ++q) (&a[k].rho)[q] = Dual(U[4 * k + q], dU[4 * k + q]);
    b = a;
    Grid<Dual> gi, go;
    for (Grid<Dual>* g : {&gi, &go}) { g->nx = nx; g->ny = ny; g->dx = dx; g->dy = dy; }
    gi.U = a.data();
    go.U = b.data();
    std::vector<Cons<Dual>> ws;
    sweep(gi, go, Dual(dt), dir, solver, ws);
    for (std::size_t k = 0; k < n; ++k)
        for (int q = 0; q < 4; ++q) { Uout[4 * k + q] = (&b[k].rho)[q].v; dUout[4 * k + q] = (&b[k].rho)[q].d; }
    return 0;
}

// One x or y sweep alone (dir 0/1) over interior cells: Uout = sweep(Uin).
int orc_sweep(int nx, int ny, double dx, double dy, const double* Uin, double* Uout, double dt, int dir, int solver) {
    Grid<double> a, b;
    for (Grid<double>* g : {&a, &b}) { g->nx = nx; g->ny = ny; g->dx = dx; g->dy = dy; }
    a.U = const_cast<Cons<double>*>(reinterpret_cast<const Cons<double>*>(Uin));
    b.U = reinterpret_cast<Cons<double>*>(Uout);
    std::vector<Cons<double>> ws;
    sweep(a, b, dt, dir, solver, ws);
    return 0;
}

}  // extern "C"
