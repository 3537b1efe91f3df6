// ref_harness.cpp -- extern "C" entry points over the reference CPU solver,
// compiled from the reference's own sources (see oracle/Makefile) into
// oracle/_ref/libmuscl_ref.so.  TEST INFRASTRUCTURE ONLY: it generates golden
// vectors and checks the oracle; nothing in the product links it.
//
// The reference's reconstruction lives in an anonymous namespace of
// src/cpu/solver_cpu.cpp, so that file is compiled as part of this
// translation unit (it is not copied: the include resolves to the reference
// tree at build time).
#include <chrono>
#include <cstring>
#include <vector>

#include "cpu/solver_cpu.cpp"  // reference: src/cpu/solver_cpu.cpp
#include "cpu/boundary_cpu.hpp"
#include "init.hpp"
#include "riemann.hpp"
#include "test_cases.hpp"

namespace {

Conserved ldc(const double* p) { return Conserved(p[0], p[1], p[2], p[3]); }
void stc(double* p, const Conserved& U) { p[0] = U.rho; p[1] = U.rhou; p[2] = U.rhov; p[3] = U.E; }

RiemannSolver solver_of(int s) {
    switch (s) {
        case 0: return RiemannSolver::HLL;
        case 1: return RiemannSolver::HLLC;
        case 2: return RiemannSolver::Exact;
        default: return RiemannSolver::FORCE;
    }
}

Direction dir_of(int d) { return d == 0 ? Direction::X : Direction::Y; }

void to_grid(Grid2D& g, const double* U) {
    std::memcpy(static_cast<void*>(g.data().data()), U, g.data().size() * sizeof(Conserved));
}
void from_grid(const Grid2D& g, double* U) {
    std::memcpy(U, static_cast<const void*>(g.data().data()), g.data().size() * sizeof(Conserved));
}

void set_left_inflow(Grid2D& g, const double* s) {
    for (int j = g.j_begin(); j < g.j_end(); ++j)
        for (int k = 0; k < g.ng(); ++k) g(g.i_begin() - 1 - k, j) = ldc(s);
}

}  // namespace

extern "C" {

void ref_riemann_flux(int n, const double* UL, const double* UR, int dir, int solver, double* F) {
    for (int k = 0; k < n; ++k)
        stc(F + 4 * k, riemann_flux(ldc(UL + 4 * k), ldc(UR + 4 * k), dir_of(dir), solver_of(solver)));
}

void ref_reconstruct(int n, const double* Um, const double* Uc, const double* Up, double dtd, int dir, double* UL,
                     double* UR) {
    for (int k = 0; k < n; ++k) {
        Conserved a, b;
        reconstruct_cell_muscl_hancock(ldc(Um + 4 * k), ldc(Uc + 4 * k), ldc(Up + 4 * k), dtd, dir_of(dir), a, b);
        stc(UL + 4 * k, a);
        stc(UR + 4 * k, b);
    }
}

void ref_exact_star(int n, const double* VL, const double* VR, int dir, double* p_star, double* u_star, int* ok) {
    for (int k = 0; k < n; ++k) {
        const Primitive L(VL[4 * k], VL[4 * k + 1], VL[4 * k + 2], VL[4 * k + 3]);
        const Primitive R(VR[4 * k], VR[4 * k + 1], VR[4 * k + 2], VR[4 * k + 3]);
        double ps = 0.0, us = 0.0;
        ok[k] = exact_star_pressure_velocity(L, R, dir_of(dir), ps, us) ? 1 : 0;
        p_star[k] = ps;
        u_star[k] = us;
    }
}

// Reference case configuration (case 0 shock_bubble, 1 blast_wave) at scale n.
int ref_case_config(int case_id, int n, int* nx, int* ny, double* bounds, double* cfl, double* t_end) {
    const CaseConfig c = get_n_case_config(case_id == 0 ? CaseId::ShockBubble : CaseId::BlastWave, n);
    *nx = c.nx; *ny = c.ny;
    bounds[0] = c.x_min; bounds[1] = c.x_max; bounds[2] = c.y_min; bounds[3] = c.y_max;
    *cfl = c.cfl; *t_end = c.t_end;
    return c.ng;
}

// Reference initial grid (make_n_grid) into U [(ny+4)*(nx+4)*4].
void ref_init_case(int case_id, int n, double* U) {
    const Grid2D g = make_n_grid(case_id == 0 ? CaseId::ShockBubble : CaseId::BlastWave, n);
    from_grid(g, U);
}

double ref_compute_dt(int nx, int ny, const double* bounds, const double* U, double cfl) {
    Grid2D g(nx, ny, 2, bounds[0], bounds[1], bounds[2], bounds[3]);
    to_grid(g, U);
    return compute_dt(g, cfl);
}

// The loop of scripts/cpu/main_cpu.cpp:221-262 on state U (in/out, AoS).
// inflow_left non-null: after the transmissive boundary, the left ghost
// columns are set to that conserved state (Dirichlet inflow).  t_end <= 0 runs
// exactly max_steps steps; dt_in non-null imposes the dt sequence.
int ref_run(int nx, int ny, const double* bounds, double* U, int solver, int order, double cfl, double t_end,
            int max_steps, const double* inflow_left, const double* dt_in, double* dt_log, double* t_out, int nsnap,
            double* snap_t, double* seconds) {
    Grid2D cur(nx, ny, 2, bounds[0], bounds[1], bounds[2], bounds[3]);
    Grid2D mid(nx, ny, 2, bounds[0], bounds[1], bounds[2], bounds[3]);
    Grid2D nxt(nx, ny, 2, bounds[0], bounds[1], bounds[2], bounds[3]);
    to_grid(cur, U);
    CpuWorkspace ws;
    ws.init(nx, ny);
    const RiemannSolver rs = solver_of(solver);
    double t = 0.0;
    int step = 0, next_snap = 0;
    const auto t0 = std::chrono::steady_clock::now();
    while ((t_end > 0.0 ? t < t_end : true) && step < max_steps) {
        apply_transmissive_boundary(cur);
        if (inflow_left) set_left_inflow(cur, inflow_left);
        double dt = dt_in ? dt_in[step] : compute_dt(cur, cfl);
        if (t_end > 0.0 && t + dt > t_end) dt = t_end - t;
        if (order == 1) advance_first_order(cur, nxt, dt, rs);
        else advance_second_order(cur, mid, nxt, dt, ws, rs);
        std::swap(cur, nxt);
        t += dt;
        if (dt_log) dt_log[step] = dt;
        ++step;
        while (t_end > 0.0 && next_snap < nsnap && t >= t_end * (double)(next_snap + 1) / (double)nsnap) {
            if (snap_t) snap_t[next_snap] = t;
            ++next_snap;
        }
    }
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    from_grid(cur, U);
    if (t_out) *t_out = t;
    return step;
}

}  // extern "C"
