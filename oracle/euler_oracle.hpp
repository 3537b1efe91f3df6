// euler_oracle.hpp -- CPU restatement of the reference MUSCL-Hancock Euler
// solver.  TEST INFRASTRUCTURE ONLY: the parity checker and the CPU baseline
// ("port") use it; the product path never links, loads or calls it.
//
// Every function follows the operation order of the reference CPU solver so
// that, built with g++ -O3 -ffp-contract=off (no FMA), it reproduces the
// reference bit for bit.  The scalar type is a template parameter: T = double
// is the oracle proper, T = Dual (dual.hpp) is the forward-mode
// operator-overloading AD reference used to check the discrete adjoint.
//
// Source map (paths relative to the reference repository):
//   physics              head/physics.hpp:11-94, head/types.hpp:13-175
//   Riemann solvers      head/riemann.hpp:24-672
//   reconstruction       src/cpu/solver_cpu.cpp:17-99
//   sweeps               src/cpu/solver_cpu.cpp:106-279, :400-499
//   dt                   src/cpu/solver_cpu.cpp:287-309
//   boundary             head/cpu/boundary_cpu.hpp:5-36
//   first-order step     src/cpu/solver_cpu.cpp:315-369
//   grid indexing        head/cpu/grid_cpu.hpp:14-88
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// Scalar helpers.  For T = double they are the libm / <algorithm> calls the
// reference makes; dual.hpp overloads them for the AD type.
// ---------------------------------------------------------------------------
inline double val(double x) { return x; }
inline bool fin(double x) { return std::isfinite(x); }
inline double lmin(double a, double b) { return std::fmin(a, b); }   // libm fmin
inline double lmax(double a, double b) { return std::fmax(a, b); }   // libm fmax
inline double smin(double a, double b) { return std::min(a, b); }    // std::min
inline double smax(double a, double b) { return std::max(a, b); }    // std::max

constexpr double kGamma = 1.4;            // head/physics.hpp:11
constexpr double kRhoFloor = 1.0e-12;     // src/cpu/solver_cpu.cpp:19
constexpr double kPFloor = 1.0e-12;       // src/cpu/solver_cpu.cpp:20

enum Dir { X = 0, Y = 1 };                                  // head/riemann.hpp:9-11
enum Solver { HLL = 0, HLLC = 1, EXACT = 2, FORCE = 3 };    // head/riemann.hpp:13-18

template <class T> struct Cons { T rho, rhou, rhov, E; };
template <class T> struct Prim { T rho, u, v, p; };

// operators of head/types.hpp:49-175 (component-wise, same association)
template <class T> inline Cons<T> operator+(const Cons<T>& a, const Cons<T>& b) { return {a.rho + b.rho, a.rhou + b.rhou, a.rhov + b.rhov, a.E + b.E}; }
template <class T> inline Cons<T> operator-(const Cons<T>& a, const Cons<T>& b) { return {a.rho - b.rho, a.rhou - b.rhou, a.rhov - b.rhov, a.E - b.E}; }
template <class T, class S> inline Cons<T> operator*(const S& s, const Cons<T>& a) { return {s * a.rho, s * a.rhou, s * a.rhov, s * a.E}; }
template <class T, class S> inline Cons<T> operator/(const Cons<T>& a, const S& s) { return {a.rho / s, a.rhou / s, a.rhov / s, a.E / s}; }
template <class T> inline Prim<T> operator+(const Prim<T>& a, const Prim<T>& b) { return {a.rho + b.rho, a.u + b.u, a.v + b.v, a.p + b.p}; }
template <class T> inline Prim<T> operator-(const Prim<T>& a, const Prim<T>& b) { return {a.rho - b.rho, a.u - b.u, a.v - b.v, a.p - b.p}; }
template <class T, class S> inline Prim<T> operator*(const S& s, const Prim<T>& a) { return {s * a.rho, s * a.u, s * a.v, s * a.p}; }

// ---------------------------------------------------------------------------
// physics (head/physics.hpp:13-74)
// ---------------------------------------------------------------------------
template <class T> inline Prim<T> cons_to_prim(const Cons<T>& U) {        // :13-20
    const T inv_rho = 1.0 / U.rho;
    const T u = U.rhou * inv_rho;
    const T v = U.rhov * inv_rho;
    const T kinetic = 0.5 * U.rho * (u * u + v * v);
    const T p = (kGamma - 1.0) * (U.E - kinetic);
    return {U.rho, u, v, p};
}

template <class T> inline Cons<T> prim_to_cons(const Prim<T>& V) {        // :22-28
    const T rhou = V.rho * V.u;
    const T rhov = V.rho * V.v;
    const T kinetic = 0.5 * V.rho * (V.u * V.u + V.v * V.v);
    const T E = V.p / (kGamma - 1.0) + kinetic;
    return {V.rho, rhou, rhov, E};
}

template <class T> inline T sound_speed(const Prim<T>& V) {                // :39-41
    using std::sqrt;
    return sqrt(kGamma * V.p / V.rho);
}

template <class T> inline Cons<T> physical_flux(const Cons<T>& U, int dir) {  // riemann.hpp:24-30
    const Prim<T> V = cons_to_prim(U);
    if (dir == X)                                                          // physics.hpp:56-63
        return {U.rhou, U.rhou * V.u + V.p, U.rhov * V.u, (U.E + V.p) * V.u};
    return {U.rhov, U.rhou * V.v, U.rhov * V.v + V.p, (U.E + V.p) * V.v};  // :65-74
}

template <class T> inline T normal_velocity(const Prim<T>& V, int dir) { return dir == X ? V.u : V.v; }
template <class T> inline T tangential_velocity(const Prim<T>& V, int dir) { return dir == X ? V.v : V.u; }

template <class T>
inline Cons<T> cons_from_normal_tangential(T rho, T un, T ut, T p, int dir) {  // riemann.hpp:40-64
    T u = 0.0, v = 0.0;
    if (dir == X) { u = un; v = ut; } else { u = ut; v = un; }
    return {rho, rho * u, rho * v, p / (kGamma - 1.0) + 0.5 * rho * (u * u + v * v)};
}

template <class T> inline bool physical_riemann(const Prim<T>& V) {        // riemann.hpp:74-81
    return fin(V.rho) && fin(V.u) && fin(V.v) && fin(V.p) && V.rho > 0.0 && V.p > 0.0;
}

// ---------------------------------------------------------------------------
// HLL (head/riemann.hpp:87-129)
// ---------------------------------------------------------------------------
template <class T> inline Cons<T> hll_flux(const Cons<T>& UL, const Cons<T>& UR, int dir) {
    const Prim<T> VL = cons_to_prim(UL);
    const Prim<T> VR = cons_to_prim(UR);
    if (!physical_riemann(VL) || !physical_riemann(VR)) {
        const Cons<T> FL = physical_flux(UL, dir);
        const Cons<T> FR = physical_flux(UR, dir);
        return 0.5 * (FL + FR);
    }
    const T aL = sound_speed(VL), aR = sound_speed(VR);
    const T unL = normal_velocity(VL, dir), unR = normal_velocity(VR, dir);
    const T SL = lmin(unL - aL, unR - aR);
    const T SR = lmax(unL + aL, unR + aR);
    const Cons<T> FL = physical_flux(UL, dir);
    const Cons<T> FR = physical_flux(UR, dir);
    if (SL >= 0.0) return FL;
    if (SR <= 0.0) return FR;
    const T denom = SR - SL;
    using std::fabs;
    if (fabs(denom) < 1.0e-14) return 0.5 * (FL + FR);
    return (SR * FL - SL * FR + (SL * SR) * (UR - UL)) / denom;
}

// ---------------------------------------------------------------------------
// HLLC (head/riemann.hpp:135-271)
// ---------------------------------------------------------------------------
template <class T>
inline Cons<T> hllc_star_state(const Cons<T>& U, const Prim<T>& V, T S, T Sstar, int dir) {
    using std::fabs;
    const T rho = V.rho, p = V.p;
    const T un = normal_velocity(V, dir), ut = tangential_velocity(V, dir);
    const T denom = S - Sstar;
    if (fabs(denom) < 1.0e-14) return U;
    const T rho_star = rho * (S - un) / denom;
    if (!fin(rho_star) || rho_star <= 0.0) return U;
    const T E_per_rho = U.E / rho;
    const T E_star = rho_star * (E_per_rho + (Sstar - un) * (Sstar + p / (rho * (S - un))));
    if (!fin(E_star)) return U;
    if (dir == X) return {rho_star, rho_star * Sstar, rho_star * ut, E_star};
    return {rho_star, rho_star * ut, rho_star * Sstar, E_star};
}

template <class T> inline Cons<T> hllc_flux(const Cons<T>& UL, const Cons<T>& UR, int dir) {
    using std::fabs;
    const Prim<T> VL = cons_to_prim(UL);
    const Prim<T> VR = cons_to_prim(UR);
    if (!physical_riemann(VL) || !physical_riemann(VR)) return hll_flux(UL, UR, dir);
    const T aL = sound_speed(VL), aR = sound_speed(VR);
    const T unL = normal_velocity(VL, dir), unR = normal_velocity(VR, dir);
    const T pL = VL.p, pR = VR.p, rhoL = VL.rho, rhoR = VR.rho;
    const T SL = lmin(unL - aL, unR - aR);
    const T SR = lmax(unL + aL, unR + aR);
    const Cons<T> FL = physical_flux(UL, dir);
    const Cons<T> FR = physical_flux(UR, dir);
    if (SL >= 0.0) return FL;
    if (SR <= 0.0) return FR;
    const T numerator = pR - pL + rhoL * unL * (SL - unL) - rhoR * unR * (SR - unR);
    const T denominator = rhoL * (SL - unL) - rhoR * (SR - unR);
    if (fabs(denominator) < 1.0e-14) return hll_flux(UL, UR, dir);
    const T Sstar = numerator / denominator;
    if (!fin(Sstar)) return hll_flux(UL, UR, dir);
    if (Sstar >= 0.0) {
        const Cons<T> UsL = hllc_star_state(UL, VL, SL, Sstar, dir);
        return FL + SL * (UsL - UL);
    }
    const Cons<T> UsR = hllc_star_state(UR, VR, SR, Sstar, dir);
    return FR + SR * (UsR - UR);
}

// ---------------------------------------------------------------------------
// FORCE-type flux with local alpha (head/riemann.hpp:284-325)
// ---------------------------------------------------------------------------
template <class T> inline Cons<T> force_flux(const Cons<T>& UL, const Cons<T>& UR, int dir) {
    using std::fabs;
    const Prim<T> VL = cons_to_prim(UL);
    const Prim<T> VR = cons_to_prim(UR);
    if (!physical_riemann(VL) || !physical_riemann(VR)) return hll_flux(UL, UR, dir);
    const T aL = sound_speed(VL), aR = sound_speed(VR);
    const T unL = normal_velocity(VL, dir), unR = normal_velocity(VR, dir);
    const T alpha = lmax(fabs(unL) + aL, fabs(unR) + aR);
    const Cons<T> FL = physical_flux(UL, dir);
    const Cons<T> FR = physical_flux(UR, dir);
    if (!fin(alpha) || alpha <= 1.0e-14) return 0.5 * (FL + FR);
    const Cons<T> F_lf = 0.5 * (FL + FR) - 0.5 * alpha * (UR - UL);
    const Cons<T> U_ri = 0.5 * (UL + UR) - 0.5 * (1.0 / alpha) * (FR - FL);
    const Prim<T> V_ri = cons_to_prim(U_ri);
    if (!physical_riemann(V_ri)) return F_lf;
    const Cons<T> F_ri = physical_flux(U_ri, dir);
    return 0.5 * (F_lf + F_ri);
}

// ---------------------------------------------------------------------------
// Exact solver (head/riemann.hpp:331-644)
// ---------------------------------------------------------------------------
template <class T>
inline void exact_pressure_function(T p, T rhoK, T pK, T aK, T& f, T& fd) {   // :331-359
    using std::pow;
    using std::sqrt;
    const double gamma = kGamma;
    if (p > pK) {
        const T A = 2.0 / ((gamma + 1.0) * rhoK);
        const T B = (gamma - 1.0) / (gamma + 1.0) * pK;
        const T q = sqrt(A / (p + B));
        f = (p - pK) * q;
        fd = q * (1.0 - 0.5 * (p - pK) / (p + B));
    } else {
        const T pratio = p / pK;
        const double expo = (gamma - 1.0) / (2.0 * gamma);
        f = 2.0 * aK / (gamma - 1.0) * (pow(pratio, expo) - 1.0);
        fd = (1.0 / (rhoK * aK)) * pow(pratio, -(gamma + 1.0) / (2.0 * gamma));
    }
}

template <class T>
inline T exact_guess_pressure(const Prim<T>& VL, const Prim<T>& VR, int dir) {  // :361-377
    const T aL = sound_speed(VL), aR = sound_speed(VR);
    const T unL = normal_velocity(VL, dir), unR = normal_velocity(VR, dir);
    const T p_pvrs = 0.5 * (VL.p + VR.p) - 0.125 * (unR - unL) * (VL.rho + VR.rho) * (aL + aR);
    return lmax(T(1.0e-12), p_pvrs);
}

template <class T>
inline bool exact_star_pressure_velocity(const Prim<T>& VL, const Prim<T>& VR, int dir,
                                         T& p_star, T& u_star, int* iters = nullptr) {  // :379-443
    using std::fabs;
    const T rhoL = VL.rho, rhoR = VR.rho, pL = VL.p, pR = VR.p;
    const T aL = sound_speed(VL), aR = sound_speed(VR);
    const T unL = normal_velocity(VL, dir), unR = normal_velocity(VR, dir);
    p_star = exact_guess_pressure(VL, VR, dir);
    int it_done = 0;
    for (int iter = 0; iter < 20; ++iter) {
        T fL = 0.0, fR = 0.0, fdL = 0.0, fdR = 0.0;
        exact_pressure_function(p_star, rhoL, pL, aL, fL, fdL);
        exact_pressure_function(p_star, rhoR, pR, aR, fR, fdR);
        const T f = fL + fR + (unR - unL);
        const T fd = fdL + fdR;
        it_done = iter + 1;
        if (fabs(fd) < 1.0e-14) { if (iters) *iters = it_done; return false; }
        const T p_old = p_star;
        p_star = p_old - f / fd;
        if (!fin(p_star) || p_star <= 0.0) p_star = 0.5 * p_old;
        if (fabs(p_star - p_old) / (0.5 * (p_star + p_old) + 1.0e-14) < 1.0e-8) break;
    }
    if (iters) *iters = it_done;
    if (!fin(p_star) || p_star <= 0.0) return false;
    T fL = 0.0, fR = 0.0, d0 = 0.0, d1 = 0.0;
    exact_pressure_function(p_star, rhoL, pL, aL, fL, d0);
    exact_pressure_function(p_star, rhoR, pR, aR, fR, d1);
    u_star = 0.5 * (unL + unR + fR - fL);
    return fin(u_star);
}

template <class T>
inline Cons<T> exact_sample_left(const Prim<T>& VL, T p_star, T u_star, int dir) {  // :445-517
    using std::pow;
    using std::sqrt;
    const double gamma = kGamma;
    const T rhoL = VL.rho, pL = VL.p;
    const T unL = normal_velocity(VL, dir), utL = tangential_velocity(VL, dir);
    const T aL = sound_speed(VL);
    const double xi = 0.0;
    T rho = rhoL, un = unL, p = pL;
    if (p_star > pL) {
        const T p_ratio = p_star / pL;
        const T SL = unL - aL * sqrt((gamma + 1.0) / (2.0 * gamma) * p_ratio + (gamma - 1.0) / (2.0 * gamma));
        if (xi <= SL) {
            rho = rhoL; un = unL; p = pL;
        } else {
            rho = rhoL * ((p_ratio + (gamma - 1.0) / (gamma + 1.0)) /
                          ((gamma - 1.0) / (gamma + 1.0) * p_ratio + 1.0));
            un = u_star; p = p_star;
        }
    } else {
        const T a_star = aL * pow(p_star / pL, (gamma - 1.0) / (2.0 * gamma));
        const T SHL = unL - aL;
        const T STL = u_star - a_star;
        if (xi <= SHL) {
            rho = rhoL; un = unL; p = pL;
        } else if (xi >= STL) {
            rho = rhoL * pow(p_star / pL, 1.0 / gamma);
            un = u_star; p = p_star;
        } else {
            const T a = 2.0 / (gamma + 1.0) * (aL + 0.5 * (gamma - 1.0) * (unL - xi));
            un = 2.0 / (gamma + 1.0) * (aL + 0.5 * (gamma - 1.0) * unL + xi);
            const T aratio = a / aL;
            rho = rhoL * pow(aratio, 2.0 / (gamma - 1.0));
            p = pL * pow(aratio, 2.0 * gamma / (gamma - 1.0));
        }
    }
    return cons_from_normal_tangential(rho, un, utL, p, dir);
}

template <class T>
inline Cons<T> exact_sample_right(const Prim<T>& VR, T p_star, T u_star, int dir) {  // :519-591
    using std::pow;
    using std::sqrt;
    const double gamma = kGamma;
    const T rhoR = VR.rho, pR = VR.p;
    const T unR = normal_velocity(VR, dir), utR = tangential_velocity(VR, dir);
    const T aR = sound_speed(VR);
    const double xi = 0.0;
    T rho = rhoR, un = unR, p = pR;
    if (p_star > pR) {
        const T p_ratio = p_star / pR;
        const T SR = unR + aR * sqrt((gamma + 1.0) / (2.0 * gamma) * p_ratio + (gamma - 1.0) / (2.0 * gamma));
        if (xi >= SR) {
            rho = rhoR; un = unR; p = pR;
        } else {
            rho = rhoR * ((p_ratio + (gamma - 1.0) / (gamma + 1.0)) /
                          ((gamma - 1.0) / (gamma + 1.0) * p_ratio + 1.0));
            un = u_star; p = p_star;
        }
    } else {
        const T a_star = aR * pow(p_star / pR, (gamma - 1.0) / (2.0 * gamma));
        const T SHR = unR + aR;
        const T STR = u_star + a_star;
        if (xi >= SHR) {
            rho = rhoR; un = unR; p = pR;
        } else if (xi <= STR) {
            rho = rhoR * pow(p_star / pR, 1.0 / gamma);
            un = u_star; p = p_star;
        } else {
            const T a = 2.0 / (gamma + 1.0) * (aR - 0.5 * (gamma - 1.0) * (unR - xi));
            un = 2.0 / (gamma + 1.0) * (-aR + 0.5 * (gamma - 1.0) * unR + xi);
            const T aratio = a / aR;
            rho = rhoR * pow(aratio, 2.0 / (gamma - 1.0));
            p = pR * pow(aratio, 2.0 * gamma / (gamma - 1.0));
        }
    }
    return cons_from_normal_tangential(rho, un, utR, p, dir);
}

template <class T>
inline Cons<T> exact_flux(const Cons<T>& UL, const Cons<T>& UR, int dir, int* iters = nullptr) {  // :601-644
    const Prim<T> VL = cons_to_prim(UL);
    const Prim<T> VR = cons_to_prim(UR);
    if (!physical_riemann(VL) || !physical_riemann(VR)) return hll_flux(UL, UR, dir);
    T p_star = 0.0, u_star = 0.0;
    const bool ok = exact_star_pressure_velocity(VL, VR, dir, p_star, u_star, iters);
    if (!ok || !fin(p_star) || !fin(u_star) || p_star <= 0.0) return hll_flux(UL, UR, dir);
    Cons<T> Us;
    if (u_star >= 0.0) Us = exact_sample_left(VL, p_star, u_star, dir);
    else Us = exact_sample_right(VR, p_star, u_star, dir);
    const Prim<T> Vs = cons_to_prim(Us);
    if (!physical_riemann(Vs)) return hll_flux(UL, UR, dir);
    return physical_flux(Us, dir);
}

template <class T>
inline Cons<T> riemann_flux(const Cons<T>& UL, const Cons<T>& UR, int dir, int solver) {  // :650-672
    switch (solver) {
        case HLL: return hll_flux(UL, UR, dir);
        case HLLC: return hllc_flux(UL, UR, dir);
        case EXACT: return exact_flux(UL, UR, dir);
        case FORCE: return force_flux(UL, UR, dir);
        default: return hll_flux(UL, UR, dir);
    }
}

// ---------------------------------------------------------------------------
// MUSCL-Hancock reconstruction (src/cpu/solver_cpu.cpp:22-99)
// ---------------------------------------------------------------------------
template <class T> inline T minmod(T a, T b) {                              // :22-28
    if (a * b <= 0.0) return T(0.0);
    return (a > 0.0) ? smin(a, b) : smax(a, b);
}

template <class T> inline bool is_physical(const Prim<T>& V) {              // :39-43
    return fin(V.rho) && fin(V.u) && fin(V.v) && fin(V.p) && (V.rho > kRhoFloor) && (V.p > kPFloor);
}

template <class T>
inline void reconstruct(const Cons<T>& Um, const Cons<T>& Uc, const Cons<T>& Up, T dt_over_d, int dir,
                        Cons<T>& UL_star, Cons<T>& UR_star) {                // :68-99
    const Prim<T> Wm = cons_to_prim(Um), Wc = cons_to_prim(Uc), Wp = cons_to_prim(Up);
    const Prim<T> a = Wc - Wm, b = Wp - Wc;
    const Prim<T> slope{minmod(a.rho, b.rho), minmod(a.u, b.u), minmod(a.v, b.v), minmod(a.p, b.p)};
    Prim<T> WL = Wc - 0.5 * slope;
    Prim<T> WR = Wc + 0.5 * slope;
    if (!is_physical(WL)) WL = Wc;
    if (!is_physical(WR)) WR = Wc;
    const Cons<T> UL = prim_to_cons(WL), UR = prim_to_cons(WR);
    const Cons<T> FL = physical_flux(UL, dir), FR = physical_flux(UR, dir);
    const Cons<T> half = 0.5 * dt_over_d * (FR - FL);
    const Cons<T> cL = UL - half, cR = UR - half;
    UL_star = is_physical(cons_to_prim(cL)) ? cL : UL;
    UR_star = is_physical(cons_to_prim(cR)) ? cR : UR;
}

// ---------------------------------------------------------------------------
// Grid: AoS with ng ghost layers, flat index j*(nx+2ng)+i (grid_cpu.hpp:86-88)
// ---------------------------------------------------------------------------
template <class T> struct Grid {
    int nx = 0, ny = 0, ng = 2;
    double dx = 0.0, dy = 0.0;
    Cons<T>* U = nullptr;  // (nx+2ng)*(ny+2ng) cells, caller-owned
    int tnx() const { return nx + 2 * ng; }
    Cons<T>& operator()(int i, int j) { return U[(std::size_t)j * tnx() + i]; }
    const Cons<T>& operator()(int i, int j) const { return U[(std::size_t)j * tnx() + i]; }
};

// Boundary kinds per edge.  TRANSMISSIVE follows head/cpu/boundary_cpu.hpp:5-36;
// INFLOW (Dirichlet) holds a fixed state in the ghost layers (new, used by
// the shock-bubble configurations with a supersonic inflow on the left).
enum Bc { TRANSMISSIVE = 0, INFLOW = 1 };
struct BcSpec {
    int kind[4] = {TRANSMISSIVE, TRANSMISSIVE, TRANSMISSIVE, TRANSMISSIVE};  // left, right, bottom, top
    double state[4][4] = {};                                                  // conserved inflow state per edge
};

template <class T> inline void apply_boundary(Grid<T>& g, const BcSpec& bc) {
    const int ng = g.ng, ib = ng, ie = ng + g.nx, jb = ng, je = ng + g.ny;
    auto inflow = [&](int e) { return Cons<T>{T(bc.state[e][0]), T(bc.state[e][1]), T(bc.state[e][2]), T(bc.state[e][3])}; };
    for (int j = jb; j < je; ++j) {
        for (int gg = 0; gg < ng; ++gg) g(ib - 1 - gg, j) = bc.kind[0] == INFLOW ? inflow(0) : g(ib, j);
        for (int gg = 0; gg < ng; ++gg) g(ie + gg, j) = bc.kind[1] == INFLOW ? inflow(1) : g(ie - 1, j);
    }
    for (int i = 0; i < g.tnx(); ++i) {
        for (int gg = 0; gg < ng; ++gg) g(i, jb - 1 - gg) = bc.kind[2] == INFLOW ? inflow(2) : g(i, jb);
        for (int gg = 0; gg < ng; ++gg) g(i, je + gg) = bc.kind[3] == INFLOW ? inflow(3) : g(i, je - 1);
    }
}

// Max signal speed over interior cells (src/cpu/solver_cpu.cpp:287-302).
template <class T> inline double max_speed(const Grid<T>& g) {
    double m = 0.0;
#pragma omp parallel for collapse(2) reduction(max : m) schedule(static)
    for (int j = g.ng; j < g.ng + g.ny; ++j)
        for (int i = g.ng; i < g.ng + g.nx; ++i) {
            const Prim<T> V = cons_to_prim(g(i, j));
            const double a = val(sound_speed(V));
            const double sx = std::abs(val(V.u)) + a;
            const double sy = std::abs(val(V.v)) + a;
            m = std::max(m, std::max(sx, sy));
        }
    return m;
}

// One x sweep Uin -> Uout over interior cells (src/cpu/solver_cpu.cpp:206-238, :437-455).
// Each cell is reconstructed once; the reference reconstructs it once per
// adjacent face, which yields the same values.
template <class T>
inline void sweep(const Grid<T>& Uin, Grid<T>& Uout, T dt, int dir, int solver, std::vector<Cons<T>>& ws) {
    const int ng = Uin.ng, nx = Uin.nx, ny = Uin.ny;
    const T dtd = dt / (dir == X ? Uin.dx : Uin.dy);
    if (dir == X) {
        const int nr = nx + 2;     // reconstructed cells i = ng-1 .. ng+nx
        ws.resize((std::size_t)3 * nr * ny);
#pragma omp parallel for schedule(static)
        for (int jj = 0; jj < ny; ++jj) {
            const int j = ng + jj;
            Cons<T>* L = &ws[(std::size_t)jj * 3 * nr];
            Cons<T>* R = L + nr;
            Cons<T>* F = R + nr;
            for (int k = 0; k < nr; ++k) {
                const int i = ng - 1 + k;
                reconstruct(Uin(i - 1, j), Uin(i, j), Uin(i + 1, j), dtd, X, L[k], R[k]);
            }
            for (int k = 0; k + 1 < nr; ++k) F[k] = riemann_flux(R[k], L[k + 1], X, solver);  // face between ng-1+k and ng+k
            for (int ii = 0; ii < nx; ++ii) Uout(ng + ii, j) = Uin(ng + ii, j) - dtd * (F[ii + 1] - F[ii]);
        }
    } else {
        const int nr = ny + 2;
        ws.resize((std::size_t)3 * nr * nx);
#pragma omp parallel for schedule(static)
        for (int ii = 0; ii < nx; ++ii) {
            const int i = ng + ii;
            Cons<T>* L = &ws[(std::size_t)ii * 3 * nr];
            Cons<T>* R = L + nr;
            Cons<T>* F = R + nr;
            for (int k = 0; k < nr; ++k) {
                const int j = ng - 1 + k;
                reconstruct(Uin(i, j - 1), Uin(i, j), Uin(i, j + 1), dtd, Y, L[k], R[k]);
            }
            for (int k = 0; k + 1 < nr; ++k) F[k] = riemann_flux(R[k], L[k + 1], Y, solver);
            for (int jj = 0; jj < ny; ++jj) Uout(i, ng + jj) = Uin(i, ng + jj) - dtd * (F[jj + 1] - F[jj]);
        }
    }
}

// Second-order split step (src/cpu/solver_cpu.cpp:400-499): x sweep Uold->Utmp,
// BC(Utmp), y sweep Utmp->Unew, BC(Unew).
template <class T>
inline void advance_second_order(const Grid<T>& Uold, Grid<T>& Utmp, Grid<T>& Unew, T dt, int solver,
                                 const BcSpec& bc, std::vector<Cons<T>>& ws) {
    sweep(Uold, Utmp, dt, X, solver, ws);
    apply_boundary(Utmp, bc);
    sweep(Utmp, Unew, dt, Y, solver, ws);
    apply_boundary(Unew, bc);
}

// First-order unsplit Godunov step (src/cpu/solver_cpu.cpp:315-369).
template <class T>
inline void advance_first_order(const Grid<T>& Uold, Grid<T>& Unew, T dt, int solver, const BcSpec& bc) {
    const int ng = Uold.ng;
    const T dtdx = dt / Uold.dx, dtdy = dt / Uold.dy;
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = ng; j < ng + Uold.ny; ++j)
        for (int i = ng; i < ng + Uold.nx; ++i) {
            const Cons<T> Fxp = riemann_flux(Uold(i, j), Uold(i + 1, j), X, solver);
            const Cons<T> Fxm = riemann_flux(Uold(i - 1, j), Uold(i, j), X, solver);
            const Cons<T> Fyp = riemann_flux(Uold(i, j), Uold(i, j + 1), Y, solver);
            const Cons<T> Fym = riemann_flux(Uold(i, j - 1), Uold(i, j), Y, solver);
            Unew(i, j) = Uold(i, j) - dtdx * (Fxp - Fxm) - dtdy * (Fyp - Fym);
        }
    apply_boundary(Unew, bc);
}

}  // namespace orc
