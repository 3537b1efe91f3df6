// cases_oracle.hpp -- initial conditions for the oracle.  TEST INFRASTRUCTURE ONLY.
//
// Reference cases restated from src/test_cases.cpp:11-101 (shock_bubble_state,
// blast_wave_state, make_conserved) and src/init.cpp:8-38 (cell-centre point
// sampling of every cell, ghosts included).  The generalised blast and the
// nondimensional shock-bubble are the same formulas with the geometry and
// states as parameters; the smoothed-fraction bubble is new (it makes the
// objective differentiable in the bubble radius).
#pragma once

#include <cmath>

#include "euler_oracle.hpp"

namespace orc {

template <class T>
inline Cons<T> make_conserved(T rho, T u, T v, T p) {   // src/test_cases.cpp:11-18
    Cons<T> U{};
    U.rho = rho;
    U.rhou = rho * u;
    U.rhov = rho * v;
    U.E = p / (kGamma - 1.0) + 0.5 * rho * (u * u + v * v);
    return U;
}

// Rankine-Hugoniot post-shock state for a shock of Mach number `mach` running
// into gas at rest (rho0, p0): src/test_cases.cpp:35-47.
inline void post_shock(double rho0, double p0, double mach, double& rho2, double& u2, double& p2) {
    const double a0 = std::sqrt(kGamma * p0 / rho0);
    const double M2 = mach * mach;
    rho2 = rho0 * ((kGamma + 1.0) * M2) / ((kGamma - 1.0) * M2 + 2.0);
    p2 = p0 * (1.0 + 2.0 * kGamma / (kGamma + 1.0) * (M2 - 1.0));
    u2 = 2.0 * a0 / (kGamma + 1.0) * (mach - 1.0 / mach);
}

// Shock-bubble with the geometry as parameters (src/test_cases.cpp:24-69).
struct ShockBubbleSpec {
    double rho0, p0, mach, shock_x;      // quiescent gas and incident shock
    double bubble_rho, bubble_x, bubble_y, bubble_r;
};

inline Cons<double> shock_bubble_state(const ShockBubbleSpec& s, double x, double y) {
    double rho2, u2, p2;
    post_shock(s.rho0, s.p0, s.mach, rho2, u2, p2);
    double rho = s.rho0, u = 0.0, v = 0.0, p = s.p0;
    if (x < s.shock_x) { rho = rho2; u = u2; v = 0.0; p = p2; }
    const double dx = x - s.bubble_x;
    const double dy = y - s.bubble_y;
    const bool in_bubble = (dx * dx + dy * dy <= s.bubble_r * s.bubble_r);
    if (in_bubble) rho = s.bubble_rho;
    return make_conserved(rho, u, v, p);
}

// Reference shock_bubble_state: air/helium in SI units (src/test_cases.cpp:24-69).
inline ShockBubbleSpec reference_shock_bubble() {
    return {1.29, 1.01325e5, 1.22, 0.005, 0.214, 0.035, 0.0445, 0.025};
}

// Blast wave with the geometry as parameters (src/test_cases.cpp:78-101).
struct BlastSpec {
    double x0, y0, r0, rho_in, p_in, rho_out, p_out;
};

inline Cons<double> blast_state(const BlastSpec& b, double x, double y) {
    const double dx = x - b.x0;
    const double dy = y - b.y0;
    const double r2 = dx * dx + dy * dy;
    if (r2 <= b.r0 * b.r0) return make_conserved(b.rho_in, 0.0, 0.0, b.p_in);
    return make_conserved(b.rho_out, 0.0, 0.0, b.p_out);
}

inline BlastSpec reference_blast() { return {0.5, 0.5, 0.1, 1.0, 100.0, 1.0, 1.0}; }

// Cell centre (src/init.cpp:8-14).
inline double cell_center(double lo, int i, int ng, double d) {
    return lo + (static_cast<double>(i - ng) + 0.5) * d;
}

// Smoothed bubble fraction: a linear ramp of width h in the distance from
// the bubble centre, f = clamp((R - r)/h + 1/2, 0, 1).  It approximates the
// cell's area fraction inside the circle and is differentiable in R almost
// everywhere (df/dR = 1/h inside the ramp).
template <class T>
inline T bubble_fraction(double r, T R, double h) {
    const T f = (R - r) / h + 0.5;
    if (f <= 0.0) return T(0.0);
    if (f >= 1.0) return T(1.0);
    return f;
}

// Shock-bubble initial state with a smoothed bubble and a per-cell bubble
// density: rho = f*rho_b + (1-f)*rho_ambient; velocity and pressure are those
// of the ambient region (the reference's bubble changes density only).
template <class T>
inline Cons<T> shock_bubble_smooth_state(const ShockBubbleSpec& s, double x, double y, double h,
                                         T bubble_R, T rho_b) {
    double rho2, u2, p2;
    post_shock(s.rho0, s.p0, s.mach, rho2, u2, p2);
    double rho_a = s.rho0, u = 0.0, v = 0.0, p = s.p0;
    if (x < s.shock_x) { rho_a = rho2; u = u2; v = 0.0; p = p2; }
    const double dx = x - s.bubble_x;
    const double dy = y - s.bubble_y;
    const double r = std::sqrt(dx * dx + dy * dy);
    const T f = bubble_fraction(r, bubble_R, h);
    const T rho = f * rho_b + (1.0 - f) * rho_a;
    return make_conserved(rho, T(u), T(v), T(p));
}

}  // namespace orc
