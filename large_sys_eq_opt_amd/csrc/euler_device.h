// euler_device.h -- FP64 physics, Riemann solvers and MUSCL-Hancock
// reconstruction as gfx950 device code.
//
// Arithmetic follows the reference operation by operation (head/physics.hpp,
// head/riemann.hpp, src/cpu/solver_cpu.cpp:22-99) and the library is built
// with -ffp-contract=off, so HLL, HLLC and FORCE reproduce the reference CPU
// solver bit for bit (the exact solver calls pow, whose last bit may differ
// between ROCm's device library and glibc).
//
// Two structural differences from the reference, neither of which changes a
// bit of the result:
//  * A primitive state that the reference recomputes from the same conserved
//    state (physical_flux after cons_to_prim, the Riemann solver's
//    cons_to_prim of a reconstructed state whose primitive the physicality
//    check already produced) is computed once and passed along.
//  * The solver and the sweep direction are template parameters, not
//    run-time switches (head/riemann.hpp:650-672 dispatches at run time).
#pragma once

#include <hip/hip_runtime.h>

namespace mh {

constexpr double kGamma = 1.4;          // head/physics.hpp:11
constexpr double kGm1 = kGamma - 1.0;
constexpr double kFloor = 1.0e-12;      // src/cpu/solver_cpu.cpp:19-20

enum : int { kHLL = 0, kHLLC = 1, kExact = 2, kForce = 3 };
// Template argument only: the exact solver with the full solve on fast tiles
// too (exact() with the fast tile's conversions and reconstruction), for
// states where few wavefronts are uniform.
constexpr int kExactFull = 4;
enum : int { kX = 0, kY = 1 };

struct Cons { double r, mu, mv, e; };   // rho, rho*u, rho*v, E
struct Prim { double r, u, v, p; };     // rho, u, v, p

__device__ __forceinline__ bool finite(double x) { return __builtin_isfinite(x); }

// Correctly rounded division and square root without the range scaling of
// the compiler's expansions.  a / b compiles to v_div_scale (numerator and
// denominator), v_rcp, two Newton steps, the quotient and its residual,
// v_div_fmas and v_div_fixup; v_div_scale changes nothing unless an operand
// is zero, denormal or near overflow, the exponents differ by 768 or more,
// or the quotient is denormal, and v_div_fixup (kept here) settles zero,
// infinite and NaN operands from the original operands.  sqrt(x) likewise
// rescales only x < 2^-767 and special values.  So div_nr / sqrt_nr return
// the same bits as a / b and sqrt(x) whenever operands and result lie well
// inside the normal range; the sweeps keep their results (F = true) only
// for tiles whose states lie in the fast range (range_key), where every
// operand does.
// div_nr_nf is div_nr without v_div_fixup, which changes nothing unless an
// operand is zero, infinite, NaN or denormal.  It serves only quotients
// whose operands are normal and nonzero wherever the result is kept: 1/rho
// of a converted state, p/(g-1) of an edge state that passed the floor
// check, g p/rho and E/rho of states that passed the physicality check,
// p/(rho (S - u)) of the HLLC star state (a zero denominator there has
// returned before).  Quotients whose numerator may be +-0 (HLL, S*, rho*)
// keep the fixup, which gives the zero its IEEE sign.
// rcp_nr(b) is the refined reciprocal div_nr_nf forms before its quotient
// steps, and div_rr(a, b, rcp_nr(b)) those steps: together exactly
// div_nr_nf(a, b).  Quotients that share a denominator (the conversion's
// 1/rho, the sound speed's gamma p/rho and HLLC's E/rho of one state) can
// so share the five instructions of the reciprocal.
__device__ __forceinline__ double rcp_nr(double b) {
    double r = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-b, r, 1.0);
    return __builtin_fma(r, e, r);
}
__device__ __forceinline__ double div_rr(double a, double b, double r) {
    const double q = a * r;
    const double t = __builtin_fma(-b, q, a);
    return __builtin_fma(t, r, q);
}
__device__ __forceinline__ double div_nr_nf(double a, double b) { return div_rr(a, b, rcp_nr(b)); }
__device__ __forceinline__ double div_nr(double a, double b) {
    return __builtin_amdgcn_div_fixup(div_nr_nf(a, b), b, a);
}
__device__ __forceinline__ double sqrt_nr(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}
// dv: a quotient whose numerator may be zero or arbitrarily small -- the
// four HLL numerators and HLLC's S* numerator, which are differences of
// momentum and energy fluxes: a velocity component in a quiescent region or
// on a symmetry line can be tiny (a momentum of 1e-307 is a normal double).
// v_div_scale rescales a nonzero numerator below 2^-969 (and a denormal
// quotient), so there the unscaled sequence can round differently.  A
// nonzero numerator below 2^-700 takes the IEEE division instead (a lane
// branch that is skipped when no lane needs it); zero keeps the unscaled
// sequence, whose v_div_fixup gives the quotient its IEEE sign.  From
// 2^-700 on, with the denominators these quotients have (at least 1e-14,
// at most 2^263 in magnitude), the quotient is normal and no scaling
// applies.
template <bool F> __device__ __forceinline__ double dv(double a, double b) {
    if constexpr (F) {
        if (fabs(a) < 0x1p-700 && a != 0.0) return a / b;
        return div_nr(a, b);
    }
    else return a / b;
}
// dvf: a quotient whose numerator is nonzero and bounded away from zero on
// every path that keeps it (HLLC's rho*: |rho (S - u)| >= rho a > 2^-115),
// keeping v_div_fixup.
template <bool F> __device__ __forceinline__ double dvf(double a, double b) {
    if constexpr (F) return div_nr(a, b);
    else return a / b;
}
template <bool F> __device__ __forceinline__ double dvn(double a, double b) {
    if constexpr (F) return div_nr_nf(a, b);
    else return a / b;
}
template <bool F> __device__ __forceinline__ double sq(double x) {
    if constexpr (F) return sqrt_nr(x);
    else return sqrt(x);
}

// libm fmin/fmax and std::min/max -- used exactly where the reference uses
// each.  Every fmin/fmax operand on this path is finite and non-zero ties
// have equal bits (the operands are sums with a positive sound speed, or the
// positive PVRS floor), so the hardware v_min/v_max_f64 gives the libm result.
__device__ __forceinline__ double lmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ double lmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

// head/physics.hpp:13-20
template <bool F = false>
__device__ __forceinline__ Prim c2p(const Cons& U) {
    const double inv_rho = dvn<F>(1.0, U.r);
    const double u = U.mu * inv_rho;
    const double v = U.mv * inv_rho;
    const double kinetic = 0.5 * U.r * (u * u + v * v);
    return {U.r, u, v, kGm1 * (U.e - kinetic)};
}
// The same; with F, rr receives rcp_nr(U.r) (the conversion's 1/rho is
// div_rr(1, rho, rr)).
template <bool F = false>
__device__ __forceinline__ Prim c2p_rr(const Cons& U, double& rr) {
    if constexpr (F) {
        rr = rcp_nr(U.r);
        const double inv_rho = div_rr(1.0, U.r, rr);
        const double u = U.mu * inv_rho;
        const double v = U.mv * inv_rho;
        const double kinetic = 0.5 * U.r * (u * u + v * v);
        return {U.r, u, v, kGm1 * (U.e - kinetic)};
    } else {
        rr = 0.0;
        return c2p<false>(U);
    }
}

// head/physics.hpp:22-28
template <bool F = false>
__device__ __forceinline__ Cons p2c(const Prim& V) {
    const double kinetic = 0.5 * V.r * (V.u * V.u + V.v * V.v);
    return {V.r, V.r * V.u, V.r * V.v, dvn<F>(V.p, kGm1) + kinetic};
}

template <bool F = false>
__device__ __forceinline__ double sound(const Prim& V) { return sq<F>(dvn<F>(kGamma * V.p, V.r)); }  // :39-41
// sound speed of V with R: from rr = rcp_nr(V.r) (fast tiles only)
template <bool F, bool R>
__device__ __forceinline__ double sound_r(const Prim& V, double rr) {
    if constexpr (F && R) return sqrt_nr(div_rr(kGamma * V.p, V.r, rr));
    else return sound<F>(V);
}

// Physical flux of U along D given V = c2p(U) (head/physics.hpp:56-74).
template <int D> __device__ __forceinline__ Cons pflux(const Cons& U, const Prim& V) {
    if (D == kX) return {U.mu, U.mu * V.u + V.p, U.mv * V.u, (U.e + V.p) * V.u};
    return {U.mv, U.mu * V.v, U.mv * V.v + V.p, (U.e + V.p) * V.v};
}

template <int D> __device__ __forceinline__ double un_of(const Prim& V) { return D == kX ? V.u : V.v; }
template <int D> __device__ __forceinline__ double ut_of(const Prim& V) { return D == kX ? V.v : V.u; }

// Physicality checks.  In a fast tile (F = true) every state that reaches
// one is finite wherever its density passes (the bounds at range_key), so
// the sign tests alone give the same answer.
template <bool F = false>
__device__ __forceinline__ bool phys_riemann(const Prim& V) {  // head/riemann.hpp:74-81
    if constexpr (F) return V.r > 0.0 && V.p > 0.0;
    return finite(V.r) && finite(V.u) && finite(V.v) && finite(V.p) && V.r > 0.0 && V.p > 0.0;
}

template <bool F = false>
__device__ __forceinline__ bool phys_recon(const Prim& V) {    // src/cpu/solver_cpu.cpp:39-43
    if constexpr (F) return (V.r > kFloor) && (V.p > kFloor);
    return finite(V.r) && finite(V.u) && finite(V.v) && finite(V.p) && (V.r > kFloor) && (V.p > kFloor);
}

__device__ __forceinline__ Cons avg(const Cons& a, const Cons& b) {  // 0.5 * (a + b)
    return {0.5 * (a.r + b.r), 0.5 * (a.mu + b.mu), 0.5 * (a.mv + b.mv), 0.5 * (a.e + b.e)};
}

// ---------------------------------------------------------------------------
// HLL (head/riemann.hpp:87-129); VL/VR = c2p(UL/UR).
// ---------------------------------------------------------------------------
// R (fast tiles): rrL, rrR are rcp_nr of the two densities.
template <int D, bool F = false, bool R = false>
__device__ __forceinline__ Cons hll(const Cons& UL, const Prim& VL, const Cons& UR, const Prim& VR, double rrL = 0.0,
                                    double rrR = 0.0) {
    const Cons FL = pflux<D>(UL, VL);
    const Cons FR = pflux<D>(UR, VR);
    if (!phys_riemann<F>(VL) || !phys_riemann<F>(VR)) return avg(FL, FR);
    const double aL = sound_r<F, R>(VL, rrL), aR = sound_r<F, R>(VR, rrR);
    const double unL = un_of<D>(VL), unR = un_of<D>(VR);
    const double SL = lmin(unL - aL, unR - aR);
    const double SR = lmax(unL + aL, unR + aR);
    if (SL >= 0.0) return FL;
    if (SR <= 0.0) return FR;
    const double denom = SR - SL;
    if (fabs(denom) < 1.0e-14) return avg(FL, FR);
    const double sls = SL * SR;
    return {dv<F>(SR * FL.r - SL * FR.r + sls * (UR.r - UL.r), denom),
            dv<F>(SR * FL.mu - SL * FR.mu + sls * (UR.mu - UL.mu), denom),
            dv<F>(SR * FL.mv - SL * FR.mv + sls * (UR.mv - UL.mv), denom),
            dv<F>(SR * FL.e - SL * FR.e + sls * (UR.e - UL.e), denom)};
}

// ---------------------------------------------------------------------------
// HLLC (head/riemann.hpp:135-271)
// ---------------------------------------------------------------------------
template <int D, bool F = false, bool R = false>
__device__ __forceinline__ Cons hllc_star(const Cons& U, const Prim& V, double S, double Sstar, double rr = 0.0) {
    const double rho = V.r, p = V.p;
    const double un = un_of<D>(V), ut = ut_of<D>(V);
    const double denom = S - Sstar;
    if (fabs(denom) < 1.0e-14) return U;
    const double rho_star = dvf<F>(rho * (S - un), denom);
    if ((!F && !finite(rho_star)) || rho_star <= 0.0) return U;   // finite in a fast tile
    double E_per_rho;
    if constexpr (F && R) E_per_rho = div_rr(U.e, rho, rr);
    else E_per_rho = dvn<F>(U.e, rho);
    const double E_star = rho_star * (E_per_rho + (Sstar - un) * (Sstar + dvn<F>(p, rho * (S - un))));
    if (!finite(E_star)) return U;
    if (D == kX) return {rho_star, rho_star * Sstar, rho_star * ut, E_star};
    return {rho_star, rho_star * ut, rho_star * Sstar, E_star};
}

template <int D, bool F = false, bool R = false>
__device__ __forceinline__ Cons hllc(const Cons& UL, const Prim& VL, const Cons& UR, const Prim& VR, double rrL = 0.0,
                                     double rrR = 0.0) {
    if (!phys_riemann<F>(VL) || !phys_riemann<F>(VR)) return hll<D, F, R>(UL, VL, UR, VR, rrL, rrR);
    const double aL = sound_r<F, R>(VL, rrL), aR = sound_r<F, R>(VR, rrR);
    const double unL = un_of<D>(VL), unR = un_of<D>(VR);
    const double pL = VL.p, pR = VR.p, rhoL = VL.r, rhoR = VR.r;
    const double SL = lmin(unL - aL, unR - aR);
    const double SR = lmax(unL + aL, unR + aR);
    const Cons FL = pflux<D>(UL, VL);
    const Cons FR = pflux<D>(UR, VR);
    if (SL >= 0.0) return FL;
    if (SR <= 0.0) return FR;
    const double numerator = pR - pL + rhoL * unL * (SL - unL) - rhoR * unR * (SR - unR);
    const double denominator = rhoL * (SL - unL) - rhoR * (SR - unR);
    if (fabs(denominator) < 1.0e-14) return hll<D, F, R>(UL, VL, UR, VR, rrL, rrR);
    const double Sstar = dv<F>(numerator, denominator);
    if (!F && !finite(Sstar)) return hll<D, F, R>(UL, VL, UR, VR, rrL, rrR);   // finite in a fast tile
    // One star state is formed, on the side the sign of S* picks
    // (neighbouring faces mostly agree, so the branch rarely diverges).
    if (Sstar >= 0.0) {
        const Cons Us = hllc_star<D, F, R>(UL, VL, SL, Sstar, rrL);
        return {FL.r + SL * (Us.r - UL.r), FL.mu + SL * (Us.mu - UL.mu), FL.mv + SL * (Us.mv - UL.mv),
                FL.e + SL * (Us.e - UL.e)};
    }
    const Cons Us = hllc_star<D, F, R>(UR, VR, SR, Sstar, rrR);
    return {FR.r + SR * (Us.r - UR.r), FR.mu + SR * (Us.mu - UR.mu), FR.mv + SR * (Us.mv - UR.mv),
            FR.e + SR * (Us.e - UR.e)};
}

// ---------------------------------------------------------------------------
// FORCE-type flux with local alpha (head/riemann.hpp:284-325)
// ---------------------------------------------------------------------------
template <int D, bool F = false, bool R = false>
__device__ __forceinline__ Cons force(const Cons& UL, const Prim& VL, const Cons& UR, const Prim& VR, double rrL = 0.0,
                                      double rrR = 0.0) {
    if (!phys_riemann<F>(VL) || !phys_riemann<F>(VR)) return hll<D, F, R>(UL, VL, UR, VR, rrL, rrR);
    const double aL = sound_r<F, R>(VL, rrL), aR = sound_r<F, R>(VR, rrR);
    const double unL = un_of<D>(VL), unR = un_of<D>(VR);
    const double alpha = lmax(fabs(unL) + aL, fabs(unR) + aR);
    const Cons FL = pflux<D>(UL, VL);
    const Cons FR = pflux<D>(UR, VR);
    if (!finite(alpha) || alpha <= 1.0e-14) return avg(FL, FR);
    const double ha = 0.5 * alpha;
    const Cons Flf = {0.5 * (FL.r + FR.r) - ha * (UR.r - UL.r), 0.5 * (FL.mu + FR.mu) - ha * (UR.mu - UL.mu),
                      0.5 * (FL.mv + FR.mv) - ha * (UR.mv - UL.mv), 0.5 * (FL.e + FR.e) - ha * (UR.e - UL.e)};
    const double hi = 0.5 * dvn<F>(1.0, alpha);   // alpha > 1e-14 and finite here
    const Cons Uri = {0.5 * (UL.r + UR.r) - hi * (FR.r - FL.r), 0.5 * (UL.mu + UR.mu) - hi * (FR.mu - FL.mu),
                      0.5 * (UL.mv + UR.mv) - hi * (FR.mv - FL.mv), 0.5 * (UL.e + UR.e) - hi * (FR.e - FL.e)};
    const Prim Vri = c2p<F>(Uri);
    if (!phys_riemann<F>(Vri)) return Flf;
    const Cons Fri = pflux<D>(Uri, Vri);
    return avg(Flf, Fri);
}

// ---------------------------------------------------------------------------
// Exact solver (head/riemann.hpp:331-644)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void exact_pfun(double p, double rhoK, double pK, double aK, double& f, double& fd) {
    constexpr double g = kGamma;
    if (p > pK) {
        const double A = 2.0 / ((g + 1.0) * rhoK);
        const double B = (g - 1.0) / (g + 1.0) * pK;
        const double q = sqrt(A / (p + B));
        f = (p - pK) * q;
        fd = q * (1.0 - 0.5 * (p - pK) / (p + B));
    } else {
        const double pratio = p / pK;
        constexpr double expo = (g - 1.0) / (2.0 * g);
        f = 2.0 * aK / (g - 1.0) * (pow(pratio, expo) - 1.0);
        fd = (1.0 / (rhoK * aK)) * pow(pratio, -(g + 1.0) / (2.0 * g));
    }
}

// Newton iteration for p*, u*; returns false where the reference does.
// `iters` receives the number of Newton iterations taken.
template <int D>
__device__ __forceinline__ bool exact_star(const Prim& VL, const Prim& VR, double aL, double aR, double& p_star,
                                           double& u_star, int& iters) {
    const double rhoL = VL.r, rhoR = VR.r, pL = VL.p, pR = VR.p;
    const double unL = un_of<D>(VL), unR = un_of<D>(VR);
    const double p_pvrs = 0.5 * (VL.p + VR.p) - 0.125 * (unR - unL) * (VL.r + VR.r) * (aL + aR);
    p_star = lmax(1.0e-12, p_pvrs);
    iters = 0;
    for (int iter = 0; iter < 20; ++iter) {
        double fL, fR, fdL, fdR;
        exact_pfun(p_star, rhoL, pL, aL, fL, fdL);
        exact_pfun(p_star, rhoR, pR, aR, fR, fdR);
        const double f = fL + fR + (unR - unL);
        const double fd = fdL + fdR;
        iters = iter + 1;
        if (fabs(fd) < 1.0e-14) return false;
        const double p_old = p_star;
        p_star = p_old - f / fd;
        if (!finite(p_star) || p_star <= 0.0) p_star = 0.5 * p_old;
        if (fabs(p_star - p_old) / (0.5 * (p_star + p_old) + 1.0e-14) < 1.0e-8) break;
    }
    if (!finite(p_star) || p_star <= 0.0) return false;
    double fL, fR, d0, d1;
    exact_pfun(p_star, rhoL, pL, aL, fL, d0);
    exact_pfun(p_star, rhoR, pR, aR, fR, d1);
    u_star = 0.5 * (unL + unR + fR - fL);
    return finite(u_star);
}

template <bool F = false>
__device__ __forceinline__ Cons cons_nt(double rho, double un, double ut, double p, int d) {  // riemann.hpp:40-64
    double u, v;
    if (d == kX) { u = un; v = ut; } else { u = ut; v = un; }
    return {rho, rho * u, rho * v, dvn<F>(p, kGm1) + 0.5 * rho * (u * u + v * v)};
}

// Sample the solution at xi = 0 on the left (side = 0) or right (side = 1)
// of the contact (head/riemann.hpp:445-591): density, normal velocity and
// pressure there (the tangential velocity is the side's own).
template <int D>
__device__ __forceinline__ void exact_sample_nt(const Prim& VK, double aK, double p_star, double u_star, int side,
                                                double& rho, double& un, double& p) {
    constexpr double g = kGamma;
    const double xi = 0.0;
    const double rhoK = VK.r, pK = VK.p, unK = un_of<D>(VK);
    rho = rhoK; un = unK; p = pK;
    if (side == 0) {
        if (p_star > pK) {
            const double p_ratio = p_star / pK;
            const double SL = unK - aK * sqrt((g + 1.0) / (2.0 * g) * p_ratio + (g - 1.0) / (2.0 * g));
            if (!(xi <= SL)) {
                rho = rhoK * ((p_ratio + (g - 1.0) / (g + 1.0)) / ((g - 1.0) / (g + 1.0) * p_ratio + 1.0));
                un = u_star; p = p_star;
            }
        } else {
            const double a_star = aK * pow(p_star / pK, (g - 1.0) / (2.0 * g));
            const double SHL = unK - aK;
            const double STL = u_star - a_star;
            if (xi <= SHL) {
            } else if (xi >= STL) {
                rho = rhoK * pow(p_star / pK, 1.0 / g);
                un = u_star; p = p_star;
            } else {
                const double a = 2.0 / (g + 1.0) * (aK + 0.5 * (g - 1.0) * (unK - xi));
                un = 2.0 / (g + 1.0) * (aK + 0.5 * (g - 1.0) * unK + xi);
                const double aratio = a / aK;
                rho = rhoK * pow(aratio, 2.0 / (g - 1.0));
                p = pK * pow(aratio, 2.0 * g / (g - 1.0));
            }
        }
    } else {
        if (p_star > pK) {
            const double p_ratio = p_star / pK;
            const double SR = unK + aK * sqrt((g + 1.0) / (2.0 * g) * p_ratio + (g - 1.0) / (2.0 * g));
            if (!(xi >= SR)) {
                rho = rhoK * ((p_ratio + (g - 1.0) / (g + 1.0)) / ((g - 1.0) / (g + 1.0) * p_ratio + 1.0));
                un = u_star; p = p_star;
            }
        } else {
            const double a_star = aK * pow(p_star / pK, (g - 1.0) / (2.0 * g));
            const double SHR = unK + aK;
            const double STR = u_star + a_star;
            if (xi >= SHR) {
            } else if (xi <= STR) {
                rho = rhoK * pow(p_star / pK, 1.0 / g);
                un = u_star; p = p_star;
            } else {
                const double a = 2.0 / (g + 1.0) * (aK - 0.5 * (g - 1.0) * (unK - xi));
                un = 2.0 / (g + 1.0) * (-aK + 0.5 * (g - 1.0) * unK + xi);
                const double aratio = a / aK;
                rho = rhoK * pow(aratio, 2.0 / (g - 1.0));
                p = pK * pow(aratio, 2.0 * g / (g - 1.0));
            }
        }
    }
}

template <int D>
__device__ __forceinline__ Cons exact_sample(const Prim& VK, double aK, double p_star, double u_star, int side) {
    double rho, un, p;
    exact_sample_nt<D>(VK, aK, p_star, u_star, side, rho, un, p);
    return cons_nt(rho, un, ut_of<D>(VK), p, D);
}

// The exact solver's result for a face whose two primitive states are equal
// (numerically; a uniform region), with 1e-12 <= p < 1e300 and rho, a below
// 1e300 (so the PVRS guess's sums stay finite): what exact()
// computes there, without the calls to pow, which all take a ratio of 1.
// The PVRS guess is p (0.5 (p + p) = p, the velocity difference is zero);
// the first Newton step evaluates both pressure functions at p = p_K
// (rarefaction branch, pow(1, e) = 1): f = +0, f' = 2 / (rho a), so
// p* = p - 0 / f' = p, the step is zero and the iteration stops;
// u* = 0.5 ((u_n + u_n + 0) - 0) (the sum as exact() forms it: +0 for a
// zero velocity); the sampled state on the side u* picks is (rho, u*, p)
// on every wave-speed branch (the head and tail speeds coincide, so no
// fan).  Then the same conserved state, conversion, check and flux.  With F
// (a fast tile) the same operations take the unscaled division, which gives
// the same bits there (the fast exact solver below); iters is left alone.
template <int D, bool F = false, bool R = false>
__device__ __forceinline__ Cons exact_uniform(const Cons& UL, const Prim& VL, const Cons& UR, const Prim& VR,
                                              double a, int& iters, double rrL = 0.0, double rrR = 0.0) {
    if constexpr (!F) iters = 1;
    const double fd1 = dvn<F>(1.0, VL.r * a) * 1.0;
    if (fabs(fd1 + fd1) < 1.0e-14) return hll<D, F, R>(UL, VL, UR, VR, rrL, rrR);
    const double unL = un_of<D>(VL), unR = un_of<D>(VR);
    const double u_star = 0.5 * (unL + unR + 0.0 - 0.0);
    if (!finite(u_star)) return hll<D, F, R>(UL, VL, UR, VR, rrL, rrR);
    const Prim& VK = u_star >= 0.0 ? VL : VR;
    const Cons Us = cons_nt<F>(VK.r, u_star, ut_of<D>(VK), VK.p, D);
    const Prim Vs = c2p<F>(Us);
    if (!phys_riemann<F>(Vs)) return hll<D, F, R>(UL, VL, UR, VR, rrL, rrR);
    return pflux<D>(Us, Vs);
}

template <int D>
__device__ __forceinline__ Cons exact(const Cons& UL, const Prim& VL, const Cons& UR, const Prim& VR, int& iters) {
    iters = 0;
    if (!phys_riemann(VL) || !phys_riemann(VR)) return hll<D>(UL, VL, UR, VR);
    const double aL = sound(VL), aR = sound(VR);
    // a wavefront whose faces all join equal states takes the closed form
    const bool uniform = VL.r == VR.r && VL.u == VR.u && VL.v == VR.v && VL.p == VR.p && VL.p >= 1.0e-12 &&
                         VL.p < 1.0e300 && VL.r < 1.0e300 && aL < 1.0e300;
    if (__all(uniform)) return exact_uniform<D>(UL, VL, UR, VR, aL, iters);
    double p_star = 0.0, u_star = 0.0;
    const bool ok = exact_star<D>(VL, VR, aL, aR, p_star, u_star, iters);
    if (!ok || !finite(p_star) || !finite(u_star) || p_star <= 0.0) return hll<D>(UL, VL, UR, VR);
    const bool left = u_star >= 0.0;
    const Cons Us = exact_sample<D>(left ? VL : VR, left ? aL : aR, p_star, u_star, left ? 0 : 1);
    const Prim Vs = c2p(Us);
    if (!phys_riemann(Vs)) return hll<D>(UL, VL, UR, VR);
    return pflux<D>(Us, Vs);
}

// The exact solver on a fast tile (F): exact()'s checks and, where every
// face of the wavefront joins equal states, its closed form with the
// unscaled division (the same bits); any other face is left to the tile's
// IEEE pass: `iters` takes kExactGaveUp, which fails the tile's verdict
// (solver_gave_up), and the returned flux is not used.  A fast tile's
// physicality checks, sound speed and closed form are those of exact() bit
// for bit, so a wavefront that passes here is the one exact() would send
// to exact_uniform.
constexpr int kExactGaveUp = 1 << 30;
template <int D, bool F, bool R>
__device__ __forceinline__ Cons exact_fast(const Cons& UL, const Prim& VL, const Cons& UR, const Prim& VR, int& iters,
                                           double rrL, double rrR) {
    if (!phys_riemann<F>(VL) || !phys_riemann<F>(VR)) return hll<D, F, R>(UL, VL, UR, VR, rrL, rrR);
    const double aL = sound_r<F, R>(VL, rrL);
    const bool uniform = VL.r == VR.r && VL.u == VR.u && VL.v == VR.v && VL.p == VR.p && VL.p >= 1.0e-12 &&
                         VL.p < 1.0e300 && VL.r < 1.0e300 && aL < 1.0e300;
    if (__all(uniform)) return exact_uniform<D, F, R>(UL, VL, UR, VR, aL, iters, rrL, rrR);
    iters |= kExactGaveUp;
    return UL;
}

// A fast tile's verdict on its Riemann solves: whether the fast exact
// solver gave up on a face (the full exact solver, which kExactFull runs on
// fast tiles, leaves its iteration count there, below that bit).
__device__ __forceinline__ bool solver_gave_up(int iters) { return (iters & kExactGaveUp) != 0; }

// F selects the unscaled division / square root (a fast tile); the exact
// solver then solves only uniform wavefronts (exact_fast) and leaves the
// rest, with the compiler's expansions and pow, to the IEEE pass.
template <int S, int D, bool F = false, bool R = false>
__device__ __forceinline__ Cons riemann(const Cons& UL, const Prim& VL, const Cons& UR, const Prim& VR, int& iters,
                                        double rrL = 0.0, double rrR = 0.0) {
    if (S == kHLL) return hll<D, F, R>(UL, VL, UR, VR, rrL, rrR);
    if (S == kHLLC) return hllc<D, F, R>(UL, VL, UR, VR, rrL, rrR);
    if (S == kForce) return force<D, F, R>(UL, VL, UR, VR, rrL, rrR);
    if constexpr (F && S == kExact) return exact_fast<D, F, R>(UL, VL, UR, VR, iters, rrL, rrR);
    return exact<D>(UL, VL, UR, VR, iters);
}

// ---------------------------------------------------------------------------
// MUSCL-Hancock reconstruction of one cell (src/cpu/solver_cpu.cpp:22-99)
// from the primitive states of the cell and its two neighbours along D.
// Outputs the evolved edge states and their primitive states.
// `half` = 0.5 * dt / d (the reference forms 0.5 * dt_over_d the same way).
// ---------------------------------------------------------------------------
// minmod with the reference's operation (a*b <= 0 -> 0; else std::min for
// a > 0, std::max for a < 0).  For a*b > 0 both branches pick the operand of
// smaller magnitude, a on ties, and a NaN operand propagates exactly as the
// reference's std::min/std::max do (a NaN a is kept; a NaN b loses), so the
// single magnitude compare is bit-identical to it.
__device__ __forceinline__ double minmod(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    return (fabs(b) < fabs(a)) ? b : a;
}

// minmod as the median of a, b and 0 (four min/max instead of a product,
// two compares and four selects).  For finite a, b whose product does not
// underflow it returns minmod's bits, except that a zero slope may come out
// as -0, which changes nothing where the slope is then added to or
// subtracted from a nonzero value.  So it serves rho and p of fast tiles:
// their differences are zero or at least 2^-72 in magnitude (in-range
// values) and their centre values are positive.
__device__ __forceinline__ double minmod_med(double a, double b) {
    return __builtin_fmax(__builtin_fmin(a, b), __builtin_fmin(__builtin_fmax(a, b), 0.0));
}

// An evolved edge state: conserved, primitive, and (fast tiles) rcp_nr of
// its density.
struct Edge { Cons U; Prim V; double rr; };

__device__ __forceinline__ Prim diff(const Prim& a, const Prim& b) { return {a.r - b.r, a.u - b.u, a.v - b.v, a.p - b.p}; }

// MUSCL-Hancock edge states of cell c (src/cpu/solver_cpu.cpp:58-99) from
// the backward difference dm = Wc - Wm; returns the forward difference
// Wp - Wc.  A sweep that marches cell by cell passes it on as the next
// cell's dm: the same subtraction of the same operands.
template <int D, bool F = false>
__device__ __forceinline__ Prim reconstruct_d(const Prim& dm, const Prim& Wc, const Prim& Wp, double half, Edge& L,
                                              Edge& R) {
    const Prim dp = diff(Wp, Wc);
    const Prim s = {F ? minmod_med(dm.r, dp.r) : minmod(dm.r, dp.r), minmod(dm.u, dp.u), minmod(dm.v, dp.v),
                    F ? minmod_med(dm.p, dp.p) : minmod(dm.p, dp.p)};
    const Prim hs = {0.5 * s.r, 0.5 * s.u, 0.5 * s.v, 0.5 * s.p};
    Prim WL = {Wc.r - hs.r, Wc.u - hs.u, Wc.v - hs.v, Wc.p - hs.p};
    Prim WR = {Wc.r + hs.r, Wc.u + hs.u, Wc.v + hs.v, Wc.p + hs.p};
    // The fallbacks (src/cpu/solver_cpu.cpp:86-87, :97-98) almost never fire:
    // when every lane of the wavefront passes a check the selects are skipped.
    // In a fast tile the first check always passes (the bounds at range_key),
    // so that pass leaves it out.
    if constexpr (!F) {
        const bool okL = phys_recon(WL), okR = phys_recon(WR);
        if (!__all(okL && okR)) {
            if (!okL) WL = Wc;
            if (!okR) WR = Wc;
        }
    }
    const Cons UL = p2c<F>(WL), UR = p2c<F>(WR);
    double rl, rr;
    const Prim VL = c2p_rr<F>(UL, rl), VR = c2p_rr<F>(UR, rr);
    const Cons FL = pflux<D>(UL, VL), FR = pflux<D>(UR, VR);
    const Cons h = {half * (FR.r - FL.r), half * (FR.mu - FL.mu), half * (FR.mv - FL.mv), half * (FR.e - FL.e)};
    L.U = {UL.r - h.r, UL.mu - h.mu, UL.mv - h.mv, UL.e - h.e};
    R.U = {UR.r - h.r, UR.mu - h.mu, UR.mv - h.mv, UR.e - h.e};
    L.V = c2p_rr<F>(L.U, L.rr);
    R.V = c2p_rr<F>(R.U, R.rr);
    const bool eL = phys_recon<F>(L.V), eR = phys_recon<F>(R.V);
    if (!__all(eL && eR)) {
        if (!eL) { L.U = UL; L.V = VL; L.rr = rl; }
        if (!eR) { R.U = UR; R.V = VR; R.rr = rr; }
    }
    return dp;
}

template <int D, bool F = false>
__device__ __forceinline__ void reconstruct(const Prim& Wm, const Prim& Wc, const Prim& Wp, double half, Edge& L,
                                            Edge& R) {
    reconstruct_d<D, F>(diff(Wc, Wm), Wc, Wp, half, L, R);
}

// Tiles whose states all lie in the fast range, swept at a dt/d in
// [2^-20, 2^20] (dt_in_fast_range), may use div_nr / sqrt_nr in the
// reconstruction, HLL/HLLC/FORCE and the signal speed: bit-identical to the
// compiler's division and square root there.  The fast range: rho, p and E
// in [2^-20, 2^20).  Then E >= (mu^2 + mv^2) / (2 rho) (p > 0) bounds
// |rho*u|, |rho*v| and |u|, |v| by 2^20.5, the reconstructed states stay
// within a factor 2 of these, their fluxes below 2^90, the evolved edge
// states below 2^110 (nonzero densities above 2^-73: a difference of
// in-range values is a multiple of their last place), and an evolved state
// that passes the physicality check (rho, p > 1e-12) has |u| <= 2^150.
// Every division whose numerator is a density, pressure, energy or 1 then
// has a nonzero numerator above 2^-300, numerator/denominator exponents
// less than 500 apart and a normal quotient, and every square root an
// argument above 2^-310 -- none of the cases in which v_div_scale or the
// square root's scaling changes anything (exponent gap >= 768, denormals,
// numerator below 2^-970, argument below 2^-767).  The range key does not
// bound momenta from below: a velocity can be arbitrarily small, and so can
// the HLL numerators and HLLC's S* numerator, which are differences of
// momentum and energy fluxes.  Those quotients go through dv, which hands a
// nonzero numerator below 2^-700 to the IEEE division.
// The same bounds make the finiteness tests of the physicality checks
// redundant there (phys_recon / phys_riemann with F = true): reconstructed
// states are finite; an evolved edge state is finite below 2^110 and, when
// its density passes (> 1e-12), so are its u, v (below 2^150) and p; a
// Riemann input is such a state or a reconstructed state converted twice
// (density > 1e-12 passed through unchanged); FORCE's intermediate density
// is a difference of a value above 2^-40 and a product, so it is zero or
// at least 2^-93 in magnitude (Sterbenz), and a positive one gives finite
// u, v, p.  A density that fails the sign test fails either form.
// The reconstruction check of WL and WR always passes there: rho and p of
// all three cells are positive and in range, and the limited slope is at
// most the centre's distance to a neighbour on its side, so each edge value
// is at least half the centre value (2^-21 > 1e-12).  In HLLC the Riemann
// inputs (rho < 2^110, |u| < 2^150, p < 2^110, a < 2^75) give |S*| < 2^461
// and |rho*| < 2^309 (both denominators are at least 1e-14 in magnitude), so
// those finiteness tests go too; E* keeps its test.
//
// The test works on the high words (sign, exponent, top of the mantissa):
// for x >= 0 the order of the high words is that of the values, so x lies
// in [2^-20, 2^20) exactly when hi(x) - hi(2^-20) < hi(2^20) - hi(2^-20) as
// unsigned integers; negative values, infinities and NaN fall outside.  A
// tile keeps the running maximum of range_key over the states it reads (an
// integer per lane rather than a lane mask: the y sweep has no scalar
// registers to spare) and compares it with kRangeKeyEnd once.
__device__ __forceinline__ unsigned hi_word(double x) { return (unsigned)(__double_as_longlong(x) >> 32); }
constexpr unsigned kRangeKeyBase = 0x3EB00000u;                   // hi_word(2^-20)
constexpr unsigned kRangeKeyEnd = 0x41300000u - kRangeKeyBase;    // hi_word(2^20) - base
__device__ __forceinline__ unsigned range_key(double rho, double e, double p) {
    const unsigned a = hi_word(rho) - kRangeKeyBase, b = hi_word(e) - kRangeKeyBase, c = hi_word(p) - kRangeKeyBase;
    return max(a, max(b, c));
}

// Cell update U - (dt/d) * (Fp - Fm)  (src/cpu/solver_cpu.cpp:452-453).
__device__ __forceinline__ Cons update(const Cons& U, double dtd, const Cons& Fm, const Cons& Fp) {
    return {U.r - dtd * (Fp.r - Fm.r), U.mu - dtd * (Fp.mu - Fm.mu), U.mv - dtd * (Fp.mv - Fm.mv),
            U.e - dtd * (Fp.e - Fm.e)};
}

// max(|u|+a, |v|+a) as src/cpu/solver_cpu.cpp:295-300 forms it.
template <bool F = false>
__device__ __forceinline__ double signal_speed(const Cons& U) {
    const Prim V = c2p<F>(U);
    double a;
    if constexpr (F) {
        // sqrt_nr(+-0) is NaN (rsq(0) = inf, times 0), sqrt(+-0) is +-0: a
        // state with p exactly 0 must still enter the maximum with |u|, |v|
        const double s = dvn<true>(kGamma * V.p, V.r);
        a = (s == 0.0) ? s : sqrt_nr(s);
    } else {
        a = sound<false>(V);
    }
    return smax(fabs(V.u) + a, fabs(V.v) + a);
}

}  // namespace mh
