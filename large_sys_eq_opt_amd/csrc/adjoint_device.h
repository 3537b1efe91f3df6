// adjoint_device.h -- hand-written reverse mode (vector-Jacobian products)
// of the device physics in euler_device.h: conversions, physical flux, sound
// speed, HLL, HLLC and FORCE with every branch, the exact solver (implicit
// derivative of p* and every sampling branch), and the MUSCL-Hancock
// reconstruction with its limiter and fallbacks.
//
// Conventions.  `b` prefixes an adjoint; every *_vjp accumulates (+=) into
// the adjoints of its inputs.  A primitive state passed next to its
// conserved state is always c2p of it, and its adjoint is returned
// separately: the caller folds it into the conserved adjoint with c2p_vjp.
// Branches are those the primal takes; a non-smooth point takes a fixed
// one-sided derivative: minmod picks the operand it returns (a on ties,
// zero when a*b <= 0), fmin/fmax the operand they return (the second on
// ties).  The time step is a constant (the adjoint of a split step at a
// given dt).
//
// F (fast tiles, see the range bounds in euler_device.h): the primal values
// recomputed here take the sweeps' fast division and square root (dv, dvf,
// dvn, sq with F), bit-identical to the IEEE ones in a fast tile, so every
// branch is the forward's; the derivative quotients take div_nr_nf / the
// refined reciprocal of the same primal denominators (normal and nonzero
// wherever the result is used), within an ulp of the IEEE quotient.
#pragma once

#include "euler_device.h"

namespace mh {

__device__ __forceinline__ void acc(Cons& a, double r, double mu, double mv, double e) {
    a.r += r; a.mu += mu; a.mv += mv; a.e += e;
}
__device__ __forceinline__ void acc(Prim& a, double r, double u, double v, double p) {
    a.r += r; a.u += u; a.v += v; a.p += p;
}
__device__ __forceinline__ Cons czero() { return {0.0, 0.0, 0.0, 0.0}; }
__device__ __forceinline__ Prim pzero() { return {0.0, 0.0, 0.0, 0.0}; }

// V = c2p(U): u = mu/r, v = mv/r, p = (g-1)(e - r(u^2+v^2)/2).
template <bool F = false>
__device__ __forceinline__ void c2p_vjp(const Cons& U, const Prim& bV, Cons& bU) {
    const double inv = dvn<F>(1.0, U.r);
    const double u = U.mu * inv, v = U.mv * inv;
    const double h = 0.5 * U.r, s = u * u + v * v;
    const double bkin = -kGm1 * bV.p;
    const double bs = h * bkin;
    const double bu = bV.u + 2.0 * u * bs, bv = bV.v + 2.0 * v * bs;
    const double binv = U.mu * bu + U.mv * bv;
    bU.r += bV.r + 0.5 * s * bkin - inv * inv * binv;
    bU.mu += inv * bu;
    bU.mv += inv * bv;
    bU.e += kGm1 * bV.p;
}

// U = p2c(V): (r, r u, r v, p/(g-1) + r(u^2+v^2)/2).
template <bool F = false>
__device__ __forceinline__ void p2c_vjp(const Prim& V, const Cons& bU, Prim& bV) {
    bV.r += bU.r + V.u * bU.mu + V.v * bU.mv + 0.5 * (V.u * V.u + V.v * V.v) * bU.e;
    bV.u += V.r * bU.mu + V.r * V.u * bU.e;
    bV.v += V.r * bU.mv + V.r * V.v * bU.e;
    bV.p += dvn<F>(bU.e, kGm1);
}

template <int D>
__device__ __forceinline__ void pflux_vjp(const Cons& U, const Prim& V, const Cons& bF, Cons& bU, Prim& bV) {
    if (D == kX) {   // (mu, mu u + p, mv u, (e+p) u)
        bU.mu += bF.r + V.u * bF.mu;
        bU.mv += V.u * bF.mv;
        bU.e += V.u * bF.e;
        bV.u += U.mu * bF.mu + U.mv * bF.mv + (U.e + V.p) * bF.e;
        bV.p += bF.mu + V.u * bF.e;
    } else {         // (mv, mu v, mv v + p, (e+p) v)
        bU.mv += bF.r + V.v * bF.mv;
        bU.mu += V.v * bF.mu;
        bU.e += V.v * bF.e;
        bV.v += U.mu * bF.mu + U.mv * bF.mv + (U.e + V.p) * bF.e;
        bV.p += bF.mv + V.v * bF.e;
    }
}

// a = sqrt(g p / r)
template <bool F = false>
__device__ __forceinline__ void sound_vjp(const Prim& V, double a, double ba, Prim& bV) {
    if constexpr (F) {
        const double rr = rcp_nr(V.r);
        const double s = div_rr(kGamma * V.p, V.r, rr);
        const double bs = dvn<true>(ba, 2.0 * a);
        bV.p += div_rr(kGamma, V.r, rr) * bs;
        bV.r -= div_rr(s, V.r, rr) * bs;
    } else {
        const double s = kGamma * V.p / V.r;
        const double bs = ba / (2.0 * a);
        bV.p += kGamma / V.r * bs;
        bV.r -= s / V.r * bs;
    }
}

template <int D> __device__ __forceinline__ void add_un(Prim& b, double x) { if (D == kX) b.u += x; else b.v += x; }
template <int D> __device__ __forceinline__ void add_ut(Prim& b, double x) { if (D == kX) b.v += x; else b.u += x; }

// S_L = fmin(unL - aL, unR - aR), S_R = fmax(unL + aL, unR + aR)
template <int D, bool F = false>
__device__ __forceinline__ void speeds_vjp(const Prim& VL, const Prim& VR, double aL, double aR, double bSL,
                                           double bSR, Prim& bVL, Prim& bVR) {
    const double unL = un_of<D>(VL), unR = un_of<D>(VR);
    double baL = 0.0, baR = 0.0;
    if ((unL - aL) < (unR - aR)) { add_un<D>(bVL, bSL); baL -= bSL; }
    else { add_un<D>(bVR, bSL); baR -= bSL; }
    if ((unL + aL) > (unR + aR)) { add_un<D>(bVL, bSR); baL += bSR; }
    else { add_un<D>(bVR, bSR); baR += bSR; }
    sound_vjp<F>(VL, aL, baL, bVL);
    sound_vjp<F>(VR, aR, baR, bVR);
}

// ---------------------------------------------------------------------------
// HLL (head/riemann.hpp:87-129)
// ---------------------------------------------------------------------------
template <int D, bool F = false>
__device__ void hll_vjp(const Cons& UL, const Prim& VL, const Cons& UR, const Prim& VR, const Cons& bF, Cons& bUL,
                        Prim& bVL, Cons& bUR, Prim& bVR) {
    const Cons FL = pflux<D>(UL, VL), FR = pflux<D>(UR, VR);
    if (!phys_riemann<F>(VL) || !phys_riemann<F>(VR)) {
        const Cons h = {0.5 * bF.r, 0.5 * bF.mu, 0.5 * bF.mv, 0.5 * bF.e};
        pflux_vjp<D>(UL, VL, h, bUL, bVL);
        pflux_vjp<D>(UR, VR, h, bUR, bVR);
        return;
    }
    const double aL = sound<F>(VL), aR = sound<F>(VR);
    const double unL = un_of<D>(VL), unR = un_of<D>(VR);
    const double SL = lmin(unL - aL, unR - aR);
    const double SR = lmax(unL + aL, unR + aR);
    if (SL >= 0.0) { pflux_vjp<D>(UL, VL, bF, bUL, bVL); return; }
    if (SR <= 0.0) { pflux_vjp<D>(UR, VR, bF, bUR, bVR); return; }
    const double denom = SR - SL;
    if (fabs(denom) < 1.0e-14) {
        const Cons h = {0.5 * bF.r, 0.5 * bF.mu, 0.5 * bF.mv, 0.5 * bF.e};
        pflux_vjp<D>(UL, VL, h, bUL, bVL);
        pflux_vjp<D>(UR, VR, h, bUR, bVR);
        return;
    }
    const double sls = SL * SR;
    const double rd = F ? rcp_nr(denom) : 0.0;   // denom >= 1e-14 in magnitude here
    auto qd = [&](double x) { return F ? div_rr(x, denom, rd) : x / denom; };
    const double fl[4] = {FL.r, FL.mu, FL.mv, FL.e}, fr[4] = {FR.r, FR.mu, FR.mv, FR.e};
    const double ul[4] = {UL.r, UL.mu, UL.mv, UL.e}, ur[4] = {UR.r, UR.mu, UR.mv, UR.e};
    const double bf[4] = {bF.r, bF.mu, bF.mv, bF.e};
    double bSL = 0.0, bSR = 0.0, bden = 0.0, bfl[4], bfr[4], bul[4], bur[4];
    for (int c = 0; c < 4; ++c) {
        const double N = SR * fl[c] - SL * fr[c] + sls * (ur[c] - ul[c]);
        const double bN = qd(bf[c]);
        bden -= qd(qd(N) * bf[c]);
        bSR += fl[c] * bN + SL * (ur[c] - ul[c]) * bN;
        bSL += -fr[c] * bN + SR * (ur[c] - ul[c]) * bN;
        bfl[c] = SR * bN;
        bfr[c] = -SL * bN;
        bur[c] = sls * bN;
        bul[c] = -sls * bN;
    }
    bSR += bden;
    bSL -= bden;
    acc(bUL, bul[0], bul[1], bul[2], bul[3]);
    acc(bUR, bur[0], bur[1], bur[2], bur[3]);
    pflux_vjp<D>(UL, VL, Cons{bfl[0], bfl[1], bfl[2], bfl[3]}, bUL, bVL);
    pflux_vjp<D>(UR, VR, Cons{bfr[0], bfr[1], bfr[2], bfr[3]}, bUR, bVR);
    speeds_vjp<D, F>(VL, VR, aL, aR, bSL, bSR, bVL, bVR);
}

// ---------------------------------------------------------------------------
// HLLC (head/riemann.hpp:135-271)
// ---------------------------------------------------------------------------
template <int D, bool F = false>
__device__ void hllc_vjp(const Cons& UL, const Prim& VL, const Cons& UR, const Prim& VR, const Cons& bF, Cons& bUL,
                         Prim& bVL, Cons& bUR, Prim& bVR) {
    if (!phys_riemann<F>(VL) || !phys_riemann<F>(VR)) { hll_vjp<D, F>(UL, VL, UR, VR, bF, bUL, bVL, bUR, bVR); return; }
    const double aL = sound<F>(VL), aR = sound<F>(VR);
    const double unL = un_of<D>(VL), unR = un_of<D>(VR);
    const double pL = VL.p, pR = VR.p, rhoL = VL.r, rhoR = VR.r;
    const double SL = lmin(unL - aL, unR - aR);
    const double SR = lmax(unL + aL, unR + aR);
    if (SL >= 0.0) { pflux_vjp<D>(UL, VL, bF, bUL, bVL); return; }
    if (SR <= 0.0) { pflux_vjp<D>(UR, VR, bF, bUR, bVR); return; }
    const double numerator = pR - pL + rhoL * unL * (SL - unL) - rhoR * unR * (SR - unR);
    const double denominator = rhoL * (SL - unL) - rhoR * (SR - unR);
    if (fabs(denominator) < 1.0e-14) { hll_vjp<D, F>(UL, VL, UR, VR, bF, bUL, bVL, bUR, bVR); return; }
    const double Ss = dv<F>(numerator, denominator);
    if (!finite(Ss)) { hll_vjp<D, F>(UL, VL, UR, VR, bF, bUL, bVL, bUR, bVR); return; }
    const bool left = Ss >= 0.0;
    const Cons& U = left ? UL : UR;
    const Prim& V = left ? VL : VR;
    const double S = left ? SL : SR;
    Cons& bU = left ? bUL : bUR;
    Prim& bV = left ? bVL : bVR;
    // star state (hllc_star) forward
    const double rho = V.r, p = V.p, un = un_of<D>(V), ut = ut_of<D>(V);
    const double dS = S - Ss;
    bool star = fabs(dS) >= 1.0e-14;
    double rs = 0.0, Epr = 0.0, Q = 0.0, P = 0.0, A = 0.0, Bq = 0.0, inner = 0.0, Es = 0.0;
    if (star) {
        Q = rho * (S - un);
        rs = dvf<F>(Q, dS);
        star = finite(rs) && rs > 0.0;
        if (star) {
            Epr = dvn<F>(U.e, rho);
            P = dvn<F>(p, Q);
            A = Ss - un;
            Bq = Ss + P;
            inner = Epr + A * Bq;
            Es = rs * inner;
            star = finite(Es);
        }
    }
    Cons Us = U;
    if (star) {
        Us.r = rs;
        Us.e = Es;
        if (D == kX) { Us.mu = rs * Ss; Us.mv = rs * ut; } else { Us.mu = rs * ut; Us.mv = rs * Ss; }
    }
    // F = F_K + S (U*_K - U_K)
    const Cons FK = pflux<D>(U, V);
    pflux_vjp<D>(U, V, bF, bU, bV);
    double bS = bF.r * (Us.r - U.r) + bF.mu * (Us.mu - U.mu) + bF.mv * (Us.mv - U.mv) + bF.e * (Us.e - U.e);
    (void)FK;
    acc(bU, -S * bF.r, -S * bF.mu, -S * bF.mv, -S * bF.e);
    const Cons bUs = {S * bF.r, S * bF.mu, S * bF.mv, S * bF.e};
    double bSs = 0.0;
    if (!star) {
        acc(bU, bUs.r, bUs.mu, bUs.mv, bUs.e);
    } else {
        const double bn = D == kX ? bUs.mu : bUs.mv;   // momentum along the face normal
        const double bt = D == kX ? bUs.mv : bUs.mu;
        double brs = bUs.r + Ss * bn + ut * bt;
        bSs += rs * bn;
        const double but = rs * bt;
        const double bEs = bUs.e;
        brs += inner * bEs;
        const double binner = rs * bEs;
        const double bEpr = binner, bA = Bq * binner, bBq = A * binner;
        double bun = 0.0, brho = 0.0, bp = 0.0, be = 0.0, bQ = 0.0;
        bSs += bA + bBq;
        bun -= bA;
        bp += dvn<F>(bBq, Q);
        bQ -= dvn<F>(P, Q) * bBq;
        be += dvn<F>(bEpr, rho);
        brho -= dvn<F>(Epr, rho) * bEpr;
        bQ += dvn<F>(brs, dS);
        const double bdS = dvn<F>(-rs, dS) * brs;
        brho += (S - un) * bQ;
        const double bD = rho * bQ;
        bS += bD;
        bun -= bD;
        bS += bdS;
        bSs -= bdS;
        bV.r += brho;
        bV.p += bp;
        add_un<D>(bV, bun);
        add_ut<D>(bV, but);
        bU.e += be;
    }
    // S* = num / den
    const double bnum = dvn<F>(bSs, denominator);
    const double bden = dvn<F>(-Ss, denominator) * bSs;
    double bSL = left ? bS : 0.0, bSR = left ? 0.0 : bS;
    bVR.p += bnum;
    bVL.p -= bnum;
    bVL.r += unL * (SL - unL) * bnum + (SL - unL) * bden;
    add_un<D>(bVL, rhoL * (SL - 2.0 * unL) * bnum - rhoL * bden);
    bSL += rhoL * unL * bnum + rhoL * bden;
    bVR.r += -unR * (SR - unR) * bnum - (SR - unR) * bden;
    add_un<D>(bVR, -rhoR * (SR - 2.0 * unR) * bnum + rhoR * bden);
    bSR += -rhoR * unR * bnum - rhoR * bden;
    speeds_vjp<D, F>(VL, VR, aL, aR, bSL, bSR, bVL, bVR);
}

// ---------------------------------------------------------------------------
// Exact solver (head/riemann.hpp:331-644)
// ---------------------------------------------------------------------------
// Partial derivatives of the pressure function f_K(p; rho_K, p_K, a_K) of
// exact_pfun (a_K treated as its own variable; the caller chains it through
// the sound speed).
__device__ __forceinline__ void exact_pfun_partials(double p, double rhoK, double pK, double aK, double& fp,
                                                    double& frho, double& fpK, double& faK) {
    constexpr double g = kGamma;
    if (p > pK) {   // shock: f = (p - pK) sqrt(A / (p + B)), A = 2/((g+1) rhoK), B = (g-1)/(g+1) pK
        const double A = 2.0 / ((g + 1.0) * rhoK);
        const double B = (g - 1.0) / (g + 1.0) * pK;
        const double q = sqrt(A / (p + B));
        const double dq = -0.5 * q / (p + B);   // dq/dp = dq/dB
        fp = q + (p - pK) * dq;
        fpK = -q + (p - pK) * dq * ((g - 1.0) / (g + 1.0));
        frho = (p - pK) * (-0.5 * q / rhoK);
        faK = 0.0;
    } else {        // rarefaction: f = 2 aK/(g-1) ((p/pK)^e - 1), e = (g-1)/(2g)
        constexpr double e = (g - 1.0) / (2.0 * g);
        const double pr = p / pK;
        const double dpr = 2.0 * aK / (g - 1.0) * e * pow(pr, e - 1.0);
        fp = dpr / pK;
        fpK = -dpr * p / (pK * pK);
        frho = 0.0;
        faK = 2.0 / (g - 1.0) * (pow(pr, e) - 1.0);
    }
}

// Adjoint of exact_sample_nt: (brho, bun, bp) of the sampled state to the
// side's primitive state (rho, un, p), its sound speed, p* and u*.  The
// branch is the one the primal takes (the wave-speed tests decide it).
template <int D>
__device__ void exact_sample_vjp(const Prim& VK, double aK, double ps, double us, int side, double brho, double bun,
                                 double bp, Prim& bVK, double& baK, double& bps, double& bus) {
    constexpr double g = kGamma;
    constexpr double g6 = (g - 1.0) / (g + 1.0);
    const double rhoK = VK.r, pK = VK.p, unK = un_of<D>(VK);
    double brhoK = 0.0, bunK = 0.0, bpK = 0.0;
    const double pr = ps / pK;
    auto ratio_back = [&](double bpr) {   // pr = ps / pK
        bps += bpr / pK;
        bpK -= bpr * ps / (pK * pK);
    };
    auto shock_star = [&]() {   // rho = rhoK (pr + g6) / (g6 pr + 1), un = u*, p = p*
        const double N = pr + g6, Dn = g6 * pr + 1.0;
        brhoK += brho * (N / Dn);
        ratio_back(brho * rhoK * (1.0 - g6 * g6) / (Dn * Dn));
        bus += bun;
        bps += bp;
    };
    auto rare_star = [&]() {    // rho = rhoK pr^(1/g), un = u*, p = p*
        brhoK += brho * pow(pr, 1.0 / g);
        ratio_back(brho * rhoK * (1.0 / g) * pow(pr, 1.0 / g - 1.0));
        bus += bun;
        bps += bp;
    };
    auto keep = [&]() { brhoK += brho; bunK += bun; bpK += bp; };
    // fan at xi = 0: a = 2/(g+1) (aK + s (g-1)/2 unK), un = 2/(g+1) (s aK + (g-1)/2 unK),
    // s = +1 on the left, -1 on the right; rho = rhoK (a/aK)^(2/(g-1)), p = pK (a/aK)^(2g/(g-1))
    auto fan = [&](double sgn) {
        const double a = 2.0 / (g + 1.0) * (aK + sgn * 0.5 * (g - 1.0) * unK);
        const double ar = a / aK;
        constexpr double m = 2.0 / (g - 1.0), n = 2.0 * g / (g - 1.0);
        brhoK += brho * pow(ar, m);
        bpK += bp * pow(ar, n);
        const double bar = brho * rhoK * m * pow(ar, m - 1.0) + bp * pK * n * pow(ar, n - 1.0);
        const double ba = bar / aK;
        baK += -bar * a / (aK * aK) + ba * (2.0 / (g + 1.0)) + bun * sgn * (2.0 / (g + 1.0));
        bunK += ba * sgn * ((g - 1.0) / (g + 1.0)) + bun * ((g - 1.0) / (g + 1.0));
    };
    if (side == 0) {
        if (ps > pK) {
            const double SL = unK - aK * sqrt((g + 1.0) / (2.0 * g) * pr + (g - 1.0) / (2.0 * g));
            if (!(0.0 <= SL)) shock_star(); else keep();
        } else {
            const double a_star = aK * pow(pr, (g - 1.0) / (2.0 * g));
            if (0.0 <= unK - aK) keep();
            else if (0.0 >= us - a_star) rare_star();
            else fan(1.0);
        }
    } else {
        if (ps > pK) {
            const double SR = unK + aK * sqrt((g + 1.0) / (2.0 * g) * pr + (g - 1.0) / (2.0 * g));
            if (!(0.0 >= SR)) shock_star(); else keep();
        } else {
            const double a_star = aK * pow(pr, (g - 1.0) / (2.0 * g));
            if (0.0 >= unK + aK) keep();
            else if (0.0 <= us + a_star) rare_star();
            else fan(-1.0);
        }
    }
    bVK.r += brhoK;
    bVK.p += bpK;
    add_un<D>(bVK, bunK);
}

// F = pflux(U*, V*) of the state sampled at xi = 0.  p* is the root of
// f(p) = f_L(p) + f_R(p) + unR - unL: its derivative is the implicit one,
// dp* = -(df/dq) / (df/dp) at the converged p* (the Newton iterations
// themselves are not differentiated), and u* = (unL + unR + f_R(p*) -
// f_L(p*)) / 2 is differentiated through p*.
template <int D>
__device__ void exact_vjp(const Cons& UL, const Prim& VL, const Cons& UR, const Prim& VR, const Cons& bF, Cons& bUL,
                          Prim& bVL, Cons& bUR, Prim& bVR) {
    if (!phys_riemann(VL) || !phys_riemann(VR)) { hll_vjp<D>(UL, VL, UR, VR, bF, bUL, bVL, bUR, bVR); return; }
    const double aL = sound(VL), aR = sound(VR);
    double ps = 0.0, us = 0.0;
    int iters = 0;
    const bool ok = exact_star<D>(VL, VR, aL, aR, ps, us, iters);
    if (!ok || !finite(ps) || !finite(us) || ps <= 0.0) {
        hll_vjp<D>(UL, VL, UR, VR, bF, bUL, bVL, bUR, bVR);
        return;
    }
    const bool left = us >= 0.0;
    const Prim& VK = left ? VL : VR;
    const double aK = left ? aL : aR;
    double rho, un, p;
    exact_sample_nt<D>(VK, aK, ps, us, left ? 0 : 1, rho, un, p);
    const double ut = ut_of<D>(VK);
    const Cons Us = cons_nt(rho, un, ut, p, D);
    const Prim Vs = c2p(Us);
    if (!phys_riemann(Vs)) { hll_vjp<D>(UL, VL, UR, VR, bF, bUL, bVL, bUR, bVR); return; }
    // back through the flux and the conversions of the sampled state
    Cons bUs = czero();
    Prim bVs = pzero();
    pflux_vjp<D>(Us, Vs, bF, bUs, bVs);
    c2p_vjp(Us, bVs, bUs);
    const Prim Ws = D == kX ? Prim{rho, un, ut, p} : Prim{rho, ut, un, p};   // cons_nt is p2c of it
    Prim bWs = pzero();
    p2c_vjp(Ws, bUs, bWs);
    const double bun = D == kX ? bWs.u : bWs.v, but = D == kX ? bWs.v : bWs.u;
    double baL = 0.0, baR = 0.0, bps = 0.0, bus = 0.0;
    Prim& bVK = left ? bVL : bVR;
    add_ut<D>(bVK, but);
    exact_sample_vjp<D>(VK, aK, ps, us, left ? 0 : 1, bWs.r, bun, bWs.p, bVK, left ? baL : baR, bps, bus);
    // u* = (unL + unR + f_R(p*) - f_L(p*)) / 2
    double fpL, frL, fpKL, faL, fpR, frR, fpKR, faR;
    exact_pfun_partials(ps, VL.r, VL.p, aL, fpL, frL, fpKL, faL);
    exact_pfun_partials(ps, VR.r, VR.p, aR, fpR, frR, fpKR, faR);
    add_un<D>(bVL, 0.5 * bus);
    add_un<D>(bVR, 0.5 * bus);
    const double bfR = 0.5 * bus, bfL = -0.5 * bus;
    bps += bfR * fpR + bfL * fpL;
    double brL = bfL * frL, bpL = bfL * fpKL, brR = bfR * frR, bpR = bfR * fpKR;
    baL += bfL * faL;
    baR += bfR * faR;
    // p*: f(p*; q) = 0
    const double lam = bps / (fpL + fpR);
    brL -= lam * frL; bpL -= lam * fpKL; baL -= lam * faL;
    brR -= lam * frR; bpR -= lam * fpKR; baR -= lam * faR;
    add_un<D>(bVL, lam);
    add_un<D>(bVR, -lam);
    bVL.r += brL; bVL.p += bpL;
    bVR.r += brR; bVR.p += bpR;
    sound_vjp(VL, aL, baL, bVL);
    sound_vjp(VR, aR, baR, bVR);
}

// ---------------------------------------------------------------------------
// FORCE with local alpha (head/riemann.hpp:284-325): F = (F_LF + F_Ri) / 2,
// F_LF = (FL + FR)/2 - (alpha/2)(UR - UL), U_Ri = (UL + UR)/2 - (1/(2 alpha))
// (FR - FL), alpha = fmax(|unL| + aL, |unR| + aR); |x| takes slope -1 only
// for x < 0.
// ---------------------------------------------------------------------------
template <int D, bool F = false>
__device__ void force_vjp(const Cons& UL, const Prim& VL, const Cons& UR, const Prim& VR, const Cons& bF, Cons& bUL,
                          Prim& bVL, Cons& bUR, Prim& bVR) {
    if (!phys_riemann<F>(VL) || !phys_riemann<F>(VR)) { hll_vjp<D, F>(UL, VL, UR, VR, bF, bUL, bVL, bUR, bVR); return; }
    const double aL = sound<F>(VL), aR = sound<F>(VR);
    const double unL = un_of<D>(VL), unR = un_of<D>(VR);
    const double alpha = lmax(fabs(unL) + aL, fabs(unR) + aR);
    const Cons FL = pflux<D>(UL, VL), FR = pflux<D>(UR, VR);
    if (!finite(alpha) || alpha <= 1.0e-14) {
        const Cons h = {0.5 * bF.r, 0.5 * bF.mu, 0.5 * bF.mv, 0.5 * bF.e};
        pflux_vjp<D>(UL, VL, h, bUL, bVL);
        pflux_vjp<D>(UR, VR, h, bUR, bVR);
        return;
    }
    const double ha = 0.5 * alpha, hi = 0.5 * dvn<F>(1.0, alpha);
    const double fl[4] = {FL.r, FL.mu, FL.mv, FL.e}, fr[4] = {FR.r, FR.mu, FR.mv, FR.e};
    const double ul[4] = {UL.r, UL.mu, UL.mv, UL.e}, ur[4] = {UR.r, UR.mu, UR.mv, UR.e};
    const double bf[4] = {bF.r, bF.mu, bF.mv, bF.e};
    const Cons Uri = {0.5 * (ul[0] + ur[0]) - hi * (fr[0] - fl[0]), 0.5 * (ul[1] + ur[1]) - hi * (fr[1] - fl[1]),
                      0.5 * (ul[2] + ur[2]) - hi * (fr[2] - fl[2]), 0.5 * (ul[3] + ur[3]) - hi * (fr[3] - fl[3])};
    const Prim Vri = c2p<F>(Uri);
    const bool ri = phys_riemann<F>(Vri);
    double blf[4], bri[4];
    for (int c = 0; c < 4; ++c) { blf[c] = ri ? 0.5 * bf[c] : bf[c]; bri[c] = ri ? 0.5 * bf[c] : 0.0; }
    double bfl[4] = {0.0, 0.0, 0.0, 0.0}, bfr[4] = {0.0, 0.0, 0.0, 0.0}, bul[4], bur[4];
    double balpha = 0.0;
    if (ri) {   // F_Ri = pflux(U_Ri, V_Ri)
        Cons bUri = czero();
        Prim bVri = pzero();
        pflux_vjp<D>(Uri, Vri, Cons{bri[0], bri[1], bri[2], bri[3]}, bUri, bVri);
        c2p_vjp<F>(Uri, bVri, bUri);
        const double bu[4] = {bUri.r, bUri.mu, bUri.mv, bUri.e};
        double bhi = 0.0;
        for (int c = 0; c < 4; ++c) {
            bhi -= bu[c] * (fr[c] - fl[c]);
            bfr[c] -= hi * bu[c];
            bfl[c] += hi * bu[c];
        }
        balpha += bhi * dvn<F>(-0.5, alpha * alpha);
        for (int c = 0; c < 4; ++c) { bul[c] = 0.5 * bu[c]; bur[c] = 0.5 * bu[c]; }
    } else {
        for (int c = 0; c < 4; ++c) { bul[c] = 0.0; bur[c] = 0.0; }
    }
    double bha = 0.0;   // F_LF
    for (int c = 0; c < 4; ++c) {
        bfl[c] += 0.5 * blf[c];
        bfr[c] += 0.5 * blf[c];
        bha -= blf[c] * (ur[c] - ul[c]);
        bur[c] -= ha * blf[c];
        bul[c] += ha * blf[c];
    }
    balpha += 0.5 * bha;
    acc(bUL, bul[0], bul[1], bul[2], bul[3]);
    acc(bUR, bur[0], bur[1], bur[2], bur[3]);
    pflux_vjp<D>(UL, VL, Cons{bfl[0], bfl[1], bfl[2], bfl[3]}, bUL, bVL);
    pflux_vjp<D>(UR, VR, Cons{bfr[0], bfr[1], bfr[2], bfr[3]}, bUR, bVR);
    double baL = 0.0, baR = 0.0;
    if ((fabs(unL) + aL) > (fabs(unR) + aR)) {
        add_un<D>(bVL, unL < 0.0 ? -balpha : balpha);
        baL += balpha;
    } else {
        add_un<D>(bVR, unR < 0.0 ? -balpha : balpha);
        baR += balpha;
    }
    sound_vjp<F>(VL, aL, baL, bVL);
    sound_vjp<F>(VR, aR, baR, bVR);
}

// F: HLL, HLLC and FORCE only (the exact solver keeps the IEEE forms).
template <int S, int D, bool F = false>
__device__ __forceinline__ void riemann_vjp(const Cons& UL, const Prim& VL, const Cons& UR, const Prim& VR,
                                            const Cons& bF, Cons& bUL, Prim& bVL, Cons& bUR, Prim& bVR) {
    if (S == kHLLC) hllc_vjp<D, F>(UL, VL, UR, VR, bF, bUL, bVL, bUR, bVR);
    else if (S == kExact) exact_vjp<D>(UL, VL, UR, VR, bF, bUL, bVL, bUR, bVR);
    else if (S == kForce) force_vjp<D, F>(UL, VL, UR, VR, bF, bUL, bVL, bUR, bVR);
    else hll_vjp<D, F>(UL, VL, UR, VR, bF, bUL, bVL, bUR, bVR);
}

// ---------------------------------------------------------------------------
// reconstruction (src/cpu/solver_cpu.cpp:22-99): adjoints of the evolved
// edge states (their conserved parts, primitive parts already folded in) to
// the primitive states of the three cells.
// ---------------------------------------------------------------------------
template <int D, bool F = false>
__device__ void reconstruct_vjp(const Prim& Wm, const Prim& Wc, const Prim& Wp, double half, const Cons& bLU,
                                const Cons& bRU, Prim& bWm, Prim& bWc, Prim& bWp) {
    const double am[4] = {Wc.r - Wm.r, Wc.u - Wm.u, Wc.v - Wm.v, Wc.p - Wm.p};
    const double bm[4] = {Wp.r - Wc.r, Wp.u - Wc.u, Wp.v - Wc.v, Wp.p - Wc.p};
    double sl[4];
    int pick[4];   // 0: zero slope, 1: a, 2: b
    for (int c = 0; c < 4; ++c) {
        if (am[c] * bm[c] <= 0.0) { sl[c] = 0.0; pick[c] = 0; }
        else if (fabs(bm[c]) < fabs(am[c])) { sl[c] = bm[c]; pick[c] = 2; }
        else { sl[c] = am[c]; pick[c] = 1; }
    }
    const Prim hs = {0.5 * sl[0], 0.5 * sl[1], 0.5 * sl[2], 0.5 * sl[3]};
    Prim WL = {Wc.r - hs.r, Wc.u - hs.u, Wc.v - hs.v, Wc.p - hs.p};
    Prim WR = {Wc.r + hs.r, Wc.u + hs.u, Wc.v + hs.v, Wc.p + hs.p};
    const bool okL = phys_recon<F>(WL), okR = phys_recon<F>(WR);
    if (!okL) WL = Wc;
    if (!okR) WR = Wc;
    const Cons UL = p2c<F>(WL), UR = p2c<F>(WR);
    const Prim VL = c2p<F>(UL), VR = c2p<F>(UR);
    const Cons FL = pflux<D>(UL, VL), FR = pflux<D>(UR, VR);
    const Cons h = {half * (FR.r - FL.r), half * (FR.mu - FL.mu), half * (FR.mv - FL.mv), half * (FR.e - FL.e)};
    const Cons cL = {UL.r - h.r, UL.mu - h.mu, UL.mv - h.mv, UL.e - h.e};
    const Cons cR = {UR.r - h.r, UR.mu - h.mu, UR.mv - h.mv, UR.e - h.e};
    const bool eL = phys_recon<F>(c2p<F>(cL)), eR = phys_recon<F>(c2p<F>(cR));
    // backward
    Cons bUL = czero(), bUR = czero(), bh = czero();
    acc(bUL, bLU.r, bLU.mu, bLU.mv, bLU.e);
    if (eL) acc(bh, -bLU.r, -bLU.mu, -bLU.mv, -bLU.e);
    acc(bUR, bRU.r, bRU.mu, bRU.mv, bRU.e);
    if (eR) acc(bh, -bRU.r, -bRU.mu, -bRU.mv, -bRU.e);
    const Cons bFR = {half * bh.r, half * bh.mu, half * bh.mv, half * bh.e};
    const Cons bFL = {-half * bh.r, -half * bh.mu, -half * bh.mv, -half * bh.e};
    Prim bVL = pzero(), bVR = pzero();
    pflux_vjp<D>(UL, VL, bFL, bUL, bVL);
    pflux_vjp<D>(UR, VR, bFR, bUR, bVR);
    c2p_vjp<F>(UL, bVL, bUL);
    c2p_vjp<F>(UR, bVR, bUR);
    Prim bWL = pzero(), bWR = pzero();
    p2c_vjp<F>(WL, bUL, bWL);
    p2c_vjp<F>(WR, bUR, bWR);
    acc(bWc, bWL.r, bWL.u, bWL.v, bWL.p);
    acc(bWc, bWR.r, bWR.u, bWR.v, bWR.p);
    double bhs[4] = {0.0, 0.0, 0.0, 0.0};
    if (okL) { bhs[0] -= bWL.r; bhs[1] -= bWL.u; bhs[2] -= bWL.v; bhs[3] -= bWL.p; }
    if (okR) { bhs[0] += bWR.r; bhs[1] += bWR.u; bhs[2] += bWR.v; bhs[3] += bWR.p; }
    double ba[4] = {0.0, 0.0, 0.0, 0.0}, bb[4] = {0.0, 0.0, 0.0, 0.0};
    for (int c = 0; c < 4; ++c) {
        const double bs = 0.5 * bhs[c];
        if (pick[c] == 1) ba[c] = bs;
        else if (pick[c] == 2) bb[c] = bs;
    }
    acc(bWc, ba[0] - bb[0], ba[1] - bb[1], ba[2] - bb[2], ba[3] - bb[3]);
    acc(bWm, -ba[0], -ba[1], -ba[2], -ba[3]);
    acc(bWp, bb[0], bb[1], bb[2], bb[3]);
}

}  // namespace mh
