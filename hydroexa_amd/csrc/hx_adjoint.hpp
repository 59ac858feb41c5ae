// hx_adjoint.hpp -- reverse-mode (adjoint) derivatives of one face of the
// plain Roe step, one face of the hydrostatic-reconstruction step and the
// per-cell Manning friction / dry clamp, written out by hand, FP64.
//
// Each *_vjp function re-evaluates its forward expression (the operation
// order of the forward kernels), keeps the intermediates in registers and
// then runs the reverse sweep: given the adjoints of the outputs it returns
// the vector-Jacobian product with respect to the inputs.  Branches are the
// ones the forward evaluation takes; at kinks the conventions are
// |x|' = +1 at x = 0 (|x| is x when x >= 0), and for max(a, b) = (a < b) ? b : a
// the derivative of the selected operand.  Terms that the forward guards
// with a wet test (h > 0) contribute nothing when the test fails, so no
// 0 * inf appears for dry states.
//
// Arithmetic: a derivative needs relative accuracy, not correct rounding,
// so every denominator is inverted once (hardware reciprocal + two Newton
// steps) and shared by the forward re-evaluation and the reverse sweep;
// square roots and their reciprocals come from one hardware rsq refined by
// two coupled Goldschmidt steps.  The branch tests of the positivity guard
// keep correctly rounded divisions so that the adjoint takes the branches
// the forward took.
//
// Forward expressions restated (reference file:line):
//   Roe face            include/solvers/kernels/Roe.H:50-309
//   homogeneous Roe     include/solvers/kernels/HydrostaticRoe.H:43-159
//   van Leer slope      src/solvers/HydrostaticRoe.cpp:173-195
//   hydrostatic face    src/solvers/HydrostaticRoe.cpp:218-396
//   Manning + dry clamp src/solvers/HydrostaticRoe.cpp:566-603
#pragma once

#include <hip/hip_runtime.h>

namespace hx {
namespace adj {

constexpr double kG = 9.81;        // include/utils/Constants.H:25
constexpr double kSqrtG = 3.132091952673165;
constexpr double kDry = 1e-6;      // include/utils/Constants.H:63
constexpr double kT12 = 1e-12;     // include/utils/Constants.H:70
constexpr double kT10 = 1e-10;     // include/utils/Constants.H:78

__device__ __forceinline__ double sgn0(double x) { return x >= 0.0 ? 1.0 : -1.0; }
__device__ __forceinline__ double fabs0(double x) { return x >= 0.0 ? x : -x; }

// 1/b for finite non-zero b: hardware estimate and two Newton steps.
__device__ __forceinline__ double rcp(double b)
{
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    e = fma(-b, r, 1.0);
    return fma(r, e, r);
}

// sqrt(x) and 1/sqrt(x) for x > 0: hardware rsq and two coupled Goldschmidt
// steps (g -> sqrt(x), h -> 1/(2 sqrt(x))).
__device__ __forceinline__ void sqrt_rsq(double x, double& sq, double& rs)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    double e = fma(-g, h, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    e = fma(-g, h, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    sq = g;
    rs = 2.0 * h;
}

// ---------------------------------------------------------------------------
// Shared Roe core: dry cut, Roe averages, entropy-fixed eigenvalues, wave
// strengths and the unrotated flux after the wave sum.
// ---------------------------------------------------------------------------
struct RoeTape {
    double hl, hr, sql, sqr, rsl, rsr, ul, vl, ur, vr, unl, utl, unr, utr;
    double ht, ct, rct, dsq, idsq, ut, vt, cgl, cgr;
    double el0, er0, el2, er2, lam0_0, lam2_0, iD0, iD2;
    double lam0, lam1, lam2, den, iden, dh, a0, a1, a2;
    double w0, w1, w2;
    double F0, F1, F2;        // unrotated, after the wave sum
    bool cutl, cutr, htact, fix0, fix2;
};

// Rotation to the frame of a face normal to direction DIR (0: x, 1: y) and
// its adjoint, with the products by the zero components of the normal left
// out: un = u nx + v ny, ut = -u ny + v nx.
template <int DIR>
struct Nrm {
    static __device__ __forceinline__ double un(double u, double v) { return DIR == 0 ? u : v; }
    static __device__ __forceinline__ double ut(double u, double v) { return DIR == 0 ? v : -u; }
    static __device__ __forceinline__ double ub(double unb, double utb)
    {
        return DIR == 0 ? unb : -utb;
    }
    static __device__ __forceinline__ double vb(double unb, double utb)
    {
        return DIR == 0 ? utb : unb;
    }
};

// UV = false: the inputs are the two states' depths and momenta; UV = true:
// depths and velocities (the hydrostatic face reconstructs velocities, and
// u = (h u) / h is u up to rounding, so the core takes them directly).
template <int DIR, bool UV>
__device__ __forceinline__ void roe_core_fwd(double hl, double hul, double hvl, double hr,
                                             double hur, double hvr, RoeTape& t)
{
    t.cutl = hl >= kDry;
    t.cutr = hr >= kDry;
    t.hl = hl = t.cutl ? hl : 0.0;
    t.hr = hr = t.cutr ? hr : 0.0;
    double s, r;
    sqrt_rsq(hl > 0.0 ? hl : 1.0, s, r);
    t.sql = hl > 0.0 ? s : 0.0;
    t.rsl = hl > 0.0 ? r : 0.0;
    sqrt_rsq(hr > 0.0 ? hr : 1.0, s, r);
    t.sqr = hr > 0.0 ? s : 0.0;
    t.rsr = hr > 0.0 ? r : 0.0;
    if (UV) {
        t.ul = hl > 0.0 ? hul : 0.0;
        t.vl = hl > 0.0 ? hvl : 0.0;
        t.ur = hr > 0.0 ? hur : 0.0;
        t.vr = hr > 0.0 ? hvr : 0.0;
    } else {
        const double ihl = t.rsl * t.rsl, ihr = t.rsr * t.rsr;
        t.ul = hul * ihl;
        t.vl = hvl * ihl;
        t.ur = hur * ihr;
        t.vr = hvr * ihr;
    }
    t.unl = Nrm<DIR>::un(t.ul, t.vl);
    t.utl = Nrm<DIR>::ut(t.ul, t.vl);
    t.unr = Nrm<DIR>::un(t.ur, t.vr);
    t.utr = Nrm<DIR>::ut(t.ur, t.vr);
    const double hmean = 0.5 * (hl + hr);
    t.htact = hmean > kDry;
    t.ht = t.htact ? hmean : kDry;
    sqrt_rsq(kG * t.ht, t.ct, t.rct);
    t.dsq = t.sql + t.sqr + kT12;
    t.idsq = rcp(t.dsq);
    t.ut = (t.unl * t.sql + t.unr * t.sqr) * t.idsq;
    t.vt = (t.utl * t.sql + t.utr * t.sqr) * t.idsq;
    t.lam0_0 = t.ut - t.ct;
    t.lam1 = t.ut;
    t.lam2_0 = t.ut + t.ct;
    t.cgl = kSqrtG * t.sql;
    t.cgr = kSqrtG * t.sqr;
    t.el0 = t.unl - t.cgl;
    t.er0 = t.unr - t.cgr;
    t.fix0 = (t.el0 < 0.0 && t.er0 > 0.0);
    t.iD0 = t.fix0 ? rcp(t.er0 - t.el0) : 0.0;
    t.lam0 = t.fix0 ? t.el0 * (t.er0 - t.lam0_0) * t.iD0 : t.lam0_0;
    t.el2 = t.unl + t.cgl;
    t.er2 = t.unr + t.cgr;
    t.fix2 = (t.el2 < 0.0 && t.er2 > 0.0);
    t.iD2 = t.fix2 ? rcp(t.er2 - t.el2) : 0.0;
    t.lam2 = t.fix2 ? t.er2 * (t.lam2_0 - t.el2) * t.iD2 : t.lam2_0;
    t.den = 2.0 * t.ct + kT12;
    t.iden = rcp(t.den);
    t.dh = hr - hl;
    const double dhu = hr * t.unr - hl * t.unl;
    const double dhv = hr * t.utr - hl * t.utl;
    t.a0 = (t.lam2 * t.dh - dhu) * t.iden;
    t.a1 = (-t.vt * t.dh + dhv) * t.rct;
    t.a2 = (-t.lam0 * t.dh + dhu) * t.iden;
    double F0 = 0.5 * (hl * t.unl + hr * t.unr);
    double F1 = 0.5 * (hl * t.unl * t.unl + 0.5 * kG * hl * hl + hr * t.unr * t.unr +
                       0.5 * kG * hr * hr);
    double F2 = 0.5 * (hl * t.unl * t.utl + hr * t.unr * t.utr);
    t.w0 = 0.5 * fabs0(t.lam0) * t.a0;
    t.w1 = 0.5 * fabs0(t.lam1) * t.a1;
    t.w2 = 0.5 * fabs0(t.lam2) * t.a2;
    F0 -= t.w0;
    F1 -= t.w0 * t.lam0;
    F2 -= t.w0 * t.vt;
    F2 -= t.w1 * t.ct;
    F0 -= t.w2;
    F1 -= t.w2 * t.lam2;
    F2 -= t.w2 * t.vt;
    t.F0 = F0;
    t.F1 = F1;
    t.F2 = F2;
}

// Adjoints of intermediates that callers (the Roe sources) add to before
// the core's reverse sweep.
struct RoeBar {
    double hl, hr, ut, vt, ct, ht, den, lam0, lam2, a0, a2;
};

// Reverse sweep of roe_core_fwd.  Fb: adjoint of the unrotated flux; A:
// extra adjoints of intermediates.  Out: adjoints of the six inputs.
template <int DIR, bool UV>
__device__ __forceinline__ void roe_core_rev(const RoeTape& t, double Fb0, double Fb1,
                                             double Fb2, RoeBar A, double* hlb_o, double* hulb,
                                             double* hvlb, double* hrb_o, double* hurb,
                                             double* hvrb)
{
    double hlb = A.hl, hrb = A.hr;
    double unlb = 0.0, utlb = 0.0, unrb = 0.0, utrb = 0.0;
    double lam0b = A.lam0, lam2b = A.lam2, utb = A.ut, vtb = A.vt, ctb = A.ct, denb = A.den;
    double a0b = A.a0, a2b = A.a2, a1b = 0.0, htb = A.ht;
    // wave sum
    {
        const double wb0 = -(Fb0 + Fb1 * t.lam0 + Fb2 * t.vt);
        const double wb1 = -(Fb2 * t.ct);
        const double wb2 = -(Fb0 + Fb1 * t.lam2 + Fb2 * t.vt);
        lam0b += -Fb1 * t.w0;
        lam2b += -Fb1 * t.w2;
        vtb += -Fb2 * (t.w0 + t.w2);
        ctb += -Fb2 * t.w1;
        a0b += wb0 * 0.5 * fabs0(t.lam0);
        a1b += wb1 * 0.5 * fabs0(t.lam1);
        a2b += wb2 * 0.5 * fabs0(t.lam2);
        lam0b += wb0 * 0.5 * t.a0 * sgn0(t.lam0);
        utb += wb1 * 0.5 * t.a1 * sgn0(t.lam1);
        lam2b += wb2 * 0.5 * t.a2 * sgn0(t.lam2);
    }
    // centred flux
    {
        const double hl = t.hl, hr = t.hr;
        hlb += 0.5 * (Fb0 * t.unl + Fb1 * (t.unl * t.unl + kG * hl) + Fb2 * t.unl * t.utl);
        hrb += 0.5 * (Fb0 * t.unr + Fb1 * (t.unr * t.unr + kG * hr) + Fb2 * t.unr * t.utr);
        unlb += 0.5 * (Fb0 * hl + Fb1 * 2.0 * hl * t.unl + Fb2 * hl * t.utl);
        unrb += 0.5 * (Fb0 * hr + Fb1 * 2.0 * hr * t.unr + Fb2 * hr * t.utr);
        utlb += 0.5 * Fb2 * hl * t.unl;
        utrb += 0.5 * Fb2 * hr * t.unr;
    }
    // wave strengths
    double dhb = 0.0, dhub = 0.0, dhvb = 0.0;
    {
        double nb = a0b * t.iden;
        denb += -nb * t.a0;
        lam2b += nb * t.dh;
        dhb += nb * t.lam2;
        dhub -= nb;
        nb = a1b * t.rct;
        ctb += -nb * t.a1;
        vtb += -nb * t.dh;
        dhb += -nb * t.vt;
        dhvb += nb;
        nb = a2b * t.iden;
        denb += -nb * t.a2;
        lam0b += -nb * t.dh;
        dhb += -nb * t.lam0;
        dhub += nb;
    }
    hrb += dhb;
    hlb -= dhb;
    hrb += dhub * t.unr;
    unrb += dhub * t.hr;
    hlb -= dhub * t.unl;
    unlb -= dhub * t.hl;
    hrb += dhvb * t.utr;
    utrb += dhvb * t.hr;
    hlb -= dhvb * t.utl;
    utlb -= dhvb * t.hl;
    ctb += 2.0 * denb;
    // entropy fix (lam2, then lam0)
    double cglb = 0.0, cgrb = 0.0, lam2_0b = lam2b, lam0_0b = lam0b;
    if (t.fix2) {
        const double M = t.lam2_0 - t.el2, L = lam2b * t.iD2;      // lam2 = er2 M / D
        const double q = L * t.er2 * M * t.iD2;
        const double er2b = L * M - q;
        const double el2b = -L * t.er2 + q;
        lam2_0b = L * t.er2;
        unlb += el2b;
        cglb += el2b;
        unrb += er2b;
        cgrb += er2b;
    }
    if (t.fix0) {
        const double M = t.er0 - t.lam0_0, L = lam0b * t.iD0;      // lam0 = el0 M / D
        const double q = L * t.el0 * M * t.iD0;
        const double el0b = L * M + q;
        const double er0b = L * t.el0 - q;
        lam0_0b = -L * t.el0;
        unlb += el0b;
        cglb -= el0b;
        unrb += er0b;
        cgrb -= er0b;
    }
    utb += lam0_0b + lam2_0b;
    ctb += lam2_0b - lam0_0b;
    // Roe averages
    double sqlb = 0.0, sqrb = 0.0;
    {
        double nb = utb * t.idsq;
        double dsqb = -nb * t.ut;
        unlb += nb * t.sql;
        sqlb += nb * t.unl;
        unrb += nb * t.sqr;
        sqrb += nb * t.unr;
        nb = vtb * t.idsq;
        dsqb += -nb * t.vt;
        utlb += nb * t.sql;
        sqlb += nb * t.utl;
        utrb += nb * t.sqr;
        sqrb += nb * t.utr;
        sqlb += dsqb;
        sqrb += dsqb;
    }
    htb += ctb * kG * 0.5 * t.rct;
    if (t.htact) {
        hlb += 0.5 * htb;
        hrb += 0.5 * htb;
    }
    // cg = sqrt(g) sqrt(h): folded into the sqrt(h) adjoint
    sqlb += kSqrtG * cglb;
    sqrb += kSqrtG * cgrb;
    // rotation to the face frame
    const double ulb = Nrm<DIR>::ub(unlb, utlb), vlb = Nrm<DIR>::vb(unlb, utlb);
    const double urb = Nrm<DIR>::ub(unrb, utrb), vrb = Nrm<DIR>::vb(unrb, utrb);
    double hulb_ = 0.0, hvlb_ = 0.0, hurb_ = 0.0, hvrb_ = 0.0;
    if (t.hl > 0.0) {
        if (UV) {
            hulb_ = ulb;
            hvlb_ = vlb;
        } else {
            const double ih = t.rsl * t.rsl;
            hulb_ = ulb * ih;
            hvlb_ = vlb * ih;
            hlb -= (ulb * t.ul + vlb * t.vl) * ih;
        }
        hlb += sqlb * 0.5 * t.rsl;
    }
    if (t.hr > 0.0) {
        if (UV) {
            hurb_ = urb;
            hvrb_ = vrb;
        } else {
            const double ih = t.rsr * t.rsr;
            hurb_ = urb * ih;
            hvrb_ = vrb * ih;
            hrb -= (urb * t.ur + vrb * t.vr) * ih;
        }
        hrb += sqrb * 0.5 * t.rsr;
    }
    *hlb_o = t.cutl ? hlb : 0.0;
    *hrb_o = t.cutr ? hrb : 0.0;
    *hulb = hulb_;
    *hvlb = hvlb_;
    *hurb = hurb_;
    *hvrb = hvrb_;
}

// ---------------------------------------------------------------------------
// Plain Roe face with split bed / friction sources (Roe.H:50-309).
// In: L, R = {h, hu, hv, z, n} (h rebuilt); adjoints of F, Sm, Sp (rotated,
// as the update uses them).  Out: Lb, Rb (5 each).
// ---------------------------------------------------------------------------
template <int DIR>
__device__ __forceinline__ void roe_face_vjp(const double* L, const double* R, double ds,
                                             const double* Fb, const double* Smb,
                                             const double* Spb, double* Lb, double* Rb)
{
    RoeTape t;
    roe_core_fwd<DIR, false>(L[0], L[1], L[2], R[0], R[1], R[2], t);
    const double hl = t.hl, hr = t.hr;
    // bed thrust and friction (Roe.H:151-189)
    const double dz = R[3] - L[3];
    const double etl = hl + L[3];
    const double etr = hr + R[3];
    const double dd = etr - etl;
    const double ta = -kG * t.ht * dz;
    int dzk;   // 0: dzt = hl, 1: dzt = hr, 2: dzt = dz
    double dzt;
    if ((dz >= 0.0) && (etl < R[3])) { dzk = 0; dzt = hl; }
    else if ((dz < 0.0) && (etr < L[3])) { dzk = 1; dzt = hr; }
    else { dzk = 2; dzt = dz; }
    const bool hupl = dz >= 0.0;
    const double hup = hupl ? hl : hr;
    const double tb = -kG * (hup - 0.5 * fabs0(dzt)) * dzt;
    const bool sel = (dz * dd >= 0.0) && (t.ut * dz > 0.0);
    const bool pick_b = !sel || (ta < tb);
    const double thrust = pick_b ? tb : ta;
    const double nt = 0.5 * (L[4] + R[4]);
    double mag, rmag;
    const double m2 = t.ut * t.ut + t.vt * t.vt;
    sqrt_rsq(m2 > 0.0 ? m2 : 1.0, mag, rmag);
    if (!(m2 > 0.0)) { mag = 0.0; rmag = 0.0; }
    const double cb = cbrt(t.ht);
    const double h43 = t.ht * cb;
    const bool dnact = kDry < h43;
    const double Dn = dnact ? h43 : kDry;
    const double iDn = rcp(Dn);
    const double sf = (nt * nt * mag * t.ut) * iDn;
    const double tau = kG * t.ht * sf * ds;
    const double B = (thrust - tau) * t.iden;
    // positivity guard (Roe.H:200-247): b0, b2 as signed picks of B, m1, m3;
    // the branch tests use correctly rounded divisions, as the forward does
    double c0[3] = {1.0, 0.0, 0.0}, c2[3] = {-1.0, 0.0, 0.0};
    double b0 = B, b2 = -B;
    {
        auto tol = [](double x) { return (fabs0(x) > kT12) ? x : (x >= 0.0 ? kT12 : -kT12); };
        const double l0s = tol(t.lam0), l2s = tol(t.lam2);
        const double h1s = hl + t.a0 - (b0 / l0s);
        const double h3s = hr - t.a2 + (b2 / l2s);
        const double m1 = -(hl + t.a0) * fabs0(t.lam0);
        const double m3 = -(hr - t.a2) * t.lam2;
        const bool d1 = (h1s < -kDry) && (h3s > kDry);
        const bool d3 = (h1s > kDry) && (h3s < -kDry);
        if (d1 || d3) {
            double mlam = fabs0(t.lam0);
            mlam = (mlam < fabs0(t.lam1)) ? fabs0(t.lam1) : mlam;
            mlam = (mlam < fabs0(t.lam2)) ? fabs0(t.lam2) : mlam;
            const double loc_dt = ds / (mlam + kT12);
            if (d1 && ((ds * 0.5 / l0s) * (hl / tol(hl - h1s)) < loc_dt)) {
                if (-m1 >= m3) { b0 = m1; c0[0] = 0.0; c0[1] = 1.0; c0[2] = 0.0; }
                b2 = -b0;
                for (int q = 0; q < 3; ++q) c2[q] = -c0[q];
            }
            if (d3 && ((ds * 0.5 / l2s) * (hr / tol(hr - h3s)) < loc_dt)) {
                if (-m3 >= m1) { b2 = m3; c2[0] = 0.0; c2[1] = 0.0; c2[2] = 1.0; }
                b0 = -b2;
                for (int q = 0; q < 3; ++q) c0[q] = -c2[q];
            }
        }
    }
    // ---- reverse ----
    auto unrot1 = [&](const double* X) { return Nrm<DIR>::un(X[1], X[2]); };
    auto unrot2 = [&](const double* X) { return Nrm<DIR>::ut(X[1], X[2]); };
    RoeBar A{};
    // sources: wave k goes to Sm when lam_k > 0, to Sp when lam_k < 0
    double betb0 = 0.0, betb2 = 0.0;
    {
        const double* S0 = (t.lam0 > 0.0) ? Smb : ((t.lam0 < 0.0) ? Spb : nullptr);
        if (S0) {
            const double s0 = S0[0], s1 = unrot1(S0), s2 = unrot2(S0);
            betb0 = s0 + s1 * t.lam0 + s2 * t.vt;
            A.lam0 += s1 * b0;
            A.vt += s2 * b0;
        }
        const double* S2 = (t.lam2 > 0.0) ? Smb : ((t.lam2 < 0.0) ? Spb : nullptr);
        if (S2) {
            const double s0 = S2[0], s1 = unrot1(S2), s2 = unrot2(S2);
            betb2 = s0 + s1 * t.lam2 + s2 * t.vt;
            A.lam2 += s1 * b2;
            A.vt += s2 * b2;
        }
    }
    const double Bb = betb0 * c0[0] + betb2 * c2[0];
    const double m1b = betb0 * c0[1] + betb2 * c2[1];
    const double m3b = betb0 * c0[2] + betb2 * c2[2];
    A.hl -= m1b * fabs0(t.lam0);
    A.a0 -= m1b * fabs0(t.lam0);
    A.lam0 -= m1b * (hl + t.a0) * sgn0(t.lam0);
    A.hr -= m3b * t.lam2;
    A.a2 += m3b * t.lam2;
    A.lam2 -= m3b * (hr - t.a2);
    const double thb = Bb * t.iden;
    const double taub = -thb;
    A.den -= thb * B;
    A.ht += taub * kG * sf * ds;
    const double sfb = taub * kG * t.ht * ds;
    const double numb = sfb * iDn;
    const double Dnb = -numb * sf;
    const double ntb = numb * 2.0 * nt * mag * t.ut;
    const double magb = numb * nt * nt * t.ut;
    A.ut += numb * nt * nt * mag;
    if (dnact) A.ht += Dnb * (cb + t.ht * rcp(3.0 * cb * cb));
    if (mag > 0.0) {
        A.ut += magb * t.ut * rmag;
        A.vt += magb * t.vt * rmag;
    }
    const double tab = pick_b ? 0.0 : thb;
    const double tbb = pick_b ? thb : 0.0;
    double dzb = 0.0;
    A.ht += -tab * kG * dz;
    dzb += -tab * kG * t.ht;
    const double hupb = -tbb * kG * dzt;
    const double dztb = -tbb * kG * (hup - fabs0(dzt));
    if (hupl) A.hl += hupb; else A.hr += hupb;
    if (dzk == 0) A.hl += dztb;
    else if (dzk == 1) A.hr += dztb;
    else dzb += dztb;
    Rb[3] = dzb;
    Lb[3] = -dzb;
    Lb[4] = 0.5 * ntb;
    Rb[4] = 0.5 * ntb;
    roe_core_rev<DIR, false>(t, Fb[0], unrot1(Fb), unrot2(Fb), A, &Lb[0], &Lb[1], &Lb[2],
                             &Rb[0], &Rb[1], &Rb[2]);
}

// ---------------------------------------------------------------------------
// van Leer slope (HydrostaticRoe.cpp:173-195): phi(r) dp with r = dm / dp,
// phi(r) = (r + |r|) / (1 + |r|), and 0 where |dp| <= 1e-12.  It equals the
// harmonic mean 2 dm dp / (dm + dp) where r >= 0 and 0 where r < 0, which
// needs one reciprocal; its partials there, 2 (dp / (dm + dp))^2 and
// 2 (dm / (dm + dp))^2, are at r = 0 the one-sided limits that the
// convention |x|' = +1 at 0 gives.  With dm, dp of one sign (or dm = 0),
// |dm + dp| >= |dp| > 1e-12.
// ---------------------------------------------------------------------------
struct VL {
    double dm, dp, iS;
    bool act;
};

__device__ __forceinline__ VL vl_eval(double qm, double q, double qp)
{
    VL v;
    v.dm = q - qm;
    v.dp = qp - q;
    v.act = (fabs0(v.dp) > kT12) && ((v.dp > 0.0) ? (v.dm >= 0.0) : (v.dm <= 0.0));
    v.iS = rcp(v.act ? v.dm + v.dp : 1.0);
    return v;
}

__device__ __forceinline__ double vl_s(const VL& v)
{
    return v.act ? 2.0 * v.dm * v.dp * v.iS : 0.0;
}

__device__ __forceinline__ void vl_rev(const VL& v, double sb, double& qmb, double& qb,
                                       double& qpb)
{
    if (!v.act) return;
    const double a = v.dp * v.iS, b = v.dm * v.iS;
    const double dmb = 2.0 * sb * a * a;
    const double dpb = 2.0 * sb * b * b;
    qb += dmb - dpb;
    qmb -= dmb;
    qpb += dpb;
}

// A cell's primitives as the hydrostatic face uses them: 1/h, u and v where
// h > 1e-12 (HydrostaticRoe.cpp:273-341), zero elsewhere.
__device__ __forceinline__ void cell_prims(double h, double hu, double hv, double& ih, double& u,
                                           double& v)
{
    ih = (h > kT12) ? rcp(h) : 0.0;
    u = hu * ih;
    v = hv * ih;
}

// ---------------------------------------------------------------------------
// Hydrostatic-reconstruction face (HydrostaticRoe.cpp:218-396), normal to
// direction DIR.  Cell k of the stencil {Lm, L, R, Rp}, quantity m of
// {h, u, v, z, 1/h} is cs[k sk + m sm] (h rebuilt; u, v and 1/h zero where
// h <= 1e-12, as cell_prims makes them).  FLb, FRb: adjoints of the two face
// fluxes.  cb (16): adjoints of {h, hu, hv, z} per cell.
// The forward keeps only the face values; the reverse reads the cells again
// and re-evaluates the slopes it differentiates, so no slope intermediate is
// live across the Roe core.  The homogeneous Roe flux takes the face depths
// and velocities: (h u) / h in the forward is u up to rounding, so the
// product and the division cancel in the derivative.
// ---------------------------------------------------------------------------
template <int DIR>
__device__ __forceinline__ void hydro_face_vjp(const double* cs, int sk, int sm,
                                               const double* FLb, const double* FRb, double* cb)
{
    auto c = [&](int k, int m) { return cs[k * sk + m * sm]; };
    for (int k = 0; k < 16; ++k) cb[k] = 0.0;
    const double zL = c(1, 3), zR = c(2, 3);
    if (fabs0(zL + 9999.0) < kT10 || fabs0(zR + 9999.0) < kT10) return;   // NODATA wall
    const bool zfR = zL < zR;
    const double zf = zfR ? zR : zL;
    double etaLf, etaRf, uLf, uRf, vLf, vRf;
    {
        const double e0 = c(0, 0) + c(0, 3), e1 = c(1, 0) + zL, e2 = c(2, 0) + zR,
                     e3 = c(3, 0) + c(3, 3);
        etaLf = e1 + 0.5 * vl_s(vl_eval(e0, e1, e2));
        etaRf = e2 - 0.5 * vl_s(vl_eval(e1, e2, e3));
        const double u0 = c(0, 1), u1 = c(1, 1), u2 = c(2, 1), u3 = c(3, 1);
        uLf = u1 + 0.5 * vl_s(vl_eval(u0, u1, u2));
        uRf = u2 - 0.5 * vl_s(vl_eval(u1, u2, u3));
        const double v0 = c(0, 2), v1 = c(1, 2), v2 = c(2, 2), v3 = c(3, 2);
        vLf = v1 + 0.5 * vl_s(vl_eval(v0, v1, v2));
        vRf = v2 - 0.5 * vl_s(vl_eval(v1, v2, v3));
    }
    const double xLu = etaLf - zL, xRu = etaRf - zR, xLf = etaLf - zf, xRf = etaRf - zf;
    const double hLu = (0.0 < xLu) ? xLu : 0.0, hRu = (0.0 < xRu) ? xRu : 0.0;
    const double hLf = (0.0 < xLf) ? xLf : 0.0, hRf = (0.0 < xRf) ? xRf : 0.0;
    RoeTape t;
    roe_core_fwd<DIR, true>(hLf, uLf, vLf, hRf, uRf, vRf, t);
    // ---- reverse ----
    // FL = rot(f) + (0, pL nx, pL ny), FR likewise
    const double fb[3] = {FLb[0] + FRb[0], FLb[1] + FRb[1], FLb[2] + FRb[2]};
    const double pLb = Nrm<DIR>::un(FLb[1], FLb[2]);
    const double pRb = Nrm<DIR>::un(FRb[1], FRb[2]);
    double hLub = pLb * kG * hLu, hRub = pRb * kG * hRu;
    double hLfb = -pLb * kG * hLf, hRfb = -pRb * kG * hRf;
    double uLfb, vLfb, uRfb, vRfb, hLfb2, hRfb2;
    RoeBar A{};
    roe_core_rev<DIR, true>(t, fb[0], Nrm<DIR>::un(fb[1], fb[2]), Nrm<DIR>::ut(fb[1], fb[2]), A,
                            &hLfb2, &uLfb, &vLfb, &hRfb2, &uRfb, &vRfb);
    hLfb += hLfb2;
    hRfb += hRfb2;
    double etaLfb = 0.0, etaRfb = 0.0, zfb = 0.0, zLb = 0.0, zRb = 0.0;
    if (0.0 < xLf) { etaLfb += hLfb; zfb -= hLfb; }
    if (0.0 < xRf) { etaRfb += hRfb; zfb -= hRfb; }
    if (0.0 < xLu) { etaLfb += hLub; zLb -= hLub; }
    if (0.0 < xRu) { etaRfb += hRub; zRb -= hRub; }
    // the cells are read again below rather than kept in registers
    asm volatile("" ::: "memory");
    // velocity reconstruction
    {
        const double u0 = c(0, 1), u1 = c(1, 1), u2 = c(2, 1), u3 = c(3, 1);
        const double v0 = c(0, 2), v1 = c(1, 2), v2 = c(2, 2), v3 = c(3, 2);
        double ub[4] = {0.0, uLfb, uRfb, 0.0}, vb[4] = {0.0, vLfb, vRfb, 0.0};
        vl_rev(vl_eval(u0, u1, u2), 0.5 * uLfb, ub[0], ub[1], ub[2]);
        vl_rev(vl_eval(u1, u2, u3), -0.5 * uRfb, ub[1], ub[2], ub[3]);
        vl_rev(vl_eval(v0, v1, v2), 0.5 * vLfb, vb[0], vb[1], vb[2]);
        vl_rev(vl_eval(v1, v2, v3), -0.5 * vRfb, vb[1], vb[2], vb[3]);
        const double uq[4] = {u0, u1, u2, u3}, vq[4] = {v0, v1, v2, v3};
        for (int q = 0; q < 4; ++q) {
            const double ih = c(q, 4);
            if (ih != 0.0) {
                cb[4 * q + 1] += ub[q] * ih;
                cb[4 * q + 2] += vb[q] * ih;
                cb[4 * q + 0] -= (ub[q] * uq[q] + vb[q] * vq[q]) * ih;
            }
        }
    }
    // free-surface reconstruction
    {
        const double e0 = c(0, 0) + c(0, 3), e1 = c(1, 0) + zL, e2 = c(2, 0) + zR,
                     e3 = c(3, 0) + c(3, 3);
        double eb[4] = {0.0, etaLfb, etaRfb, 0.0};
        vl_rev(vl_eval(e0, e1, e2), 0.5 * etaLfb, eb[0], eb[1], eb[2]);
        vl_rev(vl_eval(e1, e2, e3), -0.5 * etaRfb, eb[1], eb[2], eb[3]);
        for (int q = 0; q < 4; ++q) {
            cb[4 * q + 0] += eb[q];
            cb[4 * q + 3] += eb[q];
        }
    }
    if (zfR) zRb += zfb; else zLb += zfb;
    cb[4 + 3] += zLb;
    cb[8 + 3] += zRb;
}

// ---------------------------------------------------------------------------
// Manning friction and dry clamp of the hydrostatic update
// (HydrostaticRoe.cpp:566-603).  un: state after the flux update; ob:
// adjoint of the output.  unb: adjoint of un; returns d/dn.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double hydro_post_vjp(const double* un, double n, double dt,
                                                 const double* ob, double* unb)
{
    const double hn = un[0];
    if (hn >= kDry) {
        const double qx = un[1], qy = un[2];
        const double nsq = n * n;
        const double inv_h = rcp(hn);
        const double u = qx * inv_h, v = qy * inv_h;
        const double m2 = u * u + v * v;
        double modM, rM;
        sqrt_rsq(m2 > 0.0 ? m2 : 1.0, modM, rM);
        if (!(m2 > 0.0)) modM = 0.0;
        unb[0] = ob[0];
        if (nsq > kT12 && modM > kT12) {
            const double cb = cbrt(hn);
            const double h43 = hn * cb;
            const double ih43 = rcp(h43);
            const double tt = dt * kG * nsq * modM * ih43;
            double ff, rff;
            sqrt_rsq(1.0 + 4.0 * tt, ff, rff);
            const double i1f = rcp(1.0 + ff);
            const double fac = 2.0 * i1f;
            const double facb = ob[1] * qx + ob[2] * qy;
            double qxb = ob[1] * fac, qyb = ob[2] * fac;
            const double ffb = -facb * 2.0 * i1f * i1f;
            const double ttb = ffb * 2.0 * rff;
            const double nsqb = ttb * dt * kG * modM * ih43;
            const double modMb = ttb * dt * kG * nsq * ih43;
            const double h43b = -ttb * tt * ih43;
            double hnb = h43b * (cb + hn * rcp(3.0 * cb * cb));
            const double ub = modMb * u * rM, vb = modMb * v * rM;
            qxb += ub * inv_h;
            qyb += vb * inv_h;
            hnb -= (ub * u + vb * v) * inv_h;
            unb[0] += hnb;
            unb[1] = qxb;
            unb[2] = qyb;
            return nsqb * 2.0 * n;
        }
        unb[1] = ob[1];
        unb[2] = ob[2];
        return 0.0;
    }
    unb[0] = (hn < 0.0) ? 0.0 : ob[0];
    unb[1] = 0.0;
    unb[2] = 0.0;
    return 0.0;
}

}  // namespace adj
}  // namespace hx
