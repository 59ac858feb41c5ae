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
constexpr double kDry = 1e-6;      // include/utils/Constants.H:63
constexpr double kT12 = 1e-12;     // include/utils/Constants.H:70
constexpr double kT10 = 1e-10;     // include/utils/Constants.H:78

__host__ __device__ inline double sgn0(double x) { return x >= 0.0 ? 1.0 : -1.0; }
__host__ __device__ inline double fabs0(double x) { return x >= 0.0 ? x : -x; }

// ---------------------------------------------------------------------------
// Shared Roe core: dry cut, Roe averages, entropy-fixed eigenvalues, wave
// strengths and the unrotated flux after the wave sum.
// ---------------------------------------------------------------------------
struct RoeTape {
    double hl, hr, sql, sqr, ul, vl, ur, vr, unl, utl, unr, utr;
    double ht, ct, dsq, ut, vt, cgl, cgr;
    double el0, er0, el2, er2, lam0_0, lam2_0;
    double lam0, lam1, lam2, den, dh, a0, a1, a2;
    double w0, w1, w2;
    double F0, F1, F2;        // unrotated, after the wave sum
    bool cutl, cutr, htact, fix0, fix2;
};

__host__ __device__ inline void roe_core_fwd(double hl, double hul, double hvl, double hr,
                                             double hur, double hvr, double nx, double ny,
                                             RoeTape& t)
{
    t.cutl = hl >= kDry;
    t.cutr = hr >= kDry;
    t.hl = hl = t.cutl ? hl : 0.0;
    t.hr = hr = t.cutr ? hr : 0.0;
    t.sql = (hl > 0.0) ? sqrt(hl) : 0.0;
    t.sqr = (hr > 0.0) ? sqrt(hr) : 0.0;
    t.ul = (hl > 0.0) ? hul / hl : 0.0;
    t.vl = (hl > 0.0) ? hvl / hl : 0.0;
    t.ur = (hr > 0.0) ? hur / hr : 0.0;
    t.vr = (hr > 0.0) ? hvr / hr : 0.0;
    t.unl = t.ul * nx + t.vl * ny;
    t.utl = -t.ul * ny + t.vl * nx;
    t.unr = t.ur * nx + t.vr * ny;
    t.utr = -t.ur * ny + t.vr * nx;
    const double hmean = 0.5 * (hl + hr);
    t.htact = hmean > kDry;
    t.ht = t.htact ? hmean : kDry;
    t.ct = sqrt(kG * t.ht);
    t.dsq = t.sql + t.sqr + kT12;
    t.ut = (t.unl * t.sql + t.unr * t.sqr) / t.dsq;
    t.vt = (t.utl * t.sql + t.utr * t.sqr) / t.dsq;
    t.lam0_0 = t.ut - t.ct;
    t.lam1 = t.ut;
    t.lam2_0 = t.ut + t.ct;
    t.cgl = sqrt(kG * hl);
    t.cgr = sqrt(kG * hr);
    t.el0 = t.unl - t.cgl;
    t.er0 = t.unr - t.cgr;
    t.fix0 = (t.el0 < 0.0 && t.er0 > 0.0);
    t.lam0 = t.fix0 ? t.el0 * (t.er0 - t.lam0_0) / (t.er0 - t.el0) : t.lam0_0;
    t.el2 = t.unl + t.cgl;
    t.er2 = t.unr + t.cgr;
    t.fix2 = (t.el2 < 0.0 && t.er2 > 0.0);
    t.lam2 = t.fix2 ? t.er2 * (t.lam2_0 - t.el2) / (t.er2 - t.el2) : t.lam2_0;
    t.den = 2.0 * t.ct + kT12;
    t.dh = hr - hl;
    const double dhu = hr * t.unr - hl * t.unl;
    const double dhv = hr * t.utr - hl * t.utl;
    t.a0 = (t.lam2 * t.dh - dhu) / t.den;
    t.a1 = (-t.vt * t.dh + dhv) / t.ct;
    t.a2 = (-t.lam0 * t.dh + dhu) / t.den;
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
__host__ __device__ inline void roe_core_rev(const RoeTape& t, double nx, double ny, double Fb0,
                                             double Fb1, double Fb2, RoeBar A, double* hlb_o,
                                             double* hulb, double* hvlb, double* hrb_o,
                                             double* hurb, double* hvrb)
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
        double nb = a0b / t.den;
        denb += -a0b * t.a0 / t.den;
        lam2b += nb * t.dh;
        dhb += nb * t.lam2;
        dhub -= nb;
        nb = a1b / t.ct;
        ctb += -a1b * t.a1 / t.ct;
        vtb += -nb * t.dh;
        dhb += -nb * t.vt;
        dhvb += nb;
        nb = a2b / t.den;
        denb += -a2b * t.a2 / t.den;
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
        const double D = t.er2 - t.el2, M = t.lam2_0 - t.el2, L = lam2b;
        const double er2b = L * M / D - L * t.er2 * M / (D * D);
        const double el2b = -L * t.er2 / D + L * t.er2 * M / (D * D);
        lam2_0b = L * t.er2 / D;
        unlb += el2b;
        cglb += el2b;
        unrb += er2b;
        cgrb += er2b;
    }
    if (t.fix0) {
        const double D = t.er0 - t.el0, M = t.er0 - t.lam0_0, L = lam0b;
        const double el0b = L * M / D + L * t.el0 * M / (D * D);
        const double er0b = L * t.el0 / D - L * t.el0 * M / (D * D);
        lam0_0b = -L * t.el0 / D;
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
        double nb = utb / t.dsq;
        double dsqb = -utb * t.ut / t.dsq;
        unlb += nb * t.sql;
        sqlb += nb * t.unl;
        unrb += nb * t.sqr;
        sqrb += nb * t.unr;
        nb = vtb / t.dsq;
        dsqb += -vtb * t.vt / t.dsq;
        utlb += nb * t.sql;
        sqlb += nb * t.utl;
        utrb += nb * t.sqr;
        sqrb += nb * t.utr;
        sqlb += dsqb;
        sqrb += dsqb;
    }
    htb += ctb * kG / (2.0 * t.ct);
    if (t.htact) {
        hlb += 0.5 * htb;
        hrb += 0.5 * htb;
    }
    if (t.hl > 0.0) hlb += cglb * kG / (2.0 * t.cgl);
    if (t.hr > 0.0) hrb += cgrb * kG / (2.0 * t.cgr);
    // rotation to the face frame
    const double ulb = unlb * nx - utlb * ny, vlb = unlb * ny + utlb * nx;
    const double urb = unrb * nx - utrb * ny, vrb = unrb * ny + utrb * nx;
    double hulb_ = 0.0, hvlb_ = 0.0, hurb_ = 0.0, hvrb_ = 0.0;
    if (t.hl > 0.0) {
        hulb_ = ulb / t.hl;
        hvlb_ = vlb / t.hl;
        hlb -= (ulb * t.ul + vlb * t.vl) / t.hl;
        hlb += sqlb / (2.0 * t.sql);
    }
    if (t.hr > 0.0) {
        hurb_ = urb / t.hr;
        hvrb_ = vrb / t.hr;
        hrb -= (urb * t.ur + vrb * t.vr) / t.hr;
        hrb += sqrb / (2.0 * t.sqr);
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
__host__ __device__ inline void roe_face_vjp(const double* L, const double* R, double ds,
                                             double nx, double ny, const double* Fb,
                                             const double* Smb, const double* Spb, double* Lb,
                                             double* Rb)
{
    RoeTape t;
    roe_core_fwd(L[0], L[1], L[2], R[0], R[1], R[2], nx, ny, t);
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
    const double mag = sqrt(t.ut * t.ut + t.vt * t.vt);
    const double cb = cbrt(t.ht);
    const double h43 = t.ht * cb;
    const bool dnact = kDry < h43;
    const double Dn = dnact ? h43 : kDry;
    const double sf = (nt * nt * mag * t.ut) / Dn;
    const double tau = kG * t.ht * sf * ds;
    const double B = (thrust - tau) / t.den;
    // positivity guard (Roe.H:200-247): b0, b2 as signed picks of B, m1, m3
    double c0[3] = {1.0, 0.0, 0.0}, c2[3] = {-1.0, 0.0, 0.0};
    double b0 = B, b2 = -B;
    {
        auto tol = [](double x) { return (fabs0(x) > kT12) ? x : (x >= 0.0 ? kT12 : -kT12); };
        const double l0s = tol(t.lam0), l2s = tol(t.lam2);
        const double h1s = hl + t.a0 - (b0 / l0s);
        const double h3s = hr - t.a2 + (b2 / l2s);
        double mlam = fabs0(t.lam0);
        mlam = (mlam < fabs0(t.lam1)) ? fabs0(t.lam1) : mlam;
        mlam = (mlam < fabs0(t.lam2)) ? fabs0(t.lam2) : mlam;
        const double loc_dt = ds / (mlam + kT12);
        const double d1s = tol(hl - h1s), d3s = tol(hr - h3s);
        const double dt1 = (ds * 0.5 / l0s) * (hl / d1s);
        const double dt3 = (ds * 0.5 / l2s) * (hr / d3s);
        const double m1 = -(hl + t.a0) * fabs0(t.lam0);
        const double m3 = -(hr - t.a2) * t.lam2;
        if ((h1s < -kDry) && (h3s > kDry) && (dt1 < loc_dt)) {
            if (-m1 >= m3) { b0 = m1; c0[0] = 0.0; c0[1] = 1.0; c0[2] = 0.0; }
            b2 = -b0;
            for (int q = 0; q < 3; ++q) c2[q] = -c0[q];
        }
        if ((h1s > kDry) && (h3s < -kDry) && (dt3 < loc_dt)) {
            if (-m3 >= m1) { b2 = m3; c2[0] = 0.0; c2[1] = 0.0; c2[2] = 1.0; }
            b0 = -b2;
            for (int q = 0; q < 3; ++q) c0[q] = -c2[q];
        }
    }
    // ---- reverse ----
    auto unrot1 = [&](const double* X) { return X[1] * nx + X[2] * ny; };
    auto unrot2 = [&](const double* X) { return -X[1] * ny + X[2] * nx; };
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
    const double thb = Bb / t.den;
    const double taub = -Bb / t.den;
    A.den -= Bb * B / t.den;
    A.ht += taub * kG * sf * ds;
    const double sfb = taub * kG * t.ht * ds;
    const double numb = sfb / Dn;
    const double Dnb = -sfb * sf / Dn;
    const double ntb = numb * 2.0 * nt * mag * t.ut;
    const double magb = numb * nt * nt * t.ut;
    A.ut += numb * nt * nt * mag;
    if (dnact) A.ht += Dnb * (cb + t.ht / (3.0 * cb * cb));
    if (mag > 0.0) {
        A.ut += magb * t.ut / mag;
        A.vt += magb * t.vt / mag;
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
    roe_core_rev(t, nx, ny, Fb[0], unrot1(Fb), unrot2(Fb), A, &Lb[0], &Lb[1], &Lb[2], &Rb[0],
                 &Rb[1], &Rb[2]);
}

// ---------------------------------------------------------------------------
// van Leer slope and its reverse (HydrostaticRoe.cpp:173-195).
// ---------------------------------------------------------------------------
__host__ __device__ inline double vl_fwd(double qm, double q, double qp)
{
    const double dm = q - qm, dp = qp - q;
    if (fabs0(dp) <= kT12) return 0.0;
    const double r = dm / dp;
    return ((r + fabs0(r)) / (1.0 + fabs0(r))) * dp;
}

__host__ __device__ inline void vl_rev(double qm, double q, double qp, double sb, double& qmb,
                                       double& qb, double& qpb)
{
    const double dm = q - qm, dp = qp - q;
    if (fabs0(dp) <= kT12 || sb == 0.0) return;
    const double r = dm / dp;
    const double ar = fabs0(r), sg = sgn0(r);
    const double N = r + ar, D = 1.0 + ar;
    const double phi = N / D;
    const double phib = sb * dp;
    double dpb = sb * phi;
    const double rb = (phib / D) * (1.0 + sg) - (phib * N / (D * D)) * sg;
    const double dmb = rb / dp;
    dpb -= rb * r / dp;
    qb += dmb - dpb;
    qmb -= dmb;
    qpb += dpb;
}

// ---------------------------------------------------------------------------
// Hydrostatic-reconstruction face, forward only (HydrostaticRoe.cpp:218-396):
// the adjoint step needs the flux update of each cell before friction.
// ---------------------------------------------------------------------------
__host__ __device__ inline void hydro_face_fwd(const double* c, double nx, double ny,
                                               double* FL, double* FR)
{
    const double *Lm = c, *L = c + 4, *R = c + 8, *Rp = c + 12;
    const double zL = L[3], zR = R[3];
    if (fabs0(zL + 9999.0) < kT10 || fabs0(zR + 9999.0) < kT10) {
        for (int k = 0; k < 3; ++k) { FL[k] = 0.0; FR[k] = 0.0; }
        return;
    }
    const double zf = (zL < zR) ? zR : zL;
    const double eLm = Lm[0] + Lm[3], eL = L[0] + zL, eR = R[0] + zR, eRp = Rp[0] + Rp[3];
    const double etaLf = eL + 0.5 * vl_fwd(eLm, eL, eR);
    const double etaRf = eR - 0.5 * vl_fwd(eL, eR, eRp);
    const double* X[4] = {Lm, L, R, Rp};
    double u[4], v[4];
    for (int q = 0; q < 4; ++q) {
        const bool w = X[q][0] > kT12;
        u[q] = w ? X[q][1] / X[q][0] : 0.0;
        v[q] = w ? X[q][2] / X[q][0] : 0.0;
    }
    const double uLf = u[1] + 0.5 * vl_fwd(u[0], u[1], u[2]);
    const double uRf = u[2] - 0.5 * vl_fwd(u[1], u[2], u[3]);
    const double vLf = v[1] + 0.5 * vl_fwd(v[0], v[1], v[2]);
    const double vRf = v[2] - 0.5 * vl_fwd(v[1], v[2], v[3]);
    const double xLu = etaLf - zL, xRu = etaRf - zR, xLf = etaLf - zf, xRf = etaRf - zf;
    const double hLu = (0.0 < xLu) ? xLu : 0.0;
    const double hRu = (0.0 < xRu) ? xRu : 0.0;
    const double hLf = (0.0 < xLf) ? xLf : 0.0;
    const double hRf = (0.0 < xRf) ? xRf : 0.0;
    RoeTape t;
    roe_core_fwd(hLf, (hLf > 0.0) ? hLf * uLf : 0.0, (hLf > 0.0) ? hLf * vLf : 0.0, hRf,
                 (hRf > 0.0) ? hRf * uRf : 0.0, (hRf > 0.0) ? hRf * vRf : 0.0, nx, ny, t);
    const double f1 = t.F1 * nx - t.F2 * ny, f2 = t.F1 * ny + t.F2 * nx;
    const double pL = 0.5 * kG * (hLu * hLu - hLf * hLf);
    const double pR = 0.5 * kG * (hRu * hRu - hRf * hRf);
    FL[0] = t.F0; FL[1] = f1 + pL * nx; FL[2] = f2 + pL * ny;
    FR[0] = t.F0; FR[1] = f1 + pR * nx; FR[2] = f2 + pR * ny;
}

// ---------------------------------------------------------------------------
// Hydrostatic-reconstruction face (HydrostaticRoe.cpp:218-396).
// c: {Lm, L, R, Rp} x {h, hu, hv, z} (h rebuilt).  FLb, FRb: adjoints of
// the two face fluxes.  cb (16): adjoints of c.
// ---------------------------------------------------------------------------
__host__ __device__ inline void hydro_face_vjp(const double* c, double nx, double ny,
                                               const double* FLb, const double* FRb, double* cb)
{
    for (int k = 0; k < 16; ++k) cb[k] = 0.0;
    const double *Lm = c, *L = c + 4, *R = c + 8, *Rp = c + 12;
    const double zL = L[3], zR = R[3];
    if (fabs0(zL + 9999.0) < kT10 || fabs0(zR + 9999.0) < kT10) return;
    const bool zfR = zL < zR;
    const double zf = zfR ? zR : zL;
    const double eLm = Lm[0] + Lm[3], eL = L[0] + zL, eR = R[0] + zR, eRp = Rp[0] + Rp[3];
    const double etaLf = eL + 0.5 * vl_fwd(eLm, eL, eR);
    const double etaRf = eR - 0.5 * vl_fwd(eL, eR, eRp);
    const double* X[4] = {Lm, L, R, Rp};
    double u[4], v[4];
    bool wv[4];
    for (int q = 0; q < 4; ++q) {
        wv[q] = X[q][0] > kT12;
        u[q] = wv[q] ? X[q][1] / X[q][0] : 0.0;
        v[q] = wv[q] ? X[q][2] / X[q][0] : 0.0;
    }
    const double uLf = u[1] + 0.5 * vl_fwd(u[0], u[1], u[2]);
    const double uRf = u[2] - 0.5 * vl_fwd(u[1], u[2], u[3]);
    const double vLf = v[1] + 0.5 * vl_fwd(v[0], v[1], v[2]);
    const double vRf = v[2] - 0.5 * vl_fwd(v[1], v[2], v[3]);
    const double xLu = etaLf - zL, xRu = etaRf - zR, xLf = etaLf - zf, xRf = etaRf - zf;
    const double hLu = (0.0 < xLu) ? xLu : 0.0;
    const double hRu = (0.0 < xRu) ? xRu : 0.0;
    const double hLf = (0.0 < xLf) ? xLf : 0.0;
    const double hRf = (0.0 < xRf) ? xRf : 0.0;
    const double huLf = (hLf > 0.0) ? hLf * uLf : 0.0;
    const double hvLf = (hLf > 0.0) ? hLf * vLf : 0.0;
    const double huRf = (hRf > 0.0) ? hRf * uRf : 0.0;
    const double hvRf = (hRf > 0.0) ? hRf * vRf : 0.0;
    RoeTape t;
    roe_core_fwd(hLf, huLf, hvLf, hRf, huRf, hvRf, nx, ny, t);
    // ---- reverse ----
    // FL = rot(f) + (0, pL nx, pL ny), FR likewise
    const double fb[3] = {FLb[0] + FRb[0], FLb[1] + FRb[1], FLb[2] + FRb[2]};
    const double pLb = FLb[1] * nx + FLb[2] * ny;
    const double pRb = FRb[1] * nx + FRb[2] * ny;
    double hLub = pLb * kG * hLu, hRub = pRb * kG * hRu;
    double hLfb = -pLb * kG * hLf, hRfb = -pRb * kG * hRf;
    double huLfb, hvLfb, huRfb, hvRfb, hLfb2, hRfb2;
    RoeBar A{};
    roe_core_rev(t, nx, ny, fb[0], fb[1] * nx + fb[2] * ny, -fb[1] * ny + fb[2] * nx, A, &hLfb2,
                 &huLfb, &hvLfb, &hRfb2, &huRfb, &hvRfb);
    hLfb += hLfb2;
    hRfb += hRfb2;
    double uLfb = 0.0, vLfb = 0.0, uRfb = 0.0, vRfb = 0.0;
    if (hLf > 0.0) {
        hLfb += huLfb * uLf + hvLfb * vLf;
        uLfb = huLfb * hLf;
        vLfb = hvLfb * hLf;
    }
    if (hRf > 0.0) {
        hRfb += huRfb * uRf + hvRfb * vRf;
        uRfb = huRfb * hRf;
        vRfb = hvRfb * hRf;
    }
    double etaLfb = 0.0, etaRfb = 0.0, zfb = 0.0, zLb = 0.0, zRb = 0.0;
    if (0.0 < xLf) { etaLfb += hLfb; zfb -= hLfb; }
    if (0.0 < xRf) { etaRfb += hRfb; zfb -= hRfb; }
    if (0.0 < xLu) { etaLfb += hLub; zLb -= hLub; }
    if (0.0 < xRu) { etaRfb += hRub; zRb -= hRub; }
    // velocity reconstruction
    double ub[4] = {0.0, 0.0, 0.0, 0.0}, vb[4] = {0.0, 0.0, 0.0, 0.0};
    ub[1] += uLfb;
    vl_rev(u[0], u[1], u[2], 0.5 * uLfb, ub[0], ub[1], ub[2]);
    ub[2] += uRfb;
    vl_rev(u[1], u[2], u[3], -0.5 * uRfb, ub[1], ub[2], ub[3]);
    vb[1] += vLfb;
    vl_rev(v[0], v[1], v[2], 0.5 * vLfb, vb[0], vb[1], vb[2]);
    vb[2] += vRfb;
    vl_rev(v[1], v[2], v[3], -0.5 * vRfb, vb[1], vb[2], vb[3]);
    for (int q = 0; q < 4; ++q) {
        if (wv[q]) {
            const double h = X[q][0];
            cb[4 * q + 1] += ub[q] / h;
            cb[4 * q + 2] += vb[q] / h;
            cb[4 * q + 0] -= (ub[q] * u[q] + vb[q] * v[q]) / h;
        }
    }
    // free-surface reconstruction
    double eb[4] = {0.0, etaLfb, etaRfb, 0.0};
    vl_rev(eLm, eL, eR, 0.5 * etaLfb, eb[0], eb[1], eb[2]);
    vl_rev(eL, eR, eRp, -0.5 * etaRfb, eb[1], eb[2], eb[3]);
    for (int q = 0; q < 4; ++q) {
        cb[4 * q + 0] += eb[q];
        cb[4 * q + 3] += eb[q];
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
__host__ __device__ inline double hydro_post_vjp(const double* un, double n, double dt,
                                                 const double* ob, double* unb)
{
    const double hn = un[0];
    if (hn >= kDry) {
        const double qx = un[1], qy = un[2];
        const double nsq = n * n;
        const double inv_h = 1.0 / hn;
        const double u = qx * inv_h, v = qy * inv_h;
        const double modM = sqrt(u * u + v * v);
        unb[0] = ob[0];
        if (nsq > kT12 && modM > kT12) {
            const double cb = cbrt(hn);
            const double h43 = hn * cb;
            const double tt = dt * kG * nsq * modM / h43;
            const double ff = sqrt(1.0 + 4.0 * tt);
            const double fac = 2.0 / (1.0 + ff);
            const double facb = ob[1] * qx + ob[2] * qy;
            double qxb = ob[1] * fac, qyb = ob[2] * fac;
            const double ffb = -facb * 2.0 / ((1.0 + ff) * (1.0 + ff));
            const double ttb = ffb * 2.0 / ff;
            const double nsqb = ttb * dt * kG * modM / h43;
            const double modMb = ttb * dt * kG * nsq / h43;
            const double h43b = -ttb * tt / h43;
            double hnb = h43b * (cb + hn / (3.0 * cb * cb));
            const double ub = modMb * u / modM, vb = modMb * v / modM;
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
