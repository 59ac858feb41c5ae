// hx_device.hpp -- device-side building blocks of the fused per-level advance
// for CDNA4 (gfx950): halo-aware cell loads, per-cell hoisted quantities, the
// plain Roe face with split sources, the hydrostatic-reconstruction face, and
// the CFL accumulator.
//
// Included once per arithmetic build, inside namespace hx::HX_NS, with
// HX_FAST set:
//   HX_FAST 0 (exact): every division is a correctly rounded '/', compiled
//     with -ffp-contract=off.  Arithmetic follows the reference operation by
//     operation, so the FP64 result equals the FP64 CPU restatement bit for
//     bit (friction's cbrt aside).  It departs from the literal reference
//     expression only by
//       * hoisting a per-cell value that the reference recomputes per face
//         (same operands, same operations, same rounding),
//       * dropping a term the reference multiplies by an exact 0.0 or 1.0
//         (axis-normal rotation, the zero contact-wave source), which can
//         change at most the sign of a zero,
//       * evaluating the positivity guard's timing test only when its depth
//         test passes (no side effects, so the outcome is unchanged),
//       * multiplying by 1/dx when dx is a power of two (an exact scaling).
//   HX_FAST 1 (fast): the same expressions with divisions that share a
//     denominator turned into one Newton-refined reciprocal, and FMA
//     contraction within an expression (-ffp-contract=on).  Each quotient is within ~2 ulp; the parity tests
//     hold this build to 1e-12 relative L-infinity against the oracle.

namespace HX_NS {

template <typename R>
struct Kc {
    static constexpr R g = R(9.81);        // include/utils/Constants.H:25
    static constexpr R toldry = R(1e-6);   // include/utils/Constants.H:63
    static constexpr R tol12 = R(1e-12);   // include/utils/Constants.H:70
    static constexpr R tol10 = R(1e-10);   // include/utils/Constants.H:78
};

// std::max / std::min semantics.
template <typename R> __device__ __forceinline__ R smax(R a, R b) { return (a < b) ? b : a; }
template <typename R> __device__ __forceinline__ R smin(R a, R b) { return (b < a) ? b : a; }

#if HX_FAST
// Reciprocal from the hardware estimate r (relative error ~2^-23) and one
// third-order step: 1/b = r / (1 - e) = r (1 + e + e^2 + ...) with
// e = 1 - b r, truncated after e^2 (the e^3 term is below 2^-66).  FP32 takes
// one Newton step r + r e.  Callers only pass finite, non-zero b.
__device__ __forceinline__ double rcp(double b)
{
    const double r = __builtin_amdgcn_rcp(b);
    const double e = fma(-b, r, 1.0);
    return fma(r, fma(e, e, e), r);
}
__device__ __forceinline__ float rcp(float b)
{
    float r = __builtin_amdgcn_rcpf(b);
    const float e = fmaf(-b, r, 1.0f);
    return fmaf(r, e, r);
}
template <typename R>
struct Div {
    R r;
    __device__ __forceinline__ explicit Div(R b) : r(rcp(b)) {}
    __device__ __forceinline__ R operator()(R a) const { return a * r; }
};

// sqrt(x) and 1/sqrt(x) together for finite x > 0.  FP64: the hardware
// estimate y (relative error ~2^-23) and one third-order step,
// 1/sqrt(x) = y (1 - d)^(-1/2) = y (1 + d/2 + 3 d^2/8 + ...) with
// d = 1 - x y^2 (the d^3 term is below 2^-64), and sqrt(x) = (x y) times the
// same factor.  FP32: one Goldschmidt step g <- g + g e, h <- h + h e with
// e = 1/2 - g h, g -> sqrt(x), 2h -> 1/sqrt(x).
__device__ __forceinline__ void sqrt_rsq(double x, double& sq, double& rs)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double t = x * y;
    const double d = fma(-t, y, 1.0);
    const double c = d * fma(d, 0.375, 0.5);
    rs = fma(y, c, y);
    sq = fma(t, c, t);
}
__device__ __forceinline__ void sqrt_rsq(float x, float& sq, float& rs)
{
    const float y = __builtin_amdgcn_rsqf(x);
    float g = x * y, h = 0.5f * y;
    const float e = fmaf(-g, h, 0.5f);
    sq = fmaf(g, e, g);
    rs = 2.0f * fmaf(h, e, h);
}

// x^(-1/3) for x in [2^-100, 2^100].  Seed: exp2(-log2(x) / 3) from the
// binary32 hardware log2 and exp2 (measured relative error 7.2e-7 = 2^-20.4
// on 4M arguments from 1e-6 to 1e4).  FP64 then takes one third-order step of
// y^-3 = x: y = y0 (1 - e)^(-1/3) = y0 (1 + e/3 + 2 e^2/9 + ...),
// e = 1 - x y0^3 (the e^3 term is then below 2^-58); FP32 one Newton step
// y0 (1 + e/3).  Measured on the same arguments: 0.76 ulp (FP64), 1.12 ulp
// (FP32); y^3 is within 4 ulp of 1/x and y^4 within 6 ulp of x^(-4/3).
// Manning friction uses it for h^(-4/3) = y^4 and 1/h = y^3 without a cube
// root or a division.
__device__ __forceinline__ double inv_cbrt(double x)
{
    const float l = __builtin_amdgcn_logf((float)x);
    const double y = (double)__builtin_amdgcn_exp2f(l * (-1.0f / 3.0f));
    const double e = fma(-x, (y * y) * y, 1.0);
    const double c = e * fma(e, 2.0 / 9.0, 1.0 / 3.0);
    return fma(y, c, y);
}
__device__ __forceinline__ float inv_cbrt(float x)
{
    const float y = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (-1.0f / 3.0f));
    const float e = fmaf(-x, (y * y) * y, 1.0f);
    return fmaf(y, e * (1.0f / 3.0f), y);
}

template <typename R> struct SqrtG;
template <> struct SqrtG<double> { static constexpr double v = 3.132091952673165; };
template <> struct SqrtG<float> { static constexpr float v = 3.132092f; };
#else
template <typename R>
struct Div {
    R b;
    __device__ __forceinline__ explicit Div(R b_) : b(b_) {}
    __device__ __forceinline__ R operator()(R a) const { return a / b; }
};
#endif

// Cell as the advance sees it: hr is the rebuilt depth max(0, eta - z)
// (src/solvers/Roe.cpp:244-247, src/solvers/HydrostaticRoe.cpp:495-499),
// h the stored depth, qx/qy the momenta after the physical BC.
template <typename R>
struct Raw {
    R hr, h, qx, qy, z, n;
};

// Index mapping of one side's physical BC (include/boundaries/BCFill.H:49-82):
// the state source is the mirror cell for a reflective wall and the edge
// cell for a transmissive side; terrain is always foextrap
// (src/state/AmrMeshState_Init.cpp:62-70), i.e. the edge cell.
__device__ __forceinline__ void map_axis(int i, int n, int blo, int bhi, int& src, int& edge,
                                         bool& flip)
{
    src = i; edge = i; flip = false;
    if (i < 0) {
        if (blo >= 0) { edge = 0; flip = (blo == 0); src = flip ? (-1 - i) : 0; }
    } else if (i >= n) {
        if (bhi >= 0) { edge = n - 1; flip = (bhi == 0); src = flip ? (2 * n - 1 - i) : n - 1; }
    }
}

// Load for the advance: eta crosses the boundary with the scalar rule, the
// momenta with their vector rules, and h is rebuilt on this cell's own
// (edge) bed -- the reference's FillPatchEta + FillPatch + rebuild sequence
// (src/solvers/Roe.cpp:315-339, src/solvers/HydrostaticRoe.cpp:477-500).
template <typename R>
__device__ __forceinline__ Raw<R> load_adv(const PatchView<R>& p, int i, int j)
{
    int si, ei, sj, ej;
    bool fx, fy;
    map_axis(i, p.nx, p.bl0, p.bh0, si, ei, fx);
    map_axis(j, p.ny, p.bl1, p.bh1, sj, ej, fy);
    const int64_t s = (int64_t)sj * p.pitch + si;
    const int64_t t = (int64_t)ej * p.pitch + ei;
    Raw<R> c;
    c.h = p.h[s];
    c.z = p.z[t];
    c.n = p.n[t];
    R e;
    if (p.eta) {
        e = p.eta[s];
    } else {
        const R zs = (s == t) ? c.z : p.z[s];
        e = c.h + zs;
    }
    c.hr = smax(R(0), e - c.z);
    const R qx = p.hu[s];
    const R qy = p.hv[s];
    c.qx = fx ? -qx : qx;
    c.qy = fy ? -qy : qy;
    return c;
}

// ---------------------------------------------------------------------------
// Plain Roe (include/solvers/kernels/Roe.H:50-309).
// ---------------------------------------------------------------------------

// Per-cell quantities every face of the cell uses (Roe.H:68-82, 114-122,
// 152-153): dry-cut depth, its square root, velocities, sqrt(g h), h + z.
template <typename R>
struct RoeCell {
    R hd, sq, u, v, cg, ed, z, n;
};

template <typename R>
__device__ __forceinline__ RoeCell<R> roe_cell(R hr, R qx, R qy, R z, R n)
{
    RoeCell<R> c;
    const R hd = (hr >= Kc<R>::toldry) ? hr : R(0);
    const bool wet = hd > R(0);
    c.hd = hd;
#if HX_FAST
    // one reciprocal square root gives sqrt(h), 1/h and sqrt(g h); a dry
    // cell takes 1/h = 0 and sqrt(h) = 0 (two selects instead of four)
    R s, r;
    sqrt_rsq(wet ? hd : R(1), s, r);
    const R ih = wet ? r * r : R(0);
    c.sq = wet ? s : R(0);
    c.u = qx * ih;
    c.v = qy * ih;
    c.cg = SqrtG<R>::v * c.sq;
#else
    c.sq = wet ? sqrt(hd) : R(0);
    const Div<R> ih(wet ? hd : R(1));
    c.u = wet ? ih(qx) : R(0);
    c.v = wet ? ih(qy) : R(0);
    c.cg = sqrt(Kc<R>::g * hd);
#endif
    c.ed = hd + z;
    c.z = z;
    c.n = n;
    return c;
}

// sqrt(g h) and h + z of a cell: stored by the exact build, recomputed at
// each use by the fast build (one multiply or add instead of two live
// registers per carried cell).
template <typename R>
__device__ __forceinline__ R cell_cg(const RoeCell<R>& c)
{
#if HX_FAST
    return SqrtG<R>::v * c.sq;
#else
    return c.cg;
#endif
}
template <typename R>
__device__ __forceinline__ R cell_ed(const RoeCell<R>& c)
{
#if HX_FAST
    return c.hd + c.z;
#else
    return c.ed;
#endif
}

// Wavefront votes over the active lanes, straight from the lane mask of the
// comparison (no round trip through a VGPR).
__device__ __forceinline__ bool wave_all(bool p)
{
    return __builtin_amdgcn_ballot_w64(p) == __builtin_amdgcn_read_exec();
}
__device__ __forceinline__ bool wave_any(bool p)
{
    return __builtin_amdgcn_ballot_w64(p) != 0ull;
}

// A face before its source split: the central-minus-upwind flux in the
// face-normal frame and what the split needs (the bed/friction/positivity
// share b0 of the first acoustic wave, b2 of the third, their speeds, the
// Roe tangential velocity).
template <typename R>
struct RoeCore {
    R F0, F1, F2;
    R b0, b2, l0, l2, vt;
    bool src;   // wave-uniform: some lane has a non-zero source share
};

template <typename R>
struct RoeFace {
    R F0, F1, F2, Sm0, Sm1, Sm2, Sp0, Sp1, Sp2;
    bool src;   // wave-uniform: some lane has a non-zero source share
};

// DIR 0: x face, normal (1,0); DIR 1: y face, normal (0,1).  A is the
// low-side cell, B the high-side cell, ds the face-normal cell size.
template <typename R, int DIR>
__device__ __forceinline__ void roe_core(const RoeCell<R>& A, const RoeCell<R>& B, R ds,
                                         RoeCore<R>& o)
{
    using k = Kc<R>;
    const R hl = A.hd, hr = B.hd;
    const R unl = (DIR == 0) ? A.u : A.v;
    const R utl = (DIR == 0) ? A.v : -A.u;
    const R unr = (DIR == 0) ? B.u : B.v;
    const R utr = (DIR == 0) ? B.v : -B.u;

    const R hmean = R(0.5) * (hl + hr);
    const R ht = fmax(hmean, k::toldry);   // = (hmean > TOLDRY) ? hmean : TOLDRY
#if HX_FAST
    // ct and 1/ct from one reciprocal square root; 1/(2 ct + TOL12) from
    // 1/ct to first order in TOL12/(2 ct) (relative error below 1e-19)
    R ct, ict;
    sqrt_rsq(k::g * ht, ct, ict);
    const R qh = R(0.5) * ict;
    const R rden = qh * fma(-k::tol12, qh, R(1));
#else
    const R ct = sqrt(k::g * ht);
#endif
    const Div<R> idsq(A.sq + B.sq + k::tol12);
    const R ut = idsq(unl * A.sq + unr * B.sq);
    const R vt = idsq(utl * A.sq + utr * B.sq);

    R l0 = ut - ct;
    const R l1 = ut;
    R l2 = ut + ct;
    {
        const R cga = cell_cg(A), cgb = cell_cg(B);
        R el = unl - cga, er = unr - cgb;
        if (el < R(0) && er > R(0)) l0 = Div<R>(er - el)(el * (er - l0));
        el = unl + cga;
        er = unr + cgb;
        if (el < R(0) && er > R(0)) l2 = Div<R>(er - el)(er * (l2 - el));
    }

#if HX_FAST
    const auto iden = [rden](R a) { return a * rden; };
    const auto ictd = [ict](R a) { return a * ict; };
#else
    const Div<R> iden(R(2) * ct + k::tol12);
    const Div<R> ictd(ct);
#endif
    const R dh = hr - hl;
    const R dhu = hr * unr - hl * unl;
    const R dhv = hr * utr - hl * utl;
    const R a0 = iden(l2 * dh - dhu);
    const R a1 = ictd(-vt * dh + dhv);
    const R a2 = iden(-l0 * dh + dhu);

    // Bed thrust (Roe.H:151-176) and Manning friction (:182-188).  On a face
    // with a flat bed and zero roughness the reference's thrust is
    // -g (h - 0) * 0 and its tau a signed zero, so the source share is an
    // exact zero: when that holds on every lane of the wavefront the
    // evaluation is skipped (a wave-uniform branch).  With zero roughness
    // on a lane, tau is likewise skipped.
    const R dz = B.z - A.z;
    const R nt = R(0.5) * (A.n + B.n);
    const bool flat = wave_all((dz == R(0)) && (nt == R(0)));
    R b0 = R(0);
    if (!flat) {
        const R eda = cell_ed(A), edb = cell_ed(B);
        const R dd = edb - eda;
        const R thrust_a = -k::g * ht * dz;
        R dzt;
        if ((dz >= R(0)) && (eda < B.z)) dzt = hl;
        else if ((dz < R(0)) && (edb < A.z)) dzt = hr;
        else dzt = dz;
        const R hup = (dz >= R(0)) ? hl : hr;
        const R thrust_b = -k::g * (hup - R(0.5) * fabs(dzt)) * dzt;
        const R thrust = ((dz * dd >= R(0)) && (ut * dz > R(0))) ? smax(thrust_a, thrust_b) : thrust_b;
        R tau = R(0);
        if (nt != R(0)) {
#if HX_FAST
            // 1/max(TOLDRY, ht cbrt(ht)) = min(1/TOLDRY, ht^(-4/3)), ht >= TOLDRY
            const R y = inv_cbrt(ht);
            const R y2 = y * y;
            const R m2 = ut * ut + vt * vt;
            R sm, rm;
            sqrt_rsq(m2 > R(0) ? m2 : R(1), sm, rm);
            const R mag = (m2 > R(0)) ? sm : R(0);
            const R sf = (nt * nt * mag * ut) * smin(R(1) / k::toldry, y2 * y2);
#else
            const R sf = Div<R>(smax(k::toldry, ht * cbrt(ht)))(nt * nt * sqrt(ut * ut + vt * vt) * ut);
#endif
            tau = k::g * ht * sf * ds;
        }
        b0 = iden(thrust - tau);
    }
    R b2 = -b0;

    {   // positivity guard, Roe.H:200-247 (with b0 = 0, b/lambda is 0; the
        // clamped speeds are formed only where they are used)
        const auto clamp_speed = [](R l) {
            return (fabs(l) > k::tol12) ? l : (l >= R(0) ? k::tol12 : -k::tol12);
        };
        R h1s = hl + a0, h3s = hr - a2;
        if (!flat) {
            const R l0s = clamp_speed(l0), l2s = clamp_speed(l2);
#if HX_FAST
            const R il = rcp(l0s * l2s);   // |l0s|, |l2s| >= TOL12: no underflow
            h1s = h1s - b0 * (l2s * il);
            h3s = h3s + b2 * (l0s * il);
#else
            h1s = h1s - Div<R>(l0s)(b0);
            h3s = h3s + Div<R>(l2s)(b2);
#endif
        }
        const bool c1 = (h1s < -k::toldry) && (h3s > k::toldry);
        const bool c3 = (h1s > k::toldry) && (h3s < -k::toldry);
        if (c1 || c3) {
            R mlam = fabs(l0);
            mlam = smax(mlam, fabs(l1));
            mlam = smax(mlam, fabs(l2));
            const R loc_dt = Div<R>(mlam + k::tol12)(ds);
            const R m1 = -(hl + a0) * fabs(l0);
            const R m3 = -(hr - a2) * l2;
            const R l0s = clamp_speed(l0), l2s = clamp_speed(l2);
            if (c1) {
                const R d1 = hl - h1s;
                const R d1s = (fabs(d1) > k::tol12) ? d1 : (d1 >= R(0) ? k::tol12 : -k::tol12);
                const R dt1 = Div<R>(l0s)(ds * R(0.5)) * Div<R>(d1s)(hl);
                if (dt1 < loc_dt) {
                    if (-m1 >= m3) b0 = m1;
                    b2 = -b0;
                }
            } else {
                const R d3 = hr - h3s;
                const R d3s = (fabs(d3) > k::tol12) ? d3 : (d3 >= R(0) ? k::tol12 : -k::tol12);
                const R dt3 = Div<R>(l2s)(ds * R(0.5)) * Div<R>(d3s)(hr);
                if (dt3 < loc_dt) {
                    if (-m3 >= m1) b2 = m3;
                    b0 = -b2;
                }
            }
        }
    }

    // Central flux minus the upwind wave sum (Roe.H:255-278).  The contact
    // wave's eigenvector is (0, 0, ct), so it only touches the third row.
    R F0 = R(0.5) * (hl * unl + hr * unr);
    R F1 = R(0.5) * (hl * unl * unl + R(0.5) * k::g * hl * hl +
                     hr * unr * unr + R(0.5) * k::g * hr * hr);
    R F2 = R(0.5) * (hl * unl * utl + hr * unr * utr);
    const R w0 = R(0.5) * fabs(l0) * a0;
    const R w1 = R(0.5) * fabs(l1) * a1;
    const R w2 = R(0.5) * fabs(l2) * a2;
    F0 -= w0; F1 -= w0 * l0; F2 -= w0 * vt;
    F2 -= w1 * ct;
    F0 -= w2; F1 -= w2 * l2; F2 -= w2 * vt;

    o.F0 = F0; o.F1 = F1; o.F2 = F2;
    o.b0 = b0; o.b2 = b2; o.l0 = l0; o.l2 = l2; o.vt = vt;
    o.src = wave_any(b0 != R(0));
}

#if HX_FAST
// The common face in one pass (fast build).  On a lane where the entropy fix
// does not apply, the bed is flat and smooth and the positivity guard does
// not fire, the face is the central flux minus the upwind wave sum with no
// source share:
//   F0 = (S0 - q0 - q2) / 2,  F1 = (S1 - q0 l0 - q2 l2) / 2,
//   F2 = (S2 - vt (q0 + q2) - |ut| (dhv - vt dh)) / 2
// with S the summed physical fluxes and qk = |lk| ak (the contact wave's
// w1 ct = |ut| a1 ct / 2 with a1 ct = dhv - vt dh, so 1/ct drops out).
// Returns whether that holds on this lane; F is computed unless no lane of
// the wavefront has a flat, smooth bed (then every lane returns false).  A
// zero-sum roughness pair with non-zero members only takes the general path.
template <typename R, int DIR>
__device__ __forceinline__ bool roe_quick(const RoeCell<R>& A, const RoeCell<R>& B, R& F0, R& F1,
                                          R& F2)
{
    using k = Kc<R>;
    const bool rough = (B.z != A.z) || (A.n + B.n != R(0));
    if (wave_all(rough)) return false;
    const R hl = A.hd, hr = B.hd;
    const R unl = (DIR == 0) ? A.u : A.v;
    const R utl = (DIR == 0) ? A.v : -A.u;
    const R unr = (DIR == 0) ? B.u : B.v;
    const R utr = (DIR == 0) ? B.v : -B.u;

    // g max((hl + hr)/2, TOLDRY): scaling by g and halving commute with the
    // max and with rounding
    const R ght = fmax((hl + hr) * (R(0.5) * k::g), k::g * k::toldry);
    R ct, ict;
    sqrt_rsq(ght, ct, ict);
    const R qh = R(0.5) * ict;
    const R rden = qh * fma(-k::tol12, qh, R(1));
    const R isq = rcp(A.sq + B.sq + k::tol12);
    const R ut = (unl * A.sq + unr * B.sq) * isq;
    const R vt = (utl * A.sq + utr * B.sq) * isq;
    const R l0 = ut - ct, l2 = ut + ct;
    const R cga = cell_cg(A), cgb = cell_cg(B);
    const bool fix0 = (unl - cga < R(0)) && (unr - cgb > R(0));
    const bool fix2 = (unl + cga < R(0)) && (unr + cgb > R(0));
    const R mL = hl * unl, mR = hr * unr;
    const R dh = hr - hl;
    const R dhu = mR - mL;
    const R a0 = (l2 * dh - dhu) * rden;
    const R a2 = (dhu - l0 * dh) * rden;
    const R h1s = hl + a0, h3s = hr - a2;
    const bool guard = ((h1s < -k::toldry) && (h3s > k::toldry)) ||
                       ((h1s > k::toldry) && (h3s < -k::toldry));
    const R q0 = fabs(l0) * a0, q2 = fabs(l2) * a2;
    const R q02 = q0 + q2;
    const R dhv = hr * utr - hl * utl;
    const R S1 = mL * unl + mR * unr + (R(0.5) * k::g) * (hl * hl + hr * hr);
    const R S2 = mL * utl + mR * utr;
    F0 = R(0.5) * ((mL + mR) - q02);
    F1 = R(0.5) * (S1 - q0 * l0 - q2 * l2);
    F2 = R(0.5) * (S2 - vt * q02 - fabs(ut) * (dhv - vt * dh));
    return !(fix0 || fix2 || rough || guard);
}

// The face as the fast build's advance evaluates it, lane by lane: the
// one-pass form where it applies, roe_core elsewhere -- so the value of a
// face does not depend on how faces are grouped into wavefronts (the flux
// registers re-evaluate faces in another grouping).  Returns true when every
// lane took the one-pass form (o.F holds the fluxes, no source share);
// otherwise o is the merged core with o.src set as roe_core sets it.
template <typename R, int DIR>
__device__ __forceinline__ bool roe_core_adv(const RoeCell<R>& A, const RoeCell<R>& B, R ds,
                                             RoeCore<R>& o)
{
    const bool ok = roe_quick<R, DIR>(A, B, o.F0, o.F1, o.F2);
    if (wave_all(ok)) return true;
    RoeCore<R> g;
    roe_core<R, DIR>(A, B, ds, g);
    if (wave_any(ok)) {
        o.F0 = ok ? o.F0 : g.F0;
        o.F1 = ok ? o.F1 : g.F1;
        o.F2 = ok ? o.F2 : g.F2;
        o.b0 = ok ? R(0) : g.b0;
        o.b2 = ok ? R(0) : g.b2;
    } else {   // every lane general (a rough bed): no merge
        o.F0 = g.F0; o.F1 = g.F1; o.F2 = g.F2;
        o.b0 = g.b0; o.b2 = g.b2;
    }
    o.l0 = g.l0;
    o.l2 = g.l2;
    o.vt = g.vt;
    o.src = wave_any(o.b0 != R(0));
    return false;
}
#endif

// Upwind source split (Roe.H:281-289), without branches: each acoustic
// wave's share goes to S- when its speed is positive and to S+ when
// negative; the contact wave carries none.  Adding a selected 0 in place of
// a skipped term leaves every sum as the reference's.  Face-normal frame.
template <typename R>
__device__ __forceinline__ void src_split(const RoeCore<R>& f, R& Sm0, R& Sm1, R& Sm2, R& Sp0,
                                          R& Sp1, R& Sp2)
{
    const R k0b = f.b0 * f.l0, k0c = f.b0 * f.vt, k2b = f.b2 * f.l2, k2c = f.b2 * f.vt;
    const bool p0 = f.l0 > R(0), n0 = f.l0 < R(0), p2 = f.l2 > R(0), n2 = f.l2 < R(0);
    Sm0 = (p0 ? f.b0 : R(0)) + (p2 ? f.b2 : R(0));
    Sm1 = (p0 ? k0b : R(0)) + (p2 ? k2b : R(0));
    Sm2 = (p0 ? k0c : R(0)) + (p2 ? k2c : R(0));
    Sp0 = (n0 ? f.b0 : R(0)) + (n2 ? f.b2 : R(0));
    Sp1 = (n0 ? k0b : R(0)) + (n2 ? k2b : R(0));
    Sp2 = (n0 ? k0c : R(0)) + (n2 ? k2c : R(0));
}

#if HX_FAST
// Effective fluxes of a face that carries a source share (fast build), in
// the face-normal frame: D+ = F + S+ for the cell on the low side and
// D- = F - S- for the cell on the high side.  Where every lane of the
// wavefront has l0 < 0 < l2 (subcritical faces), S- is the third wave's
// share and S+ the first's, so the selects of src_split drop out; the
// products are rounded on their own, as in src_split, and the two forms
// agree up to the sign of a zero whatever the grouping into wavefronts.
template <typename R>
__device__ __forceinline__ void face_dpm(const RoeCore<R>& f, R& Dp0, R& Dp1, R& Dp2, R& Dm0,
                                         R& Dm1, R& Dm2)
{
    if (wave_all((f.l0 < R(0)) && (f.l2 > R(0)))) {
        const R k0b = f.b0 * f.l0, k0c = f.b0 * f.vt, k2b = f.b2 * f.l2, k2c = f.b2 * f.vt;
        Dp0 = f.F0 + f.b0; Dp1 = f.F1 + k0b; Dp2 = f.F2 + k0c;
        Dm0 = f.F0 - f.b2; Dm1 = f.F1 - k2b; Dm2 = f.F2 - k2c;
    } else {
        R Sm0, Sm1, Sm2, Sp0, Sp1, Sp2;
        src_split(f, Sm0, Sm1, Sm2, Sp0, Sp1, Sp2);
        Dp0 = f.F0 + Sp0; Dp1 = f.F1 + Sp1; Dp2 = f.F2 + Sp2;
        Dm0 = f.F0 - Sm0; Dm1 = f.F1 - Sm1; Dm2 = f.F2 - Sm2;
    }
}
#endif

// The face with its split sources, rotated back to the x-y frame.
template <typename R, int DIR>
__device__ __forceinline__ RoeFace<R> roe_face(const RoeCell<R>& A, const RoeCell<R>& B, R ds)
{
    RoeCore<R> f;
    roe_core<R, DIR>(A, B, ds, f);
    RoeFace<R> o;
    o.src = f.src;
    R Sm0 = R(0), Sm1 = R(0), Sm2 = R(0), Sp0 = R(0), Sp1 = R(0), Sp2 = R(0);
    if (f.src) src_split(f, Sm0, Sm1, Sm2, Sp0, Sp1, Sp2);
    o.F0 = f.F0; o.Sm0 = Sm0; o.Sp0 = Sp0;
    if (DIR == 0) {
        o.F1 = f.F1; o.F2 = f.F2; o.Sm1 = Sm1; o.Sm2 = Sm2; o.Sp1 = Sp1; o.Sp2 = Sp2;
    } else {   // rotate back with n = (0,1): (a, b) -> (-b, a)
        o.F1 = -f.F2; o.F2 = f.F1; o.Sm1 = -Sm2; o.Sm2 = Sm1; o.Sp1 = -Sp2; o.Sp2 = Sp1;
    }
    return o;
}

// The face with its split sources as the advance evaluates it (the one-pass
// form where it applies in the fast build; roe_face in the exact build).
template <typename R, int DIR>
__device__ __forceinline__ RoeFace<R> roe_face_adv(const RoeCell<R>& A, const RoeCell<R>& B, R ds)
{
#if HX_FAST
    RoeCore<R> f;
    const bool quick = roe_core_adv<R, DIR>(A, B, ds, f);
    RoeFace<R> o;
    o.src = !quick && f.src;
    R Sm0 = R(0), Sm1 = R(0), Sm2 = R(0), Sp0 = R(0), Sp1 = R(0), Sp2 = R(0);
    if (o.src) src_split(f, Sm0, Sm1, Sm2, Sp0, Sp1, Sp2);
    o.F0 = f.F0; o.Sm0 = Sm0; o.Sp0 = Sp0;
    if (DIR == 0) {
        o.F1 = f.F1; o.F2 = f.F2; o.Sm1 = Sm1; o.Sm2 = Sm2; o.Sp1 = Sp1; o.Sp2 = Sp2;
    } else {
        o.F1 = -f.F2; o.F2 = f.F1; o.Sm1 = -Sm2; o.Sm2 = Sm1; o.Sp1 = -Sp2; o.Sp2 = Sp1;
    }
    return o;
#else
    return roe_face<R, DIR>(A, B, ds);
#endif
}

// ---------------------------------------------------------------------------
// Hydrostatic reconstruction (src/solvers/HydrostaticRoe.cpp:173-396,
// include/solvers/kernels/HydrostaticRoe.H:43-159).
// ---------------------------------------------------------------------------

// van Leer slope (HydrostaticRoe.cpp:173-195); |r| on the floating value.
// The fast build uses the equivalent harmonic-mean form 2 dm dp / (dm + dp)
// for dm dp > 0 (one reciprocal instead of two quotients).
template <typename R>
__device__ __forceinline__ R vl_slope(R qm, R q, R qp)
{
    const R dm = q - qm;
    const R dp = qp - q;
    if (fabs(dp) <= Kc<R>::tol12) return R(0);
#if HX_FAST
    const R pr = dm * dp;
    return (pr > R(0)) ? R(2) * pr * rcp(dm + dp) : R(0);
#else
    const R r = dm / dp;
    return ((r + fabs(r)) / (R(1) + fabs(r))) * dp;
#endif
}

// Per-cell reconstruction inputs: eta = h_ext + z_ext, u and v guarded by
// h_ext > TOL12 (HydrostaticRoe.cpp:257-337), and the cell's limited slopes
// along the face normal (each computed once, used by both faces).
template <typename R>
struct HydroCell {
    R eta, u, v, z, se, su, sv;
};

template <typename R>
__device__ __forceinline__ void hydro_prims(R hr, R qx, R qy, R z, R& eta, R& u, R& v)
{
    eta = hr + z;
    const bool ok = hr > Kc<R>::tol12;
    const Div<R> ih(ok ? hr : R(1));
    u = ok ? ih(qx) : R(0);
    v = ok ? ih(qy) : R(0);
}

#if !HX_FAST
// Homogeneous Roe flux on face states (kernels/HydrostaticRoe.H:43-159),
// in the reference's operation order (exact build).
template <typename R, int DIR>
__device__ __forceinline__ void hroe(R hi, R hui, R hvi, R hj, R huj, R hvj, R& F0, R& F1, R& F2)
{
    using k = Kc<R>;
    hi = (hi >= k::toldry) ? hi : R(0);
    hj = (hj >= k::toldry) ? hj : R(0);
    const bool wl = hi > R(0), wr = hj > R(0);
    const R sql = wl ? sqrt(hi) : R(0);
    const R sqr = wr ? sqrt(hj) : R(0);
    const Div<R> il(wl ? hi : R(1)), ir(wr ? hj : R(1));
    const R ul = wl ? il(hui) : R(0);
    const R vl = wl ? il(hvi) : R(0);
    const R ur = wr ? ir(huj) : R(0);
    const R vr = wr ? ir(hvj) : R(0);
    const R unl = (DIR == 0) ? ul : vl;
    const R utl = (DIR == 0) ? vl : -ul;
    const R unr = (DIR == 0) ? ur : vr;
    const R utr = (DIR == 0) ? vr : -ur;
    const R hmean = R(0.5) * (hi + hj);
    const R ht = (hmean > k::toldry) ? hmean : k::toldry;
    const R ct = sqrt(k::g * ht);
    const Div<R> idsq(sql + sqr + k::tol12);
    const R ut = idsq(unl * sql + unr * sqr);
    const R vt = idsq(utl * sql + utr * sqr);
    R l0 = ut - ct;
    const R l1 = ut;
    R l2 = ut + ct;
    {
        const R cl = sqrt(k::g * hi), cr = sqrt(k::g * hj);
        R el = unl - cl, er = unr - cr;
        if (el < R(0) && er > R(0)) l0 = Div<R>(er - el)(el * (er - l0));
        el = unl + cl;
        er = unr + cr;
        if (el < R(0) && er > R(0)) l2 = Div<R>(er - el)(er * (l2 - el));
    }
    const Div<R> iden(R(2) * ct + k::tol12);
    const Div<R> ictd(ct);
    const R dh = hj - hi;
    const R dhu = hj * unr - hi * unl;
    const R dhv = hj * utr - hi * utl;
    const R a0 = iden(l2 * dh - dhu);
    const R a1 = ictd(-vt * dh + dhv);
    const R a2 = iden(-l0 * dh + dhu);
    R f0 = R(0.5) * (hi * unl + hj * unr);
    R f1 = R(0.5) * (hi * unl * unl + R(0.5) * k::g * hi * hi +
                     hj * unr * unr + R(0.5) * k::g * hj * hj);
    R f2 = R(0.5) * (hi * unl * utl + hj * unr * utr);
    const R w0 = R(0.5) * fabs(l0) * a0;
    const R w1 = R(0.5) * fabs(l1) * a1;
    const R w2 = R(0.5) * fabs(l2) * a2;
    f0 -= w0; f1 -= w0 * l0; f2 -= w0 * vt;
    f2 -= w1 * ct;
    f0 -= w2; f1 -= w2 * l2; f2 -= w2 * vt;
    F0 = f0;
    if (DIR == 0) { F1 = f1; F2 = f2; }
    else { F1 = -f2; F2 = f1; }
}
#endif

#if HX_FAST
// The homogeneous face from the face depths and velocities (fast build): a
// wet side's velocity is the MUSCL face velocity itself instead of the
// reference's (h u) / h, which equals it to within 2 ulp, so the face needs
// neither the face momenta nor the reciprocal depths.  sql, sqr: sqrt of the
// dry-cut depths; unl .. utr: the dry-cut velocities in the face frame.
template <typename R>
__device__ __forceinline__ void hroe_core(R hi, R hj, R sql, R sqr, R unl, R utl, R unr, R utr,
                                          R& f0, R& f1, R& f2)
{
    using k = Kc<R>;
    // g max((hi + hj)/2, TOLDRY): scaling by g and halving commute with the
    // max and with rounding
    R ct, ict;
    sqrt_rsq(fmax((hi + hj) * (R(0.5) * k::g), k::g * k::toldry), ct, ict);
    const R qh = R(0.5) * ict;
    const R rden = qh * fma(-k::tol12, qh, R(1));
    const Div<R> idsq(sql + sqr + k::tol12);
    const R ut = idsq(unl * sql + unr * sqr);
    const R vt = idsq(utl * sql + utr * sqr);
    R l0 = ut - ct;
    const R l1 = ut;
    R l2 = ut + ct;
    {
        const R cl = SqrtG<R>::v * sql, cr = SqrtG<R>::v * sqr;
        R el = unl - cl, er = unr - cr;
        if (el < R(0) && er > R(0)) l0 = Div<R>(er - el)(el * (er - l0));
        el = unl + cl;
        er = unr + cr;
        if (el < R(0) && er > R(0)) l2 = Div<R>(er - el)(er * (l2 - el));
    }
    // central flux minus the upwind wave sum with qk = |lk| ak and the
    // contact term |ut| a1 ct = |ut| (dhv - vt dh) (1/ct drops out)
    const R mL = hi * unl, mR = hj * unr;
    const R dh = hj - hi;
    const R dhu = mR - mL;
    const R dhv = hj * utr - hi * utl;
    const R a0 = (l2 * dh - dhu) * rden;
    const R a2 = (dhu - l0 * dh) * rden;
    const R q0 = fabs(l0) * a0, q2 = fabs(l2) * a2;
    const R q02 = q0 + q2;
    const R S1 = mL * unl + mR * unr + (R(0.5) * k::g) * (hi * hi + hj * hj);
    const R S2 = mL * utl + mR * utr;
    f0 = R(0.5) * ((mL + mR) - q02);
    f1 = R(0.5) * (S1 - q0 * l0 - q2 * l2);
    f2 = R(0.5) * (S2 - vt * q02 - fabs(l1) * (dhv - vt * dh));
}

template <typename R, int DIR>
__device__ __forceinline__ void hroe_uv(R hi, R ui, R vi, R hj, R uj, R vj, R& F0, R& F1, R& F2)
{
    using k = Kc<R>;
    const bool wl = hi >= k::toldry, wr = hj >= k::toldry;
    R f0, f1, f2;
    if (wave_all(wl && wr)) {
        // every face of the wavefront wet on both sides: no dry cuts
        R sl, rl, sr, rr;
        sqrt_rsq(hi, sl, rl);
        sqrt_rsq(hj, sr, rr);
        hroe_core<R>(hi, hj, sl, sr, (DIR == 0) ? ui : vi, (DIR == 0) ? vi : -ui,
                     (DIR == 0) ? uj : vj, (DIR == 0) ? vj : -uj, f0, f1, f2);
    } else {
        hi = wl ? hi : R(0);
        hj = wr ? hj : R(0);
        // max(h, TOLDRY) is h on a wet side and a positive finite argument
        // on a dry side (whose values are not used)
        R sl, rl, sr, rr;
        sqrt_rsq(fmax(hi, k::toldry), sl, rl);
        sqrt_rsq(fmax(hj, k::toldry), sr, rr);
        const R sql = wl ? sl : R(0);
        const R sqr = wr ? sr : R(0);
        const R ul = wl ? ui : R(0);
        const R vl = wl ? vi : R(0);
        const R ur = wr ? uj : R(0);
        const R vr = wr ? vj : R(0);
        hroe_core<R>(hi, hj, sql, sqr, (DIR == 0) ? ul : vl, (DIR == 0) ? vl : -ul,
                     (DIR == 0) ? ur : vr, (DIR == 0) ? vr : -ur, f0, f1, f2);
    }
    F0 = f0;
    if (DIR == 0) { F1 = f1; F2 = f2; }
    else { F1 = -f2; F2 = f1; }
}
#endif

// One face: NODATA wall, face bed, MUSCL face states, Audusse depths,
// homogeneous Roe, hydrostatic pressure correction (HydrostaticRoe.cpp:234-396).
// Only the normal component of FL/FR differs from the Roe flux, by pL/pR.
// The MUSCL states of a side (eta, u, v extrapolated half a cell towards the
// face) come in with its bed: the x face takes the left side's from the lane
// that owns that cell (4 values instead of its whole reconstruction).
template <typename R>
struct HSide {
    R z, eta, u, v;
};

template <typename R>
__device__ __forceinline__ HSide<R> side_low(const HydroCell<R>& c)    // seen from the face above
{
    return HSide<R>{c.z, c.eta + R(0.5) * c.se, c.u + R(0.5) * c.su, c.v + R(0.5) * c.sv};
}
template <typename R>
__device__ __forceinline__ HSide<R> side_high(const HydroCell<R>& c)   // seen from the face below
{
    return HSide<R>{c.z, c.eta - R(0.5) * c.se, c.u - R(0.5) * c.su, c.v - R(0.5) * c.sv};
}

template <typename R, int DIR>
__device__ __forceinline__ void hydro_face_st(const HSide<R>& L, const HSide<R>& Rc,
                                              R& F0, R& F1, R& F2, R& pL, R& pR)
{
    using k = Kc<R>;
    // a NODATA side makes the face a wall with no flux and no pressure: the
    // face is evaluated on every lane and those lanes take zeros at the end
    // (a wavefront without NODATA cells only votes)
    const bool wall = fabs(L.z + R(9999)) < k::tol10 || fabs(Rc.z + R(9999)) < k::tol10;
#if HX_FAST
    const R zf = fmax(L.z, Rc.z);   // terrain is never NaN
#else
    const R zf = smax(L.z, Rc.z);
#endif
    const R etaLf = L.eta, etaRf = Rc.eta;
    const R uLf = L.u, uRf = Rc.u;
    const R vLf = L.v, vRf = Rc.v;
    const R hLu = smax(R(0), etaLf - L.z);
    const R hRu = smax(R(0), etaRf - Rc.z);
    const R hLf = smax(R(0), etaLf - zf);
    const R hRf = smax(R(0), etaRf - zf);
#if HX_FAST
    hroe_uv<R, DIR>(hLf, uLf, vLf, hRf, uRf, vRf, F0, F1, F2);
#else
    const R huLf = (hLf > R(0)) ? hLf * uLf : R(0);
    const R hvLf = (hLf > R(0)) ? hLf * vLf : R(0);
    const R huRf = (hRf > R(0)) ? hRf * uRf : R(0);
    const R hvRf = (hRf > R(0)) ? hRf * vRf : R(0);
    hroe<R, DIR>(hLf, huLf, hvLf, hRf, huRf, hvRf, F0, F1, F2);
#endif
    pL = R(0.5) * k::g * (hLu * hLu - hLf * hLf);
    pR = R(0.5) * k::g * (hRu * hRu - hRf * hRf);
    if (__builtin_expect(wave_any(wall), 0)) {
        F0 = wall ? R(0) : F0;
        F1 = wall ? R(0) : F1;
        F2 = wall ? R(0) : F2;
        pL = wall ? R(0) : pL;
        pR = wall ? R(0) : pR;
    }
}

template <typename R, int DIR>
__device__ __forceinline__ void hydro_face(const HydroCell<R>& L, const HydroCell<R>& Rc,
                                           R& F0, R& F1, R& F2, R& pL, R& pR)
{
    hydro_face_st<R, DIR>(side_low(L), side_high(Rc), F0, F1, F2, pL, pR);
}

// ---------------------------------------------------------------------------
// CFL accumulation of one cell (src/solvers/Roe.cpp:38-74 and
// HydrostaticRoe.cpp:37-63, identical cell functions).  The maximum of
// a = (|u|+c)/dx + (|v|+c)/dy is kept; see slot_dt_est.
// ---------------------------------------------------------------------------
template <typename R>
struct CflAcc {
    R amax, vmax;
    unsigned flags;
    __device__ __forceinline__ void init() { amax = R(0); vmax = R(0); flags = 0u; }
    __device__ __forceinline__ void add(R h, R qx, R qy, R z, R dx, R dy, R rdx, R rdy, bool mul)
    {
        if (fabs(z - R(-9999)) < Kc<R>::tol10) { flags |= CFL_NODATA; return; }
        if (h > R(1e-12)) {
#if HX_FAST
            R s, r;
            sqrt_rsq(h, s, r);
            const R ih = r * r;
            const R u = qx * ih, v = qy * ih;
            const R c = SqrtG<R>::v * s;
#else
            const Div<R> ih(h);
            const R u = ih(qx), v = ih(qy);
            const R c = sqrt(Kc<R>::g * h);
#endif
            const R sx = fabs(u) + c, sy = fabs(v) + c;
            const R a = mul ? (sx * rdx + sy * rdy) : (sx / dx + sy / dy);
            // v_max_f64: the same value as the select for these non-negative
            // operands (a NaN operand leaves the other, as the select does)
            amax = fmax(amax, a);
            vmax = fmax(vmax, fmax(fabs(u), fabs(v)));
            flags |= CFL_WET;
        } else {
            flags |= CFL_DRY;
        }
    }
};

__device__ __forceinline__ unsigned long long dbits(double x)
{
    return (unsigned long long)__double_as_longlong(x);
}

// Sub-slot of the CU that runs the calling wave: (XCC_ID bits 3:0) x 8 +
// (HW_ID CU_ID bits 11:8) mod 8.  Each shader engine
// numbers its CUs with 8 consecutive ids, so on 256 CUs every one of 64
// sub-slots serves 4 CUs.  Read where it is used, it holds no SGPR over the
// march, as blockIdx.x would.
__device__ __forceinline__ unsigned slot_sub_here()
{
    constexpr int kHwIdReg = 4, kCuIdOffset = 8, kCuIdBits = 4;
    constexpr int kXccIdReg = 20, kXccIdBits = 4;
    static_assert(kCflSubs == 64, "sub-slot index: 8 XCDs x 8 CU ids");
    const unsigned cu = (unsigned)__builtin_amdgcn_s_getreg(((kCuIdBits - 1) << 11) |
                                                            (kCuIdOffset << 6) | kHwIdReg);
    const unsigned x = (unsigned)__builtin_amdgcn_s_getreg(((kXccIdBits - 1) << 11) | kXccIdReg);
    return ((x & 7u) << 3) | (cu & 7u);
}

// A slot's value read by a whole wavefront (every lane active): lane l loads
// a and the flags of sub-slot l with one 16-byte load (and the velocity when
// want_v, a wave-uniform condition: it only feeds the history one wavefront
// writes); four DPP steps fold each row of 16 lanes (quad_perm [1,0,3,2]:
// lane ^ 1, quad_perm [2,3,0,1]: lane ^ 2, row_half_mirror: i <-> 7 - i,
// row_mirror: i <-> 15 - i) and two lane exchanges the four rows; a and the
// flags are made wave-uniform, the velocity is left in every lane.  With
// spread false (wave-uniform) every lane loads sub-slot 0 and nothing is
// folded.  Scalar loads of the sub-slots would hold SGPRs the advance
// kernels have not got.
template <int CTRL>
__device__ __forceinline__ unsigned long long dpp64(unsigned long long x)
{
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)x, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(x >> 32), CTRL, 0xf, 0xf, false);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}
template <int CTRL>
__device__ __forceinline__ void fold_dpp(unsigned long long& x)
{
    const unsigned long long o = dpp64<CTRL>(x);
    x = o > x ? o : x;
}
template <int CTRL>
__device__ __forceinline__ void fold_dpp(unsigned& f)
{
    f |= (unsigned)__builtin_amdgcn_mov_dpp((int)f, CTRL, 0xf, 0xf, false);
}
template <typename T>
__device__ __forceinline__ void fold_wave(T& x)
{
    fold_dpp<0xB1>(x);
    fold_dpp<0x4E>(x);
    fold_dpp<0x141>(x);
    fold_dpp<0x140>(x);
#pragma unroll
    for (int off = 16; off < 64; off <<= 1) {
        if constexpr (sizeof(T) == 8) {
            const T o = __shfl_xor(x, off, 64);
            x = o > x ? o : x;
        } else {
            x |= __shfl_xor(x, off, 64);
        }
    }
}
__device__ __forceinline__ CflValue slot_value_wave(const CflSlot* s, bool want_v, bool spread = true)
{
    const CflSub& q = s->sub[spread ? (threadIdx.x & (kCflSubs - 1)) : 0u];
    unsigned long long a = q.amax;
    unsigned f = q.flags;
    unsigned long long v = want_v ? q.vmax : 0ull;
    if (spread) {
        if (want_v) fold_wave(v);
        fold_wave(a);
        fold_wave(f);
    }
    CflValue r;
    r.amax = ((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(a >> 32)) << 32) |
             (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)a);
    r.vmax = v;
    r.flags = (unsigned)__builtin_amdgcn_readfirstlane((int)f);
    return r;
}

// Fold one partial (max a, max velocity, flags) into sub-slot k.  The
// sub-slot only grows while partials fold into it, so it is read at device
// scope first and the atomics that cannot change it are skipped:
// same-address device-scope atomics serialize, and after the first partials
// nearly all skip.
__device__ __forceinline__ void slot_fold(CflSlot* slot, unsigned k, double a, double v, unsigned f)
{
    CflSub& q = slot->sub[k];
    const unsigned long long ca = __hip_atomic_load(&q.amax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long cv = __hip_atomic_load(&q.vmax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned cf = __hip_atomic_load(&q.flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (dbits(a) > ca) atomicMax(&q.amax, dbits(a));
    if (dbits(v) > cv) atomicMax(&q.vmax, dbits(v));
    if ((cf | f) != cf) atomicOr(&q.flags, f);
}

// Reduce a wavefront's accumulator and fold one block into its sub-slot of
// the slot (sub-slot 0 when not spread).  Every thread of the block must
// call this.
template <typename R>
__device__ __forceinline__ void cfl_commit(CflAcc<R> acc, CflSlot* slot, bool spread = true)
{
    double a = (double)acc.amax, v = (double)acc.vmax;
    unsigned f = acc.flags;
    for (int off = 32; off > 0; off >>= 1) {
        const double ao = __shfl_xor(a, off, 64);
        const double vo = __shfl_xor(v, off, 64);
        const unsigned fo = __shfl_xor(f, off, 64);
        a = (a < ao) ? ao : a;
        v = (v < vo) ? vo : v;
        f |= fo;
    }
    __shared__ double sa[16], sv[16];
    __shared__ unsigned sf[16];
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) { sa[w] = a; sv[w] = v; sf[w] = f; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < nw; ++q) {
            a = (a < sa[q]) ? sa[q] : a;
            v = (v < sv[q]) ? sv[q] : v;
            f |= sf[q];
        }
        if (f) slot_fold(slot, spread ? slot_sub_here() : 0u, a, v, f);
    }
}

}  // namespace HX_NS
