// hx_device.hpp -- device-side building blocks of the fused per-level advance
// for CDNA4 (gfx950): halo-aware cell loads, per-cell hoisted quantities, the
// plain Roe face with split sources, the hydrostatic-reconstruction face, and
// the CFL accumulator.
//
// Arithmetic follows the reference operation by operation so that a build
// with -ffp-contract=off reproduces the FP64 CPU restatement bit for bit
// (friction's cbrt aside).  Where this file departs from the literal
// reference expression it is only by
//   * hoisting a per-cell value that the reference recomputes per face
//     (same operands, same operations, same rounding),
//   * dropping a term that the reference multiplies by an exact 0.0 or 1.0
//     (the face-normal rotation with axis normals, the zero contact-wave
//     source), which can change at most the sign of a zero,
//   * evaluating the positivity-guard timing test only when its depth test
//     passes (no side effects, so the outcome is unchanged).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <cstring>

namespace hx {

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

// Device view of one box (hx_patch with a typed element).
template <typename R>
struct PatchView {
    const R* h;
    const R* hu;
    const R* hv;
    const R* eta;   // may be null: eta = h + z
    const R* z;
    const R* n;
    int64_t pitch;
    int nx, ny;
    int bl0, bl1, bh0, bh1;   // per side: 0 reflect, 1 transmit, -1 stored ghosts
};

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
    c.sq = wet ? sqrt(hd) : R(0);
    c.u = wet ? qx / hd : R(0);
    c.v = wet ? qy / hd : R(0);
    c.cg = sqrt(Kc<R>::g * hd);
    c.ed = hd + z;
    c.z = z;
    c.n = n;
    return c;
}

template <typename R>
struct RoeFace {
    R F0, F1, F2, Sm0, Sm1, Sm2, Sp0, Sp1, Sp2;
};

// DIR 0: x face, normal (1,0); DIR 1: y face, normal (0,1).  A is the
// low-side cell, B the high-side cell, ds the face-normal cell size.
template <typename R, int DIR>
__device__ __forceinline__ RoeFace<R> roe_face(const RoeCell<R>& A, const RoeCell<R>& B, R ds)
{
    using k = Kc<R>;
    const R hl = A.hd, hr = B.hd;
    const R unl = (DIR == 0) ? A.u : A.v;
    const R utl = (DIR == 0) ? A.v : -A.u;
    const R unr = (DIR == 0) ? B.u : B.v;
    const R utr = (DIR == 0) ? B.v : -B.u;

    const R hmean = R(0.5) * (hl + hr);
    const R ht = (hmean > k::toldry) ? hmean : k::toldry;
    const R ct = sqrt(k::g * ht);
    const R dsq = A.sq + B.sq + k::tol12;
    const R ut = (unl * A.sq + unr * B.sq) / dsq;
    const R vt = (utl * A.sq + utr * B.sq) / dsq;

    R l0 = ut - ct;
    const R l1 = ut;
    R l2 = ut + ct;
    {
        R el = unl - A.cg, er = unr - B.cg;
        if (el < R(0) && er > R(0)) l0 = el * (er - l0) / (er - el);
        el = unl + A.cg;
        er = unr + B.cg;
        if (el < R(0) && er > R(0)) l2 = er * (l2 - el) / (er - el);
    }

    const R den = R(2) * ct + k::tol12;
    const R dh = hr - hl;
    const R dhu = hr * unr - hl * unl;
    const R dhv = hr * utr - hl * utl;
    const R a0 = (l2 * dh - dhu) / den;
    const R a1 = (-vt * dh + dhv) / ct;
    const R a2 = (-l0 * dh + dhu) / den;

    // Bed thrust (Roe.H:151-176) and Manning friction (:182-188).  With
    // zero roughness the reference's tau is a signed zero, so it is skipped.
    const R dz = B.z - A.z;
    const R dd = B.ed - A.ed;
    const R thrust_a = -k::g * ht * dz;
    R dzt;
    if ((dz >= R(0)) && (A.ed < B.z)) dzt = hl;
    else if ((dz < R(0)) && (B.ed < A.z)) dzt = hr;
    else dzt = dz;
    const R hup = (dz >= R(0)) ? hl : hr;
    const R thrust_b = -k::g * (hup - R(0.5) * fabs(dzt)) * dzt;
    const R thrust = ((dz * dd >= R(0)) && (ut * dz > R(0))) ? smax(thrust_a, thrust_b) : thrust_b;
    const R nt = R(0.5) * (A.n + B.n);
    R tau = R(0);
    if (nt != R(0)) {
        const R sf = (nt * nt * sqrt(ut * ut + vt * vt) * ut) / smax(k::toldry, ht * cbrt(ht));
        tau = k::g * ht * sf * ds;
    }
    R b0 = (thrust - tau) / den;
    R b2 = -b0;

    {   // positivity guard, Roe.H:200-247
        const R l0s = (fabs(l0) > k::tol12) ? l0 : (l0 >= R(0) ? k::tol12 : -k::tol12);
        const R l2s = (fabs(l2) > k::tol12) ? l2 : (l2 >= R(0) ? k::tol12 : -k::tol12);
        const R h1s = hl + a0 - (b0 / l0s);
        const R h3s = hr - a2 + (b2 / l2s);
        const bool c1 = (h1s < -k::toldry) && (h3s > k::toldry);
        const bool c3 = (h1s > k::toldry) && (h3s < -k::toldry);
        if (c1 || c3) {
            R mlam = fabs(l0);
            mlam = smax(mlam, fabs(l1));
            mlam = smax(mlam, fabs(l2));
            const R loc_dt = ds / (mlam + k::tol12);
            const R m1 = -(hl + a0) * fabs(l0);
            const R m3 = -(hr - a2) * l2;
            if (c1) {
                const R d1 = hl - h1s;
                const R d1s = (fabs(d1) > k::tol12) ? d1 : (d1 >= R(0) ? k::tol12 : -k::tol12);
                const R dt1 = (ds * R(0.5) / l0s) * (hl / d1s);
                if (dt1 < loc_dt) {
                    if (-m1 >= m3) b0 = m1;
                    b2 = -b0;
                }
            } else {
                const R d3 = hr - h3s;
                const R d3s = (fabs(d3) > k::tol12) ? d3 : (d3 >= R(0) ? k::tol12 : -k::tol12);
                const R dt3 = (ds * R(0.5) / l2s) * (hr / d3s);
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

    // Upwind source split (Roe.H:281-289); the contact wave carries none.
    R Sm0 = R(0), Sm1 = R(0), Sm2 = R(0), Sp0 = R(0), Sp1 = R(0), Sp2 = R(0);
    if (l0 > R(0)) { Sm0 += b0; Sm1 += b0 * l0; Sm2 += b0 * vt; }
    else if (l0 < R(0)) { Sp0 += b0; Sp1 += b0 * l0; Sp2 += b0 * vt; }
    if (l2 > R(0)) { Sm0 += b2; Sm1 += b2 * l2; Sm2 += b2 * vt; }
    else if (l2 < R(0)) { Sp0 += b2; Sp1 += b2 * l2; Sp2 += b2 * vt; }

    RoeFace<R> o;
    o.F0 = F0; o.Sm0 = Sm0; o.Sp0 = Sp0;
    if (DIR == 0) {
        o.F1 = F1; o.F2 = F2; o.Sm1 = Sm1; o.Sm2 = Sm2; o.Sp1 = Sp1; o.Sp2 = Sp2;
    } else {   // rotate back with n = (0,1): (a, b) -> (-b, a)
        o.F1 = -F2; o.F2 = F1; o.Sm1 = -Sm2; o.Sm2 = Sm1; o.Sp1 = -Sp2; o.Sp2 = Sp1;
    }
    return o;
}

// ---------------------------------------------------------------------------
// Hydrostatic reconstruction (src/solvers/HydrostaticRoe.cpp:173-396,
// include/solvers/kernels/HydrostaticRoe.H:43-159).
// ---------------------------------------------------------------------------

// van Leer slope (HydrostaticRoe.cpp:173-195); |r| on the floating value.
template <typename R>
__device__ __forceinline__ R vl_slope(R qm, R q, R qp)
{
    const R dm = q - qm;
    const R dp = qp - q;
    if (fabs(dp) <= Kc<R>::tol12) return R(0);
    const R r = dm / dp;
    return ((r + fabs(r)) / (R(1) + fabs(r))) * dp;
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
    u = ok ? qx / hr : R(0);
    v = ok ? qy / hr : R(0);
}

// Homogeneous Roe flux on face states (kernels/HydrostaticRoe.H:43-159).
template <typename R, int DIR>
__device__ __forceinline__ void hroe(R hi, R hui, R hvi, R hj, R huj, R hvj, R& F0, R& F1, R& F2)
{
    using k = Kc<R>;
    hi = (hi >= k::toldry) ? hi : R(0);
    hj = (hj >= k::toldry) ? hj : R(0);
    const bool wl = hi > R(0), wr = hj > R(0);
    const R sql = wl ? sqrt(hi) : R(0);
    const R sqr = wr ? sqrt(hj) : R(0);
    const R ul = wl ? hui / hi : R(0);
    const R vl = wl ? hvi / hi : R(0);
    const R ur = wr ? huj / hj : R(0);
    const R vr = wr ? hvj / hj : R(0);
    const R unl = (DIR == 0) ? ul : vl;
    const R utl = (DIR == 0) ? vl : -ul;
    const R unr = (DIR == 0) ? ur : vr;
    const R utr = (DIR == 0) ? vr : -ur;
    const R hmean = R(0.5) * (hi + hj);
    const R ht = (hmean > k::toldry) ? hmean : k::toldry;
    const R ct = sqrt(k::g * ht);
    const R dsq = sql + sqr + k::tol12;
    const R ut = (unl * sql + unr * sqr) / dsq;
    const R vt = (utl * sql + utr * sqr) / dsq;
    R l0 = ut - ct;
    const R l1 = ut;
    R l2 = ut + ct;
    {
        const R cl = sqrt(k::g * hi), cr = sqrt(k::g * hj);
        R el = unl - cl, er = unr - cr;
        if (el < R(0) && er > R(0)) l0 = el * (er - l0) / (er - el);
        el = unl + cl;
        er = unr + cr;
        if (el < R(0) && er > R(0)) l2 = er * (l2 - el) / (er - el);
    }
    const R den = R(2) * ct + k::tol12;
    const R dh = hj - hi;
    const R dhu = hj * unr - hi * unl;
    const R dhv = hj * utr - hi * utl;
    const R a0 = (l2 * dh - dhu) / den;
    const R a1 = (-vt * dh + dhv) / ct;
    const R a2 = (-l0 * dh + dhu) / den;
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

// One face: NODATA wall, face bed, MUSCL face states, Audusse depths,
// homogeneous Roe, hydrostatic pressure correction (HydrostaticRoe.cpp:234-396).
// Only the normal component of FL/FR differs from the Roe flux.
template <typename R, int DIR>
__device__ __forceinline__ void hydro_face(const HydroCell<R>& L, const HydroCell<R>& Rc,
                                           R& F0, R& F1, R& F2, R& pL, R& pR)
{
    using k = Kc<R>;
    if (fabs(L.z + R(9999)) < k::tol10 || fabs(Rc.z + R(9999)) < k::tol10) {
        F0 = R(0); F1 = R(0); F2 = R(0); pL = R(0); pR = R(0);
        return;
    }
    const R zf = smax(L.z, Rc.z);
    const R etaLf = L.eta + R(0.5) * L.se;
    const R etaRf = Rc.eta - R(0.5) * Rc.se;
    const R uLf = L.u + R(0.5) * L.su;
    const R uRf = Rc.u - R(0.5) * Rc.su;
    const R vLf = L.v + R(0.5) * L.sv;
    const R vRf = Rc.v - R(0.5) * Rc.sv;
    const R hLu = smax(R(0), etaLf - L.z);
    const R hRu = smax(R(0), etaRf - Rc.z);
    const R hLf = smax(R(0), etaLf - zf);
    const R hRf = smax(R(0), etaRf - zf);
    const R huLf = (hLf > R(0)) ? hLf * uLf : R(0);
    const R hvLf = (hLf > R(0)) ? hLf * vLf : R(0);
    const R huRf = (hRf > R(0)) ? hRf * uRf : R(0);
    const R hvRf = (hRf > R(0)) ? hRf * vRf : R(0);
    hroe<R, DIR>(hLf, huLf, hvLf, hRf, huRf, hvRf, F0, F1, F2);
    pL = R(0.5) * k::g * (hLu * hLu - hLf * hLf);
    pR = R(0.5) * k::g * (hRu * hRu - hRf * hRf);
}

// ---------------------------------------------------------------------------
// CFL accumulation of one cell (src/solvers/Roe.cpp:38-74 and
// HydrostaticRoe.cpp:37-63, identical cell functions).  Instead of
// min over cells of 1/a, the maximum a is kept: correctly rounded division
// is monotone, so fl(1/max a) = min fl(1/a) exactly.  Dry cells contribute
// the 1e6 and NODATA cells the 1e30 of the reference through flags.
// ---------------------------------------------------------------------------
enum : unsigned { CFL_WET = 1u, CFL_DRY = 2u, CFL_NODATA = 4u };

struct CflSlot {
    unsigned long long amax;   // bits of a non-negative double
    unsigned long long vmax;   // bits of a non-negative double
    unsigned int flags;
    unsigned int pad;
};

template <typename R>
struct CflAcc {
    R amax, vmax;
    unsigned flags;
    __device__ __forceinline__ void init() { amax = R(0); vmax = R(0); flags = 0u; }
    __device__ __forceinline__ void add(R h, R qx, R qy, R z, R dx, R dy)
    {
        if (fabs(z - R(-9999)) < Kc<R>::tol10) { flags |= CFL_NODATA; return; }
        if (h > R(1e-12)) {
            const R u = qx / h, v = qy / h;
            const R c = sqrt(Kc<R>::g * h);
            const R sx = fabs(u) + c, sy = fabs(v) + c;
            const R a = sx / dx + sy / dy;
            amax = smax(amax, a);
            vmax = smax(vmax, smax(fabs(u), fabs(v)));
            flags |= CFL_WET;
        } else {
            flags |= CFL_DRY;
        }
    }
};

// dt estimate (before cfl) from a reduced slot.
template <typename R>
__host__ __device__ inline R slot_dt_est(unsigned long long abits, unsigned flags)
{
    double a;
    memcpy(&a, &abits, sizeof(a));
    R est = (sizeof(R) == 8) ? R(1.7976931348623157e308) : R(3.40282347e38f);
    if (flags & CFL_WET) est = R(1) / R(a);
    if ((flags & CFL_DRY) && R(1.e6) < est) est = R(1.e6);
    if ((flags & CFL_NODATA) && R(1.e30) < est) est = R(1.e30);
    return est;
}

__device__ __forceinline__ unsigned long long dbits(double x)
{
    return (unsigned long long)__double_as_longlong(x);
}

// Reduce a wavefront's accumulator and fold one block into the slot with
// one atomic per quantity.  Every thread of the block must call this.
template <typename R>
__device__ __forceinline__ void cfl_commit(CflAcc<R> acc, CflSlot* slot)
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
        if (f) {
            atomicMax(&slot->amax, dbits(a));
            atomicMax(&slot->vmax, dbits(v));
            atomicOr(&slot->flags, f);
        }
    }
}

}  // namespace hx
