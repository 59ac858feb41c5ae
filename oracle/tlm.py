"""Tangent-linear reference for the adjoint: a vectorised numpy restatement of
the plain Roe and hydrostatic-reconstruction steps, generic over the element
type, differentiated by the complex step.

TEST INFRASTRUCTURE ONLY (the gradient checks use it as the checker).

The same code runs on float64 (then it reproduces the C restatement in
``hx_oracle.c`` step for step) and on complex128: with every input x
replaced by x + i * eps * dx (eps = 1e-30), the imaginary part of each output
divided by eps is the directional derivative d out / d x . dx to machine
precision, with no subtraction error (Squire and Trapp, SIAM Review 40(1),
1998).  Branches, comparisons, max/min and |x| act on real parts, so the
derivative is that of the branch the forward step takes; the conventions at
kinks are sign(0) = +1 for |x| (|x| is x when x >= 0) and, for max(a, b) =
(a < b) ? b : a, the derivative of the operand selected.

Citations follow the C restatement: src/solvers/Roe.cpp:250-434,
include/solvers/kernels/Roe.H:50-309, src/solvers/HydrostaticRoe.cpp:173-606,
include/solvers/kernels/HydrostaticRoe.H:43-159,
include/boundaries/BCFill.H:49-132.
"""
from __future__ import annotations

import numpy as np

GRAV = 9.81
TOLDRY = 1e-6
TOL12 = 1e-12
TOL10 = 1e-10
G = 2
EPS = 1e-30


def _re(x):
    return np.real(x)


def _where(c, a, b):
    return np.where(c, a, b)


def _fabs(x):
    return _where(_re(x) >= 0.0, x, -x)


def _smax(a, b):          # std::max(a, b) = (a < b) ? b : a
    return _where(_re(a) < _re(b), b, a)


def _sqrt(x):
    return np.sqrt(x)


def _cbrt(x):
    if np.iscomplexobj(x):
        return np.power(x, 1.0 / 3.0)
    return np.cbrt(x)


def _safe(x, cond, fill=1.0):
    """x where cond holds, fill elsewhere (keeps untaken branches finite)."""
    return _where(cond, x, fill)


# -- ghost fill and rebuild (hx_oracle.c fill_ghosts / build_work) -----------

def _fill(a, kind, bclo, bchi):
    """a: padded (ny + 4, nx + 4); kind 0 scalar, 1 normal-x, 2 normal-y,
    3 foextrap.  x ghosts of valid rows first, then y ghosts of all columns."""
    ny, nx = a.shape[0] - 2 * G, a.shape[1] - 2 * G
    rows = slice(G, G + ny)
    for g in (1, 2):
        refl = kind != 3 and bclo[0] == 0
        src = (g - 1) if refl else 0
        sgn = -1.0 if (refl and kind == 1) else 1.0
        a[rows, G - g] = sgn * a[rows, G + src]
        refl = kind != 3 and bchi[0] == 0
        src = (nx - g) if refl else (nx - 1)
        sgn = -1.0 if (refl and kind == 1) else 1.0
        a[rows, G + nx - 1 + g] = sgn * a[rows, G + src]
    for g in (1, 2):
        refl = kind != 3 and bclo[1] == 0
        src = (g - 1) if refl else 0
        sgn = -1.0 if (refl and kind == 2) else 1.0
        a[G - g, :] = sgn * a[G + src, :]
        refl = kind != 3 and bchi[1] == 0
        src = (ny - g) if refl else (ny - 1)
        sgn = -1.0 if (refl and kind == 2) else 1.0
        a[G + ny - 1 + g, :] = sgn * a[G + src, :]
    return a


def _work(h, hu, hv, z, n, bclo, bchi):
    ny, nx = h.shape
    dt = np.result_type(h, hu, hv, z, n)
    P = lambda: np.zeros((ny + 2 * G, nx + 2 * G), dtype=dt)
    eta, pu, pv, pz, pn = P(), P(), P(), P(), P()
    v = (slice(G, G + ny), slice(G, G + nx))
    eta[v] = h + z
    pu[v], pv[v], pz[v], pn[v] = hu, hv, z, n
    _fill(eta, 0, bclo, bchi)
    _fill(pu, 1, bclo, bchi)
    _fill(pv, 2, bclo, bchi)
    _fill(pz, 3, bclo, bchi)
    _fill(pn, 3, bclo, bchi)
    he = _where(_re(eta - pz) > 0.0, eta - pz, 0.0 * eta)     # smax(0, eta - z)
    return he, pu, pv, pz, pn


# -- shared Roe core (Roe.H:68-142 / HydrostaticRoe.H:59-150) -----------------

def _roe_core(hl, hul, hvl, hr, hur, hvr, nx, ny):
    hl = _where(_re(hl) >= TOLDRY, hl, 0.0 * hl)
    hr = _where(_re(hr) >= TOLDRY, hr, 0.0 * hr)
    wl, wr = _re(hl) > 0.0, _re(hr) > 0.0
    sql = _where(wl, _sqrt(_safe(hl, wl)), 0.0 * hl)
    sqr = _where(wr, _sqrt(_safe(hr, wr)), 0.0 * hr)
    ul = _where(wl, hul / _safe(hl, wl), 0.0 * hl)
    vl = _where(wl, hvl / _safe(hl, wl), 0.0 * hl)
    ur = _where(wr, hur / _safe(hr, wr), 0.0 * hr)
    vr = _where(wr, hvr / _safe(hr, wr), 0.0 * hr)
    unl = ul * nx + vl * ny
    utl = -ul * ny + vl * nx
    unr = ur * nx + vr * ny
    utr = -ur * ny + vr * nx
    hmean = 0.5 * (hl + hr)
    ht = _where(_re(hmean) > TOLDRY, hmean, TOLDRY + 0.0 * hmean)
    ct = _sqrt(GRAV * ht)
    dsq = sql + sqr + TOL12
    ut = (unl * sql + unr * sqr) / dsq
    vt = (utl * sql + utr * sqr) / dsq
    lam0 = ut - ct
    lam1 = ut
    lam2 = ut + ct
    el = unl - _sqrt(GRAV * hl)
    er = unr - _sqrt(GRAV * hr)
    fix = (_re(el) < 0.0) & (_re(er) > 0.0)
    d = _safe(er - el, fix)
    lam0 = _where(fix, el * (er - lam0) / d, lam0)
    el = unl + _sqrt(GRAV * hl)
    er = unr + _sqrt(GRAV * hr)
    fix = (_re(el) < 0.0) & (_re(er) > 0.0)
    d = _safe(er - el, fix)
    lam2 = _where(fix, er * (lam2 - el) / d, lam2)
    den = 2.0 * ct + TOL12
    dh = hr - hl
    dhu = hr * unr - hl * unl
    dhv = hr * utr - hl * utl
    a0 = (lam2 * dh - dhu) / den
    a1 = (-vt * dh + dhv) / ct
    a2 = (-lam0 * dh + dhu) / den
    F0 = 0.5 * (hl * unl + hr * unr)
    F1 = 0.5 * (hl * unl * unl + 0.5 * GRAV * hl * hl + hr * unr * unr + 0.5 * GRAV * hr * hr)
    F2 = 0.5 * (hl * unl * utl + hr * unr * utr)
    lam = (lam0, lam1, lam2)
    alp = (a0, a1, a2)
    for k in range(3):
        e0 = 0.0 if k == 1 else 1.0
        e1 = lam0 if k == 0 else (0.0 if k == 1 else lam2)
        e2 = ct if k == 1 else vt
        w = 0.5 * _fabs(lam[k]) * alp[k]
        F0 = F0 - w * e0
        F1 = F1 - w * e1
        F2 = F2 - w * e2
    return dict(hl=hl, hr=hr, ut=ut, vt=vt, ct=ct, ht=ht, den=den, lam=lam, alp=alp,
                F=[F0, F1, F2])


def _rot(a1, a2, nx, ny):
    return a1 * nx - a2 * ny, a1 * ny + a2 * nx


def roe_faces(L, R, ds, nx, ny):
    """Plain Roe faces, vectorised (Roe.H:50-309).  L, R: tuples of arrays
    (h, hu, hv, z, n).  Returns F, Sm, Sp (lists of 3 arrays)."""
    c = _roe_core(L[0], L[1], L[2], R[0], R[1], R[2], nx, ny)
    hl, hr, ut, vt, ct, ht, den = c["hl"], c["hr"], c["ut"], c["vt"], c["ct"], c["ht"], c["den"]
    lam0, lam1, lam2 = c["lam"]
    a0, a1, a2 = c["alp"]
    zl, zr = L[3], R[3]
    dz = zr - zl
    etl = hl + zl
    etr = hr + zr
    dd = etr - etl
    thrust_a = -GRAV * ht * dz
    b1c = (_re(dz) >= 0.0) & (_re(etl) < _re(zr))
    b2c = (_re(dz) < 0.0) & (_re(etr) < _re(zl))
    dzt = _where(b1c, hl, _where(b2c, hr, dz))
    hup = _where(_re(dz) >= 0.0, hl, hr)
    thrust_b = -GRAV * (hup - 0.5 * _fabs(dzt)) * dzt
    sel = (_re(dz * dd) >= 0.0) & (_re(ut * dz) > 0.0)
    thrust = _where(sel, _smax(thrust_a, thrust_b), thrust_b)
    nt = 0.5 * (L[4] + R[4])
    mag = _sqrt(ut * ut + vt * vt)
    h43 = ht * _cbrt(ht)
    sf = (nt * nt * mag * ut) / _smax(TOLDRY + 0.0 * h43, h43)
    tau = GRAV * ht * sf * ds
    b0 = (thrust - tau) / den
    b2 = -b0
    # positivity guard (Roe.H:200-247)
    tol = lambda x: _where(_re(_fabs(x)) > TOL12, x,
                           _where(_re(x) >= 0.0, TOL12 + 0.0 * x, -TOL12 + 0.0 * x))
    l0s, l2s = tol(lam0), tol(lam2)
    h1s = hl + a0 - (b0 / l0s)
    h3s = hr - a2 + (b2 / l2s)
    mlam = _smax(_smax(_fabs(lam0), _fabs(lam1)), _fabs(lam2))
    loc_dt = ds / (mlam + TOL12)
    d1s, d3s = tol(hl - h1s), tol(hr - h3s)
    dt1 = (ds * 0.5 / l0s) * (hl / d1s)
    dt3 = (ds * 0.5 / l2s) * (hr / d3s)
    m1 = -(hl + a0) * _fabs(lam0)
    m3 = -(hr - a2) * lam2
    c1 = (_re(h1s) < -TOLDRY) & (_re(h3s) > TOLDRY) & (_re(dt1) < _re(loc_dt))
    b0 = _where(c1 & (_re(-m1) >= _re(m3)), m1, b0)
    b2 = _where(c1, -b0, b2)
    c2 = (_re(h1s) > TOLDRY) & (_re(h3s) < -TOLDRY) & (_re(dt3) < _re(loc_dt))
    b2 = _where(c2 & (_re(-m3) >= _re(m1)), m3, b2)
    b0 = _where(c2, -b2, b0)
    F = c["F"]
    z0 = 0.0 * F[0]
    Sm, Sp = [z0, z0, z0], [z0, z0, z0]
    for k, (lk, bk) in enumerate(((lam0, b0), (lam2, b2))):
        e = (1.0, lk, vt)
        pos, neg = _re(lk) > 0.0, _re(lk) < 0.0
        for q in range(3):
            Sm[q] = Sm[q] + _where(pos, bk * e[q], z0)
            Sp[q] = Sp[q] + _where(neg, bk * e[q], z0)
    F[1], F[2] = _rot(F[1], F[2], nx, ny)
    Sm[1], Sm[2] = _rot(Sm[1], Sm[2], nx, ny)
    Sp[1], Sp[2] = _rot(Sp[1], Sp[2], nx, ny)
    return F, Sm, Sp


def vanleer(qm, q, qp):
    """van Leer slope (HydrostaticRoe.cpp:173-195), vectorised."""
    dm = q - qm
    dp = qp - q
    flat = _re(_fabs(dp)) <= TOL12
    dps = _safe(dp, ~flat)
    r = dm / dps
    ar = _fabs(r)
    s = ((r + ar) / (1.0 + ar)) * dps
    return _where(flat, 0.0 * s, s)


def hydro_faces(Lm, L, R, Rp, nx, ny):
    """Hydrostatic-reconstruction faces (HydrostaticRoe.cpp:218-396); each
    cell is (h, hu, hv, z) with h rebuilt.  Returns FL, FR."""
    zL, zR = L[3], R[3]
    nod = (np.abs(_re(zL) + 9999.0) < TOL10) | (np.abs(_re(zR) + 9999.0) < TOL10)
    zf = _smax(zL, zR)
    eLm, eL, eR, eRp = Lm[0] + Lm[3], L[0] + zL, R[0] + zR, Rp[0] + Rp[3]
    etaLf = eL + 0.5 * vanleer(eLm, eL, eR)
    etaRf = eR - 0.5 * vanleer(eL, eR, eRp)

    def vel(X, k):
        w = _re(X[0]) > TOL12
        return _where(w, X[k] / _safe(X[0], w), 0.0 * X[0])

    uLm, uL, uR, uRp = vel(Lm, 1), vel(L, 1), vel(R, 1), vel(Rp, 1)
    vLm, vL, vR, vRp = vel(Lm, 2), vel(L, 2), vel(R, 2), vel(Rp, 2)
    uLf = uL + 0.5 * vanleer(uLm, uL, uR)
    uRf = uR - 0.5 * vanleer(uL, uR, uRp)
    vLf = vL + 0.5 * vanleer(vLm, vL, vR)
    vRf = vR - 0.5 * vanleer(vL, vR, vRp)
    pos = lambda x: _where(_re(x) > 0.0, x, 0.0 * x)     # smax(0, x)
    hLu, hRu = pos(etaLf - zL), pos(etaRf - zR)
    hLf, hRf = pos(etaLf - zf), pos(etaRf - zf)
    wL, wR = _re(hLf) > 0.0, _re(hRf) > 0.0
    huLf = _where(wL, hLf * uLf, 0.0 * hLf)
    hvLf = _where(wL, hLf * vLf, 0.0 * hLf)
    huRf = _where(wR, hRf * uRf, 0.0 * hRf)
    hvRf = _where(wR, hRf * vRf, 0.0 * hRf)
    f = _roe_core(hLf, huLf, hvLf, hRf, huRf, hvRf, nx, ny)["F"]
    f[1], f[2] = _rot(f[1], f[2], nx, ny)
    pL = 0.5 * GRAV * (hLu * hLu - hLf * hLf)
    pR = 0.5 * GRAV * (hRu * hRu - hRf * hRf)
    FL = [f[0], f[1] + pL * nx, f[2] + pL * ny]
    FR = [f[0], f[1] + pR * nx, f[2] + pR * ny]
    z0 = 0.0 * f[0]
    return [_where(nod, z0, x) for x in FL], [_where(nod, z0, x) for x in FR]


def step(solver, h, hu, hv, z, n, dt, dx, dy, bc_lo=(1, 1), bc_hi=(1, 1)):
    """One step (solver 0 plain Roe, 1 hydrostatic) with a given dt."""
    ny, nx = h.shape
    he, pu, pv, pz, pn = _work(h, hu, hv, z, n, bc_lo, bc_hi)
    vj, vi = slice(G, G + ny), slice(G, G + nx)
    cell = lambda jj, ii: (he[jj, ii], pu[jj, ii], pv[jj, ii], pz[jj, ii], pn[jj, ii])
    if solver == 0:
        # x faces i = 0..nx between padded columns G+i-1 and G+i
        L = cell(vj, slice(G - 1, G + nx))
        R = cell(vj, slice(G, G + nx + 1))
        Fx, Smx, Spx = roe_faces(L, R, dx, 1.0, 0.0)
        L = cell(slice(G - 1, G + ny), vi)
        R = cell(slice(G, G + ny + 1), vi)
        Fy, Smy, Spy = roe_faces(L, R, dy, 0.0, 1.0)
        uo = (he[vj, vi], hu, hv)
        out = []
        for k in range(3):
            out.append(uo[k]
                       - (dt / dx) * (Fx[k][:, 1:] - Fx[k][:, :-1])
                       - (dt / dy) * (Fy[k][1:, :] - Fy[k][:-1, :])
                       - (dt / dx) * (Spx[k][:, 1:] + Smx[k][:, :-1])
                       - (dt / dy) * (Spy[k][1:, :] + Smy[k][:-1, :]))
        return out
    c4 = lambda jj, ii: (he[jj, ii], pu[jj, ii], pv[jj, ii], pz[jj, ii])
    FLx, FRx = hydro_faces(*[c4(vj, slice(G - 2 + q, G + nx - 1 + q)) for q in range(4)], 1.0, 0.0)
    FLy, FRy = hydro_faces(*[c4(slice(G - 2 + q, G + ny - 1 + q), vi) for q in range(4)], 0.0, 1.0)
    un = []
    for k, uo in enumerate((h, hu, hv)):
        un.append(uo
                  - (dt / dx) * (FLx[k][:, 1:] - FRx[k][:, :-1])
                  - (dt / dy) * (FLy[k][1:, :] - FRy[k][:-1, :]))
    hn, qx, qy = un
    wet = _re(hn) >= TOLDRY
    hs = _safe(hn, wet)
    nsq = n * n
    inv_h = 1.0 / hs
    u, v = qx * inv_h, qy * inv_h
    modM = _sqrt(u * u + v * v)
    act = wet & (_re(nsq) > TOL12) & (_re(modM) > TOL12)
    ms = _safe(modM, act)
    h43 = hs * _cbrt(hs)
    tt = dt * GRAV * nsq * ms / h43
    ff = _sqrt(1.0 + 4.0 * tt)
    fac = 2.0 / (1.0 + ff)
    z0 = 0.0 * hn
    ho = _where(wet, hn, _where(_re(hn) < 0.0, z0, hn))
    huo = _where(wet, _where(act, qx * fac, qx), z0)
    hvo = _where(wet, _where(act, qy * fac, qy), z0)
    return [ho, huo, hvo]


def rollout(solver, h, hu, hv, z, n, dts, dx, dy, bc_lo, bc_hi, gauges):
    """Frozen-dt rollout: returns (series[s, g] of h at the flat gauge
    indices after step s + 1, final state (h, hu, hv))."""
    series = []
    for dt in dts:
        h, hu, hv = step(solver, h, hu, hv, z, n, dt, dx, dy, bc_lo, bc_hi)
        series.append(h.reshape(-1)[gauges])
    return np.array(series), (h, hu, hv)


def directional_derivative(solver, state, z, n, dts, dx, dy, bc_lo, bc_hi, gauges,
                           d_state=None, d_z=None, d_n=None):
    """Complex-step tangent of (series, final state) along the direction
    (d_state = (dh, dhu, dhv), d_z, d_n)."""
    def pert(x, d):
        x = np.asarray(x, dtype=np.complex128)
        return x if d is None else x + 1j * EPS * np.asarray(d)

    h, hu, hv = state
    ds = d_state if d_state is not None else (None, None, None)
    s, fin = rollout(solver, pert(h, ds[0]), pert(hu, ds[1]), pert(hv, ds[2]), pert(z, d_z),
                     pert(n, d_n), dts, dx, dy, bc_lo, bc_hi, gauges)
    return np.imag(s) / EPS, tuple(np.imag(f) / EPS for f in fin), (np.real(s), tuple(np.real(f) for f in fin))
