"""Pin the oracle's per-face functions with identities the mathematics fixes.

None of these retypes the oracle's formula: each check is a property of the
Roe linearisation, of the well-balanced source split, of the van Leer
limiter or of the hydrostatic reconstruction, chosen so that a dropped term,
a wrong sign or a swapped operand fails it.
"""
import numpy as np
import pytest

import oracle as O

G = 9.81


def phys_flux_x(h, hu, hv):
    u, v = hu / h, hv / h
    return np.array([hu, hu * u + 0.5 * G * h * h, hu * v])


def rand_state(rng, hmin=0.2, hmax=2.0, umax=0.5):
    h = rng.uniform(hmin, hmax)
    return h, h * rng.uniform(-umax, umax), h * rng.uniform(-umax, umax)


# ---------------------------------------------------------------- plain Roe

def test_roe_consistency():
    """F(U, U) is the physical flux and carries no source on a flat,
    frictionless bed."""
    rng = np.random.default_rng(1)
    for _ in range(200):
        h, hu, hv = rand_state(rng)
        s = [h, hu, hv, 0.3, 0.0]
        F, Sm, Sp = O.roe_face(s, s, 1e-3, 0.01, 1.0, 0.0)
        np.testing.assert_allclose(F, phys_flux_x(h, hu, hv), rtol=1e-12, atol=1e-13)
        assert np.all(Sm == 0) and np.all(Sp == 0)


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_roe_supersonic_upwinding(sign):
    """With every eigenvalue of one sign (and no entropy fix active) the Roe
    flux is exactly the upwind physical flux: this only holds when the Roe
    averages, eigenvectors and wave strengths are all right."""
    rng = np.random.default_rng(2)
    for _ in range(200):
        hl, hr = rng.uniform(0.5, 1.5), rng.uniform(0.5, 1.5)
        ul, ur = sign * rng.uniform(5.0, 8.0), sign * rng.uniform(5.0, 8.0)
        vl, vr = rng.uniform(-1, 1), rng.uniform(-1, 1)
        L = [hl, hl * ul, hl * vl, 0.0, 0.0]
        R = [hr, hr * ur, hr * vr, 0.0, 0.0]
        F, _, _ = O.roe_face(L, R, 1e-3, 0.01, 1.0, 0.0)
        up = phys_flux_x(*(L[:3] if sign > 0 else R[:3]))
        # the TOL12 regularisations perturb at ~1e-12 relative
        np.testing.assert_allclose(F, up, rtol=1e-10, atol=1e-10)


def test_roe_mirror_symmetry():
    """Reflecting the problem (swap sides, negate normal velocity) negates
    the mass and transverse fluxes and keeps the normal momentum flux; the
    flux seen by the left cell, F + S+, becomes the one seen by the right
    cell, F - S-, so the source halves swap with the opposite pattern.

    The reference's bed thrust is antisymmetric only on its thrust_b branch
    with dztilde = dz (kernels/Roe.H:161-176: the max() branch and the
    emerged-step branch are not), so the states keep u dz < 0 and every
    depth above the step height."""
    rng = np.random.default_rng(3)
    flip = np.array([-1.0, 1.0, -1.0])
    for _ in range(200):
        sgn = rng.choice([-1.0, 1.0])
        hl, hr = rng.uniform(0.2, 2.0), rng.uniform(0.2, 2.0)
        ul, ur = sgn * rng.uniform(0.01, 0.5), sgn * rng.uniform(0.01, 0.5)
        vl, vr = rng.uniform(-.5, .5), rng.uniform(-.5, .5)
        zl = rng.uniform(0, 0.1)
        zr = zl - sgn * rng.uniform(0, 0.1)
        nl, nr = rng.uniform(0, 0.05), rng.uniform(0, 0.05)
        L = [hl, hl * ul, hl * vl, zl, nl]
        R = [hr, hr * ur, hr * vr, zr, nr]
        F, Sm, Sp = O.roe_face(L, R, 1e-3, 0.01, 1.0, 0.0)
        Fm, Smm, Spm = O.roe_face([hr, -hr * ur, hr * vr, zr, nr], [hl, -hl * ul, hl * vl, zl, nl],
                                  1e-3, 0.01, 1.0, 0.0)
        np.testing.assert_allclose(Fm, flip * F, rtol=1e-11, atol=1e-12)
        np.testing.assert_allclose(Smm, -flip * Sp, rtol=1e-11, atol=1e-12)
        np.testing.assert_allclose(Spm, -flip * Sm, rtol=1e-11, atol=1e-12)


def test_roe_rotation():
    """A y face of (h, hu, hv) is the x face of (h, hv, -hu) rotated back."""
    rng = np.random.default_rng(4)
    for _ in range(100):
        L = [*rand_state(rng), rng.uniform(0, .2), rng.uniform(0, .05)]
        R = [*rand_state(rng), rng.uniform(0, .2), rng.uniform(0, .05)]
        Fy, Smy, Spy = O.roe_face(L, R, 1e-3, 0.02, 0.0, 1.0)
        rot = lambda s: [s[0], s[2], -s[1], s[3], s[4]]
        Fx, Smx, Spx = O.roe_face(rot(L), rot(R), 1e-3, 0.02, 1.0, 0.0)
        back = lambda a: np.array([a[0], -a[2], a[1]])
        np.testing.assert_allclose(Fy, back(Fx), rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(Smy, back(Smx), rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(Spy, back(Spx), rtol=1e-14, atol=1e-15)


def test_roe_lake_at_rest_balance():
    """C-property at one face: with h + z constant and no motion, the flux
    seen by the left cell (F + S+) is that cell's hydrostatic pressure and
    the flux seen by the right cell (F - S-) is the right cell's."""
    rng = np.random.default_rng(5)
    for _ in range(200):
        eta = rng.uniform(1.0, 2.0)
        zl, zr = rng.uniform(0, 0.8), rng.uniform(0, 0.8)
        hl, hr = eta - zl, eta - zr
        F, Sm, Sp = O.roe_face([hl, 0, 0, zl, 0], [hr, 0, 0, zr, 0], 1e-3, 0.01, 1.0, 0.0)
        np.testing.assert_allclose(F + Sp, [0, 0.5 * G * hl * hl, 0], atol=1e-12 * G * eta ** 2)
        np.testing.assert_allclose(F - Sm, [0, 0.5 * G * hr * hr, 0], atol=1e-12 * G * eta ** 2)


def test_roe_manning_closed_form():
    """Uniform flow on a flat bed: the two source halves of a face net to the
    Manning source -g n^2 u|u| / h^(1/3) per unit (dt/dx) in the momentum
    equation and nothing in continuity."""
    rng = np.random.default_rng(6)
    dx = 0.05
    for _ in range(100):
        h = rng.uniform(0.1, 3.0)
        u, v = rng.uniform(-2, 2), rng.uniform(-0.3, 0.3)
        n = rng.uniform(0.01, 0.08)
        s = [h, h * u, h * v, 0.2, n]
        F, Sm, Sp = O.roe_face(s, s, 1e-3, dx, 1.0, 0.0)
        net = (Sp + Sm) / dx   # the update of a cell with identical faces
        c = np.sqrt(G * h)
        speed = np.sqrt(u * u + v * v)
        expect = -G * n * n * speed * u / np.cbrt(h) * (2 * c / (2 * c + 1e-12))
        assert abs(net[0]) < 1e-13
        if abs(u) < c:   # both acoustic waves carry a share only when subsonic
            np.testing.assert_allclose(-net[1], expect, rtol=1e-11, atol=1e-15)


def test_roe_dry_dry_face():
    F, Sm, Sp = O.roe_face([0, 0, 0, 0, 0], [5e-7, 0, 0, 0, 0], 1e-3, 0.01, 1.0, 0.0)
    assert np.all(F == 0) and np.all(Sm == 0) and np.all(Sp == 0)


def stationary_jump(h_up, h_dn, sign=1.0):
    """Two states joined by a stationary Rankine-Hugoniot discontinuity: equal
    mass flux q and equal momentum flux q^2/h + g h^2/2 on both sides, so
    q^2 = g h_up h_dn (h_up + h_dn) / 2.  The jump is an eigenvector of the Roe
    matrix with eigenvalue u~ - c~ = 0 (sign +1) or u~ + c~ = 0 (sign -1)."""
    q = sign * np.sqrt(G * h_up * h_dn * (h_up + h_dn) / 2.0)
    return q, [h_up, q, 0.0, 0.0, 0.0], [h_dn, q, 0.0, 0.0, 0.0]


def test_roe_keeps_an_admissible_stationary_shock():
    """A stationary hydraulic jump (shallow fast water into deep slow water)
    is a Lax shock: no eigenvalue changes sign from negative to positive
    across it, so the entropy fix must stay off and the Roe flux resolves the
    jump exactly (F = f_L = f_R).  An entropy-fix test with its sign
    conditions swapped would fire here."""
    for hl, hr in ((0.5, 1.0), (0.1, 0.9), (0.3, 2.0)):
        q, L, R = stationary_jump(hl, hr)
        F, Sm, Sp = O.roe_face(L, R, 1e-3, 0.01, 1.0, 0.0)
        np.testing.assert_allclose(F, phys_flux_x(*L[:3]), rtol=1e-11, atol=1e-12)
        np.testing.assert_allclose(F, phys_flux_x(*R[:3]), rtol=1e-11, atol=1e-12)
        # the mirrored jump (flow to the left) is admissible too
        F2, _, _ = O.roe_face([hr, -q, 0, 0, 0], [hl, -q, 0, 0, 0], 1e-3, 0.01, 1.0, 0.0)
        np.testing.assert_allclose(F2, phys_flux_x(hr, -q, 0), rtol=1e-11, atol=1e-12)


def test_roe_entropy_fix_breaks_a_stationary_expansion_shock():
    """The same stationary discontinuity run backwards (deep slow water into
    shallow fast water) is an expansion shock: u - c goes from negative on the
    left to positive on the right.  Without a fix the Roe flux would keep it
    (F = f_L, since the jump's Roe eigenvalue is 0).  The reference applies
    Harten-Hyman's fixed eigenvalue l* = e_L (e_R - l) / (e_R - e_L),
    e = u -/+ sqrt(g h) (kernels/Roe.H:113-127) before the wave strengths, so
    with l = 0 exactly the strengths are alpha0 = dh and alpha2 = -l* dh / 2c~
    (wave 1), and the mass flux becomes q - |l*| dh: q + |l*| (h_L - h_R).
    The mirrored problem (flow to the left, wave 3) gives q - l* (h_R - h_L)."""
    for hl, hr in ((1.0, 0.5), (0.9, 0.1), (2.0, 0.3)):
        q, L, R = stationary_jump(hl, hr)
        el, er = q / hl - np.sqrt(G * hl), q / hr - np.sqrt(G * hr)
        assert el < 0.0 < er
        lstar = el * er / (er - el)
        F, _, _ = O.roe_face(L, R, 1e-3, 0.01, 1.0, 0.0)
        np.testing.assert_allclose(F[0], q + abs(lstar) * (hl - hr), rtol=1e-10)
        assert abs(F[0] - q) > 0.05 * q          # the expansion shock is not kept
        # wave 3: left state shallow and fast to the left, right deep
        el3, er3 = -q / hr + np.sqrt(G * hr), -q / hl + np.sqrt(G * hl)
        assert el3 < 0.0 < er3
        lstar3 = -el3 * er3 / (er3 - el3)
        F3, _, _ = O.roe_face([hr, -q, 0, 0, 0], [hl, -q, 0, 0, 0], 1e-3, 0.01, 1.0, 0.0)
        np.testing.assert_allclose(F3[0], -q - lstar3 * (hl - hr), rtol=1e-10)


def test_roe_emerged_step_is_a_wall():
    """Water at rest beside a dry step that rises above its surface, step on
    the right (dz >= 0, d_i = h_i + z_i < z_j): the depth-limited thrust
    (dztilde = h_i, kernels/Roe.H:161-163) makes the face a wall.  The wet
    cell sees exactly its hydrostatic pressure (F + S+ = (0, g h^2 / 2, 0))
    and the dry cell receives nothing (F - S- = 0).  A centred thrust, a
    wrong branch test or dztilde = dz leaves a net force."""
    rng = np.random.default_rng(11)
    for _ in range(200):
        h = rng.uniform(0.05, 2.0)
        zw = rng.uniform(-0.5, 0.5)
        zs = zw + h + rng.uniform(1e-3, 1.0)          # step crest above the surface
        n = rng.uniform(0.0, 0.05)
        wall = [0.0, 0.5 * G * h * h, 0.0]
        F, Sm, Sp = O.roe_face([h, 0, 0, zw, n], [0, 0, 0, zs, n], 1e-3, 0.01, 1.0, 0.0)
        np.testing.assert_allclose(F + Sp, wall, rtol=1e-11, atol=1e-12 * G * h * h)
        np.testing.assert_allclose(F - Sm, 0.0, atol=1e-12 * G * h * h)


def test_roe_emerged_step_on_the_left_leaks():
    """The mirrored step (dz < 0, d_j < z_i) takes dztilde = h_j
    (kernels/Roe.H:164-165), which is positive although dz is negative, so the
    thrust -g (h_j - h_j/2) h_j = -g h_j^2 / 2 has the sign of the other
    branch.  Worked through the Roe face (h~ = h/2, c~ = sqrt(g h / 2),
    u~ = 0, alpha0 = alpha2 = h/2, beta0 = -c~ h / 2): the wet right cell
    sees F - S- = (-c~ h, 0, 0) -- no pressure from the wall and a mass flux
    onto the step -- and the dry left cell F + S+ = (-c~ h, g h^2 / 2, 0).
    This is the reference's behaviour and the oracle keeps it; dztilde = -h_j
    (the wall) or dztilde = dz would fail here."""
    rng = np.random.default_rng(12)
    for _ in range(200):
        h = rng.uniform(0.05, 2.0)
        zw = rng.uniform(-0.5, 0.5)
        zs = zw + h + rng.uniform(1e-3, 1.0)
        n = rng.uniform(0.0, 0.05)
        c = np.sqrt(G * h / 2.0)
        F, Sm, Sp = O.roe_face([0, 0, 0, zs, n], [h, 0, 0, zw, n], 1e-3, 0.01, 1.0, 0.0)
        np.testing.assert_allclose(F - Sm, [-c * h, 0.0, 0.0], rtol=1e-11, atol=1e-12 * G * h * h)
        np.testing.assert_allclose(F + Sp, [-c * h, 0.5 * G * h * h, 0.0], rtol=1e-11,
                                   atol=1e-12 * G * h * h)


def star_depths(hi, ui, hj, uj, zj, n, dx):
    """The reference's star depths h1* = h_i + alpha0 - beta0/l0 and
    h3* = h_j - alpha2 + beta2/l2 (kernels/Roe.H:209-210) read back from one
    face's outputs.  With l0 < 0 < l2 and the entropy fix off, S+ = beta0 e0
    and S- = beta2 e2 with e = (1, l, v~) (so l = S[1]/S[0]), and the Roe
    property f_j - f_i = sum l_k alpha_k e_k makes F0 = q_i + l0 alpha0 =
    q_j - l2 alpha2.  So h1* = h_i + (F0 - q_i - S+0) / l0 and
    h3* = h_j - (q_j - F0 - S-0) / l2: the depths the guard limits, with
    whatever beta the face ended with."""
    F, Sm, Sp = O.roe_face([hi, hi * ui, 0, 0, n], [hj, hj * uj, 0, zj, n], 1e-3, dx, 1.0, 0.0)
    l0, l2 = Sp[1] / Sp[0], Sm[1] / Sm[0]
    assert l0 < 0.0 < l2
    h1 = hi + (F[0] - hi * ui - Sp[0]) / l0
    h3 = hj - (hj * uj - F[0] - Sm[0]) / l2
    return h1, h3


# (h_i, u_i, h_j, u_j, z_j, n, dx), z_i = 0, same roughness on both sides.
# Rough diverging or converging flow over a small step, entropy fix off and
# |l0|, |l2| > 0.1.  GUARD_LEFT: without the guard h1* < -0.8 while h3* > 0.1;
# GUARD_RIGHT: h3* < -1 while h1* > 0.1, and the 3-wave reaches the cell
# within the face's local step dx / max|l|; LATE_RIGHT: the same sign pattern
# but the 3-wave would arrive after 1.2-2.5 local steps, so the guard holds back.
GUARD_LEFT = [(0.9, -0.7, 1.68, 2.8, 0.09, 0.23, 50.0), (0.23, 1.0, 0.94, 2.4, 0.37, 0.24, 1.0),
              (0.76, 1.2, 0.13, -1.3, -0.01, 0.27, 50.0)]
GUARD_RIGHT = [(0.3, 1.9, 0.63, -2.4, 0.1, 0.22, 50.0), (0.91, -2.6, 1.93, -1.7, 0.06, 0.23, 50.0),
               (0.23, -1.3, 1.12, -2.4, -0.5, 0.27, 50.0)]
LATE_RIGHT = [(0.31, -1.2, 0.74, -2.5, -0.26, 0.16, 1.0), (0.91, -0.6, 1.92, -3.0, -0.36, 0.11, 50.0),
              (0.17, 0.7, 1.74, -2.1, -0.27, 0.1, 50.0)]


@pytest.mark.parametrize("case", GUARD_LEFT)
def test_roe_positivity_guard_empties_the_left_cell(case):
    """Where the source share would drive the left star depth negative while
    the right one stays wet, the guard limits beta0 to the value that just
    empties the cell (kernels/Roe.H:228-236): h1* = 0 to round-off."""
    h1, h3 = star_depths(*case)
    assert abs(h1) <= 1e-11 * max(case[0], case[2]), h1
    assert h3 > 1e-6


@pytest.mark.parametrize("case", GUARD_RIGHT)
def test_roe_positivity_guard_empties_the_right_cell(case):
    """The mirrored branch (kernels/Roe.H:238-246) limits beta2: h3* = 0."""
    h1, h3 = star_depths(*case)
    assert abs(h3) <= 1e-11 * max(case[0], case[2]), h3
    assert h1 > 1e-6


@pytest.mark.parametrize("case", LATE_RIGHT)
def test_roe_positivity_guard_waits_for_waves_beyond_the_local_step(case):
    """A negative star depth whose wave would not reach the cell within the
    local step (dt3* >= dx / max|l|) is left alone: the unlimited h3* stays."""
    h1, h3 = star_depths(*case)
    assert h3 < -0.2, h3
    assert h1 > 1e-6


# ---------------------------------------------------------- homogeneous Roe

def test_hroe_consistency_and_upwinding():
    rng = np.random.default_rng(7)
    for _ in range(200):
        h, hu, hv = rand_state(rng)
        F = O.hroe_face(h, hu, hv, h, hu, hv, 1.0, 0.0)
        np.testing.assert_allclose(F, phys_flux_x(h, hu, hv), rtol=1e-12, atol=1e-13)
    for _ in range(100):
        hl, hr = rng.uniform(0.5, 1.5), rng.uniform(0.5, 1.5)
        ul, ur = rng.uniform(5, 8), rng.uniform(5, 8)
        F = O.hroe_face(hl, hl * ul, 0.1, hr, hr * ur, 0.2, 1.0, 0.0)
        np.testing.assert_allclose(F, phys_flux_x(hl, hl * ul, 0.1), rtol=1e-10, atol=1e-10)


def test_hroe_equals_plain_roe_flux_on_flat_bed():
    """On a flat frictionless bed the two kernels share their flux."""
    rng = np.random.default_rng(8)
    for _ in range(100):
        L, R = rand_state(rng), rand_state(rng)
        F1 = O.hroe_face(*L, *R, 0.0, 1.0)
        F2, _, _ = O.roe_face([*L, 0, 0], [*R, 0, 0], 1e-3, 0.01, 0.0, 1.0)
        np.testing.assert_array_equal(F1, F2)


# ----------------------------------------------------------------- van Leer

def test_vanleer_harmonic_mean_form():
    """phi(r) dp with phi = (r+|r|)/(1+|r|) is van Leer's harmonic mean
    2 dm dp / (dm + dp) when dm dp > 0 and zero otherwise."""
    rng = np.random.default_rng(9)
    for _ in range(2000):
        q = rng.normal(size=3) * rng.choice([1e-3, 1.0, 1e3])
        dm, dp = q[1] - q[0], q[2] - q[1]
        s = O.vanleer_slope(*q)
        if abs(dp) <= 1e-12:
            assert s == 0.0
        elif dm * dp > 0:
            np.testing.assert_allclose(s, 2 * dm * dp / (dm + dp), rtol=1e-13)
        else:
            assert s == 0.0


def test_vanleer_limits():
    assert O.vanleer_slope(0.0, 1.0, 2.0) == pytest.approx(1.0)        # r = 1
    assert O.vanleer_slope(0.0, 1.0, 1.0 + 1e-13) == 0.0                # tiny dp
    assert O.vanleer_slope(2.0, 1.0, 2.0) == 0.0                        # extremum
    big = O.vanleer_slope(0.0, 1e6, 1e6 + 1.0)                          # r -> inf
    assert big == pytest.approx(2.0, rel=1e-5)


# -------------------------------------------------------- hydrostatic face

def test_hydro_face_lake_at_rest():
    """Constant free surface and no motion: the two face fluxes are the
    hydrostatic pressures of their own cells (Audusse reconstruction)."""
    rng = np.random.default_rng(10)
    for _ in range(200):
        eta = rng.uniform(1.0, 2.0)
        z = rng.uniform(0, 0.9, size=4)
        cells = np.array([[eta - zi, 0, 0, zi] for zi in z])
        FL, FR = O.hydro_face(cells, 1.0, 0.0)
        hL, hR = eta - z[1], eta - z[2]
        np.testing.assert_allclose(FL, [0, 0.5 * G * hL * hL, 0], atol=1e-13 * G * eta ** 2)
        np.testing.assert_allclose(FR, [0, 0.5 * G * hR * hR, 0], atol=1e-13 * G * eta ** 2)


def test_hydro_face_constant_state_flat_bed():
    rng = np.random.default_rng(11)
    for _ in range(100):
        h, hu, hv = rand_state(rng)
        cells = np.array([[h, hu, hv, 0.1]] * 4)
        FL, FR = O.hydro_face(cells, 1.0, 0.0)
        np.testing.assert_allclose(FL, phys_flux_x(h, hu, hv), rtol=1e-12, atol=1e-13)
        np.testing.assert_array_equal(FL, FR)


def test_hydro_face_emerged_step_is_wall():
    """A bed step above the free surface: the face depth is zero, no mass
    crosses, and each side sees its own hydrostatic pressure."""
    eta = 1.0
    cells = np.array([[eta - 0.2, 0, 0, 0.2], [eta - 0.2, 0, 0, 0.2], [0.0, 0, 0, 1.5], [0.0, 0, 0, 1.5]])
    FL, FR = O.hydro_face(cells, 1.0, 0.0)
    assert FL[0] == 0.0 and FR[0] == 0.0
    np.testing.assert_allclose(FL[1], 0.5 * G * 0.8 ** 2, rtol=1e-14)
    assert FR[1] == 0.0


def test_hydro_face_nodata_wall():
    cells = np.array([[1, .1, 0, 0], [1, .1, 0, 0], [0, 0, 0, -9999.0], [0, 0, 0, -9999.0]])
    FL, FR = O.hydro_face(cells, 1.0, 0.0)
    assert np.all(FL == 0) and np.all(FR == 0)
