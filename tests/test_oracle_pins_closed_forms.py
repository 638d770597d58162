"""Closed-form pins of the oracle's PPM fluxes and limiters, vertical PPM
limiters, shortwave column and the C/D-grid damping and pressure-gradient
terms.  Each expectation is derived independently of the oracle's code
(hand algebra on special profiles, a boundary-value solve of the two-stream
equations, discrete Fourier eigenvalues, the finite-volume contour identity)
and each test also shows that a plausible wrong variant misses it.
"""
import numpy as np
import pytest

import oracle as O

W7_12 = float(np.float32(7.0 / 12.0))  # binary32 literals (Q1)
W1_12 = float(np.float32(1.0 / 12.0))
CP = 1005.0


# ---- PPM fluxes (mod_sw_dyn.f90:781-919) ---------------------------------------
def _k12(s):
    return (1 - 2 * s) * (1 - s), s * (s - 1)  # the reference's weights (Q3)


def test_ppm_flux_on_a_linear_profile():
    """q = a + b i: dqi = dqi_mono = b/2, so both edge anomalies are -+b/2
    in every variant (no limiter acts), and the upwind face value is
    q(i-1) + (b/2)(1-s)^2 for c > 0, q(i) - (b/2)(1-s)^2 for c < 0 with the
    Q3 weights k1 = (1-2s)(1-s), k2 = s(s-1).  The textbook upwind PPM value
    q(i-1) + (b/2)(1-s), or k1 and k2 swapped, miss it."""
    n, a, b = 18, 3.0, 0.7
    q = a + b * np.arange(n)
    i = np.arange(3, n - 2)
    for var in (O.PPM_UNCONSTRAINED, O.PPM_CONSTRAINED, O.PPM_NONNEGATIVE):
        for s in (0.1, 0.35, 0.8):
            for sg in (1.0, -1.0):
                f = O.flux_pencil(q, np.full(n, sg * s), var)[i]
                up = q[i - 1] if sg > 0 else q[i]
                expect = up + sg * (b / 2) * (1 - s) ** 2
                np.testing.assert_allclose(f, expect, rtol=1e-14, atol=1e-14)
                textbook = up + sg * (b / 2) * (1 - s)
                k1, k2 = _k12(s)
                swapped = up + sg * (b / 2) * (k2 - k1)
                assert np.max(np.abs(f - textbook)) > 1e-3 and np.max(np.abs(f - swapped)) > 1e-3


def test_ppm_flux_on_a_quadratic_profile():
    """Cell means of x^2 (q_i = x_i^2 + 1/12, x_i = i + 5): dqi_mono = x_i and
    the edge anomalies are those of the parabola itself, L = -x + 1/6,
    R = x + 1/6 (the 4th-order edge values), so for c > 0 the face value is
    q(i-1) + k1 R(i-1) + k2 L(i-1).  Edges without the (dqm(i-1) - dqm(i))/3
    curvature term (L = -x + 1/2) miss it."""
    n = 16
    x = np.arange(n) + 5.0
    q = x ** 2 + 1.0 / 12.0
    i = np.arange(3, n - 2)
    for s in (0.2, 0.6, 0.9):
        f = O.flux_pencil(q, np.full(n, s), O.PPM_UNCONSTRAINED)[i]
        k1, k2 = _k12(s)
        xm = x[i - 1]
        expect = q[i - 1] + k1 * (xm + 1 / 6) + k2 * (-xm + 1 / 6)
        np.testing.assert_allclose(f, expect, rtol=1e-13)
        no_curv = q[i - 1] + k1 * (xm + 1 / 2) + k2 * (-xm + 1 / 2)
        assert np.max(np.abs(f - no_curv) / np.abs(expect)) > 1e-4


def test_ppm_constrained_falls_back_to_upwind():
    """q = 0 0 0 0 0 1 1 0, face between cells 4 and 5 (0-based).  By hand:
    dqi_mono is 0 in cells 3..6 (each sits at a one-sided extremum of its
    three values); cell 4: L0 = 0, R0 = 1/2, after the constraint L = -0,
    R = 1/2; cell 5: L0 = -1/2, R0 = 0 -> L = -1/2, R = 0.  Both cells satisfy
    3|L+R| = 3/2 > |L-R| = 1/2, so PPM_CONSTRAINED returns the upwind cell
    value (flux 821-827): 0 for c > 0, 1 for c < 0.  The constrained
    parabola without the fallback gives k1/2 (c > 0) and 1 - k1/2 (c < 0)."""
    q = np.array([0, 0, 0, 0, 0, 1, 1, 0], dtype=float)
    for s in (0.3, 0.4, 0.7):
        k1, _ = _k12(s)
        fp = O.flux_pencil(q, np.full(8, s), O.PPM_CONSTRAINED)[5]
        fm = O.flux_pencil(q, np.full(8, -s), O.PPM_CONSTRAINED)[5]
        assert fp == 0.0 and fm == 1.0
        assert abs(O.flux_pencil(q, np.full(8, s), O.PPM_UNCONSTRAINED)[5] - k1 / 2) < 1e-15
        assert abs(O.flux_pencil(q, np.full(8, -s), O.PPM_UNCONSTRAINED)[5] - (1 - k1 / 2)) < 1e-15
        assert abs(k1 / 2) > 1e-3  # the fallback changes the value


def test_ppm_nonnegative_filter_on_a_spike():
    """A single cell of 1 among zeros.  By hand: every dqi_mono is 0; the spike
    cell has L0 = R0 = -1/2, constrained to 0 (dqi = 0 there); its left
    neighbour has L = -0, R = 1/2, where the negative-minimum test holds
    (|a| = 1/2 < -b = 3/2 and 0 + a^2/(4b) + b/12 < 0) with L R = 0 and L < R,
    so R becomes 2|L| = 0 (ppm 900-909).  Face into the spike with c = 0.8 > 0:
    PPM_NONNEGATIVE gives 0; the unconstrained and constrained variants give
    k1/2 = -0.06 (a negative face value); face out of the spike: 1 for the
    constrained variants, 1 - (k1 + k2)/2 = 1.14 unconstrained."""
    n, k = 12, 6
    q = np.zeros(n)
    q[k] = 1.0
    s = 0.8
    k1, k2 = _k12(s)
    c = np.full(n, s)
    nn = O.flux_pencil(q, c, O.PPM_NONNEGATIVE)
    un = O.flux_pencil(q, c, O.PPM_UNCONSTRAINED)
    co = O.flux_pencil(q, c, O.PPM_CONSTRAINED)
    assert nn[k] == 0.0
    assert abs(un[k] - k1 / 2) < 1e-15 and abs(co[k] - k1 / 2) < 1e-15 and k1 / 2 < 0
    assert nn[k + 1] == 1.0 and co[k + 1] == 1.0
    assert abs(un[k + 1] - (1 - (k1 + k2) / 2)) < 1e-14


def test_ppm_nonnegative_keeps_faces_non_negative():
    """The purpose of PPM_NONNEGATIVE (edge floors at -q and the negative-
    minimum filter, ppm 890-911): for non-negative pencils every face value
    is >= 0 for any Courant number; the other variants produce negative
    face values on the same pencils (so a dropped filter branch fails)."""
    rng = np.random.default_rng(4)
    n = 24
    neg_u = neg_c = 0
    for _ in range(400):
        q = np.maximum(rng.standard_normal(n) * 2, 0) * rng.uniform(0.1, 10)
        c = rng.uniform(-1, 1, n)
        f = O.flux_pencil(q, c, O.PPM_NONNEGATIVE)[3:n - 2]
        assert np.all(f >= 0.0)
        neg_u += np.any(O.flux_pencil(q, c, O.PPM_UNCONSTRAINED)[3:n - 2] < 0)
        neg_c += np.any(O.flux_pencil(q, c, O.PPM_CONSTRAINED)[3:n - 2] < 0)
    assert neg_u > 300 and neg_c > 300


# ---- vertical PPM limiters (mod_remap.f90:287-356) ---------------------------
def _uniform_gam(L):
    """gam of the elimination on a uniform grid: alpha = 1 everywhere, so
    gam(L) = (1 + 2.5)/1.5 = 7/3 and gam(k-1) = 1/(4 - gam(k)) -> 2 - sqrt 3:
    all positive (the two-sided monotonicity clamp applies everywhere)."""
    g = np.zeros(L + 2)
    g[L] = 7.0 / 3.0
    for k in range(L, 1, -1):
        g[k - 1] = 1.0 / (4.0 - g[k])
    return g


def test_vertical_ppm_clamps():
    """On a uniform grid with jumps in the data (the tridiagonal values then
    overshoot): level 1 is not limited; level 2 is clipped to [min, max] of
    its two layers (316-317); level 3 (the first monotonicity clamp, 325-334,
    gam(4) gam(2) > 0) likewise; layers 1-2 are outside the Huynh
    constraint, so those are their final edges.  Each clamp fires in these
    cases (the raw value lies outside its range), so a dropped clamp fails.
    The curvature is 6 v - 3 (top + btm) (354)."""
    rng = np.random.default_rng(9)
    L = 12
    coord = 1e5 - 1000.0 * np.arange(L + 1)
    assert np.all(_uniform_gam(L)[1:L + 1] > 0)
    fired2 = fired3 = 0
    for _ in range(200):
        v = rng.choice([0.0, 1.0, 5.0, -3.0], size=L) + rng.normal(0, 0.01, L)
        top, btm, curv, raw = O.vertical_ppm(v, coord)
        assert btm[0] == raw[0]
        c2 = min(max(raw[1], min(v[0], v[1])), max(v[0], v[1]))
        assert top[0] == c2 and btm[1] == c2
        c3 = min(max(raw[2], min(v[1], v[2])), max(v[1], v[2]))
        assert top[1] == c3
        fired2 += c2 != raw[1]
        fired3 += c3 != raw[2]
        np.testing.assert_allclose(curv, 6 * v - 3 * (top + btm), rtol=0, atol=1e-12 * np.max(np.abs(v)))
    assert fired2 > 50 and fired3 > 50


def test_vertical_ppm_huynh_uses_gam_as_slopes():
    """Q8: the Huynh constraint (340-352) adds gam (dimensionless elimination
    coefficients, ~0.27 on a uniform grid) to the layer values as if they
    were slopes, so from layer 3 up the top edge stays within
    2|gam(k-1)| + 1.5|gam(k-2)| of its layer value (the bottom edge within
    2|gam(k)| + 1.5|gam(k-1)|) however large the data,
    and the limited edges are not homogeneous in the data (scaling the
    layer values by 1000 does not scale the edges).  A Huynh constraint with
    data slopes keeps the exact edges of linear data and scales with them."""
    L = 12
    coord = 1e5 - 1000.0 * np.arange(L + 1)
    gam = _uniform_gam(L)
    v = 1e4 - 300.0 * np.arange(L)  # linear data: the raw edges are exact (v +- 150)
    top, btm, curv, raw = O.vertical_ppm(v, coord)
    for k in range(3, L + 1):  # 1-based layers
        bound_t = 2 * abs(gam[k - 1]) + 1.5 * abs(gam[k - 2]) + 1e-9  # top: gam(k-1), gam(k-2)
        bound_b = 2 * abs(gam[k]) + 1.5 * abs(gam[k - 1]) + 1e-9  # bottom: gam(k), gam(k-1)
        assert abs(top[k - 1] - v[k - 1]) <= bound_t and abs(btm[k - 1] - v[k - 1]) <= bound_b
    assert np.max(np.abs(raw[3:L] - (v[3:L] + v[2:L - 1]) / 2)) < 1e-6  # raw edges exact
    # decreasing data: the top edges are clamped to v - 2 gam(k-1), which does
    # not scale with the data
    np.testing.assert_allclose(top[2:L], v[2:L] - 2 * gam[2:L], rtol=0, atol=1e-9)
    t2, b2, _, _ = O.vertical_ppm(1000.0 * v, coord)
    assert np.max(np.abs(t2 - 1000.0 * top)) > 100.0


# ---- shortwave column (mod_radiation.f90:158-268) ----------------------------
def _sw_two_stream_bvp(w, tau):
    """Direct-to-diffuse reflectance and transmittance of one layer from the
    two-stream equations with a direct-beam source (Meador & Weaver 1980):
    d/dtau (F_up, F_dn) = [[g1, -g2], [g2, -g1]] (F_up, F_dn) + w e^-tau (-g3, g4),
    F_dn(0) = 0, F_up(tau) = 0, mu0 = 1, asymmetry 0 (so g1 = (7 - 4w)/4,
    g2 = (4w - 1)/4, g3 = g4 = 1/2), for a unit incident beam; solved with the
    matrix exponential and the particular solution c e^-tau."""
    from scipy.linalg import expm
    g1, g2 = (7 - 4 * w) / 4, (4 * w - 1) / 4
    A = np.array([[g1, -g2], [g2, -g1]])
    b = w * np.array([-0.5, 0.5])
    c = -np.linalg.solve(A + np.eye(2), b)
    E = expm(A * tau)
    e = np.exp(-tau)
    u0 = ((E @ c)[0] - c[0] * e) / E[0, 0]
    y = E @ (np.array([u0, 0.0]) - c) + c * e
    return u0, y[1]


def _sw_closed_form_R(w, tau, alpha2):
    """The Meador-Weaver closed form of the direct reflectance (mu0 = 1,
    asymmetry 0: g3 = g4 = 1/2, lambda = sqrt(3 (1 - w))) for a given alpha2."""
    g1, g2 = (7 - 4 * w) / 4, (4 * w - 1) / 4
    lam = np.sqrt(3 * (1 - w))
    em = np.exp(-lam * tau)
    t1 = w / ((1 - lam ** 2) * (lam + g1 + (lam - g1) * em ** 2))
    return t1 * ((1 - lam) * (alpha2 + lam / 2) - (1 + lam) * (alpha2 - lam / 2) * em ** 2
                 - 2 * lam * (0.5 - alpha2) * em * np.exp(-tau))


def _solve_adding(R, T, Su, Sd, toa, sfc_up, alb):
    """Fluxes of a stack of layers as the solution of the two-stream linear
    system (level 0 = surface)."""
    L = len(R)
    n = L + 1
    A = np.zeros((2 * n, 2 * n))
    b = np.zeros(2 * n)
    r = 0
    for k in range(L):
        A[r, k + 1] = 1; A[r, n + k + 1] = -R[k]; A[r, k] = -T[k]; b[r] = Su[k]; r += 1
        A[r, n + k] = 1; A[r, n + k + 1] = -T[k]; A[r, k] = -R[k]; b[r] = Sd[k]; r += 1
    A[r, n + L] = 1; b[r] = toa; r += 1
    A[r, 0] = 1; A[r, n] = -alb; b[r] = sfc_up
    x = np.linalg.solve(A, b)
    return x[:n], x[n:]


def _sw_expected(g, play, dm, Tl, irr, alpha2_as_reference=True, q6=True):
    """solve_shortwave_column at the reference's settings (asymmetry 0, mu0 = 1,
    surface albedo binary32 0.3), assembled from independent pieces:
    gas optics (pinned against the compiled reference), per-layer direct
    R/T from the two-stream boundary-value problem, diffuse R/T of the
    hemispheric-constant approximation, the beam exp(-sum tau), and the
    fluxes as the solution of the two-stream linear system.  Reference
    properties reproduced on purpose:
      * alpha2 = g2 g3 + g2 g4 (mod_radiation.f90:214) where the two-stream
        solution has g1 g3 + g2 g4: only the direct reflectance changes, by
        the closed form's alpha2 dependence (checked against the BVP);
      * Q6: the diffuse R/T use e^(-lambda tau) of the direct-beam lambda;
      * the diffuse downward flux at the top is the incident beam (264), and
        the direct beam reaching the surface is not reflected (sfc_up 0)."""
    L = len(play)
    R, T, Su, Sd = np.zeros(L), np.zeros(L), np.zeros(L), np.zeros(L)
    beam = np.zeros(L + 1)
    beam[L] = irr
    ksca = O.sw_sca(g)
    for k in range(L - 1, -1, -1):
        kabs, err = O.sw_abs(g, play[k], Tl[k])
        assert err == 0
        tau = (kabs + ksca) * dm[k]
        w = ksca / (kabs + ksca)
        Rd, Td = _sw_two_stream_bvp(w, tau)
        if alpha2_as_reference:  # closed form with the reference's alpha2 = g2
            g2 = (4 * w - 1) / 4
            Rd = Rd + _sw_closed_form_R(w, tau, g2) - _sw_closed_form_R(w, tau, 0.5 * ((7 - 4 * w) / 4 + g2))
        lam_dir = np.sqrt(3 * (1 - w))
        g1d, g2d = 2 - w, w  # backscatter 1/2 (Ritter & Geleyn, asymmetry 0)
        lamd = np.sqrt(g1d ** 2 - g2d ** 2)
        e = np.exp(-(lam_dir if q6 else lamd) * tau)
        den = lamd + g1d + (lamd - g1d) * e ** 2
        R[k] = g2d * (1 - e ** 2) / den
        T[k] = 2 * lamd * e / den
        beam[k] = beam[k + 1] * np.exp(-tau)
        Su[k], Sd[k] = Rd * beam[k + 1], Td * beam[k + 1]
    fu, fd = _solve_adding(R, T, Su, Sd, beam[L], 0.0, float(np.float32(0.3)))
    return fu, fd + beam


def test_two_stream_bvp_matches_the_closed_form():
    """The boundary-value solution above equals the Meador-Weaver closed form
    with alpha2 = g1 g3 + g2 g4 (to rounding), while the reference's
    alpha2 = g2 g3 + g2 g4 changes the reflectance (and only it)."""
    for w, tau in ((0.3, 0.5), (0.9, 2.0), (0.05, 0.01), (0.6, 7.0)):
        g1, g2 = (7 - 4 * w) / 4, (4 * w - 1) / 4
        Rb, _ = _sw_two_stream_bvp(w, tau)
        assert abs(_sw_closed_form_R(w, tau, 0.5 * (g1 + g2)) - Rb) <= 1e-13 * abs(Rb)
        assert abs(_sw_closed_form_R(w, tau, g2) - Rb) > 1e-4 * abs(Rb)


def test_shortwave_column_end_to_end():
    """orc_sw_column against the independently assembled column, every
    g-point, random in-range columns: 1e-12 norm-relative.  The column with
    the two-stream alpha2, or with Q6 'fixed' (the diffuse lambda in the
    exponential), misses it."""
    rng = np.random.default_rng(13)
    L = 10
    plev = np.exp(np.linspace(np.log(1e5), np.log(5e3), L + 1))
    play = np.sqrt(plev[:-1] * plev[1:])
    dm = (plev[:-1] - plev[1:]) / 9.807
    for trial in range(2):
        Tl = np.maximum(285.0 * (play / 1e5) ** 0.19, 215.0) + rng.normal(0, 1.5, L)
        for g in range(1, 17):
            irr = 600.0 * O.solar_fraction(g)
            fu, fd, err = O.sw_column(g, play, dm, Tl, irr)
            assert err == 0
            eu, ed = _sw_expected(g, play, dm, Tl, irr)
            scale = max(np.max(np.abs(eu)), np.max(np.abs(ed)))
            rel = max(np.max(np.abs(fu - eu)), np.max(np.abs(fd - ed))) / scale
            assert rel < 1e-12, (trial, g, rel)
            if trial == 0 and O.sw_sca(g) > 0:
                for kw in (dict(alpha2_as_reference=False), dict(q6=False)):
                    wu, wd = _sw_expected(g, play, dm, Tl, irr, **kw)
                    wrong = max(np.max(np.abs(fu - wu)), np.max(np.abs(fd - wd))) / scale
                    assert wrong > 1e-9, (g, kw, wrong)


# ---- C- and D-grid terms through whole steps ----------------------------------
def _oracle_state(nx, ny, nlay, dx, u, v, pt=None, gzs=None, f=0.0, cdiv=0.15, cvort=0.005, bump=500.0):
    from atmos import driver
    from atmos.config import Config
    from atmos.ic import make_initial_state
    ic = make_initial_state(nx, ny, nlay, dx, bump=bump)
    ic["u"] = np.broadcast_to(u, ic["dp"].shape).copy()
    ic["v"] = np.broadcast_to(v, ic["dp"].shape).copy()
    if pt is not None:
        ic["pt"] = np.broadcast_to(pt, ic["dp"].shape).copy()
    if gzs is not None:
        ic["gzs"] = gzs
    o = O.Oracle(nx, ny, nlay, dx, dx, coriolis=f, divergence_damping=cdiv, vorticity_damping=cvort)
    cfg = Config(dt_max=60.0, t_final=1e12, dt_output=1e12, dt_radiation=1e12, dt_physics=1e12, dt_remap=1e12,
                 radiation_on=False, physics_on=False)
    driver.init_model(o, ic, cfg)
    return o, ic


def _step(o, dt):
    o.cgrid_dynamics_step(dt / 2)
    o.cgrid_halo_exchange()
    o.dgrid_dynamics_step(dt)
    o.dgrid_halo_exchange()


@pytest.mark.parametrize("m", [8, 16])
def test_divergence_damping_is_the_del4_eigenvalue(m):
    """A divergent wave u = A cos(2 pi m x / Lx) (uniform in y and height),
    v = 0, f = 0.  The fourth-order divergence damping (682-702, coefficient
    (c_div dx dy)^2, 1017) enters only the D-grid wind update and only
    through the start state's winds, so switching it off changes u by
    exactly dt (c_div dx dy)^2 lambda^3 u, lambda = -(4/dx^2) sin^2(pi m/nx)
    the 5-point Laplacian's eigenvalue.  The wrong sign, a second-order
    (lambda^2) damping or the coefficient without its square miss it."""
    nx, ny, nlay, dx, dt, A, cdiv = 32, 16, 6, 100e3, 60.0, 10.0, 0.15
    x = np.arange(nx)
    u = A * np.cos(2 * np.pi * m * x / nx)[None, None, :]
    ua = []
    for c in (cdiv, 0.0):
        o, _ = _oracle_state(nx, ny, nlay, dx, u, 0.0, cdiv=c, bump=0.0)
        _step(o, dt)
        ua.append(o.interior(O.UD).copy())
    lam = -(4 / dx ** 2) * np.sin(np.pi * m / nx) ** 2
    divdamp = (cdiv * dx * dx) ** 2
    expect = dt * divdamp * lam ** 3 * np.broadcast_to(u, ua[0].shape)
    got = ua[0] - ua[1]
    np.testing.assert_allclose(got, expect, rtol=1e-5, atol=1e-6 * np.max(np.abs(expect)))
    for wrong in (-expect, dt * divdamp * lam ** 2 * np.broadcast_to(u, got.shape),
                  dt * cdiv * dx * dx * lam ** 3 * np.broadcast_to(u, got.shape)):
        assert np.max(np.abs(got - wrong)) > 0.5 * np.max(np.abs(expect))


def test_vorticity_damping_is_the_del2_eigenvalue():
    """A shear wave u = A sin(2 pi m y / Ly), v = 0, f = 0, uniform dp and pt:
    to first order in A the D-grid u update's vorticity damping
    dt vy (lap(j) - lap(j-1)) (724-738) of the start state's vorticity
    zeta = -(u(j+1) - u(j))/dy is -dt vy dy lambda^2 u, vy = c_vort dx^2
    (1019-1021), lambda = -(4/dy^2) sin^2(pi m/ny); the other terms are
    second order.  The difference of runs with and without the damping
    matches it to the size of the second-order terms; the wrong sign, a
    fourth-order operator or the wrong Laplacian eigenvalue miss it."""
    nx, ny, nlay, dx, dt, A, m, cv = 16, 32, 4, 100e3, 60.0, 1e-4, 8, 0.005
    y = np.arange(ny)
    u = A * np.sin(2 * np.pi * m * y / ny)[None, :, None]
    res = []
    for c in (cv, 0.0):
        o, ic = _oracle_state(nx, ny, nlay, dx, u, 0.0, cvort=c, bump=0.0, pt=None)
        _step(o, dt)
        res.append(o.interior(O.UD).copy())
    lam = -(4 / dx ** 2) * np.sin(np.pi * m / ny) ** 2
    vy = cv * dx * dx
    u_full = np.broadcast_to(u, res[0].shape)
    expect = -dt * vy * dx * lam ** 2 * u_full
    got = res[0] - res[1]
    err = np.max(np.abs(got - expect)) / np.max(np.abs(expect))
    assert err < 1e-3, err
    for wrong in (-expect, expect * lam * dx * dx, expect * 0.5):
        assert np.max(np.abs(got - wrong)) > 0.3 * np.max(np.abs(expect))


def _hydrostatic_offset(o, theta):
    """G = gz + cp theta pkap on each level: for a uniform theta the
    hydrostatic relation gz(k+1) - gz(k) = cp theta (pkap(k) - pkap(k+1))
    makes it the same on every level."""
    gz, pk = o.field(O.GZ), o.field(O.PKAP)
    return gz + CP * theta * pk


def test_pressure_gradient_is_exact_for_a_hydrostatic_tilt():
    """At rest with f = 0, no vorticity damping, a uniform potential
    temperature theta, an uneven surface pressure and an uneven surface
    geopotential, gz is G(x, y) - cp theta pkap with G the same on every
    level.  The finite-volume pressure gradient (contour integral over the
    cell's two faces and two levels) is then exact: the C step's u is
    (dt/2)/dx (G(i-1) - G(i)) (mod_sw_dyn.f90:325-339, 369-382) on every
    level, whatever the pressures.  A pressure gradient from the geopotential
    difference on a level (gz(i-1) - gz(i)) misses it.  (The D step's
    B-grid form cannot be checked this way: its inner advective step
    .5 (q + denom (q - rA (f(i+1) - f(i)))) (470-511) does not keep a
    uniform theta uniform where the C winds diverge.)"""
    nx, ny, nlay, dx, dt, theta = 32, 16, 8, 100e3, 60.0, 30.0
    x, y = np.arange(nx), np.arange(ny)
    gzs = 400.0 * np.cos(2 * np.pi * x / nx)[None, :] + 250.0 * np.sin(4 * np.pi * y / ny)[:, None]
    # vorticity damping off: it also damps dp (167-171, 492-498), which moves
    # pt = dp theta / dpc off the uniform theta the identity needs
    o, ic = _oracle_state(nx, ny, nlay, dx, 0.0, 0.0, pt=theta, gzs=gzs, cvort=0.0)
    o.cgrid_dynamics_step(dt / 2)
    assert np.max(np.abs(o.field(O.PTC)[:, 5:-5, 5:-5] - theta)) <= 4e-15 * theta  # (dp theta) / dp
    G = _hydrostatic_offset(o, theta)
    h = 7
    sl = (slice(h, h + ny), slice(h, h + nx))
    Gc = G[:nlay]
    assert np.max(np.abs(Gc[:, sl[0], sl[1]] - Gc[0, sl[0], sl[1]])) < 1e-9 * np.max(np.abs(Gc[0]))
    uc = o.field(O.UC)[:, sl[0], sl[1]]
    expect = (dt / 2) / dx * (G[0, sl[0], h - 1:h + nx - 1] - G[0, sl[0], sl[1]])
    np.testing.assert_allclose(uc, np.broadcast_to(expect, uc.shape), rtol=0, atol=1e-9 * np.max(np.abs(expect)))
    level_diff = (dt / 2) / dx * (o.field(O.GZ)[:nlay, sl[0], h - 1:h + nx - 1] - o.field(O.GZ)[:nlay, sl[0], sl[1]])
    assert np.max(np.abs(uc - level_diff)) > 0.1 * np.max(np.abs(expect))


def test_corner_interpolation_of_a_wave():
    """interpolate_to_corners (921-956, Q19) of G = A cos(k x + phi), uniform
    in y: the x-then-y pass gives TX/2 and the y-then-x pass 2 (7/12 - 1/12)
    times the 4-point Lagrange value, so the corner value is
    A cos(k (x - dx/2) + phi) [w7 cos t - w1 cos 3t + 2 (w7 - w1)(9/16 cos t
    - 1/16 cos 3t)], t = k dx / 2, with the binary32 7/12 and 1/12 (Q1: a
    constant comes out times 0.99999995).  Exact 7/12 and 1/12 miss it."""
    ni, nj = 40, 12
    for m in (1, 3, 7):
        k = 2 * np.pi * m / ni
        phi = 0.3
        x = np.arange(ni)
        q = np.broadcast_to(2.5 * np.cos(k * x + phi), (nj, ni)).copy()
        out = O.interpolate_to_corners(q)
        t = k / 2

        def corner(w7, w1):
            return 2.5 * np.cos(k * (x - 0.5) + phi) * (w7 * np.cos(t) - w1 * np.cos(3 * t) + 2 * (w7 - w1) * (
                9 / 16 * np.cos(t) - 1 / 16 * np.cos(3 * t)))

        got = out[2:nj - 1, 2:ni - 1]
        np.testing.assert_allclose(got, np.broadcast_to(corner(W7_12, W1_12)[2:ni - 1], got.shape), rtol=0,
                                   atol=1e-14)
        assert np.max(np.abs(got - corner(7 / 12, 1 / 12)[2:ni - 1])) > 1e-8
