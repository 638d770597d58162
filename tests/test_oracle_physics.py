"""Pins of the oracle's physics (src/mod_physics.f90) restatement.

* a statically unstable two-layer column mixes completely: both layers end
  at the mass-weighted mean wind (closed form), so u_tend, v_tend follow;
* a stable column with no shear is left alone: zero wind tendencies and a
  potential-temperature tendency at round-off (the energy round trip);
* the mixing sweep conserves column momentum (sum of dp * tendency);
* surface heating: ts advances by dt * net_flux(1) / (41680 * 5);
* apply_physics_tendencies interpolates with the binary32 weights 7/12 and
  1/12 (Q1) along j for u and along i for v: an x<->y transpose of the
  state transposes every result;
* the Richardson-number mixing law for 0 < Ri < 1/4 from its shear
  dependence at fixed stratification (the binary32 shear floor included).
"""
import numpy as np

import oracle as O

N = 8
H = O.HALO


def _column_state(dp, pt, u, v, ptop=1000.0, nx=N, ny=N):
    L = len(dp)
    o = O.Oracle(nx, ny, L, 100e3, 100e3)
    ones = np.ones((L, ny, nx))
    o.put("dp", ones * np.asarray(dp)[:, None, None])
    o.put("pt", ones * np.asarray(pt)[:, None, None])
    o.put("ud", ones * np.asarray(u)[:, None, None])
    o.put("vd", ones * np.asarray(v)[:, None, None])
    o.put("gzs", np.zeros((ny, nx)))
    o.put("ts", np.full((ny, nx), 280.0))
    ak = np.full(L + 1, ptop)
    bk = np.linspace(1.0, 0.0, L + 1)
    o.put("ak", ak)
    o.put("bk", bk)
    o.init_prognostic_fields()
    o.initial_halo_exchange()
    o.allocate_physics_arrays()
    return o


def test_unstable_two_layers_mix_to_mass_weighted_mean():
    dp = np.array([40000.0, 25000.0])
    pt = np.array([320.0, 300.0])  # potential temperature falls with height: unstable
    u, v = np.array([12.0, -3.0]), np.array([1.5, 7.25])
    o = _column_state(dp, pt, u, v)
    o.calculate_physics_tendencies()
    ut, vt, _ = o.physics_column_tendencies(3, 4)
    for w, t in ((u, ut), (v, vt)):
        mean = (dp[0] * w[0] + dp[1] * w[1]) / (dp[0] + dp[1])
        np.testing.assert_allclose(t, (mean - w) / 600.0, rtol=1e-13, atol=1e-16)


def test_stable_column_without_shear_is_untouched():
    L = 6
    dp = np.full(L, 15000.0)
    pt = 300.0 + 10.0 * np.arange(L)  # stable
    o = _column_state(dp, pt, np.full(L, 5.0), np.full(L, -2.0))
    o.calculate_physics_tendencies()
    for j in (1, 4, N):
        ut, vt, tt = o.physics_column_tendencies(2, j)
        assert np.all(ut == 0.0) and np.all(vt == 0.0)
        assert np.max(np.abs(tt)) * 600.0 < 1e-12 * np.max(pt)


def test_mixing_conserves_column_momentum():
    rng = np.random.default_rng(11)
    L = 20
    dp = rng.uniform(2000.0, 8000.0, L)
    pt = 300.0 + np.cumsum(rng.normal(0.0, 3.0, L))  # mixed stability
    u, v = rng.normal(0, 10, L), rng.normal(0, 10, L)
    o = _column_state(dp, pt, u, v)
    o.calculate_physics_tendencies()
    ut, vt, _ = o.physics_column_tendencies(5, 2)
    assert np.any(ut != 0.0)  # some interfaces mixed
    for t, w in ((ut, u), (vt, v)):
        assert abs(np.sum(dp * t)) * 600.0 <= 1e-12 * np.sum(dp * np.abs(w))


def test_surface_heating_closed_form():
    o = _column_state([30000.0, 30000.0], [300.0, 310.0], [0.0, 0.0], [0.0, 0.0])
    nf = np.zeros((3, N, N))
    nf[0] = np.linspace(-50.0, 120.0, N * N).reshape(N, N)
    o.put("net_flux", nf)
    ts0 = o.get("ts")
    o.calculate_physics_tendencies()
    dt = 60.0
    o.apply_physics_tendencies(dt)
    np.testing.assert_array_equal(o.get("ts"), ts0 + dt * (nf[0] / float(41680 * 5)))


def test_apply_weights_binary32_and_transpose_symmetry():
    rng = np.random.default_rng(5)
    L, n = 8, 12
    dp = rng.uniform(5000.0, 15000.0, (L, n, n))
    pt = 300.0 + np.cumsum(rng.normal(0.0, 2.0, (L, n, n)), axis=0)
    u = rng.normal(0, 8, (L, n, n))
    v = rng.normal(0, 8, (L, n, n))

    def run(dp, pt, u, v):
        o = O.Oracle(n, n, L, 100e3, 100e3)
        o.put("dp", dp)
        o.put("pt", pt)
        o.put("ud", u)
        o.put("vd", v)
        o.put("gzs", np.zeros((n, n)))
        o.put("ts", np.full((n, n), 280.0))
        o.put("ak", np.full(L + 1, 1000.0))
        o.put("bk", np.linspace(1.0, 0.0, L + 1))
        o.init_prognostic_fields()
        o.initial_halo_exchange()
        o.allocate_physics_arrays()
        o.calculate_physics_tendencies()
        o.apply_physics_tendencies(30.0)
        return {k: o.get(k) for k in ("dp", "pt", "ud", "vd")}

    # D-grid u(i, j) sits at the y-face j - 1/2 and v(i, j) at the x-face
    # i - 1/2: transposing moves u to v with the same face offsets
    tr = lambda a: np.ascontiguousarray(np.swapaxes(a, 1, 2))
    a = run(dp, pt, u, v)
    b = run(tr(dp), tr(pt), tr(v), tr(u))
    np.testing.assert_array_equal(a["dp"], tr(b["dp"]))
    np.testing.assert_allclose(a["pt"], tr(b["pt"]), rtol=1e-14)
    np.testing.assert_allclose(a["ud"], tr(b["vd"]), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(a["vd"], tr(b["ud"]), rtol=1e-13, atol=1e-13)
    # uniform tendency: the 4-point weights sum to 2 * (7/12 - 1/12) in
    # binary32, 1 - 5.96e-8, not 1
    o = _column_state([40000.0, 25000.0], [320.0, 300.0], [12.0, -3.0], [1.5, 7.25])
    o.calculate_physics_tendencies()
    ut, _, _ = o.physics_column_tendencies(3, 3)
    u0 = o.get("ud")
    o.apply_physics_tendencies(10.0)
    w7, w1 = float(np.float32(7.0 / 12.0)), float(np.float32(1.0 / 12.0))
    expect = u0[:, 3, 3] + 10.0 * (w7 * (ut + ut) - w1 * (ut + ut))
    np.testing.assert_array_equal(o.get("ud")[:, 3, 3], expect)
    assert abs((w7 * 2 - w1 * 2) - 1.0) > 1e-8


def _fraction(du, dpt=1e-8):
    """Mixing fraction f of a stable two-layer column with shear du (u of the
    lower layer du, upper 0): the lower layer's wind tendency is
    -f dp2/(dp1 + dp2) du / 600, the upper's +f dp1/(dp1 + dp2) du / 600."""
    dp = np.array([40000.0, 25000.0])
    o = _column_state(dp, [300.0, 300.0 + dpt], [du, 0.0], [0.0, 0.0])
    o.calculate_physics_tendencies()
    ut, _, _ = o.physics_column_tendencies(3, 4)
    f0 = -ut[0] * 600.0 * (dp[0] + dp[1]) / (dp[1] * du)
    f1 = ut[1] * 600.0 * (dp[0] + dp[1]) / (dp[0] * du)
    return f0, f1


def test_richardson_mixing_law():
    """For 0 < Ri < 1/4 the exchange is (1 - Ri/Ric)^2 dp1 dp2/(dp1 + dp2) times
    the difference (mod_physics.f90:99-111) with Ri proportional to
    1/(du^2 + dv^2 + min_v2): at fixed stratification 1 - sqrt(f) =
    4 A/(du^2 + min_v2) with one A for every shear.  The floor min_v2 is the
    binary32 1e-4 (Q1); a linear law, no floor, or the double 1e-4 give a
    different A at each shear.  Just past the threshold there is no mixing."""
    floor = float(np.float32(1e-4))
    shears = [0.025, 0.03, 0.04, 0.06, 0.1, 0.3]
    s = []
    for du in shears:
        f0, f1 = _fraction(du)
        assert 0.0 < f0 < 1.0 and abs(f0 - f1) <= 1e-12 * f0
        s.append(1.0 - np.sqrt(f0))
    s = np.array(s)
    d2 = np.array(shears) ** 2
    spread = lambda a: float(np.max(np.abs(a / a[0] - 1.0)))  # noqa: E731
    A = s * (d2 + floor) / 4.0
    assert spread(A) < 1e-11, A
    # the alternatives the data rule out
    assert spread(s * (d2 + 1e-4)) > 1e-9                      # the double 1e-4
    assert spread(s * d2) > 1e-3                               # no floor
    lin = np.array([1.0 - _fraction(du)[0] for du in shears]) * (d2 + floor)
    assert spread(lin) > 1e-3                                  # (1 - Ri/Ric) unsquared
    # Ri = 1/4 at du_c^2 = 4 A - min_v2: no exchange just below du_c, some just above
    du_c = np.sqrt(4.0 * A[0] - floor)
    assert _fraction(0.999 * du_c)[0] == 0.0
    assert _fraction(1.001 * du_c)[0] > 0.0


def test_richardson_number_is_the_bulk_richardson_number():
    """The stability factor of the local Richardson number (mod_physics.f90:
    99-104) pinned by its physical meaning.  For two layers in hydrostatic
    balance the bulk Richardson number across their interface is
    (dtheta / theta_m) * g dz / S^2: g dz is the mass-weighted mean of the
    layers' geopotential thicknesses, each cp theta (p_bot^kappa - p_top^kappa)
    for an isentropic layer, theta_m the mass-weighted mean potential
    temperature and S^2 the squared shear plus the binary32 floor.  The
    oracle's mixing fractions at several shears, read through the law
    (1 - Ri/Ric)^2, give A = Ri S^2; it must equal that product from the
    column's pressures alone.  One layer's thickness in place of the mean,
    the lower layer's theta in place of the mean, or the sum of the two
    thicknesses, each miss by > 1e-3."""
    dp = np.array([40000.0, 25000.0])
    ptop, dpt = 1000.0, 3.0
    th = np.array([300.0, 300.0 + dpt])
    kappa = float(np.float32(287.052874)) / 1005.0
    floor = float(np.float32(1e-4))
    p = np.array([ptop + dp[0] + dp[1], ptop + dp[1], ptop])  # surface, interface, top
    dgz = 1005.0 * th * (p[:-1] ** kappa - p[1:] ** kappa)
    A_phys = dpt / ((dp @ th) / dp.sum()) * (dp @ dgz) / dp.sum()

    def frac(du):
        o = _column_state(dp, th, [du, 0.0], [0.0, 0.0], ptop=ptop)
        o.calculate_physics_tendencies()
        ut, _, _ = o.physics_column_tendencies(3, 4)
        return -ut[0] * 600.0 * (dp[0] + dp[1]) / (dp[1] * du)

    du_c = np.sqrt(4.0 * A_phys - floor)  # Ri = 1/4
    A = []
    for du in (1.2 * du_c, 1.5 * du_c, 2.5 * du_c):
        f = frac(du)
        assert 0.0 < f < 1.0
        A.append((1.0 - np.sqrt(f)) * (du * du + floor) / 4.0)
    A = np.array(A)
    assert float(np.max(np.abs(A / A_phys - 1.0))) < 1e-9, (A, A_phys)
    # the threshold sits where the physical bulk Richardson number is 1/4
    assert frac(0.999 * du_c) == 0.0 and frac(1.001 * du_c) > 0.0
    # plausible misreadings of the stability factor
    for wrong in (dpt / ((dp @ th) / dp.sum()) * dgz[0],
                  dpt / th[0] * (dp @ dgz) / dp.sum(),
                  dpt / ((dp @ th) / dp.sum()) * dgz.sum()):
        assert abs(wrong / A_phys - 1.0) > 1e-3
