"""The fitted target's evaluator and the target cache's fill in the oracle
(src/fit.jl:182-275,699-755,1057-1151), pinned without the reference's
runtime:

* the fit's solid harmonics are the shape series' real Condon–Shortley
  harmonics at unit vectors (the reference's own check, test/fit_tests.jl:
  100-117, and scipy's orthonormal Y_lm in the reference's real convention),
  scale as ρ^l, and at the center only ỹ_00 survives;
* the radial orders and a moving center against a numpy evaluation of the
  ansatz f_v = Σ S_lm(ξ) Σ_k C ρ^2k from scipy's harmonics;
* state_from_fit inverts the fit variables of a numpy restatement of
  fit_variables (plain and Π̃), the reference's round trip (test/fit_tests.jl:
  260-290), on Kerr–Schild states to 16 eps of the state;
* fit_state is state_from_fit then bounds_project into the fit's bounds: the
  identity where nothing fires, the projection's own result where it does;
* a least-squares fit of the a = 0.7 Kerr–Schild hole (built here in numpy
  from scipy's harmonics) reproduces the analytic state in its shell;
* fill_target!'s continuation A = F_a + (t − t_a) S, S = κ(F_a − F_b) on two
  constant fits, the fill sphere, clear and the untouched points, a fit at
  rest whose rate variant is bit for bit the plain one, and a moving fit
  whose slope is the centred difference of fills a little earlier and later
  (test/moving_tests.jl:120-170).
"""
import numpy as np
import pytest

import oracle_lib
import tgh_desc as TD
from test_oracle_fitted import _real_harmonics

EPS = np.finfo(np.float64).eps


def _solid(L, xi):
    """S_lm(ξ) = |ξ|^l ỹ_lm(ξ/|ξ|) from scipy's harmonics."""
    r = float(np.linalg.norm(xi))
    if r == 0.0:
        out = np.zeros((L + 1) ** 2)
        out[0] = 1.0 / np.sqrt(4 * np.pi)
        return out
    y = _real_harmonics(L, np.asarray(xi) / r)
    ls = np.array([int(np.sqrt(s)) for s in range((L + 1) ** 2)])
    return y * r ** ls


def _slots_at(L, x, rbar=1.0, center=(0.0, 0.0, 0.0)):
    """Every S_lm at x through the oracle's evaluator: twenty slots at a time,
    each one-hot in its own variable."""
    NS = (L + 1) ** 2
    out = np.zeros(NS)
    for s0 in range(0, NS, 20):
        C = np.zeros((20, 1, NS))
        for v in range(min(20, NS - s0)):
            C[v, 0, s0 + v] = 1.0
        f = TD.make_fit(C, lmax=L, cont=0, rbar=rbar, center=center)
        vals = oracle_lib.fit_variables_at(f, x, 0.0)
        n = min(20, NS - s0)
        out[s0:s0 + n] = vals[:n]
    return out


def test_solid_harmonics_are_the_shape_series_and_scale_as_rho_l():
    rng = np.random.default_rng(20260923)
    for L in (4, 8):
        worst_shape = worst_scale = worst_scipy = 0.0
        ls = np.array([int(np.sqrt(s)) for s in range((L + 1) ** 2)])
        for _ in range(12):
            n = rng.normal(size=3)
            n /= np.linalg.norm(n)
            c = rng.normal(size=3)
            S1 = _slots_at(L, c + 0.75 * n, rbar=0.75, center=c)   # ξ = n
            S2 = _slots_at(L, c + 1.5 * n, rbar=0.75, center=c)    # ξ = 2n
            for s in range(len(S1)):
                e = np.zeros((L + 1) ** 2)
                e[s] = 1.0
                worst_shape = max(worst_shape, abs(S1[s] - oracle_lib.shape_series(e, L, n)))
            worst_scale = max(worst_scale, float(np.abs(S2 / 2.0 ** ls - S1).max()))
            worst_scipy = max(worst_scipy, float(np.abs(S1 - _real_harmonics(L, n)).max()))
        print(f"L = {L}: shape series {worst_shape:.2e}, rho^l {worst_scale:.2e}, scipy {worst_scipy:.2e}")
        # the reference's bar (test/fit_tests.jl:115): 32 eps
        assert worst_shape <= 32 * EPS and worst_scale <= 32 * EPS
        assert worst_scipy <= 1e-13
    S0 = _slots_at(8, np.array([0.3, -0.2, 0.1]), center=(0.3, -0.2, 0.1))
    assert S0[0] == pytest.approx(1 / np.sqrt(4 * np.pi), rel=4 * EPS)
    assert np.all(S0[1:] == 0.0)


def test_radial_orders_and_the_moving_center():
    rng = np.random.default_rng(5)
    L, cont, rbar = 6, 2, 1.3
    NS = (L + 1) ** 2
    C = rng.normal(size=(20, cont + 1, NS)) / (1 + np.arange(NS))
    c0, v = np.array([0.2, -0.1, 0.05]), np.array([0.3, -0.2, 0.1])
    f = TD.make_fit(C, lmax=L, cont=cont, rbar=rbar, center=c0, velocity=v)
    worst = 0.0
    for _ in range(30):
        x, t = rng.uniform(-1.5, 1.5, size=3), float(rng.uniform(-2, 2))
        xi = (x - (c0 + t * v)) / rbar
        S = _solid(L, xi)
        r2 = xi @ xi
        ref = np.einsum("s,vks,k->v", S, C, np.array([1.0, r2, r2 * r2]))
        scale = np.einsum("s,vks,k->v", np.abs(S), np.abs(C), np.array([1.0, r2, r2 * r2]))
        got = oracle_lib.fit_variables_at(f, x, t)
        worst = max(worst, float((np.abs(got - ref) / scale).max()))
    print(f"the ansatz against scipy's harmonics: {worst:.2e} of the terms' size")
    assert worst <= 1e-13


def _fit_variables(u, tilde=False):
    """fit_variables (src/fit.jl:114-169, the values only), restated in numpy:
    (log α, β^i, γ_ij − δ_ij, Π or (α/√γ) Π) of the packed state u."""
    h, P = u[:10], u[10:]
    g = np.array([[1 + h[4], h[5], h[6]], [h[5], 1 + h[7], h[8]], [h[6], h[8], 1 + h[9]]])
    b = h[1:4]
    beta = np.linalg.solve(g, b)
    a2 = (1 - h[0]) + b @ beta
    s = np.sqrt(a2 / np.linalg.det(g)) if tilde else 1.0
    return np.concatenate([[np.log(a2) / 2], beta, h[4:10], s * P])


def _ks_states(n, rng, a=0.7, rlo=1.0, rhi=1.6):
    d = TD.kerr_schild_hole_case(N=8, nblocks=4, halfwidth=2.0, a=a)
    out, xs = [], []
    for _ in range(n):
        m = rng.normal(size=3)
        x = m / np.linalg.norm(m) * rng.uniform(rlo, rhi)
        e = np.zeros(20)
        oracle_lib.lib().orc_case_state(oracle_lib.C.byref(d), 0.0, np.ascontiguousarray(x), e)
        out.append(e)
        xs.append(x)
    return np.array(xs), np.array(out)


def test_state_from_fit_inverts_fit_variables():
    rng = np.random.default_rng(11)
    _, U = _ks_states(200, rng)
    for tilde in (False, True):
        worst = 0.0
        for u in U:
            h, P = oracle_lib.state_from_fit(_fit_variables(u, tilde), tilde)
            worst = max(worst, float(np.abs(np.concatenate([h, P]) - u).max() / np.abs(u).max()))
        print(f"tilde = {tilde}: round trip {worst:.2e} of the state")
        assert worst <= 16 * EPS
    # a sample that is not a metric (det γ < 0) has no √det γ: Π̃ comes back NaN
    v = np.zeros(20)
    v[4] = -2.0
    v[10:] = 1.0
    h, P = oracle_lib.state_from_fit(v, True)
    assert np.all(np.isnan(P)) and np.all(np.isfinite(h))


def _const_fit(vals, t=0.0, tilde=False, bounds=TD.WIDE_BOUNDS, velocity=(0.0, 0.0, 0.0)):
    """A fit whose twenty variables are the constants vals everywhere."""
    C = np.zeros((20, 1, 1))
    C[:, 0, 0] = np.asarray(vals) * np.sqrt(4 * np.pi)
    return TD.make_fit(C, lmax=0, cont=0, t=t, tilde=tilde, bounds=bounds, velocity=velocity)


def test_fit_state_projects_into_the_bounds():
    rng = np.random.default_rng(3)
    _, U = _ks_states(5, rng)
    x = np.array([0.4, 0.1, -0.2])
    for u in U:
        v = _fit_variables(u, True)
        f = _const_fit(v, tilde=True)
        h, P = oracle_lib.fit_state(f, x, 0.0)
        h0, P0 = oracle_lib.state_from_fit(oracle_lib.fit_variables_at(f, x, 0.0), True)
        assert np.array_equal(h, h0) and np.array_equal(P, P0)  # nothing fires: the identity
        # a lapse of 1/100 under a floor of 1/50: the projection's own result
        v2 = v.copy()
        v2[0] = np.log(0.01)
        bnd = (0.02, 50.0, 0.01, 1000.0, 10.0, 100.0)
        f2 = _const_fit(v2, tilde=True, bounds=bnd)
        h2, P2 = oracle_lib.fit_state(f2, x, 0.0)
        hs, Ps = oracle_lib.state_from_fit(oracle_lib.fit_variables_at(f2, x, 0.0), True)
        hp, Pp, hit, _ = oracle_lib.bounds_project(hs, Ps, bnd + (1.0,))
        assert hit and np.array_equal(h2, hp) and np.array_equal(P2, Pp)
        assert not np.array_equal(h2, hs)


def _ks_fit(L=8, cont=2, rbar=2.0, n=1500, seed=17, tilde=True):
    """A least-squares fit of the a = 0.7 hole's fit variables in the thin
    shell 1.9 ≤ |x| ≤ 2.1 about its horizon (equatorial radius 1.85), built
    here from scipy's harmonics: the coefficients (20, cont+1, NS)."""
    rng = np.random.default_rng(seed)
    xs, U = _ks_states(n, rng, rlo=1.9, rhi=2.1)
    NS = (L + 1) ** 2
    X = np.zeros((n, NS * (cont + 1)))
    for p, x in enumerate(xs):
        xi = x / rbar
        S = _solid(L, xi)
        r2 = xi @ xi
        for k in range(cont + 1):
            X[p, k * NS:(k + 1) * NS] = S * r2 ** k
    Y = np.array([_fit_variables(u, tilde) for u in U])
    Csol, *_ = np.linalg.lstsq(X, Y, rcond=None)
    return Csol.T.reshape(20, cont + 1, NS)


def test_a_fit_of_the_hole_reproduces_it():
    L, cont, rbar = 8, 2, 2.0
    C = _ks_fit(L, cont, rbar)
    bnd = (0.02, 50.0, 0.01, 1000.0, 10.0, 100.0)
    f = TD.make_fit(C, lmax=L, cont=cont, rbar=rbar, tilde=True, bounds=bnd)
    rng = np.random.default_rng(99)
    xs, U = _ks_states(300, rng, rlo=1.91, rhi=2.09)
    worst = 0.0
    for x, u in zip(xs, U):
        h, P = oracle_lib.fit_state(f, x, 0.0)
        worst = max(worst, float(np.abs(np.concatenate([h, P]) - u).max() / np.abs(u).max()))
    print(f"the a = 0.7 hole by a fit of degree 8 (cont 2) in its shell: {worst:.2e} of the state")
    # measured 9e-4 (the least squares' own error); a wrong sign or slot in
    # the harmonics, a dropped √det γ or a wrong radial power is off by the
    # data's size
    assert worst <= 4e-3


def _smooth_fit_coeffs(L, cont, seed=8, amp=0.05):
    """Coefficients of a near-flat fit that varies in space: every variable
    small, so its states are metrics the bounds leave alone."""
    rng = np.random.default_rng(seed)
    NS = (L + 1) ** 2
    return amp * rng.normal(size=(20, cont + 1, NS)) / (1.0 + np.arange(NS))


def _fitted_problem(N=16, nb=4, hw=2.0, r_h=1.5, offset=0.2, thickness=0.5, v=(0.0, 0.0, 0.0)):
    d = TD.kerr_schild_hole_case(N=N, nblocks=nb, halfwidth=hw, a=0.5, r0=0.75, r1=1.25,
                                 gauge=TD.TGH_GAUGE_KERR_SCHILD)
    d.velocity[:] = list(v)
    d.shaped, d.shape_lmax = 1, 0
    d.shape[0] = r_h * np.sqrt(4 * np.pi)
    d.r_in, d.r_out, d.offset, d.thickness = r_h, r_h, offset, thickness
    return d, oracle_lib.OracleProblem(d)


def _positions(d, o):
    N, P = d.N, d.N + 1
    h = (d.hi[0] - d.lo[0]) / (d.nblocks[0] * N)
    coords = o.block_coords()
    i = np.arange(P)
    X = []
    for B in coords:
        zz, yy, xx = np.meshgrid(i, i, i, indexing="ij")
        X.append(np.stack([d.lo[0] + (B[0] * N + xx) * h, d.lo[1] + (B[1] * N + yy) * h,
                           d.lo[2] + (B[2] * N + zz) * h], axis=-1).reshape(-1, 3))
    return np.array(X), h  # (nblocks, P^3, 3)


def test_fill_target_continuation_sphere_and_clear():
    d, o = _fitted_problem()
    X, h = _positions(d, o)
    nb, P3 = X.shape[0], X.shape[1]
    rng = np.random.default_rng(21)
    _, U = _ks_states(2, rng)
    va, vb = _fit_variables(U[0], True), _fit_variables(U[1], True)
    fa, fb = _const_fit(va, t=1.0, tilde=True), _const_fit(vb, t=0.5, tilde=True)
    t = 1.2
    out = o.fill_target(np.full(2 * o.n, 7.0), fa, fb, t, h, rate=False, clear=True)
    W = out.reshape(nb, 40, P3)
    Ha, Pa = oracle_lib.fit_state(fa, np.zeros(3), t)
    Hb, Pb = oracle_lib.fit_state(fb, np.zeros(3), t)
    Fa, Fb = np.concatenate([Ha, Pa]), np.concatenate([Hb, Pb])
    kap = 1.0 / (1.0 - 0.5)
    S = kap * (Fa - Fb)
    A = Fa + (t - 1.0) * S
    inside = np.linalg.norm(X, axis=-1) < (d.r_out - d.offset) + h
    assert 0 < inside.sum() < inside.size
    for v in range(20):
        assert np.array_equal(W[:, v][inside], np.full(inside.sum(), A[v]))
        assert np.array_equal(W[:, 20 + v][inside], np.full(inside.sum(), S[v]))
    assert np.all(W.transpose(0, 2, 1)[~inside] == 0.0)
    # without clear the points outside keep what they held
    out2 = o.fill_target(np.full(2 * o.n, 7.0), fa, fb, t, h, rate=False, clear=False)
    W2 = out2.reshape(nb, 40, P3)
    assert np.all(W2.transpose(0, 2, 1)[~inside] == 7.0)
    assert np.array_equal(W2.transpose(0, 2, 1)[inside], W.transpose(0, 2, 1)[inside])
    # a previous fit no older than the latest does not count: S = 0, A = F_a
    out3 = o.fill_target(np.zeros(2 * o.n), fa, _const_fit(vb, t=1.0, tilde=True), t, h)
    W3 = out3.reshape(nb, 40, P3)
    for v in range(20):
        assert np.all(W3[:, v][inside] == Fa[v]) and np.all(W3[:, 20 + v][inside] == 0.0)


def test_fill_target_rate_at_rest_and_moving():
    d, o = _fitted_problem()
    X, h = _positions(d, o)
    nb, P3 = X.shape[0], X.shape[1]
    L, cont, rbar = 4, 1, 1.5
    C = _smooth_fit_coeffs(L, cont)
    bnd = TD.WIDE_BOUNDS
    rest = TD.make_fit(C, lmax=L, cont=cont, rbar=rbar, tilde=True, bounds=bnd)
    plain = o.fill_target(np.zeros(2 * o.n), rest, None, 0.0, h, rate=False)
    rated = o.fill_target(np.zeros(2 * o.n), rest, None, 0.0, h, rate=True)
    assert np.array_equal(plain, rated)  # a fit at rest: no rate, bit for bit
    mv = TD.make_fit(C, lmax=L, cont=cont, rbar=rbar, tilde=True, bounds=bnd, velocity=(0.3, 0.0, 0.0))
    dl = 1e-3
    cr = o.fill_target(np.zeros(2 * o.n), mv, None, 0.0, h, rate=True).reshape(nb, 40, P3)
    cp = o.fill_target(np.zeros(2 * o.n), mv, None, dl, h).reshape(nb, 40, P3)
    cm = o.fill_target(np.zeros(2 * o.n), mv, None, -dl, h).reshape(nb, 40, P3)
    diff = (cp[:, :20] - cm[:, :20]) / (2 * dl)
    inside = np.linalg.norm(X, axis=-1) < (d.r_out - d.offset) + h
    sel = np.broadcast_to(inside[:, None, :], diff.shape)
    worst = float(np.abs(cr[:, 20:] - diff)[sel].max())
    scale = float(np.abs(diff)[sel].max())
    print(f"moving fit: slope against the fills' centred difference {worst:.2e} of {scale:.2e}")
    # the two differences have steps of h/8 / 0.3 and 1e-3 in time; they agree
    # to the first one's truncation (the reference's bar, 1e-3 of the scale)
    assert scale > 1e-2 and worst <= 1e-3 * scale
    # the values are the plain fill's
    assert np.array_equal(cr[:, :20], o.fill_target(np.zeros(2 * o.n), mv, None, 0.0, h).reshape(nb, 40, P3)[:, :20])
