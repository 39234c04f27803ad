"""The oracle's pointwise algebra, pinned by mathematics rather than by a
retyped formula: inverse-metric and ADM identities, offset precision against
50-digit arithmetic, the chain rule against a forward-mode dual pass, the
vacuum Einstein equations on exact backgrounds (the residual of
d_t Pi - d_i F^i - msrc falls 16x per halving of a 4th-order difference,
test/pointwise_identity_tests.jl:164, CODE.md "The equations"), and the two
spellings of the source against each other (test/pointwise_identity_tests.jl:78-217).
"""
import ctypes as C

import mpmath
import numpy as np
import pytest

import oracle_lib as O
import tgh_desc as TD

L = O.lib()
PAIRS = [(0, 0), (1, 0), (2, 0), (3, 0), (1, 1), (2, 1), (3, 1), (2, 2), (3, 2), (3, 3)]
ETA = np.diag([-1.0, 1.0, 1.0, 1.0])


def sym4(v):
    m = np.zeros((4, 4))
    for n, (a, b) in enumerate(PAIRS):
        m[a, b] = m[b, a] = v[n]
    return m


def random_state(rng, amp=0.3):
    h = rng.uniform(-amp, amp, 10)
    return h


def test_metric_quantities_identities():
    rng = np.random.default_rng(7)
    for _ in range(50):
        h = random_state(rng)
        m = O.metric_quantities(h)
        g = ETA + sym4(h)
        gi = np.linalg.inv(g)
        np.testing.assert_allclose(m["g4"], g, atol=0)
        np.testing.assert_allclose(m["gu4"], gi, rtol=1e-13, atol=1e-14)
        np.testing.assert_allclose(m["alpha"], 1 / np.sqrt(-gi[0, 0]), rtol=1e-14)
        np.testing.assert_allclose(m["beta"], -gi[0, 1:] / gi[0, 0], rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(m["gamu"], np.linalg.inv(g[1:, 1:]), rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(m["sqrtg"], np.sqrt(np.linalg.det(g[1:, 1:])), rtol=1e-14)


def test_offsets_keep_relative_precision_at_tiny_h():
    """pointwise.jl:145-184: g^{ab} − η^{ab}, det(γ) − 1 are formed as offsets."""
    mpmath.mp.dps = 50
    rng = np.random.default_rng(3)
    for _ in range(10):
        h = 1e-13 * rng.uniform(-1, 1, 10)
        m = O.metric_quantities(h)
        g = mpmath.matrix(4, 4)
        H = sym4(h)
        for a in range(4):
            for b in range(4):
                g[a, b] = mpmath.mpf(ETA[a, b]) + mpmath.mpf(H[a, b])
        gi = g ** -1
        for a in range(4):
            for b in range(4):
                exact = gi[a, b] - ETA[a, b]
                if abs(exact) > 1e-20:
                    assert abs((m["guo"][a, b] - exact) / exact) < 1e-13
        gam = mpmath.matrix([[g[i, j] for j in range(1, 4)] for i in range(1, 4)])
        detm1 = mpmath.det(gam) - 1
        sq_m1 = mpmath.sqrt(1 + detm1) - 1
        # sqrt(gamma) itself is within an ulp of the 50-digit value
        assert abs(m["sqrtg"] - float(1 + sq_m1)) <= 2.3e-16


def test_metric_derivatives_are_the_chain_rule():
    """The closed forms (pointwise.jl:462-500) against a forward-mode dual pass
    through metric_quantities (test/pointwise_tests.jl:186)."""
    rng = np.random.default_rng(11)
    for _ in range(40):
        h = random_state(rng)
        dh = rng.uniform(-1, 1, 30)
        md = np.zeros(39)
        L.orc_metric_derivatives(h, dh, md)
        dalpha, dbeta, dA = md[:3], md[3:12].reshape(3, 3), md[12:].reshape(3, 3, 3)
        m = O.metric_quantities(h)
        for i in range(3):
            jvp = np.zeros(14)
            L.orc_metric_quantities_jvp(h, np.ascontiguousarray(dh[10 * i:10 * i + 10]), jvp)
            da, db, ds, dg = jvp[0], jvp[1:4], jvp[4], jvp[5:14].reshape(3, 3)
            scale = max(1.0, np.abs(dh).max())
            assert abs(dalpha[i] - da) <= 1e-13 * scale
            assert np.abs(dbeta[i] - db).max() <= 1e-13 * scale
            dAi = da * m["sqrtg"] * m["gamu"] + m["alpha"] * ds * m["gamu"] + m["alpha"] * m["sqrtg"] * dg
            assert np.abs(dA[i] - dAi).max() <= 1e-12 * scale
        div = np.zeros(4)
        L.orc_metric_divergences(h, dh, div)
        assert abs(div[0] - np.trace(dbeta)) <= 1e-13 * np.abs(dbeta).max()
        for j in range(3):
            ref = dA[0, 0, j] + dA[1, 1, j] + dA[2, 2, j]
            assert abs(div[1 + j] - ref) <= 1e-13 * max(1.0, np.abs(dA).max())


def test_lean_source_is_the_reference_spelling():
    """gh_node_source_lean vs gh_node_source on 200 random states
    (CODE.md:6315-6316: 8e-16 relative)."""
    rng = np.random.default_rng(5)
    worst = 0.0
    for _ in range(200):
        h = random_state(rng)
        dth = rng.uniform(-1, 1, 10)
        dh = rng.uniform(-1, 1, 30)
        Hl = rng.uniform(-1, 1, 4)
        dHl = rng.uniform(-1, 1, 16)
        g0, g2 = rng.uniform(0, 2), rng.uniform(-0.5, 0.5)
        dg = np.zeros((4, 4, 4))
        dg[0] = sym4(dth)
        for i in range(3):
            dg[i + 1] = sym4(dh[10 * i:10 * i + 10])
        ref = np.zeros(10)
        lean = np.zeros(10)
        L.orc_gh_node_source(h, np.ascontiguousarray(dg.ravel()), Hl, dHl, g0, g2, ref)
        L.orc_gh_node_source_lean(h, dth, dh, Hl, dHl, g0, g2, lean)
        worst = max(worst, np.abs(lean - ref).max() / max(1.0, np.abs(ref).max()))
    assert worst < 1e-14, worst


def analytic(desc, t, x):
    h, P, dh = np.zeros(10), np.zeros(10), np.zeros(30)
    L.orc_background_state(C.byref(desc), t, np.ascontiguousarray(x, dtype=float), h, P, dh)
    return h, P, dh


def gauge(desc, t, x):
    Hl, dHl = np.zeros(4), np.zeros(16)
    L.orc_gauge_source_grad(C.byref(desc), t, np.ascontiguousarray(x, dtype=float), Hl, dHl)
    return Hl, dHl


def flux_and_source(desc, t, x, g0, g2):
    h, P, dh = analytic(desc, t, x)
    Hl, dHl = gauge(desc, t, x)
    dtg, F, msrc = np.zeros(10), np.zeros(30), np.zeros(10)
    L.orc_gh_node_rhs_flux(h, P, dh, Hl, dHl, g0, g2, dtg, F, msrc)
    return F.reshape(3, 10), msrc, P


def vacuum_residual(desc, t, x, delta, g0=1.0, g2=0.0):
    """d_t Pi - d_i F^i - msrc with 4th-order central differences of step delta."""
    c = [(-2, 1 / 12), (-1, -2 / 3), (1, 2 / 3), (2, -1 / 12)]
    divF = np.zeros(10)
    for i in range(3):
        for o, w in c:
            xs = np.array(x, dtype=float)
            xs[i] += o * delta
            F, _, _ = flux_and_source(desc, t, xs, g0, g2)
            divF += w * F[i] / delta
    dtP = np.zeros(10)
    for o, w in c:
        _, _, P = flux_and_source(desc, t + o * delta, x, g0, g2)
        dtP += w * P / delta
    _, msrc, _ = flux_and_source(desc, t, x, g0, g2)
    return dtP - divF - msrc, max(np.abs(divF).max(), np.abs(msrc).max(), np.abs(dtP).max())


@pytest.mark.parametrize("name,desc,points", [
    ("kerr-schild a=0.7", TD.kerr_schild_hole_case(a=0.7, gauge=TD.TGH_GAUGE_SAMPLED),
     [(2.3, 1.7, -1.1), (-1.9, 2.7, 1.3)]),
    ("kerr-schild a=0", TD.kerr_schild_hole_case(a=0.0, gauge=TD.TGH_GAUGE_SAMPLED),
     [(3.1, -0.7, 0.4)]),
    ("gauge wave", TD.gauge_wave_case(A=0.1, d=1.0), [(0.3, 0.2, 0.1)]),
    ("shifted minkowski", TD.make_desc(background=TD.TGH_BG_SHIFTED_MINKOWSKI, wave_amp=0.5, wave_len=2.0),
     [(0.7, 0.1, -0.3)]),
])
def test_source_solves_the_vacuum_einstein_equations(name, desc, points):
    """The reduced source and the damping are the vacuum Einstein equation in
    this gauge: on exact data the residual is the difference's truncation
    error and falls 16x per halving (CODE.md, "The equations")."""
    for x in points:
        r1, s = vacuum_residual(desc, 0.37, x, 1e-2)
        r2, _ = vacuum_residual(desc, 0.37, x, 5e-3)
        e1, e2 = np.abs(r1).max(), np.abs(r2).max()
        if e1 < 1e-11 * s:  # exact to roundoff (e.g. the gauge wave in x - t)
            continue
        ratio = e1 / e2
        assert 12 < ratio < 20, (name, x, e1, e2, ratio)
        assert e2 < 1e-6 * s


def test_constraint_damping_vanishes_on_exact_data():
    """Z_ab = 0 where C_a = Γ_a + H_a = 0 (pointwise.jl:382-398)."""
    desc = TD.kerr_schild_hole_case(a=0.7, gauge=TD.TGH_GAUGE_SAMPLED)
    for x in [(2.3, 1.7, -1.1), (0.2, -1.8, 2.2)]:
        h, P, dh = analytic(desc, 0.0, x)
        Hl, dHl = gauge(desc, 0.0, x)
        m = O.metric_quantities(h)
        dtg = m["beta"][0] * dh[:10] + m["beta"][1] * dh[10:20] + m["beta"][2] * dh[20:] + \
            (m["alpha"] / m["sqrtg"]) * P
        a, b = np.zeros(10), np.zeros(10)
        L.orc_gh_node_source_lean(h, dtg, dh, Hl, dHl, 0.0, 0.0, a)
        L.orc_gh_node_source_lean(h, dtg, dh, Hl, dHl, 1.0, -0.5, b)
        assert np.abs(a - b).max() <= 64 * np.finfo(float).eps * max(1, np.abs(a).max())


def test_kerr_schild_closed_form_source_is_minus_gamma():
    """KerrSchildSource (gauge.jl:293-326) against H_a = −Γ_a of the metric by
    a nested dual pass, and ∂_c H_a likewise."""
    desc = TD.kerr_schild_hole_case(a=0.7)
    rng = np.random.default_rng(1)
    for _ in range(20):
        x = rng.uniform(-4, 4, 3)
        if np.linalg.norm(x) < 1.5:
            continue
        h, P, dh = analytic(desc, 0.0, x)
        Hl, dHl = gauge(desc, 0.0, x)
        Hk, dHk = np.zeros(4), np.zeros(16)
        L.orc_ks_algebraic_source(1.0, np.array([0, 0, 0, 0.7]), np.array([1.0, 0, 0, 0]), h,
                                  np.zeros(10), dh, Hk, dHk)
        assert np.abs(Hl - Hk).max() <= 1e-14 * np.abs(Hl).max()
        assert np.abs(dHl - dHk).max() <= 1e-13 * np.abs(dHl).max()


def test_harmonic_backgrounds_have_no_gauge_source():
    """isharmonic (gauge.jl:62-104): Minkowski and the gauge wave are harmonic."""
    for desc in (TD.gauge_wave_case(A=0.1), TD.make_desc()):
        Hl, dHl = gauge(desc, 0.3, (0.31, 0.2, 0.7))
        assert np.abs(Hl).max() < 1e-15 and np.abs(dHl).max() < 1e-14


def test_matter_source_assembles_the_stress_energy():
    """matter_source (pointwise.jl:667-698) against T_ab = ρ n_a n_b + 2 S_(a n_b)
    + S_ab built from the 4-metric and the unit normal."""
    rng = np.random.default_rng(2)
    for _ in range(10):
        h = random_state(rng, 0.2)
        rho = rng.uniform(0, 1)
        S = rng.uniform(-1, 1, 3)
        Sij = sym4(np.concatenate([[0, 0, 0, 0], rng.uniform(-1, 1, 6)]))[1:, 1:]
        out = np.zeros(10)
        L.orc_matter_source(h, rho, S, np.array([Sij[0, 0], Sij[1, 0], Sij[2, 0], Sij[1, 1], Sij[2, 1], Sij[2, 2]]), out)
        m = O.metric_quantities(h)
        g = m["g4"]
        alpha, beta = m["alpha"], m["beta"]
        n_low = np.array([-alpha, 0, 0, 0])
        # projector gamma_a^i for covariant spatial tensors: S_a = (β·S, S_i)
        S4 = np.concatenate([[beta @ S], S])
        Sab = np.zeros((4, 4))
        Sab[1:, 1:] = Sij
        Sab[0, 1:] = Sab[1:, 0] = Sij @ beta
        Sab[0, 0] = beta @ Sij @ beta
        T = rho * np.outer(n_low, n_low) + np.outer(n_low, S4) + np.outer(S4, n_low) + Sab
        gu = np.linalg.inv(g)
        trT = np.einsum("ab,ab->", gu, T)
        ref = 16 * np.pi * alpha * m["sqrtg"] * (T - g * trT / 2)
        expect = np.array([ref[a, b] for a, b in PAIRS])
        assert np.abs(out - expect).max() <= 1e-12 * max(1, np.abs(expect).max())
