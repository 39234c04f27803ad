"""The moving hole in the oracle: translate(boost(bg, -v), c0) (SpacetimeMetrics'
boost, unvendored; the convention the reference states at src/interior.jl:66-72
and src/initialdata.jl:93-96), the moving core's ADM relaxation
(src/interior.jl:450-503) and the targets' rates (src/initialdata.jl:1697-1764).

Pinned by what fixes them: the boosted metric at a point riding with the hole
does not change (the hole moves at +v, not -v); it still solves the vacuum
equations; a boost of the harmonic chart is harmonic; the reference's own
closed-form KerrSchildSource with a velocity (src/gauge.jl:263-278: u = gamma
(1, v), S the boosted spin) is -Gamma_a of the boosted Kerr-Schild metric, which
a wrong boost sign or a wrong spin transformation breaks at O(v); the ADM
relaxation reduces to the target's rate at h = h_t and to -rho (h - h_t) to
first order; the rates match finer differences."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
import tgh_desc as TD
from test_oracle_pointwise import vacuum_residual

L = oracle_lib.lib()
dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
L.orc_background_state.argtypes = [C.c_void_p, C.c_double, dp, dp, dp, dp]
L.orc_gauge_source_grad.argtypes = [C.c_void_p, C.c_double, dp, dp, dp]
L.orc_dmetric.argtypes = [C.c_void_p, C.c_double, dp, dp, dp]
L.orc_background_state_rate.argtypes = [C.c_void_p, C.c_double, dp, dp, dp, dp, dp]
L.orc_core_target_rate.argtypes = [C.c_void_p, C.c_double, dp, dp, dp]
L.orc_core_relaxation_rate.argtypes = [dp, dp, C.c_double, dp, dp]
L.orc_layer_rate_step.argtypes = [C.c_void_p]
L.orc_layer_rate_step.restype = C.c_double
L.orc_ks_algebraic_source.argtypes = [C.c_double, dp, dp, dp, dp, dp, dp, dp]

V = (0.3, 0.0, 0.0)
C0 = (0.5, -0.25, 0.0)


def moving(bg, a, v=V, c0=C0, **kw):
    gauge = TD.TGH_GAUGE_NONE if bg == TD.TGH_BG_HARMONIC else TD.TGH_GAUGE_KERR_SCHILD
    return TD.make_desc(background=bg, gauge=gauge, M=1.0, a=a, periodic=(0, 0, 0), center=c0, velocity=v,
                        **kw)


def state(d, t, x):
    h, P, dh = np.zeros(10), np.zeros(10), np.zeros(30)
    L.orc_background_state(C.byref(d), t, np.array(x, float), h, P, dh)
    return h, P, dh.reshape(3, 10)


def dmetric(d, t, x):
    g, dg = np.zeros(16), np.zeros(64)
    L.orc_dmetric(C.byref(d), t, np.array(x, float), g, dg)
    return g.reshape(4, 4), dg.reshape(4, 4, 4)


CHARTS = [(TD.TGH_BG_KERR_SCHILD, 0.7), (TD.TGH_BG_KERR_SCHILD, 0.0), (TD.TGH_BG_HARMONIC, 0.5)]


@pytest.mark.parametrize("bg,a", CHARTS)
def test_the_boosted_hole_moves_at_v(bg, a):
    d = moving(bg, a)
    v = np.array(V)
    for off in [(2.1, 0.7, -0.4), (-1.3, 2.2, 0.9), (0.2, -0.3, 2.5)]:
        x0 = np.array(C0) + off
        g0, dg0 = dmetric(d, 0.0, x0)
        for t in (0.37, 1.9):
            g, dg = dmetric(d, t, x0 + v * t)
            assert np.abs(g - g0).max() < 1e-13
            # and the comoving derivative vanishes: d_t g + v.grad g = 0
            adv = dg[:, :, 0] + dg[:, :, 1] * v[0] + dg[:, :, 2] * v[1] + dg[:, :, 3] * v[2]
            assert np.abs(adv).max() < 1e-13 * max(1.0, np.abs(dg).max())
    # a hole moving the other way would not pass: the point that rides at -v changes
    g, _ = dmetric(d, 1.9, np.array(C0) + (2.1, 0.7, -0.4) - v * 1.9)
    g0, _ = dmetric(d, 0.0, np.array(C0) + (2.1, 0.7, -0.4))
    assert np.abs(g - g0).max() > 1e-3


@pytest.mark.parametrize("bg,a", CHARTS)
def test_boosted_backgrounds_solve_the_vacuum_equations(bg, a):
    d = moving(bg, a)
    for x in [(2.6, 1.3, -0.9), (-1.7, 2.4, 1.1)]:
        r1, s = vacuum_residual(d, 0.37, x, 1e-2)
        r2, _ = vacuum_residual(d, 0.37, x, 5e-3)
        e1, e2 = np.abs(r1).max(), np.abs(r2).max()
        assert 12 < e1 / e2 < 20, (x, e1, e2)
        assert e2 < 1e-6 * s


def test_a_boost_of_the_harmonic_chart_is_harmonic():
    d = moving(TD.TGH_BG_HARMONIC, 0.5, v=(0.2, -0.1, 0.25))
    for x in [(2.6, 1.3, -0.9), (-1.7, 2.4, 1.1), (0.4, -0.3, 2.8)]:
        Hl, dHl = np.zeros(4), np.zeros(16)
        L.orc_gauge_source_grad(C.byref(d), 0.8, np.array(x, float), Hl, dHl)
        h, P, dh = state(d, 0.8, x)
        assert np.abs(Hl).max() < 1e-13 * max(1.0, np.abs(dh).max())
        assert np.abs(dHl).max() < 1e-12 * max(1.0, np.abs(dh).max())


@pytest.mark.parametrize("v", [(0.3, 0.0, 0.0), (0.0, 0.0, 0.4), (0.2, -0.15, 0.1)])
def test_moving_kerr_schild_source_is_minus_gamma_of_the_boosted_metric(v):
    """KerrSchildSource(T; M, spin = (0, 0, a), velocity = v) (src/gauge.jl:263-278),
    evaluated on the boosted state, is the boosted metric's own -Gamma_a."""
    a = 0.7
    d = moving(TD.TGH_BG_KERR_SCHILD, a, v=v)
    vv = np.array(v)
    v2 = vv @ vv
    gam = 1 / np.sqrt(1 - v2)
    av = np.array([0.0, 0.0, a])
    va = vv @ av
    S = np.concatenate([[gam * va], av + ((gam - 1) * va / v2) * vv])
    u = np.concatenate([[gam], gam * vv])
    for x in [(2.6, 1.3, -0.9), (-1.7, 2.4, 1.1), (0.4, -2.3, 2.8)]:
        x = np.array(C0) + x
        t = 0.6
        g, dg = dmetric(d, t, x)
        h = np.array([g[i, j] - (-1.0 if i == j == 0 else float(i == j))
                      for i, j in [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]])
        pk = [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]
        dth = np.array([dg[i, j, 0] for i, j in pk])
        dh = np.array([[dg[i, j, c] for i, j in pk] for c in (1, 2, 3)]).ravel()
        Hc, dHc = np.zeros(4), np.zeros(16)
        L.orc_ks_algebraic_source(1.0, S, u, h, dth, dh, Hc, dHc)
        Hn, dHn = np.zeros(4), np.zeros(16)
        L.orc_gauge_source_grad(C.byref(d), t, x, Hn, dHn)
        assert np.abs(Hc - Hn).max() < 1e-13, (x, Hc, Hn)
        assert np.abs(dHc - dHn).max() < 1e-12 * max(1.0, np.abs(dHn).max())
        # the unboosted spin would not match a motion along the spin, nor would the
        # source of a hole moving the other way
        if v[2] != 0:
            Hs, dHs = np.zeros(4), np.zeros(16)
            L.orc_ks_algebraic_source(1.0, np.array([0, 0, 0, a]), u, h, dth, dh, Hs, dHs)
            assert np.abs(Hs - Hn).max() > 1e-4
        um = np.concatenate([[gam], -gam * vv])
        Hs, dHs = np.zeros(4), np.zeros(16)
        L.orc_ks_algebraic_source(1.0, S, um, h, dth, dh, Hs, dHs)
        assert np.abs(Hs - Hn).max() > 1e-3


def test_core_relaxation_rate_limits():
    rng = np.random.default_rng(7)
    d = moving(TD.TGH_BG_KERR_SCHILD, 0.7)
    ht, _, _ = state(d, 0.3, np.array(C0) + (0.9, 0.2, 0.1))
    dht = rng.normal(size=10) * 0.3
    out = np.zeros(10)
    # at the target the ADM rate maps back to the target's own rate
    L.orc_core_relaxation_rate(ht, ht, 7.0, dht, out)
    assert np.abs(out - dht).max() < 1e-13 * np.abs(dht).max()
    # without a target rate, near the target: -rho (h - h_t) to first order
    k = rng.normal(size=10)
    for eps in (1e-4, 1e-5):
        L.orc_core_relaxation_rate(ht + eps * k, ht, 3.0, np.zeros(10), out)
        rel = np.abs(out + 3.0 * eps * k).max() / (3.0 * eps * np.abs(k).max())
        assert rel < 200 * eps
    # and it is not that in h far from the target (the ADM path, not the straight line)
    L.orc_core_relaxation_rate(ht + 0.3 * k, ht, 3.0, np.zeros(10), out)
    assert np.abs(out + 3.0 * 0.3 * k).max() > 1e-3


@pytest.mark.parametrize("bg,a", CHARTS)
def test_background_state_rate_matches_finer_differences(bg, a):
    d = moving(bg, a, r0=0.71, r1=1.21, interior=TD.TGH_INTERIOR_DAMPED, rho_max=4.0)
    delta = L.orc_layer_rate_step(C.byref(d))
    assert delta == np.cbrt(np.finfo(float).eps) * 0.71
    x = np.array(C0) + (1.1, 0.6, -0.3)
    t = 0.4
    h, P, dth, dtP = (np.zeros(10) for _ in range(4))
    L.orc_background_state_rate(C.byref(d), t, x, h, P, dth, dtP)
    h0, P0, _ = state(d, t, x)
    assert np.array_equal(h, h0) and np.array_equal(P, P0)

    def cd(f, e):
        return (f(t + e) - f(t - e)) / (2 * e)

    def rich(f):
        return (4 * cd(f, 5e-4) - cd(f, 1e-3)) / 3

    dth_ref = rich(lambda s: state(d, s, x)[0])
    dtP_ref = rich(lambda s: state(d, s, x)[1])
    assert np.abs(dth - dth_ref).max() < 1e-9 * max(1.0, np.abs(dth_ref).max())
    assert np.abs(dtP - dtP_ref).max() < 1e-8 * max(1.0, np.abs(dtP_ref).max())


def test_core_target_rate():
    d = moving(TD.TGH_BG_KERR_SCHILD, 0.7, r0=0.8, r1=1.3, interior=TD.TGH_INTERIOR_DAMPED, rho_max=4.0)
    t = 0.5
    c = np.array(C0) + np.array(V) * t
    dth, dtP = np.zeros(10), np.zeros(10)
    # within r_0/4 of the center: no rate, no pass
    L.orc_core_target_rate(C.byref(d), t, c + (0.1, 0.05, 0.0), dth, dtP)
    assert not dth.any() and not dtP.any()
    # in the outer half: the weighted rate of u_exact(t, c(t) + r_0 n(t))
    x = c + (0.0, 0.6, 0.2)
    L.orc_core_target_rate(C.byref(d), t, x, dth, dtP)
    r = np.linalg.norm(x - c)
    s = L.orc_smoothstep(4 * r / 0.8 - 1)
    assert s == 1.0

    def target(tt):
        cc = np.array(C0) + np.array(V) * tt
        n = (x - cc) / np.linalg.norm(x - cc)
        return state(d, tt, cc + 0.8 * n)

    e = 1e-4
    ref_h = (target(t + e)[0] - target(t - e)[0]) / (2 * e)
    ref_P = (target(t + e)[1] - target(t - e)[1]) / (2 * e)
    assert np.abs(dth - ref_h).max() < 1e-6 * max(1.0, np.abs(ref_h).max())
    assert np.abs(dtP - ref_P).max() < 1e-5 * max(1.0, np.abs(ref_P).max())
    # a hole at rest: the target does not move
    d0 = moving(TD.TGH_BG_KERR_SCHILD, 0.7, v=(0.0, 0.0, 0.0), r0=0.8, r1=1.3, interior=TD.TGH_INTERIOR_DAMPED,
                rho_max=4.0)
    L.orc_core_target_rate(C.byref(d0), t, np.array(C0) + (0.0, 0.6, 0.2), dth, dtP)
    assert not dth.any() and not dtP.any()
