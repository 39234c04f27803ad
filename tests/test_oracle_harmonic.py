"""The harmonic Kerr background (SpacetimeMetrics' Harmonic(M, a), unvendored;
restated from Cook & Scheel 1997) in the oracle, pinned by what fixes it: the
coordinates are harmonic (Gamma^a = 0, so H_a = -Gamma_a vanishes with its
derivatives -- a wrong time function or angle breaks this at O(1)); the vacuum
Einstein equations hold (the residual falls 16x per halving of a 4th-order
difference); the horizon R = sqrt(M^2 - a^2) is a null surface (it lies at that
radius on the axis and at M on the equator); far away the metric is the
Kerr-Schild one to O(M/r) -- both are asymptotically Minkowski. The sense of
rotation beyond these invariants is parity-unpinned, as for Kerr-Schild."""
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

M, A = 1.0, 0.7


def desc(a=A):
    return TD.make_desc(background=TD.TGH_BG_HARMONIC, M=M, a=a, periodic=(0, 0, 0))


def state(d, x, t=0.0):
    h, P, dh = np.zeros(10), np.zeros(10), np.zeros(30)
    L.orc_background_state(C.byref(d), t, np.array(x, float), h, P, dh)
    return h, P, dh


def gauge(d, x, t=0.0):
    Hl, dHl = np.zeros(4), np.zeros(16)
    L.orc_gauge_source_grad(C.byref(d), t, np.array(x, float), Hl, dHl)
    return Hl, dHl


def radius(x, a=A):
    """Harmonic R: the Kerr-Schild quartic's root in the harmonic position."""
    rho2 = x[0] ** 2 + x[1] ** 2 + x[2] ** 2
    hb = (rho2 - a * a) / 2
    return np.sqrt(hb + np.sqrt(hb * hb + a * a * x[2] ** 2))


@pytest.mark.parametrize("a", [0.0, 0.5, 0.7])
def test_harmonic_coordinates_have_no_gauge_source(a):
    rng = np.random.default_rng(3)
    d = desc(a)
    for _ in range(25):
        x = rng.uniform(-4, 4, 3)
        if radius(x, a) < 1.2:
            continue
        Hl, dHl = gauge(d, x)
        h, P, dh = state(d, x)
        scale = max(1.0, np.abs(dh).max())
        assert np.abs(Hl).max() < 1e-13 * scale, (a, x, Hl)
        assert np.abs(dHl).max() < 1e-12 * scale, (a, x, dHl)
    # Kerr-Schild's own coordinates are not harmonic: the test can tell
    Hk, _ = gauge(TD.make_desc(background=TD.TGH_BG_KERR_SCHILD, M=M, a=a, periodic=(0, 0, 0)), (2.3, 1.7, -1.1))
    assert np.abs(Hk).max() > 1e-3


@pytest.mark.parametrize("x", [(2.3, 1.7, -1.1), (-1.9, 2.7, 1.3), (0.4, -0.6, 1.9)])
def test_harmonic_kerr_solves_the_vacuum_einstein_equations(x):
    d = desc()
    r1, s = vacuum_residual(d, 0.37, x, 1e-2)
    r2, _ = vacuum_residual(d, 0.37, x, 5e-3)
    e1, e2 = np.abs(r1).max(), np.abs(r2).max()
    assert 12 < e1 / e2 < 20, (x, e1, e2)
    assert e2 < 1e-6 * s


def test_horizon_is_a_null_surface_at_the_harmonic_radius():
    """g^{ij} d_i R d_j R = 0 on R = sqrt(M^2 - a^2): the axis value sqrt(M^2 - a^2)
    and the equatorial value M (src/interior.jl:740-765)."""
    d = desc()
    Rp = np.sqrt(M * M - A * A)
    rng = np.random.default_rng(5)
    for k in range(12):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        # the point on the ray n at harmonic radius Rp: rho^2 = Rp^2 + a^2 (1 - nz^2 ...) by bisection
        lo, hi = 0.1, 3.0
        for _ in range(200):
            mid = (lo + hi) / 2
            lo, hi = (mid, hi) if radius(mid * n) < Rp else (lo, mid)
        x = lo * n
        eps = 1e-6
        grad = np.array([(radius(x + eps * e) - radius(x - eps * e)) / (2 * eps) for e in np.eye(3)])
        h, P, dh = state(d, x)
        gu = oracle_lib.metric_quantities(h)["gu4"]
        q = grad @ gu[1:, 1:] @ grad
        assert abs(q) < 1e-8 * np.abs(gu).max() * (grad @ grad), (x, q)
        # off the horizon the normal is spacelike
        h2, _, _ = state(d, 1.5 * x)
        assert oracle_lib.metric_quantities(h2)["gu4"][1:, 1:].dot(grad).dot(grad) > 1e-3
    assert abs(radius(np.array([0, 0, Rp])) - Rp) < 1e-15
    assert abs(radius(np.array([M, 0, 0])) - Rp) < 1e-15


def test_far_field_agrees_with_kerr_schild_to_order_m_over_r():
    dk = TD.make_desc(background=TD.TGH_BG_KERR_SCHILD, M=M, a=A, periodic=(0, 0, 0))
    for rad in (50.0, 100.0, 200.0):
        x = rad * np.array([0.6, 0.48, 0.64])
        hh, _, _ = state(desc(), x)
        hk, _, _ = state(dk, x)
        assert np.abs(hh - hk).max() < 4 * M / rad, rad
