"""The sense of rotation of the oracle's Kerr charts: "spin along +z, J = +Ma"
(the convention the device restates). SpacetimeMetrics' own convention is not
vendored and stays unpinned; what is pinned here is the oracle's claim, by the
frame-dragging rate omega = -g_tphi / g_phiphi on the equator, which for a
hole whose angular momentum is +Ma along z is positive and tends to the
Lense-Thirring value 2 J / r^3 far away -- in both charts (Kerr-Schild and
harmonic), with the sign following a's. A chart whose spin entered with the
wrong sign (a flipped in l_x, l_y, or in the harmonic map) fails it."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
import tgh_desc as TD

L = oracle_lib.lib()
dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
L.orc_dmetric.argtypes = [C.c_void_p, C.c_double, dp, dp, dp]


def metric(d, x):
    g, dg = np.zeros(16), np.zeros(64)
    L.orc_dmetric(C.byref(d), 0.0, np.array(x, float), g, dg)
    return g.reshape(4, 4)


def omega(d, x, y, z=0.0):
    """-g_tphi / g_phiphi at (x, y, z) with d/dphi = -y d/dx + x d/dy."""
    g = metric(d, (x, y, z))
    e = np.array([0.0, -y, x, 0.0])
    gtphi = g[0] @ e
    gpp = e @ g @ e
    return -gtphi / gpp


@pytest.mark.parametrize("bg", [TD.TGH_BG_KERR_SCHILD, TD.TGH_BG_HARMONIC], ids=["kerr_schild", "harmonic"])
@pytest.mark.parametrize("a", [0.7, -0.5])
def test_frame_dragging_follows_the_spin_along_z(bg, a):
    gauge = TD.TGH_GAUGE_NONE if bg == TD.TGH_BG_HARMONIC else TD.TGH_GAUGE_KERR_SCHILD
    d = TD.make_desc(background=bg, gauge=gauge, M=1.0, a=a, periodic=(0, 0, 0))
    # every direction on the equator, near and far: the sign of a
    for r in (3.0, 6.0, 12.0):
        for ph in np.linspace(0.0, 2 * np.pi, 7, endpoint=False):
            w = omega(d, r * np.cos(ph), r * np.sin(ph))
            assert np.sign(w) == np.sign(a), (r, ph, w)
    # Lense-Thirring far away: omega r^3 / (2 M a) -> 1 with O(M / r) corrections
    ratios = []
    for r in (50.0, 100.0, 200.0):
        ratios.append(omega(d, r, 0.0) * r ** 3 / (2.0 * a))
    print(f"bg {bg} a {a}: omega r^3 / 2Ma = {ratios}")
    assert abs(ratios[-1] - 1.0) < 0.03
    assert abs(ratios[-1] - 1.0) < abs(ratios[0] - 1.0) or abs(ratios[0] - 1.0) < 1e-3


def test_a_flipped_spin_is_told_apart():
    """The same check run on the chart with a and -a: the rates are opposite,
    so a sign error anywhere in the spin terms cannot pass the test above."""
    dp_ = TD.make_desc(background=TD.TGH_BG_KERR_SCHILD, gauge=TD.TGH_GAUGE_KERR_SCHILD, M=1.0, a=0.7,
                       periodic=(0, 0, 0))
    dm = TD.make_desc(background=TD.TGH_BG_KERR_SCHILD, gauge=TD.TGH_GAUGE_KERR_SCHILD, M=1.0, a=-0.7,
                      periodic=(0, 0, 0))
    for x, y in ((4.0, 0.0), (0.0, 5.0), (-3.0, 2.0)):
        wp, wm = omega(dp_, x, y), omega(dm, x, y)
        assert wp > 0 > wm
        assert abs(wp + wm) <= 1e-12 * abs(wp)
