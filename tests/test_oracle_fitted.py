"""The tracked geometry and the :fitted interior in the oracle
(src/interior.jl:1175-1990, src/evolution.jl:500-700,1244-1272,
src/fit.jl:1050-1055), pinned without the reference's runtime:

* shape_series against scipy's orthonormal Condon–Shortley harmonics in the
  reference's real convention (ỹ_l0 = Y_l0, ỹ^c = √2 Re Y, ỹ^s = −√2 Im Y);
* a FittedInterior holding a sphere is bit for bit the sphere's Interior on a
  right-hand side (the assertion of test/tracking_tests.jl the reference's
  _layer_profiles docstring cites);
* the :fitted core and layer against the :damped ones with the target cache
  holding the exact state (equal in the layer, −ρ_max(u − A) in the core),
  the target's rate entering as (1 − w) S in the layer and S in the core,
  and the trailing-side ramp leaving the leading side untouched.
"""
import numpy as np
import pytest
import scipy.special as sp

import oracle_lib
import tgh_desc as TD


def _ks(N=8, nblocks=4, hw=2.0, a=0.5, r0=0.75, r1=1.25, variant=TD.TGH_INTERIOR_DAMPED, v=(0.0, 0.0, 0.0)):
    d = TD.kerr_schild_hole_case(N=N, nblocks=nblocks, halfwidth=hw, a=a, r0=r0, r1=r1,
                                 gauge=TD.TGH_GAUGE_KERR_SCHILD)
    d.interior = variant
    d.velocity[:] = list(v)
    return d


def _shaped(d, shape, lmax, r_in, r_out, offset, thickness):
    d.shaped, d.shape_lmax = 1, lmax
    for i, a in enumerate(shape):
        d.shape[i] = a
    d.r_in, d.r_out, d.offset, d.thickness = r_in, r_out, offset, thickness
    return d


def _real_harmonics(L, n):
    th, ph = np.arccos(n[2]), np.arctan2(n[1], n[0])
    out = np.zeros((L + 1) ** 2)
    for l in range(L + 1):
        for m in range(l + 1):
            Y = sp.sph_harm_y(l, m, th, ph) if hasattr(sp, "sph_harm_y") else sp.sph_harm(m, l, ph, th)
            if m == 0:
                out[l * l + l] = Y.real
            else:
                out[l * l + l + m] = np.sqrt(2) * Y.real
                out[l * l + l - m] = -np.sqrt(2) * Y.imag
    return out


def test_shape_series_is_the_real_condon_shortley_series(tgh):
    rng = np.random.default_rng(7)
    for L in (0, 1, 3, 8):
        a = rng.normal(size=(L + 1) ** 2)
        for _ in range(20):
            n = rng.normal(size=3)
            n /= np.linalg.norm(n)
            ref = float(a @ _real_harmonics(L, n))
            assert abs(oracle_lib.shape_series(a, L, n) - ref) <= 1e-13 * max(1.0, abs(ref))
            # the package's host copy (shape_bounds samples it) agrees
            assert abs(float(tgh.shape_series(a, L, n)) - ref) <= 1e-13 * max(1.0, abs(ref))
    # the poles: only m = 0 survives, Y_l0(z = ±1) = ±^l sqrt((2l+1)/4pi)
    a = np.zeros(81)
    a[4 * 4 + 4] = 1.0
    assert oracle_lib.shape_series(a, 8, np.array([0.0, 0.0, -1.0])) == pytest.approx(np.sqrt(9 / (4 * np.pi)))


def test_tracked_geometry_radii_and_fast_paths():
    d = _shaped(_ks(), [1.5 * np.sqrt(4 * np.pi)], 0, 1.4, 1.6, 0.25, 0.5)
    # l = 0: r_h = 1.5 along every ray (inside the clamp)
    g = oracle_lib.interior_geo(d, 0.0, np.array([0.3, 0.2, 0.9]))
    assert g[2] == pytest.approx(1.25, abs=1e-15) and g[1] == pytest.approx(0.75, abs=1e-15)
    # outside r_out - offset: the bounding sphere's radii, no series
    g = oracle_lib.interior_geo(d, 0.0, np.array([1.4, 0.0, 0.0]))
    assert tuple(g[1:]) == (1.35 - 0.5, 1.35)
    # inside r_in - offset - thickness: the inner bounding sphere's
    g = oracle_lib.interior_geo(d, 0.0, np.array([0.1, 0.0, 0.0]))
    assert tuple(g[1:]) == ((1.4 - 0.25) - 0.5, 1.4 - 0.25)


def test_fitted_interior_holding_a_sphere_is_the_sphere():
    """A tracked geometry whose clamp pins r_h = 1.5 with offset 1/4 and
    thickness 1/2 is the sphere r_0 = 3/4, r_1 = 5/4, bit for bit."""
    t = 0.0
    sph = _ks()
    shp = _shaped(_ks(), [0.0], 0, 1.5, 1.5, 0.25, 0.5)
    o1, o2 = oracle_lib.OracleProblem(sph), oracle_lib.OracleProblem(shp)
    u = o1.fill_exact(t)
    o1.add_noise(u, 0xF17, 1e-6)
    np.testing.assert_array_equal(o2.fill_exact(t), o1.fill_exact(t))
    np.testing.assert_array_equal(o1.gh_rhs(u, t, 8), o2.gh_rhs(u, t, 8))
    assert o1.error_norms(u, t) == o2.error_norms(u, t)


def _geometry_masks(o, d, t):
    """(core, layer) masks over owned points, from the oracle's geometry."""
    N, P = d.N, d.N + 1
    B = o.block_coords()
    h = (d.hi[0] - d.lo[0]) / (d.nblocks[0] * N)
    core = np.zeros((len(B), P, P, P), bool)
    layer = np.zeros_like(core)
    for b, (bx, by, bz) in enumerate(B):
        for k in range(P):
            for j in range(P):
                for i in range(P):
                    x = np.array([d.lo[0] + (bx * N + i) * h, d.lo[1] + (by * N + j) * h,
                                  d.lo[2] + (bz * N + k) * h])
                    r, r0, r1 = oracle_lib.interior_geo(d, t, x)
                    core[b, k, j, i] = r < r0
                    layer[b, k, j, i] = r0 <= r < r1
    return core, layer


def test_fitted_core_and_layer_against_damped():
    t = 0.1
    rng = np.random.default_rng(3)
    dd = _shaped(_ks(), [1.5 * np.sqrt(4 * np.pi), 0.0, 0.05, 0.02], 1, 1.45, 1.55, 0.25, 0.5)
    df = _shaped(_ks(variant=TD.TGH_INTERIOR_FITTED), [1.5 * np.sqrt(4 * np.pi), 0.0, 0.05, 0.02], 1,
                 1.45, 1.55, 0.25, 0.5)
    od, of = oracle_lib.OracleProblem(dd), oracle_lib.OracleProblem(df)
    A = od.fill_exact(t)
    u = A.copy()
    od.add_noise(u, 0xF18, 1e-6)
    nb = len(A) // (20 * 9 ** 3)
    tw = np.concatenate([A.reshape(nb, 20, -1), np.zeros((nb, 20, 9 ** 3))], axis=1).ravel()
    of.set_target(tw, t)
    rd, rf = od.gh_rhs(u, t, 8), of.gh_rhs(u, t, 8)
    core, layer = _geometry_masks(od, dd, t)
    assert core.any() and layer.any()
    V = lambda x: x.reshape(nb, 20, 9, 9, 9)
    ev = ~core
    np.testing.assert_array_equal(V(rf).transpose(1, 0, 2, 3, 4)[:, ev], V(rd).transpose(1, 0, 2, 3, 4)[:, ev])
    rho = dd.rho_max
    np.testing.assert_array_equal(V(rf).transpose(1, 0, 2, 3, 4)[:, core],
                                  0.0 - rho * (V(u) - V(A)).transpose(1, 0, 2, 3, 4)[:, core])
    # the target's rate: S in the core, (1 - w) S in the layer, nothing outside
    S = rng.normal(size=A.shape)
    tw2 = np.concatenate([A.reshape(nb, 20, -1), S.reshape(nb, 20, -1)], axis=1).ravel()
    of.set_target(tw2, t, rate=True)
    rr = of.gh_rhs(u, t, 8)
    diff = V(rr - rf).transpose(1, 0, 2, 3, 4)
    Sv = V(S).transpose(1, 0, 2, 3, 4)
    np.testing.assert_allclose(diff[:, core], Sv[:, core], rtol=0, atol=1e-15)
    out = ~(core | layer)
    assert np.all(diff[:, out] == 0.0)
    ratio = diff[:, layer] / Sv[:, layer]
    assert np.all(ratio >= -1e-12) and np.all(ratio <= 1 + 1e-12)
    assert np.ptp(ratio, axis=0).max() <= 1e-9  # one 1 - w per point


def test_trailing_ramp_narrows_only_behind_the_hole():
    t = 0.0
    v = (0.3, 0.0, 0.0)
    df = _shaped(_ks(variant=TD.TGH_INTERIOR_FITTED, v=v), [1.5 * np.sqrt(4 * np.pi)], 0, 1.4, 1.6, 0.25, 0.5)
    o = oracle_lib.OracleProblem(df)
    A = o.fill_exact(t)
    u = A.copy()
    o.add_noise(u, 0xF19, 1e-6)
    nb = len(A) // (20 * 9 ** 3)
    tw = np.concatenate([A.reshape(nb, 20, -1), np.zeros((nb, 20, 9 ** 3))], axis=1).ravel()
    o.set_target(tw, t, trail=0.0)
    r0 = o.gh_rhs(u, t, 8)
    o.set_target(tw, t, trail=0.5)
    r1 = o.gh_rhs(u, t, 8)
    B = o.block_coords()
    N, h = df.N, (df.hi[0] - df.lo[0]) / (df.nblocks[0] * df.N)
    d = (r1 - r0).reshape(nb, 20, 9, 9, 9)
    i = np.arange(9)
    lead = trail_side = 0
    for b, (bx, by, bz) in enumerate(B):
        x = df.lo[0] + (bx * N + i) * h
        lead_mask = (x > 1e-12)[None, None, :] * np.ones((9, 9, 9), bool)
        assert np.all(d[b][:, lead_mask] == 0.0)   # ahead of the hole: unchanged
        lead += 1
        trail_side += int(np.any(d[b][:, ~lead_mask] != 0.0))
    assert trail_side > 0
