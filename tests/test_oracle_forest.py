"""Frozen hierarchies: the forest builders, the oracle's ghost fill on levels
and the interface-order rule (test/interface_tests.jl), on the CPU.

TreeAMR's prolongation is not vendored with the reference, so the oracle's
fill is pinned here by what fixes it without the source: the midpoint
Lagrange weights (textbook values), exact reproduction of polynomials of
degree p - 1 per axis (and failure at degree p), every mesh point owned by
exactly one primary of the finest leaf holding it, and the reference's own
measured interface rates (rate 3.98 at p = 6, 3.18 at p = 4 on the two-level
gauge-wave mesh, test/interface_tests.jl:64-96; the brackets asserted there).
"""
import ctypes
import math

import numpy as np
import pytest

import oracle_lib
import tgh_desc as TD

IFACE_Q = 4
IFACE_NS = (8, 10, 12)
IFACE_T_END = 1 / 16


def _gw(tgh):
    return tgh.gauge_wave_case(eps_ko=0.0, gamma0=1.0, gamma2=0.0)


def test_two_level_forest_is_the_frozen_hierarchy(tgh):
    """test/interface_tests.jl:150-165: 15 leaves, one level, the same layout
    at every N, the finest spacing 1 / (4N), and the unrefined control 8."""
    case = _gw(tgh)
    forests = [tgh.gh_forest(None, case, N=N, roots=2, refined=True) for N in IFACE_NS]
    assert all(f.nleaves() == 15 for f in forests)
    assert all(f.maxlevel() == 1 for f in forests)
    assert all(f.leaves == forests[0].leaves for f in forests)
    assert forests[0].isbalanced()
    assert [f.minimum_spacing() for f in forests] == pytest.approx([1 / (2 * 2 * N) for N in IFACE_NS])
    assert tgh.gh_forest(None, case, N=8, roots=2).nleaves() == 8
    # the refined root is the one holding (3/8, 3/8, 3/8)
    assert sorted(lf for lf in forests[0].leaves if lf[0] == 1) == sorted(
        (1, i, j, k) for i in (0, 1) for j in (0, 1) for k in (0, 1))


def test_balance_refines_until_neighbours_differ_by_one_level(tgh):
    """balance!: a leaf two levels finer than a neighbour (across a corner
    and a periodic seam too) refines the neighbour."""
    case = _gw(tgh)
    f = tgh.gh_forest(None, case, N=8, roots=2)
    f.refine([(0, 0, 0, 0)])
    f.refine([(1, 0, 0, 0)])  # level 2 at the box's corner: touches (0,1,1,1) across seams
    assert not f.isbalanced()
    f.balance()
    assert f.isbalanced()
    # every root touches the corner leaf across the periodic seams, so all refine
    assert all(lf[0] >= 1 for lf in f.leaves)
    with pytest.raises(tgh.ArgumentError):
        f.refine([(0, 0, 0, 0)])


def test_hole_forest_nests_shells(tgh):
    """hole_forest (src/initialdata.jl:1469-1515): shells refine the leaves
    meeting them, level by level; the radii must nest."""
    case = tgh.kerr_schild_case(M=1.0, a=0.5, halfwidth=4.0, r_0=0.6, r_1=1.1)
    f = tgh.hole_forest(None, case, N=8, roots=4, radii=(2.5, 1.0))
    assert f.maxlevel() == 2 and f.isbalanced()
    lev2 = [lf for lf in f.leaves if lf[0] == 2]
    assert lev2
    for lf in lev2:  # each one's parent meets the ball of radius 1 about the center
        ext = f.block_extent((1, lf[1] >> 1, lf[2] >> 1, lf[3] >> 1))
        d2 = sum(max(ext[d][0], 0.0, -ext[d][1]) ** 2 for d in range(3))
        assert d2 <= 1.0
    # the cube shell of the same radius refines no fewer leaves
    fc = tgh.hole_forest(None, case, N=8, roots=4, radii=(2.5, 1.0), shape="cube")
    assert fc.nleaves() >= f.nleaves()
    # a swept shell along +x covers the segment's end
    fp = tgh.hole_forest(None, case, N=8, roots=4, radii=(1.0,), path=(2.0, 0.0, 0.0))
    assert any(fp.block_extent(lf)[0][0] <= 2.0 <= fp.block_extent(lf)[0][1] and lf[0] == 1 for lf in fp.leaves)
    with pytest.raises(tgh.ArgumentError):
        tgh.hole_forest(None, case, N=8, roots=4, radii=(1.0, 2.0))
    with pytest.raises(tgh.ArgumentError):
        tgh.hole_forest(None, case, N=8, roots=4, radii=(1.0,), shape="sphere")


@pytest.mark.parametrize("p,ref", [(2, [1, 1]), (4, [-1, 9, 9, -1]), (6, [3, -25, 150, 150, -25, 3]),
                                   (8, [-5, 49, -245, 1225, 1225, -245, 49, -5])])
def test_midpoint_weights_are_the_textbook_ones(tgh, p, ref):
    case = _gw(tgh)
    q = max(4, p - 2)
    f = tgh.gh_forest(None, case, N=2 * (q // 2 + 1), roots=2, refined=True)
    o = oracle_lib.OracleProblem(TD.gauge_wave_case(N=f.N, roots=2, q=q), f.leaves, p)
    den = {2: 2, 4: 16, 6: 256, 8: 2048}[p]
    np.testing.assert_array_equal(o.prolong_weights(p) * den, np.array(ref, dtype=float))


def _owned_positions(o, f):
    """Finest-level integer coordinates of every owned point, per block."""
    L = f.maxlevel()
    N = f.N
    out = []
    for b, lf in enumerate(f.leaves):
        s = 1 << (L - lf[0])
        out.append([((lf[1] * N + i) * s, (lf[2] * N + j) * s, (lf[3] * N + k) * s, b, i, j, k)
                    for k in range(N + 1) for j in range(N + 1) for i in range(N + 1)])
    return out


def test_every_point_has_one_primary_in_its_finest_leaf(tgh):
    case = _gw(tgh)
    f = tgh.hole_forest(None, case, N=8, roots=2, center=(0.3, 0.3, 0.3), radii=(0.2,))
    o = oracle_lib.OracleProblem(TD.gauge_wave_case(N=8, roots=2, q=4), f.leaves, 6)
    L = f.maxlevel()
    tot = 2 * 8 << L
    owners = {}
    levels = [lf[0] for lf in f.leaves]
    finest = {}
    for blk in _owned_positions(o, f):
        for (X, Y, Z, b, i, j, k) in blk:
            key = (X % tot, Y % tot, Z % tot)
            finest[key] = max(finest.get(key, -1), levels[b])
            if lib_primary(o, b, i, j, k):
                assert key not in owners, (key, owners.get(key), b)
                owners[key] = b
    assert set(owners) == set(finest)
    for key, b in owners.items():
        assert levels[b] == finest[key]


def lib_primary(o, b, i, j, k):
    return oracle_lib.lib().orc_is_primary(o.p, b, i, j, k) == 1


def _poly_state(o, f, deg):
    """Each variable v a product over axes of (x_d + a_vd)^deg."""
    u = o.zeros()
    N = f.N
    P = N + 1
    h0 = f.root_spacing()
    for b, lf in enumerate(f.leaves):
        h = h0 / 2 ** lf[0]
        idx = np.arange(P)
        x = (lf[1] * N + idx) * h
        y = (lf[2] * N + idx) * h
        z = (lf[3] * N + idx) * h
        for v in range(20):
            a = (0.1 + 0.03 * v, -0.2 + 0.05 * v, 0.3 - 0.02 * v)
            val = ((x[None, None, :] + a[0]) ** deg * (y[None, :, None] + a[1]) ** deg *
                   (z[:, None, None] + a[2]) ** deg)
            u[(b * 20 + v) * P ** 3:(b * 20 + v + 1) * P ** 3] = val.ravel()
    return u


@pytest.mark.parametrize("p", [4, 6])
def test_prolongation_reproduces_polynomials_of_degree_p_minus_1(tgh, p):
    """The ghosts that prolongation fills, on the fine blocks of a refined
    root in the middle of a 4^3 brick (no periodic seam in reach), take a
    per-axis polynomial of degree p - 1 exactly and miss one of degree p."""
    case = _gw(tgh)
    N = 8
    f = tgh.gh_forest(None, case, N=N, roots=4)
    f.refine([(0, 1, 1, 1)])
    f.balance()
    o = oracle_lib.OracleProblem(TD.gauge_wave_case(N=N, roots=4, q=4), f.leaves, p)
    G, S = 3, N + 7
    h0 = f.root_spacing()
    errs = {}
    for deg in (p - 1, p):
        o.prepare(_poly_state(o, f, deg), 0.0)
        W = o.work()
        worst = 0.0
        for b, lf in enumerate(f.leaves):
            if lf[0] != 1:
                continue
            h = h0 / 2
            for k in range(-G, N + G + 1):
                for j in range(-G, N + G + 1):
                    for i in range(-G, N + G + 1):
                        kind, _ = o.classify(b, i, j, k)
                        if kind != 3:
                            continue
                        xyz = ((lf[1] * N + i) * h, (lf[2] * N + j) * h, (lf[3] * N + k) * h)
                        for v in (0, 7, 19):
                            a = (0.1 + 0.03 * v, -0.2 + 0.05 * v, 0.3 - 0.02 * v)
                            exact = ((xyz[0] + a[0]) ** deg * (xyz[1] + a[1]) ** deg * (xyz[2] + a[2]) ** deg)
                            got = W[(b * 20 + v) * S ** 3 + (i + G) + S * ((j + G) + S * (k + G))]
                            worst = max(worst, abs(got - exact) / max(1.0, abs(exact)))
        errs[deg] = worst
    assert errs[p - 1] < 1e-13, errs
    assert errs[p] > 1e-9, errs


def _rates(tgh, p, refined, eps_ko=0.0):
    case = tgh.gauge_wave_case(eps_ko=eps_ko, gamma0=1.0, gamma2=0.0)
    hs, l2, li = [], [], []
    for N in IFACE_NS:
        f = tgh.gh_forest(None, case, N=N, roots=2, refined=refined)
        o = oracle_lib.OracleProblem(TD.gauge_wave_case(N=N, roots=2, q=IFACE_Q, eps_ko=eps_ko), f.leaves, p)
        u = o.fill_exact(0.0)
        dt = 0.25 * f.minimum_spacing() / o.max_speed(u, 0.0)
        n = max(1, math.ceil(IFACE_T_END / dt))
        dt = IFACE_T_END / n
        for s in range(n):
            o.rk4_step(u, s * dt, dt, nthreads=8)
        e2, ei = o.error_norms(u, IFACE_T_END)
        hs.append(f.minimum_spacing())
        l2.append(e2)
        li.append(ei)
    x, y, z = np.log(hs), np.log(l2), np.log(li)
    rate = lambda a, b: float(((a - a.mean()) * (b - b.mean())).sum() / ((a - a.mean()) ** 2).sum())
    return dict(l2=l2, linf=li, rate_l2=rate(x, y), rate_linf=rate(x, z))


def test_interface_order_rule_on_the_oracle(tgh):
    """test/interface_tests.jl:64-96: at q = 4 the two-level mesh converges at
    about 3 with an order-4 prolongation and at 4 with an order-6 one."""
    r4 = _rates(tgh, 4, True)
    r6 = _rates(tgh, 6, True)
    assert 2.5 <= r4["rate_l2"] <= 3.6 and 2.5 <= r4["rate_linf"] <= 3.6, r4
    assert r6["rate_l2"] >= 3.75 and r6["rate_linf"] >= 3.75, r6
    assert r4["l2"] == sorted(r4["l2"], reverse=True)
    assert r6["l2"] == sorted(r6["l2"], reverse=True)
    assert all(a < b for a, b in zip(r6["l2"], r4["l2"]))


def test_unrefined_control_never_prolongates(tgh):
    """test/interface_tests.jl:120-135: with no coarse-fine face the operator
    order is the same computation, bit for bit, and the rate is 4."""
    c4 = _rates(tgh, 4, False)
    c6 = _rates(tgh, 6, False)
    assert c4["rate_l2"] >= 3.75
    assert c4["l2"] == c6["l2"]
