"""The regrid's state transfer (refine!/coarsen! and the state carried over,
src/refinement.jl:973-1021) in the oracle, on the CPU.

TreeAMR's transfer is not vendored with the reference, so the oracle's is
pinned by what fixes it without the source: onto the same forest it is the
identity (away from periodic seams, where a non-periodic test polynomial
has two values) and idempotent; coarsening is injection, bit for bit; on
leaves one level finer it reproduces a per-axis polynomial of degree p - 1
exactly and misses one of degree p (the order-p prolongation).
"""
import numpy as np
import pytest

import oracle_lib
import tgh_desc as TD
from test_oracle_forest import _poly_state

N = 8
ROOTS = 4


def _forests(tgh):
    case = tgh.gauge_wave_case(eps_ko=0.0, gamma0=1.0, gamma2=0.0)
    f0 = tgh.gh_forest(None, case, N=N, roots=ROOTS)
    f0.refine([(0, 1, 1, 1)])
    f0.balance()
    f1 = tgh.gh_forest(None, case, N=N, roots=ROOTS)
    f1.refine([(0, 1, 1, 1)])
    f1.balance()
    f1.refine([(1, 2, 2, 2)])
    f1.balance()
    return f0, f1


def _off_seams(f):
    """Per state entry: True where the point is on no periodic seam."""
    P = N + 1
    idx = np.arange(P)
    out = []
    for lf in f.leaves:
        tot = ROOTS * N << lf[0]
        on = np.zeros((P, P, P), bool)
        for d, ax in ((1, 2), (2, 1), (3, 0)):
            g = lf[d] * N + idx
            sh = [1, 1, 1]
            sh[ax] = P
            on |= ((g == 0) | (g == tot)).reshape(sh)
        out.append(np.tile(~on.ravel(), 20))
    return np.concatenate(out)


def _pair(f0, f1, p):
    d = TD.gauge_wave_case(N=N, roots=ROOTS, q=4)
    return oracle_lib.OracleProblem(d, f0.leaves, p), oracle_lib.OracleProblem(d, f1.leaves, p)


def test_forests_differ_by_one_level(tgh):
    f0, f1 = _forests(tgh)
    assert f0.maxlevel() == 1 and f1.maxlevel() == 2
    assert f1.isbalanced() and f1.nleaves() > f0.nleaves()
    assert (1, 2, 2, 2) in f0.leaves and (1, 2, 2, 2) not in f1.leaves


@pytest.mark.parametrize("p", [4, 6])
def test_transfer_onto_the_same_forest_is_the_identity(tgh, p):
    f0, _ = _forests(tgh)
    o0, _ = _pair(f0, f0, p)
    m = _off_seams(f0)
    u = _poly_state(o0, f0, p - 1)
    v = o0.transfer_from(o0, u, 0.0)
    np.testing.assert_array_equal(v[m], u[m])
    np.testing.assert_array_equal(o0.transfer_from(o0, v, 0.0), v)


@pytest.mark.parametrize("p", [4, 6])
def test_coarsening_is_injection(tgh, p):
    f0, f1 = _forests(tgh)
    o0, o1 = _pair(f0, f1, p)
    m = _off_seams(f0)
    for deg in (p - 1, p):
        u0, u1 = _poly_state(o0, f0, deg), _poly_state(o1, f1, deg)
        np.testing.assert_array_equal(o0.transfer_from(o1, u1, 0.0)[m], u0[m])


@pytest.mark.parametrize("p", [4, 6])
def test_refinement_reproduces_polynomials_of_degree_p_minus_1(tgh, p):
    """On the new level-2 leaves (every point a copy of a level-1 vertex or
    a prolongation whose stencil stays off the seams)."""
    f0, f1 = _forests(tgh)
    o0, o1 = _pair(f0, f1, p)
    P3 = (N + 1) ** 3
    new = np.concatenate([np.full(20 * P3, lf[0] == 2) for lf in f1.leaves])
    errs = {}
    for deg in (p - 1, p):
        u0, u1 = _poly_state(o0, f0, deg), _poly_state(o1, f1, deg)
        got = o1.transfer_from(o0, u0, 0.0)
        errs[deg] = float(np.max(np.abs(got - u1)[new] / np.maximum(1.0, np.abs(u1[new]))))
    assert errs[p - 1] < 1e-13, errs
    assert errs[p] > 1e-11, errs


def test_transfer_refuses_a_jump_of_two_levels(tgh):
    """A new leaf two levels finer than the old forest where it lies has no
    old leaf one level coarser to interpolate from."""
    case = tgh.gauge_wave_case(eps_ko=0.0, gamma0=1.0, gamma2=0.0)
    f0, f1 = _forests(tgh)
    flat = tgh.gh_forest(None, case, N=N, roots=ROOTS)
    flat.refine([])
    d = TD.gauge_wave_case(N=N, roots=ROOTS, q=4)
    o_flat = oracle_lib.OracleProblem(d, flat.leaves, 4)
    o1 = oracle_lib.OracleProblem(d, f1.leaves, 4)
    with pytest.raises(RuntimeError):
        o1.transfer_from(o_flat, o_flat.zeros(), 0.0)


def test_forest_coarsen_undoes_refine(tgh):
    f0, _ = _forests(tgh)
    g = tgh.Forest(f0.case, f0.N, f0.roots, f0.leaves)
    g.refine([(1, 2, 2, 2)])
    assert g.nleaves() == f0.nleaves() + 7
    g.coarsen([(1, 2, 2, 2)])
    assert sorted(g.leaves) == sorted(f0.leaves)
    with pytest.raises(tgh.ArgumentError):
        g.coarsen([(1, 2, 2, 2)])


def test_regrid_forest_refines_marked_leaves_and_balances(tgh):
    f0, _ = _forests(tgh)
    leaves = f0.leaves
    flags = [1 if lf == (1, 2, 2, 2) else 0 for lf in leaves]
    g = tgh.regrid_forest(f0, leaves, flags)
    assert g.isbalanced() and g.maxlevel() == 2
    assert all((2, 4 + (c & 1), 4 + (c >> 1 & 1), 4 + (c >> 2 & 1)) in g.leaves for c in range(8))
    with pytest.raises(tgh.ArgumentError):
        tgh.regrid_forest(f0, leaves, flags[:-1])


def test_regrid_forest_coarsens_only_where_the_parent_stays_balanced(tgh):
    """Every level-1 leaf of the three-level forest marked to coarsen: each
    complete group there touches the level-2 corner (balance! made them), so
    none merges -- merging any one unbalances the forest -- while on the
    two-level forest every group merges."""
    _, f1 = _forests(tgh)
    leaves = f1.leaves
    g = tgh.regrid_forest(f1, leaves, [-1 if lf[0] == 1 else 0 for lf in leaves])
    assert g.isbalanced()
    assert {lf for lf in leaves if lf[0] == 2} <= set(g.leaves)
    assert sorted(g.leaves) == sorted(leaves)
    kept = {(lf[0] - 1, lf[1] >> 1, lf[2] >> 1, lf[3] >> 1) for lf in g.leaves if lf[0] == 1}
    assert kept
    for par in kept:
        kids = [(1, 2 * par[1] + (c & 1), 2 * par[2] + (c >> 1 & 1), 2 * par[3] + (c >> 2 & 1)) for c in range(8)]
        if not all(k in g.leaves for k in kids):
            continue
        h = tgh.Forest(g.case, g.N, g.roots, g.leaves)
        h.coarsen([par])
        assert not h.isbalanced(), par
    # the whole coarse level goes where nothing finer is marked to stay
    f0, _ = _forests(tgh)
    u = tgh.regrid_forest(f0, f0.leaves, [-1 if lf[0] == 1 else 0 for lf in f0.leaves])
    assert u.maxlevel() == 0 and u.nleaves() == ROOTS ** 3
