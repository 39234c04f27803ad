"""The regrid's marking on the CPU: Refinement's checks, the level floor and
ceiling (level_bounds, block_level_bounds in src/refinement.jl:430-668) and
refine_flags' clamp (:670-730), on hand-made indicator values."""
import pytest


def _case(tgh, **kw):
    # a = 0.5: r_h,min = 1 + sqrt(3)/2, r_h,max = sqrt(r_h,min^2 + 1/4); the
    # layer r_0 = 0.7, r_1 = 1.2 needs h <= min((r_h,min - r_1)/8, 0.5/8) =
    # 1/16 at q = 4: level 1 at the root spacing 1/8 of roots = 4 over [-4, 4]^3
    return tgh.kerr_schild_case(M=1.0, a=0.5, halfwidth=4.0, r_0=0.7, r_1=1.2, gauge_source="algebraic", **kw)


def _ref(tgh, **kw):
    args = dict(refine_tol=0.5, coarsen_tol=0.1, maxlevel_cap=2, floor_margin=0.0)
    args.update(kw)
    return tgh.Refinement(**args)


def test_refinement_refuses_thresholds_the_reference_refuses(tgh):
    for kw in (dict(coarsen_tol=0.0), dict(coarsen_tol=0.5), dict(refine_tol=1.0, coarsen_tol=0.1),
               dict(maxlevel_cap=-1), dict(floor_margin=-0.1), dict(ceiling_level=3),
               dict(ceiling_cells=-1), dict(eps=0.0)):
        with pytest.raises(tgh.ArgumentError):
            _ref(tgh, **kw)


def test_the_floor_holds_the_horizon_shell_and_the_ceiling_the_boundary(tgh):
    case = _case(tgh)
    f = tgh.gh_forest(None, case, N=16, roots=4)
    ref = _ref(tgh)
    bounds = tgh.level_bounds(f, case, 4, ref, 0.0)
    r_lo = 1.2
    r_hi = tgh.horizon_max_radius(case.background)
    for lf, (flo, cap) in zip(f.leaves, bounds):
        ext = f.block_extent(lf)
        near, far = tgh._box_radii(ext, (0.0, 0.0, 0.0))
        assert flo == (1 if near <= r_hi and far >= r_lo else 0)
        boundary = any(ext[d][0] <= -4.0 + 0.125 or ext[d][1] >= 4.0 - 0.125 for d in range(3))
        assert cap == (0 if boundary else 2)
    assert sum(1 for b in bounds if b[0] == 1) == 8


def test_a_quiet_horizon_shell_does_not_coarsen(tgh):
    """Every leaf quiet (τ = 0): the level-1 leaves the floor's shell meets
    keep their level, the other level-1 leaves ask to coarsen, level 0 stays."""
    case = _case(tgh)
    f = tgh.hole_forest(None, case, N=16, roots=4, radii=(3.0,))
    ref = _ref(tgh, maxlevel_cap=1)
    leaves = f.leaves
    bounds = tgh.level_bounds(f, case, 4, ref, 0.0)
    flags = tgh.refine_flags(leaves, [0.0] * len(leaves), bounds, ref)
    kept = [lf for lf, fl in zip(leaves, flags) if lf[0] == 1 and fl == 0]
    assert len(kept) == 64
    for lf, fl, (flo, _) in zip(leaves, flags, bounds):
        assert fl == (0 if lf[0] == 0 or flo == 1 else -1)
    g = tgh.regrid_forest(f, leaves, flags)
    assert all(lf in g.leaves for lf in kept) and g.maxlevel() == 1


def test_a_boundary_block_does_not_refine(tgh):
    """τ above refine_tol everywhere: the leaves within the ceiling's margin
    keep level 0, the others refine once."""
    case = _case(tgh)
    f = tgh.gh_forest(None, case, N=16, roots=4)
    ref = _ref(tgh)
    bounds = tgh.level_bounds(f, case, 4, ref, 0.0)
    flags = tgh.refine_flags(f.leaves, [0.9] * f.nleaves(), bounds, ref)
    for fl, (_, cap) in zip(flags, bounds):
        assert fl == (0 if cap == 0 else 1)
    assert 0 < sum(flags) < f.nleaves()


def test_impossible_bounds_are_errors(tgh):
    case = _case(tgh)
    f = tgh.gh_forest(None, case, N=16, roots=4)
    # the interior needs level 1, the cap allows 0
    with pytest.raises(tgh.ArgumentError, match="needs level 1"):
        tgh.level_bounds(f, case, 4, _ref(tgh, maxlevel_cap=0), 0.0)
    # a shell reaching the ceiling's margin: floor 1 against cap 0
    with pytest.raises(tgh.ArgumentError, match="caps it"):
        tgh.level_bounds(f, case, 4, _ref(tgh, floor_margin=1.0), 0.0)
    # a finer layer's demand at the default margin m = 8
    with pytest.raises(tgh.ArgumentError, match="needs level"):
        tgh.level_bounds(f, _case(tgh, margin=20), 4, _ref(tgh, maxlevel_cap=1), 0.0)
