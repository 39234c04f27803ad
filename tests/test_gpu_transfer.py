"""The regrid's state transfer on the device (tgh_transfer) against the oracle,
and regrid end to end (refine_flags then refine!/coarsen!/balance! and the
transfer, src/refinement.jl:738-795,973-1021), through the C ABI.

Per component |GPU − oracle| / max(1, |oracle|) ≤ 1e-12: the copies and the
interpolation sums are the oracle's own order, and the Dirichlet ghosts a
stencil reads near a non-periodic face are the background evaluated on
both sides.
"""
import numpy as np
import pytest

import oracle_lib
import tgh_desc as TD
from test_oracle_transfer import _forests

pytestmark = pytest.mark.gpu

TOL = 1e-12


def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    return torch


def _problem(tgh, f, case, p):
    dev = tgh.GHProblem(f, case, q=4, ops=tgh.Operators(prolongation=p))
    desc = TD.TghDesc.from_buffer_copy(dev.desc)
    return dev, oracle_lib.OracleProblem(desc, dev.leaves(), p)


def _noisy_state(tgh, torch, dev, t, seed):
    u = dev.statevector()
    tgh.initial_state(u, dev, t)
    tgh.add_noise(u, dev, seed, 1e-3)
    return u


def _close(got, want):
    err = np.abs(got - want) / np.maximum(1.0, np.abs(want))
    return float(err.max())


@pytest.mark.parametrize("p", [4, 6])
def test_transfer_matches_the_oracle_on_the_gauge_wave(tgh, p):
    """Refinement, coarsening and the identity between two periodic forests."""
    torch = torch_cuda()
    f0, f1 = _forests(tgh)
    case = tgh.gauge_wave_case(eps_ko=0.0, gamma0=1.0, gamma2=0.0)
    d0, o0 = _problem(tgh, f0, case, p)
    d1, o1 = _problem(tgh, f1, case, p)
    t = 0.1
    u0 = _noisy_state(tgh, torch, d0, t, 11)
    u1 = _noisy_state(tgh, torch, d1, t, 12)
    h0, h1 = u0.cpu().numpy(), u1.cpu().numpy()
    for src_d, src_o, u_src, h_src, dst_d, dst_o in ((d0, o0, u0, h0, d1, o1), (d1, o1, u1, h1, d0, o0),
                                                     (d0, o0, u0, h0, d0, o0)):
        got = dst_d.statevector()
        tgh.transfer(src_d, u_src, t, dst_d, got)
        want = dst_o.transfer_from(src_o, h_src, t)
        assert _close(got.cpu().numpy(), want) <= TOL, (src_d.nblocks, dst_d.nblocks)


def test_transfer_from_a_uniform_problem_to_a_hole_forest(tgh):
    """A Kerr-Schild hole in a Dirichlet box (N = 16: the layer's width is at
    least 2(G+1)h on the root level): the uniform problem's blocks (its
    own order) to a two-level hole_forest, prolongation stencils reading the
    hook's ghosts at the faces."""
    torch = torch_cuda()
    case = tgh.kerr_schild_case(M=1.0, a=0.5, halfwidth=4.0, r_0=0.7, r_1=1.7, gauge_source="algebraic")
    fu = tgh.gh_forest(None, case, N=16, roots=4)
    fh = tgh.hole_forest(None, case, N=16, roots=4, radii=(1.5,))
    du, ou = _problem(tgh, fu, case, 6)
    dh, oh = _problem(tgh, fh, case, 6)
    u = _noisy_state(tgh, torch, du, 0.0, 5)
    got = dh.statevector()
    tgh.transfer(du, u, 0.0, dh, got)
    want = oh.transfer_from(ou, u.cpu().numpy(), 0.0)
    assert _close(got.cpu().numpy(), want) <= TOL
    # back again: injection onto the uniform blocks
    back = du.statevector()
    tgh.transfer(dh, got, 0.0, du, back)
    assert _close(back.cpu().numpy(), ou.transfer_from(oh, want, 0.0)) <= TOL


def test_transfer_refuses_other_bricks_and_two_level_jumps(tgh):
    torch_cuda()
    case = tgh.gauge_wave_case(eps_ko=0.0, gamma0=1.0, gamma2=0.0)
    f0, f1 = _forests(tgh)
    flat = tgh.gh_forest(None, case, N=8, roots=4)
    dflat = tgh.GHProblem(flat, case, q=4)
    d1 = tgh.GHProblem(f1, case, q=4, ops=tgh.Operators(prolongation=6))
    with pytest.raises(tgh.ArgumentError, match="more than one level finer"):
        tgh.transfer(dflat, dflat.statevector(), 0.0, d1, d1.statevector())
    other = tgh.GHProblem(tgh.gh_forest(None, case, N=8, roots=2), case, q=4)
    with pytest.raises(tgh.ArgumentError, match="same root brick"):
        tgh.transfer(other, other.statevector(), 0.0, dflat, dflat.statevector())


def _hole_case(tgh, v=0.0, halfwidth=4.0):
    """A Kerr-Schild hole (a = 0.5, the closed-form source) whose layer
    r_0 = 0.7, r_1 = 1.2 needs level-1 spacing 1/16 at root spacing 1/8
    (r_1 - r_0 >= 2(G+1)h with G = 3), so the refinement's level floor is 1
    (src/refinement.jl:436-470) on the shell from r_1 to the horizon's
    largest radius (1.93) plus the floor's margin."""
    return tgh.kerr_schild_case(M=1.0, a=0.5, v=v, axis=1, halfwidth=halfwidth, r_0=0.7, r_1=1.2,
                                gauge_source="algebraic")


def _tau(tgh, p, u, t):
    return tgh.refinement_indicator(p, u, t, scale=tgh.field_scale(p, u, t), eps=0.01)


def test_regrid_refines_where_the_indicator_fires_and_carries_the_state(tgh):
    """A Kerr-Schild hole in [-6, 6]^3 on 6^3 root blocks of 16^3 whose eight
    central blocks hold the layer at level 1 (hole_forest): with the floor's
    shell reaching 1/2 past the horizon's largest radius, the 24 root blocks
    face-adjacent to the centre are held at level 1 by the floor, so they
    refine, and so do the root blocks whose Löhner maximum passes refine_tol
    (the upper quartile); the state on the new leaves is the oracle's
    transfer, and the new problem steps."""
    torch = torch_cuda()
    case = _hole_case(tgh, halfwidth=6.0)
    f = tgh.hole_forest(None, case, N=16, roots=6, radii=(1.5,))
    assert f.nleaves() == 272 and f.maxlevel() == 1
    dev, orc = _problem(tgh, f, case, 6)
    u = dev.statevector()
    tgh.initial_state(u, dev, 0.0)
    tau = _tau(tgh, dev, u, 0.0)
    thr = sorted(tau)[len(tau) * 3 // 4]
    ref = tgh.Refinement(refine_tol=thr, coarsen_tol=thr * 1e-3, maxlevel_cap=1, floor_margin=0.5)
    bounds = tgh.level_bounds(dev.forest, dev.case, dev.q, ref, 0.0, dev.leaves())
    held = [lf for lf, (flo, _) in zip(dev.leaves(), bounds) if flo == 1 and lf[0] == 0]
    assert len(held) == 24  # the root blocks face-adjacent to the central eight
    new, u_new, flags = tgh.regrid(dev, u, 0.0, refinement=ref)
    fired = [lf for lf, fl in zip(dev.leaves(), flags) if fl > 0]
    assert 24 <= len(fired) < f.nleaves()
    for lf, (flo, _), t in zip(dev.leaves(), bounds, tau):
        assert (lf in fired) == (lf[0] == 0 and (flo == 1 or t > thr)), (lf, flo, t, thr)
    for lf in fired:
        assert lf not in new.forest.leaves
        assert (1, 2 * lf[1], 2 * lf[2], 2 * lf[3]) in new.forest.leaves
    assert new.forest.isbalanced()
    desc = TD.TghDesc.from_buffer_copy(new.desc)
    o_new = oracle_lib.OracleProblem(desc, new.forest.leaves, 6)
    want = o_new.transfer_from(orc, u.cpu().numpy(), 0.0)
    assert _close(u_new.cpu().numpy(), want) <= TOL
    dt = 0.25 * new.forest.minimum_spacing()
    tgh.gh_solve(new, u_new, (0.0, 2 * dt), dt=dt)
    assert torch.isfinite(u_new).all()


def test_regrid_coarsens_quiet_leaves_outside_the_floor(tgh):
    """The same hole on a two-level hole_forest that reaches the outer faces,
    with coarsen_tol above every block's indicator: the level-1 leaves the
    floor's shell meets keep their level, the sibling groups outside it merge
    (the ceiling caps the ones at the boundary at level 0 anyway), and the
    state is injected."""
    torch_cuda()
    case = _hole_case(tgh)
    fh = tgh.hole_forest(None, case, N=16, roots=4, radii=(3.0,))
    dev, orc = _problem(tgh, fh, case, 6)
    u = dev.statevector()
    tgh.initial_state(u, dev, 0.0)
    ref = tgh.Refinement(refine_tol=0.9999, coarsen_tol=0.999, maxlevel_cap=1, floor_margin=0.0)
    new, u_new, flags = tgh.regrid(dev, u, 0.0, refinement=ref)
    bounds = tgh.level_bounds(fh, case, 4, ref, 0.0, dev.leaves())
    for lf, fl, (flo, _) in zip(dev.leaves(), flags, bounds):
        assert fl == (0 if lf[0] == 0 or flo == 1 else -1), (lf, fl, flo)
    kept = [lf for lf in dev.leaves() if lf[0] == 1 and lf in new.forest.leaves]
    assert len(kept) == 64 and new.forest.maxlevel() == 1
    desc = TD.TghDesc.from_buffer_copy(new.desc)
    o_new = oracle_lib.OracleProblem(desc, new.forest.leaves, 6)
    assert _close(u_new.cpu().numpy(), o_new.transfer_from(orc, u.cpu().numpy(), 0.0)) <= TOL


def test_regrid_cadence_follows_a_boosted_hole(tgh):
    """Config 5's loop in small: a Kerr-Schild hole boosted to v = 0.3 along
    x on a two-level hole_forest, regridded every 8 steps with each chunk's
    step sized on its own mesh (cfl 1/4), the upper quartile of the first
    indicator refining. The loop is deterministic, reaches level 2 and keeps
    the masked error within twice that of the fixed forest over the same
    span."""
    torch = torch_cuda()
    case = _hole_case(tgh, v=0.3)
    fh = tgh.hole_forest(None, case, N=16, roots=4, radii=(1.5,))
    ops = tgh.Operators(prolongation=6)
    fixed = tgh.GHProblem(fh, case, q=4, ops=ops)
    u = fixed.statevector()
    tgh.initial_state(u, fixed, 0.0)
    tau = _tau(tgh, fixed, u, 0.0)
    thr = sorted(tau)[len(tau) * 3 // 4]
    ref = tgh.Refinement(refine_tol=thr, coarsen_tol=thr * 1e-3, maxlevel_cap=2, floor_margin=0.0)
    t1 = 24 * tgh.gh_dt(fixed, u, cfl=0.25)
    runs = []
    for _ in range(2):
        p0 = tgh.GHProblem(fh, case, q=4, ops=ops)
        u0 = p0.statevector()
        tgh.initial_state(u0, p0, 0.0)
        runs.append(tgh.gh_solve_regrid(p0, u0, (0.0, t1), cfl=0.25, every=8, refinement=ref))
    (p, uu, log), (p2, uu2, log2) = runs
    print(f"regrids: {log}")
    assert len(log) >= 2 and log == log2
    assert p.forest.maxlevel() == 2 and p.forest.isbalanced()
    # once level 2 exists the step halves: the chunks after it are on it
    assert log[-1][2] < log[0][2]
    assert torch.equal(uu, uu2)
    assert torch.isfinite(uu).all()
    tgh.gh_solve(fixed, u, (0.0, t1), dt=tgh.gh_dt(fixed, u, cfl=0.25))
    e_fixed = tgh.error_norms(fixed, u, t1)[0]
    e_regrid = tgh.error_norms(p, uu, t1)[0]
    assert e_regrid <= 2 * e_fixed, (e_regrid, e_fixed)
