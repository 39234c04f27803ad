"""The gauge constraint monitor on the device against the oracle, through the
C ABI (uniform and refined meshes, closed-form and sampled sources, a
harmonic background), and its masked norms' convergence on the frozen
hierarchy."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
import tgh_desc as TD

pytestmark = pytest.mark.gpu


def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    return torch


def _compare(tgh, dev, orc, u, t):
    torch = torch_cuda()
    ud = torch.from_numpy(u.copy()).cuda()
    l2g, lig = tgh.gauge_constraint_norms(dev, ud, t)
    l2o, lio = orc.gauge_constraint(u, t)
    print(f"gauge constraint: GPU l2 {l2g} linf {lig}; oracle l2 {list(l2o)} linf {list(lio)}")
    np.testing.assert_allclose(l2g, l2o, rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(lig, lio, rtol=1e-9, atol=1e-15)


@pytest.mark.parametrize("gauge_source", [None, "algebraic"])
def test_gauge_constraint_matches_oracle_on_a_uniform_hole(tgh, gauge_source):
    case = tgh.kerr_schild_case(M=1.0, a=0.5, halfwidth=2.0, r_0=0.71, r_1=1.21, gauge_source=gauge_source)
    dev = tgh.GHProblem(tgh.gh_forest(None, case, N=16, roots=4), case, q=4)
    orc = oracle_lib.OracleProblem(TD.TghDesc.from_buffer_copy(dev.desc))
    u = orc.fill_exact(0.0)
    orc.add_noise(u, 0xC0A5, 1e-6)
    _compare(tgh, dev, orc, u, 0.0)


def test_gauge_constraint_matches_oracle_on_a_forest(tgh):
    case = tgh.kerr_schild_case(M=1.0, a=0.5, halfwidth=4.0, r_0=0.7, r_1=1.7, gauge_source="algebraic")
    f = tgh.hole_forest(None, case, N=8, roots=4, radii=(2.5, 1.0))
    dev = tgh.GHProblem(f, case, q=4, ops=tgh.Operators(prolongation=6))
    orc = oracle_lib.OracleProblem(TD.TghDesc.from_buffer_copy(dev.desc), f.leaves, 6)
    u = orc.fill_exact(0.0)
    orc.add_noise(u, 0xC0A6, 1e-6)
    _compare(tgh, dev, orc, u, 0.0)


def test_gauge_constraint_matches_oracle_on_the_gauge_wave_forest(tgh):
    case = tgh.gauge_wave_case(eps_ko=0.0, gamma0=1.0, gamma2=0.0)
    f = tgh.gh_forest(None, case, N=8, roots=2, refined=True)
    dev = tgh.GHProblem(f, case, q=4, ops=tgh.Operators(prolongation=6))
    orc = oracle_lib.OracleProblem(TD.TghDesc.from_buffer_copy(dev.desc), f.leaves, 6)
    u = orc.fill_exact(0.1)
    orc.add_noise(u, 0xC0A7, 1e-6)
    _compare(tgh, dev, orc, u, 0.1)


def test_gauge_constraint_converges_on_the_frozen_hierarchy(tgh):
    """test/driver_tests.jl:209-245's rate_gauge (the record's gauge_l2, the
    largest component's L2, src/driver.jl:1545) on the static-hole fixture
    at q = 2: evolved to t = 3/20, falling at order q."""
    torch = torch_cuda()
    case = tgh.kerr_schild_case(M=1.0, a=0.0, halfwidth=2.5, r_0=0.4, r_1=1.15)
    hs, gl = [], []
    for N in (6, 8, 10):
        f = tgh.hole_forest(None, case, N=N, roots=1, radii=(3.0, 3.0, 1.0))
        p = tgh.GHProblem(f, case, q=2, ops=tgh.Operators(prolongation=4))
        u = p.statevector()
        tgh.initial_state(u, p, 0.0)
        dt = tgh.gh_dt(p, u, cfl=0.25)
        tgh.gh_solve(p, u, (0.0, 3 / 20), dt=dt, graph=True)
        l2, _ = tgh.gauge_constraint_norms(p, u, 3 / 20)
        hs.append(f.minimum_spacing())
        gl.append(max(l2))
    x, y = np.log(hs), np.log(gl)
    rate = float(((x - x.mean()) * (y - y.mean())).sum() / ((x - x.mean()) ** 2).sum())
    print(f"gauge constraint on the frozen hierarchy, q=2: {gl}, rate {rate:.3f}")
    assert rate > 2 - 0.25


def _compare_adm(tgh, dev, orc, u, t):
    torch = torch_cuda()
    ud = torch.from_numpy(u.copy()).cuda()
    l2g, lig = tgh.adm_constraint_norms(dev, ud, t)
    ham, ml2, mli = orc.adm_constraint(u, t)
    l2o, lio = [ham[0]] + list(ml2), [ham[1]] + list(mli)
    print(f"ADM constraints: GPU l2 {l2g} linf {lig}; oracle l2 {l2o} linf {lio}")
    np.testing.assert_allclose(l2g, l2o, rtol=1e-8, atol=1e-14)
    np.testing.assert_allclose(lig, lio, rtol=1e-8, atol=1e-14)


@pytest.mark.parametrize("gauge_source", [None, "algebraic"])
def test_adm_constraints_match_oracle_on_a_uniform_hole(tgh, gauge_source):
    case = tgh.kerr_schild_case(M=1.0, a=0.5, halfwidth=2.0, r_0=0.71, r_1=1.21, gauge_source=gauge_source)
    dev = tgh.GHProblem(tgh.gh_forest(None, case, N=16, roots=4), case, q=4)
    orc = oracle_lib.OracleProblem(TD.TghDesc.from_buffer_copy(dev.desc))
    u = orc.fill_exact(0.0)
    orc.add_noise(u, 0xADA1, 1e-6)
    _compare_adm(tgh, dev, orc, u, 0.0)


def test_adm_constraints_match_oracle_on_a_forest(tgh):
    case = tgh.kerr_schild_case(M=1.0, a=0.5, halfwidth=4.0, r_0=0.7, r_1=1.7, gauge_source="algebraic")
    f = tgh.hole_forest(None, case, N=8, roots=4, radii=(2.5, 1.0))
    dev = tgh.GHProblem(f, case, q=4, ops=tgh.Operators(prolongation=6))
    orc = oracle_lib.OracleProblem(TD.TghDesc.from_buffer_copy(dev.desc), f.leaves, 6)
    u = orc.fill_exact(0.0)
    orc.add_noise(u, 0xADA2, 1e-6)
    _compare_adm(tgh, dev, orc, u, 0.0)


def test_refinement_indicator_matches_oracle(tgh):
    """gh_tau! + per-block maxima on the device against the oracle, on a noisy
    hole (uniform) and on a forest, with the field scale as the reference
    amplitude."""
    torch = torch_cuda()
    case = tgh.kerr_schild_case(M=1.0, a=0.5, halfwidth=4.0, r_0=0.7, r_1=1.7, gauge_source="algebraic")
    small = tgh.kerr_schild_case(M=1.0, a=0.5, halfwidth=2.0, r_0=0.7, r_1=1.7, gauge_source="algebraic")
    for case, forest in ((small, tgh.gh_forest(None, small, N=16, roots=4)),
                         (case, tgh.hole_forest(None, case, N=8, roots=4, radii=(2.5, 1.0)))):
        if forest.refined:
            dev = tgh.GHProblem(forest, case, q=4, ops=tgh.Operators(prolongation=6))
            orc = oracle_lib.OracleProblem(TD.TghDesc.from_buffer_copy(dev.desc), forest.leaves, 6)
        else:
            dev = tgh.GHProblem(forest, case, q=4)
            orc = oracle_lib.OracleProblem(TD.TghDesc.from_buffer_copy(dev.desc))
        u = orc.fill_exact(0.0)
        orc.add_noise(u, 0x7A0, 1e-4)
        ud = torch.from_numpy(u.copy()).cuda()
        s_d = tgh.field_scale(dev, ud, 0.0)
        s_o = orc.field_scale(u, 0.0)
        assert s_d == s_o
        tg = np.array(tgh.refinement_indicator(dev, ud, 0.0, scale=s_o, eps=0.01))
        to = orc.tau_blocks(u, 0.0, s_o, 0.01)
        print(f"{forest.nleaves()} leaves: tau_max GPU {tg.max():.6e} oracle {to.max():.6e}, "
              f"max |diff| {np.abs(tg - to).max():.3e}")
        np.testing.assert_allclose(tg, to, rtol=1e-13, atol=0)


def test_record_norms_is_the_three_monitors_against_the_oracle(tgh):
    """record_norms (src/constraints.jl:868-881) on the a = 0.5 hole with noise:
    its gauge, ham and mom rows are the oracle's monitors' and its error rows
    the oracle's error rows (the drift over the horizon shell), each to the
    monitors' own tolerance; without the ADM pass its ham and mom fields are
    empty, as the driver's rows between ADM chunks are."""
    torch = torch_cuda()
    case = tgh.kerr_schild_case(M=1.0, a=0.5, halfwidth=2.0, r_0=0.71, r_1=1.21, gauge_source="algebraic")
    dev = tgh.GHProblem(tgh.gh_forest(None, case, N=16, roots=4), case, q=4)
    orc = oracle_lib.OracleProblem(TD.TghDesc.from_buffer_copy(dev.desc))
    u = orc.fill_exact(0.0)
    orc.add_noise(u, 0xC0A9, 1e-6)
    ud = torch.from_numpy(u.copy()).cuda()
    shell = (1.21, tgh.horizon_max_radius(case.background))
    rec = tgh.record_norms(dev, ud, 0.05, shell=shell)
    l2o, lio = orc.gauge_constraint(u, 0.05)
    np.testing.assert_allclose(rec["gauge_l2"], l2o, rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(rec["gauge_linf"], lio, rtol=1e-9, atol=1e-15)
    ham, ml2, mli = orc.adm_constraint(u, 0.05)
    np.testing.assert_allclose([rec["ham_l2"], rec["ham_linf"]], ham, rtol=1e-8, atol=1e-14)
    np.testing.assert_allclose(rec["mom_l2"], ml2, rtol=1e-8, atol=1e-14)
    np.testing.assert_allclose(rec["mom_linf"], mli, rtol=1e-8, atol=1e-14)
    rows = orc.error_rows(u, 0.05, shell)
    print(f"record: {rec}; oracle error rows {rows}")
    np.testing.assert_allclose([rec["err_l2"], rec["err_linf"], rec["residual"], rec["drift"]], rows,
                               rtol=1e-9, atol=1e-15)
    assert rec["drift"] > 0 and rec["residual"] > 0
    lean = tgh.record_norms(dev, ud, 0.05, adm=False)
    assert lean["ham_l2"] is None and lean["mom_linf"] is None and lean["drift"] == 0.0
    assert lean["gauge_l2"] == rec["gauge_l2"] and lean["err_l2"] == rec["err_l2"]
