"""The gauge constraint monitor in the oracle (gh_constraint_kernel!,
src/constraints.jl:211-309; gauge_constraint_at_node, src/pointwise.jl:
947-960), pinned by what the mathematics fixes: C_a vanishes on flat space
and (to roundoff) on the gauge wave, which is harmonic; on the exact
Kerr-Schild state with its own source H_a = -Gamma_a it is the truncation
error of the first derivatives, falling 2^q per halving of h, whichever form
the source takes (a wrong sign or index in Gamma or in raising H leaves an
O(1) constraint that does not fall)."""
import numpy as np

import oracle_lib
import tgh_desc as TD


def test_flat_space_and_the_gauge_wave_satisfy_it():
    o = oracle_lib.OracleProblem(TD.minkowski_noise_case(N=8))
    l2, li = o.gauge_constraint(o.fill_exact(0.0), 0.0)
    assert np.all(l2 == 0.0) and np.all(li == 0.0)
    o = oracle_lib.OracleProblem(TD.gauge_wave_case(N=8, roots=2, q=4))
    l2, li = o.gauge_constraint(o.fill_exact(0.1), 0.1)
    assert np.all(li <= 1e-15)


def test_kerr_schild_constraint_converges_at_order_q_with_either_source():
    res = {}
    for gauge in (TD.TGH_GAUGE_KERR_SCHILD, TD.TGH_GAUGE_SAMPLED):
        for N in (8, 16):
            d = TD.kerr_schild_hole_case(N=N, nblocks=4, halfwidth=4.0, a=0.5, gauge=gauge, r0=0.7, r1=1.7)
            o = oracle_lib.OracleProblem(d)
            res[gauge, N] = o.gauge_constraint(o.fill_exact(0.0), 0.0)
    for gauge in (TD.TGH_GAUGE_KERR_SCHILD, TD.TGH_GAUGE_SAMPLED):
        rate = np.log2(res[gauge, 8][0] / res[gauge, 16][0])
        assert np.all(rate > 3.7) and np.all(rate < 4.3), rate
        assert np.all(res[gauge, 8][1] > 0)
    # at the exact state the closed form and the sample are the same H
    np.testing.assert_allclose(res[TD.TGH_GAUGE_SAMPLED, 16][0], res[TD.TGH_GAUGE_KERR_SCHILD, 16][0], rtol=1e-10)


def test_adm_constraints_vanish_on_flat_space_and_converge_on_kerr_schild():
    """adm_constraint_kernel! (src/constraints.jl:100-210,432-534): zero on
    flat space, roundoff on the gauge wave (flat space in another slicing),
    and the truncation error of the second derivatives on the exact
    Kerr-Schild state, falling about 2^q per halving (a wrong term in the
    Ricci tensor or the time derivatives leaves an O(1) constraint)."""
    o = oracle_lib.OracleProblem(TD.minkowski_noise_case(N=8))
    ham, ml2, mli = o.adm_constraint(o.fill_exact(0.0), 0.0)
    assert np.all(ham == 0.0) and np.all(ml2 == 0.0) and np.all(mli == 0.0)
    o = oracle_lib.OracleProblem(TD.gauge_wave_case(N=8, roots=2, q=4))
    ham, ml2, mli = o.adm_constraint(o.fill_exact(0.1), 0.1)
    assert ham[1] <= 1e-12 and np.all(mli <= 1e-12)
    res = {}
    for N in (8, 16):
        d = TD.kerr_schild_hole_case(N=N, nblocks=4, halfwidth=4.0, a=0.5, gauge=TD.TGH_GAUGE_KERR_SCHILD,
                                     r0=0.7, r1=1.7)
        o = oracle_lib.OracleProblem(d)
        res[N] = o.adm_constraint(o.fill_exact(0.0), 0.0)
    rate_h = np.log2(res[8][0][0] / res[16][0][0])
    rate_m = np.log2(res[8][1] / res[16][1])
    assert 3.5 < rate_h < 4.5, rate_h
    assert np.all(rate_m > 3.5) and np.all(rate_m < 4.5), rate_m
