"""The refinement indicator in the oracle (lohner, cell_tau, gh_tau_kernel!,
src/refinement.jl:152-219,312-365): zero on linear data, and at a unit jump
1 / (1 + 4 eps |scale|) -- the ratio's numerator is the jump and its
denominator the jump plus the noise floor -- in exactly the blocks that hold
a point beside the jump; field_scale the largest |h_v| over evolved points."""
import numpy as np

import oracle_lib
import tgh_desc as TD


def _line_problem():
    # eight blocks of 8^3 cells along x over [0, 8] x [0, 1]^2, periodic
    d = TD.make_desc(q=4, N=8, nblocks=(8, 1, 1), periodic=(1, 1, 1), lo=(0, 0, 0), hi=(8, 1, 1))
    return d, oracle_lib.OracleProblem(d)


def _state(o, d, f):
    u = o.zeros()
    P = d.N + 1
    B = o.block_coords()
    i = np.arange(P)
    for b, (bx, by, bz) in enumerate(B):
        x = (bx * d.N + i) / d.N
        vals = np.broadcast_to(f(x)[None, None, :], (P, P, P)).ravel()
        for v in range(20):
            u[(b * 20 + v) * P ** 3:(b * 20 + v + 1) * P ** 3] = vals
    return u, B


def test_indicator_on_linear_data_and_at_a_jump():
    d, o = _line_problem()
    u, B = _state(o, d, lambda x: np.full_like(x, 0.25))
    assert np.all(o.tau_blocks(u, 0.0, 1.0) == 0.0)
    u, B = _state(o, d, lambda x: (x >= 4.0).astype(float))
    assert o.field_scale(u, 0.0) == 1.0
    eps = 0.01
    tau = o.tau_blocks(u, 0.0, 1.0, eps)
    jump = 1.0 / (1.0 + 4 * eps)
    for b, (bx, _, _) in enumerate(B):
        # the jumps at x = 4 (blocks 3 and 4) and across the periodic seam (0 and 7)
        expect = jump if bx in (0, 3, 4, 7) else 0.0
        assert abs(tau[b] - expect) <= 1e-15, (bx, tau[b])
    # linear data in the middle blocks: only the seam's wrap shows
    u, B = _state(o, d, lambda x: 0.5 * x)
    tau = o.tau_blocks(u, 0.0, 4.0, eps)
    for b, (bx, _, _) in enumerate(B):
        if bx in (1, 2, 3, 4, 5, 6):
            assert tau[b] <= 1e-14
