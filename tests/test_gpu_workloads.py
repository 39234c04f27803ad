"""The workloads at their own sizes or in small, on the device.

* The single-hole workload at full size: the a = 0.7 Kerr-Schild hole in
  [-8M, 8M]^3 on 4096 blocks of 16^3 (h = M/16, q = 4, :damped layer,
  Gaussian damping), 100 RK4 steps of the brick stepper's hipGraph at cfl
  1/4, for both gauge-source kinds (the closed form and the sample taken at
  creation). Too large for the oracle, so checked through properties that do
  not depend on size: the state stays finite, the masked error and the
  gauge constraint stay at the truncation level, and they grow no faster
  than a power of t (checked every 25 steps: the growth ratio over the last
  quarter is below the one over the second, where an instability's ratio
  over equal spans stays constant or grows).
* The batching workload in small, against the oracle: a three-level octree
  around the a = 0.7 hole (hole_forest, levels 0-2, p = 6) and four
  independent holes a = 0, 0.3, 0.5, 0.7 on uneven uniform forests (27, 64,
  27 and 125 blocks of 16^3), stepped together by one batched hipGraph (the
  block layout: one fused launch over every block of every problem per stage
  and each problem's planned fill), two RK4 steps; each problem's state is
  then compared with the oracle's two RK4 steps of the same problem alone."""
import numpy as np
import pytest

import oracle_lib
import tgh_desc as TD

pytestmark = pytest.mark.gpu

NTHREADS = 16


def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    return torch


@pytest.mark.parametrize("gauge", ["closed_form", "sampled"])
def test_single_hole_workload_at_full_size_stays_at_truncation(tgh, gauge):
    torch = torch_cuda()
    from test_gpu_parity import tgh_problem_from_desc
    kind = TD.TGH_GAUGE_KERR_SCHILD if gauge == "closed_form" else TD.TGH_GAUGE_SAMPLED
    desc = TD.kerr_schild_hole_case(N=16, nblocks=16, halfwidth=8.0, a=0.7, gauge=kind)
    p = tgh_problem_from_desc(tgh, desc)
    assert p.nblocks == 4096
    u = torch.empty(p.state_length, dtype=torch.float64, device="cuda")
    tgh.initial_state(u, p, 0.0)
    dt = 0.25 * (1 / 16) / tgh.max_speed(p, u, 0.0)
    L = tgh._load()
    e2, einf, c2 = [], [], []
    for k in range(4):
        tgh._check(L.tgh_rk4(p.ptr, tgh.C.c_void_p(u.data_ptr()), 25 * k * dt, dt, 25, 1))
        torch.cuda.synchronize()
        assert bool(torch.isfinite(u).all()), f"not finite after {25 * (k + 1)} steps"
        t = 25 * (k + 1) * dt
        rec = tgh.error_norms(p, u, t, shell=(1.21, 2.0))
        g2, _ = tgh.gauge_constraint_norms(p, u, t)
        e2.append(rec["err_l2"])
        einf.append(rec["err_linf"])
        c2.append(max(g2))
        print(f"{gauge}: step {25 * (k + 1)}, t = {t:.4f}: err_l2 {rec['err_l2']:.6e} err_linf "
              f"{rec['err_linf']:.6e} residual {rec['residual']:.6e} drift {rec['drift']:.6e} "
              f"gauge_l2 {max(g2):.6e}", flush=True)
    # at the truncation level of h = M/16, q = 4 (bounds a few times what the
    # two source kinds give: 7.3e-4 / 6.4e-4, 0.29 / 0.23, 1.1e-4 / 1.2e-4)
    assert e2[-1] < 2e-3 and einf[-1] < 1.0 and c2[-1] < 5e-4
    # and growing no faster than a power of t: the ratio over the last
    # quarter below the ratio over the second (an instability's ratio over
    # equal spans stays constant or grows)
    assert e2[3] / e2[2] < e2[1] / e2[0] and c2[3] / c2[2] < c2[1] / c2[0]


def test_octree_and_four_holes_in_one_batched_graph_match_the_oracle(tgh):
    torch = torch_cuda()
    ops = tgh.Operators(prolongation=6)
    probs, orcs = [], []
    case = tgh.kerr_schild_case(M=1.0, a=0.7, halfwidth=2.5, r_0=0.71, r_1=1.21, gauge_source="algebraic")
    f = tgh.hole_forest(None, case, N=16, roots=2, radii=(2.0, 1.0))
    assert f.maxlevel() == 2 and f.nleaves() == 120
    p = tgh.GHProblem(f, case, q=4, ops=ops)
    probs.append(p)
    orcs.append(oracle_lib.OracleProblem(TD.TghDesc.from_buffer_copy(p.desc), f.leaves, 6))
    for a, roots in ((0.0, 3), (0.3, 4), (0.5, 3), (0.7, 5)):
        c = tgh.kerr_schild_case(M=1.0, a=a, halfwidth=roots / 2, r_0=0.71, r_1=1.21, gauge_source="algebraic")
        q = tgh.GHProblem(tgh.gh_forest(None, c, N=16, roots=roots), c, q=4)
        probs.append(q)
        orcs.append(oracle_lib.OracleProblem(TD.TghDesc.from_buffer_copy(q.desc)))
    hs, us = [], []
    for i, (q, o) in enumerate(zip(probs, orcs)):
        u = o.fill_exact(0.0)
        o.add_noise(u, 0xC0F3 + i, 1e-8)
        hs.append(u)
        us.append(torch.from_numpy(u.copy()).cuda())
    dt = min(tgh.gh_dt(q, u, cfl=0.25) for q, u in zip(probs, us))
    batch = tgh.GHBatch(probs)
    tgh.gh_solve_batch(batch, us, (0.0, 2 * dt), dt=dt, graph=True)
    torch.cuda.synchronize()
    for i, (o, u) in enumerate(zip(orcs, hs)):
        for n in range(2):
            o.rk4_step(u, n * dt, dt, NTHREADS)
        got = us[i].cpu().numpy()
        diff = float(np.abs(got - u).max())
        print(f"problem {i} ({probs[i].nblocks} blocks): max |batch - oracle| = {diff:.3e}")
        assert np.isfinite(got).all()
        assert diff <= 1e-12 * max(1.0, float(np.abs(u).max()))
