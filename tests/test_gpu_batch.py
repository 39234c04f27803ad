"""The batched stepper (several independent holes of uneven block counts, one
head and one Pi launch per RK4 stage) and the plain per-block form, through
the C ABI. A problem in a batch runs the same kernel code on the same arrays
as the problem alone, so it must reproduce the brick stepper's bits; the
per-block form runs the fused block kernel one block at a time and must
reproduce the fused block form's bits (both are checked against the oracle
in test_gpu_parity.py)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    return torch


def ks_problem(tgh, a, roots, gauge_source="algebraic"):
    """kerr_schild_case on roots^3 blocks of 16^3 at h = 1/16."""
    case = tgh.kerr_schild_case(M=1.0, a=a, halfwidth=roots / 2, r_0=0.71, r_1=1.21,
                                gauge_source=gauge_source)
    return tgh.GHProblem(tgh.gh_forest(None, case, N=16, roots=roots), case, q=4)


def harmonic_problem(tgh, a, roots):
    case = tgh.harmonic_kerr_case(M=1.0, a=a, halfwidth=roots / 2, r_0=0.35, r_1=0.9)
    return tgh.GHProblem(tgh.gh_forest(None, case, N=16, roots=roots), case, q=4)


def noisy_states(tgh, probs, amplitude=1e-8):
    us = []
    for i, p in enumerate(probs):
        u = p.statevector()
        tgh.initial_state(u, p, 0.0)
        tgh.add_noise(u, p, 0xBA7C0 + i, amplitude)
        us.append(u)
    return us


def check_batch_equals_alone(tgh, probs, nsteps, graph, roundoff=False):
    """Every problem of the batch computes the bits it computes alone; with
    roundoff, to an ulp of the state's scale (a moving hole's head is its own
    kernel alone and in a batch, and their arithmetic may round apart)."""
    torch = torch_cuda()
    us = noisy_states(tgh, probs)
    ref = [u.clone() for u in us]
    dt = min(tgh.gh_dt(p, u, cfl=0.25) for p, u in zip(probs, us))
    for p, u in zip(probs, ref):
        tgh.gh_solve(p, u, (0.0, nsteps * dt), dt=dt, graph=graph)
    batch = tgh.GHBatch(probs)
    tgh.gh_solve_batch(batch, us, (0.0, nsteps * dt), dt=dt, graph=graph)
    torch.cuda.synchronize()
    for i, (u, r) in enumerate(zip(us, ref)):
        d = float((u - r).abs().max())
        print(f"problem {i} ({probs[i].nblocks} blocks): max |batch - alone| = {d:.3e}")
        assert torch.isfinite(u).all()
        if roundoff:
            assert d <= 4e-16 * float(r.abs().max()), f"problem {i} differs by {d}"
        else:
            assert torch.equal(u, r), f"problem {i} differs by {d}"
    # a second call on the same batch (tables and graph reused)
    tgh.gh_solve_batch(batch, us, (0.0, dt), dt=dt, graph=graph)
    for p, r in zip(probs, ref):
        tgh.gh_solve(p, r, (0.0, dt), dt=dt, graph=graph)
    for u, r in zip(us, ref):
        if roundoff:
            assert float((u - r).abs().max()) <= 4e-16 * float(r.abs().max())
        else:
            assert torch.equal(u, r)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_of_kerr_schild_holes_steps_each_as_alone(tgh, graph):
    probs = [ks_problem(tgh, 0.0, 3), ks_problem(tgh, 0.3, 4), ks_problem(tgh, 0.7, 5)]
    check_batch_equals_alone(tgh, probs, 3, graph)


def test_batch_of_sampled_source_holes_steps_each_as_alone(tgh):
    probs = [ks_problem(tgh, 0.5, 4, gauge_source=None), ks_problem(tgh, 0.0, 3, gauge_source=None)]
    check_batch_equals_alone(tgh, probs, 2, True)


def test_batch_of_harmonic_holes_steps_each_as_alone(tgh):
    probs = [harmonic_problem(tgh, 0.0, 4), harmonic_problem(tgh, 0.3, 3)]
    check_batch_equals_alone(tgh, probs, 2, True)


def test_batch_of_moving_frames_steps_each_as_alone(tgh):
    """Gauge waves: periodic frames refilled every stage, tables re-made every step."""
    probs = []
    for roots, N in ((2, 8), (3, 8)):
        case = tgh.gauge_wave_case(eps_ko=0.5, gamma0=1.0, gamma2=0.0)
        probs.append(tgh.GHProblem(tgh.gh_forest(None, case, N=N, roots=roots), case, q=4))
    check_batch_equals_alone(tgh, probs, 3, False)
    check_batch_equals_alone(tgh, probs, 2, True)


def test_batch_refuses_mixed_specialisations(tgh):
    torch_cuda()
    with pytest.raises(tgh.ArgumentError):
        tgh.GHBatch([ks_problem(tgh, 0.0, 3), harmonic_problem(tgh, 0.0, 3)])
    p = ks_problem(tgh, 0.0, 3)
    with pytest.raises(tgh.ArgumentError):
        tgh.GHBatch([p, p])


def test_per_block_form_equals_fused_block_form(tgh):
    torch = torch_cuda()
    p = ks_problem(tgh, 0.7, 3)
    u = noisy_states(tgh, [p])[0]
    v = u.clone()
    dt = tgh.gh_dt(p, u, cfl=0.25)
    L = tgh._load()
    tgh._check(L.tgh_rk4(p.ptr, C.c_void_p(u.data_ptr()), 0.0, dt, 2, tgh.TGH_RK4_PER_BLOCK))
    tgh._check(L.tgh_rk4(p.ptr, C.c_void_p(v.data_ptr()), 0.0, dt, 2, tgh.TGH_RK4_BLOCKS | tgh.TGH_RK4_FUSED))
    torch.cuda.synchronize()
    assert torch.equal(u, v), float((u - v).abs().max())
    assert L.tgh_rk4(p.ptr, C.c_void_p(u.data_ptr()), 0.0, dt, 1,
                     tgh.TGH_RK4_PER_BLOCK | tgh.TGH_RK4_SPLIT) == 1


def test_batch_of_moving_holes_steps_each_as_alone(tgh):
    """Moving harmonic holes (translate(boost(Harmonic, -v), x0)): time-dependent
    frames and tables, the moving core and layer from each problem's layer box,
    in one launch per kernel; equal to each hole stepped alone to an ulp of
    the state (measured: two and six entries of a near-zero Pi component one
    ulp apart, outside the layer)."""
    probs = []
    for a, roots, v in ((0.0, 3, 0.3), (0.3, 4, -0.2)):
        case = tgh.harmonic_kerr_case(M=1.0, a=a, v=v, x0=0.25, axis=1, halfwidth=roots / 2, r_0=0.35, r_1=0.9)
        probs.append(tgh.GHProblem(tgh.gh_forest(None, case, N=16, roots=roots), case, q=4))
    check_batch_equals_alone(tgh, probs, 2, True, roundoff=True)


def _forest_problem(tgh, a, halfwidth, roots, radii):
    case = tgh.kerr_schild_case(M=1.0, a=a, halfwidth=halfwidth, r_0=0.71, r_1=1.21, gauge_source="algebraic")
    f = tgh.hole_forest(None, case, N=16, roots=roots, radii=radii)
    return tgh.GHProblem(f, case, q=4, ops=tgh.Operators(prolongation=6))


@pytest.mark.parametrize("graph", [False, True])
def test_block_batch_of_forests_and_a_uniform_hole_steps_each_as_alone(tgh, graph):
    """Forests with levels among the problems: the batch runs on the block
    layout, its blocks run as bricks -- one head launch and one Pi launch over
    every block of every problem per stage (all levels of all holes) -- and
    each problem's result is bit for bit its own block-layout step in the same
    form."""
    torch = torch_cuda()
    probs = [_forest_problem(tgh, 0.5, 2.0, 4, (1.2,)), _forest_problem(tgh, 0.7, 1.5, 3, (0.9,)),
             ks_problem(tgh, 0.3, 3)]
    assert probs[0].forest.maxlevel() == 1 and probs[1].forest.maxlevel() == 1
    us = noisy_states(tgh, probs)
    ref = [u.clone() for u in us]
    dt = min(tgh.gh_dt(p, u, cfl=0.25) for p, u in zip(probs, us))
    L = tgh._load()
    flags = (4 | 8) | (1 if graph else 0)  # the block layout, blocks run as bricks, as the batch runs it
    for p, r in zip(probs, ref):
        tgh._check(L.tgh_rk4(p.ptr, C.c_void_p(r.data_ptr()), 0.0, dt, 3, flags))
    batch = tgh.GHBatch(probs)
    uarr = (C.c_void_p * 3)(*[u.data_ptr() for u in us])
    tgh._check(L.tgh_batch_rk4(batch.ptr, uarr, 0.0, dt, 3, 1 if graph else 0))
    torch.cuda.synchronize()
    for i, (u, r) in enumerate(zip(us, ref)):
        d = float((u - r).abs().max())
        print(f"problem {i} ({probs[i].nblocks} blocks): max |batch - alone| = {d:.3e}")
        assert torch.isfinite(u).all()
        assert torch.equal(u, r), f"problem {i} differs by {d}"
    # again on the same batch (tables and graph reused), and on swapped states
    tgh._check(L.tgh_batch_rk4(batch.ptr, uarr, 3 * dt, dt, 1, 1 if graph else 0))
    for p, r in zip(probs, ref):
        tgh._check(L.tgh_rk4(p.ptr, C.c_void_p(r.data_ptr()), 3 * dt, dt, 1, flags))
    torch.cuda.synchronize()
    for u, r in zip(us, ref):
        assert torch.equal(u, r)


def test_block_batch_refuses_moving_holes(tgh):
    torch_cuda()
    moving = tgh.kerr_schild_case(M=1.0, a=0.3, v=0.2, halfwidth=2.0, r_0=0.71, r_1=1.21, gauge_source="algebraic")
    fm = tgh.hole_forest(None, moving, N=16, roots=4, radii=(1.2,))
    pm = tgh.GHProblem(fm, moving, q=4, ops=tgh.Operators(prolongation=6))
    with pytest.raises(NotImplementedError):
        tgh.GHBatch([pm])
