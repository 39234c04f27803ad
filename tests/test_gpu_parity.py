"""Parity of the HIP path against the CPU oracle, through the C ABI.

Every comparison is per component against the reference's rhs_scale
(test/evolution_tests.jl:97-108): |GPU − oracle| / scale ≤ 1e-12, the bar of
the reference's own kernel test (test/evolution_tests.jl:141-179). Sizes are
ones the oracle finishes in seconds.
"""
import math
import os

import numpy as np
import pytest

import oracle_lib
import tgh_desc as TD

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)
TOL = 1e-12


def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    return torch


def rhs_scale_all(work, desc):
    """rhs_scale at every owned point, from a work array (numpy, vectorised)."""
    N, q = desc.N, desc.q
    G = q // 2 + 1
    S = N + 1 + 2 * G
    nb = len(work) // (S ** 3 * 20)
    W = work.reshape(nb, 20, S, S, S)  # (b, v, k, j, i)
    _, _, w2 = oracle_lib.weights(q, 2)
    r = q // 2
    h = (desc.hi[0] - desc.lo[0]) / (desc.nblocks[0] * N)
    own = slice(G, G + N + 1)
    scale = np.abs(W[:, :, own, own, own]).max(axis=1)
    for ax in (2, 3, 4):
        acc = 0
        for kk, wk in enumerate(w2):
            o = kk - r
            sl = [slice(None), slice(None), own, own, own]
            sl[ax] = slice(G + o, G + o + N + 1)
            acc = acc + wk * W[tuple(sl)]
        scale = np.maximum(scale, (np.abs(acc) / (h * h)).max(axis=1))
    scale = np.maximum(scale, 1.0)  # (b, k, j, i)
    return scale


def state_view(u, desc):
    N = desc.N
    P = N + 1
    nb = len(u) // (P ** 3 * 20)
    return u.reshape(nb, 20, P, P, P)


def make_pair(tgh, desc):
    torch = torch_cuda()
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    assert p.state_length == o.n
    return torch, p, o


def tgh_problem_from_desc(tgh, desc):
    """A GHProblem-like handle over a raw descriptor (the C ABI directly)."""
    import ctypes as C

    class Handle:
        pass

    h = Handle()
    h.desc = desc
    ptr = C.c_void_p()
    L = tgh._load()
    rc = L.tgh_problem_create(C.byref(tgh.TghDesc.from_buffer_copy(desc)), C.byref(ptr))
    tgh._check(rc)
    import torch
    tgh._check(L.tgh_set_stream(ptr, C.c_void_p(torch.cuda.current_stream().cuda_stream)))
    h._ptr = ptr
    h.ptr = ptr
    h.state_length = int(L.tgh_state_length(ptr))
    h.work_length = int(L.tgh_work_length(ptr))
    h.nblocks = desc.nblocks[0] * desc.nblocks[1] * desc.nblocks[2]
    h.case = None

    def work():
        import torch
        w = torch.empty(h.work_length, dtype=torch.float64, device="cuda")
        tgh._check(L.tgh_copy_work(ptr, C.c_void_p(w.data_ptr())))
        return w

    def coords():
        out = (C.c_int32 * (3 * h.nblocks))()
        tgh._check(L.tgh_block_coords(ptr, out))
        return np.array(list(out)).reshape(h.nblocks, 3)

    h.work = work
    h.block_coords = coords
    h.__del__ = lambda: L.tgh_problem_destroy(ptr)
    return h


def core_mask(desc):
    """(b, k, j, i) mask of owned points inside the frozen core r < r_0."""
    o = oracle_lib.OracleProblem(desc)
    coords = o.block_coords()
    P, N = desc.N + 1, desc.N
    h = (desc.hi[0] - desc.lo[0]) / (desc.nblocks[0] * N)
    i = np.arange(P)
    out = np.zeros((len(coords), P, P, P), dtype=bool)
    for b, (bx, by, bz) in enumerate(coords):
        x = desc.lo[0] + (bx * N + i) * h - desc.center[0]
        y = desc.lo[1] + (by * N + i) * h - desc.center[1]
        z = desc.lo[2] + (bz * N + i) * h - desc.center[2]
        r = np.sqrt(z[:, None, None] ** 2 + y[None, :, None] ** 2 + x[None, None, :] ** 2)
        out[b] = r < desc.r0
    return out


def rel_err(a, b, scale_state, desc):
    d = np.abs(state_view(a, desc) - state_view(b, desc))  # (b, v, k, j, i)
    return float((d / scale_state[:, None]).max())


def check_rhs(tgh, desc, t, noise=None, tol=TOL):
    torch, p, o = make_pair(tgh, desc)
    np.testing.assert_array_equal(p.block_coords(), o.block_coords())
    u = o.fill_exact(t)
    if noise is not None:
        o.add_noise(u, noise[0], noise[1])
    ud = torch.from_numpy(u.copy()).cuda()
    du = torch.full_like(ud, float("nan"))
    tgh.gh_rhs(du, ud, p, t)
    ref = o.gh_rhs(u, t, NTHREADS)
    work_o = o.work()
    work_g = p.work().cpu().numpy()
    scale = rhs_scale_all(work_o, desc)
    assert np.isfinite(du.cpu().numpy()).all()
    err = rel_err(du.cpu().numpy(), ref, scale, desc)
    return err, work_o, work_g, p, o, u, ud


GW_ROWS = [(2, 0.5, 8), (4, 0.0, 8), (4, 0.5, 8), (6, 0.5, 10), (8, 0.5, 10)]


@pytest.mark.parametrize("q,eps,N", GW_ROWS)
def test_gauge_wave_rhs_matches_oracle(tgh, q, eps, N):
    desc = TD.gauge_wave_case(N=N, roots=2, q=q, eps_ko=eps, gamma0=1.0, gamma2=-0.5)
    err, wo, wg, *_ = check_rhs(tgh, desc, 0.25)
    print(f"gauge wave q={q} eps={eps}: max rel err {err:.3e}")
    assert err <= TOL
    # the periodic fill copies values: the work arrays agree bit for bit
    np.testing.assert_array_equal(wg, wo)


def test_minkowski_noise_rhs_matches_oracle(tgh):
    desc = TD.minkowski_noise_case(N=16, q=4)
    err, wo, wg, *_ = check_rhs(tgh, desc, 0.0, noise=(0x5EED20260917, 1e-6))
    print(f"Minkowski + noise: max rel err {err:.3e}")
    assert err <= TOL
    np.testing.assert_array_equal(wg, wo)


def test_minkowski_is_exactly_stationary(tgh):
    torch = torch_cuda()
    for q in (2, 4, 6):
        N = 10 if q == 6 else 8
        desc = TD.make_desc(q=q, N=N, nblocks=(2, 2, 2), gamma0=1.0, gamma2=-0.5, eps_ko=0.5)
        p = tgh_problem_from_desc(tgh, desc)
        u = torch.zeros(p.state_length, dtype=torch.float64, device="cuda")
        du = torch.full_like(u, float("nan"))
        tgh.gh_rhs(du, u, p, 3 / 7)
        assert bool((du == 0).all())
        tgh.gh_solve(p, u, (0.0, 0.5), dt=0.05)
        assert bool((u == 0).all())


def ks_small(gauge=TD.TGH_GAUGE_KERR_SCHILD, interior=TD.TGH_INTERIOR_DAMPED):
    d = TD.kerr_schild_hole_case(N=16, nblocks=4, halfwidth=2.0, a=0.7, gauge=gauge)
    d.interior = interior
    return d


@pytest.mark.parametrize("interior", [TD.TGH_INTERIOR_DAMPED, TD.TGH_INTERIOR_FROZEN])
def test_kerr_schild_hole_rhs_matches_oracle(tgh, interior):
    desc = ks_small(interior=interior)
    err, wo, wg, p, o, u, ud = check_rhs(tgh, desc, 0.0)
    print(f"KS hole a=0.7 interior={interior}: max rel err {err:.3e}")
    assert err <= TOL
    # Dirichlet ghosts: the analytic state at this t, to roundoff of a dual pass
    denom = np.maximum(np.abs(wo), 1.0)
    assert float((np.abs(wg - wo) / denom).max()) <= 1e-13


def test_kerr_schild_noise_rhs_matches_oracle(tgh):
    desc = ks_small()
    err, *_ = check_rhs(tgh, desc, 0.0, noise=(12345, 1e-6))
    assert err <= TOL


def test_dirichlet_hook_is_the_exact_state_at_t(tgh):
    """test/evolution_tests.jl:239-288: a moving background, a non-zero t."""
    torch = torch_cuda()
    desc = TD.make_desc(q=4, N=8, nblocks=(1, 1, 1), periodic=(0, 0, 0), lo=(0, 0, 0), hi=(2, 2, 2),
                        background=TD.TGH_BG_GAUGE_WAVE, wave_amp=1 / 20, wave_len=2.0)
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    t = 1 / 3
    u = o.fill_exact(t)
    ud = torch.from_numpy(u.copy()).cuda()
    tgh.prepare(p, ud, t)
    o.prepare(u, t)
    wg, wo = p.work().cpu().numpy(), o.work()
    assert float(np.abs(wg - wo).max()) <= 1e-15
    # and a hook at another time would not have been zero
    o.prepare(o.fill_exact(0.0), 0.0)
    assert float(np.abs(o.work() - wo).max()) > 1e-2


# tgh_rk4's forms (include/tgh.h flags): the brick layout and the block
# layout, each with the right-hand side as one fused kernel or as head + Pi.
RK4_FORMS = {"brick-fused": 2, "brick-split": 8, "blocks-split": 4 | 8, "blocks-fused": 4 | 2,
             "per-block": 16}


@pytest.mark.parametrize("form", list(RK4_FORMS))
def test_rk4_config1_ten_steps_match_oracle(tgh, form):
    """The CPU-runnable workload: Minkowski + 1e-6 noise, one periodic 16^3
    block, 10 RK4 steps at cfl 1/4."""
    torch = torch_cuda()
    desc = TD.minkowski_noise_case(N=16, q=4)
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    u = o.fill_exact(0.0)
    o.add_noise(u, 0x5EED20260917, 1e-6)
    ud = torch.from_numpy(u.copy()).cuda()
    lam_o = o.max_speed(u)
    lam_g = tgh.max_speed(p, ud, 0.0)
    assert abs(lam_g - lam_o) <= 1e-14 * lam_o
    dt = 0.25 * (1 / 16) / lam_o
    for n in range(10):
        o.rk4_step(u, n * dt, dt, NTHREADS)
    tgh._check(tgh._load().tgh_rk4(p.ptr, tgh.C.c_void_p(ud.data_ptr()), 0.0, dt, 10, RK4_FORMS[form]))
    diff = float(np.abs(ud.cpu().numpy() - u).max())
    print(f"config 1, 10 RK4 steps ({form}): max |GPU - oracle| = {diff:.3e}")
    assert diff <= 1e-12 * max(1.0, float(np.abs(u).max()))


@pytest.mark.parametrize("form", list(RK4_FORMS))
def test_rk4_kerr_schild_two_steps_match_oracle(tgh, form):
    torch = torch_cuda()
    desc = ks_small()
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    u = o.fill_exact(0.0)
    ud = torch.from_numpy(u.copy()).cuda()
    lam = o.max_speed(u)
    assert abs(tgh.max_speed(p, ud, 0.0) - lam) <= 1e-13 * lam
    dt = 0.25 * (1 / 16) / lam
    for n in range(2):
        o.rk4_step(u, n * dt, dt, NTHREADS)
    tgh._check(tgh._load().tgh_rk4(p.ptr, tgh.C.c_void_p(ud.data_ptr()), 0.0, dt, 2, RK4_FORMS[form]))
    g = ud.cpu().numpy()
    scale = np.maximum(np.abs(u), 1.0)
    diff = float((np.abs(g - u) / scale).max())
    print(f"KS hole, 2 RK4 steps ({form}): max rel diff = {diff:.3e}")
    assert diff <= 1e-12
    l2g, lig = tgh.error_norms(p, ud, 2 * dt)
    l2o, lio = o.error_norms(u, 2 * dt)
    assert abs(l2g - l2o) <= 1e-9 * l2o and abs(lig - lio) <= 1e-9 * lio


@pytest.mark.parametrize("form", list(RK4_FORMS))
def test_fused_stepper_equals_verbs_and_graph_replay(tgh, form):
    """The fused RK4 (stage update in the RHS epilogue) against RK4 written
    with the verbs, and the hipGraph replay against the launch sequence."""
    torch = torch_cuda()
    desc = TD.gauge_wave_case(N=8, roots=2, q=4, eps_ko=0.5, gamma0=1.0, gamma2=0.0)
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    u0 = torch.from_numpy(o.fill_exact(0.0)).cuda()
    dt = 0.01
    flags = RK4_FORMS[form]
    a = u0.clone()
    tgh._check(tgh._load().tgh_rk4(p.ptr, tgh.C.c_void_p(a.data_ptr()), 0.0, dt, 3, flags))
    b = u0.clone()
    for rep in range(2):  # replays, with eager work in between
        b.copy_(u0)
        tgh._check(tgh._load().tgh_rk4(p.ptr, tgh.C.c_void_p(b.data_ptr()), 0.0, dt, 3, flags | 1))
        torch.cuda.synchronize()
        assert torch.equal(a, b), f"graph replay {rep} differs"
    c = u0.clone()
    k = [torch.empty_like(c) for _ in range(4)]
    for n in range(3):
        t = n * dt
        tgh.gh_rhs(k[0], c, p, t)
        y = c + (dt / 2) * k[0]
        tgh.gh_rhs(k[1], y, p, t + dt / 2)
        y = c + (dt / 2) * k[1]
        tgh.gh_rhs(k[2], y, p, t + dt / 2)
        y = c + dt * k[2]
        tgh.gh_rhs(k[3], y, p, t + dt)
        c = c + (dt / 6) * (k[0] + 2 * k[1] + 2 * k[2] + k[3])
    assert float((a - c).abs().max()) <= 1e-14


def test_rhs_is_pure_and_repeatable(tgh):
    torch = torch_cuda()
    desc = ks_small()
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    u = torch.from_numpy(o.fill_exact(0.0)).cuda()
    u0 = u.clone()
    d1, d2 = torch.empty_like(u), torch.empty_like(u)
    tgh.gh_rhs(d1, u, p, 0.125)
    tgh.gh_rhs(d2, u, p, 0.125)
    assert torch.equal(u, u0)
    assert torch.equal(d1, d2)
    assert bool(torch.isfinite(d1).all())


def test_initial_state_and_noise_match_oracle(tgh):
    """initial_state! (the exact state with the core rule) and add_noise! (the
    SplitMix64 draw of the position bits) on the device."""
    torch = torch_cuda()
    desc = ks_small()
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    u = torch.full((p.state_length,), float("nan"), dtype=torch.float64, device="cuda")
    tgh.initial_state(u, p, 0.0)
    ref = o.fill_exact(0.0)
    g = u.cpu().numpy()
    # Outside the core the state is the background's at the point, to 1e-12 of
    # itself (the bar of the kernel's parity; near r = 0.71 the a = 0.7 ring
    # singularity is 0.01 away and the dual pass is ill-conditioned at ~1e-13).
    # Inside the core the rule evaluates the background on the sphere r_0, where
    # one ulp of the projected position moves the metric by more.
    rel = state_view(np.abs(g - ref) / np.maximum(np.abs(ref), 1.0), desc)
    core = core_mask(desc)
    assert float(rel[~core[:, None].repeat(20, 1)].max()) <= 1e-12
    assert float(rel.max()) <= 1e-11
    tgh.add_noise(u, p, 0x5EED20260917, 1e-6)
    o.add_noise(ref, 0x5EED20260917, 1e-6)
    rel = state_view(np.abs(u.cpu().numpy() - ref) / np.maximum(np.abs(ref), 1.0), desc)
    assert float(rel[~core[:, None].repeat(20, 1)].max()) <= 1e-12


def test_max_speed_finds_a_planted_maximum(tgh):
    """The speed reduction visits every owned point: a shift planted at each
    of several points is the maximum wherever it sits."""
    torch = torch_cuda()
    desc = TD.make_desc(q=4, N=8, nblocks=(2, 1, 1), hi=(2.0, 1.0, 1.0), gamma0=1.0)
    p = tgh_problem_from_desc(tgh, desc)
    P = 9
    for (b, i, j, k) in ((0, 0, 0, 0), (1, 8, 8, 8), (0, 2, 3, 7), (1, 5, 0, 4)):
        u = torch.zeros(p.state_length, dtype=torch.float64, device="cuda")
        idx = ((b * 20 + 1) * P + k) * P * P + j * P + i  # h_tx
        u[idx] = 0.3
        lam = tgh.max_speed(p, u, 0.0)
        o = oracle_lib.OracleProblem(desc)
        assert abs(lam - o.max_speed(u.cpu().numpy())) <= 1e-14 * lam
        assert lam > 1.9


@pytest.mark.parametrize("interior", [TD.TGH_INTERIOR_DAMPED])
def test_kerr_schild_sampled_gauge_source_matches_oracle(tgh, interior):
    """The reference's default for a hole: H_a = −Γ_a sampled once
    (src/gauge.jl:156-184) through the core rule; second derivatives of the
    metric by jets on the device, by nested duals in the oracle."""
    desc = ks_small(gauge=TD.TGH_GAUGE_SAMPLED, interior=interior)
    err, *_ = check_rhs(tgh, desc, 0.0, noise=(777, 1e-6))
    print(f"KS hole, sampled Hsrc: max rel err {err:.3e}")
    assert err <= TOL


def test_shifted_minkowski_sampled_source_through_a_dirichlet_face(tgh):
    desc = TD.make_desc(q=4, N=8, nblocks=(2, 2, 2), periodic=(0, 1, 1), lo=(-2, -2, -2), hi=(2, 2, 2),
                        background=TD.TGH_BG_SHIFTED_MINKOWSKI, gauge=TD.TGH_GAUGE_SAMPLED,
                        wave_amp=0.5, wave_len=2.0, gamma0=1.0, gamma2=-0.5, eps_ko=0.5)
    err, *_ = check_rhs(tgh, desc, 0.3)
    assert err <= TOL


def _noise_growth(tgh, eps):
    """gh_noise_growth (test/evolution_cases.jl:167-184): 1e-8 noise on flat
    space, q = 4, one 8^3 block, 1000 RK4 steps at cfl 1/4."""
    torch = torch_cuda()
    desc = TD.make_desc(q=4, N=8, nblocks=(1, 1, 1), gamma0=1.0, gamma2=0.0, eps_ko=eps)
    p = tgh_problem_from_desc(tgh, desc)
    u = torch.zeros(p.state_length, dtype=torch.float64, device="cuda")
    tgh.add_noise(u, p, 20260917, 1e-8)
    lam = tgh.max_speed(p, u, 0.0)
    dt = 0.25 * (1 / 8) / lam
    l2_0, li_0 = tgh.error_norms(p, u, 0.0)
    tgh.gh_solve(p, u, (0.0, 1000 * dt), dt=dt, graph=True)
    l2_1, li_1 = tgh.error_norms(p, u, 1000 * dt)
    return 1000 * dt, l2_1 / l2_0, li_1 / li_0, li_1, bool(torch.isfinite(u).all())


def test_white_noise_stays_bounded_with_dissipation(tgh):
    """test/noise_tests.jl:23-48: with eps_KO = 1/2 the L2 ratio < 2, Linf < 5."""
    t_end, r2, ri, _, fin = _noise_growth(tgh, 0.5)
    print(f"eps=0.5: t_end={t_end:.2f} l2 x{r2:.3f} linf x{ri:.3f}")
    assert fin and t_end > 15 and r2 < 2 and ri < 5


def test_without_dissipation_the_noise_grows(tgh):
    """test/noise_tests.jl:50-62: the negative control, L2 grows by more than 2."""
    t_end, r2, ri, li, fin = _noise_growth(tgh, 0.0)
    print(f"eps=0: t_end={t_end:.2f} l2 x{r2:.3f} linf x{ri:.3f}")
    assert fin and r2 > 2 and li < 1e-5


@pytest.mark.parametrize("brick_form", ["brick-fused", "brick-split"])
@pytest.mark.parametrize("case", ["shifted-mixed-frame", "gauge-wave-dirichlet", "ks-hole"])
def test_brick_stepper_equals_block_stepper(tgh, case, brick_form):
    """The brick layout (every point once, a ghost frame around the brick)
    against the block layout (per-block ghosts, a fill per stage), 3 steps:
    a Dirichlet face beside periodic ones with a sampled gauge source, a
    moving Dirichlet hook (the frame refilled at every stage time), and the
    hole with its static frame, replayed from a hipGraph."""
    torch = torch_cuda()
    if case == "shifted-mixed-frame":
        desc = TD.make_desc(q=4, N=8, nblocks=(2, 2, 2), periodic=(0, 1, 1), lo=(-2, -2, -2), hi=(2, 2, 2),
                            background=TD.TGH_BG_SHIFTED_MINKOWSKI, gauge=TD.TGH_GAUGE_SAMPLED,
                            wave_amp=0.5, wave_len=2.0, gamma0=1.0, gamma2=-0.5, eps_ko=0.5)
        flags = 0
    elif case == "gauge-wave-dirichlet":
        desc = TD.make_desc(q=4, N=8, nblocks=(2, 1, 1), periodic=(0, 0, 0), lo=(0, 0, 0), hi=(2, 1, 1),
                            background=TD.TGH_BG_GAUGE_WAVE, wave_amp=1 / 20, wave_len=2.0, eps_ko=0.5)
        flags = 0
    else:
        desc = ks_small()
        flags = 1
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    u = o.fill_exact(0.0)
    o.add_noise(u, 77, 1e-7)
    u0 = torch.from_numpy(u).cuda()
    lam = tgh.max_speed(p, u0, 0.0)
    dt = 0.25 * (desc.hi[0] - desc.lo[0]) / (desc.nblocks[0] * desc.N) / lam
    a, b = u0.clone(), u0.clone()
    tgh._check(tgh._load().tgh_rk4(p.ptr, tgh.C.c_void_p(a.data_ptr()), 0.1, dt, 3,
                                   RK4_FORMS[brick_form] | flags))
    tgh._check(tgh._load().tgh_rk4(p.ptr, tgh.C.c_void_p(b.data_ptr()), 0.1, dt, 3,
                                   RK4_FORMS["blocks-fused"] | flags))
    scale = torch.clamp(b.abs(), min=1.0)
    diff = float(((a - b).abs() / scale).max())
    print(f"{case}: {brick_form} vs blocks, 3 steps: max rel diff = {diff:.3e}")
    assert diff <= 1e-13


def test_rk4_flags_are_checked(tgh):
    torch = torch_cuda()
    desc = TD.minkowski_noise_case(N=8, q=4)
    p = tgh_problem_from_desc(tgh, desc)
    u = torch.zeros(p.state_length, dtype=torch.float64, device="cuda")
    L = tgh._load()
    assert L.tgh_rk4(p.ptr, tgh.C.c_void_p(u.data_ptr()), 0.0, 0.01, 1, 2 | 8) == 1  # FUSED with SPLIT
    assert "exclude" in L.tgh_last_error().decode()
    assert L.tgh_rk4(p.ptr, tgh.C.c_void_p(u.data_ptr()), 0.0, 0.01, 1, 64) == 1
    assert "unknown" in L.tgh_last_error().decode()


def horizon_profile_case():
    """A spinning hole with the HorizonDissipation profile (eps_out 1/2 outside the
    horizon's axis radius, eps_in 3 inside r_1)."""
    a = 0.5
    rh = 1.0 + (1.0 - a * a) ** 0.5
    return TD.make_desc(q=4, N=16, nblocks=(3, 3, 3), periodic=(0, 0, 0), lo=(-3, -3, -3), hi=(3, 3, 3),
                        background=TD.TGH_BG_KERR_SCHILD, gauge=TD.TGH_GAUGE_KERR_SCHILD,
                        damping=TD.TGH_DAMPING_GAUSSIAN, gamma0_near=1.0, gamma0_far=0.1, gamma0_width=3.0,
                        interior=TD.TGH_INTERIOR_DAMPED, M=1.0, a=a, r0=0.71, r1=1.71, rho_max=4.0,
                        eps_ko=0.5, dissipation=TD.TGH_DISSIPATION_HORIZON, eps_in=3.0, eps_r1=1.71, eps_rh=rh)


def test_horizon_dissipation_rhs_matches_oracle(tgh):
    err, *_ = check_rhs(tgh, horizon_profile_case(), 0.0, noise=(4242, 1e-6))
    print(f"horizon dissipation profile: max rel err {err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("form", ["brick-split", "blocks-fused"])
def test_horizon_dissipation_rk4_matches_oracle(tgh, form):
    torch = torch_cuda()
    desc = horizon_profile_case()
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    u = o.fill_exact(0.0)
    o.add_noise(u, 4242, 1e-6)
    ud = torch.from_numpy(u.copy()).cuda()
    lam = o.max_speed(u)
    dt = 0.25 * 0.125 / lam
    for n in range(2):
        o.rk4_step(u, n * dt, dt, NTHREADS)
    tgh._check(tgh._load().tgh_rk4(p.ptr, tgh.C.c_void_p(ud.data_ptr()), 0.0, dt, 2, RK4_FORMS[form]))
    g = ud.cpu().numpy()
    diff = float((np.abs(g - u) / np.maximum(np.abs(u), 1.0)).max())
    print(f"horizon dissipation, 2 RK4 steps ({form}): max rel diff = {diff:.3e}")
    assert diff <= 1e-12



def bounded_hole_case():
    """ks_small with default_bounds(M = 1) and default_gate r_1 - 2 G h (src/bounds.jl:138-184)."""
    d = ks_small()
    h = (d.hi[0] - d.lo[0]) / (d.nblocks[0] * d.N)
    d.alpha_min, d.alpha_max, d.lambda_min, d.lambda_max, d.beta_max, d.K_max = 1 / 50, 50.0, 1 / 100, 1000.0, 10.0, 100.0
    d.r_gate = d.r1 - 2 * (d.q // 2 + 1) * h
    return d


def test_range_projection_matches_oracle_bitwise(tgh):
    """apply_bounds! on the device against the oracle on a state with wild
    values (and non-finite ones) scattered through the gate: the same points
    move, to the same bits (both evaluate the map without contraction)."""
    torch = torch_cuda()
    desc = bounded_hole_case()
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    u = o.fill_exact(0.0)
    rng = np.random.default_rng(7)
    pick = rng.uniform(size=u.size) < 0.03
    u[pick] += rng.uniform(-6, 6, pick.sum()) * 10 ** rng.uniform(-2, 1, pick.sum())
    u[rng.choice(u.size, 5, replace=False)] = np.nan
    ud = torch.from_numpy(u.copy()).cuda()
    n_o = o.apply_bounds(u)
    import ctypes as C
    n, nf, rm = C.c_int64(), C.c_int64(), C.c_double()
    tgh._check(tgh._load().tgh_apply_bounds(p.ptr, C.c_void_p(ud.data_ptr()), 0.0, C.byref(n), C.byref(nf),
                                            C.byref(rm)))
    g = ud.cpu().numpy()
    print(f"range projection: {n.value} points moved (oracle {n_o}), {nf.value} non-finite, r_max {rm.value:.3f}")
    assert n.value == n_o and n_o > 100
    same = (g == u) | (np.isnan(g) & np.isnan(u))
    assert same.all(), f"{(~same).sum()} entries differ"
    assert 0 < rm.value < desc.r_gate


@pytest.mark.parametrize("form", ["brick-split", "blocks-fused"])
def test_rk4_with_range_projection_matches_oracle(tgh, form):
    torch = torch_cuda()
    desc = bounded_hole_case()
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    u = o.fill_exact(0.0)
    o.add_noise(u, 99, 1e-6)
    ud = torch.from_numpy(u.copy()).cuda()
    lam = o.max_speed(u)
    dt = 0.25 * (1 / 16) / lam
    for n in range(2):
        o.rk4_step(u, n * dt, dt, NTHREADS)
    tgh._check(tgh._load().tgh_rk4(p.ptr, tgh.C.c_void_p(ud.data_ptr()), 0.0, dt, 2, RK4_FORMS[form]))
    g = ud.cpu().numpy()
    diff = float((np.abs(g - u) / np.maximum(np.abs(u), 1.0)).max())
    print(f"range projection, 2 RK4 steps ({form}): max rel diff = {diff:.3e}")
    assert diff <= 1e-12


@pytest.mark.parametrize("q", [2, 6, 8])
def test_brick_stepper_other_orders_equal_block_stepper(tgh, q):
    """The brick stepper's LDS tiles at q = 2, 6, 8 (halo G = q/2 + 1) against the
    block stepper, 3 steps of the gauge wave on a Dirichlet box with noise."""
    torch = torch_cuda()
    N = 10 if q >= 6 else 8
    desc = TD.make_desc(q=q, N=N, nblocks=(2, 2, 1), periodic=(0, 0, 1), lo=(0, 0, 0), hi=(2, 2, 1),
                        background=TD.TGH_BG_GAUGE_WAVE, wave_amp=1 / 20, wave_len=2.0, eps_ko=0.5,
                        gamma0=1.0, gamma2=-0.5)
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    u = o.fill_exact(0.0)
    o.add_noise(u, 31, 1e-7)
    u0 = torch.from_numpy(u).cuda()
    dt = 0.25 * (desc.hi[0] - desc.lo[0]) / (desc.nblocks[0] * desc.N) / tgh.max_speed(p, u0, 0.0)
    res = {}
    for form in ("brick-split", "brick-fused", "blocks-fused"):
        w = u0.clone()
        tgh._check(tgh._load().tgh_rk4(p.ptr, tgh.C.c_void_p(w.data_ptr()), 0.2, dt, 3, RK4_FORMS[form]))
        res[form] = w
    for form in ("brick-split", "brick-fused"):
        diff = float(((res[form] - res["blocks-fused"]).abs() / torch.clamp(res["blocks-fused"].abs(), min=1.0)).max())
        print(f"q={q} {form} vs blocks-fused: {diff:.3e}")
        assert diff <= 1e-13


# --- the Kerr hole in harmonic coordinates (harmonic_kerr_case) -------------

@pytest.mark.parametrize("a", [0.0, 0.3])
def test_harmonic_hole_rhs_matches_oracle(tgh, a):
    """harmonic_kerr_case (src/initialdata.jl:1300-1326): no gauge source; the
    Dirichlet ghosts and the layer's target are the harmonic chart's state."""
    desc = TD.harmonic_hole_case(a=a)
    err, wo, wg, *_ = check_rhs(tgh, desc, 0.0)
    print(f"harmonic hole a={a}: max rel err {err:.3e}")
    assert err <= TOL
    denom = np.maximum(np.abs(wo), 1.0)
    assert float((np.abs(wg - wo) / denom).max()) <= 1e-13


@pytest.mark.parametrize("eps", [0.0, 0.5])
def test_harmonic_kerr_off_centre_rhs_matches_oracle(tgh, eps):
    """The reference's own harmonic row (test/evolution_tests.jl:110-179):
    Harmonic(1, 9/10) on the box (2, 4)^3, N = 8, q = 4, gamma0 = 1,
    gamma2 = -1/2, t = 1/4, no interior, Dirichlet faces."""
    desc = TD.make_desc(q=4, N=8, nblocks=(1, 1, 1), periodic=(0, 0, 0), lo=(2, 2, 2), hi=(4, 4, 4),
                        background=TD.TGH_BG_HARMONIC, gauge=TD.TGH_GAUGE_NONE, M=1.0, a=0.9,
                        gamma0=1.0, gamma2=-0.5, eps_ko=eps)
    err, wo, wg, *_ = check_rhs(tgh, desc, 0.25)
    print(f"harmonic Kerr a=0.9 off centre eps={eps}: max rel err {err:.3e}")
    assert err <= TOL
    assert float(np.abs(wg - wo).max()) <= 1e-15


@pytest.mark.parametrize("form", ["brick-split", "brick-fused", "blocks-split", "blocks-fused"])
def test_rk4_harmonic_hole_two_steps_match_oracle(tgh, form):
    torch = torch_cuda()
    desc = TD.harmonic_hole_case(a=0.3)
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    u = o.fill_exact(0.0)
    o.add_noise(u, 777, 1e-7)
    ud = torch.from_numpy(u.copy()).cuda()
    lam = o.max_speed(u)
    assert abs(tgh.max_speed(p, ud, 0.0) - lam) <= 1e-13 * lam
    dt = 0.25 * (1 / 16) / lam
    for n in range(2):
        o.rk4_step(u, n * dt, dt, NTHREADS)
    tgh._check(tgh._load().tgh_rk4(p.ptr, tgh.C.c_void_p(ud.data_ptr()), 0.0, dt, 2, RK4_FORMS[form]))
    g = ud.cpu().numpy()
    scale = np.maximum(np.abs(u), 1.0)
    diff = float((np.abs(g - u) / scale).max())
    print(f"harmonic hole, 2 RK4 steps ({form}): max rel diff = {diff:.3e}")
    assert diff <= 1e-12


# --- the moving hole (translate(boost(bg, -v), x0), the moving core) -------

MOVING = {"harmonic-x": dict(background=TD.TGH_BG_HARMONIC, axis=1),
          "kerr-schild-z": dict(background=TD.TGH_BG_KERR_SCHILD, axis=3),
          "kerr-schild-x": dict(background=TD.TGH_BG_KERR_SCHILD, axis=1)}


@pytest.mark.parametrize("name", list(MOVING))
def test_moving_hole_rhs_matches_oracle(tgh, name):
    """The boosted background in the Dirichlet ghosts and the layer's target,
    the layer's (1 - w) d_t u_e, and the moving :damped core's ADM relaxation
    with its target's rate (src/evolution.jl:526-583, 652-680), at t = 0.3."""
    desc = TD.moving_hole_case(**MOVING[name])
    err, wo, wg, *_ = check_rhs(tgh, desc, 0.3)
    print(f"moving hole {name}: max rel err {err:.3e}")
    assert err <= TOL
    denom = np.maximum(np.abs(wo), 1.0)
    assert float((np.abs(wg - wo) / denom).max()) <= 1e-13


@pytest.mark.parametrize("form", ["brick-split", "blocks-fused", "blocks-split"])
@pytest.mark.parametrize("case", ["moving-hole", "gauge-wave"])
def test_time_dependent_graph_replays_equal_launches(tgh, case, form):
    """The stage clock: a step of a time-dependent case (a moving hole's frame,
    layer and core; the gauge wave's frame) captured once and replayed step
    after step -- each replay reading its time from a device scalar set just
    before it -- equals the same steps launched one by one, bit for bit; a
    second call that continues from the end reuses the graph."""
    torch = torch_cuda()
    desc = TD.moving_hole_case(**MOVING["harmonic-x"]) if case == "moving-hole" else TD.gauge_wave_case(N=8, roots=2)
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    t0 = 0.2
    u = o.fill_exact(t0)
    dt = 0.25 * (1 / 16) / o.max_speed(u, t0)
    L = tgh._load()
    a = torch.from_numpy(u.copy()).cuda()
    b = a.clone()
    flags = RK4_FORMS[form]
    for t, n in ((t0, 4), (t0 + 4 * dt, 2)):
        tgh._check(L.tgh_rk4(p.ptr, tgh.C.c_void_p(a.data_ptr()), t, dt, n, flags | tgh.TGH_RK4_GRAPH))
        tgh._check(L.tgh_rk4(p.ptr, tgh.C.c_void_p(b.data_ptr()), t, dt, n, flags))
    torch.cuda.synchronize()
    d = float((a - b).abs().max())
    print(f"{case} ({form}): 6 replayed steps against launches, max |diff| = {d:.3e}")
    assert torch.equal(a, b)
    assert float((a - torch.from_numpy(u).cuda()).abs().max()) > 0  # the state moved


@pytest.mark.parametrize("form", ["brick-split", "blocks-fused"])
@pytest.mark.parametrize("name", ["harmonic-x", "kerr-schild-z"])
def test_rk4_moving_hole_two_steps_match_oracle(tgh, name, form):
    torch = torch_cuda()
    desc = TD.moving_hole_case(**MOVING[name])
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    t0 = 0.2
    u = o.fill_exact(t0)
    ud = torch.from_numpy(u.copy()).cuda()
    lam = o.max_speed(u, t0)
    dt = 0.25 * (1 / 16) / lam
    for n in range(2):
        o.rk4_step(u, t0 + n * dt, dt, NTHREADS)
    tgh._check(tgh._load().tgh_rk4(p.ptr, tgh.C.c_void_p(ud.data_ptr()), t0, dt, 2, RK4_FORMS[form]))
    g = ud.cpu().numpy()
    scale = np.maximum(np.abs(u), 1.0)
    diff = float((np.abs(g - u) / scale).max())
    print(f"moving hole {name}, 2 RK4 steps ({form}): max rel diff = {diff:.3e}")
    assert diff <= 1e-12
