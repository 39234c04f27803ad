"""The oracle's mesh path: the streaming kernel against the one-expression
assembly, the exact statements about flat space, the ghost fill, the
interior's branches, and the scheme's convergence against the numbers the
reference publishes (CODE.md:5318-5337)."""
import math

import numpy as np
import pytest

import oracle_lib as O
import tgh_desc as TD


def state_view(u, d):
    P = d.N + 1
    return u.reshape(-1, 20, P, P, P)


def test_minkowski_du_is_exactly_zero():
    """test/evolution_tests.jl:181-206."""
    for q in (2, 4, 6):
        for eps in (0.0, 0.5):
            N = 10 if q == 6 else 8
            d = TD.make_desc(q=q, N=N, nblocks=(2, 2, 2), gamma0=1.0, gamma2=-0.5, eps_ko=eps)
            p = O.OracleProblem(d)
            u = p.fill_exact(0.0)
            assert not u.any()
            du = p.gh_rhs(u, 3 / 7)
            assert not du.any()


CASES = {
    "gauge wave": TD.gauge_wave_case(N=8, roots=2, q=4, eps_ko=0.5, gamma0=1.0, gamma2=-0.5),
    "gauge wave q=2": TD.gauge_wave_case(N=8, roots=2, q=2, eps_ko=0.5, gamma0=1.0, gamma2=-0.5),
    "gauge wave q=6": TD.gauge_wave_case(N=10, roots=1, q=6, eps_ko=0.5, gamma0=1.0, gamma2=-0.5),
    "shifted minkowski": TD.make_desc(q=4, N=8, nblocks=(1, 1, 1), periodic=(0, 1, 1), lo=(-2, -2, -2),
                                      hi=(2, 2, 2), background=TD.TGH_BG_SHIFTED_MINKOWSKI,
                                      gauge=TD.TGH_GAUGE_SAMPLED, wave_amp=0.5, wave_len=2.0,
                                      gamma0=1.0, gamma2=-0.5, eps_ko=0.5),
    "kerr-schild off centre": TD.make_desc(q=4, N=8, nblocks=(1, 1, 1), periodic=(0, 0, 0),
                                           lo=(2, 2, 2), hi=(4, 4, 4), background=TD.TGH_BG_KERR_SCHILD,
                                           gauge=TD.TGH_GAUGE_KERR_SCHILD, M=1.0, a=0.9,
                                           gamma0=1.0, gamma2=-0.5, eps_ko=0.5),
}


@pytest.mark.parametrize("name", list(CASES))
def test_streaming_kernel_is_the_expanded_rhs(name):
    """test/evolution_tests.jl:141-179: < 1e-12 of rhs_scale at sample points."""
    d = CASES[name]
    p = O.OracleProblem(d)
    u = p.fill_exact(0.25)
    du = p.gh_rhs(u, 0.25)
    dv = state_view(du, d)
    N = d.N
    worst = 0.0
    nb = dv.shape[0]
    for b in range(nb):
        for (i, j, k) in ((0, 0, 0), (1, N // 2, N - 1), (N - 1, N - 1, N - 1), (N // 2, 2, N - 2), (N, N, N)):
            ref = p.host_rhs_at(b, i, j, k, 0.25)
            scale = p.rhs_scale(b, i, j, k)
            worst = max(worst, float(np.abs(dv[b, :, k, j, i] - ref).max()) / scale)
    assert worst < 1e-12, worst


def test_face_copies_carry_their_primaries_rhs():
    """Closed blocks: a face point is an own point of every block holding it,
    and its copies evaluate to the same bits."""
    d = TD.gauge_wave_case(N=8, roots=2, q=4, eps_ko=0.5, gamma0=1.0, gamma2=0.0)
    p = O.OracleProblem(d)
    u = p.fill_exact(0.1)
    du = state_view(p.gh_rhs(u, 0.1), d)
    coords = p.block_coords()
    idx = {tuple(c): b for b, c in enumerate(coords)}
    b0, b1 = idx[(0, 0, 0)], idx[(1, 0, 0)]
    np.testing.assert_array_equal(du[b0, :, :, :, 8], du[b1, :, :, :, 0])
    # and the periodic image: plane 8 of block (1,0,0) is plane 0 of (0,0,0)
    np.testing.assert_array_equal(du[b1, :, :, :, 8], du[b0, :, :, :, 0])


def test_dirichlet_hook_fills_ghosts_with_the_state_at_t():
    d = TD.make_desc(q=4, N=8, nblocks=(1, 1, 1), periodic=(0, 0, 0), lo=(0, 0, 0), hi=(2, 2, 2),
                     background=TD.TGH_BG_GAUGE_WAVE, wave_amp=1 / 20, wave_len=2.0)
    p = O.OracleProblem(d)
    t = 1 / 3
    u = p.fill_exact(t)
    p.prepare(u, t)
    S = 8 + 1 + 6
    W = p.work().reshape(1, 20, S, S, S)
    h = 0.25
    L = O.lib()
    import ctypes as C
    worst = 0.0
    for (i, j, k) in ((0, 4, 4), (1, 4, 4), (2, 4, 4), (4, 0, 4), (4, 4, 1), (12, 4, 4), (4, 14, 4), (0, 0, 0)):
        x = np.array([(i - 3) * h, (j - 3) * h, (k - 3) * h])
        s = np.zeros(20)
        L.orc_case_state(C.byref(d), t, x, s)
        worst = max(worst, float(np.abs(W[0, :, k, j, i] - s).max()))
    assert worst == 0.0


def test_dissipation_switch_is_exactly_eps_zero():
    """test/evolution_tests.jl:322-360."""
    dus = []
    for eps in (0.0, 1e-300):
        d = TD.gauge_wave_case(N=8, roots=1, q=4, eps_ko=eps, gamma0=1.0, gamma2=0.0)
        p = O.OracleProblem(d)
        u = p.fill_exact(0.0)
        dus.append(p.gh_rhs(u, 0.0))
    assert np.all(np.abs(dus[0] - dus[1]) <= 64 * np.finfo(float).eps * np.abs(dus[0]) + 1e-290)


def test_interior_branches():
    """test/interior_tests.jl:584-770: the core is exactly 0, outside r_1 is F,
    the layer is w F − ρ (u − u_exact)."""
    d = TD.kerr_schild_hole_case(N=16, nblocks=2, halfwidth=1.5, a=0.7, r0=0.71, r1=1.21)
    d.N, d.nblocks[:] = 12, [2, 2, 2]  # h = 1/8: r1 - r0 = 0.5 >= 2(G+1)h
    p = O.OracleProblem(d)
    u = p.fill_exact(0.0)
    p.add_noise(u, 2026, 1e-3)  # a function of position: face copies agree
    du = state_view(p.gh_rhs(u, 0.0), d)
    # the same state with no interior at all: F itself
    d2 = TD.TghDesc.from_buffer_copy(d)
    d2.interior = TD.TGH_INTERIOR_NONE
    F = state_view(O.OracleProblem(d2).gh_rhs(u, 0.0), d)
    uv = state_view(u, d)
    coords = p.block_coords()
    h = 3.0 / 24
    import ctypes as C
    L = O.lib()
    n_core = n_layer = n_out = 0
    for b, (bx, by, bz) in enumerate(coords):
        for k in range(13):
            for j in range(13):
                for i in range(13):
                    x = np.array([-1.5 + (bx * 12 + i) * h, -1.5 + (by * 12 + j) * h, -1.5 + (bz * 12 + k) * h])
                    r = np.linalg.norm(x)
                    got = du[b, :, k, j, i]
                    if r < 0.71:
                        assert not got.any()
                        n_core += 1
                    elif r >= 1.21:
                        np.testing.assert_array_equal(got, F[b, :, k, j, i])
                        n_out += 1
                    else:
                        s = (r - 0.71) / 0.5
                        w = O.lib().orc_smoothstep(s / 0.5)
                        rho = 4.0 * O.lib().orc_smoothstep((1 - s) / 0.5)
                        he, Pe, dh = np.zeros(10), np.zeros(10), np.zeros(30)
                        L.orc_background_state(C.byref(d), 0.0, x, he, Pe, dh)
                        ue = np.concatenate([he, Pe])
                        ref = w * F[b, :, k, j, i] - rho * (uv[b, :, k, j, i] - ue)
                        assert np.abs(got - ref).max() <= 1e-12 * max(1, np.abs(F[b, :, k, j, i]).max())
                        n_layer += 1
    assert n_core > 0 and n_layer > 0 and n_out > 0


def test_smoothstep_is_exact_at_its_ends():
    s = O.lib().orc_smoothstep
    assert s(0.0) == 0.0 and s(-1.0) == 0.0 and s(1.0) == 1.0 and s(2.0) == 1.0
    assert s(1 - 2 ** -53) <= 1.0
    assert abs(s(0.5) - 0.5) < 1e-15


def run_gauge_wave(roots, q=4, N=8, eps=0.0, gamma2=0.0, t_end=1 / 8, threads=8):
    """gh_errors (test/evolution_cases.jl:60-89)."""
    d = TD.gauge_wave_case(N=N, roots=roots, q=q, eps_ko=eps, gamma0=1.0, gamma2=gamma2)
    p = O.OracleProblem(d)
    u = p.fill_exact(0.0)
    dt = 0.25 * (1.0 / (N * roots)) / p.max_speed(u)
    n = max(1, math.ceil(t_end / dt))
    dt = t_end / n
    for s in range(n):
        p.rk4_step(u, s * dt, dt, threads)
    ue = p.fill_exact(t_end)
    err = state_view(u - ue, d)
    # volume_weighted_norm over the primaries, all 20 variables
    prim = err[:, :, :N, :N, :N]
    l2 = math.sqrt(float((prim ** 2).mean()))
    return 1.0 / (N * roots), l2


def test_gauge_wave_reproduces_the_published_errors():
    """CODE.md:5318-5337 and 4413-4416: q = 4, N = 8, roots 1, 2, 3, an eighth
    of a crossing at cfl 1/4: L2 errors 2.00e-4, 1.30e-5, 2.6e-6, rate 3.95."""
    res = [run_gauge_wave(r) for r in (1, 2, 3)]
    hs, l2 = [r[0] for r in res], [r[1] for r in res]
    assert f"{l2[0]:.2e}" == "2.00e-04"
    assert f"{l2[1]:.2e}" == "1.30e-05"
    assert f"{l2[2]:.1e}" == "2.6e-06"
    rate = np.polyfit(np.log(hs), np.log(l2), 1)[0]
    assert f"{rate:.2f}" == "3.95"


def test_gauge_wave_q2_rate():
    """CODE.md:5324: q = 2 converges at 1.99 in L2."""
    res = [run_gauge_wave(r, q=2) for r in (1, 2, 3)]
    rate = np.polyfit(np.log([r[0] for r in res]), np.log([r[1] for r in res]), 1)[0]
    assert abs(rate - 1.99) < 0.03


def test_static_kerr_schild_is_stationary_to_truncation_order():
    """F(u_exact) on a static hole is the truncation error: it falls ~2^q per
    halving of h at an evolved point away from the boundary."""
    errs = []
    for N, nb in ((8, 2), (16, 2)):
        d = TD.make_desc(q=4, N=N, nblocks=(nb,) * 3, periodic=(0, 0, 0), lo=(2.0, 1.0, -0.5),
                         hi=(4.0, 3.0, 1.5), background=TD.TGH_BG_KERR_SCHILD,
                         gauge=TD.TGH_GAUGE_KERR_SCHILD, M=1.0, a=0.7, gamma0=1.0, eps_ko=0.0)
        p = O.OracleProblem(d)
        u = p.fill_exact(0.0)
        du = state_view(p.gh_rhs(u, 0.0), d)
        coords = [tuple(c) for c in p.block_coords()]
        b = coords.index((1, 1, 1))
        errs.append(float(np.abs(du[b, :, 0, 0, 0]).max()))  # the point (3, 2, 0.5)
    ratio = errs[0] / errs[1]
    assert 10 < ratio < 24, (errs, ratio)
