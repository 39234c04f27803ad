"""HorizonDissipation (src/gauge.jl:1192-1264) in the oracle, pinned by what the
profile fixes rather than by retyping it: exactly eps_out at and outside r_h and
exactly eps_in at and inside r_1 (the right-hand side there is bitwise the one of
the constant amplitude), a profile whose two values agree is that number bit for
bit, and in between d_t h -- linear in the amplitude -- interpolates the two
constant-amplitude right-hand sides with the quintic smoothstep of
(r_h - r)/(r_h - r_1), computed here independently."""
import numpy as np
import pytest

import oracle_lib
import tgh_desc as TD

A, M = 0.5, 1.0
R_H = M + (M * M - A * A) ** 0.5  # horizon_min_radius, src/interior.jl:760
R0, R1 = 0.71, 1.71


def desc(**kw):
    d = TD.make_desc(q=4, N=16, nblocks=(3, 3, 3), periodic=(0, 0, 0), lo=(-3, -3, -3), hi=(3, 3, 3),
                     background=TD.TGH_BG_KERR_SCHILD, gauge=TD.TGH_GAUGE_KERR_SCHILD,
                     damping=TD.TGH_DAMPING_GAUSSIAN, gamma0_near=1.0, gamma0_far=0.1, gamma0_width=3.0,
                     interior=TD.TGH_INTERIOR_DAMPED, M=M, a=A, r0=R0, r1=R1, rho_max=4.0, **kw)
    return d


def smoothstep(s):  # src/interior.jl:117-120
    u = np.clip(s, 0.0, 1.0)
    return u * u * u * (10.0 + u * (6.0 * u - 15.0))


@pytest.fixture(scope="module")
def state():
    o = oracle_lib.OracleProblem(desc(eps_ko=0.5))
    u = o.fill_exact(0.0)
    o.add_noise(u, 4242, 1e-6)
    return u


def rhs(d, u):
    return oracle_lib.OracleProblem(d).gh_rhs(u, 0.0, oracle_lib.lib().orc_max_threads())


def radii(d):
    """r = |x| at every entry of the state vector (u[i, j, k, v, b], column-major)."""
    o = oracle_lib.OracleProblem(d)
    P, N = d.N + 1, d.N
    h = (d.hi[0] - d.lo[0]) / (d.nblocks[0] * N)
    B = o.block_coords()
    i = np.arange(P)
    out = []
    for b in range(B.shape[0]):
        x = d.lo[0] + (B[b, 0] * N + i) * h
        y = d.lo[1] + (B[b, 1] * N + i) * h
        z = d.lo[2] + (B[b, 2] * N + i) * h
        r = np.sqrt(x[None, None, :] ** 2 + y[None, :, None] ** 2 + z[:, None, None] ** 2)  # (k, j, i)
        out.append(np.broadcast_to(r[None], (20, P, P, P)))
    return np.concatenate([o.ravel() for o in out])


def horizon(eps_out, eps_in):
    return desc(eps_ko=eps_out, dissipation=TD.TGH_DISSIPATION_HORIZON, eps_in=eps_in, eps_r1=R1, eps_rh=R_H)


def test_profile_is_each_constant_beyond_its_radius(state):
    eo, ei = 0.5, 3.0
    fp = rhs(horizon(eo, ei), state)
    fo = rhs(desc(eps_ko=eo), state)
    fi = rhs(desc(eps_ko=ei), state)
    r = radii(desc())
    out, inn = r >= R_H, r <= R1
    assert out.sum() > 1000 and inn.sum() > 1000
    assert np.array_equal(fp[out], fo[out])
    assert np.array_equal(fp[inn], fi[inn])
    assert not np.array_equal(fo[inn], fi[inn])  # the two amplitudes do differ there


def test_profile_with_equal_values_is_that_number_bitwise(state):
    assert np.array_equal(rhs(horizon(0.5, 0.5), state), rhs(desc(eps_ko=0.5), state))


def test_blend_between_r1_and_rh_is_the_quintic_smoothstep(state):
    eo, ei = 0.5, 3.0
    d = desc()
    r = radii(d)
    P3 = (d.N + 1) ** 3
    hmask = np.zeros(20, bool)
    hmask[:10] = True
    comp = np.tile(np.repeat(hmask, P3), r.size // (20 * P3))  # the d_t h components
    f0, f1 = rhs(desc(eps_ko=0.0), state), rhs(desc(eps_ko=1.0), state)
    fp = rhs(horizon(eo, ei), state)
    eps = np.where(r >= R_H, eo, np.where(r <= R1, ei, eo + (ei - eo) * smoothstep((R_H - r) / (R_H - R1))))
    sel = comp & (r > R1) & (r < R_H)
    assert sel.sum() > 100
    pred = f0 + eps * (f1 - f0)
    scale = np.abs(f1 - f0)[sel].max()
    err = np.abs(fp - pred)[sel].max()
    assert err <= 1e-12 * scale
    # a profile centred on r_1 instead, or a reversed blend, would miss by O(1)
    wrong = f0 + (eo + (ei - eo) * smoothstep((r - R1) / (R_H - R1))) * (f1 - f0)
    assert np.abs(fp - wrong)[sel].max() > 1e-3 * scale
