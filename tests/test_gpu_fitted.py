"""The tracked geometry and the :fitted interior on the device against the
oracle, through the C ABI: the right-hand side per component against
rhs_scale (≤ 1e-12, test/evolution_tests.jl:97-108), two RK4 steps, the
masked error norms, and the refusal of a :fitted problem without a target.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib
import tgh_desc as TD
from test_gpu_parity import rhs_scale_all, tgh_problem_from_desc, torch_cuda

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)
SHAPE = [1.5 * np.sqrt(4 * np.pi), 0.0, 0.04, 0.03, 0.0, 0.01, -0.02, 0.0, 0.015]  # lmax = 2


def _desc(variant, v=(0.0, 0.0, 0.0)):
    d = TD.kerr_schild_hole_case(N=16, nblocks=4, halfwidth=2.0, a=0.5, r0=0.65, r1=1.35,
                                 gauge=TD.TGH_GAUGE_KERR_SCHILD)
    d.interior = variant
    d.velocity[:] = list(v)
    d.shaped, d.shape_lmax = 1, 2
    for i, a in enumerate(SHAPE):
        d.shape[i] = a
    d.r_in, d.r_out, d.offset, d.thickness = 1.4, 1.6, 0.25, 0.5
    return d


def _target(o, u_exact, rng, rate_scale):
    n = len(u_exact)
    nb = n // (20 * 17 ** 3)
    S = rate_scale * rng.normal(size=n)
    return np.concatenate([u_exact.reshape(nb, 20, -1), S.reshape(nb, 20, -1)], axis=1).ravel()


def _pair(tgh, d, tw=None, t_f=0.0, rate=False, trail=0.0):
    torch = torch_cuda()
    p = tgh_problem_from_desc(tgh, d)
    o = oracle_lib.OracleProblem(d)
    tw_d = None
    if tw is not None:
        o.set_target(tw, t_f, rate=rate, trail=trail)
        tw_d = torch.from_numpy(tw).cuda()
        tgh._check(tgh._load().tgh_set_target(p.ptr, C.c_void_p(tw_d.data_ptr()), t_f, 1 if rate else 0, trail))
    return torch, p, o, tw_d


@pytest.mark.parametrize("which", ["shaped-damped", "fitted-rate", "fitted-moving-trail"])
def test_general_interior_rhs_matches_oracle(tgh, which):
    rng = np.random.default_rng(11)
    t = 0.05
    if which == "shaped-damped":
        d = _desc(TD.TGH_INTERIOR_DAMPED)
        torch, p, o, tw = _pair(tgh, d)
    else:
        v = (0.3, 0.0, 0.0) if which == "fitted-moving-trail" else (0.0, 0.0, 0.0)
        d = _desc(TD.TGH_INTERIOR_FITTED, v)
        o0 = oracle_lib.OracleProblem(d)
        tw = _target(o0, o0.fill_exact(0.0), rng, 0.1)
        torch, p, o, tw = _pair(tgh, d, tw, t_f=0.0, rate=True,
                                trail=0.5 if which == "fitted-moving-trail" else 0.0)
    u = o.fill_exact(t)
    o.add_noise(u, 0xF20, 1e-6)
    ud = torch.from_numpy(u.copy()).cuda()
    du = torch.full_like(ud, float("nan"))
    tgh.gh_rhs(du, ud, p, t)
    ref = o.gh_rhs(u, t, NTHREADS)
    scale = rhs_scale_all(o.work(), d)
    P = d.N + 1
    nb = len(u) // (20 * P ** 3)
    dd = np.abs(du.cpu().numpy() - ref).reshape(nb, 20, P, P, P) / scale[:, None]
    print(f"{which}: rhs max |GPU - oracle| / rhs_scale = {dd.max():.3e}")
    assert np.isfinite(du.cpu().numpy()).all()
    assert dd.max() <= 1e-12


@pytest.mark.parametrize("flags", [0, 4 | 2, 16, 1])
def test_fitted_rk4_matches_oracle(tgh, flags):
    rng = np.random.default_rng(12)
    d = _desc(TD.TGH_INTERIOR_FITTED, (0.3, 0.0, 0.0))
    o0 = oracle_lib.OracleProblem(d)
    tw = _target(o0, o0.fill_exact(0.0), rng, 0.05)
    torch, p, o, twd = _pair(tgh, d, tw, t_f=0.0, rate=True, trail=0.5)
    u = o.fill_exact(0.0)
    ud = torch.from_numpy(u.copy()).cuda()
    lam = o.max_speed(u)
    dt = 0.25 * 0.0625 / lam
    for n in range(2):
        o.rk4_step(u, n * dt, dt, NTHREADS)
    tgh._check(tgh._load().tgh_rk4(p.ptr, C.c_void_p(ud.data_ptr()), 0.0, dt, 2, flags))
    g = ud.cpu().numpy()
    diff = float((np.abs(g - u) / np.maximum(np.abs(u), 1.0)).max())
    print(f"fitted moving hole, 2 RK4 steps (flags {flags}): max rel diff = {diff:.3e}")
    assert diff <= 1e-12
    l2g, lig = tgh.error_norms(p, ud, 2 * dt)
    l2o, lio = o.error_norms(u, 2 * dt)
    assert abs(l2g - l2o) <= 1e-9 * l2o and abs(lig - lio) <= 1e-9 * lio


def test_fitted_without_target_is_refused(tgh):
    torch = torch_cuda()
    d = _desc(TD.TGH_INTERIOR_FITTED)
    p = tgh_problem_from_desc(tgh, d)
    u = torch.zeros(tgh._load().tgh_state_length(p.ptr), dtype=torch.float64, device="cuda")
    du = torch.empty_like(u)
    rc = tgh._load().tgh_gh_rhs(p.ptr, C.c_void_p(du.data_ptr()), C.c_void_p(u.data_ptr()), 0.0, None)
    assert rc == 1 and b"target" in tgh._load().tgh_last_error()


def test_batch_of_fitted_problems_follows_new_targets(tgh):
    """Two :fitted holes in one batch (the block layout) step as each alone,
    and a target refilled between calls is what the next batch call reads."""
    torch = torch_cuda()
    rng = np.random.default_rng(13)
    L = tgh._load()
    d = _desc(TD.TGH_INTERIOR_FITTED)
    o = oracle_lib.OracleProblem(d)
    u0 = o.fill_exact(0.0)
    probs, tws, us = [], [], []
    for i in range(2):
        p = tgh_problem_from_desc(tgh, d)
        tw = torch.from_numpy(_target(o, u0, rng, 0.05)).cuda()
        tgh._check(L.tgh_set_target(p.ptr, C.c_void_p(tw.data_ptr()), 0.0, 1, 0.0))
        probs.append(p)
        tws.append(tw)
        us.append(torch.from_numpy(u0.copy()).cuda())
    refs = [u.clone() for u in us]
    dt = 0.25 * 0.0625 / o.max_speed(u0)
    batch = tgh.GHBatch(probs)
    uarr = (C.c_void_p * 2)(*[u.data_ptr() for u in us])
    for rnd in range(2):
        if rnd == 1:  # new targets for both
            for p, tw in zip(probs, tws):
                tw.mul_(0.5)
                tgh._check(L.tgh_set_target(p.ptr, C.c_void_p(tw.data_ptr()), 0.0, 1, 0.0))
        for p, r in zip(probs, refs):
            tgh._check(L.tgh_rk4(p.ptr, C.c_void_p(r.data_ptr()), rnd * dt, dt, 1, 4 | 2 | 1))
        tgh._check(L.tgh_batch_rk4(batch.ptr, uarr, rnd * dt, dt, 1, 1))
        torch.cuda.synchronize()
        for u, r in zip(us, refs):
            assert torch.equal(u, r)
