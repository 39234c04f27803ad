"""The stepper's variants the hole workloads do not reach, and the verbs on a
segment of a longer state vector, on the device against the oracle.

* The :pasted limiter (gh_paste_kernel!, src/evolution.jl:880-905) run by
  tgh_rk4 on every stage value and on the step's result (the step limiter,
  src/evolution.jl:1438-1466), in each of the stepper's forms, against the
  oracle's RK4 with its stage limiter.
* Matter (the in-kernel adapter of src/evolution.jl:1295-1311) through
  tgh_gh_rhs and through tgh_rk4 after tgh_set_matter (MatterRHS,
  src/stepping.jl:150-154), against the oracle's right-hand side with the
  same fields added to the source.
* The contract of test/verbs_tests.jl:24-101: prepare, rhs (with matter),
  stage_limit, step_limit, max_speed and initial_state called on this
  package's segment of a longer vector (after another field set of three
  cell-centred fields, as the reference's script lays it out) are the
  problem's own run -- the verbs' RK4 equals tgh_rk4 on a vector of its own
  -- and nothing outside the segment is written.

Tolerances: 1e-12 of max(1, |oracle|) per component after two RK4 steps (the
oracle's bar elsewhere); between the verbs and tgh_rk4 1e-13, since both are
the device's arithmetic and differ only in where the RK4 combine rounds (the
stepper folds it into the right-hand side's epilogue)."""
import numpy as np
import pytest

import oracle_lib
import tgh_desc as TD
from test_gpu_parity import bounded_hole_case, ks_small, tgh_problem_from_desc

pytestmark = pytest.mark.gpu

NTHREADS = 16
FORMS = {"brick-fused": 2, "brick-split": 8, "blocks-fused": 4 | 2, "blocks-split": 4 | 8,
         "brick-split-graph": 8 | 1, "blocks-fused-graph": 4 | 2 | 1}


def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    return torch


def _rk4(tgh, p, ud, t0, dt, n, flags):
    tgh._check(tgh._load().tgh_rk4(p.ptr, tgh.C.c_void_p(ud.data_ptr()), t0, dt, n, flags))


def _rel(got, want):
    return float((np.abs(got - want) / np.maximum(np.abs(want), 1.0)).max())


def _radii(desc, coords):
    """(b, k, j, i) coordinate radius |x - center| of every owned point."""
    P, N = desc.N + 1, desc.N
    h = (desc.hi[0] - desc.lo[0]) / (desc.nblocks[0] * N)
    i = np.arange(P)
    out = np.zeros((len(coords), P, P, P))
    for b, (bx, by, bz) in enumerate(coords):
        x = desc.lo[0] + (bx * N + i) * h - desc.center[0]
        y = desc.lo[1] + (by * N + i) * h - desc.center[1]
        z = desc.lo[2] + (bz * N + i) * h - desc.center[2]
        out[b] = np.sqrt(z[:, None, None] ** 2 + y[None, :, None] ** 2 + x[None, None, :] ** 2)
    return out


def _matter_fields(desc, coords):
    """A smooth matter distribution, a function of position only (so the two
    copies of a point shared by neighbouring closed blocks agree): a Gaussian
    blob off the centre in rho, a momentum density along a rotation, and an
    anisotropic stress. Layout of tgh_matter: rho P^3 x nb, S P^3 x 3 x nb,
    Sij P^3 x 6 x nb (xx, xy, xz, yy, yz, zz)."""
    P, N = desc.N + 1, desc.N
    h = (desc.hi[0] - desc.lo[0]) / (desc.nblocks[0] * N)
    i = np.arange(P)
    nb = len(coords)
    rho = np.zeros((nb, P, P, P))
    S = np.zeros((nb, 3, P, P, P))
    Sij = np.zeros((nb, 6, P, P, P))
    c0 = (0.3, -0.2, 0.25)
    for b, (bx, by, bz) in enumerate(coords):
        x = (desc.lo[0] + (bx * N + i) * h)[None, None, :] + np.zeros((P, P, P))
        y = (desc.lo[1] + (by * N + i) * h)[None, :, None] + np.zeros((P, P, P))
        z = (desc.lo[2] + (bz * N + i) * h)[:, None, None] + np.zeros((P, P, P))
        g = np.exp(-((x - c0[0]) ** 2 + (y - c0[1]) ** 2 + (z - c0[2]) ** 2))
        rho[b] = 0.02 * g
        S[b] = 0.005 * np.stack([-y * g, x * g, 0.3 * z * g])
        xs = (x, y, z)
        pairs = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))
        Sij[b] = np.stack([0.004 * ((1.0 if a == c else 0.0) + 0.5 * xs[a] * xs[c]) * g for a, c in pairs])
    return rho.ravel(), S.ravel(), Sij.ravel()


def _device_matter(torch, fields):
    return tuple(torch.from_numpy(f.copy()).cuda() for f in fields)


# -- :pasted in the stepper ---------------------------------------------------

@pytest.mark.parametrize("form", list(FORMS))
def test_rk4_pasted_matches_oracle(tgh, form):
    """The a = 0.7 hole with the :pasted interior and 1e-6 noise everywhere:
    two RK4 steps of tgh_rk4 against the oracle's, and every point inside
    r_1 is the exact state afterwards (the paste has the last word, so the
    noise planted there is gone)."""
    torch = torch_cuda()
    desc = ks_small(interior=TD.TGH_INTERIOR_PASTED)
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    u = o.fill_exact(0.0)
    o.add_noise(u, 0x9A57ED, 1e-6)
    ud = torch.from_numpy(u.copy()).cuda()
    lam = o.max_speed(u)
    dt = 0.25 * (1 / 16) / lam
    for n in range(2):
        o.rk4_step(u, n * dt, dt, NTHREADS)
    _rk4(tgh, p, ud, 0.0, dt, 2, FORMS[form])
    g = ud.cpu().numpy()
    diff = _rel(g, u)
    print(f":pasted, 2 RK4 steps ({form}): max rel diff = {diff:.3e}")
    assert diff <= 1e-12
    inside = _radii(desc, o.block_coords()) < desc.r1
    assert inside.sum() > 1000
    ex = o.fill_exact(2 * dt).reshape(-1, 20, *inside.shape[1:])
    gv = g.reshape(ex.shape)
    # per component against max(1, |exact|): the paste's exact state and the
    # oracle's are two dual passes, equal to a few ulps of values up to ~2e3
    # near the core
    rel = (np.abs(gv - ex) / np.maximum(np.abs(ex), 1.0)).transpose(1, 0, 2, 3, 4)
    assert float(rel[:, inside].max()) <= 1e-12
    outside = np.abs(gv - ex).transpose(1, 0, 2, 3, 4)[:, ~inside]
    assert float(outside.max()) > 1e-8  # the noise outside r_1 is evolved, not pasted


def test_pasted_is_refused_by_the_batch(tgh):
    """A batch does not run the :pasted limiter: it says so instead of
    stepping without it."""
    torch_cuda()
    desc = ks_small(interior=TD.TGH_INTERIOR_PASTED)
    p = tgh_problem_from_desc(tgh, desc)
    import ctypes as C
    arr = (C.c_void_p * 1)(p.ptr.value)
    out = C.c_void_p()
    rc = tgh._load().tgh_batch_create(arr, 1, C.byref(out))
    assert rc == tgh.TGH_EUNSUPPORTED
    assert "pasted" in tgh._load().tgh_last_error().decode()


# -- matter -------------------------------------------------------------------

MATTER_CASES = {
    "kerr_schild_hole": lambda: ks_small(),
    "gauge_wave": lambda: TD.gauge_wave_case(N=8, roots=2, q=4, eps_ko=0.5, gamma0=1.0, gamma2=-0.5),
}


def _matter_pair(tgh, torch, name):
    desc = MATTER_CASES[name]()
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    fields = _matter_fields(desc, o.block_coords())
    o.set_matter(*fields)
    return desc, p, o, fields, _device_matter(torch, fields)


@pytest.mark.parametrize("name", list(MATTER_CASES))
def test_rhs_with_matter_matches_oracle(tgh, name):
    """tgh_gh_rhs with the matter fields against the oracle's right-hand side
    with the same fields, per component against rhs_scale; and the matter's
    contribution itself (with matter minus without) is resolved, so a dropped
    or mis-signed term fails."""
    torch = torch_cuda()
    from test_gpu_parity import rel_err, rhs_scale_all
    desc, p, o, fields, mat = _matter_pair(tgh, torch, name)
    t = 0.05
    u = o.fill_exact(t)
    o.add_noise(u, 0x3A77E2, 1e-6)
    ud = torch.from_numpy(u.copy()).cuda()
    du = torch.full_like(ud, float("nan"))
    tgh.gh_rhs(du, ud, p, t, matter=mat)
    ref = o.gh_rhs(u, t, NTHREADS)
    scale = rhs_scale_all(o.work(), desc)
    got = du.cpu().numpy()
    assert np.isfinite(got).all()
    err = rel_err(got, ref, scale, desc)
    du0 = torch.empty_like(ud)
    tgh.gh_rhs(du0, ud, p, t)
    contrib = float(np.abs(got - du0.cpu().numpy()).max())
    print(f"{name}: rhs with matter, max rel err {err:.3e}; matter's part up to {contrib:.3e}")
    assert err <= 1e-12
    assert contrib > 1e-3


@pytest.mark.parametrize("name", list(MATTER_CASES))
@pytest.mark.parametrize("form", ["blocks-fused", "blocks-split", "blocks-fused-graph"])
def test_rk4_with_matter_matches_oracle(tgh, name, form):
    """tgh_set_matter then two RK4 steps of tgh_rk4 against the oracle's RK4
    with the same fields; tgh_set_matter(NULL) returns the problem to vacuum."""
    torch = torch_cuda()
    desc, p, o, fields, mat = _matter_pair(tgh, torch, name)
    u = o.fill_exact(0.0)
    ud = torch.from_numpy(u.copy()).cuda()
    u_vac = ud.clone()
    dt = 0.25 * ((desc.hi[0] - desc.lo[0]) / (desc.nblocks[0] * desc.N)) / o.max_speed(u)
    for n in range(2):
        o.rk4_step(u, n * dt, dt, NTHREADS)
    tgh.set_matter(p, mat)
    _rk4(tgh, p, ud, 0.0, dt, 2, FORMS[form])
    g = ud.cpu().numpy()
    diff = _rel(g, u)
    tgh.set_matter(p, None)
    _rk4(tgh, p, u_vac, 0.0, dt, 2, FORMS[form])
    o.set_matter(None, None, None)
    v = o.fill_exact(0.0)
    for n in range(2):
        o.rk4_step(v, n * dt, dt, NTHREADS)
    vac = _rel(u_vac.cpu().numpy(), v)
    moved = float(np.abs(g - u_vac.cpu().numpy()).max())
    print(f"{name}, matter, 2 RK4 steps ({form}): max rel diff {diff:.3e}; vacuum again {vac:.3e}; "
          f"matter moved the state by {moved:.3e}")
    assert diff <= 1e-12 and vac <= 1e-12
    assert moved > 1e-6


# -- the verbs on a segment ---------------------------------------------------

def _segments(torch, n_other, n_gh, fill):
    """A vector of another field set (n_other entries) followed by this
    package's segment (n_gh), with the other set's entries and a tail of
    sentinels after the segment, every entry `fill`; returns (vector,
    segment view)."""
    n = n_other + n_gh + 64
    v = torch.full((n,), fill, dtype=torch.float64, device="cuda")
    return v, v[n_other:n_other + n_gh]


def test_verbs_on_a_segment_are_the_stepper(tgh):
    """The a = 0.7 hole with the :pasted interior and matter, on 64 blocks of
    16^3: a script's RK4 through prepare / rhs(matter) / stage_limit /
    step_limit on the segment after three cell-centred fields of another
    set equals tgh_rk4 (with tgh_set_matter) on a vector of its own, to 1e-13;
    max_speed and initial_state on the segment are the vector's own, bit for
    bit; settle on the segment leaves the stepper's closed-block state as it
    is; the other set's entries and the sentinels after the segment keep
    their bits, and rhs writes no entry of du outside the segment."""
    torch = torch_cuda()
    desc = ks_small(interior=TD.TGH_INTERIOR_PASTED)
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    mat = _device_matter(torch, _matter_fields(desc, o.block_coords()))
    n_gh = p.state_length
    n_other = 3 * desc.N ** 3 * p.nblocks  # three cell-centred fields (an even count: 16-byte aligned)
    assert n_other % 2 == 0
    u, seg = _segments(torch, n_other, n_gh, 0.0)
    gen = torch.Generator(device="cuda").manual_seed(20261109)
    u[:n_other] = torch.rand(n_other, dtype=torch.float64, device="cuda", generator=gen)
    u[n_other + n_gh:] = 7.25
    before = u.clone()

    # the initial state on the segment is the vector's own
    tgh.initial_state(seg, p, 0.0)
    u0 = torch.empty(n_gh, dtype=torch.float64, device="cuda")
    tgh.initial_state(u0, p, 0.0)
    assert torch.equal(seg, u0)
    tgh.add_noise(seg, p, 0xBEEF, 1e-7)
    tgh.add_noise(u0, p, 0xBEEF, 1e-7)
    assert torch.equal(seg, u0)
    start = u0.cpu().numpy()

    # the speed of the segment, and the step of a script
    lam = tgh.max_speed(p, seg, 0.0)
    assert lam == tgh.max_speed(p, u0, 0.0)
    dt = 0.25 * (1 / 16) / lam
    nsteps = 3

    # rhs writes its segment of du only
    du, dseg = _segments(torch, n_other, n_gh, float("nan"))
    tgh.prepare(p, seg, 0.0)
    tgh.rhs(dseg, p, 0.0, matter=mat)
    assert torch.isnan(du[:n_other]).all() and torch.isnan(du[n_other + n_gh:]).all()
    assert torch.isfinite(dseg).all()
    d0 = torch.empty_like(u0)
    tgh.gh_rhs(d0, u0, p, 0.0, matter=mat)
    assert torch.equal(dseg, d0)

    # the script's RK4 over the segment
    ks = [_segments(torch, n_other, n_gh, float("nan")) for _ in range(4)]
    y, yseg = _segments(torch, n_other, n_gh, float("nan"))
    for n in range(nsteps):
        t = n * dt
        for s, (c, ts) in enumerate(((0.0, t), (0.5, t + dt / 2), (0.5, t + dt / 2), (1.0, t + dt))):
            src = seg if s == 0 else yseg
            tgh.prepare(p, src, ts)
            tgh.rhs(ks[s][1], p, ts, matter=mat)
            if s < 3:
                cn = (0.5, 0.5, 1.0)[s]
                yseg.copy_(seg + (cn * dt) * ks[s][1])
                tgh.stage_limit(yseg, p, t + cn * dt)
        seg.copy_(seg + (dt / 6) * (ks[0][1] + 2 * ks[1][1] + 2 * ks[2][1] + ks[3][1]))
        tgh.step_limit(seg, p, t + dt)
    tgh.set_matter(p, mat)
    _rk4(tgh, p, u0, 0.0, dt, nsteps, 4 | 2)
    tgh.set_matter(p, None)
    ref = u0.cpu().numpy()
    got = seg.cpu().numpy()
    diff = _rel(got, ref)
    print(f"verbs on a segment vs tgh_rk4, {nsteps} steps: max rel diff {diff:.3e}; "
          f"|u_T - u_0| up to {float(np.abs(ref - start).max()):.3e}")
    assert diff <= 1e-13
    settled = seg.clone()
    tgh.settle(p, seg, nsteps * dt)
    assert torch.equal(seg, settled)
    assert torch.equal(u[:n_other], before[:n_other])
    assert torch.equal(u[n_other + n_gh:], before[n_other + n_gh:])
    for k, _ in ks:
        assert torch.isnan(k[:n_other]).all() and torch.isnan(k[n_other + n_gh:]).all()


@pytest.mark.parametrize("which", ["pasted", "bounds"])
def test_limiters_on_a_segment_are_the_vectors_own(tgh, which):
    """stage_limit on a segment at an offset, with failures planted in it (the
    :pasted overwrite inside r_1; the range projection on wild values in the
    gate): the segment comes out as the same call leaves a vector of its own,
    bit for bit, step_limit is stage_limit, and nothing outside the segment
    changes."""
    torch = torch_cuda()
    desc = ks_small(interior=TD.TGH_INTERIOR_PASTED) if which == "pasted" else bounded_hole_case()
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    n_gh = p.state_length
    n_other = 3 * desc.N ** 3 * p.nblocks
    u, seg = _segments(torch, n_other, n_gh, -3.5)
    h = o.fill_exact(0.1)
    rng = np.random.default_rng(5)
    pick = rng.random(h.shape) < 0.02
    h[pick] *= rng.choice([-40.0, 60.0, 1e3], size=int(pick.sum()))
    seg.copy_(torch.from_numpy(h))
    before = u.clone()
    own = torch.from_numpy(h.copy()).cuda()
    tgh.stage_limit(seg, p, 0.1)
    tgh.stage_limit(own, p, 0.1)
    assert torch.equal(seg, own)
    assert not torch.equal(seg, before[n_other:n_other + n_gh])  # the planted failures were limited
    want = h.copy()
    o.stage_limit(want, 0.1)
    assert _rel(seg.cpu().numpy(), want) <= 1e-12
    seg2 = before[n_other:n_other + n_gh].clone()
    tgh.step_limit(seg2, p, 0.1)
    assert torch.equal(seg2, seg)
    assert torch.equal(u[:n_other], before[:n_other])
    assert torch.equal(u[n_other + n_gh:], before[n_other + n_gh:])
