"""Reflecting faces on the device (the mirrors of src/boundaries.jl:84-142,
test/reflection_tests.jl) against the oracle, through the C ABI:

* the ghost fill of the octant bit for bit (Minkowski with noise: mirrored
  ghosts are the interior's values times the parities, the wall's odd
  variables zero) and to the Dirichlet data's roundoff on a Kerr-Schild hole;
* the right-hand side and RK4 on the octant of the a = 0 hole and on the
  bitant of the a = 0.7 hole, in each stepper form;
* the reference's own check without the oracle: the device's octant equals
  the device's mirrored box to roundoff after 32 steps;
* add_noise's projection of the low walls, bit for bit, and the settled
  walls tgh_rk4 returns;
* the rotating seam (src/boundaries.jl:145-205, test/rotation_tests.jl): the
  fill and RK4 on the rotating octant of the a = 3/10 hole against the
  oracle, the device's octant against its box, the seam planes' noise;
* both on a two-level forest (the plan's mirrored and turned entries): the
  fill and RK4 against the oracle;
* the refusals: a wrapped and mirrored axis, half-open blocks, a batch.
Tolerances as the other parity tests: 1e-12 of rhs_scale or of max(1, |u|)."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
import tgh_desc as TD
from test_gpu_parity import RK4_FORMS, check_rhs, tgh_problem_from_desc

pytestmark = pytest.mark.gpu

NTHREADS = 16


def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    return torch


def minkowski_octant(N=8, nb=2, L=4.0):
    return TD.make_desc(q=4, N=N, nblocks=(nb,) * 3, periodic=(0, 0, 0), lo=(0.0,) * 3, hi=(L,) * 3,
                        background=TD.TGH_BG_MINKOWSKI, gamma0=1.0, gamma2=0.0, eps_ko=0.5,
                        reflecting=((1, 0), (1, 0), (1, 0)))


def hole_octant():
    """The a = 0 Kerr-Schild hole on the octant [0, 2]^3, 8 blocks of 16^3
    (hole_case(; octant = true), the closed-form source, :damped layer)."""
    d = TD.kerr_schild_hole_case(N=16, nblocks=2, halfwidth=2.0, a=0.0, r0=0.4, r1=1.15)
    d.lo[:] = [0.0] * 3
    d.hi[:] = [2.0] * 3
    for dd in range(3):
        d.reflecting[dd][0] = 1
    return d


def hole_bitant():
    """The a = 0.7 hole on the half z >= 0 of [-2, 2]^2 x [0, 2] (hole_case(;
    octant = :bitant)), 4 x 4 x 2 blocks of 16^3."""
    d = TD.kerr_schild_hole_case(N=16, nblocks=4, halfwidth=2.0, a=0.7)
    d.nblocks[2] = 2
    d.lo[2] = 0.0
    d.reflecting[2][0] = 1
    return d


def _noisy(o, seed, amp):
    u = o.fill_exact(0.0)
    o.add_noise(u, seed, amp)
    return u


@pytest.mark.parametrize("case", ["minkowski_octant", "hole_octant"])
def test_fill_mirrors_the_interior_with_the_parities(tgh, case):
    torch = torch_cuda()
    desc = minkowski_octant() if case == "minkowski_octant" else hole_octant()
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    u = _noisy(o, 0x0C7A, 1e-6)
    ud = torch.from_numpy(u.copy()).cuda()
    tgh.prepare(p, ud, 0.0)
    o.prepare(u, 0.0)
    wg, wo = p.work().cpu().numpy(), o.work()
    err = float((np.abs(wg - wo) / np.maximum(np.abs(wo), 1.0)).max())
    print(f"{case}: fill |GPU - oracle| / max(1, |oracle|) = {err:.3e}")
    if case == "minkowski_octant":
        assert np.array_equal(wg, wo)  # copies, signs and zeros: no arithmetic to round
    else:
        assert err <= 1e-13


@pytest.mark.parametrize("case", ["hole_octant", "hole_bitant"])
def test_rhs_on_a_mirrored_hole_matches_oracle(tgh, case):
    desc = hole_octant() if case == "hole_octant" else hole_bitant()
    err, *_ = check_rhs(tgh, desc, 0.0, noise=(0x7E11, 1e-6))
    print(f"{case}: rhs max rel err {err:.3e}")
    assert err <= 1e-12


@pytest.mark.parametrize("form", ["brick-split", "brick-fused", "blocks-fused", "blocks-split"])
@pytest.mark.parametrize("case", ["hole_octant", "hole_bitant"])
def test_rk4_on_a_mirrored_hole_matches_oracle(tgh, case, form):
    torch = torch_cuda()
    desc = hole_octant() if case == "hole_octant" else hole_bitant()
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    u = _noisy(o, 0x7E12, 1e-6)
    ud = torch.from_numpy(u.copy()).cuda()
    lam = o.max_speed(u)
    dt = 0.25 * (1 / 16) / lam
    for n in range(2):
        o.rk4_step(u, n * dt, dt, NTHREADS)
    tgh._check(tgh._load().tgh_rk4(p.ptr, tgh.C.c_void_p(ud.data_ptr()), 0.0, dt, 2, RK4_FORMS[form]))
    g = ud.cpu().numpy()
    diff = float((np.abs(g - u) / np.maximum(np.abs(u), 1.0)).max())
    print(f"{case}, 2 RK4 steps ({form}): max rel diff {diff:.3e}")
    assert diff <= 1e-12


def _walls(desc, o, arr):
    """The largest |odd variable| on the reflecting walls of a state."""
    L = oracle_lib.lib()
    P, N = desc.N + 1, desc.N
    U = arr.reshape(-1, 20, P, P, P)
    worst = 0.0
    for b, bc in enumerate(o.block_coords()):
        for d in range(3):
            if not desc.reflecting[d][0] or bc[d] != 0:
                continue
            sl = [slice(None)] * 3
            sl[2 - d] = 0  # (k, j, i) order: axis d is index 2 - d
            for v in range(20):
                if (L.orc_parity_odd_mask(v) >> d) & 1:
                    worst = max(worst, float(np.abs(U[b, v][tuple(sl)]).max()))
    return worst


def test_noise_projection_and_settled_walls(tgh):
    """tgh_add_noise projects the low walls onto the parity as the oracle's
    add_noise does (bit for bit); after tgh_rk4 the walls' odd variables are
    zero (the settled state gh_solve returns) while the oracle's unsettled
    stages leave them at roundoff."""
    torch = torch_cuda()
    desc = minkowski_octant()
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    u = o.fill_exact(0.0)
    ud = torch.from_numpy(u.copy()).cuda()
    o.add_noise(u, 0xBEE5, 1e-8)
    tgh.add_noise(ud, p, 0xBEE5, 1e-8)
    assert np.array_equal(ud.cpu().numpy(), u)
    assert _walls(desc, o, u) == 0.0
    dt = 0.25 / (4 * np.sqrt(3.0))
    for n in range(4):
        o.rk4_step(u, n * dt, dt, NTHREADS)
    tgh._check(tgh._load().tgh_rk4(p.ptr, tgh.C.c_void_p(ud.data_ptr()), 0.0, dt, 4, 0))
    g = ud.cpu().numpy()
    print(f"odd on the walls after 4 steps: GPU {_walls(desc, o, g):.3e}, oracle {_walls(desc, o, u):.3e}")
    assert _walls(desc, o, g) == 0.0
    assert _walls(desc, o, u) <= 64 * np.finfo(float).eps * float(np.abs(u).max())
    assert float(np.abs(g - u).max()) <= 1e-12 * max(1.0, float(np.abs(u).max()))


@pytest.mark.parametrize("form", ["brick-split", "blocks-fused"])
def test_the_octant_is_the_mirrored_box_on_the_device(tgh, form):
    """test/reflection_tests.jl:124-187 on the device alone: the octant
    [0, 4]^3 with noise inside [0, 1]^3 and the box [-4, 4]^3 with the same
    noise extended by parity, 32 steps of dt = h / (4 sqrt 3); every owned
    point of the octant equals the box's at its position to 64 eps of the
    data's scale."""
    torch = torch_cuda()
    od = minkowski_octant()
    bd = TD.make_desc(q=4, N=8, nblocks=(4,) * 3, periodic=(0, 0, 0), lo=(-4.0,) * 3, hi=(4.0,) * 3,
                      background=TD.TGH_BG_MINKOWSKI, gamma0=1.0, gamma2=0.0, eps_ko=0.5)
    po, pb = tgh_problem_from_desc(tgh, od), tgh_problem_from_desc(tgh, bd)
    oo, ob = oracle_lib.OracleProblem(od), oracle_lib.OracleProblem(bd)
    from test_oracle_reflect import _odd, _points
    L = oracle_lib.lib()
    P, h = od.N + 1, 0.25
    uo = _noisy(oo, 1, 1e-8)
    xo = _points(od, oo.block_coords())
    Uo = uo.reshape(-1, 20, P, P, P)
    for v in range(20):
        Uo[:, v][xo.max(axis=-1) > 1.0] = 0.0
    odd = _odd(L)
    at = {}
    for b in range(xo.shape[0]):
        for idx in np.ndindex(P, P, P):
            at.setdefault(tuple(int(round(c / h)) for c in xo[b][idx]), (b, idx))
    ub = ob.fill_exact(0.0)
    Ub = ub.reshape(-1, 20, P, P, P)
    xb = _points(bd, ob.block_coords())
    for b in range(xb.shape[0]):
        for idx in np.ndindex(P, P, P):
            x = xb[b][idx]
            if np.abs(x).max() >= 4.0:
                continue
            bo, io = at[tuple(int(round(abs(c) / h)) for c in x)]
            for v in range(20):
                s = -1.0 if sum(1 for d in range(3) if x[d] < 0 and odd[v, d]) % 2 else 1.0
                Ub[b, v][idx] = s * Uo[bo, v][io]
    go, gb = torch.from_numpy(uo.copy()).cuda(), torch.from_numpy(ub.copy()).cuda()
    dt = h / (4 * np.sqrt(3.0))
    for p, g in ((po, go), (pb, gb)):
        tgh._check(tgh._load().tgh_rk4(p.ptr, tgh.C.c_void_p(g.data_ptr()), 0.0, dt, 32, RK4_FORMS[form]))
    Go = go.cpu().numpy().reshape(-1, 20, P, P, P)
    Gb = gb.cpu().numpy().reshape(-1, 20, P, P, P)
    box_at = {}
    for b in range(xb.shape[0]):
        for idx in np.ndindex(P, P, P):
            box_at.setdefault(tuple(int(round(c / h)) for c in xb[b][idx]), (b, idx))
    diff = 0.0
    for b in range(xo.shape[0]):
        for idx in np.ndindex(P, P, P):
            bb, ib = box_at[tuple(int(round(c / h)) for c in xo[b][idx])]
            diff = max(diff, float(np.abs(Go[b][(slice(None),) + idx] - Gb[bb][(slice(None),) + ib]).max()))
    scale = float(np.abs(Go).max())
    print(f"device octant vs device box ({form}): max|diff| {diff:.3e}, scale {scale:.3e}")
    assert scale > 1e-9 and diff <= 64 * np.finfo(float).eps * scale


def test_mirrors_are_refused_where_they_are_not_built(tgh):
    torch_cuda()
    L = tgh._load()
    d = minkowski_octant()
    d.periodic[0] = 1
    out = C.c_void_p()
    assert L.tgh_problem_create(C.byref(tgh.TghDesc.from_buffer_copy(d)), C.byref(out)) == tgh.TGH_EINVAL
    assert "periodic" in L.tgh_last_error().decode()
    d = minkowski_octant()
    d.ownership = tgh.Forest.OWNERSHIP["half_open"]
    leaves = (C.c_int32 * 32)(*[x for k in range(2) for j in range(2) for i in range(2) for x in (0, i, j, k)])
    rc = L.tgh_problem_create_forest(C.byref(tgh.TghDesc.from_buffer_copy(d)), leaves, 8, 6, C.byref(out))
    assert rc == tgh.TGH_EUNSUPPORTED  # the wall plane must be owned: closed blocks only
    p = tgh_problem_from_desc(tgh, minkowski_octant())
    arr = (C.c_void_p * 1)(p.ptr.value)
    assert L.tgh_batch_create(arr, 1, C.byref(out)) == tgh.TGH_EUNSUPPORTED


# -- the rotating seam ----------------------------------------------------------

def rotating_hole(nb=2, octant=True):
    """The a = 3/10 hole (q = 2, the sampled source, h = 1/16) on the rotating
    octant [0, 2]^3 (x = 0 glued to y = 0 by a quarter turn about z, z = 0
    mirrored), or in the box [-2, 2]^3."""
    from test_oracle_reflect import _rotating_hole
    return _rotating_hole(nb, octant)


def test_seam_fill_matches_oracle(tgh):
    """The fill of the rotating octant (ghosts turned across the seam, twice
    at the axis's corner, mirrored across z = 0, the hook's) against the
    oracle's: the turned copies are signed permutations, so equal to the
    Dirichlet data's roundoff."""
    torch = torch_cuda()
    desc = rotating_hole()
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    u = _noisy(o, 0x5EA, 1e-6)
    ud = torch.from_numpy(u.copy()).cuda()
    tgh.prepare(p, ud, 0.0)
    o.prepare(u, 0.0)
    wg, wo = p.work().cpu().numpy(), o.work()
    err = float((np.abs(wg - wo) / np.maximum(np.abs(wo), 1.0)).max())
    print(f"rotating octant: fill |GPU - oracle| / max(1, |oracle|) = {err:.3e}")
    assert err <= 1e-13


@pytest.mark.parametrize("form", ["brick-split", "brick-fused", "blocks-fused", "blocks-split"])
def test_rk4_on_the_rotating_octant_matches_oracle(tgh, form):
    torch = torch_cuda()
    desc = rotating_hole()
    p = tgh_problem_from_desc(tgh, desc)
    o = oracle_lib.OracleProblem(desc)
    u = _noisy(o, 0x5EB, 1e-6)
    ud = torch.from_numpy(u.copy()).cuda()
    dt = 0.25 * (1 / 16) / o.max_speed(u)
    for n in range(2):
        o.rk4_step(u, n * dt, dt, NTHREADS)
    tgh._check(tgh._load().tgh_rk4(p.ptr, tgh.C.c_void_p(ud.data_ptr()), 0.0, dt, 2, RK4_FORMS[form]))
    g = ud.cpu().numpy()
    diff = float((np.abs(g - u) / np.maximum(np.abs(u), 1.0)).max())
    print(f"rotating octant, 2 RK4 steps ({form}): max rel diff {diff:.3e}")
    assert diff <= 1e-12


def test_the_rotating_octant_is_the_hole_in_the_box_on_the_device(tgh):
    """test/rotation_tests.jl:148-190 on the device alone: the spinning hole
    on the rotating octant and in the whole box, 8 steps from the exact
    state: every owned point outside the core agrees to 64 eps of the
    state's scale; and the seam's noise leaves both seam planes alone."""
    torch = torch_cuda()
    from test_oracle_reflect import _points
    do, db = rotating_hole(2, True), rotating_hole(4, False)
    po, pb = tgh_problem_from_desc(tgh, do), tgh_problem_from_desc(tgh, db)
    oo, ob = oracle_lib.OracleProblem(do), oracle_lib.OracleProblem(db)
    go = torch.from_numpy(oo.fill_exact(0.0)).cuda()
    gb = torch.from_numpy(ob.fill_exact(0.0)).cuda()
    dt = 0.25 * (1 / 16) / tgh.max_speed(pb, gb, 0.0)
    for p, g in ((po, go), (pb, gb)):
        tgh._check(tgh._load().tgh_rk4(p.ptr, tgh.C.c_void_p(g.data_ptr()), 0.0, dt, 8, 0))
    P, h = do.N + 1, 1 / 16
    xo, xb = _points(do, oo.block_coords()), _points(db, ob.block_coords())
    Go, Gb = go.cpu().numpy().reshape(-1, 20, P, P, P), gb.cpu().numpy().reshape(-1, 20, P, P, P)
    at = {}
    for b in range(xb.shape[0]):
        for idx in np.ndindex(P, P, P):
            at.setdefault(tuple(int(round(c / h)) for c in xb[b][idx]), (b, idx))
    diff, scale = 0.0, 0.0
    for b in range(xo.shape[0]):
        for idx in np.ndindex(P, P, P):
            x = xo[b][idx]
            if x @ x < 0.35 ** 2:
                continue
            bb, ib = at[tuple(int(round(c / h)) for c in x)]
            diff = max(diff, float(np.abs(Go[b][(slice(None),) + idx] - Gb[bb][(slice(None),) + ib]).max()))
            scale = max(scale, float(np.abs(Go[b][(slice(None),) + idx]).max()))
    print(f"device rotating octant vs device box: max|diff| {diff:.3e}, scale {scale:.3e}")
    assert diff <= 64 * np.finfo(float).eps * scale
    # the noise: both seam planes keep their data, the device as the oracle
    u = oo.fill_exact(0.0)
    ud = torch.from_numpy(u.copy()).cuda()
    oo.add_noise(u, 2, 1e-8)
    tgh.add_noise(ud, po, 2, 1e-8)
    g = ud.cpu().numpy()
    # the same draws (the device's sum may be contracted: one rounding apart)
    assert float((np.abs(g - u) / np.maximum(np.abs(u), 1.0)).max()) <= 1e-15
    # (off the z = 0 wall, whose odd variables the projection zeroes)
    seam = ((xo[..., 0] == 0.0) | (xo[..., 1] == 0.0)) & (xo[..., 2] > 0.0)
    U0 = oo.fill_exact(0.0).reshape(-1, 20, P, P, P)
    for arr in (u, g):
        moved = np.abs(arr.reshape(-1, 20, P, P, P) - U0).transpose(1, 0, 2, 3, 4)
        assert float(moved[:, seam].max()) == 0.0 and float(moved[:, ~seam].max()) > 9e-9



# -- on a forest with levels -----------------------------------------------------

def _forest_case(tgh, kind):
    """The a = 0 hole on the mirrored octant or the a = 3/10 hole on the
    rotating octant, [0, 2]^3, on the two-level forest of the oracle's test
    (the root block at the corner refined), q = 2, the sampled source, p = 4."""
    from test_oracle_reflect import _octant_leaves
    a = 0.0 if kind == "reflecting" else 0.3
    d = TD.kerr_schild_hole_case(N=8, nblocks=2, halfwidth=2.0, a=a, r0=0.35, r1=1.1,
                                 gauge=TD.TGH_GAUGE_SAMPLED, q=2)
    d.lo[:] = [0.0] * 3
    if kind == "reflecting":
        for dd in range(3):
            d.reflecting[dd][0] = 1
    else:
        d.reflecting[2][0] = 1
        d.rotating = 1
    leaves = _octant_leaves()
    L = tgh._load()
    arr = (C.c_int32 * (4 * len(leaves)))(*[x for lf in leaves for x in lf])
    ptr = C.c_void_p()
    tgh._check(L.tgh_problem_create_forest(C.byref(tgh.TghDesc.from_buffer_copy(d)), arr, len(leaves), 4,
                                           C.byref(ptr)))
    import torch
    tgh._check(L.tgh_set_stream(ptr, C.c_void_p(torch.cuda.current_stream().cuda_stream)))

    class Handle:
        pass

    h = Handle()
    h.ptr = ptr
    h.state_length = int(L.tgh_state_length(ptr))
    h.work_length = int(L.tgh_work_length(ptr))
    o = oracle_lib.OracleProblem(d, leaves, 4)
    assert o.n == h.state_length
    return d, h, o


@pytest.mark.parametrize("kind", ["reflecting", "rotating"])
def test_forest_fill_and_rk4_across_the_walls_match_oracle(tgh, kind):
    torch = torch_cuda()
    d, h, o = _forest_case(tgh, kind)
    L = tgh._load()
    u = _noisy(o, 0xF0E5, 1e-6)
    ud = torch.from_numpy(u.copy()).cuda()
    tgh._check(L.tgh_prepare(h.ptr, C.c_void_p(ud.data_ptr()), 0.0))
    w = torch.empty(h.work_length, dtype=torch.float64, device="cuda")
    tgh._check(L.tgh_copy_work(h.ptr, C.c_void_p(w.data_ptr())))
    o.prepare(u, 0.0)
    wg, wo = w.cpu().numpy(), o.work()
    ferr = float((np.abs(wg - wo) / np.maximum(np.abs(wo), 1.0)).max())
    dt = 0.25 * (2.0 / 32) / o.max_speed(u)
    for n in range(2):
        o.rk4_step(u, n * dt, dt, NTHREADS)
    tgh._check(L.tgh_rk4(h.ptr, C.c_void_p(ud.data_ptr()), 0.0, dt, 2, 0))
    g = ud.cpu().numpy()
    L.tgh_problem_destroy(h.ptr)
    diff = float((np.abs(g - u) / np.maximum(np.abs(u), 1.0)).max())
    print(f"{kind} forest: fill {ferr:.3e}, 2 RK4 steps {diff:.3e}")
    assert ferr <= 1e-13 and diff <= 1e-12
