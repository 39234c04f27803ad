"""Reflecting faces in the oracle (the mirrors of src/boundaries.jl:84-142 and
test/reflection_tests.jl): flat space with noise, because noise has every
component nonzero, so a wrong parity in any of the twenty variables shows.

Pinned by: the reference's own parity table (test/reflection_tests.jl:
53-69); the octant [0, 4]^3 against the box [-4, 4]^3 with Dirichlet data
everywhere and the octant's noise extended by parity, which is the one
problem whose answer is known without a mirror (test/reflection_tests.jl:
124-187: the octant is the box's symmetric part to roundoff); the odd
variables staying zero on their walls; and add_noise's projection of the
low walls (src/initialdata.jl:1938-1990). A plausible slip -- a parity
taken from the wrong index, a mirror about the wrong plane (-g - 1 instead
of -g), a missing sign -- moves the octant off the box by the data's own
size."""
import numpy as np
import pytest

import oracle_lib
import tgh_desc as TD

NTHREADS = 8
E, O = 0, 1


def test_parities_are_the_tensors():
    """Component ab of h and Pi is odd across x^d = 0 when x^d occurs an odd
    number of times in ab: the reference's table, entry for entry."""
    L = oracle_lib.lib()
    table = {0: (E, E, E), 1: (O, E, E), 2: (E, O, E), 3: (E, E, O), 4: (E, E, E),
             5: (O, O, E), 6: (O, E, O), 7: (E, E, E), 8: (E, O, O), 9: (E, E, E)}
    for v in range(20):
        m = L.orc_parity_odd_mask(v)
        assert tuple((m >> d) & 1 for d in range(3)) == table[v % 10], v


def _minkowski(lo, hi, nb, reflecting):
    return TD.make_desc(q=4, N=8, nblocks=(nb,) * 3, periodic=(0, 0, 0), lo=(lo,) * 3, hi=(hi,) * 3,
                        background=TD.TGH_BG_MINKOWSKI, gamma0=1.0, gamma2=0.0, eps_ko=0.5,
                        reflecting=reflecting)


def _points(desc, coords):
    """(b, k, j, i, 3) positions of the owned points."""
    P, N = desc.N + 1, desc.N
    h = (desc.hi[0] - desc.lo[0]) / (desc.nblocks[0] * N)
    i = np.arange(P)
    x = np.zeros((len(coords), P, P, P, 3))
    for b, (bx, by, bz) in enumerate(coords):
        x[b, ..., 0] = (desc.lo[0] + (bx * N + i) * h)[None, None, :]
        x[b, ..., 1] = (desc.lo[1] + (by * N + i) * h)[None, :, None]
        x[b, ..., 2] = (desc.lo[2] + (bz * N + i) * h)[:, None, None]
    return x


def _odd(L):
    return np.array([[(L.orc_parity_odd_mask(v) >> d) & 1 for d in range(3)] for v in range(20)])


def _octant_and_box():
    oct_d = _minkowski(0.0, 4.0, 2, ((1, 0), (1, 0), (1, 0)))
    box_d = _minkowski(-4.0, 4.0, 4, ((0, 0), (0, 0), (0, 0)))
    return oct_d, oracle_lib.OracleProblem(oct_d), box_d, oracle_lib.OracleProblem(box_d)


def test_noise_is_projected_onto_the_parity_on_the_low_walls():
    L = oracle_lib.lib()
    oct_d, o, _, _ = _octant_and_box()
    u = o.fill_exact(0.0)
    o.add_noise(u, 7, 1e-8)
    P = oct_d.N + 1
    U = u.reshape(-1, 20, P, P, P)
    x = _points(oct_d, o.block_coords())
    odd = _odd(L)
    wall_odd, wall_even = 0.0, 0.0
    for d in range(3):
        on = x[..., d] == 0.0  # (b, k, j, i)
        for v in range(20):
            vals = np.abs(U[:, v][on])
            if odd[v, d]:
                wall_odd = max(wall_odd, float(vals.max()))
            else:
                wall_even = max(wall_even, float(vals.max()))
    assert wall_odd == 0.0 and wall_even > 9e-9
    assert float(np.abs(u).max()) <= 1e-8


def test_the_octant_is_the_mirrored_box_to_roundoff():
    """32 steps of dt = h / (4 sqrt 3) from noise put inside [0, 1]^3 (so the
    box's evolved plane x = -4, which the octant has as its Dirichlet plane,
    stays out of the data's reach): every owned point of the octant equals
    the box's at the same position to 64 eps of the data's scale; the odd
    variables on the walls stay at roundoff; and a control with the parities
    ignored (every variable extended evenly into the box) is off by the
    data's own size, so the comparison is not vacuous."""
    L = oracle_lib.lib()
    oct_d, o, box_d, bx = _octant_and_box()
    P = oct_d.N + 1
    uo = o.fill_exact(0.0)
    o.add_noise(uo, 1, 1e-8)
    xo = _points(oct_d, o.block_coords())
    far = xo.max(axis=-1) > 1.0
    Uo = uo.reshape(-1, 20, P, P, P)
    for v in range(20):
        Uo[:, v][far] = 0.0
    odd = _odd(L)
    xb = _points(box_d, bx.block_coords())
    # the octant's owned point at a position: index by rounded coordinates
    h = 0.25  # 4 / (2 blocks x 8 cells), the box's too
    lookup = {}
    for b in range(xo.shape[0]):
        for idx in np.ndindex(P, P, P):
            key = tuple(int(round(c / h)) for c in xo[b][idx])
            lookup.setdefault(key, (b, idx))

    def box_state(sign_by_parity):
        ub = bx.fill_exact(0.0)
        Ub = ub.reshape(-1, 20, P, P, P)
        for b in range(xb.shape[0]):
            for idx in np.ndindex(P, P, P):
                x = xb[b][idx]
                if np.abs(x).max() >= 4.0:
                    continue
                bo, io = lookup[tuple(int(round(abs(c) / h)) for c in x)]
                for v in range(20):
                    s = 1.0
                    if sign_by_parity:
                        for d in range(3):
                            if x[d] < 0 and odd[v, d]:
                                s = -s
                    Ub[b, v][idx] = s * Uo[bo, v][io]
        return ub

    ub = box_state(True)
    uc = box_state(False)
    dt = h / (4 * np.sqrt(3.0))
    for n in range(32):
        o.rk4_step(uo, n * dt, dt, NTHREADS)
        bx.rk4_step(ub, n * dt, dt, NTHREADS)
        bx.rk4_step(uc, n * dt, dt, NTHREADS)
    Ub = ub.reshape(-1, 20, P, P, P)
    Uc = uc.reshape(-1, 20, P, P, P)
    box_at = {}
    for b in range(xb.shape[0]):
        for idx in np.ndindex(P, P, P):
            box_at.setdefault(tuple(int(round(c / h)) for c in xb[b][idx]), (b, idx))
    diff, ctrl = 0.0, 0.0
    for b in range(xo.shape[0]):
        for idx in np.ndindex(P, P, P):
            bb, ib = box_at[tuple(int(round(c / h)) for c in xo[b][idx])]
            diff = max(diff, float(np.abs(Uo[b][(slice(None),) + idx] - Ub[bb][(slice(None),) + ib]).max()))
            ctrl = max(ctrl, float(np.abs(Uo[b][(slice(None),) + idx] - Uc[bb][(slice(None),) + ib]).max()))
    scale = float(np.abs(uo).max())
    wall = 0.0
    for d in range(3):
        on = xo[..., d] == 0.0
        for v in range(20):
            if odd[v, d]:
                wall = max(wall, float(np.abs(Uo[:, v][on]).max()))
    print(f"octant vs mirrored box: max|diff| {diff:.3e}, scale {scale:.3e}; control {ctrl:.3e}; "
          f"odd on the walls {wall:.3e}")
    assert scale > 1e-9
    assert diff <= 64 * np.finfo(float).eps * scale
    assert ctrl > 1e-3 * scale
    assert wall <= 64 * np.finfo(float).eps * scale


# -- the rotating seam (src/boundaries.jl:145-205, test/rotation_tests.jl) -----

ROT = [1, -3, 2, 4, 8, -6, -9, 5, 7, 10]  # the reference's state_rotation, tt .. zz


def test_the_quarter_turn_is_the_tensors():
    """The map every seam transfer turns the state with: the reference's table
    (test/rotation_tests.jl:54-58), for h and for Pi, and against the
    solution itself: u(R p) = Q u(p) for the spinning hole a = 3/10 at random
    points outside its core, R (x, y, z) = (-y, x, z)."""
    L = oracle_lib.lib()
    r = [L.orc_rotation(v) for v in range(20)]
    assert r[:10] == ROT
    assert r[10:] == [int(np.sign(x)) * (abs(x) + 10) for x in ROT]
    d = TD.kerr_schild_hole_case(N=12, nblocks=2, halfwidth=2.5, a=0.3, r0=0.35, r1=1.1,
                                 gauge=TD.TGH_GAUGE_SAMPLED, q=2)
    rng = np.random.default_rng(4)
    worst = 0.0
    import ctypes as C
    for _ in range(20):
        p = np.array([2 * rng.random() + 0.5, 2 * rng.random() + 0.5, 2 * rng.random() - 1.0])
        u, uR = np.zeros(20), np.zeros(20)
        L.orc_case_state(C.byref(d), 0.0, p, u)
        L.orc_case_state(C.byref(d), 0.0, np.array([-p[1], p[0], p[2]]), uR)
        want = np.array([np.sign(r[v]) * u[abs(r[v]) - 1] for v in range(20)])
        worst = max(worst, float(np.abs(uR - want).max() / max(1.0, float(np.abs(u).max()))))
    print(f"u(R p) against Q u(p): {worst:.3e}")
    assert worst <= 64 * np.finfo(float).eps


def _rotating_hole(nb, octant):
    """The a = 3/10 hole, q = 2, the sampled source, r_0 = 0.35, r_1 = 1.1, at
    h = 1/16 (exact in binary, so that the octant's and the box's points sit
    at the same positions to the bit): the octant [0, 2]^3 on 2^3 blocks of
    16^3, or the box [-2, 2]^3 on 4^3."""
    d = TD.kerr_schild_hole_case(N=16, nblocks=nb, halfwidth=2.0, a=0.3, r0=0.35, r1=1.1,
                                 gauge=TD.TGH_GAUGE_SAMPLED, q=2)
    if octant:
        d.lo[:] = [0.0] * 3
        d.reflecting[2][0] = 1
        d.rotating = 1
    return d


def test_the_seams_ghosts_are_the_solution_a_quarter_turn_away():
    """After one fill of the exact state every stored point of every block --
    owned, mirrored across z = 0, turned across the seam, turned twice at the
    axis's corner, or the hook's -- holds the case's exact state at its own
    position (test/rotation_tests.jl:115-146), outside the core, to the
    roundoff of the solution's own symmetry."""
    import ctypes as C
    L = oracle_lib.lib()
    d = _rotating_hole(2, True)
    o = oracle_lib.OracleProblem(d)
    o.prepare(o.fill_exact(0.0), 0.0)
    W = o.work()
    G, N = d.q // 2 + 1, d.N
    S = N + 1 + 2 * G
    h = 1 / 16
    Wv = W.reshape(-1, 20, S, S, S)
    worst, scale = 0.0, 0.0
    e = np.zeros(20)
    for b, (bx, by, bz) in enumerate(o.block_coords()):
        for k in range(S):
            for j in range(S):
                for i in range(S):
                    x = np.array([(bx * N + i - G) * h, (by * N + j - G) * h, (bz * N + k - G) * h])
                    if x @ x < 0.35 ** 2:
                        continue
                    L.orc_case_state(C.byref(d), 0.0, x, e)
                    worst = max(worst, float(np.abs(Wv[b, :, k, j, i] - e).max()))
                    scale = max(scale, float(np.abs(e).max()))
    print(f"seam ghosts against the exact state: {worst / scale:.3e} relative")
    assert worst <= 256 * np.finfo(float).eps * scale


def test_a_spinning_hole_on_the_rotating_octant_is_the_hole_in_the_box():
    """The a = 3/10 hole (q = 2, h = 1/16, the sampled source, :damped layer)
    on the rotating octant and in the whole box [-2, 2]^3, 8 RK4 steps
    from the exact state: every owned point of the octant outside the core
    equals the box's at its position to 64 eps of the state's scale
    (test/rotation_tests.jl:148-190)."""
    do, db = _rotating_hole(2, True), _rotating_hole(4, False)
    o, bx = oracle_lib.OracleProblem(do), oracle_lib.OracleProblem(db)
    uo, ub = o.fill_exact(0.0), bx.fill_exact(0.0)
    dt = 0.25 * (1 / 16) / bx.max_speed(ub)
    for n in range(8):
        o.rk4_step(uo, n * dt, dt, NTHREADS)
        bx.rk4_step(ub, n * dt, dt, NTHREADS)
    P, h = do.N + 1, 1 / 16
    xo, xb = _points(do, o.block_coords()), _points(db, bx.block_coords())
    Uo, Ub = uo.reshape(-1, 20, P, P, P), ub.reshape(-1, 20, P, P, P)
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
            a, c = Uo[b][(slice(None),) + idx], Ub[bb][(slice(None),) + ib]
            diff = max(diff, float(np.abs(a - c).max()))
            scale = max(scale, float(np.abs(a).max()))
    moved = float(np.abs(uo - o.fill_exact(0.0)).max())
    print(f"rotating octant vs box after 8 steps: max|diff| {diff:.3e}, scale {scale:.3e}, moved {moved:.3e}")
    assert moved > 1e-6
    assert diff <= 64 * np.finfo(float).eps * scale


# -- mirrors and the seam on forests with levels ---------------------------------

def _forest_problem(desc, leaves, p=4):
    return oracle_lib.OracleProblem(desc, leaves, p)


def _octant_leaves():
    """Two levels about the corner of a 2^3 root brick: the root block at the
    origin refined (8 level-1 leaves) and the other seven roots as they are,
    2:1 balanced (a face, edge and corner of the refined block touch each of
    the three reflecting walls and the seam)."""
    leaves = [(0, i, j, k) for k in range(2) for j in range(2) for i in range(2) if (i, j, k) != (0, 0, 0)]
    leaves += [(1, i, j, k) for k in range(2) for j in range(2) for i in range(2)]
    return leaves


@pytest.mark.parametrize("kind", ["reflecting", "rotating"])
def test_forest_ghosts_across_the_walls_are_the_exact_state(kind):
    """The a = 0 (mirrored octant) or a = 3/10 (rotating octant) hole on a
    two-level forest, one fill of the exact state: every stored point that
    is a copy -- plain, mirrored, turned, or both -- holds the exact state at
    its own position to roundoff, every Dirichlet point too, and every
    prolongation to the order-4 interpolation's error; the classification
    reports mirrored and turned copies, so the check covers them."""
    import ctypes as C
    L = oracle_lib.lib()
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
    o = _forest_problem(d, leaves)
    o.prepare(o.fill_exact(0.0), 0.0)
    G, N = d.q // 2 + 1, d.N
    S = N + 1 + 2 * G
    Wv = o.work().reshape(-1, 20, S, S, S)
    h0 = 2.0 / (2 * N)
    coords = o.block_coords()
    worst_copy, worst_prol, scale = 0.0, 0.0, 0.0
    seen = {"mirrored": 0, "turned": 0}
    e = np.zeros(20)
    for b in range(len(coords)):
        lv = o.block_level(b)
        h = h0 / 2 ** lv
        for k in range(-G, N + G + 1):
            for j in range(-G, N + G + 1):
                for i in range(-G, N + G + 1):
                    x = np.array([(coords[b][0] * N + i) * h, (coords[b][1] * N + j) * h, (coords[b][2] * N + k) * h])
                    if x @ x < 0.5 ** 2:
                        continue
                    c, src = o.classify(b, i, j, k)
                    L.orc_case_state(C.byref(d), 0.0, x, e)
                    err = float(np.abs(Wv[b, :, k + G, j + G, i + G] - e).max())
                    scale = max(scale, float(np.abs(e).max()))
                    if c == 3:
                        worst_prol = max(worst_prol, err)
                    else:
                        worst_copy = max(worst_copy, err)
                    if src[5] & 7:
                        seen["mirrored"] += 1
                    if src[5] >> 3:
                        seen["turned"] += 1
    print(f"{kind}: copies {worst_copy / scale:.3e}, prolongations {worst_prol / scale:.3e} of {scale:.3e}; {seen}")
    assert seen["mirrored"] > 0 and (kind == "reflecting" or seen["turned"] > 0)
    assert worst_copy <= 256 * np.finfo(float).eps * scale
    assert worst_prol <= 1e-2 * scale


def test_on_a_refined_octant_the_odd_components_stay_zero_on_the_walls():
    """test/reflection_tests.jl:190-231: flat space with noise on a two-level
    octant, 20 steps: an odd variable that starts zero on its wall stays zero
    there to roundoff (the mirrored ghosts give u(-h) = -u(h) and centred
    stencils respect it), while the data themselves move."""
    L = oracle_lib.lib()
    d = TD.make_desc(q=4, N=8, nblocks=(2,) * 3, periodic=(0, 0, 0), lo=(0.0,) * 3, hi=(4.0,) * 3,
                     background=TD.TGH_BG_MINKOWSKI, gamma0=1.0, gamma2=0.0, eps_ko=0.5,
                     reflecting=((1, 0), (1, 0), (1, 0)))
    o = _forest_problem(d, _octant_leaves(), 6)
    u = o.fill_exact(0.0)
    o.add_noise(u, 3, 1e-8)
    u0 = u.copy()
    dt = (0.25 / 2) / (4 * np.sqrt(3.0))
    for n in range(20):
        o.rk4_step(u, n * dt, dt, NTHREADS)
    P = d.N + 1
    U = u.reshape(-1, 20, P, P, P)
    coords = o.block_coords()
    odd = _odd(L)
    wall = 0.0
    for b, bc in enumerate(coords):
        for dd in range(3):
            if bc[dd] != 0:
                continue
            sl = [slice(None)] * 3
            sl[2 - dd] = 0
            for v in range(20):
                if odd[v, dd]:
                    wall = max(wall, float(np.abs(U[b, v][tuple(sl)]).max()))
    scale = float(np.abs(u).max())
    print(f"refined octant after 20 steps: odd on the walls {wall:.3e}, scale {scale:.3e}, "
          f"moved {float(np.abs(u - u0).max()):.3e}")
    assert float(np.abs(u - u0).max()) > 1e-9
    assert wall <= 64 * np.finfo(float).eps * scale
