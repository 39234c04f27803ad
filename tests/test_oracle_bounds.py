"""The range projection (src/bounds.jl:260-589) in the oracle, pinned by what
the map must do rather than by its formulae: the eigen-solver against numpy
(double eigenvalues included), the identity -- bit for bit -- on healthy states,
idempotence of state and flag on wild states, every range met afterwards, the
non-finite repair, the lapse raise keeping (alpha/sqrt(gamma)) Pi, only the
fired ADM block moving, and the gate on the mesh."""
import numpy as np
import pytest

import oracle_lib
import tgh_desc as TD

EPS = np.finfo(float).eps
DEFAULT = (1 / 50, 50.0, 1 / 100, 1000.0, 10.0, 100.0, 1.0)  # default_bounds(M = 1), bounds.jl:138-160


def unpack(h):
    """g_ab (4x4) from the packed offset h (tt,tx,ty,tz,xx,xy,xz,yy,yz,zz)."""
    eta = np.diag([-1.0, 1, 1, 1])
    idx = [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]
    g = eta.copy()
    for n, (a, b) in enumerate(idx):
        g[a, b] += h[n]
        if a != b:
            g[b, a] += h[n]
    return g


def adm(h):
    g = unpack(h)
    gam = g[1:, 1:]
    bl = g[0, 1:]
    bu = np.linalg.solve(gam, bl)
    a2 = bl @ bu - g[0, 0]
    return gam, bl, bu, a2


def test_jacobi_eigen_matches_numpy_including_double_eigenvalues():
    rng = np.random.default_rng(1)
    mats = []
    for _ in range(300):
        B = rng.normal(size=(3, 3)) * rng.uniform(0.1, 10)
        mats.append(B + B.T)
    # Kerr-Schild's gamma = delta + (2M/r) l l: a double eigenvalue 1
    for r in (0.8, 1.5, 4.0):
        l = rng.normal(size=3)
        l /= np.linalg.norm(l)
        mats.append(np.eye(3) + (2 / r) * np.outer(l, l))
    mats.append(np.diag([0.01, 0.01, 3.0]))
    for A in mats:
        lam, V = oracle_lib.sym_eigen3(A)
        ref = np.linalg.eigvalsh(A)
        nrm = np.abs(A).max()
        assert np.allclose(np.sort(lam), ref, atol=8 * EPS * nrm * 4, rtol=0)
        assert np.allclose(V.T @ V, np.eye(3), atol=1e-14)
        assert np.allclose(A @ V, V * lam, atol=16 * EPS * nrm)


def ks_states(r_min=1.0):
    """The spinning hole's exact states (a = 0.7) at the mesh points with r >= r_min:
    outside the frozen core, whose states come from the sphere r_0 = 0.71 beside
    the ring singularity (momentum scale ~134/M there, above K_max)."""
    d = TD.kerr_schild_hole_case(N=8, nblocks=2, halfwidth=2.0, a=0.7)
    o = oracle_lib.OracleProblem(d)
    P = d.N + 1
    u = o.fill_exact(0.0).reshape(-1, 20, P ** 3)
    B = o.block_coords()
    h = (d.hi[0] - d.lo[0]) / (d.nblocks[0] * d.N)
    i = np.arange(P)
    rows = []
    for b in range(B.shape[0]):
        x = d.lo[0] + (B[b, 0] * d.N + i) * h
        y = d.lo[1] + (B[b, 1] * d.N + i) * h
        z = d.lo[2] + (B[b, 2] * d.N + i) * h
        r = np.sqrt(x[None, None, :] ** 2 + y[None, :, None] ** 2 + z[:, None, None] ** 2).ravel()
        rows.append(u[b].T[r >= r_min])
    return np.concatenate(rows)


def test_identity_bit_for_bit_on_healthy_states():
    states = ks_states()
    assert len(states) > 3000
    for s in states:
        h, P = s[:10], s[10:]
        ho, Po, hit, nf = oracle_lib.bounds_project(h, P, DEFAULT)
        assert not hit and not nf
        assert np.array_equal(ho, h) and np.array_equal(Po, P)


def wild(rng, n):
    for _ in range(n):
        yield rng.uniform(-5, 5, 10) * rng.uniform(0, 1), rng.uniform(-100, 100, 10) * rng.uniform(0, 1)


def test_idempotent_on_state_and_flag_and_ranges_met():
    rng = np.random.default_rng(2)
    bnd = DEFAULT
    amin, amax, lmin, lmax, bmax, Kmax, _ = bnd
    fired = 0
    for h, P in wild(rng, 4000):
        ho, Po, hit, nf = oracle_lib.bounds_project(h, P, bnd)
        fired += hit
        h2, P2, hit2, _ = oracle_lib.bounds_project(ho, Po, bnd)
        assert not hit2, (h, P)
        assert np.array_equal(h2, ho) and np.array_equal(P2, Po)
        gam, bl, bu, a2 = adm(ho)
        lam = np.linalg.eigvalsh(gam)
        tol = 1e-12 * max(1.0, np.abs(lam).max())
        assert lam.min() >= lmin - tol and lam.max() <= lmax + tol
        assert bl @ bu <= bmax ** 2 * (1 + 1e-11)
        sc = 1e-11 * (1 + abs(ho[0]) + abs(bl @ bu))
        assert amin ** 2 - sc <= a2 <= amax ** 2 + sc
        k = np.sqrt(a2) / np.sqrt(np.linalg.det(gam)) * np.abs(Po).max()
        assert k <= Kmax * (1 + 1e-11)
    assert fired > 1000  # the wild states do exercise the map


def test_nonfinite_components_take_minkowski_values():
    s = ks_states()[100]
    h, P = s[:10].copy(), s[10:].copy()
    h[3], P[7] = np.nan, np.inf
    ho, Po, hit, nf = oracle_lib.bounds_project(h, P, DEFAULT)
    assert hit and nf
    assert ho[3] == 0.0 and Po[7] == 0.0
    keep = [n for n in range(10) if n != 3]
    assert np.array_equal(ho[keep], h[keep])  # the repaired state is healthy: nothing else moved


def test_raised_lapse_keeps_the_momentum_term():
    rng = np.random.default_rng(3)
    h = np.zeros(10)
    h[0] = 1 - 0.01 ** 2  # alpha = 0.01 < alpha_min = 0.02, flat gamma, no shift
    P = rng.uniform(-0.5, 0.5, 10)
    ho, Po, hit, nf = oracle_lib.bounds_project(h, P, DEFAULT)
    assert hit and not nf
    gam, bl, bu, a2 = adm(ho)
    assert abs(np.sqrt(a2) - 1 / 50) < 1e-14  # h_tt = 1 - alpha^2 cancels: ~eps/alpha
    assert np.allclose(np.sqrt(a2) * Po, 0.01 * P, rtol=1e-12, atol=0)
    assert np.array_equal(ho[1:], h[1:])  # only h_tt moved


def test_only_the_fired_block_moves():
    s = ks_states()[300]
    h, P = s[:10].copy(), s[10:].copy()
    h[1:4] *= 40.0  # a shift far above beta_max, gamma untouched
    ho, Po, hit, _ = oracle_lib.bounds_project(h, P, DEFAULT)
    assert hit
    assert np.array_equal(ho[4:], h[4:])  # the spatial block keeps its bits
    gam, bl, bu, a2 = adm(ho)
    assert abs(np.sqrt(bl @ bu) - 10.0) < 1e-12


def test_mesh_projection_respects_the_gate():
    d = TD.kerr_schild_hole_case(N=8, nblocks=2, halfwidth=2.0, a=0.7)
    d.alpha_min, d.alpha_max, d.lambda_min, d.lambda_max, d.beta_max, d.K_max = DEFAULT[:6]
    d.r_gate = 1.2
    o = oracle_lib.OracleProblem(d)
    u = o.fill_exact(0.0)
    rng = np.random.default_rng(4)
    v = u.copy()
    v += rng.uniform(-3, 3, v.size) * (rng.uniform(size=v.size) < 0.02)
    w = v.copy()
    n = o.apply_bounds(w)
    P = d.N + 1
    B = o.block_coords()
    h = (d.hi[0] - d.lo[0]) / (d.nblocks[0] * d.N)
    i = np.arange(P)
    moved = 0
    for b in range(B.shape[0]):
        x = d.lo[0] + (B[b, 0] * d.N + i) * h
        y = d.lo[1] + (B[b, 1] * d.N + i) * h
        z = d.lo[2] + (B[b, 2] * d.N + i) * h
        r = np.sqrt(x[None, None, :] ** 2 + y[None, :, None] ** 2 + z[:, None, None] ** 2)
        vb = v.reshape(-1, 20, P, P, P)[b]
        wb = w.reshape(-1, 20, P, P, P)[b]
        ch = (vb != wb).any(axis=0)
        assert not ch[r >= 1.2].any()
        moved += int(ch.sum())
    assert n == moved and n > 0
