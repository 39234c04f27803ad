"""The oracle's tangent pass: the reference's CPU Float64 right-hand side run
with a forward-mode dual as its float type (the check the reference's own
ForwardDiff.Dual runs make), pinned against central differences of the
plain double pass, in the state and in each of the five parameters the
adjoint differentiates: the damping's near and far rates, gamma_2, and the
layer's rho_max and rho ramp."""
import numpy as np
import pytest

import oracle_lib
import tgh_desc as TD

PARAMS = ("gamma0_near", "gamma0_far", "gamma2", "rho_max", "rho_ramp")


def _hole(tgh, **kw):
    # the adjoint's workload in small: the a = 0.7 Kerr-Schild hole with the
    # closed-form source, :damped r_0 = 0.71, r_1 = 1.21, Gaussian damping
    # (1, 0.1, 3), on 4^3 blocks of 8^3 over [-1.75, 1.75]^3 (h = 7/64) at q = 2
    case = tgh.kerr_schild_case(M=1.0, a=0.7, halfwidth=1.75, r_0=0.71, r_1=1.21, gauge_source="algebraic",
                                gamma2=kw.pop("gamma2", 0.1), rho_max=kw.pop("rho_max", 4.0),
                                rho_ramp=kw.pop("rho_ramp", 0.5))
    if "near" in kw or "far" in kw:
        case.gamma0 = tgh.GaussianDamping(near=kw.pop("near", 1.0), far=kw.pop("far", 0.1), width=3.0)
    f = tgh.gh_forest(None, case, N=8, roots=4)
    desc = TD.TghDesc.from_buffer_copy(tgh._desc(f, case, 2))
    return oracle_lib.OracleProblem(desc), desc


def _state(orc, seed):
    u = orc.fill_exact(0.05)
    orc.add_noise(u, seed, 1e-4)
    return u


def test_state_tangent_matches_central_differences(tgh):
    orc, _ = _hole(tgh)
    u = _state(orc, 0xA1)
    v = np.random.default_rng(3).standard_normal(u.shape)
    F, dF = oracle_lib.gh_rhs_jvp(orc, u, v, 0.05, np.zeros(5))
    assert np.array_equal(F, orc.gh_rhs(u, 0.05, 8))  # the dual's value part is the plain pass
    eps = 1e-6
    fd = (orc.gh_rhs(u + eps * v, 0.05, 8) - orc.gh_rhs(u - eps * v, 0.05, 8)) / (2 * eps)
    err = np.abs(fd - dF).max() / np.abs(dF).max()
    print(f"state tangent vs central difference: {err:.2e}")
    assert err < 1e-6
    # linear in the direction, to roundoff
    _, d2 = oracle_lib.gh_rhs_jvp(orc, u, 2 * v, 0.05, np.zeros(5))
    assert np.abs(d2 - 2 * dF).max() <= 1e-12 * np.abs(dF).max()


@pytest.mark.parametrize("k", range(5), ids=PARAMS)
def test_parameter_tangent_matches_central_differences(tgh, k):
    base = dict(near=1.0, far=0.1, gamma2=0.1, rho_max=4.0, rho_ramp=0.5)
    orc, _ = _hole(tgh, **dict(base))
    u = _state(orc, 0xA2)
    pd = np.zeros(5)
    pd[k] = 1.0
    _, dF = oracle_lib.gh_rhs_jvp(orc, u, np.zeros_like(u), 0.05, pd)
    name = ("near", "far", "gamma2", "rho_max", "rho_ramp")[k]
    # the right-hand side is linear in the first four (gamma0 = far + (near -
    # far) e^(-r^2/2w^2) enters the damping term linearly, so does 1 + gamma2,
    # rho = rho_max smoothstep(.)): a wide step loses nothing to truncation;
    # rho_ramp enters through the smoothstep, a narrow step
    eps = (1e-2, 1e-2, 1e-2, 1.0, 1e-5)[k]
    plus, minus = dict(base), dict(base)
    plus[name] += eps
    minus[name] -= eps
    op, _ = _hole(tgh, **plus)
    om, _ = _hole(tgh, **minus)
    fd = (op.gh_rhs(u, 0.05, 8) - om.gh_rhs(u, 0.05, 8)) / (2 * eps)
    scale = np.abs(fd).max()
    assert scale > 0
    err = np.abs(fd - dF).max() / scale
    print(f"{PARAMS[k]} tangent vs central difference: {err:.2e} (scale {scale:.3e})")
    assert err < (1e-9 if k < 4 else 1e-6)


@pytest.mark.parametrize("k", [2, 4], ids=["gamma2", "rho_ramp"])
def test_tangent_linear_rk4_matches_central_differences_of_the_error(tgh, k):
    """Three RK4 steps carried with their tangent, and the masked L2 error's
    directional derivative along it: d J / d theta_k against central
    differences of J after the same steps of the plain pass."""
    base = dict(near=1.0, far=0.1, gamma2=0.1, rho_max=4.0, rho_ramp=0.5)
    orc, _ = _hole(tgh, **dict(base))
    dt, n = 0.02, 3
    u0 = orc.fill_exact(0.0)
    pd = np.zeros(5)
    pd[k] = 1.0
    u, ud = u0.copy(), np.zeros_like(u0)
    for s in range(n):
        oracle_lib.rk4_step_jvp(orc, u, ud, s * dt, dt, pd)
    dJ = oracle_lib.error_tangent(orc, u, ud, n * dt)
    name = ("near", "far", "gamma2", "rho_max", "rho_ramp")[k]
    eps = 1e-5

    def J(v):
        args = dict(base)
        args[name] = v
        o, _ = _hole(tgh, **args)
        uu = u0.copy()
        for s in range(n):
            o.rk4_step(uu, s * dt, dt, 8)
        return o.error_norms(uu, n * dt)[0]

    fd = (J(base[name] + eps) - J(base[name] - eps)) / (2 * eps)
    print(f"dJ/d{PARAMS[k]}: tangent {dJ:.10e} central difference {fd:.10e}")
    assert abs(dJ - fd) <= 1e-6 * abs(fd)
