"""The adjoint on the device: the vector-Jacobian product of the fused
right-hand side against the oracle's dual-number tangent of the same
right-hand side (<lambda, J v> = <J^T lambda, v>, to roundoff), the gradient
of the masked L2 error after checkpointed RK4 against the oracle's tangent
linear model (the reference's CPU path run with a forward-mode dual), and, at
the differentiation workload's size (the a = 0.7 hole on 64 blocks of 16^3,
200 steps, a checkpoint every 20), against central differences of the
device's own forward runs.

theta = (gamma0_near, gamma0_far, gamma2, rho_max, rho_ramp): GaussianDamping
(src/gauge.jl:1094-1156) and the interior layer (src/interior.jl:178-238)."""
import numpy as np
import pytest

import oracle_lib
import tgh_desc as TD

pytestmark = pytest.mark.gpu


def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    return torch


def _hole_case(tgh, gauge="algebraic", halfwidth=1.75, **kw):
    return tgh.kerr_schild_case(M=1.0, a=0.7, halfwidth=halfwidth, r_0=0.71, r_1=1.21, gauge_source=gauge,
                                gamma2=kw.get("gamma2", 0.1), rho_max=kw.get("rho_max", 4.0),
                                rho_ramp=kw.get("rho_ramp", 0.5))


def _pair(tgh, case, N, roots, q):
    f = tgh.gh_forest(None, case, N=N, roots=roots)
    dev = tgh.GHProblem(f, case, q=q)
    orc = oracle_lib.OracleProblem(TD.TghDesc.from_buffer_copy(dev.desc))
    return dev, orc


# (case, N, roots, q): boxes of half-width 3/2 about the hole, the layer's
# thickness 1/2 at least 2(G+1) spacings
CASES = {
    "kerr_schild_closed_form_q4": lambda tgh: (_hole_case(tgh, halfwidth=1.5), 16, 3, 4),
    "kerr_schild_sampled_q2": lambda tgh: (_hole_case(tgh, gauge=None, halfwidth=1.5), 12, 3, 2),
    "gauge_wave_q4": lambda tgh: (tgh.gauge_wave_case(eps_ko=0.5, gamma0=1.0, gamma2=-0.5), 8, 2, 4),
}


@pytest.mark.parametrize("name", list(CASES))
def test_vjp_is_the_transpose_of_the_oracle_tangent(tgh, name):
    """<lambda, dF(u)[v, pd]> from the oracle's dual pass equals <u_bar, v> +
    <theta_bar, pd> from the device's transposed right-hand side, for random
    lambda (face copies included), v and parameter directions pd."""
    torch = torch_cuda()
    case, N, roots, q = CASES[name](tgh)
    dev, orc = _pair(tgh, case, N, roots, q)
    t = 0.05
    u = orc.fill_exact(t)
    orc.add_noise(u, 0xAD1, 1e-4)
    rng = np.random.default_rng(7)
    lam = rng.standard_normal(u.shape)
    v = rng.standard_normal(u.shape)
    pd = rng.standard_normal(5)
    _, dF = oracle_lib.gh_rhs_jvp(orc, u, v, t, pd)
    ub, thb = tgh.vjp_rhs(dev, torch.from_numpy(u).cuda(), t, torch.from_numpy(lam).cuda())
    ub = ub.cpu().numpy()
    lhs = float(np.dot(lam, dF))
    rhs = float(np.dot(ub, v) + np.dot(thb, pd))
    scale = float(np.abs(lam * dF).sum())
    print(f"{name}: <lam, J v> = {lhs:.15e}, <J^T lam, v> = {rhs:.15e}, |diff| / scale = {abs(lhs - rhs) / scale:.2e}")
    assert abs(lhs - rhs) <= 1e-12 * scale
    if name == "gauge_wave_q4":  # no layer, constant damping: only gamma2 acts
        assert thb[0] == thb[1] == thb[3] == thb[4] == 0.0


def test_adjoint_gradient_matches_the_oracle_tangent_linear_model(tgh):
    """The masked L2 error of the a = 0.7 hole (27 blocks of 12^3 at q = 2,
    closed-form source, :damped layer) after 8 steps, checkpoints every 3
    (a short last segment): each component of the device's reverse-sweep
    gradient equals the oracle's forward tangent, d J = <dJ/du_T, u_dot_T>
    with u_dot carried by the dual pass through every stage."""
    torch = torch_cuda()
    case = _hole_case(tgh, halfwidth=1.5)
    dev, orc = _pair(tgh, case, 12, 3, 2)
    dt, nsteps = 0.01, 8
    u0 = orc.fill_exact(0.0)
    ud = torch.from_numpy(u0.copy()).cuda()
    J, grad, _ = tgh.adjoint(dev, ud, 0.0, dt, nsteps, every=3)
    uT = ud.cpu().numpy()
    want = []
    for k in range(5):
        pd = np.zeros(5)
        pd[k] = 1.0
        u = u0.copy()
        udot = np.zeros_like(u)
        for n in range(nsteps):
            oracle_lib.rk4_step_jvp(orc, u, udot, n * dt, dt, pd)
        if k == 0:
            l2, _ = orc.error_norms(u, nsteps * dt)
            print(f"J device {J:.15e} oracle {l2:.15e}; |u_T device - oracle| {np.abs(uT - u).max():.2e}")
            assert abs(J - l2) <= 1e-12 * l2
        want.append(oracle_lib.error_tangent(orc, u, udot, nsteps * dt))
    for k in range(5):
        print(f"dJ/d{tgh.ADJOINT_PARAMS[k]}: device {grad[k]:.12e} oracle {want[k]:.12e}")
    for k in range(5):
        assert abs(grad[k] - want[k]) <= 1e-9 * max(abs(want[k]), 1e-3 * max(abs(w) for w in want))


def test_adjoint_gradient_matches_central_differences_at_the_workload_size(tgh):
    """The differentiation workload: the a = 0.7 hole on 64 blocks of 16^3
    ([-2, 2]^3, h = 1/16, q = 4, the closed-form source, Gaussian damping),
    200 steps of dt = 1/100 to T = 2M, a checkpoint every 20: the adjoint's
    gradient of the masked L2 error against central differences of the
    device's forward runs, agreeing to 1e-6 relative. The difference step is
    1e-4 times the parameter: its truncation error is O(1e-8) relative, and
    the forward run's roundoff (about 1e-16 of J ~ 1e-2 after 200 steps)
    divided by the step stays below 1e-7 of the smallest derivative
    (dJ/dgamma0_far ~ 3e-4); a step of 1e-6 times gamma0_far = 0.1 measures
    that roundoff instead (2e-6 relative)."""
    torch = torch_cuda()
    base = dict(gamma2=0.1, rho_max=4.0, rho_ramp=0.5)

    def problem(**kw):
        args = dict(base)
        args.update({k: v for k, v in kw.items() if k in base})
        case = _hole_case(tgh, halfwidth=2.0, **args)
        case.gamma0 = tgh.GaussianDamping(near=kw.get("near", 1.0), far=kw.get("far", 0.1), width=3.0)
        return tgh.GHProblem(tgh.gh_forest(None, case, N=16, roots=4), case, q=4)

    dt, nsteps = 0.01, 200
    p = problem()
    u = p.statevector()
    tgh.initial_state(u, p, 0.0)
    u0 = u.clone()
    J, grad, (fwd, rev) = tgh.adjoint(p, u, 0.0, dt, nsteps, every=20)
    print(f"J = {J:.15e}; forward {fwd:.1f} ms, reverse {rev:.1f} ms, ratio {rev / fwd:.2f}")

    def J_of(**kw):
        q = problem(**kw)
        uu = u0.clone()
        tgh.gh_solve(q, uu, (0.0, nsteps * dt), dt=dt)
        return tgh.error_norms(q, uu, nsteps * dt)[0]

    J0 = J_of()
    assert abs(J0 - J) <= 1e-13 * J
    names = ("near", "far", "gamma2", "rho_max", "rho_ramp")
    values = (1.0, 0.1, 0.1, 4.0, 0.5)
    for k, (nm, v0) in enumerate(zip(names, values)):
        eps = 1e-4 * abs(v0)
        fd = (J_of(**{nm: v0 + eps}) - J_of(**{nm: v0 - eps})) / (2 * eps)
        rel = abs(grad[k] - fd) / max(abs(fd), 1e-300)
        print(f"dJ/d{tgh.ADJOINT_PARAMS[k]}: adjoint {grad[k]:.10e} central difference {fd:.10e} rel {rel:.2e}")
        assert rel <= 1e-6
