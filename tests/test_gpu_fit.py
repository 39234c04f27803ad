"""The target cache's fill (fit_target_kernel!, src/fit.jl:1057-1151) on the
device against the oracle, through the C ABI and the Python mirror:

* a near-flat fit that varies in space, alone, with clear: every one of the
  40 slots at every owned point to 1e-13 of the data (the device's exp may
  round apart from the host's by an ulp; everything else is the same
  arithmetic without contraction);
* the same fit moving, with an older previous fit and the translation rate,
  into a cache whose outside points keep what they held;
* a least-squares fit of the a = 0.7 Kerr–Schild hole (degree 8, cont 2, Π̃,
  bounds that fire where the fit is extrapolated toward the center);
* tgh.fill_target through GHProblem with a FittedInterior equals the C call;
* the refusals: no tracked geometry, a moving fit's rate without a spacing,
  a degree beyond 32.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
import tgh_desc as TD
from test_gpu_parity import tgh_problem_from_desc, torch_cuda
from test_oracle_fit import _fitted_problem, _ks_fit, _smooth_fit_coeffs

pytestmark = pytest.mark.gpu


def _dev_fit(torch, host_fit):
    """The same tgh_fit with its coefficients on the device."""
    f = TD.TghFit.from_buffer_copy(host_fit)
    dev = torch.from_numpy(host_fit._host.copy()).cuda()
    f.coeffs = dev.data_ptr()
    f._dev = dev
    return f


def _both(tgh, d, o, fa, fb, t, h, rate, clear, fill=0.0):
    torch = torch_cuda()
    p = tgh_problem_from_desc(tgh, d)
    L = tgh._load()
    out = torch.full((2 * o.n,), fill, dtype=torch.float64, device="cuda")
    da = _dev_fit(torch, fa)
    db = _dev_fit(torch, fb) if fb is not None else None
    tgh._check(L.tgh_fill_target(p.ptr, C.c_void_p(out.data_ptr()), C.byref(tgh.TghFit.from_buffer_copy(da)),
                                 C.byref(tgh.TghFit.from_buffer_copy(db)) if db is not None else None,
                                 float(t), float(h), 1 if rate else 0, 1 if clear else 0))
    ref = o.fill_target(np.full(2 * o.n, fill), fa, fb, t, h, rate=rate, clear=clear)
    g = out.cpu().numpy()
    L.tgh_problem_destroy(p.ptr)
    return g, ref


def _rel(g, ref):
    """max |g − ref| over the largest |ref|, NaN where exactly one side is NaN."""
    both = np.isnan(g) & np.isnan(ref)
    if np.any(np.isnan(g) != np.isnan(ref)):
        return float("nan")
    d = np.where(both, 0.0, g - ref)
    return float(np.abs(d).max() / max(np.abs(np.where(both, 0.0, ref)).max(), 1e-300))


def test_fill_target_matches_oracle(tgh):
    d, o = _fitted_problem()
    h = (d.hi[0] - d.lo[0]) / (d.nblocks[0] * d.N)
    C0 = _smooth_fit_coeffs(4, 1)
    fa = TD.make_fit(C0, lmax=4, cont=1, rbar=1.5, tilde=True, t=1.0)
    g, ref = _both(tgh, d, o, fa, None, 1.0, h, rate=False, clear=True)
    e1 = _rel(g, ref)
    # moving, with an older fit and the rate, into a cache that keeps its outside
    fm = TD.make_fit(C0, lmax=4, cont=1, rbar=1.5, tilde=True, t=1.0, velocity=(0.3, -0.1, 0.05),
                     center=(0.05, 0.0, -0.02))
    fb = TD.make_fit(_smooth_fit_coeffs(4, 1, seed=9), lmax=4, cont=1, rbar=1.5, tilde=True, t=0.75)
    g2, ref2 = _both(tgh, d, o, fm, fb, 1.1, h, rate=True, clear=False, fill=7.0)
    e2 = _rel(g2, ref2)
    print(f"fill: alone {e1:.2e}, moving + previous + rate {e2:.2e}")
    assert e1 <= 1e-13 and e2 <= 1e-13
    assert np.count_nonzero(g2 == 7.0) == np.count_nonzero(ref2 == 7.0) > 0


def test_fill_target_of_a_hole_fit_matches_oracle(tgh):
    d, o = _fitted_problem(r_h=2.2)
    h = (d.hi[0] - d.lo[0]) / (d.nblocks[0] * d.N)
    bnd = (0.02, 50.0, 0.01, 1000.0, 10.0, 100.0)
    fa = TD.make_fit(_ks_fit(8, 2, 2.0), lmax=8, cont=2, rbar=2.0, tilde=True, bounds=bnd)
    g, ref = _both(tgh, d, o, fa, None, 0.0, h, rate=False, clear=True)
    e = _rel(g, ref)
    print(f"the hole's fit: {e:.2e} of the data")
    assert np.array_equal(np.isnan(g), np.isnan(ref)) and e <= 1e-12


def test_fill_target_through_the_python_mirror(tgh):
    torch = torch_cuda()
    case = tgh.kerr_schild_case(M=1.0, a=0.5, halfwidth=2.0, r_0=0.75, r_1=1.25, interior="damped",
                                gauge_source="algebraic")
    it = tgh.FittedInterior(center=(0.0, 0.0, 0.0), shape=[1.5 * np.sqrt(4 * np.pi)], offset=0.2,
                            thickness=0.5, r_in=1.5, r_out=1.5, h=0.0625)
    case = tgh.with_interior(case, it)
    p = tgh.GHProblem(tgh.gh_forest(None, case, N=16, roots=4), case, q=4)
    bnd = tgh.StateBounds(alpha_min=1e-3, alpha_max=1e3, lambda_min=1e-4, lambda_max=1e4, beta_max=1e3,
                          K_max=1e6, r_gate=1.0)
    C0 = _smooth_fit_coeffs(4, 1)
    prm = tgh.FitParams(lmax=4, cont=1, rbar=1.5, center=((0.0, 0.0, 0.0), (0.3, 0.0, 0.0)), bounds=bnd,
                        tilde=True)
    fit = tgh.InteriorFit(prm, torch.from_numpy(C0).cuda(), 0.0)
    tw = tgh.fill_target(tgh.target_cache(p), p, (fit, None), 0.0, rate=True)
    d, o = _fitted_problem()
    assert d.N == 16 and tuple(d.nblocks) == (4, 4, 4)
    fa = TD.make_fit(C0, lmax=4, cont=1, rbar=1.5, tilde=True, velocity=(0.3, 0.0, 0.0),
                     bounds=TD.WIDE_BOUNDS)
    ref = o.fill_target(np.zeros(2 * o.n), fa, None, 0.0, 0.0625, rate=True)
    e = _rel(tw.cpu().numpy(), ref)
    print(f"tgh.fill_target against the oracle: {e:.2e}")
    assert e <= 1e-13


def test_fill_target_refusals(tgh):
    torch = torch_cuda()
    L = tgh._load()
    d, o = _fitted_problem()
    C0 = _smooth_fit_coeffs(2, 0)
    fa = _dev_fit(torch, TD.make_fit(C0, lmax=2, cont=0, velocity=(0.2, 0.0, 0.0)))
    out = torch.zeros(2 * o.n, dtype=torch.float64, device="cuda")
    plain = TD.kerr_schild_hole_case(N=16, nblocks=4, halfwidth=2.0, a=0.5, r0=0.75, r1=1.25)
    for desc, f, h, rate, what in [(plain, fa, 0.1, 0, "not shaped"), (d, fa, 0.0, 1, "moving without h")]:
        p = tgh_problem_from_desc(tgh, desc)
        rc = L.tgh_fill_target(p.ptr, C.c_void_p(out.data_ptr()), C.byref(tgh.TghFit.from_buffer_copy(f)), None,
                               0.0, h, rate, 1)
        assert rc == tgh.TGH_EINVAL, what
        L.tgh_problem_destroy(p.ptr)
    p = tgh_problem_from_desc(tgh, d)
    bad = tgh.TghFit.from_buffer_copy(fa)
    bad.lmax = 33
    assert L.tgh_fill_target(p.ptr, C.c_void_p(out.data_ptr()), C.byref(bad), None, 0.0, 0.1, 0, 1) == tgh.TGH_EINVAL
    L.tgh_problem_destroy(p.ptr)
