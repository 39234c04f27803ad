"""tgh_gauge_constraint_noalloc, the gauge monitor that allocates nothing
(its partials in the library's static device array), called through the C
ABI against tgh_gauge_constraint on the same state: the norms agree bit for
bit (the same kernels, the same fold order) on two problems in a row, the
second smaller than the first (so the second call reads only what it wrote,
over stale partials of the first), each called twice; and a problem beyond
the array's limit is refused with the limit named."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def _case(tgh, halfwidth):
    return tgh.kerr_schild_case(M=1.0, a=0.7, halfwidth=halfwidth, r_0=0.71, r_1=1.21, gauge_source="algebraic")


def test_noalloc_monitor_equals_the_allocating_one(tgh):
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    L = tgh._load()
    # 64 blocks of 16^3 at q = 4, then 27 blocks of 12^3 at q = 2
    for halfwidth, N, roots, q in ((2.0, 16, 4, 4), (1.5, 12, 3, 2)):
        case = _case(tgh, halfwidth)
        p = tgh.GHProblem(tgh.gh_forest(None, case, N=N, roots=roots), case, q=q)
        u = p.statevector()
        tgh.initial_state(u, p, 0.0)
        tgh.add_noise(u, p, 0x5A5A, 1e-6)
        for rep in range(2):
            a2, ai = (C.c_double * 4)(), (C.c_double * 4)()
            tgh._check(L.tgh_gauge_constraint_noalloc(p.ptr, C.c_void_p(u.data_ptr()), 0.1, a2, ai))
            l2, li = tgh.gauge_constraint_norms(p, u, 0.1)
            print(f"{roots ** 3} blocks of {N}^3, q = {q}, call {rep}: no-alloc {list(a2)} allocating {l2}",
                  flush=True)
            assert list(a2) == l2 and list(ai) == li
            assert all(x > 0 for x in l2)


def test_noalloc_monitor_refuses_a_problem_beyond_its_scratch(tgh):
    """6^3 x 3^3 = 5832 blocks of 16^3 cells need 116 640 workgroup partials,
    above the array's 116 506: refused before any launch, the limit named."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    case = tgh.gauge_wave_case(eps_ko=0.0, gamma0=1.0, gamma2=0.0)
    p = tgh.GHProblem(tgh.gh_forest(None, case, N=16, roots=18), case, q=2)
    assert p.nblocks == 5832
    u = p.statevector()
    tgh.initial_state(u, p, 0.0)
    L = tgh._load()
    a2, ai = (C.c_double * 4)(), (C.c_double * 4)()
    rc = L.tgh_gauge_constraint_noalloc(p.ptr, C.c_void_p(u.data_ptr()), 0.0, a2, ai)
    assert rc == tgh.TGH_EINVAL
    msg = L.tgh_last_error().decode()
    assert "116506" in msg and "5825" in msg, msg
