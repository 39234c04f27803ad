"""The brick Pi kernel's tile shapes and the autotuner that picks one: every
shape stages the same points and computes each with the same arithmetic, so
an RK4 step must give the same bits whatever the shape (checked against the
oracle for the default shape in test_gpu_parity.py)."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    return torch


@pytest.mark.parametrize("q", [4, 8])
def test_every_tile_shape_gives_the_same_bits(tgh, q):
    torch = torch_cuda()
    case = tgh.kerr_schild_case(M=1.0, a=0.5, halfwidth=2.0, r_0=0.71, r_1=1.5, gauge_source="algebraic")
    p = tgh.GHProblem(tgh.gh_forest(None, case, N=16, roots=4), case, q=q)
    u0 = p.statevector()
    tgh.initial_state(u0, p, 0.0)
    tgh.add_noise(u0, p, 0x711E, 1e-8)
    dt = tgh.gh_dt(p, u0, cfl=0.25)
    L = tgh._load()
    ref = None
    for tile in range(tgh.TGH_PI_TILES):
        tgh._check(L.tgh_set_tile(p.ptr, tile))
        for graph in (0, 1):
            u = u0.clone()
            tgh._check(L.tgh_rk4(p.ptr, C.c_void_p(u.data_ptr()), 0.0, dt, 2, 8 | graph))
            torch.cuda.synchronize()
            assert torch.isfinite(u).all()
            if ref is None:
                ref = u
            d = float((u - ref).abs().max())
            print(f"q={q} tile {tile} graph {graph}: max |u - u(tile 0)| = {d:.3e}")
            assert torch.equal(u, ref)


def test_autotune_keeps_the_fastest_shape(tgh):
    torch = torch_cuda()
    case = tgh.kerr_schild_case(M=1.0, a=0.7, halfwidth=3.0, r_0=0.71, r_1=1.21, gauge_source="algebraic")
    p = tgh.GHProblem(tgh.gh_forest(None, case, N=16, roots=6), case, q=4)
    u = p.statevector()
    tgh.initial_state(u, p, 0.0)
    dt = tgh.gh_dt(p, u, cfl=0.25)
    before = u.clone()
    ms = tgh.autotune(p, u, dt)
    print("Pi ms per tile shape:", ms)
    assert len(ms) == tgh.TGH_PI_TILES and all(m > 0 for m in ms)
    assert torch.equal(u, before)  # the tuner steps a copy
    with pytest.raises(tgh.ArgumentError):
        tgh._check(tgh._load().tgh_set_tile(p.ptr, tgh.TGH_PI_TILES))
