"""The C ABI library loads and exports every entry point include/tgh.h
declares (no compute call: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "tgh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tgh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_verbs():
    syms = declared_symbols()
    for verb in ("tgh_prepare", "tgh_rhs", "tgh_gh_rhs", "tgh_settle", "tgh_stage_limit",
                 "tgh_step_limit", "tgh_max_speed", "tgh_initial_state", "tgh_rk4",
                 "tgh_problem_create", "tgh_last_error"):
        assert verb in syms


def test_library_exports_every_declared_symbol(tgh):
    path = tgh.library_path()
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_descriptor_layout_matches_the_header(tgh, tmp_path):
    """Every field of the ctypes tgh_desc sits at the offset the C compiler gives
    it in include/tgh.h, and the sizes agree."""
    import subprocess
    import tgh_desc
    assert [f[0] for f in tgh.TghDesc._fields_] == [f[0] for f in tgh_desc.TghDesc._fields_]
    names = [f[0] for f in tgh.TghDesc._fields_]
    src = tmp_path / "layout.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "tgh.h"\nint main(void) {\n'
                   '  printf("%zu\\n", sizeof(tgh_desc));\n' +
                   "".join(f'  printf("%zu\\n", offsetof(tgh_desc, {n}));\n' for n in names) + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == ctypes.sizeof(tgh.TghDesc)
    assert out[1:] == [getattr(tgh.TghDesc, n).offset for n in names]


def test_descriptor_checks_refuse_bad_dissipation_profiles(tgh):
    """HorizonDissipation's refusals (src/gauge.jl:1229-1252) at problem creation:
    the descriptor is checked before any device call."""
    import tgh_desc as TD
    L = tgh._load()
    for kw in (dict(eps_in=-1.0), dict(eps_in=5.0), dict(eps_r1=2.0), dict(eps_r1=0.0), dict(dissipation=7)):
        d = TD.make_desc(q=4, N=16, nblocks=(3, 3, 3), periodic=(0, 0, 0), lo=(-3, -3, -3), hi=(3, 3, 3),
                         background=TD.TGH_BG_KERR_SCHILD, gauge=TD.TGH_GAUGE_KERR_SCHILD, M=1.0, a=0.5,
                         eps_ko=0.5, dissipation=TD.TGH_DISSIPATION_HORIZON, eps_in=3.0, eps_r1=1.71,
                         eps_rh=1.866)
        for k, v in kw.items():
            setattr(d, k, v)
        ptr = ctypes.c_void_p()
        assert L.tgh_problem_create(ctypes.byref(tgh.TghDesc.from_buffer_copy(d)), ctypes.byref(ptr)) == 1, kw
        assert L.tgh_last_error()


def test_descriptor_checks_refuse_bad_harmonic_cases(tgh):
    """The harmonic chart takes no gauge source (src/gauge.jl:64) and its frozen core
    contains the singular disk |x| <= |a| (src/interior.jl:1000, 1120-1126)."""
    import tgh_desc as TD
    L = tgh._load()
    good = TD.harmonic_hole_case()
    ptr = ctypes.c_void_p()
    rc = L.tgh_problem_create(ctypes.byref(tgh.TghDesc.from_buffer_copy(good)), ctypes.byref(ptr))
    assert rc != 1, L.tgh_last_error()  # passes the checks (no device here: rc 2)
    if rc == 0:
        L.tgh_problem_destroy(ptr)
    for kw in (dict(gauge=TD.TGH_GAUGE_SAMPLED), dict(gauge=TD.TGH_GAUGE_KERR_SCHILD), dict(r0=0.25),
               dict(background=5)):
        d = tgh.TghDesc.from_buffer_copy(good)
        for k, v in kw.items():
            setattr(d, k, v)
        ptr = ctypes.c_void_p()
        assert L.tgh_problem_create(ctypes.byref(d), ctypes.byref(ptr)) == 1, kw
        assert L.tgh_last_error()


def test_python_harmonic_case_builds_the_descriptor(tgh):
    import tgh_desc as TD
    case = tgh.harmonic_kerr_case(M=1.0, a=0.3, halfwidth=2.0, r_0=0.35, r_1=0.9)
    assert abs(tgh.horizon_min_radius(case.background) - 0.91 ** 0.5) < 1e-15
    d = tgh._desc(tgh.gh_forest(None, case, N=16, roots=4), case, 4)
    ref = TD.harmonic_hole_case()
    ref.wave_amp, ref.wave_len = d.wave_amp, d.wave_len  # the gauge wave's, unused here
    assert bytes(d) == bytes(ref)
    with pytest.raises(tgh.ArgumentError):
        tgh.Harmonic(1.0, 1.0)


def test_python_case_builds_the_horizon_profile(tgh):
    case = tgh.kerr_schild_case(M=1.0, a=0.5, halfwidth=3.0, r_0=0.71, r_1=1.71)
    hd = tgh.horizon_dissipation(case, eps_in=3.0)
    assert hd.r_1 == 1.71 and abs(hd.r_h - (1 + 0.75 ** 0.5)) < 1e-15 and hd.eps_out == 0.5
    case2 = tgh.kerr_schild_case(M=1.0, a=0.5, halfwidth=3.0, r_0=0.71, r_1=1.71, eps_ko=hd)
    d = tgh._desc(tgh.gh_forest(None, case2, N=16, roots=3), case2, 4)
    assert (d.dissipation, d.eps_ko, d.eps_in, d.eps_r1) == (1, 0.5, 3.0, 1.71)
    with pytest.raises(tgh.ArgumentError):
        tgh.HorizonDissipation(eps_out=0.5, eps_in=5.0, r_1=1.0, r_h=2.0)
    with pytest.raises(tgh.ArgumentError):
        tgh.HorizonDissipation(eps_out=0.5, eps_in=1.0, r_1=2.0, r_h=1.0)


def test_float_type_first_refuses_other_types(tgh):
    with pytest.raises(tgh.ArgumentError):
        tgh.gauge_wave_case("Float32", eps_ko=0.0, gamma0=1.0, gamma2=0.0)
    with pytest.raises(tgh.ArgumentError):
        tgh.Interior(r_0=1.0, r_1=0.5)


def test_forest_checks_refuse_bad_leaf_lists(tgh):
    """tgh_problem_create_forest checks the leaves before any device call: they
    tile the box, are 2:1 balanced (across periodic seams too), N is even and
    at least 2G, and the prolongation order fits the ghosts."""
    import tgh_desc as TD
    L = tgh._load()
    good = [(1, i, j, k) for k in (0, 1) for j in (0, 1) for i in (0, 1)] + \
           [(0, i, j, k) for k in (0, 1) for j in (0, 1) for i in (0, 1) if (i, j, k) != (0, 0, 0)]

    def create(leaves, N=8, p=6):
        d = tgh.TghDesc.from_buffer_copy(TD.gauge_wave_case(N=N, roots=2, q=4))
        flat = [x for lf in leaves for x in lf]
        arr = (ctypes.c_int32 * len(flat))(*flat)
        ptr = ctypes.c_void_p()
        rc = L.tgh_problem_create_forest(ctypes.byref(d), arr, len(leaves), p, ctypes.byref(ptr))
        if rc == 0:
            L.tgh_problem_destroy(ptr)
        return rc, L.tgh_last_error().decode()

    rc, msg = create(good)
    assert rc != 1, msg  # passes the checks (no device here: rc 2)
    overlap = good + [(0, 0, 0, 0)]
    assert create(overlap) == (1, create(overlap)[1]) and "overlap" in create(overlap)[1]
    hole = good[1:]
    assert create(hole)[0] == 1 and "cover" in create(hole)[1]
    # a level-2 leaf beside a level-0 one across the periodic seam
    deep = [lf for lf in good if lf != (1, 0, 0, 0)] + \
           [(2, i, j, k) for k in (0, 1) for j in (0, 1) for i in (0, 1)]
    rc, msg = create(deep)
    assert rc == 1 and "balanced" in msg
    assert create(good, N=9)[0] == 1
    assert create(good, N=4)[0] == 1
    assert create(good, p=8)[0] == 1 and "prolongation" in create(good, p=8)[1]
    assert create(good, p=3)[0] == 1
    assert create([(0, 2, 0, 0)] + good[1:])[0] == 1
