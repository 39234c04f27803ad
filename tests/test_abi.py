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


def test_descriptor_layout_matches_the_header(tgh):
    import tgh_desc
    assert ctypes.sizeof(tgh.TghDesc) == ctypes.sizeof(tgh_desc.TghDesc)
    assert [f[0] for f in tgh.TghDesc._fields_] == [f[0] for f in tgh_desc.TghDesc._fields_]
    # 13 int32 (incl. the reserved pad) then 29 doubles, no hidden padding
    assert ctypes.sizeof(tgh.TghDesc) == 13 * 4 + 4 + 29 * 8 - 4 or ctypes.sizeof(tgh.TghDesc) % 8 == 0


def test_float_type_first_refuses_other_types(tgh):
    with pytest.raises(tgh.ArgumentError):
        tgh.gauge_wave_case("Float32", eps_ko=0.0, gamma0=1.0, gamma2=0.0)
    with pytest.raises(tgh.ArgumentError):
        tgh.Interior(r_0=1.0, r_1=0.5)
