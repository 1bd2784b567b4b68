"""The C-ABI library loads, exports exactly what include/rbg.h declares, does
not link the oracle, and its host-only entry points behave (no GPU needed)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "rbg.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"//[^\n]*", "", src)
    src = re.sub(r"#[^\n]*", "", src)
    return sorted(set(re.findall(r"\b(rbg_[a-z0-9_]+)\s*\(", src)))


def exported_functions(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return sorted({ln.split()[-1] for ln in out.splitlines() if " T rbg_" in ln})


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("rbg_open", "rbg_close", "rbg_sketch_winners", "rbg_build_graph", "rbg_asm_create",
                 "rbg_asm_build_graph", "rbg_edge_weights", "rbg_calibrate", "rbg_sketch_tables"):
        assert must in names


def test_library_exports_exactly_the_header():
    from rabbitbin_amd import LIB_PATH, api
    api.lib()  # loads with ctypes, no device call
    assert exported_functions(LIB_PATH) == declared_functions()
    L = C.CDLL(LIB_PATH)
    for name in declared_functions():
        assert hasattr(L, name), name


def test_library_does_not_link_the_oracle():
    from rabbitbin_amd import LIB_PATH
    needed = subprocess.run(["readelf", "-d", LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert "oracle" not in needed and "ref_probmh" not in needed
    assert "libamdhip64" in needed
    syms = subprocess.run(["nm", "-D", LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert "orc_" not in syms and "ref_pmh" not in syms


def test_default_params_match_reference_cli_defaults():
    from rabbitbin_amd import api
    p = api.default_params()
    assert p.m == 500 and p.max_edges == 200  # --sketch-m 500 --max-edges 200
    assert p.min_edge_weight == 0.70 and p.coverage == 0.95 and p.min_cv == 1.0
    assert p.neg_depth_thr == -0.3 and p.pmh_seed == 42 and p.regime == 0


def test_host_entry_points_reject_bad_arguments():
    from rabbitbin_amd import api
    L = api.lib()
    out = C.c_double()
    rm = np.zeros(4, np.uint32)
    assert L.rbg_calibrate(None, 4, 500, 0.0, 0.95, 1, 1, C.byref(out)) != 0
    assert L.rbg_calibrate(rm.ctypes.data_as(api.u32p), 4, 500, 0.0, 0.95, 1, 3, C.byref(out)) != 0
    assert L.rbg_sketch_tables(1, 42, None, None, None) != 0  # m must be >= 2
    with pytest.raises(api.RbgError):
        api.calibrate(np.zeros(10, np.uint32), 1, 0.0, 1)
