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


def test_env_semantics_follow_reference_parsing(monkeypatch):
    """rbg_env_semantics reports the RABBIT_* variables whose value changes
    the path's results, parsed as the reference parses them (empty values
    are unset for rb_getenv variables and set for getenv ones); the default
    parameters carry the mask, and RABBIT_NEG_DEPTH becomes neg_depth_thr."""
    from rabbitbin_amd import api
    for k in api.SEMANTICS:
        monkeypatch.delenv(k, raising=False)
    monkeypatch.delenv("RABBIT_NEG_DEPTH", raising=False)
    assert api.env_semantics() == (0, "")
    assert api.default_params().semantics == 0
    # values that keep the defaults
    for k, v in (("RABBIT_MUTUAL_KNN", "1"), ("RABBIT_WIN_BITS", "16"), ("RABBIT_PMHK", "4"),
                 ("RABBIT_GC_NORM_CAP", "0.5"), ("RABBIT_EDGE_CUT", "fixed"), ("RABBIT_DEPTH_SIM", "fuse"),
                 ("RABBIT_DEPTH_FUSE", "min"), ("RABBIT_TILE", "64"), ("RABBIT_NO_COMP_EE", "1"),
                 ("RABBIT_PMH", "")):
        monkeypatch.setenv(k, v)
    assert api.env_semantics() == (0, "")
    # values that change the results
    cases = [("RABBIT_MUTUAL_KNN", "0"), ("RABBIT_WIN_BITS", "32"), ("RABBIT_WIN_BITS", ""),
             ("RABBIT_NO_ABDFIRST", ""), ("RABBIT_PMHK", "5"), ("RABBIT_GC_NORM_CAP", "10"),
             ("RABBIT_EDGE_CUT", "otsu"), ("RABBIT_DEPTH_SIM", "wjac"), ("RABBIT_DEPTH_FUSE", "geo"),
             ("RABBIT_W_COMP", "0.5"), ("RABBIT_EDGE_POWER", "2"), ("RABBIT_PMH", "0")]
    for k, v in cases:
        old = os.environ.get(k)
        monkeypatch.setenv(k, v)
        mask, names = api.env_semantics()
        assert mask == api.SEMANTICS[k] and names == k, (k, v, mask, names)
        assert api.default_params().semantics == mask
        if old is None:
            monkeypatch.delenv(k)
        else:
            monkeypatch.setenv(k, old)
    monkeypatch.setenv("RABBIT_NEG_DEPTH", "-0.5")
    assert api.default_params().neg_depth_thr == -0.5
