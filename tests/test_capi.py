"""The C-ABI library loads without a GPU and exports every entry point the
header declares; argument errors come back as status codes, not exits."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "glimpse_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gh_[a-z_0-9]+)\s*\(", txt)))


def test_header_symbols_exported():
    from glimpse_amd import hip
    lib = hip.lib()
    names = _declared()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bad_arguments_return_status():
    from glimpse_amd import hip
    lib = hip.lib()
    rc = lib.gh_panel_create(None, None)
    assert rc == -1
    assert b"bad descriptor" in lib.gh_last_error()


def test_default_params_match_reference_defaults():
    from glimpse_amd import hip
    p = hip.default_params()
    assert (p.burnin, p.main_iter, p.ne, p.pbwt_depth, p.kinit, p.kpbwt, p.seed) == (5, 15, 100000, 12, 1000, 2000, 15052011)
    assert abs(p.min_gl - 1e-10) < 1e-16 and abs(p.err_imp - 1e-12) < 1e-18 and abs(p.err_phase - 1e-4) < 1e-10
    assert abs(p.pbwt_modulo_cm - 0.1) < 1e-7
