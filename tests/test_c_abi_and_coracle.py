"""C ABI library exports and the C oracle against the NumPy oracle (CPU only)."""
import ctypes
import os
import re

import numpy as np
import pytest

from chemgen_amd.kernels import load_bundled
from chemgen_amd.synth import make_states
from oracle.getrates_np import getrates
from oracle.metrics import parity_stats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "chemgen_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cg_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from chemgen_amd.runtime import LIB_PATH
    if not os.path.exists(LIB_PATH):
        import __graft_entry__
        __graft_entry__.build_runtime()
    lib = ctypes.CDLL(LIB_PATH)
    syms = _declared_symbols()
    assert "cg_getrates" in syms and "cg_mech_load" in syms
    for s in syms:
        assert hasattr(lib, s), s


def test_error_path_without_gpu():
    """Argument errors are reported through cg_last_error, not crashes."""
    from chemgen_amd.runtime import lib
    L = lib()
    assert L.cg_getrates(None, None, None, 0, 0, None, None, None) == -1
    assert b"null handle" in L.cg_last_error()
    h = ctypes.c_void_p()
    assert L.cg_mech_load(None, None, 0, 0, ctypes.byref(h)) == -1


@pytest.mark.parametrize("name", ["h2_burke", "gri30", "synth_gri53"])
def test_c_oracle_matches_numpy_oracle(name):
    from oracle.build import COracle, build_one, so_path
    m = load_bundled(name)
    if not os.path.exists(so_path(name)):
        build_one(m, name)
    P, T, Y = make_states(m.n_species, 512, seed=11)
    ref, qf, qb = getrates(m, P, T, Y, return_parts=True)
    w = COracle(name).getrates(P, T, Y)
    st = parity_stats(m, w, ref, qf, qb)
    assert st["max_gross"] <= 1e-13, st


def test_surrogate_entry_points_reject_bad_arguments_without_gpu():
    """Negative sizes and missing buffers are refused before any launch."""
    from chemgen_amd import surrogate
    from chemgen_amd.runtime import lib
    surrogate._bind()
    L = lib()
    x = 16                                   # a non-null stand-in: never dereferenced on these paths
    cases = [
        ("cg_mlp_targets", (x, 8, -1, 4, x, x, None)),
        ("cg_mlp_targets", (x, 4, 8, 4, x, x, None)),           # leading dimension below ng
        ("cg_mlp_targets", (None, 8, 8, 4, x, x, None)),
        ("cg_mlp_mse_grad", (x, x, -1, 1.0, x, None, None)),
        ("cg_mlp_silu_back", (x, x, -5, x, None)),
        ("cg_mlp_colsum", (x, -1, 4, x, None)),
        ("cg_mlp_colsum", (None, 4, 4, x, None)),
        ("cg_mlp_adam", (x, x, x, x, -1, 1e-3, 0.9, 0.999, 1e-8, 1, None, None)),
        ("cg_mlp_adam", (x, x, x, x, 8, 1e-3, 0.9, 0.999, 1e-8, 0, None, None)),
        ("cg_mlp_to_bf16", (x, -1, x, None)),
        ("cg_mlp_decode", (x, -1, 4, x, x, 2, 1e-2, x, 8, x, x, None)),
        ("cg_mlp_decode", (x, 8, 4, x, x, 2, 1e-2, x, 8, None, x, None)),
        ("cg_mlp_compact", (x, -1, 64, x, x, None, None)),
        ("cg_mlp_compact", (None, 8, 64, x, x, None, None)),
        ("cg_mlp_audit", (x, 4, x, 4, 8, 4, x, 1e-2, 0.05, x, x, x, None, None)),   # leading dimension below n
        ("cg_mlp_audit", (x, 8, x, 8, 8, 4, x, 1e-2, 0.05, None, x, x, None, None)),
        ("cg_mlp_audit", (x, 8, x, 8, 8, 4, x, 1e-2, 0.05, x, x, None, None, None)),  # no workspace
        ("cg_mlp_fallback_plan", (x, x, 0, 0.5, 64, x, x, x, x, None)),
        ("cg_mlp_fallback_plan", (x, x, 8, 0.5, 0, x, x, x, x, None)),
        ("cg_mlp_fallback_plan", (x, None, 8, 0.5, 64, x, x, x, x, None)),
        ("cg_count_tiles", (x, 8, 0, x, x, None)),
        ("cg_count_tiles", (None, 8, 64, x, x, None)),
        ("cg_mlp_gather", (x, x, x, 8, x, x, -1, 4, x, x, x, 8, None)),
        ("cg_mlp_gather", (x, x, x, 8, x, x, 8, 4, x, x, x, 4, None)),
        ("cg_mlp_scatter", (x, 8, x, x, -1, 4, x, 8, None)),
        ("cg_mlp_scatter", (x, 8, None, x, 8, 4, x, 8, None)),
    ]
    for name, args in cases:
        assert getattr(L, name)(*args) == -1, (name, args)
        assert b"bad argument" in L.cg_last_error(), name
    assert L.cg_mlp_decode(x, 8, 4, x, x, 9, 1e-2, x, 8, x, x, None) != 0
    assert b"at most 8" in L.cg_last_error()
