"""The reference's coefficient layouts at every parallel level, pinned by hand-derived arrays.

The arrays below are written out by hand from the rules of chemgen/write_constants.py
(arrhenius :14-113, maps :196-236, coefficients :238-277, efficiencies :279-314, rocBLAS
matrices :350-406) and chemgen/chemistry.py (MAX_SP :124-133, dummy reactions :143-183) for a
two-reaction mechanism -- an elementary reaction with three distinct reactants (MAX_SP = 3,
rounded to 4 at level 3) and a three-body reaction with two efficiencies -- padded to two
reactions per type.  They pin this build's coefficient-module emitter (chemgen_amd/legacy_coef.py)
and, through it, the layouts the compatibility libraries read (tests/test_gpu_legacy_abi.py).
The compatibility libraries' converters are checked on the host for every level, and each
refuses another level's layout.
"""
import ctypes
import math
import os

import numpy as np
import pytest

from chemgen_amd.kernels import load_bundled
from chemgen_amd.legacy_coef import emit_coef_module, max_species, parse_coef_module
from chemgen_amd.mechanism import Mechanism, Reaction

RC = 1.9872155832
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LEGACY_DIR = os.path.join(ROOT, "chemgen_amd", "_kernels", "legacy_h2_burke")


def _two_reaction_mech():
    h2 = load_bundled("h2_burke")
    by = {s.name: s for s in h2.species}
    sp = [by[n] for n in ("H", "O2", "O", "OH", "HO2", "N2")]       # indices 0..5
    r0 = Reaction("H + O + OH <=> HO2 + H", {"H": 1.0, "O": 1.0, "OH": 1.0}, {"HO2": 1.0, "H": 1.0},
                  "standard", True, (2.0e13, 0.5, 1000.0))
    r1 = Reaction("H + O2 + M <=> HO2 + M", {"H": 1.0, "O2": 1.0}, {"HO2": 1.0}, "third_body", True,
                  (6.366e20, -1.72, 524.8), third_body={"O2": -0.22, "N2": 0.5})
    return Mechanism("two", ["H", "O", "N"], sp, [r0, r1])


def _close(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, (a.shape, b.shape)
    assert np.allclose(a, b, rtol=1e-15, atol=0.0), (a, b)


def test_max_sp_rounding():
    m = _two_reaction_mech()
    assert [max_species(m, lv) for lv in (1, 2, 3, 4)] == [3, 3, 4, 4]
    h2 = load_bundled("h2_burke")
    assert max_species(h2, 1) == 3 and max_species(h2, 3) == 4


@pytest.mark.parametrize("level", [2, 3])
def test_maps_levels_2_and_3(level):
    m = _two_reaction_mech()
    a = parse_coef_module(emit_coef_module(m, level, 2))
    # Arrhenius: ln A, then (beta, -Ea/Rc) pairs; the dummy reaction has A = 1, beta = Ea = 0
    _close(a["A_h"], [math.log(2.0e13), 0.0])
    _close(a["B_h"], [0.5, -1000.0 / RC, 0.0, 0.0])
    _close(a["A_third_h"], [math.log(6.366e20), 0.0])
    _close(a["B_third_h"], [-1.72, -524.8 / RC, 0.0, 0.0])
    if level == 2:   # MAX_SP = 3: reactants then products, each padded with species 0
        assert list(a["map_r_h"]) == [0, 2, 3, 0, 0, 0]
        assert list(a["map_p_h"]) == [4, 0, 0, 0, 0, 0]
        assert list(a["sk_map_h"]) == [0, 2, 3, 4, 0, 0, 0, 0, 0, 0, 0, 0]
        _close(a["sk_coef_h"], [-1, -1, -1, 1, 1, 0, 0, 0, 0, 0, 0, 0])
        _close(a["coef_r_h"], [1, 1, 1, 0, 0, 0])
        _close(a["coef_p_h"], [1, 1, 0, 0, 0, 0])
        assert list(a["map_r_third_h"]) == [0, 1, 0, 0, 0, 0]
        assert list(a["map_p_third_h"]) == [4, 0, 0, 0, 0, 0]
        _close(a["sk_coef_third_h"], [-1, -1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0])
    else:            # MAX_SP = 4
        assert list(a["map_r_h"]) == [0, 2, 3, 0, 0, 0, 0, 0]
        assert list(a["map_p_h"]) == [4, 0, 0, 0, 0, 0, 0, 0]
        assert list(a["sk_map_h"]) == [0, 2, 3, 0, 4, 0, 0, 0] + [0] * 8
        _close(a["sk_coef_h"], [-1, -1, -1, 0, 1, 1, 0, 0] + [0] * 8)
        _close(a["coef_r_h"], [1, 1, 1, 0, 0, 0, 0, 0])
        _close(a["coef_p_h"], [1, 1, 0, 0, 0, 0, 0, 0])
        assert list(a["map_r_third_h"]) == [0, 1, 0, 0, 0, 0, 0, 0]
        assert list(a["map_p_third_h"]) == [4, 0, 0, 0, 0, 0, 0, 0]
        _close(a["sk_coef_third_h"], [-1, -1, 0, 0, 1, 0, 0, 0] + [0] * 8)
    # dense 1 + (eff - 1) per species, the dummy's efficiencies all 1
    _close(a["eff_fac_third_h"], [1.0, 0.78, 1.0, 1.0, 1.0, 1.5] + [1.0] * 6)
    assert "wdot_coef_h" not in a


def test_dense_matrices_level_4():
    m = _two_reaction_mech()
    a = parse_coef_module(emit_coef_module(m, 4, 2))
    assert "map_r_h" not in a and "sk_map_h" not in a
    # [reaction][species]: H is on both sides of reaction 0, so its net coefficient is 0
    _close(a["sk_coef_h"], [0, 0, -1, -1, 1, 0] + [0] * 6)
    _close(a["coef_r_h"], [1, 0, 1, 1, 0, 0] + [0] * 6)
    _close(a["coef_p_h"], [1, 0, 0, 0, 1, 0] + [0] * 6)
    # wdot_coef [species][reaction]
    _close(a["wdot_coef_h"], [0, 0, 0, 0, -1, 0, -1, 0, 1, 0, 0, 0])
    _close(a["sk_coef_third_h"], [-1, -1, 0, 0, 1, 0] + [0] * 6)
    _close(a["coef_r_third_h"], [1, 1, 0, 0, 0, 0] + [0] * 6)
    _close(a["coef_p_third_h"], [0, 0, 0, 0, 1, 0] + [0] * 6)
    _close(a["wdot_coef_third_h"], [-1, 0, -1, 0, 0, 0, 0, 0, 1, 0, 0, 0])
    _close(a["eff_fac_third_h"], [1.0, 0.78, 1.0, 1.0, 1.0, 1.5] + [1.0] * 6)


def test_level_1_without_padding():
    m = _two_reaction_mech()
    a = parse_coef_module(emit_coef_module(m, 1))
    _close(a["A_h"], [math.log(2.0e13)])
    assert list(a["map_r_h"]) == [0, 2, 3]
    _close(a["eff_fac_third_h"], [1.0, 0.78, 1.0, 1.0, 1.0, 1.5])


# ---- the compatibility libraries' converters (host only) ----

_LIBS = {1: "liblegacy_h2_burke.so", 2: "liblegacy_h2_burke.so", 3: "liblegacy_h2_burke_v3.so",
         4: "liblegacy_h2_burke_v4.so"}


def level_args(mech, level, rpb=8, pad_level_1=False):
    """copyConstantsToDevice's arguments, in order, from this build's coefficient module of the
    level (what chemgen_m passes); level 1 unpadded unless ``pad_level_1`` (the reference's
    gpu_coef mode pads it too, with its default of 8 reactions per block)."""
    from chemgen_amd.legacy_abi import _params, _params4
    arrs = parse_coef_module(emit_coef_module(mech, level, rpb if level > 1 or pad_level_1 else None))
    present = [t for t in ("standard", "troe", "third_body") if any(r.rtype == t for r in mech.reactions)]
    names = [p.split()[-1] for p in (_params(present) if level < 4 else _params4(present))]
    return [np.ascontiguousarray(arrs[n]) for n in names]


def load_level_library(level):
    path = os.path.join(LEGACY_DIR, _LIBS[level])
    if not os.path.exists(path):
        pytest.skip("compatibility libraries not built (run __graft_entry__.build())")
    from chemgen_amd.runtime import lib
    lib()
    return ctypes.CDLL(path)


@pytest.mark.parametrize("level", [1, 2, 3, 4])
def test_converter_reads_each_level(level):
    from chemgen_amd.codegen import build_coef
    L = load_level_library(level)
    m = load_bundled("h2_burke")
    args = level_args(m, level)
    ref, _, _ = build_coef(m)
    cf = np.full(len(ref), np.nan)
    rc = L.cg_legacy_convert(*[ctypes.c_void_p(a.ctypes.data) for a in args], ctypes.c_void_p(cf.ctypes.data))
    assert rc == 0
    scale = np.maximum(np.abs(ref), 1.0)
    assert np.all(np.isfinite(cf) | np.isinf(ref))
    assert np.max(np.abs(np.where(np.isinf(ref), 0.0, cf - ref)) / scale) < 1e-13


def test_converter_reads_padded_level_1():
    """The reference's default coefficient module (level 1, each type padded to 8 with dummy
    reactions) converts to the same blob as the unpadded one."""
    from chemgen_amd.codegen import build_coef
    L = load_level_library(1)
    m = load_bundled("h2_burke")
    ref, _, _ = build_coef(m)
    outs = []
    for pad in (False, True):
        args = level_args(m, 1, pad_level_1=pad)
        cf = np.full(len(ref), np.nan)
        assert L.cg_legacy_convert(*[ctypes.c_void_p(a.ctypes.data) for a in args],
                                   ctypes.c_void_p(cf.ctypes.data)) == 0
        outs.append(cf)
    assert np.array_equal(outs[0], outs[1], equal_nan=True)


def test_converter_refuses_other_level():
    """The level-1 library refuses level-3 arrays (other MAX_SP stride) and vice versa."""
    L1, L3 = load_level_library(1), load_level_library(3)
    m = load_bundled("h2_burke")
    cf = np.zeros(4096)
    for L, lv in ((L1, 3), (L3, 1)):
        args = level_args(m, lv)
        assert L.cg_legacy_convert(*[ctypes.c_void_p(a.ctypes.data) for a in args],
                                   ctypes.c_void_p(cf.ctypes.data)) == 2
