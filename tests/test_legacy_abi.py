"""The reference's C ABI over the generated kernel (compatibility library): the conversion of
the reference's coefficient arrays into this build's gpu_coef blob, checked on the host."""
import ctypes
import os

import numpy as np
import pytest

from chemgen_amd.codegen import build_coef
from chemgen_amd.kernels import load_bundled
from oracle.legacy_coef import legacy_arrays

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LEGACY = os.path.join(ROOT, "chemgen_amd", "_kernels", "legacy_h2_burke", "liblegacy_h2_burke.so")
SYMBOLS = ["copyConstantsToDevice", "getrates_hip_wrapper_v1", "getrates_hip_wrapper_v2",
           "getrates_hip_wrapper_v3", "getrates_hip_wrapper_v4", "allocate_intermediate_DeviceMemory",
           "cg_legacy_convert"]


def _load():
    if not os.path.exists(LEGACY):
        pytest.skip("compatibility library not built (run __graft_entry__.build())")
    from chemgen_amd.runtime import lib
    lib()  # the shared runtime first (its soname satisfies the compatibility library)
    return ctypes.CDLL(LEGACY)


def _ptrs(arrs):
    return [ctypes.c_void_p(a.ctypes.data) for a in arrs.values()]


def test_exports_reference_symbols():
    L = _load()
    for s in SYMBOLS:
        assert hasattr(L, s), s


def test_convert_reproduces_gpu_coef():
    """The reference's arrays (ln A, (beta, -Ea/Rc), Troe centre, dense efficiencies, NASA-7
    transform, molecular weights, all printed to 16 digits) convert into the blob this
    build derives from the mechanism directly, to the rounding of those 16 digits."""
    L = _load()
    m = load_bundled("h2_burke")
    arrs = legacy_arrays(m)
    ref, _, _ = build_coef(m)
    cf = np.full(len(ref), np.nan)
    rc = L.cg_legacy_convert(*_ptrs(arrs), ctypes.c_void_p(cf.ctypes.data))
    assert rc == 0
    assert np.all(np.isfinite(cf) | np.isinf(ref))
    scale = np.maximum(np.abs(ref), 1.0)
    assert np.max(np.abs(cf - ref) / scale) < 1e-13


def test_convert_rejects_other_mechanism():
    """Arrays whose species maps do not match the compiled mechanism are refused (code 2)."""
    L = _load()
    m = load_bundled("h2_burke")
    arrs = legacy_arrays(m)
    arrs["map_r_h"] = arrs["map_r_h"].copy()
    arrs["map_r_h"][0] = (arrs["map_r_h"][0] + 1) % m.n_species
    cf = np.zeros(len(build_coef(m)[0]))
    assert L.cg_legacy_convert(*_ptrs(arrs), ctypes.c_void_p(cf.ctypes.data)) == 2
