"""The reference's entry points (copyConstantsToDevice + getrates_hip_wrapper_v1..v4) driving
the generated kernel on the GPU, with the reference's Fortran layout Y(ng, Ns)."""
import ctypes

import numpy as np
import pytest

from chemgen_amd.kernels import load_bundled
from chemgen_amd.synth import make_states
from oracle.getrates_np import getrates as oracle_getrates
from oracle.legacy_coef import legacy_arrays
from oracle.metrics import parity_stats
from tests.test_legacy_abi import _load, _ptrs

pytestmark = pytest.mark.gpu


def test_legacy_entry_points_match_oracle():
    import torch
    L = _load()
    m = load_bundled("h2_burke")
    arrs = legacy_arrays(m)
    L.copyConstantsToDevice(*_ptrs(arrs))
    ng = 1000   # not a multiple of any block size: the reference would drop the tail
    P, T, Y = make_states(m.n_species, ng, seed=17)
    ref, qf, qb = oracle_getrates(m, P, T, Y, return_parts=True)
    d = torch.device("cuda:0")
    Pd, Td, Yd = (torch.from_numpy(np.ascontiguousarray(x)).to(d) for x in (P, T, Y))
    for v in (1, 2, 3, 4):
        print(f"[legacy] getrates_hip_wrapper_v{v}", flush=True)
        W = torch.full_like(Yd, float("nan"))
        f = getattr(L, f"getrates_hip_wrapper_v{v}")
        f(ctypes.c_void_p(Pd.data_ptr()), ctypes.c_void_p(Td.data_ptr()), ctypes.c_int(ng),
          ctypes.c_void_p(Yd.data_ptr()), ctypes.c_void_p(W.data_ptr()), None, None)
        st = parity_stats(m, W.cpu().numpy(), ref, qf, qb)
        assert st["finite"] and st["max_gross"] <= 1e-12, (v, st)


@pytest.mark.parametrize("level", [1, 2, 3, 4])
def test_legacy_entry_points_each_level(level):
    """chemgen_m over the compatibility library of each parallel level: copyConstantsToDevice
    with the arrays of this build's coefficient module for that level (each reaction type padded
    to 8 with the reference's dummy reactions at levels 2-4; MAX_SP = 4 at level 3; the dense
    rocBLAS matrices and level 4's own argument list), then getrates_hip_wrapper_v1..v4 against
    the oracle."""
    import torch
    from tests.test_legacy_levels import level_args, load_level_library
    L = load_level_library(level)
    m = load_bundled("h2_burke")
    args = level_args(m, level)
    L.copyConstantsToDevice(*[ctypes.c_void_p(a.ctypes.data) for a in args])
    ng = 777
    P, T, Y = make_states(m.n_species, ng, seed=23)
    ref, qf, qb = oracle_getrates(m, P, T, Y, return_parts=True)
    d = torch.device("cuda:0")
    Pd, Td, Yd = (torch.from_numpy(np.ascontiguousarray(x)).to(d) for x in (P, T, Y))
    for v in (1, 2, 3, 4):
        print(f"[legacy] level {level} getrates_hip_wrapper_v{v}", flush=True)
        W = torch.full_like(Yd, float("nan"))
        getattr(L, f"getrates_hip_wrapper_v{v}")(
            ctypes.c_void_p(Pd.data_ptr()), ctypes.c_void_p(Td.data_ptr()), ctypes.c_int(ng),
            ctypes.c_void_p(Yd.data_ptr()), ctypes.c_void_p(W.data_ptr()), None, None)
        st = parity_stats(m, W.cpu().numpy(), ref, qf, qb)
        assert st["finite"] and st["max_gross"] <= 1e-12, (level, v, st)
