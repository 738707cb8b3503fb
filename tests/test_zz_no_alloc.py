"""cg_getrates_batched_noalloc (no device allocation; tile list in the library's fixed device
array) against cg_getrates_batched through the C ABI, on two calls in a row, the second with a
smaller batch: bit-identical wdot for every kernel shape (the table kernel sums a species'
terms in a fixed order too), and for the table kernel both also within the parity bar of the
oracle (gross-normalised error <= 1e-12) on a sample of every patch.  Batches beyond the
array's limits are refused."""
import ctypes

import numpy as np
import pytest

from chemgen_amd.kernels import load_bundled
from chemgen_amd.synth import make_states
from oracle.getrates_np import getrates as oracle_getrates
from oracle.metrics import parity_stats

pytestmark = pytest.mark.gpu


def _call(L, dm, P, T, Y, W, offs, grid, noalloc):
    f = L.cg_getrates_batched_noalloc if noalloc else L.cg_getrates_batched
    o = np.ascontiguousarray(offs, dtype=np.int64)
    return f(dm.handle, P.data_ptr(), T.data_ptr(), Y.data_ptr(), W.data_ptr(), Y.stride(0),
             o.ctypes.data, len(o) - 1, grid, None)


@pytest.mark.parametrize("name,grid", [("synth_gri53", 0), ("synth_gri53", -1), ("synth_jetsurf348", 0)])
def test_noalloc_matches_batched(name, grid):
    import torch
    from chemgen_amd.autotune import best_variant
    from chemgen_amd.runtime import DeviceMechanism, lib
    L = lib()
    m = load_bundled(name)
    dm = DeviceMechanism(m, best_variant(m), build=False)
    d = torch.device("cuda:0")
    big = np.concatenate([[0], np.cumsum([700, 0, 1, 64, 3000, 129, 5])]).astype(np.int64)
    small = np.concatenate([[0], np.cumsum([90, 17, 0, 400])]).astype(np.int64)
    ng = int(big[-1])
    P, T, Y = make_states(m.n_species, ng, seed=31, T_range=(600.0, 2800.0))
    Pd, Td, Yd = (torch.from_numpy(x).to(d) for x in (P, T, Y))
    for offs in (big, small):
        print(f"[noalloc] {name} grid={grid} points={int(offs[-1])}", flush=True)
        ref = torch.full_like(Yd, -7.0)
        out = torch.full_like(Yd, -7.0)
        assert _call(L, dm, Pd, Td, Yd, ref, offs, grid, False) == 0, L.cg_last_error()
        assert _call(L, dm, Pd, Td, Yd, out, offs, grid, True) == 0, L.cg_last_error()
        torch.cuda.synchronize()
        assert torch.equal(out, ref)
        if dm.variant.table:
            pick = np.concatenate([np.arange(a, min(a + 16, b)) for a, b in zip(offs[:-1], offs[1:])])
            w0, qf, qb = oracle_getrates(m, P[pick], T[pick], Y[:, pick], return_parts=True)
            for w in (ref, out):
                st = parity_stats(m, w[:, torch.from_numpy(pick).to(d)].cpu().numpy(), w0, qf, qb)
                assert st["finite"] and st["max_gross"] <= 1e-12, st


def test_noalloc_refuses_beyond_limits():
    import torch
    from chemgen_amd.runtime import DeviceMechanism, lib
    L = lib()
    m = load_bundled("h2_burke")
    dm = DeviceMechanism(m, build=False)
    d = torch.device("cuda:0")
    Y = torch.zeros((m.n_species, 1000), dtype=torch.float64, device=d)
    P = torch.ones(1000, dtype=torch.float64, device=d)
    many = np.arange(0, 258, dtype=np.int64)            # 257 patches > 256
    assert _call(L, dm, P, P, Y, Y, many, 0, True) == -5
    assert b"patches" in L.cg_last_error()
    huge = np.array([0, (1 << 19) * dm.block + 1], dtype=np.int64)  # one tile too many
    Yh = torch.zeros((m.n_species, 1), dtype=torch.float64, device=d)
    o = np.ascontiguousarray(huge)
    rc = L.cg_getrates_batched_noalloc(dm.handle, P.data_ptr(), P.data_ptr(), Yh.data_ptr(), Yh.data_ptr(),
                                       int(huge[-1]), o.ctypes.data, 1, 0, None)
    assert rc == -5 and b"tiles" in L.cg_last_error()
