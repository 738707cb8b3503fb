"""The tuned kernels at the benchmark sizes (2^20 points GRI-sized, a 2^22-point ragged batch of
the n-heptane-sized mechanism, 2^19 points JetSurf-sized): oracle parity on a random sample of
each full-size run, plus size-independent properties -- a repeated launch is bit-identical
(every kernel shape sums in a fixed order, the table kernel included), and the ragged batch
equals the contiguous launch of the same kernel bit for bit."""
import numpy as np
import pytest

from chemgen_amd.kernels import load_bundled
from oracle.getrates_np import getrates as oracle_getrates
from oracle.metrics import parity_stats

pytestmark = pytest.mark.gpu
TOL_GROSS = 1e-12


def _sample_parity(m, P, T, Y, W, n, seed):
    import torch
    idx = np.sort(np.random.default_rng(seed).choice(P.shape[0], n, replace=False))
    it = torch.from_numpy(idx).to(P.device)
    p, t = P[it].cpu().numpy(), T[it].cpu().numpy()
    y, w = Y[:, it].cpu().numpy(), W[:, it].cpu().numpy()
    ref, qf, qb = oracle_getrates(m, p, t, y, return_parts=True)
    return parity_stats(m, w, ref, qf, qb)


def test_gri_sized_full_size():
    import torch
    from chemgen_amd.runtime import DeviceMechanism
    from chemgen_amd.surrogate import random_states
    m = load_bundled("synth_gri53")
    dm = DeviceMechanism(m, build=False)
    print(f"[fullsize] synth_gri53 {dm.variant.tag} 2^20", flush=True)
    P, T, Y = random_states(m.n_species, 1 << 20, seed=21, T_range=(300.0, 2500.0))
    W1, W2 = torch.empty_like(Y), torch.empty_like(Y)
    dm.getrates_(P, T, Y, W1)
    dm.getrates_(P, T, Y, W2)
    torch.cuda.synchronize()
    assert torch.equal(W1, W2)
    st = _sample_parity(m, P, T, Y, W1, 2048, 1)
    assert st["finite"] and st["max_gross"] <= TOL_GROSS, st


def test_ragged_batch_full_size_equals_contiguous():
    import torch
    from chemgen_amd.autotune import best_batched
    from chemgen_amd.runtime import DeviceMechanism
    from chemgen_amd.surrogate import random_states
    from chemgen_amd.synth import ragged_patches
    m = load_bundled("synth_nhept88")
    v, grid = best_batched(m)
    dm = DeviceMechanism(m, v, build=False)
    offs = ragged_patches()
    ng = int(offs[-1])
    print(f"[fullsize] synth_nhept88 {v.tag} grid={grid} {ng} points in {len(offs) - 1} patches", flush=True)
    P, T, Y = random_states(m.n_species, ng, seed=22, T_range=(600.0, 2800.0))
    Wb, Wc = torch.empty_like(Y), torch.empty_like(Y)
    dm.getrates_batched_(P, T, Y, Wb, offs, grid=grid)
    dm.getrates_(P, T, Y, Wc)
    torch.cuda.synchronize()
    assert torch.equal(Wb, Wc)
    st = _sample_parity(m, P, T, Y, Wb, 1024, 2)
    assert st["finite"] and st["max_gross"] <= TOL_GROSS, st


def test_jetsurf_sized_full_size():
    import torch
    from chemgen_amd.runtime import DeviceMechanism
    from chemgen_amd.surrogate import random_states
    m = load_bundled("synth_jetsurf348")
    dm = DeviceMechanism(m, build=False)
    print(f"[fullsize] synth_jetsurf348 {dm.variant.tag} 2^19", flush=True)
    P, T, Y = random_states(m.n_species, 1 << 19, seed=23, T_range=(600.0, 2800.0))
    W, W2 = torch.empty_like(Y), torch.empty_like(Y)
    dm.getrates_(P, T, Y, W)
    dm.getrates_(P, T, Y, W2)
    torch.cuda.synchronize()
    assert torch.equal(W, W2)          # the table kernel's fixed-order species sums
    st = _sample_parity(m, P, T, Y, W, 256, 3)
    assert st["finite"] and st["max_gross"] <= TOL_GROSS, st
