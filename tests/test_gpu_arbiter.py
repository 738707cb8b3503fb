"""GPU getrates and the FP64 oracle, both graded against the 40-digit arbiter.

The gate is the gross-normalised error |dw| / (|w| + sum_j (nu'+nu'')(|q_f| + |q_b|)) against
the exact value of the reference formula on the same float64 inputs (oracle/arbiter_mp.py),
<= 1e-12.  The raw relative error |dw| / |w| of both FP64 paths is printed beside it: for a
species near steady state it is set by the cancellation between the reaction terms, and
the arbiter shows whether the GPU or the oracle is closer to the exact value.
"""
import numpy as np
import pytest

from chemgen_amd.kernels import load_bundled
from chemgen_amd.synth import make_states
from oracle.arbiter_mp import arbiter_stats, getrates_mp
from oracle.getrates_np import getrates as oracle_getrates

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,n", [("h2_burke", 64), ("gri30", 32), ("synth_gri53", 32),
                                    ("synth_plog53", 32), ("synth_ext53", 16), ("synth_jetsurf348", 8)])
def test_gpu_against_arbiter(name, n):
    import torch
    from chemgen_amd.runtime import DeviceMechanism
    m = load_bundled(name)
    print(f"[arbiter] {name} n={n}", flush=True)
    P, T, Y = make_states(m.n_species, n, seed=21)
    dm = DeviceMechanism(m, build=False)
    d = torch.device("cuda:0")
    w = dm.getrates(*(torch.from_numpy(x).to(d) for x in (P, T, Y)))
    torch.cuda.synchronize()
    w = w.cpu().numpy()
    wm, G = getrates_mp(m, P, T, Y)
    gpu = arbiter_stats(w, wm, G)
    cpu = arbiter_stats(oracle_getrates(m, P, T, Y), wm, G)
    print(f"{name} gpu {gpu} oracle {cpu}", flush=True)
    assert np.isfinite(w).all()
    assert gpu["max_gross"] <= 1e-12, gpu
