"""GPU parity of the QSS kernels (chemgen_amd/qssa.py) with the QSS oracle, through the C ABI."""
import numpy as np
import pytest

from chemgen_amd.synth import make_states
from oracle.qssa_np import qss_gross_scale, qss_getrates

pytestmark = pytest.mark.gpu
TOL_GROSS = 1e-12


@pytest.mark.parametrize("name", ["h2_burke", "gri30"])
@pytest.mark.parametrize("baked", [False, True])
def test_qss_kernel_vs_oracle(name, baked):
    import torch
    from chemgen_amd.codegen import Variant
    from chemgen_amd.qssa import bundled_qss
    from chemgen_amd.runtime import DeviceMechanism
    qm = bundled_qss(name)
    v = Variant(64, 1, "lean", "gh", "horner", False, 4, 1, 0, baked)
    dm = DeviceMechanism(qm, v, build=False)
    assert dm.n_species == qm.n_transported
    P, T, Y = make_states(qm.n_transported, 1000, seed=12, T_range=(800.0, 2500.0))
    print(f"[qss] {qm.name} {v.tag} ng={P.shape[0]}", flush=True)
    d = torch.device("cuda:0")
    Pd, Td, Yd = (torch.from_numpy(np.ascontiguousarray(x)).to(d) for x in (P, T, Y))
    w = dm.getrates(Pd, Td, Yd)
    torch.cuda.synchronize()
    w = w.cpu().numpy()
    ref, x, qf, qb, _ = qss_getrates(qm, qm.qss, P, T, Y, return_parts=True)
    G = qss_gross_scale(qm, qm.qss, qf, qb)
    assert np.isfinite(w).all()
    d = np.abs(w - ref)
    assert np.all(d[G == 0.0] == 0.0)                       # inert species
    assert np.max(d[G > 0] / (np.abs(ref) + G)[G > 0]) <= TOL_GROSS
    # the transported rates conserve elements only when the QSS balance holds
    a = qm.element_matrix()[:, qm.transported]
    assert np.max(np.abs(a @ w) / np.maximum(np.abs(a) @ G, 1e-300)) <= 1e-11
