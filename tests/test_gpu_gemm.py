"""GPU parity of the GEMM formulation (cg_gemm_create / cg_getrates_gemm: FP64 DGEMMs on the
matrix cores around two element-wise kernels; the reference's parallel level 4) against the
CPU oracle, through the C ABI."""
import numpy as np
import pytest

from chemgen_amd.kernels import load_bundled
from chemgen_amd.synth import make_states
from oracle.getrates_np import getrates as oracle_getrates
from oracle.metrics import parity_stats

pytestmark = pytest.mark.gpu

TOL_GROSS = 1e-12


@pytest.mark.parametrize("name,ng,slab", [
    ("h2_burke", 4096, 1 << 16),
    ("gri30", 2048, 1 << 16),
    ("synth_gri53", 1000, 384),          # three slabs, the last one partial
    ("synth_jetsurf348", 512, 1 << 16),
    ("synth_plog53", 1024, 1 << 16),
    ("synth_ext53", 1024, 1 << 16),
    ("hashemi2016", 1024, 1 << 16),
])
def test_gemm_formulation_vs_oracle(name, ng, slab):
    import torch
    from chemgen_amd.gemm import GemmGetRates
    m = load_bundled(name)
    P, T, Y = make_states(m.n_species, ng, seed=12)
    print(f"[gemm] {name} ng={ng} slab={slab}", flush=True)
    g = GemmGetRates(m, slab=slab)
    d = torch.device("cuda:0")
    Pd, Td, Yd = (torch.from_numpy(x).to(d) for x in (P, T, Y))
    w = g.getrates(Pd, Td, Yd)
    torch.cuda.synchronize()
    ref, qf, qb = oracle_getrates(m, P, T, Y, return_parts=True)
    st = parity_stats(m, w.cpu().numpy(), ref, qf, qb)
    assert st["finite"] and st["max_gross"] <= TOL_GROSS, st
    g.close()


def test_gemm_strided_rows_and_empty_call():
    """A leading dimension larger than ng (wdot written only in the first ng columns of each
    row) and ng = 0."""
    import torch
    from chemgen_amd.gemm import GemmGetRates
    m = load_bundled("gri30")
    ng, ld = 700, 1024
    P, T, Y = make_states(m.n_species, ng, seed=13)
    d = torch.device("cuda:0")
    Yb = torch.zeros((m.n_species, ld), dtype=torch.float64, device=d)
    Yb[:, :ng] = torch.from_numpy(Y).to(d)
    Wb = torch.full((m.n_species, ld), 7.0, dtype=torch.float64, device=d)
    g = GemmGetRates(m, slab=512)
    Pd, Td = torch.from_numpy(P).to(d), torch.from_numpy(T).to(d)
    g.getrates_(Pd, Td, Yb[:, :ng], Wb[:, :ng])
    g.getrates_(Pd[:0], Td[:0], Yb[:, :0], Wb[:, :0])
    torch.cuda.synchronize()
    W = Wb.cpu().numpy()
    assert (W[:, ng:] == 7.0).all()
    ref, qf, qb = oracle_getrates(m, P, T, Y, return_parts=True)
    st = parity_stats(m, W[:, :ng], ref, qf, qb)
    assert st["max_gross"] <= TOL_GROSS, st
    g.close()
