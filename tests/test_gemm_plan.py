"""The GEMM formulation's plan (dense stoichiometric and efficiency matrices, per-reaction
records into the gpu_coef blob; the reference's parallel level 4, hip_v4_cpp.j2:1-19) checked
on the host: evaluated in NumPy by oracle/gemm_np.py it must reproduce the oracle's wdot on
every kind of rate the generator reads, and its arrays must hold what the mechanism says."""
import numpy as np
import pytest

from chemgen_amd.codegen import build_coef
from chemgen_amd.gemm import gemm_plan
from chemgen_amd.kernels import load_bundled
from chemgen_amd.synth import make_states
from oracle.gemm_np import gemm_formulation_np
from oracle.getrates_np import getrates as oracle_getrates
from oracle.metrics import parity_stats

TOL_GROSS = 1e-12


@pytest.mark.parametrize("name", ["h2_burke", "h2_burke2012", "gri30", "synth_gri53", "synth_jetsurf348",
                                  "synth_plog53", "synth_ext53", "hashemi2016"])
def test_plan_reproduces_oracle(name):
    """Third-body, Lindemann, Troe with and without T2, SRI (JetSurf-sized), PLOG (synthetic and
    Hashemi 2016 with negative prefactors), non-integer coefficients (synth_ext53)."""
    m = load_bundled(name)
    P, T, Y = make_states(m.n_species, 200, seed=11)
    ref, qf, qb = oracle_getrates(m, P, T, Y, return_parts=True)
    st = parity_stats(m, gemm_formulation_np(m, P, T, Y), ref, qf, qb)
    assert st["finite"] and st["max_gross"] <= TOL_GROSS, st


def test_plan_arrays_gri30():
    """GRI-Mech 3.0: the matrices and records hold the mechanism's numbers."""
    m = load_bundled("gri30")
    pl = gemm_plan(m)
    idx = {s.name: i for i, s in enumerate(m.species)}
    coef, _, roffs = build_coef(m)
    assert pl.nu_r.shape == pl.nu_p.shape == (m.n_reactions, m.n_species)
    n_tb = sum(r.rtype in ("third_body", "lindemann", "troe", "sri") for r in m.reactions)
    assert pl.eff.shape == (n_tb, m.n_species)
    kinds = {"standard": 0, "third_body": 1, "lindemann": 2}
    for j, r in enumerate(m.reactions):
        for s, n in r.reactants.items():
            assert pl.nu_r[j, idx[s]] == n
        for s, n in r.products.items():
            assert pl.nu_p[j, idx[s]] == n
        assert pl.nu_r[j].sum() == sum(r.reactants.values())
        assert pl.dnu[j] == pytest.approx(sum(r.products.values()) - sum(r.reactants.values()))
        assert pl.rec[j, 1] == roffs[j] and pl.rec[j, 3] == int(r.reversible)
        if r.rtype in kinds:
            assert pl.rec[j, 0] == kinds[r.rtype]
        if r.rtype == "troe":
            assert pl.rec[j, 0] == (4 if len(r.troe) == 4 else 3)
        tb = pl.rec[j, 2]
        if r.rtype == "standard":
            assert tb == -1
        else:
            # efficiency 1 for unlisted species, the listed efficiency otherwise
            want = np.ones(m.n_species)
            for s, e in r.third_body.items():
                want[idx[s]] = 1.0 + e
            np.testing.assert_allclose(pl.eff[tb], want, rtol=1e-15)
    # 2 O + M <=> O2 + M (GRI-Mech 3.0 reaction 1): H2 2.4, H2O 15.4, AR 0.83
    j = next(j for j, r in enumerate(m.reactions) if r.rtype == "third_body" and r.reactants == {"O": 2.0})
    row = pl.eff[pl.rec[j, 2]]
    assert (row[idx["H2"]], row[idx["H2O"]], row[idx["AR"]]) == pytest.approx((2.4, 15.4, 0.83))
    assert coef[pl.rec[j, 1]] == pytest.approx(np.log2(1.2e17))
