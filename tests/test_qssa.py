"""QSS species (chemgen_amd/qssa.py; oracle/qssa_np.py): the oracle against a closed form,
the QSS balance, element conservation and the plain oracle; the linearity check; the
generated kernel's structure."""
import numpy as np
import pytest

from chemgen_amd.kernels import load_bundled
from chemgen_amd.mechanism import Mechanism
from chemgen_amd.qssa import pick_qss, qss_errors, with_qss
from chemgen_amd.synth import make_states
from oracle.getrates_np import PATM, R0, _arrh, _kb, _smh, getrates, gross_scale, r15
from oracle.qssa_np import qss_gross_scale, qss_getrates

from chemgen_amd.qssa import BUNDLED_QSS

GRI_QSS = list(BUNDLED_QSS["gri30"])


def _states(mech, qss, ng, seed):
    P, T, Y = make_states(mech.n_species - len(qss), ng, seed=seed, T_range=(800.0, 2500.0))
    return P, T, Y


def test_single_qss_species_closed_form():
    """O in H + O2 <=> O + OH, O + H2 <=> H + OH, H2 + OH <=> H2O + H: its balance gives
    C_O = (F0 + B1) / (B0 + F1) with F0 = kf0 C_H C_O2, B0 = kb0 C_OH, F1 = kf1 C_H2,
    B1 = kb1 C_H C_OH, and every transported rate follows from it."""
    h2 = load_bundled("h2_burke")
    m = Mechanism("h2_three", list(h2.elements), list(h2.species), [h2.reactions[j] for j in (0, 1, 3)])
    P, T, Y = _states(m, ["O"], 257, 3)
    w, x, _, _, resid = qss_getrates(m, ["O"], P, T, Y, return_parts=True)
    names = [s.name for s in m.species]
    R = [n for n in names if n != "O"]
    idx = {n: i for i, n in enumerate(names)}
    W = np.array([r15(m.species[idx[n]].mw) for n in R])
    yw = Y / W[:, None]
    C = {n: yw[i] * P / (R0 * T * yw.sum(axis=0)) for i, n in enumerate(R)}
    EG = np.array([np.exp(_smh(s.thermo, T, np.log(T))) for s in m.species])
    pfac = PATM / (R0 * T)
    k = []
    for r in m.reactions:
        kf = _arrh(*r.arrh, T)
        k.append((kf, _kb(r, kf, EG, idx, pfac)))
    F0, B0 = k[0][0] * C["H"] * C["O2"], k[0][1] * C["OH"]
    F1, B1 = k[1][0] * C["H2"], k[1][1] * C["H"] * C["OH"]
    xO = (F0 + B1) / (B0 + F1)
    np.testing.assert_allclose(x[0], xO, rtol=1e-13)
    q0, q1 = F0 - B0 * xO, F1 * xO - B1
    q2 = k[2][0] * C["H2"] * C["OH"] - k[2][1] * C["H2O"] * C["H"]
    exp = {"H": -q0 + q1 + q2, "O2": -q0, "OH": q0 + q1 - q2, "H2": -q1 - q2, "H2O": q2}
    for n, v in exp.items():
        g = np.abs(F0) + np.abs(B0 * xO) + np.abs(F1 * xO) + np.abs(B1) + np.abs(k[2][0] * C["H2"] * C["OH"]) \
            + np.abs(k[2][1] * C["H2O"] * C["H"])
        assert np.max(np.abs(w[R.index(n)] - v) / g) <= 1e-13, n
    assert np.max(np.abs(resid[0]) / (np.abs(F0) + np.abs(B1) + np.abs(B0 * xO) + np.abs(F1 * xO))) <= 1e-13


@pytest.mark.parametrize("name,qss", [("h2_burke", ["H2O2"]), ("gri30", GRI_QSS), ("synth_gri53", None)])
def test_qss_balance_and_element_conservation(name, qss):
    """The solved QSS concentrations zero each QSS species' net rate, and the transported
    rates then conserve every element (a wrong row, sign or index in the balance breaks
    both)."""
    m = load_bundled(name)
    if qss is None:
        qss = pick_qss(m, [s.name for s in m.species][::-1], 10)
        assert len(qss) >= 5
    P, T, Y = _states(m, qss, 300, 5)
    w, x, qf, qb, resid = qss_getrates(m, qss, P, T, Y, return_parts=True)
    assert np.isfinite(w).all() and np.isfinite(x).all()
    names = [s.name for s in m.species]
    gq = np.zeros_like(resid)
    for j, r in enumerate(m.reactions):
        for sp in list(r.reactants) + list(r.products):
            if sp in qss:
                gq[qss.index(sp)] += qf[j] + qb[j]
    assert np.max(np.abs(resid) / gq) <= 1e-11
    a = m.element_matrix()
    R = [k for k, n in enumerate(names) if n not in qss]
    G = qss_gross_scale(m, qss, qf, qb)
    bal = a[:, R] @ w
    scale = np.abs(a[:, R]) @ G
    assert np.max(np.abs(bal) / np.maximum(scale, 1e-300)) <= 1e-11


@pytest.mark.parametrize("name", ["h2_burke", "gri30"])
def test_empty_qss_set_is_the_plain_oracle(name):
    m = load_bundled(name)
    P, T, Y = make_states(m.n_species, 200, seed=2)
    w = qss_getrates(m, [], P, T, Y)
    ref, qf, qb = getrates(m, P, T, Y, return_parts=True)
    G = gross_scale(m, qf, qb)
    d = np.abs(w - ref)
    assert np.all(d[G == 0.0] == 0.0)                       # inert species
    assert np.max(d[G > 0] / (np.abs(ref) + G)[G > 0]) <= 1e-14


def test_linearity_check():
    h2 = load_bundled("h2_burke")
    assert qss_errors(h2, ["H2O2"]) == []
    assert any("coefficient 2.0" in e for e in qss_errors(h2, ["O"]))          # O + O + M <=> O2 + M
    assert any("two QSS species" in e for e in qss_errors(h2, ["H2O2", "O", ]))
    assert qss_errors(h2, ["XX"]) == ["XX: not a species"]
    with pytest.raises(ValueError):
        with_qss(h2, ["O"])
    g = load_bundled("gri30")
    assert pick_qss(g, GRI_QSS + ["CH2"], 16) == GRI_QSS                          # CH2 + CH2 excluded


def test_qss_kernel_source_structure():
    """The generated kernel holds Y/wdot rows for the transported species only, one
    elimination step per QSS unknown, and the two stages for every reaction involving a
    QSS species."""
    from chemgen_amd.codegen import Variant, emit_hip, structure_hash
    g = load_bundled("gri30")
    qm = with_qss(g, GRI_QSS)
    src, coef, meta = emit_hip(qm, Variant())
    assert meta["ns"] == g.n_species - len(GRI_QSS) and meta["n_qss"] == len(GRI_QSS)
    assert src.count("QSS stage 1") == src.count("QSS stage 2") == meta["qss_reactions"]
    assert all(f"const double x{i} =" in src for i in range(len(GRI_QSS)))
    assert f"W[{meta['ns'] - 1} * ld + i]" in src and f"W[{meta['ns']} * ld + i]" not in src
    assert structure_hash(qm) != structure_hash(g)


def test_generator_writes_a_qss_module(tmp_path):
    """generate_chemistry --qssa: the generated kernel source and the saved metadata carry the
    QSS set, and the module's species list is the transported one."""
    import json
    import os
    from chemgen_amd import generate_chemistry
    from chemgen_amd.kernels import DATA_DIR
    out = tmp_path / "h2q.py"
    args = generate_chemistry.parse_arguments(["--mode", "getrates", "--language", "hip", "--mech",
                                               os.path.join(DATA_DIR, "h2_burke.inp"), "--output", str(out),
                                               "--qssa", "H2O2"])
    mech = generate_chemistry._load(args)
    assert tuple(mech.qss) == ("H2O2",) and mech.n_transported == 8
    from chemgen_amd.api import save_generated
    from chemgen_amd.codegen import emit_hip
    v = generate_chemistry.variant_for(args, mech)
    src, _, meta = emit_hip(mech, v)
    assert meta["ns"] == 8 and "quasi-steady-state" in src
    stem = str(tmp_path / "h2q")
    save_generated(stem, mech, v, stem + ".hsaco")
    assert json.load(open(stem + ".json"))["qss"] == ["H2O2"]
