"""Mechanism front-end: Chemkin/YAML readers, reference semantics, synthetic mechanisms."""
import os

import numpy as np
import pytest

from chemgen_amd.kernels import load_bundled
from chemgen_amd.mechanism import (ATOMIC_WEIGHTS, load_yaml, parse_chemkin, parse_equation,
                                   write_chemkin)
from chemgen_amd.synth import SPECS, make_states

REF = "/root/reference/ck_files"
needs_ref = pytest.mark.skipif(not os.path.isdir(REF), reason="reference mechanism files absent")


def test_equation_parser_forms():
    sp = {"H", "O2", "HO2", "O", "OH", "CH2(S)", "AR", "H2O"}
    r, p, rev, m, coll = parse_equation("H+O2(+M)<=>HO2(+M)", sp)
    assert (r, p, rev, m, coll) == ({"H": 1.0, "O2": 1.0}, {"HO2": 1.0}, True, False, "M")
    r, p, rev, m, coll = parse_equation("2O+M=O2+M", sp)
    assert r == {"O": 2.0} and p == {"O2": 1.0} and m and coll is None
    r, p, rev, m, coll = parse_equation("O+CH2(S)=>H+OH", sp)
    assert r == {"O": 1.0, "CH2(S)": 1.0} and not rev and not m
    r, p, rev, m, coll = parse_equation("H+O2(+AR)<=>HO2(+AR)", sp)
    assert coll == "AR"
    r, p, rev, m, coll = parse_equation("H2O+H2O=H+OH+H2O", sp | {"H2O"})
    assert r == {"H2O": 2.0} and p == {"H": 1.0, "OH": 1.0, "H2O": 1.0}
    # coefficients as Cantera's converters write them (ck_files/CH4_Methane/TwoStep/chem.yaml)
    r, p, rev, m, coll = parse_equation("CO + 5.00E-01 O2 <=> CO2", sp | {"CO", "CO2"})
    assert r == {"CO": 1.0, "O2": 0.5} and p == {"CO2": 1.0} and rev
    r, p, rev, m, coll = parse_equation("2.5e+00 O2 + 0.5 H => 1.5E1 O", sp)
    assert r == {"O2": 2.5, "H": 0.5} and p == {"O": 15.0} and not rev
    # a species whose name reads like an exponent stays a species
    r, p, rev, m, coll = parse_equation("2E2 + O => OH", sp | {"E2"})
    assert r == {"E2": 2.0, "O": 1.0}


def test_molecular_weights_cantera_table():
    # Cantera's element weights: H2O = 2*1.008 + 15.999
    m = load_bundled("h2_burke")
    w = {s.name: s.mw for s in m.species}
    assert w["H2O"] == pytest.approx(2 * 1.008 + 15.999, rel=0, abs=1e-12)
    assert w["N2"] == pytest.approx(2 * 14.007, abs=1e-12)
    assert ATOMIC_WEIGHTS["Ar"] == 39.95


def test_bundled_counts():
    c = {n: load_bundled(n) for n in ("h2_burke", "h2_burke2012", "gri30", "synth_gri53", "synth_nhept88",
                                     "synth_jetsurf348")}
    assert (c["h2_burke"].n_species, c["h2_burke"].n_reactions) == (9, 23)
    assert (c["h2_burke2012"].n_species, c["h2_burke2012"].n_reactions) == (13, 27)
    assert (c["gri30"].n_species, c["gri30"].n_reactions) == (53, 325)
    g = c["synth_gri53"]
    assert (g.n_species, g.n_reactions) == (53, 325)
    assert g.counts()["troe"] + g.counts()["lindemann"] == 29
    assert g.counts()["third_body"] == 40
    assert (c["synth_nhept88"].n_species, c["synth_nhept88"].n_reactions) == (88, 387)
    j = c["synth_jetsurf348"]
    assert (j.n_species, j.n_reactions) == (348, 2163)
    assert j.counts()["troe"] + j.counts()["sri"] == 173
    f = load_bundled("fuller2021")
    assert (f.n_species, f.n_reactions) == (835, 4289)
    assert f.counts() == {"standard": 3658, "third_body": 35, "troe": 63, "lindemann": 10, "plog": 523}
    e = load_bundled("synth_ext53")
    assert (e.n_species, e.n_reactions) == (53, 132)


@pytest.mark.parametrize("name", ["synth_gri53", "synth_nhept88", "synth_jetsurf348"])
def test_synthetic_element_balance(name):
    m = load_bundled(name)
    a = m.element_matrix()
    idx = {s.name: i for i, s in enumerate(m.species)}
    for r in m.reactions:
        nu = np.zeros(m.n_species)
        for s, n in r.reactants.items():
            nu[idx[s]] -= n
        for s, n in r.products.items():
            nu[idx[s]] += n
        assert np.all(a @ nu == 0), r.equation
    for r in m.reactions:
        for e in r.third_body.values():
            assert -0.5 <= e <= 11.0


def test_chemkin_round_trip(tmp_path):
    m = load_bundled("synth_gri53")
    p = tmp_path / "m.inp"
    write_chemkin(m, str(p))
    m2 = parse_chemkin(str(p))
    assert m2.species_names == m.species_names
    for a, b in zip(m.reactions, m2.reactions):
        assert a.reactants == b.reactants and a.products == b.products and a.rtype == b.rtype
        assert a.arrh == b.arrh and a.low == b.low and a.troe == b.troe and a.third_body == b.third_body


def test_states_recipe():
    P, T, Y = make_states(13, 1000, seed=0)
    assert Y.shape == (13, 1000)
    assert np.allclose(Y.sum(0), 1.0, atol=1e-15)
    assert T.min() >= 300 and T.max() <= 2500
    assert P.min() >= 1013250.0 * (1 - 1e-15) and P.max() <= 40 * 1013250.0 * (1 + 1e-15)
    P2, T2, Y2 = make_states(13, 1000, seed=0)
    assert np.array_equal(Y, Y2) and np.array_equal(T, T2)


@needs_ref
@pytest.mark.parametrize("inp,therm,yml", [
    ("H2_burke/chem.inp", "H2_burke/therm.dat", "H2_burke/chem.yaml"),
    ("GRI30/chem.inp", "GRI30/therm.dat", "GRI30/chem.yaml"),
])
def test_chemkin_matches_yaml(inp, therm, yml):
    """The two readers agree on the reference's own Chemkin/YAML pairs."""
    a = parse_chemkin(os.path.join(REF, inp), os.path.join(REF, therm))
    b = load_yaml(os.path.join(REF, yml))
    assert a.species_names == b.species_names
    for sa, sb in zip(a.species, b.species):
        assert sa.mw == pytest.approx(sb.mw, rel=1e-15)
        assert sa.thermo.low == sb.thermo.low and sa.thermo.high == sb.thermo.high
    assert len(a.reactions) == len(b.reactions)
    for ra, rb in zip(a.reactions, b.reactions):
        assert (ra.reactants, ra.products, ra.rtype, ra.reversible) == (rb.reactants, rb.products, rb.rtype,
                                                                       rb.reversible)
        assert np.allclose(ra.arrh, rb.arrh, rtol=1e-14, atol=0)
        assert ra.third_body == pytest.approx(rb.third_body)
        assert ra.troe == rb.troe


@needs_ref
def test_reference_compat_drops_types():
    """compat='reference' keeps only the reference's types (chemgen/parser.py:44-49,146-147)."""
    full = load_yaml(os.path.join(REF, "GRI30/chem.yaml"))
    ref = load_yaml(os.path.join(REF, "GRI30/chem.yaml"), compat="reference")
    assert full.counts().get("lindemann") == 3
    assert "lindemann" not in ref.counts()
    assert ref.n_reactions == full.n_reactions - 3


def test_yaml_no_is_a_species(tmp_path):
    """PyYAML would read species NO as False; the loader keeps it a string."""
    doc = """
units: {length: cm, time: s, quantity: mol, activation-energy: cal/mol}
phases:
- name: gas
  elements: [O, N]
  species: [N, NO, O]
species:
- name: N
  composition: {N: 1}
  thermo: {model: NASA7, temperature-ranges: [300.0, 5000.0], data: [[2.5, 0, 0, 0, 0, 56104.6, 4.19]]}
- name: NO
  composition: {N: 1, O: 1}
  thermo: {model: NASA7, temperature-ranges: [300.0, 1000.0, 5000.0],
           data: [[4.2, -4.6e-3, 1.1e-5, -9.3e-9, 2.8e-12, 9844.6, 2.28], [3.26, 1.19e-3, -4.3e-7, 6.9e-11, -4.0e-15, 9920.97, 6.37]]}
- name: O
  composition: {O: 1}
  thermo: {model: NASA7, temperature-ranges: [300.0, 5000.0], data: [[2.5, 0, 0, 0, 0, 29230.0, 4.92]]}
reactions:
- equation: N + NO <=> N2 + O
  rate-constant: {A: 2.7e+13, b: 0.0, Ea: 355.0}
"""
    p = tmp_path / "m.yaml"
    p.write_text(doc.replace("N + NO <=> N2 + O", "N + NO <=> NO + N"))
    m = load_yaml(str(p))
    assert m.species_names == ["N", "NO", "O"]
    assert m.species[0].thermo.t_mid == 5000.0  # single range expanded to two identical ranges
