"""The 40-digit arbiter (oracle/arbiter_mp.py) and the FP64 oracle against it.

* The arbiter passes the closed-form pins of tests/test_oracle_pins.py (Y -> C, Arrhenius,
  third-body M, Troe / SRI / Lindemann broadening, PLOG), evaluated through it instead of the
  FP64 oracle: it is pinned by the same facts, not by agreement with the oracle.
* Its float64 results do not move between 40 and 60 digits.
* The FP64 oracle is within rounding of it (gross-normalised error <= 1e-13) on every
  mechanism family: elementary/third-body/Troe with T2/Lindemann (GRI30), SRI (JetSurf-sized),
  PLOG, extended records with coefficients 0.5 and sides of five slots, negative prefactors.
"""
import numpy as np
import pytest

import tests.test_oracle_pins as pins
from chemgen_amd.kernels import load_bundled
from chemgen_amd.synth import make_states
from oracle.arbiter_mp import arbiter_stats, getrates_mp
from oracle.getrates_np import getrates


def _mp_getrates(mech, P, T, Y, return_parts=False):
    assert not return_parts
    return getrates_mp(mech, P, T, Y)[0]


def test_arbiter_passes_closed_form_pins(monkeypatch):
    monkeypatch.setattr(pins, "getrates", _mp_getrates)
    pins.test_ytoc_ideal_gas()
    pins.test_arrhenius_temperature_law()
    pins.test_third_body_efficiencies_closed_form()
    pins.test_troe_center_closed_form()
    pins.test_lindemann_limits()
    pins.test_sri_at_unit_reduced_pressure()
    pins.test_plog_interpolation_closed_form()
    pins._falloff_pins()


def test_arbiter_precision_converged():
    m = load_bundled("gri30")
    P, T, Y = make_states(m.n_species, 4, seed=11)
    w40, _ = getrates_mp(m, P, T, Y, dps=40)
    w60, _ = getrates_mp(m, P, T, Y, dps=60)
    nz = w60 != 0.0
    assert np.all(np.abs(w40[nz] - w60[nz]) <= 2.3e-16 * np.abs(w60[nz]))


@pytest.mark.parametrize("name,n", [("h2_burke", 64), ("h2_burke2012", 32), ("gri30", 12),
                                    ("synth_gri53", 12), ("synth_plog53", 12), ("synth_ext53", 8),
                                    ("hashemi2016", 3), ("synth_jetsurf348", 2)])
def test_oracle_within_rounding_of_arbiter(name, n):
    m = load_bundled(name)
    P, T, Y = make_states(m.n_species, n, seed=5)
    wm, G = getrates_mp(m, P, T, Y)
    st = arbiter_stats(getrates(m, P, T, Y), wm, G)
    assert st["max_gross"] <= 1e-13, st
