"""Pins of the CPU oracle by facts that do not come from its own formulas.

* NASA-7 smh at 298.15 K against CODATA key values (S and formation enthalpy), and its
  temperature derivative against the NASA enthalpy polynomial (d(-G/RT)/dT = H/RT^2).
* Y -> C against the ideal-gas law (sum C = P/(R0 T), mole fractions).
* Arrhenius temperature dependence with Ea in cal/mol.
* Detailed balance: with concentrations at equilibrium, built from Kp = exp(-dG/RT) with
  H and S evaluated separately, the net rate vanishes, for dnu = -1, 0, +1, +0.5 (a
  coefficient 0.5) and a side of four species with one species on both sides.
* Third-body M: all efficiencies one -> M = P/(R0 T); an efficiency zero removes C_k.
* Duplicates with a negative prefactor add up like one reaction with the summed prefactor.
* Troe/Lindemann/SRI special cases with closed forms: Fcent = a, f1 = 0 -> F = a; f1 = +-1 ->
  F = Fcent^(1/2); n = 0 and the |log10 Pr| -> inf limit -> F = Fcent^(1/(1+(1/0.14)^2)); the T2
  term; Lindemann F = 1; SRI X = 1, 1/2, 1/5, 1/10 at Pr = 1, 10^+-1, 10^+-2, 10^3.
* Pressure scaling q ~ P^order at fixed T and composition.
* Element and mass conservation for every bundled mechanism.
* Planted mistakes in the oracle make the pins fail.
"""
import math

import numpy as np
import pytest

import oracle.getrates_np as og
from chemgen_amd.kernels import load_bundled
from chemgen_amd.mechanism import Mechanism, Reaction, Species
from chemgen_amd.synth import make_states
from oracle.getrates_np import PATM, R0, RC, getrates, gross_scale

R_J = 8.314462618  # J/mol/K


def _sp(name):
    return next(s for s in load_bundled("h2_burke").species if s.name == name)


def _mech(species_names, reactions):
    return Mechanism("t", ["H", "O", "N"], [_sp(n) for n in species_names], reactions)


def _state(mech, C, T):
    """Y, P for given concentrations C[k] (mol/cm^3) at T (ideal gas)."""
    W = mech.mw_array()
    C = np.asarray(C, dtype=np.float64)
    rho = np.sum(C * W)
    Y = (C * W / rho)[:, None]
    P = np.array([np.sum(C) * R0 * T])
    return P, np.array([T]), Y


def _h_rt(a, T):
    return a[0] + a[1] * T / 2 + a[2] * T ** 2 / 3 + a[3] * T ** 3 / 4 + a[4] * T ** 4 / 5 + a[5] / T


def _s_r(a, T):
    return a[0] * math.log(T) + a[1] * T + a[2] * T ** 2 / 2 + a[3] * T ** 3 / 3 + a[4] * T ** 4 / 4 + a[6]


def _coefs(th, T):
    return th.low if T < th.t_mid else th.high


# CODATA key values at 298.15 K: S [J/mol/K], formation enthalpy [kJ/mol]
CODATA = {"H2": (130.680, 0.0), "O2": (205.152, 0.0), "N2": (191.609, 0.0), "H2O": (188.835, -241.826)}


@pytest.mark.parametrize("name", list(CODATA))
def test_smh_codata_298(name):
    th = _sp(name).thermo
    T = np.array([298.15])
    smh = og._smh(th, T, np.log(T))[0]
    S, H = CODATA[name]
    expect = S / R_J - H * 1e3 / (R_J * 298.15)
    assert abs(smh - expect) < 0.02, (smh, expect)  # NASA fit error at 298 K is ~1e-3 relative


@pytest.mark.parametrize("name", ["H", "H2", "O", "OH", "H2O", "O2", "HO2", "H2O2", "N2"])
@pytest.mark.parametrize("T0", [450.0, 800.0, 1500.0, 2600.0])
def test_smh_derivative_is_enthalpy(name, T0):
    """d(smh)/dT = H/(R T^2), with H from the NASA enthalpy polynomial."""
    th = _sp(name).thermo
    h = 1e-3 * T0
    T = np.array([T0 - 2 * h, T0 - h, T0 + h, T0 + 2 * h])
    s = og._smh(th, T, np.log(T))
    d = (s[0] - 8 * s[1] + 8 * s[2] - s[3]) / (12 * h)  # fourth-order central difference
    expect = _h_rt(_coefs(th, T0), T0) / T0
    assert d == pytest.approx(expect, rel=1e-6, abs=1e-9)


def _irrev(reac, prod, arrh, rtype="standard", **kw):
    r = Reaction("", reac, prod, rtype, False, arrh, **kw)
    return r


def test_ytoc_ideal_gas():
    m = _mech(["H", "O2", "HO2", "N2"], [_irrev({"H": 1.0, "O2": 1.0}, {"HO2": 1.0}, (1.0, 0.0, 0.0))])
    P, T, Y = make_states(4, 64, seed=3)
    w = getrates(m, P, T, Y)
    W = m.mw_array()
    n = Y / W[:, None]
    X = n / n.sum(0)
    ctot = P / (R0 * T)
    assert np.allclose(-w[0], X[0] * X[1] * ctot ** 2, rtol=1e-14, atol=0)
    assert np.allclose(w[2], X[0] * X[1] * ctot ** 2, rtol=1e-14, atol=0)


def test_arrhenius_temperature_law():
    A, b, Ea = 3.0e12, 1.5, 10000.0
    m = _mech(["H", "O2", "HO2", "N2"], [_irrev({"H": 1.0, "O2": 1.0}, {"HO2": 1.0}, (A, b, Ea))])
    C = np.array([1e-6, 2e-6, 0.0, 5e-6])
    q = []
    for T in (700.0, 1900.0):
        P, Tt, Y = _state(m, C, T)
        q.append(-getrates(m, P, Tt, Y)[0, 0])
    ratio = (700.0 / 1900.0) ** b * math.exp(-Ea / RC * (1 / 700.0 - 1 / 1900.0))
    assert q[0] / q[1] == pytest.approx(ratio, rel=1e-13)
    assert q[1] == pytest.approx(A * 1900.0 ** b * math.exp(-Ea / (RC * 1900.0)) * 1e-6 * 2e-6, rel=1e-13)


def _kc(species, nu, T):
    """Kc from textbook thermodynamics: Kp = exp(-dG/RT), dG/RT = dH/RT - dS/R."""
    dg = 0.0
    for name, n in nu.items():
        a = _coefs(_sp(name).thermo, T)
        dg += n * (_h_rt(a, T) - _s_r(a, T))
    dnu = sum(nu.values())
    return math.exp(-dg) * (PATM / (R0 * T)) ** dnu


@pytest.mark.parametrize("reac,prod,rtype", [
    ({"H": 1.0, "O2": 1.0}, {"HO2": 1.0}, "standard"),          # dnu = -1
    ({"H": 1.0, "O2": 1.0}, {"O": 1.0, "OH": 1.0}, "standard"),  # dnu = 0
    ({"H2O": 1.0}, {"H": 1.0, "OH": 1.0}, "third_body"),        # dnu = +1
    ({"H2O2": 1.0}, {"OH": 2.0}, "standard"),                   # nu = 2
    ({"H2O2": 0.5}, {"OH": 1.0}, "standard"),                   # nu = 0.5, dnu = +0.5
    ({"H": 1.0, "O2": 1.0, "H2": 1.0, "O": 1.0}, {"HO2": 1.0, "H": 1.0, "OH": 1.0}, "third_body"),  # 4 slots
])
@pytest.mark.parametrize("T", [600.0, 1200.0, 2400.0])
def test_detailed_balance(reac, prod, rtype, T):
    names = ["H", "O", "OH", "H2O", "O2", "HO2", "H2O2", "H2", "N2"]
    r = Reaction("", reac, prod, rtype, True, (2.0e13, 0.3, 5000.0))
    m = _mech(names, [r])
    idx = {n: i for i, n in enumerate(names)}
    nu = {k: -v for k, v in reac.items()}
    for k, v in prod.items():
        nu[k] = nu.get(k, 0.0) + v
    C = np.full(len(names), 1e-7)
    # set the last product so that prod(C_p^nu)/prod(C_r^nu) = Kc
    kc = _kc(names, nu, T)
    lastp, lastnu = list(prod.items())[-1]
    others = 1.0
    for k, v in prod.items():
        if k != lastp:
            others *= C[idx[k]] ** v
    rprod = np.prod([C[idx[k]] ** v for k, v in reac.items()])
    C[idx[lastp]] = (kc * rprod / others) ** (1.0 / lastnu)
    P, Tt, Y = _state(m, C, T)
    w, qf, qb = getrates(m, P, Tt, Y, return_parts=True)
    assert abs(w[idx[lastp], 0]) <= 1e-12 * qf[0, 0] * lastnu


def test_negative_prefactor_duplicates_sum():
    """Duplicate reactions with prefactors A1 > 0 and A2 < 0 (same beta, Ea) rate as one reaction
    with A1 + A2 (real mechanisms write deviations from Arrhenius form this way)."""
    arr = lambda A: (A, 0.7, 8000.0)
    pair = [Reaction("", {"H": 1.0, "O2": 1.0}, {"O": 1.0, "OH": 1.0}, "standard", True, arr(3.0e13)),
            Reaction("", {"H": 1.0, "O2": 1.0}, {"O": 1.0, "OH": 1.0}, "standard", True, arr(-1.0e13))]
    one = [Reaction("", {"H": 1.0, "O2": 1.0}, {"O": 1.0, "OH": 1.0}, "standard", True, arr(2.0e13))]
    names = ["H", "O", "OH", "O2", "N2"]
    P, T, Y = make_states(len(names), 64, seed=3)
    w2 = getrates(_mech(names, pair), P, T, Y)
    w1 = getrates(_mech(names, one), P, T, Y)
    assert np.allclose(w2, w1, rtol=1e-13, atol=0.0)


def test_third_body_efficiencies_closed_form():
    base = _irrev({"H": 1.0, "OH": 1.0}, {"H2O": 1.0}, (1.0e12, 0.0, 0.0), rtype="third_body")
    zero = _irrev({"H": 1.0, "OH": 1.0}, {"H2O": 1.0}, (1.0e12, 0.0, 0.0), rtype="third_body")
    zero.third_body = {"N2": -1.0}
    C = np.array([1e-6, 2e-6, 3e-6, 4e-6])
    for r, M_expect in ((base, C.sum()), (zero, C.sum() - C[3])):
        m = _mech(["H", "OH", "H2O", "N2"], [r])
        P, T, Y = _state(m, C, 1000.0)
        w = getrates(m, P, T, Y)
        assert -w[0, 0] == pytest.approx(1e12 * C[0] * C[1] * M_expect, rel=1e-13)


def test_troe_center_closed_form():
    """T3 -> 0, T1 -> inf gives Fcent = a; at log10(Pr) = 0.4 + 0.67 log10(a), f1 = 0, F = a."""
    a = 0.43
    A, A0 = 2.0e12, 5.0e17
    r = _irrev({"H": 1.0, "O2": 1.0}, {"HO2": 1.0}, (A, 0.0, 0.0), rtype="troe",
               low=(A0, 0.0, 0.0), troe=(a, 1e-30, 1e30))
    m = _mech(["H", "O2", "HO2", "N2"], [r])
    pr = 10.0 ** (0.4 + 0.67 * math.log10(a))
    ctot = pr * A / A0
    C = np.array([0.1, 0.2, 0.0, 0.7]) * ctot
    P, T, Y = _state(m, C, 1500.0)
    q = -getrates(m, P, T, Y)[0, 0]
    assert q == pytest.approx(A * pr / (1 + pr) * a * C[0] * C[1], rel=1e-12)


def _falloff_q(rtype, pr, T=1500.0, A=2.0e12, **kw):
    """-wdot of H for one irreversible falloff reaction H + O2 (+M) => HO2 at reduced pressure
    ``pr`` (M = ctot; k0 and kinf temperature-independent, so Pr = A0 ctot / A)."""
    ctot = 1.0e-5
    A0 = pr * A / ctot
    r = _irrev({"H": 1.0, "O2": 1.0}, {"HO2": 1.0}, (A, 0.0, 0.0), rtype=rtype, low=(A0, 0.0, 0.0), **kw)
    m = _mech(["H", "O2", "HO2", "N2"], [r])
    C = np.array([0.1, 0.2, 0.0, 0.7]) * ctot
    P, Tt, Y = _state(m, C, T)
    return -getrates(m, P, Tt, Y)[0, 0], A * C[0] * C[1]


# Troe broadening, textbook definition (Gilbert, Luther & Troe 1983; Kee, Coltrin & Glarborg,
# Chemically Reacting Flow, eq. 9.99): log10 F = log10 Fcent / (1 + f1^2),
# f1 = (log10 Pr + c) / (n - 0.14 (log10 Pr + c)), c = -0.4 - 0.67 log10 Fcent,
# n = 0.75 - 1.27 log10 Fcent.
def _troe_lp_for_f1(fcent, f1):
    """log10 Pr at which f1 takes the value +1 or -1: x = lp + c with x = n / 1.14 (f1 = 1) or
    x = -n / 0.86 (f1 = -1)."""
    lf = math.log10(fcent)
    c, n = -0.4 - 0.67 * lf, 0.75 - 1.27 * lf
    x = n / 1.14 if f1 > 0 else -n / 0.86
    return x - c


@pytest.mark.parametrize("a", [0.2, 0.43, 0.8])
@pytest.mark.parametrize("f1", [1, -1])
def test_troe_half_power_at_unit_f1(a, f1):
    """At f1 = +-1 the broadening is F = Fcent^(1/2): pins 0.4, 0.67, 0.75, 1.27 and 0.14 together
    (Fcent = a with T3 -> 0, T1 -> inf)."""
    pr = 10.0 ** _troe_lp_for_f1(a, f1)
    q, kc = _falloff_q("troe", pr, troe=(a, 1e-30, 1e30))
    assert q == pytest.approx(kc * pr / (1 + pr) * math.sqrt(a), rel=1e-12)


@pytest.mark.parametrize("lp", [2.0, -3.0, 0.5])
def test_troe_limit_exponent_at_zero_n(lp):
    """With Fcent = 10^(0.75/1.27), n = 0 and f1 = -1/0.14 for every Pr, so
    F = Fcent^(1/(1 + (1/0.14)^2)) exactly: pins 0.14 (and the ratio 0.75 / 1.27)."""
    a = 10.0 ** (0.75 / 1.27)
    pr = 10.0 ** lp
    q, kc = _falloff_q("troe", pr, troe=(a, 1e-30, 1e30))
    F = a ** (1.0 / (1.0 + (1.0 / 0.14) ** 2))
    assert q == pytest.approx(kc * pr / (1 + pr) * F, rel=1e-12)


@pytest.mark.parametrize("lp", [250.0, -250.0])
def test_troe_large_pr_limit(lp):
    """|log10 Pr| -> inf: f1 -> -1/0.14, F -> Fcent^(1/(1 + (1/0.14)^2)) for any Fcent.  At
    |log10 Pr| = 250 the distance of f1 from its limit is |n / (0.14 x)| < 4%, so log10 F is
    within 10% of the limit (a wrong 0.14 moves it by 40% or more)."""
    a = 0.43
    pr = 10.0 ** lp
    q, kc = _falloff_q("troe", pr, troe=(a, 1e-30, 1e30))
    logF = math.log10(q / (kc * pr / (1 + pr)))
    limit = math.log10(a) / (1.0 + (1.0 / 0.14) ** 2)
    assert logF == pytest.approx(limit, rel=0.1)


@pytest.mark.parametrize("T2,T", [(1000.0, 1500.0), (4000.0, 800.0), (250.0, 2200.0)])
def test_troe_t2_term(T2, T):
    """T3 -> 0, T1 -> inf: Fcent = a + exp(-T2/T); at f1 = 0 (log10 Pr = -c) F = Fcent."""
    a = 0.3
    fcent = a + math.exp(-T2 / T)
    lf = math.log10(fcent)
    pr = 10.0 ** (0.4 + 0.67 * lf)
    q, kc = _falloff_q("troe", pr, T=T, troe=(a, 1e-30, 1e30, T2))
    assert q == pytest.approx(kc * pr / (1 + pr) * fcent, rel=1e-12)


@pytest.mark.parametrize("pr,X", [(10.0, 0.5), (0.1, 0.5), (100.0, 0.2), (1e-2, 0.2), (1e3, 0.1)])
@pytest.mark.parametrize("sri", [(0.6, 300.0, 900.0, 1.3, 0.2), (0.45, 1200.0, 250.0)])
def test_sri_exponent_closed_form(pr, X, sri):
    """SRI (Kee, Coltrin & Glarborg, eq. 9.101): X = 1/(1 + log10(Pr)^2) is 1/2 at Pr = 10 and
    0.1, 1/5 at Pr = 100 and 0.01, 1/10 at Pr = 1000; F = d (a e^(-b/T) + e^(-T/c))^X T^e
    (d = 1, e = 0 for three parameters)."""
    T = 1400.0
    q, kc = _falloff_q("sri", pr, T=T, sri=sri)
    a, b, c = sri[:3]
    d, e = (sri[3], sri[4]) if len(sri) == 5 else (1.0, 0.0)
    F = d * (a * math.exp(-b / T) + math.exp(-T / c)) ** X * T ** e
    assert q == pytest.approx(kc * pr / (1 + pr) * F, rel=1e-12)


def test_lindemann_limits():
    A, A0 = 2.0e12, 5.0e17
    r = _irrev({"H": 1.0, "O2": 1.0}, {"HO2": 1.0}, (A, 0.0, 0.0), rtype="lindemann", low=(A0, 0.0, 0.0))
    m = _mech(["H", "O2", "HO2", "N2"], [r])
    for pr in (1e-7, 1.0, 1e7):
        ctot = pr * A / A0
        C = np.array([0.1, 0.2, 0.0, 0.7]) * ctot
        P, T, Y = _state(m, C, 1500.0)
        q = -getrates(m, P, T, Y)[0, 0]
        assert q == pytest.approx(A * pr / (1 + pr) * C[0] * C[1], rel=1e-13)
    # low-pressure limit: kf -> k0 M
    ctot = 1e-7 * A / A0
    C = np.array([0.1, 0.2, 0.0, 0.7]) * ctot
    P, T, Y = _state(m, C, 1500.0)
    assert -getrates(m, P, T, Y)[0, 0] == pytest.approx(A0 * ctot * C[0] * C[1], rel=2e-7)


def test_sri_at_unit_reduced_pressure():
    A, A0 = 2.0e12, 5.0e17
    sri = (0.6, 300.0, 900.0, 1.3, 0.2)
    r = _irrev({"H": 1.0, "O2": 1.0}, {"HO2": 1.0}, (A, 0.0, 0.0), rtype="sri", low=(A0, 0.0, 0.0), sri=sri)
    m = _mech(["H", "O2", "HO2", "N2"], [r])
    T = 1400.0
    ctot = A / A0  # Pr = 1 -> X = 1
    C = np.array([0.1, 0.2, 0.0, 0.7]) * ctot
    P, Tt, Y = _state(m, C, T)
    F = sri[3] * (sri[0] * math.exp(-sri[1] / T) + math.exp(-T / sri[2])) * T ** sri[4]
    assert -getrates(m, P, Tt, Y)[0, 0] == pytest.approx(A * 0.5 * F * C[0] * C[1], rel=1e-13)


def test_plog_interpolation_closed_form():
    """PLOG (reactions/plog.j2:3-38): the rate at a listed pressure is that point's sum of
    Arrhenius terms, between two points ln k is linear in ln P, outside the range the end
    point's rate holds."""
    atm = PATM
    T = 1300.0
    pts = [(1.0, [(2.0e12, 0.5, 8000.0)]), (10.0, [(3.0e12, 0.2, 6000.0), (4.0e11, 1.0, 2000.0)]),
           (100.0, [(9.0e12, 0.0, 5000.0)])]
    plog = [(p * atm, A, b, Ea) for p, rates in pts for (A, b, Ea) in rates]
    r = _irrev({"H": 1.0, "O2": 1.0}, {"HO2": 1.0}, plog[0][1:], rtype="plog", plog=plog)
    m = _mech(["H", "O2", "HO2", "N2"], [r])

    def k_at(p_atm):
        X = np.array([0.2, 0.3, 0.0, 0.5])
        C = X * p_atm * atm / (R0 * T)
        P, Tt, Y = _state(m, C, T)
        return -getrates(m, P, Tt, Y)[0, 0] / (C[0] * C[1])

    def ksum(rates):
        return sum(A * T ** b * math.exp(-Ea / (RC * T)) for (A, b, Ea) in rates)

    k = [ksum(rates) for _, rates in pts]
    assert k_at(0.3) == pytest.approx(k[0], rel=1e-12)
    assert k_at(10.0) == pytest.approx(k[1], rel=1e-12)
    assert k_at(100.0) == pytest.approx(k[2], rel=1e-12)
    assert k_at(5000.0) == pytest.approx(k[2], rel=1e-12)
    assert math.log(k_at(math.sqrt(10.0))) == pytest.approx(0.5 * (math.log(k[0]) + math.log(k[1])), rel=1e-12)
    w = 0.3  # ln P a fraction w of the way from 10 to 100 atm
    assert math.log(k_at(10.0 ** (1.0 + w))) == pytest.approx((1 - w) * math.log(k[1]) + w * math.log(k[2]),
                                                              rel=1e-12)


def test_pressure_scaling():
    r1 = _irrev({"H": 1.0, "O2": 1.0}, {"O": 1.0, "OH": 1.0}, (1e13, 0.5, 3000.0))
    r2 = _irrev({"H": 1.0, "OH": 1.0}, {"H2O": 1.0}, (1e16, -1.0, 0.0), rtype="third_body")
    for r, order in ((r1, 2), (r2, 3)):
        m = _mech(["H", "O", "OH", "H2O", "O2", "N2"], [r])
        _, T, Y = make_states(6, 8, seed=5)
        P1 = np.full(8, 1013250.0)
        w1 = getrates(m, P1, T, Y)
        w2 = getrates(m, 7.0 * P1, T, Y)
        assert np.allclose(w2[0], 7.0 ** order * w1[0], rtol=1e-13)


@pytest.mark.parametrize("name", ["h2_burke", "h2_burke2012", "gri30", "synth_gri53", "synth_nhept88",
                                  "synth_jetsurf348"])
def test_conservation(name):
    m = load_bundled(name)
    P, T, Y = make_states(m.n_species, 256, seed=7)
    w, qf, qb = getrates(m, P, T, Y, return_parts=True)
    G = gross_scale(m, qf, qb)
    a = m.element_matrix()
    assert np.all(np.abs(a @ w) <= 1e-13 * (a @ G) + 1e-300)
    W = m.mw_array()
    assert np.all(np.abs(W @ w) <= 1e-13 * (W @ G))


def test_planted_mistakes_fail(monkeypatch):
    """A wrong pressure constant breaks detailed balance for dnu != 0; a wrong NASA
    coefficient breaks the enthalpy derivative pin."""
    monkeypatch.setattr(og, "PATM", 1.01325e5)
    with pytest.raises(AssertionError):
        test_detailed_balance({"H": 1.0, "O2": 1.0}, {"HO2": 1.0}, "standard", 1200.0)
    monkeypatch.undo()
    real = og._smh

    def bad_smh(th, T, lnT):
        s = real(th, T, lnT)
        a = th.low
        return s + np.where(T < th.t_mid, a[2] * T ** 2 / 6.0, 0.0)  # a2/3 instead of a2/6

    monkeypatch.setattr(og, "_smh", bad_smh)
    with pytest.raises(AssertionError):
        test_smh_derivative_is_enthalpy("H2O", 450.0)


def _falloff_pins():
    for a in (0.2, 0.43, 0.8):
        for f1 in (1, -1):
            test_troe_half_power_at_unit_f1(a, f1)
    for lp in (2.0, -3.0, 0.5):
        test_troe_limit_exponent_at_zero_n(lp)
    for lp in (250.0, -250.0):
        test_troe_large_pr_limit(lp)
    for T2, T in ((1000.0, 1500.0), (4000.0, 800.0), (250.0, 2200.0)):
        test_troe_t2_term(T2, T)
    for pr, X in ((10.0, 0.5), (0.1, 0.5), (100.0, 0.2), (1e-2, 0.2), (1e3, 0.1)):
        for sri in ((0.6, 300.0, 900.0, 1.3, 0.2), (0.45, 1200.0, 250.0)):
            test_sri_exponent_closed_form(pr, X, sri)


@pytest.mark.parametrize("plant", [
    ("TROE_N1", -1.25),                     # n = 0.75 - 1.25 log10 Fcent
    ("TROE_D", 0.11),                       # 0.11 in place of 0.14
    ("TROE_C1", -0.76),                     # c = -0.4 - 0.76 log10 Fcent
    ("TROE_N0", 0.7),                       # n = 0.70 - 1.27 log10 Fcent
    ("troe_t2_term", lambda T2, T: np.exp(T2 / T)),              # exp(+T2/T)
    ("sri_exponent", lambda lp: 1.0 / (1.0 + 2.0 * lp ** 2)),     # X = 1/(1 + 2 lp^2)
])
def test_planted_falloff_mistakes_fail(monkeypatch, plant):
    """Each mistake planted in the oracle's falloff broadening makes a falloff pin fail, and the
    unmodified oracle passes them all."""
    _falloff_pins()
    monkeypatch.setattr(og, *plant)
    with pytest.raises(AssertionError):
        _falloff_pins()
