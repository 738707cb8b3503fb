"""CPU FP64 oracle for getrates -- TEST INFRASTRUCTURE ONLY.

This module restates, in NumPy, the arithmetic of the reference's Python-language
getrates, i.e. the text the reference renders from its templates:

* header constants and ``pfac = Patm/(R0*T)``     chemgen/templates/python_header.j2:47-56,
                                                 chemgen/chemistry.py:204-206
* Y -> C, ctot                                    chemgen/templates/ytoc.j2:3-18
* NASA-7 smh and EG = exp(smh)                    chemgen/templates/exp_g.j2:2-27
* elementary reactions                            chemgen/templates/reactions/standard.j2:3-27
* third-body reactions                            chemgen/templates/reactions/third_body.j2:3-36
* Troe falloff                                    chemgen/templates/reactions/troe.j2:3-56
* PLOG                                            chemgen/templates/reactions/plog.j2:1-60
* coeff_sum                                       chemgen/chemistry.py:276

Coefficients are rounded to 16 significant digits (``"%.15e"``) exactly where the
templates print them that way, and every expression is evaluated in the order the
rendered Python evaluates it.  Lindemann and SRI falloff are not in the reference; they
follow the Troe template with the textbook broadening factors (F = 1 for Lindemann;
F = d [a exp(-b/T) + exp(-T/c)]^X T^e, X = 1/(1 + log10(Pr)^2) for SRI, Kee, Coltrin &
Glarborg, Chemically Reacting Flow, eq. 9.101).

Only tests, the smoke entry point and the benchmark's CPU-baseline leg import this file,
as a checker.  It is never the product path.

Arrays are species-major (SoA): ``Y[k, i]`` and ``wdot[k, i]`` with ``i`` the grid point.
"""
from __future__ import annotations

import numpy as np

R0 = 8.314510e+07
RC = 1.9872155832
PATM = 1013250.0

# Troe broadening constants (chemgen/templates/reactions/troe.j2:29-36; Gilbert, Luther & Troe,
# Ber. Bunsenges. Phys. Chem. 87 (1983) 169): c = C0 + C1 log10 Fcent, n = N0 + N1 log10 Fcent,
# f1 = (log10 Pr + c) / (n - D (log10 Pr + c)).  Module-level so that tests can plant mistakes.
TROE_C0, TROE_C1 = -0.4, -0.67
TROE_N0, TROE_N1 = 0.75, -1.27
TROE_D = 0.14


def troe_t2_term(T2, T):
    """The optional third Fcent term exp(-T2/T) (troe.j2:29-31)."""
    return np.exp(r15(-T2) / T)


def sri_exponent(lp):
    """SRI exponent X = 1 / (1 + log10(Pr)^2) (Kee, Coltrin & Glarborg, eq. 9.101)."""
    return 1.0 / (1.0 + lp ** 2)


def r15(x: float) -> float:
    """A coefficient as the templates print it: ``"%.15e" % x``."""
    return float("%.15e" % x)


def _smh(th, T, lnT):
    """chemgen/templates/exp_g.j2:7-24: low set if T < T_mid else high set."""
    out = np.empty_like(T)
    lo = T < th.t_mid
    for mask, a in ((lo, th.low), (~lo, th.high)):
        if not mask.any():
            continue
        t = T[mask]
        s = (r15(a[6] - a[0])
             + np.copysign(r15(abs(a[0])), a[0]) * lnT[mask]
             + np.copysign(r15(abs(a[1] / 2.0)), a[1]) * t
             + np.copysign(r15(abs(a[2] / 6.0)), a[2]) * t ** 2
             + np.copysign(r15(abs(a[3] / 12.0)), a[3]) * np.power(t, 3.0)
             + np.copysign(r15(abs(a[4] / 20.0)), a[4]) * np.power(t, 4.0))
        s = s - np.copysign(r15(abs(a[5])), a[5]) / t
        out[mask] = s
    return out


def _arrh(A, beta, Ea, T):
    """``A * T**beta * np.exp(-Ea/(Rc*T))`` as printed by reactions/standard.j2:5."""
    k = r15(A)
    if beta != 0.0:
        k = k * np.power(T, float(beta))
    return k * np.exp(r15(-1.0 * Ea) / (RC * T))


def _kb(r, kf, EG, idx, pfac, always_pfac_branch=False):
    """Reverse rate constant, reactions/standard.j2:7-16 (troe.j2:41-45)."""
    if not r.reversible:
        return np.zeros_like(kf)
    kb = kf
    for sp, nu in r.reactants.items():
        e = EG[idx[sp]]
        kb = kb * (e if nu == 1.0 else e ** float(nu))
    for sp, nu in r.products.items():
        e = EG[idx[sp]]
        kb = kb / (e if nu == 1.0 else e ** float(nu))
    cs = r.coeff_sum
    if cs != 0.0:
        kb = kb / pfac ** float(cs)
    return kb


def _prod(k, st, C, idx):
    out = k
    for sp, nu in st.items():
        c = C[idx[sp]]
        out = out * (c if nu == 1.0 else c ** float(nu))
    return out


def _m(r, ctot, C, idx):
    """``M = ctot +/- |eff-1| * C ...``, reactions/third_body.j2:18-25."""
    M = ctot
    for sp, e in r.third_body.items():
        M = M + r15(abs(e)) * C[idx[sp]] if e > 0.0 else M - r15(abs(e)) * C[idx[sp]]
    return M


def _falloff_kf(r, T, M):
    """Falloff kf (reactions/troe.j2:19-39; Lindemann/SRI extensions)."""
    A, b, Ea = r.arrh
    A0, b0, Ea0 = r.low
    k0 = _arrh(A0, b0, Ea0, T)
    kinf = _arrh(A, b, Ea, T)
    Pr = (k0 * M) / kinf
    with np.errstate(divide="ignore", over="ignore", under="ignore"):
        if r.rtype == "troe":
            tp = r.troe
            a, T3, T1 = tp[0], tp[1], tp[2]
            t3 = np.exp(-T / r15(T3)) if T3 != 0.0 else np.zeros_like(T)
            t1 = np.exp(-T / r15(T1)) if T1 != 0.0 else np.zeros_like(T)
            Fcent = r15(1.0 - a) * t3 + r15(a) * t1
            if len(tp) == 4 and tp[3]:
                Fcent = Fcent + troe_t2_term(tp[3], T)
            lf = np.log10(Fcent)
            C1 = TROE_C0 + TROE_C1 * lf
            N1 = TROE_N0 + TROE_N1 * lf
            lp = np.log10(Pr)
            f1 = (lp + C1) / (N1 - TROE_D * (lp + C1))
            F = 10.0 ** (lf / (1.0 + f1 ** 2))
        elif r.rtype == "sri":
            sp = tuple(r.sri) + ((1.0, 0.0) if len(r.sri) == 3 else ())
            a, b_, c, d, e = sp
            lp = np.log10(Pr)
            X = sri_exponent(lp)
            F = r15(d) * (r15(a) * np.exp(r15(-b_) / T) + np.exp(-T / r15(c))) ** X
            if e != 0.0:
                F = F * np.power(T, float(e))
        else:  # lindemann
            F = 1.0
    return kinf * (Pr / (1.0 + Pr)) * F


def _plog_kf(r, T, P):
    """reactions/plog.j2:3-38."""
    pts = sorted(set(p for p, *_ in r.plog))
    by = {p: [c[1:] for c in r.plog if c[0] == p] for p in pts}

    def ksum(p):
        k = 0.0
        for (A, b, Ea) in by[p]:
            k = k + (r15(A) * np.power(T, float(b)) * np.exp(r15(-1.0 * Ea) / (RC * T)))
        return k
    kfl = np.empty_like(T)
    kfh = np.empty_like(T)
    lpl = np.empty_like(T)
    lph = np.empty_like(T)
    done = np.zeros(T.shape, bool)
    for j, p in enumerate(pts):
        m = (~done) & (P < r15(p))
        if m.any():
            if j == 0:
                kl = ksum(p)
                kfl[m] = np.broadcast_to(kl, T.shape)[m]
                kfh[m] = kfl[m]
                lpl[m] = 1.0
                lph[m] = r15(np.log(p))
            else:
                kl, kh = ksum(pts[j - 1]), ksum(p)
                kfl[m] = np.broadcast_to(kl, T.shape)[m]
                kfh[m] = np.broadcast_to(kh, T.shape)[m]
                lpl[m] = r15(np.log(pts[j - 1]))
                lph[m] = r15(np.log(p))
        done |= m
    m = ~done
    if m.any():
        kl = ksum(pts[-1])
        kfl[m] = np.broadcast_to(kl, T.shape)[m]
        kfh[m] = kfl[m]
        lpl[m] = r15(np.log(pts[-1]))
        lph[m] = 100.0
    return np.exp(np.log(kfl) + (np.log(kfh) - np.log(kfl)) * (np.log(P) - lpl) / (lph - lpl))


def getrates(mech, P, T, Y, return_parts=False):
    """wdot[Ns, ng] (mol/cm^3/s) for P [dyn/cm^2], T [K], Y[Ns, ng] mass fractions."""
    P = np.asarray(P, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    ns = mech.n_species
    assert Y.shape[0] == ns, "Y must be species-major [Ns, ng]"
    idx = {s.name: i for i, s in enumerate(mech.species)}
    # Y -> C (ytoc.j2)
    C = np.empty_like(Y)
    for k, s in enumerate(mech.species):
        C[k] = Y[k] / r15(s.mw)
    ctot = np.zeros_like(T)
    for k in range(ns):
        ctot = ctot + C[k]
    for k in range(ns):
        C[k] = C[k] * P / (R0 * ctot * T)
    if any(r.rtype in ("third_body", "troe", "lindemann", "sri") for r in mech.reactions):
        ctot = np.zeros_like(T)
        for k in range(ns):
            ctot = ctot + C[k]
    pfac = PATM / (R0 * T)
    lnT = np.log(T)
    EG = np.empty((ns,) + T.shape)
    for k, s in enumerate(mech.species):
        EG[k] = np.exp(_smh(s.thermo, T, lnT))
    wdot = np.zeros_like(Y)
    qf_abs = np.zeros((mech.n_reactions,) + T.shape) if return_parts else None
    qb_abs = np.zeros_like(qf_abs) if return_parts else None
    for j, r in enumerate(mech.reactions):
        if r.rtype == "standard":
            kf = _arrh(*r.arrh, T)
            kb = _kb(r, kf, EG, idx, pfac)
            qf, qb = _prod(kf, r.reactants, C, idx), _prod(kb, r.products, C, idx)
            rr = qf - qb
        elif r.rtype == "third_body":
            kf = _arrh(*r.arrh, T)
            kb = _kb(r, kf, EG, idx, pfac)
            M = _m(r, ctot, C, idx)
            qf, qb = _prod(kf, r.reactants, C, idx), _prod(kb, r.products, C, idx)
            rr = M * (qf - qb)
            qf, qb = M * qf, M * qb
        elif r.rtype in ("troe", "lindemann", "sri"):
            M = _m(r, ctot, C, idx)
            kf = _falloff_kf(r, T, M)
            kb = _kb(r, kf, EG, idx, pfac)
            qf, qb = _prod(kf, r.reactants, C, idx), _prod(kb, r.products, C, idx)
            rr = qf - qb
        elif r.rtype == "plog":
            kf = _plog_kf(r, T, P)
            kb = _kb(r, kf, EG, idx, pfac)
            qf, qb = _prod(kf, r.reactants, C, idx), _prod(kb, r.products, C, idx)
            rr = qf - qb
        else:
            raise ValueError(r.rtype)
        for sp, nu in r.reactants.items():
            wdot[idx[sp]] = wdot[idx[sp]] - (rr if nu == 1.0 else rr * float(nu))
        for sp, nu in r.products.items():
            wdot[idx[sp]] = wdot[idx[sp]] + (rr if nu == 1.0 else rr * float(nu))
        if return_parts:
            qf_abs[j] = np.abs(qf)
            qb_abs[j] = np.abs(qb)
    if return_parts:
        return wdot, qf_abs, qb_abs
    return wdot


def gross_scale(mech, qf_abs, qb_abs):
    """G[k, i] = sum_j (nu'_kj + nu''_kj) (|q_f,j| + |q_b,j|): the size of the terms the
    reference adds into wdot_k.  Both sides count, because a species on both sides of a
    reaction (a spectator or collider such as ``H2O + H2O <=> H + OH + H2O``) receives
    ``-rr`` and ``+rr`` separately (reactions/standard.j2:21-27), which leaves rounding
    noise of size ulp(rr) in wdot_k even though its net coefficient is zero."""
    ns = mech.n_species
    idx = {s.name: i for i, s in enumerate(mech.species)}
    G = np.zeros((ns,) + qf_abs.shape[1:])
    for j, r in enumerate(mech.reactions):
        g = qf_abs[j] + qb_abs[j]
        for st in (r.reactants, r.products):
            for sp, n in st.items():
                G[idx[sp]] += abs(n) * g
    return G
