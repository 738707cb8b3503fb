"""40-digit arbiter for getrates -- TEST INFRASTRUCTURE ONLY.

Evaluates the reference's Python-language getrates (the formulas of oracle/getrates_np.py:
chemgen/templates/ytoc.j2:3-18, exp_g.j2:2-27, reactions/{standard,third_body,troe,plog}.j2,
chemgen/chemistry.py:204-206,276) in mpmath at ``dps`` significant decimal digits.  Every
input -- P, T, Y and each coefficient after the reference's ``"%.15e"`` printing -- enters as
the exact value of its float64, and only the arithmetic is carried at high precision.  The
result is the exact value (to ~1e-35 relative) of the reference formula on the float64
inputs the FP64 paths see, so both the FP64 oracle's and the GPU kernel's rounding errors can
be measured against it; where they disagree with each other, the arbiter says which one is
closer.

Only tests and the benchmark's parity leg (as a checker) use it; it is pure Python, one
point at a time, for samples of at most a few hundred points.
"""
from __future__ import annotations

import math

import mpmath as mp
import numpy as np

from oracle import getrates_np as og
from oracle.getrates_np import PATM, R0, RC, r15


def _f(x) -> mp.mpf:
    return mp.mpf(float(x))


def _cs(x: float) -> float:
    """A NASA coefficient with the sign-preserving 16-digit rounding of exp_g.j2."""
    return math.copysign(r15(abs(x)), x)


class _Mech:
    """Per-mechanism constants as exact mpf values of their float64 forms."""

    def __init__(self, mech):
        self.mech = mech
        self.idx = {s.name: i for i, s in enumerate(mech.species)}
        self.mw = [_f(r15(s.mw)) for s in mech.species]
        self.th = []
        for s in mech.species:
            t = s.thermo
            rows = []
            for a in (t.low, t.high):
                rows.append((_f(r15(a[6] - a[0])), _f(_cs(a[0])), _f(_cs(a[1] / 2.0)), _f(_cs(a[2] / 6.0)),
                             _f(_cs(a[3] / 12.0)), _f(_cs(a[4] / 20.0)), _f(_cs(a[5]))))
            self.th.append((float(t.t_mid), rows[0], rows[1]))


def _smh(row, T, lnT):
    c0, a0, a1, a2, a3, a4, a5 = row
    return c0 + a0 * lnT + a1 * T + a2 * T ** 2 + a3 * T ** 3 + a4 * T ** 4 - a5 / T


def _arrh(A, beta, Ea, T):
    k = _f(r15(A))
    if beta != 0.0:
        k = k * mp.power(T, _f(beta))
    return k * mp.exp(_f(r15(-1.0 * Ea)) / (_f(RC) * T))


def _falloff_kf(r, T, M):
    k0 = _arrh(*r.low, T)
    kinf = _arrh(*r.arrh, T)
    Pr = (k0 * M) / kinf
    if r.rtype == "troe":
        a, T3, T1 = r.troe[0], r.troe[1], r.troe[2]
        t3 = mp.exp(-T / _f(r15(T3))) if T3 != 0.0 else mp.mpf(0)
        t1 = mp.exp(-T / _f(r15(T1))) if T1 != 0.0 else mp.mpf(0)
        Fcent = _f(r15(1.0 - a)) * t3 + _f(r15(a)) * t1
        if len(r.troe) == 4 and r.troe[3]:
            Fcent = Fcent + mp.exp(_f(r15(-r.troe[3])) / T)
        lf = mp.log10(Fcent)
        C1 = _f(og.TROE_C0) + _f(og.TROE_C1) * lf
        N1 = _f(og.TROE_N0) + _f(og.TROE_N1) * lf
        lp = mp.log10(Pr)
        f1 = (lp + C1) / (N1 - _f(og.TROE_D) * (lp + C1))
        F = mp.power(10, lf / (1 + f1 ** 2))
    elif r.rtype == "sri":
        sp = tuple(r.sri) + ((1.0, 0.0) if len(r.sri) == 3 else ())
        a, b, c, d, e = sp
        lp = mp.log10(Pr)
        X = 1 / (1 + lp ** 2)
        F = _f(r15(d)) * mp.power(_f(r15(a)) * mp.exp(_f(r15(-b)) / T) + mp.exp(-T / _f(r15(c))), X)
        if e != 0.0:
            F = F * mp.power(T, _f(e))
    else:
        F = mp.mpf(1)
    return kinf * (Pr / (1 + Pr)) * F


def _plog_kf(r, T, P):
    pts = sorted(set(p for p, *_ in r.plog))
    by = {p: [c[1:] for c in r.plog if c[0] == p] for p in pts}

    def ksum(p):
        k = mp.mpf(0)
        for (A, b, Ea) in by[p]:
            k = k + _f(r15(A)) * mp.power(T, _f(b)) * mp.exp(_f(r15(-1.0 * Ea)) / (_f(RC) * T))
        return k

    lg = lambda p: _f(r15(float(np.log(p))))
    Pf = float(P)
    j = next((j for j, p in enumerate(pts) if Pf < r15(p)), None)
    if j == 0:
        kl = kh = ksum(pts[0])
        lpl, lph = mp.mpf(1), lg(pts[0])
    elif j is None:
        kl = kh = ksum(pts[-1])
        lpl, lph = lg(pts[-1]), mp.mpf(100)
    else:
        kl, kh = ksum(pts[j - 1]), ksum(pts[j])
        lpl, lph = lg(pts[j - 1]), lg(pts[j])
    return mp.exp(mp.log(kl) + (mp.log(kh) - mp.log(kl)) * (mp.log(P) - lpl) / (lph - lpl))


def _point(cm: _Mech, P, T, Y):
    """wdot (list of mpf) and the gross scale G_k = sum_j (nu'_kj + nu''_kj)(|q_f| + |q_b|)."""
    mech = cm.mech
    ns = mech.n_species
    idx = cm.idx
    C = [_f(Y[k]) / cm.mw[k] for k in range(ns)]
    ctot = mp.fsum(C)
    C = [c * P / (_f(R0) * ctot * T) for c in C]
    ctot = mp.fsum(C)
    pfac = _f(PATM) / (_f(R0) * T)
    lnT = mp.log(T)
    EG = []
    for (tmid, lo, hi) in cm.th:
        EG.append(mp.exp(_smh(lo if float(T) < tmid else hi, T, lnT)))
    w = [mp.mpf(0)] * ns
    G = [mp.mpf(0)] * ns

    def prod(k, side):
        out = k
        for sp, nu in side.items():
            out = out * (C[idx[sp]] if nu == 1.0 else mp.power(C[idx[sp]], _f(nu)))
        return out

    for r in mech.reactions:
        M = None
        if r.rtype in ("third_body", "troe", "lindemann", "sri"):
            M = ctot
            for sp, e in r.third_body.items():
                M = M + _f(r15(abs(e))) * C[idx[sp]] if e > 0.0 else M - _f(r15(abs(e))) * C[idx[sp]]
        if r.rtype in ("standard", "third_body"):
            kf = _arrh(*r.arrh, T)
        elif r.rtype == "plog":
            kf = _plog_kf(r, T, P)
        else:
            kf = _falloff_kf(r, T, M)
        if r.reversible:
            kb = kf
            for sp, nu in r.reactants.items():
                kb = kb * (EG[idx[sp]] if nu == 1.0 else mp.power(EG[idx[sp]], _f(nu)))
            for sp, nu in r.products.items():
                kb = kb / (EG[idx[sp]] if nu == 1.0 else mp.power(EG[idx[sp]], _f(nu)))
            if r.coeff_sum != 0.0:
                kb = kb / mp.power(pfac, _f(r.coeff_sum))
        else:
            kb = mp.mpf(0)
        qf, qb = prod(kf, r.reactants), prod(kb, r.products)
        if r.rtype == "third_body":
            qf, qb = M * qf, M * qb
        rr = qf - qb
        g = abs(qf) + abs(qb)
        for sp, nu in r.reactants.items():
            w[idx[sp]] -= rr * _f(nu)
            G[idx[sp]] += g * abs(_f(nu))
        for sp, nu in r.products.items():
            w[idx[sp]] += rr * _f(nu)
            G[idx[sp]] += g * abs(_f(nu))
    return w, G


def getrates_mp(mech, P, T, Y, dps: int = 40):
    """(wdot, G) as float64 arrays [Ns, n] rounded from ``dps``-digit values: wdot of the
    reference formula and the gross scale of its terms, for P [n], T [n], Y [Ns, n]."""
    P = np.asarray(P, dtype=np.float64).reshape(-1)
    T = np.asarray(T, dtype=np.float64).reshape(-1)
    Y = np.asarray(Y, dtype=np.float64)
    n = P.shape[0]
    ns = mech.n_species
    out = np.empty((ns, n))
    gross = np.empty((ns, n))
    with mp.workdps(dps):
        cm = _Mech(mech)
        for i in range(n):
            w, G = _point(cm, _f(P[i]), _f(T[i]), Y[:, i])
            out[:, i] = [float(x) for x in w]
            gross[:, i] = [float(x) for x in G]
    return out, gross


def arbiter_stats(w, w_mp, G):
    """Errors of an FP64 result ``w`` against the arbiter: max raw relative error (entries
    with a non-zero exact value), the fraction of entries within 1e-12 raw, and the max
    gross-normalised error |dw| / (|w_mp| + G)."""
    w = np.asarray(w, dtype=np.float64)
    d = np.abs(w - w_mp)
    nz = w_mp != 0.0
    with np.errstate(divide="ignore", invalid="ignore"):
        raw = np.where(nz, d / np.abs(w_mp), 0.0)
        gross = d / (np.abs(w_mp) + G)
    gross = np.where(np.isfinite(gross), gross, np.where(d == 0.0, 0.0, np.inf))
    return {
        "max_raw": float(raw[nz].max()) if nz.any() else 0.0,
        "frac_raw_le_1e-12": float(np.mean(raw[nz] <= 1e-12)) if nz.any() else 1.0,
        "max_gross": float(gross.max()) if gross.size else 0.0,
    }
