"""The reference's gpu_coef arrays for a mechanism -- TEST INFRASTRUCTURE.

Restates, for parallel level 1 (no dummy-reaction padding), the host arrays that the
reference's Fortran coefficient module hands to ``copyConstantsToDevice``
(chemgen/write_constants.py:14-347, argument order chemgen/constants.py:119-240), with the
same text rounding the module's Fortran literals carry: ``%21.15E`` for reals
(write_constants.py:30-33,124,152-180,310,337), ``%.2E`` for T_mid (:142), ``%.1f`` for
stoichiometric coefficients (:272), 0-based species indices (:213-216).

Used only by tests, to drive the legacy entry points of a generated compatibility library.
"""
from __future__ import annotations

from collections import OrderedDict
from typing import Dict, List

import numpy as np

RC = 1.9872155832


def _e15(x: float) -> float:
    return float(f"{x:21.15E}")


def _arh(A: float, beta: float, Ea: float):
    """(ln A, beta, -Ea/Rc) as write_constants.py:25-33 prints them."""
    return _e15(np.log(A)), _e15(beta), _e15(-Ea / RC)


def legacy_arrays(mech) -> "OrderedDict[str, np.ndarray]":
    """Arrays in copyConstantsToDevice's argument order for the reaction types present.
    The mechanism must hold only standard, troe and third_body reactions."""
    idx = {s.name: i for i, s in enumerate(mech.species)}
    ns = mech.n_species
    by_type: Dict[str, List] = {"standard": [], "troe": [], "third_body": []}
    for r in mech.reactions:
        if r.rtype not in by_type:
            raise ValueError(f"the reference's HIP coefficient arrays hold no {r.rtype} reactions")
        by_type[r.rtype].append(r)
    max_sp = max(max(len(r.reactants), len(r.products)) for r in mech.reactions)
    out: "OrderedDict[str, np.ndarray]" = OrderedDict()

    def maps(rs):
        sk_map, map_r, map_p, sk_coef, coef_r, coef_p = [], [], [], [], [], []
        for r in rs:
            for side, mp in ((r.reactants, map_r), (r.products, map_p)):
                sk_map += [idx[s] for s in side] + [0] * (max_sp - len(side))
                mp += [idx[s] for s in side] + [0] * (max_sp - len(side))
            sk_coef += [float(f"{-n:.1f}") for n in r.reactants.values()] + [0.0] * (max_sp - len(r.reactants))
            sk_coef += [float(f"{n:.1f}") for n in r.products.values()] + [0.0] * (max_sp - len(r.products))
            coef_r += [float(f"{n:.1f}") for n in r.reactants.values()] + [0.0] * (max_sp - len(r.reactants))
            coef_p += [float(f"{n:.1f}") for n in r.products.values()] + [0.0] * (max_sp - len(r.products))
        i32 = lambda a: np.array(a, dtype=np.int32)
        f64 = lambda a: np.array(a, dtype=np.float64)
        return i32(sk_map), f64(sk_coef), i32(map_r), f64(coef_r), i32(map_p), f64(coef_p)

    def eff(rs):
        return np.array([_e15(1.0 + r.third_body.get(s.name, 0.0)) for r in rs for s in mech.species],
                        dtype=np.float64)

    def ab(rs, which="arrh"):
        A, B = [], []
        for r in rs:
            a, b, e = _arh(*(r.arrh if which == "arrh" else r.low))
            A.append(a)
            B += [b, e]
        return np.array(A), np.array(B)

    if by_type["standard"]:
        rs = by_type["standard"]
        A, B = ab(rs)
        sk_map, sk_coef, map_r, coef_r, map_p, coef_p = maps(rs)
        out.update(A_h=A, B_h=B, sk_map_h=sk_map, sk_coef_h=sk_coef, map_r_h=map_r, coef_r_h=coef_r,
                   map_p_h=map_p, coef_p_h=coef_p)
    if by_type["troe"]:
        rs = by_type["troe"]
        A0, B0 = ab(rs, "low")
        Ai, Bi = ab(rs)
        sk_map, sk_coef, map_r, coef_r, map_p, coef_p = maps(rs)
        fc = []
        for r in rs:
            t = r.troe
            fc += [_e15(1.0 - t[0]), _e15(t[1]), _e15(t[0]), _e15(t[2])]
            fc += [_e15(1.0), _e15(t[3])] if len(t) > 3 else [_e15(0.0), _e15(0.0)]
        out.update(A_0_troe_h=A0, B_0_troe_h=B0, A_inf_troe_h=Ai, B_inf_troe_h=Bi, sk_map_troe_h=sk_map,
                   sk_coef_troe_h=sk_coef, map_r_troe_h=map_r, coef_r_troe_h=coef_r, map_p_troe_h=map_p,
                   coef_p_troe_h=coef_p, eff_fac_troe_h=eff(rs), fcent_coef_troe_h=np.array(fc))
    if by_type["third_body"]:
        rs = by_type["third_body"]
        A, B = ab(rs)
        sk_map, sk_coef, map_r, coef_r, map_p, coef_p = maps(rs)
        out.update(A_third_h=A, B_third_h=B, sk_map_third_h=sk_map, sk_coef_third_h=sk_coef,
                   map_r_third_h=map_r, coef_r_third_h=coef_r, map_p_third_h=map_p, coef_p_third_h=coef_p,
                   eff_fac_third_h=eff(rs))
    out["mw_h"] = np.array([_e15(s.mw) for s in mech.species])
    smh, tmid = [], []
    for s in mech.species:
        th = s.thermo
        for a in (th.low, th.high):
            smh += [_e15(a[6] - a[0]), _e15(a[0]), _e15(a[1] / 2.0), _e15(a[2] / 6.0), _e15(a[3] / 12.0),
                    _e15(a[4] / 20.0), _e15(-a[5])]
        tmid.append(float(f"{th.t_mid:.2E}"))
    out["smh_coef_h"] = np.array(smh)
    out["T_mid_h"] = np.array(tmid)
    assert len(out["mw_h"]) == ns
    return out
