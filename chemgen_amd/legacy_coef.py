"""The reference's Fortran coefficient module ``coef_m_vN.f90`` for one mechanism.

The reference's ``chemgen_m`` (chemgen/templates/chemgen_mod.j2:8-16) ``use``s the module
``coef_m_vN`` written by ``write_coef_module`` (chemgen/write_constants.py:456-493) and
passes its arrays to ``copyConstantsToDevice`` (chemgen/constants.py:119-240).  With the
compatibility library of :mod:`chemgen_amd.legacy_abi` in place of the reference's HIP code,
this module is the last file ``chemgen_m`` needs; :func:`emit_coef_module` writes it with the
reference's array names, types, sizes and values, so that no reference tooling is required:

* per reaction type present (standard, falloff-Troe, three-body, in that order):
  ``A*_h`` = ln A and ``B*_h`` = (beta, -Ea/Rc) pairs (write_constants.py:14-113; for Troe
  ``A_inf_troe_h``/``B_inf_troe_h`` from the high-pressure and ``A_0_troe_h``/``B_0_troe_h``
  from the low-pressure limit), the species maps ``sk_map*_h``, ``map_r*_h``, ``map_p*_h``
  (0-based, each side padded with 0 to MAX_SP entries, :196-237), the stoichiometric
  coefficients ``sk_coef*_h`` (-nu reactants, +nu products), ``coef_r*_h``, ``coef_p*_h``
  (:239-275), the dense efficiencies ``eff_fac_troe_h``/``eff_fac_third_h`` = 1 + (eff - 1)
  per species (:277-312) and ``fcent_coef_troe_h`` = (1 - a, T3, a, T1, 1 | 0, T2 | 0)
  (:314-347);
* ``mw_h``, ``T_mid_h`` and ``smh_coef_h`` = (a6 - a0, a0, a1/2, a2/6, a3/12, a4/20, -a5)
  for the low then the high range of every species (:116-194).

Numbers are printed as the reference prints them: ``%21.15E`` with a ``D`` exponent for
reals, ``%.2E`` for T_mid, ``%.1f`` for stoichiometric coefficients.  Level 3 rounds MAX_SP up
to a power of two; level 4 replaces the maps by the dense matrices of the reference's rocBLAS
path; with reactions per block given, each type is padded with the reference's dummy reactions
(write_constants.py:85-90).
"""
from __future__ import annotations

import math
from typing import Dict, List, Sequence, Tuple

from .mechanism import Mechanism

RC = 1.9872155832
_SUFFIX = {"standard": "", "troe": "_troe", "third_body": "_third"}


def _d(x: float) -> str:
    return f"{x:21.15E}".replace("E", "D").strip()


class _Dummy:
    """The reference's padding reaction (chemgen/chemistry.py:143-183): the first species with
    coefficient 0 on both sides, A = 1, beta = Ea = 0; Troe low limit (1, 0, 0) and centre
    (1, 1, 1e30); three-body efficiencies {first species: 0} (eff - 1)."""

    def __init__(self, mech, rtype):
        sp0 = mech.species[0].name
        self.rtype = rtype
        self.reactants = {sp0: 0.0}
        self.products = {sp0: 0.0}
        self.arrh = (1.0, 0.0, 0.0)
        self.low = (1.0, 0.0, 0.0)
        self.troe = (1.0, 1.0, 1e30)
        self.third_body = {sp0: 0.0} if rtype == "third_body" else {}


def max_species(mech: Mechanism, level: int) -> int:
    """MAX_SP of the level's arrays: the most species on one side of a reaction, rounded up to a
    power of two at level 3 (chemgen/chemistry.py:124-133, called with level > 2 by
    write_constants.py:202,244,285)."""
    m = max(max(len(r.reactants), len(r.products)) for r in mech.reactions)
    return 2 ** math.ceil(math.log2(m)) if level > 2 else m


def reactions_by_type(mech: Mechanism, rpb=None) -> Dict[str, list]:
    """The reactions of each type the reference's HIP arrays hold, in mechanism order, each
    type padded with dummy reactions to a multiple of ``rpb`` when it is given
    (write_constants.py:85-90, repeated in every writer)."""
    by_type: Dict[str, list] = {t: [] for t in _SUFFIX}
    for r in mech.reactions:
        if r.rtype not in by_type:
            raise ValueError(f"the reference's coefficient module holds no {r.rtype} reactions")
        by_type[r.rtype].append(r)
    if rpb:
        for t, rs in by_type.items():
            if rs and len(rs) % rpb:
                rs += [_Dummy(mech, t) for _ in range(rpb - len(rs) % rpb)]
    return by_type


def _arrays(mech: Mechanism, level: int = 1, rpb=None) -> List[Tuple[str, str, List[str]]]:
    """(Fortran type, name, literal values) in the module's order (write_constants.py:456-493)."""
    if level not in (1, 2, 3, 4):
        raise ValueError("parallel level must be 1..4")
    if level > 1 and not rpb:
        raise ValueError("parallel levels 2-4 need reactions per block (the reference asserts it)")
    idx = {s.name: i for i, s in enumerate(mech.species)}
    ns = mech.n_species
    by_type = reactions_by_type(mech, rpb)
    present = [t for t in _SUFFIX if by_type[t]]
    max_sp = max_species(mech, level)
    out: List[Tuple[str, str, List[str]]] = []
    real, integer = "real(kind=8)", "integer"

    def arrhenius(params: Sequence[float]) -> Tuple[str, List[str]]:
        A, beta, Ea = params
        return _d(math.log(A)), [_d(beta), _d(-Ea / RC)]

    for t in present:
        rs = by_type[t]
        if t == "troe":
            hi = [arrhenius(r.arrh) for r in rs]
            lo = [arrhenius(r.low) for r in rs]
            out.append((real, "A_inf_troe_h", [a for a, _ in hi]))
            out.append((real, "B_inf_troe_h", [b for _, bb in hi for b in bb]))
            out.append((real, "A_0_troe_h", [a for a, _ in lo]))
            out.append((real, "B_0_troe_h", [b for _, bb in lo for b in bb]))
        else:
            ab = [arrhenius(r.arrh) for r in rs]
            out.append((real, f"A{_SUFFIX[t]}_h", [a for a, _ in ab]))
            out.append((real, f"B{_SUFFIX[t]}_h", [b for _, bb in ab for b in bb]))
    if level < 4:
        for t in present:
            sk, mr, mp = [], [], []
            for r in by_type[t]:
                for side, m in ((r.reactants, mr), (r.products, mp)):
                    ks = [str(idx[s]) for s in side] + ["0"] * (max_sp - len(side))
                    sk += ks
                    m += ks
            sfx = _SUFFIX[t]
            out += [(integer, f"sk_map{sfx}_h", sk), (integer, f"map_r{sfx}_h", mr), (integer, f"map_p{sfx}_h", mp)]
        for t in present:
            sk, cr, cp = [], [], []
            for r in by_type[t]:
                sk += [f"{-n:.1f}" for n in r.reactants.values()] + ["0.0"] * (max_sp - len(r.reactants))
                sk += [f"{n:.1f}" for n in r.products.values()] + ["0.0"] * (max_sp - len(r.products))
                cr += [f"{n:.1f}" for n in r.reactants.values()] + ["0.0"] * (max_sp - len(r.reactants))
                cp += [f"{n:.1f}" for n in r.products.values()] + ["0.0"] * (max_sp - len(r.products))
            sfx = _SUFFIX[t]
            out += [(real, f"sk_coef{sfx}_h", sk), (real, f"coef_r{sfx}_h", cr), (real, f"coef_p{sfx}_h", cp)]
    else:
        # level 4: dense matrices for the rocBLAS products (write_constants.py:350-406):
        # sk_coef/coef_r/coef_p [reaction][species], wdot_coef [species][reaction]
        for t in present:
            rs = by_type[t]
            nr = len(rs)
            sk, cr, cp, wd = [0.0] * (nr * ns), [0.0] * (nr * ns), [0.0] * (nr * ns), [0.0] * (nr * ns)
            for q, r in enumerate(rs):
                for side, sign in ((r.reactants, -1.0), (r.products, 1.0)):
                    for sp, nu in side.items():
                        sk[q * ns + idx[sp]] += sign * nu
                        wd[q + idx[sp] * nr] += sign * nu
                for sp, nu in r.reactants.items():
                    cr[q * ns + idx[sp]] += nu
                for sp, nu in r.products.items():
                    cp[q * ns + idx[sp]] += nu
            sfx = _SUFFIX[t]
            f1 = lambda a: [f"{v:.1f}" for v in a]
            out += [(real, f"sk_coef{sfx}_h", f1(sk)), (real, f"coef_r{sfx}_h", f1(cr)),
                    (real, f"coef_p{sfx}_h", f1(cp)), (real, f"wdot_coef{sfx}_h", f1(wd))]
    for t in ("troe", "third_body"):
        if by_type[t]:
            vals = [_d(1.0 + r.third_body.get(s.name, 0.0)) for r in by_type[t] for s in mech.species]
            out.append((real, f"eff_fac{_SUFFIX[t]}_h", vals))
    if by_type["troe"]:
        vals = []
        for r in by_type["troe"]:
            a, t3, t1 = r.troe[0], r.troe[1], r.troe[2]
            vals += [_d(1.0 - a), _d(t3), _d(a), _d(t1)]
            vals += [_d(1.0), _d(r.troe[3])] if len(r.troe) > 3 else [_d(0.0), _d(0.0)]
        out.append((real, "fcent_coef_troe_h", vals))
    out.append((real, "mw_h", [_d(s.mw) for s in mech.species]))
    out.append((real, "T_mid_h", [f"{s.thermo.t_mid:.2E}".replace("E", "D") for s in mech.species]))
    smh = []
    for s in mech.species:
        for a in (s.thermo.low, s.thermo.high):
            smh += [_d(a[6] - a[0]), _d(a[0]), _d(a[1] / 2.0), _d(a[2] / 6.0), _d(a[3] / 12.0),
                    _d(a[4] / 20.0), _d(-a[5])]
    out.append((real, "smh_coef_h", smh))
    return out


def emit_coef_module(mech: Mechanism, level: int = 1, rpb=None) -> str:
    """Free-form Fortran source of ``module coef_m_v<level>`` (see the module docstring):
    levels 1-3 with species maps (MAX_SP rounded to a power of two at level 3), level 4 with the
    dense matrices of the rocBLAS path; each reaction type padded with the reference's dummy
    reactions to a multiple of ``rpb`` when it is given (its CLI always gives it, default 8)."""
    lines = [f"module coef_m_v{level}", "  implicit none", ""]
    for ftype, name, vals in _arrays(mech, level, rpb):
        lines.append(f"  {ftype} :: {name}({len(vals)})=(/&")
        per = 4
        for i in range(0, len(vals), per):
            chunk = ", ".join(vals[i:i + per])
            last = i + per >= len(vals)
            lines.append(f"    {chunk}{'' if last else ','}&")
        lines.append("  /)")
    lines += ["", f"end module coef_m_v{level}", ""]
    return "\n".join(lines)


def parse_coef_module(text: str) -> Dict[str, "object"]:
    """The arrays of a coefficient module as NumPy arrays (int32 for ``integer``, float64 for
    ``real(kind=8)``), by name: what ``chemgen_m`` passes to copyConstantsToDevice."""
    import re
    import numpy as np
    out = {}
    for m in re.finditer(r"(integer|real\(kind=8\))\s*::\s*(\w+)\((\d+)\)\s*=\s*\(/(.*?)/\)", text, re.S):
        ftype, name, n, body = m.group(1), m.group(2), int(m.group(3)), m.group(4)
        toks = [t for t in re.split(r"[,\s&]+", body) if t]
        if ftype == "integer":
            arr = np.array([int(t) for t in toks], dtype=np.int32)
        else:
            arr = np.array([float(t.replace("D", "E").replace("d", "e")) for t in toks], dtype=np.float64)
        if arr.size != n:
            raise ValueError(f"{name}: {arr.size} values for size {n}")
        out[name] = arr
    return out
