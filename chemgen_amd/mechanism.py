"""Mechanism front-end: Chemkin (.inp + therm.dat) and Cantera-YAML readers.

This replaces the reference's Cantera/pint based ``ckparser`` (chemgen/parser.py:18-187)
and the ``chemistry`` model (chemgen/chemistry.py:22-183) without Cantera.  The
semantics the reference relies on are reproduced:

* reaction types ``standard``, ``third_body``, ``troe``, ``plog`` with the names of
  chemgen/parser.py:42-49; ``lindemann`` and ``sri`` falloff are added (the reference
  drops them, chemgen/parser.py:146-147) and can be dropped again with
  ``compat="reference"``;
* three-body efficiencies are kept as ``eff - 1`` for every species whose efficiency
  differs from one, so that ``M = ctot + sum (eff-1) C`` (chemgen/parser.py:97,131);
* an explicit-collider three-body reaction (no ``M``) is an elementary reaction with
  the collider on both sides (chemgen/parser.py:107-130);
* pre-exponential factors are in cm-mol-s with reaction order ``sum(nu_r)`` (+1 for
  third-body and for the low-pressure limit), activation energies in cal/mol
  (chemgen/parser.py:152-159);
* molecular weights come from Cantera's atomic-weight table, which is what
  ``species.molecular_weight`` returns to chemgen/chemistry.py:260;
* single-range NASA-7 entries are expanded to two identical ranges, since the
  reference templates index ``data[1]`` (chemgen/templates/exp_g.j2:5).
"""
from __future__ import annotations

import hashlib
import math
import os
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

# Cantera's atomic weights (IUPAC conventional values) for the elements used here.
ATOMIC_WEIGHTS = {
    "H": 1.008, "He": 4.002602, "C": 12.011, "N": 14.007, "O": 15.999,
    "Ne": 20.1797, "Ar": 39.95, "F": 18.998403163, "Cl": 35.45, "S": 32.06,
    "Si": 28.085, "Kr": 83.798, "Xe": 131.293, "E": 5.48579909065e-4,
}

# Gas constant in cal/mol/K as Cantera + pint see it (8.314462618 J/mol/K, 1 cal = 4.184 J).
_R_CAL = 8.314462618 / 4.184
_NA = 6.02214076e23
_ATM_CGS = 1013250.0  # dyn/cm^2


def _norm_element(sym: str) -> str:
    s = sym.strip()
    if not s:
        return s
    return s[0].upper() + s[1:].lower()


@dataclass
class Nasa7:
    """NASA 7-coefficient polynomials: ``low`` applies for T < t_mid, else ``high``."""
    t_min: float
    t_mid: float
    t_max: float
    low: Tuple[float, ...]
    high: Tuple[float, ...]


@dataclass
class Species:
    name: str
    composition: Dict[str, float]
    thermo: Nasa7
    mw: float = 0.0

    def __post_init__(self):
        if not self.mw:
            self.mw = molecular_weight(self.composition)


def molecular_weight(composition: Dict[str, float]) -> float:
    w = 0.0
    for el, n in composition.items():
        w += n * ATOMIC_WEIGHTS[_norm_element(el)]
    return w


@dataclass
class Reaction:
    """One reaction in reference-compatible form.

    ``rtype`` is one of ``standard``, ``third_body``, ``troe``, ``lindemann``, ``sri``,
    ``plog``.  ``arrh`` is (A [cm-mol-s], beta, Ea [cal/mol]); for falloff reactions it is
    the high-pressure limit and ``low`` the low-pressure limit.  ``third_body`` maps
    species -> (efficiency - 1) for efficiencies different from one.
    """
    equation: str
    reactants: Dict[str, float]
    products: Dict[str, float]
    rtype: str
    reversible: bool = True
    arrh: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    low: Optional[Tuple[float, float, float]] = None
    troe: Optional[Tuple[float, ...]] = None      # (a, T3, T1[, T2])
    sri: Optional[Tuple[float, ...]] = None       # (a, b, c[, d, e])
    third_body: Dict[str, float] = field(default_factory=dict)
    plog: Optional[List[Tuple[float, float, float, float]]] = None  # (P dyn/cm2, A, b, Ea)
    duplicate: bool = False
    collider: Optional[str] = None                # explicit (+SP) collider of a falloff

    @property
    def coeff_sum(self) -> float:
        """sum(nu_products) - sum(nu_reactants), chemgen/chemistry.py:276."""
        return sum(self.products.values()) - sum(self.reactants.values())


@dataclass
class Mechanism:
    name: str
    elements: List[str]
    species: List[Species]
    reactions: List[Reaction]

    @property
    def n_species(self) -> int:
        return len(self.species)

    @property
    def n_reactions(self) -> int:
        return len(self.reactions)

    @property
    def species_names(self) -> List[str]:
        return [s.name for s in self.species]

    def species_index(self, name: str) -> int:
        return self._index()[name]

    def _index(self) -> Dict[str, int]:
        idx = getattr(self, "_idx_cache", None)
        if idx is None or len(idx) != len(self.species):
            idx = {s.name: i for i, s in enumerate(self.species)}
            object.__setattr__(self, "_idx_cache", idx)
        return idx

    def counts(self) -> Dict[str, int]:
        c: Dict[str, int] = {}
        for r in self.reactions:
            c[r.rtype] = c.get(r.rtype, 0) + 1
        return c

    def element_matrix(self):
        """a[e, k] = atoms of element e in species k (float)."""
        import numpy as np
        els = [_norm_element(e) for e in self.elements]
        a = np.zeros((len(els), self.n_species))
        for k, s in enumerate(self.species):
            for el, n in s.composition.items():
                a[els.index(_norm_element(el)), k] = n
        return a

    def mw_array(self):
        import numpy as np
        return np.array([s.mw for s in self.species], dtype=np.float64)

    def digest(self) -> str:
        """Content hash over species, thermo and reactions (used to key built kernels)."""
        h = hashlib.sha256()
        for s in self.species:
            h.update(repr((s.name, sorted(s.composition.items()), s.mw, s.thermo.t_mid,
                           s.thermo.low, s.thermo.high)).encode())
        for r in self.reactions:
            h.update(repr((sorted(r.reactants.items()), sorted(r.products.items()), r.rtype,
                           r.reversible, r.arrh, r.low, r.troe, r.sri,
                           sorted(r.third_body.items()), r.plog)).encode())
        return h.hexdigest()[:16]


# ----------------------------------------------------------------------------------
# equation parsing (shared by Chemkin and YAML readers)
# ----------------------------------------------------------------------------------

_COEF_RE = re.compile(r"^(\d+\.\d*|\.\d+|\d+)\s*(.+)$")
# a coefficient in exponent form ("5.00E-01 O2", written by Cantera's converters) must be
# followed by white space, so that a species name that starts with E is not read as one
_COEF_EXP_RE = re.compile(r"^((?:\d+\.?\d*|\.\d+)[eE][-+]?\d+)\s+(.+)$")
_FALLOFF_RE = re.compile(r"\(\s*\+\s*([^()\s]+(?:\([^()]*\))?)\s*\)")


def _split_terms(side: str, species: set) -> List[str]:
    raw = [t.strip() for t in side.split("+")]
    # re-join a coefficient that a "+" of its exponent split ("5.0E+01 O2")
    joined: List[str] = []
    for t in raw:
        if joined and re.fullmatch(r"(?:\d+\.?\d*|\.\d+)[eE]", joined[-1]) and re.match(r"\d", t):
            joined[-1] = joined[-1] + "+" + t
        else:
            joined.append(t)
    raw = joined
    out: List[str] = []
    for t in raw:
        if t == "" and out and (out[-1] + "+") in species:
            out[-1] = out[-1] + "+"
            continue
        if t == "":
            continue
        out.append(t)
    return out


def _parse_side(side: str, species: set) -> Tuple[Dict[str, float], bool]:
    """Return ({species: nu}, has_M) for one side of an equation."""
    stoich: Dict[str, float] = {}
    has_m = False
    for term in _split_terms(side, species):
        term = term.strip()
        if term.upper() == "M":
            has_m = True
            continue
        if term in species:
            name, nu = term, 1.0
        else:
            m = _COEF_EXP_RE.match(term)
            if not m or m.group(2).strip() not in species:
                m = _COEF_RE.match(term)
            if not m or m.group(2).strip() not in species:
                raise ValueError(f"unknown species term '{term}' in '{side}'")
            name, nu = m.group(2).strip(), float(m.group(1))
        stoich[name] = stoich.get(name, 0.0) + nu
    return stoich, has_m


def parse_equation(eq: str, species: set):
    """Parse ``eq`` -> (reactants, products, reversible, has_M, falloff_collider).

    ``falloff_collider`` is None for no ``(+X)``, ``"M"`` for ``(+M)``, else the species.
    """
    e = eq.strip()
    if "<=>" in e:
        lhs, rhs = e.split("<=>")
        rev = True
    elif "=>" in e:
        lhs, rhs = e.split("=>")
        rev = False
    elif "=" in e:
        lhs, rhs = e.split("=", 1)
        rev = True
    else:
        raise ValueError(f"no '=' in equation '{eq}'")
    coll = None
    fl = _FALLOFF_RE.findall(lhs)
    fr = _FALLOFF_RE.findall(rhs)
    if fl or fr:
        if len(fl) != 1 or len(fr) != 1 or fl[0].upper() != fr[0].upper():
            raise ValueError(f"inconsistent falloff collider in '{eq}'")
        coll = "M" if fl[0].upper() == "M" else fl[0]
        lhs = _FALLOFF_RE.sub("", lhs)
        rhs = _FALLOFF_RE.sub("", rhs)
    r, ml = _parse_side(lhs, species)
    p, mr = _parse_side(rhs, species)
    if ml != mr:
        raise ValueError(f"third body 'M' on one side only in '{eq}'")
    return r, p, rev, ml, coll


def _fmt_side(st: Dict[str, float]) -> str:
    parts = []
    for s, n in st.items():
        if n == 1.0:
            parts.append(s)
        elif float(n).is_integer():
            parts.append(f"{int(n)} {s}")
        else:
            parts.append(f"{n} {s}")
    return " + ".join(parts)


# ----------------------------------------------------------------------------------
# Chemkin reader
# ----------------------------------------------------------------------------------

def _strip_comment(line: str) -> str:
    i = line.find("!")
    return line if i < 0 else line[:i]


def _ffloat(s: str) -> float:
    s = s.strip().replace("D", "E").replace("d", "e")
    return float(s)


def parse_therm_lines(lines: List[str], default_tmid: float = 1000.0,
                      default_trange=(300.0, 5000.0)) -> Dict[str, Tuple[Dict[str, float], Nasa7]]:
    """Parse fixed-column NASA-7 entries (Chemkin THERMO format)."""
    out: Dict[str, Tuple[Dict[str, float], Nasa7]] = {}
    i = 0
    n = len(lines)
    while i < n:
        ln = lines[i].rstrip("\n")
        if not ln.strip() or ln.lstrip().startswith("!"):
            i += 1
            continue
        if len(ln) >= 80 and ln[79] == "1" and i + 3 < n:
            l1, l2, l3, l4 = ln, lines[i + 1], lines[i + 2], lines[i + 3]
            name = l1[:18].split()[0]
            comp: Dict[str, float] = {}
            slots = [l1[24 + 5 * j: 29 + 5 * j] for j in range(4)]
            if len(l1) >= 78:
                slots.append(l1[73:78])
            for sl in slots:
                el = sl[:2].strip()
                cnt = sl[2:].strip()
                if el and cnt:
                    try:
                        c = float(cnt)
                    except ValueError:
                        continue
                    if c != 0.0:
                        comp[_norm_element(el)] = comp.get(_norm_element(el), 0.0) + c
            tlo = _ffloat(l1[45:55]) if l1[45:55].strip() else default_trange[0]
            thi = _ffloat(l1[55:65]) if l1[55:65].strip() else default_trange[1]
            tmid_s = l1[65:73].strip()
            tmid = _ffloat(tmid_s) if tmid_s else default_tmid
            vals = []
            for L in (l2, l3, l4):
                for j in range(5):
                    f = L[15 * j: 15 * j + 15]
                    if f.strip():
                        vals.append(_ffloat(f))
            high = tuple(vals[0:7])
            low = tuple(vals[7:14])
            out[name] = (comp, Nasa7(tlo, tmid, thi, low, high))
            i += 4
            continue
        i += 1
    return out


_AUX_RE = re.compile(r"([^\s/]+)\s*/([^/]*)/")


def _ea_factor(units: List[str]) -> Tuple[float, float]:
    """(multiplier to cal/mol for Ea, A molecules->mol flag) from REACTIONS line units."""
    f = 1.0
    molec = False
    for u in units:
        U = u.upper()
        if U in ("CAL/MOLE", "CAL/MOL"):
            f = 1.0
        elif U in ("KCAL/MOLE", "KCAL/MOL"):
            f = 1000.0
        elif U in ("JOULES/MOLE", "J/MOL", "JOULES/MOL"):
            f = 1.0 / 4.184
        elif U in ("KJOULES/MOLE", "KJ/MOL", "KJOULES/MOL"):
            f = 1000.0 / 4.184
        elif U in ("KELVINS", "KELVIN"):
            f = _R_CAL
        elif U == "EVOLTS":
            f = 1.602176634e-19 * _NA / 4.184
        elif U == "MOLECULES":
            molec = True
    return f, molec


def parse_chemkin(mech_file: str, therm_file: Optional[str] = None, compat: str = "full",
                  name: Optional[str] = None) -> Mechanism:
    """Read a Chemkin mechanism (+ optional thermo database file)."""
    with open(mech_file, "r", errors="replace") as fh:
        raw = fh.read().splitlines()
    thermo: Dict[str, Tuple[Dict[str, float], Nasa7]] = {}
    if therm_file:
        with open(therm_file, "r", errors="replace") as fh:
            tl = fh.read().splitlines()
        thermo.update(_parse_thermo_section(tl))

    elements: List[str] = []
    species: List[str] = []
    section = None
    rxn_lines: List[str] = []
    th_lines: List[str] = []
    ea_f, molec = 1.0, False
    for line in raw:
        s = _strip_comment(line)
        st = s.strip()
        if not st:
            if section == "THERMO":
                th_lines.append(line)
            continue
        key = st.split()[0].upper()
        if section is None or key in ("ELEMENTS", "ELEM", "SPECIES", "SPEC", "THERMO",
                                       "REACTIONS", "REAC"):
            if key in ("ELEMENTS", "ELEM"):
                section = "ELEMENTS"
                rest = st.split()[1:]
                st = " ".join(rest)
            elif key in ("SPECIES", "SPEC"):
                section = "SPECIES"
                st = " ".join(st.split()[1:])
            elif key == "THERMO":
                section = "THERMO"
                th_lines = []
                continue
            elif key in ("REACTIONS", "REAC"):
                section = "REACTIONS"
                ea_f, molec = _ea_factor(st.split()[1:])
                continue
        if st.upper() == "END" or st.upper().startswith("END "):
            if section == "THERMO":
                thermo.update(_parse_thermo_section(th_lines))
            section = None if section != "REACTIONS" else "DONE"
            continue
        if section == "ELEMENTS":
            for tok in st.split():
                if tok.upper() == "END":
                    section = None
                    break
                el = tok.split("/")[0]
                elements.append(_norm_element(el))
        elif section == "SPECIES":
            for tok in st.split():
                if tok.upper() == "END":
                    section = None
                    break
                if tok not in species:
                    species.append(tok)
        elif section == "THERMO":
            th_lines.append(line)
        elif section == "REACTIONS":
            rxn_lines.append(st)
    if section == "THERMO":
        thermo.update(_parse_thermo_section(th_lines))

    sp_objs = []
    for s in species:
        if s not in thermo:
            raise ValueError(f"no thermo data for species {s}")
        comp, th = thermo[s]
        sp_objs.append(Species(s, comp, th))
    spset = set(species)
    reactions = _parse_reaction_lines(rxn_lines, spset, ea_f, molec, compat)
    mname = name or os.path.splitext(os.path.basename(mech_file))[0]
    return Mechanism(mname, elements, sp_objs, reactions)


def _parse_thermo_section(lines: List[str]):
    tmid = 1000.0
    tr = (300.0, 5000.0)
    body = []
    for ln in lines:
        st = ln.strip()
        if not st or st.startswith("!"):
            continue
        if st.upper().startswith("THERMO"):
            continue
        if st.upper() == "END":
            break
        body.append(ln)
    # optional temperature-range line
    if body:
        toks = body[0].split()
        if len(toks) == 3 and not (len(body[0]) >= 80 and body[0][79] == "1"):
            try:
                tr = (float(toks[0]), float(toks[2]))
                tmid = float(toks[1])
                body = body[1:]
            except ValueError:
                pass
    return parse_therm_lines(body, tmid, tr)


def _order(st: Dict[str, float]) -> float:
    return sum(st.values())


def _parse_reaction_lines(lines: List[str], spset: set, ea_f: float, molec: bool,
                          compat: str) -> List[Reaction]:
    num_re = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eEdD][+-]?\d+)?$")
    blocks: List[Tuple[str, List[str]]] = []
    for ln in lines:
        toks = ln.split()
        is_rxn = (len(toks) >= 4 and all(num_re.match(t) for t in toks[-3:])
                  and "=" in "".join(toks[:-3]) and "/" not in "".join(toks[:-3]))
        if is_rxn:
            blocks.append((ln, []))
        else:
            if not blocks:
                raise ValueError(f"auxiliary line before first reaction: {ln}")
            blocks[-1][1].append(ln)
    out: List[Reaction] = []
    for head, aux in blocks:
        toks = head.split()
        eq = "".join(toks[:-3])
        A, b, Ea = (_ffloat(t) for t in toks[-3:])
        reac, prod, rev, has_m, coll = parse_equation(eq, spset)
        kv: Dict[str, List[str]] = {}
        effs: Dict[str, float] = {}
        dup = False
        plog: List[Tuple[float, float, float, float]] = []
        for a in aux:
            au = a.upper().strip()
            if au.startswith("DUP"):
                dup = True
                rest = a.strip()[len(a.strip().split()[0]):]
                if not rest.strip():
                    continue
                a = rest
            for m in _AUX_RE.finditer(a):
                k, v = m.group(1), m.group(2).split()
                K = k.upper()
                if K in ("LOW", "TROE", "SRI", "REV", "HIGH", "FORD", "RORD", "UNITS", "LT",
                         "RLT", "TDEP", "XSMI", "MOME", "EXCI", "JAN", "FIT1", "CHEB", "TCHEB",
                         "PCHEB"):
                    kv[K] = v
                elif K == "PLOG":
                    P, pA, pb, pE = (float(x.replace("D", "E")) for x in v)
                    plog.append((P, pA, pb, pE))
                elif K.startswith("DUP"):
                    dup = True
                elif k in spset:
                    effs[k] = float(v[0].replace("D", "E"))
                else:
                    raise ValueError(f"unsupported auxiliary keyword '{k}' for '{eq}'")
            if au.startswith("DUP"):
                dup = True
        for bad in ("HIGH", "FORD", "RORD", "CHEB", "LT", "RLT", "TDEP", "XSMI", "MOME", "EXCI",
                    "JAN", "FIT1"):
            if bad in kv:
                raise ValueError(f"auxiliary keyword {bad} not supported ('{eq}')")
        order = _order(reac)
        # A unit conversion for MOLECULES: k [cm^3/molecule]^(n-1) -> [cm^3/mol]^(n-1)
        def conv_A(Aval, n_extra):
            if not molec:
                return Aval
            return Aval * _NA ** (order + n_extra - 1)
        Ea_c = Ea * ea_f
        if coll is not None:
            low = kv.get("LOW")
            if low is None:
                raise ValueError(f"falloff reaction without LOW: {eq}")
            lowt = (conv_A(_ffloat(low[0]), 1), _ffloat(low[1]), _ffloat(low[2]) * ea_f)
            if "TROE" in kv:
                tp = tuple(_ffloat(x) for x in kv["TROE"])
                rtype = "troe"
            elif "SRI" in kv:
                tp = tuple(_ffloat(x) for x in kv["SRI"])
                rtype = "sri"
            else:
                tp = None
                rtype = "lindemann"
            r = Reaction(eq, reac, prod, rtype, rev, (conv_A(A, 0), b, Ea_c), low=lowt,
                         duplicate=dup)
            if rtype == "troe":
                r.troe = _troe_params(tp)
            elif rtype == "sri":
                r.sri = tp
            if coll == "M":
                r.third_body = {s: e - 1.0 for s, e in effs.items() if e != 1.0}
            else:
                r.collider = coll
                # the reference keeps (eff-1) of Cantera's efficiency map {coll: 1.0}: empty
                r.third_body = {}
            out.append(r)
        elif has_m:
            r = Reaction(eq, reac, prod, "third_body", rev, (conv_A(A, 1), b, Ea_c), duplicate=dup)
            r.third_body = {s: e - 1.0 for s, e in effs.items() if e != 1.0}
            out.append(r)
        elif plog:
            r = Reaction(eq, reac, prod, "plog", rev, (conv_A(A, 0), b, Ea_c), duplicate=dup)
            r.plog = [(P * _ATM_CGS, conv_A(pA, 0), pb, pE * ea_f) for (P, pA, pb, pE) in plog]
            out.append(r)
        else:
            if "REV" in kv:
                # explicit reverse parameters: forward and reverse become two irreversible
                # reactions (what ck2yaml writes for REV)
                rv = kv["REV"]
                fwd = Reaction(eq, reac, prod, "standard", False, (conv_A(A, 0), b, Ea_c),
                               duplicate=dup)
                bwd = Reaction(eq + " (reverse)", dict(prod), dict(reac), "standard", False,
                               (_ffloat(rv[0]) * (_NA ** (_order(prod) - 1) if molec else 1.0),
                                _ffloat(rv[1]), _ffloat(rv[2]) * ea_f), duplicate=dup)
                out.extend([fwd, bwd])
                continue
            out.append(Reaction(eq, reac, prod, "standard", rev, (conv_A(A, 0), b, Ea_c),
                                duplicate=dup))
    return _apply_compat(out, compat)


def _troe_params(tp):
    """Normalise Troe parameters like Cantera: T3/T1 == 0 map to an infinite 1/T;
    T2 is kept only when non-zero (chemgen/templates/reactions/troe.j2:7,30)."""
    a, T3, T1 = tp[0], tp[1], tp[2]
    T2 = tp[3] if len(tp) > 3 else 0.0
    if T2 != 0.0:
        return (a, T3, T1, T2)
    return (a, T3, T1)


def _apply_compat(reactions: List[Reaction], compat: str) -> List[Reaction]:
    if compat == "full":
        return reactions
    if compat == "reference":
        # the reference keeps only standard/troe/third_body/plog (chemgen/parser.py:44-49)
        return [r for r in reactions if r.rtype in ("standard", "troe", "third_body", "plog")]
    raise ValueError(f"unknown compat mode {compat}")


# ----------------------------------------------------------------------------------
# Cantera YAML reader
# ----------------------------------------------------------------------------------

def _yaml_loader():
    import yaml

    class _L(yaml.SafeLoader):
        pass
    # YAML-1.1 booleans (NO, ON, Y, ...) are species names here: only true/false are bools
    _L.yaml_implicit_resolvers = {
        k: [(tag, rx) for (tag, rx) in v if tag != "tag:yaml.org,2002:bool"]
        for k, v in yaml.SafeLoader.yaml_implicit_resolvers.items()
    }
    _L.add_implicit_resolver("tag:yaml.org,2002:bool", re.compile(r"^(?:true|True|TRUE|false|False|FALSE)$"),
                             list("tTfF"))
    return _L


_LEN = {"cm": 1.0, "m": 100.0, "mm": 0.1}
_QTY = {"mol": 1.0, "kmol": 1000.0, "molec": 1.0 / _NA, "molecule": 1.0 / _NA}
_EA = {"cal/mol": 1.0, "kcal/mol": 1000.0, "J/mol": 1 / 4.184, "kJ/mol": 1000 / 4.184,
       "J/kmol": 1 / 4184.0, "kJ/kmol": 1 / 4.184, "cal/kmol": 1e-3, "kcal/kmol": 1.0,
       "K": _R_CAL, "eV": 1.602176634e-19 * _NA / 4.184}


def _qty_value(v, default_unit: str, table: Dict[str, float]) -> float:
    if isinstance(v, (int, float)):
        return float(v) * table[default_unit]
    parts = str(v).split(None, 1)
    val = float(parts[0])
    unit = parts[1].strip() if len(parts) > 1 else default_unit
    if unit not in table:
        raise ValueError(f"unsupported unit '{unit}'")
    return val * table[unit]


def load_yaml(path: str, compat: str = "full", name: Optional[str] = None) -> Mechanism:
    """A Cantera YAML mechanism.  Non-stoichiometric reaction ``orders`` are not read: the
    reference raises concentrations to the stoichiometric coefficients (Cantera's
    ``reaction.reactants``, chemgen/parser.py:73-75, reactions/standard.j2:19) and converts A
    with the same coefficients, and so does this build."""
    import yaml
    with open(path) as fh:
        doc = yaml.load(fh, Loader=_yaml_loader())
    units = doc.get("units", {}) or {}
    lf = _LEN[units.get("length", "m")]
    qf = _QTY[units.get("quantity", "kmol")]
    ea_unit = units.get("activation-energy", None)
    if ea_unit is None:
        eu = units.get("energy", "J")
        ea_unit = f"{eu}/{units.get('quantity', 'kmol')}"
    conc = lf ** 3 / qf  # one (length^3/quantity) in cm^3/mol
    phase = doc["phases"][0]
    names = [str(s) for s in phase["species"]] if isinstance(phase["species"], list) else None
    sp_entries = {str(s["name"]): s for s in doc["species"]}
    if names is None or (names and isinstance(phase["species"][0], dict)):
        names = list(sp_entries.keys())
    species = []
    for nm in names:
        e = sp_entries[nm]
        th = e["thermo"]
        if th.get("model", "NASA7") != "NASA7":
            raise ValueError(f"thermo model {th.get('model')} unsupported ({nm})")
        tr = [float(x) for x in th["temperature-ranges"]]
        data = [tuple(float(x) for x in d) for d in th["data"]]
        if len(data) == 1:
            nasa = Nasa7(tr[0], tr[-1], tr[-1], data[0], data[0])
        else:
            nasa = Nasa7(tr[0], tr[1], tr[2], data[0], data[1])
        comp = {_norm_element(str(k)): float(v) for k, v in (e.get("composition") or {}).items()}
        species.append(Species(nm, comp, nasa))
    spset = set(names)
    rx = []
    for e in doc.get("reactions", []):
        eq = str(e["equation"])
        reac, prod, rev, has_m, coll = parse_equation(eq, spset)
        typ = e.get("type", "elementary")
        dup = bool(e.get("duplicate", False))

        def arr(rc, extra):
            order = sum(reac.values()) + extra
            A = _qty_value(rc["A"], "1", {"1": 1.0}) if not isinstance(rc["A"], str) else float(str(rc["A"]).split()[0])
            A = A * conc ** (order - 1)
            Ea = _qty_value(rc.get("Ea", 0.0), ea_unit, _EA)
            return (A, float(rc.get("b", 0.0)), Ea)
        effs = {str(k): float(v) for k, v in (e.get("efficiencies") or {}).items()}
        if typ in ("falloff",):
            low = arr(e["low-P-rate-constant"], 1)
            high = arr(e["high-P-rate-constant"], 0)
            if "Troe" in e:
                t = e["Troe"]
                tp = (float(t["A"]), float(t["T3"]), float(t["T1"])) + ((float(t["T2"]),) if "T2" in t else ())
                r = Reaction(eq, reac, prod, "troe", rev, high, low=low, troe=_troe_params(tp), duplicate=dup)
            elif "SRI" in e:
                t = e["SRI"]
                tp = (float(t["A"]), float(t["B"]), float(t["C"])) + (
                    (float(t.get("D", 1.0)), float(t.get("E", 0.0))) if ("D" in t or "E" in t) else ())
                r = Reaction(eq, reac, prod, "sri", rev, high, low=low, sri=tp, duplicate=dup)
            else:
                r = Reaction(eq, reac, prod, "lindemann", rev, high, low=low, duplicate=dup)
            if coll is not None and coll != "M":
                r.collider = coll
                r.third_body = {}
            else:
                r.third_body = {s: v - 1.0 for s, v in effs.items() if v != 1.0}
            rx.append(r)
        elif typ == "three-body" or has_m:
            r = Reaction(eq, reac, prod, "third_body", rev, arr(e["rate-constant"], 1), duplicate=dup)
            r.third_body = {s: v - 1.0 for s, v in effs.items() if v != 1.0}
            rx.append(r)
        elif typ == "pressure-dependent-Arrhenius":
            pl = []
            for rc in e["rate-constants"]:
                P = rc["P"]
                if isinstance(P, str):
                    val, unit = P.split(None, 1)
                    P = float(val) * {"atm": _ATM_CGS, "Pa": 10.0, "bar": 1e6, "dyn/cm^2": 1.0}[unit.strip()]
                else:
                    P = float(P) * 10.0  # SI Pa
                a = arr(rc, 0)
                pl.append((P, a[0], a[1], a[2]))
            r = Reaction(eq, reac, prod, "plog", rev, pl[0][1:], plog=pl, duplicate=dup)
            rx.append(r)
        elif typ in ("elementary", "Arrhenius"):
            rx.append(Reaction(eq, reac, prod, "standard", rev, arr(e["rate-constant"], 0), duplicate=dup))
        else:
            if compat == "reference":
                continue
            raise ValueError(f"reaction type '{typ}' not supported: {eq}")
    elements = [_norm_element(str(x)) for x in phase.get("elements", [])]
    mname = name or os.path.basename(os.path.dirname(os.path.abspath(path)))
    return Mechanism(mname, elements, species, _apply_compat(rx, compat))


def load_mechanism(mech: Optional[str] = None, therm: Optional[str] = None,
                   yaml_file: Optional[str] = None, compat: str = "full",
                   prefer_yaml: bool = False, name: Optional[str] = None) -> Mechanism:
    """Resolve inputs the way ``generate_chemistry.py`` does.

    The reference reads ``yaml_file`` if it exists, else the ``.yaml`` next to the
    Chemkin file, and converts Chemkin only when neither exists
    (chemgen/parser.py:55-66).  ``prefer_yaml=True`` reproduces that resolution;
    by default an explicit Chemkin file is parsed directly.
    """
    if yaml_file and os.path.exists(yaml_file):
        return load_yaml(yaml_file, compat, name)
    if mech is None:
        raise ValueError("either mech or yaml_file must be given")
    side = os.path.splitext(mech)[0] + ".yaml"
    if prefer_yaml and os.path.exists(side):
        return load_yaml(side, compat, name)
    if mech.endswith((".yaml", ".yml")):
        return load_yaml(mech, compat, name)
    return parse_chemkin(mech, therm, compat, name)


def write_chemkin(mech: Mechanism, path: str, therm_path: Optional[str] = None) -> None:
    """Write a mechanism as Chemkin text (THERMO inline unless ``therm_path`` is given).

    A and Ea are printed with 17 significant digits so that parsing the file back
    gives the same doubles.
    """
    def g(x):
        return repr(float(x))
    lines = ["ELEMENTS", " ".join(e.upper() for e in mech.elements), "END", "SPECIES"]
    names = mech.species_names
    for i in range(0, len(names), 8):
        lines.append(" ".join(names[i:i + 8]))
    lines.append("END")
    th = _thermo_text(mech)
    if therm_path:
        with open(therm_path, "w") as fh:
            fh.write("\n".join(th) + "\n")
    else:
        lines.extend(th)
    lines.append("REACTIONS CAL/MOLE MOLES")
    for r in mech.reactions:
        arrow = "<=>" if r.reversible else "=>"
        lhs, rhs = _fmt_side(r.reactants).replace(" ", ""), _fmt_side(r.products).replace(" ", "")
        if r.rtype in ("troe", "lindemann", "sri"):
            c = "+M" if r.collider is None else "+" + r.collider
            eq = f"{lhs}({c}){arrow}{rhs}({c})"
        elif r.rtype == "third_body":
            eq = f"{lhs}+M{arrow}{rhs}+M"
        else:
            eq = f"{lhs}{arrow}{rhs}"
        A, b, Ea = r.arrh
        lines.append(f"{eq}   {g(A)} {g(b)} {g(Ea)}")
        if r.rtype in ("troe", "lindemann", "sri"):
            lines.append("    LOW/ " + " ".join(g(x) for x in r.low) + " /")
            if r.rtype == "troe":
                lines.append("    TROE/ " + " ".join(g(x) for x in r.troe) + " /")
            elif r.rtype == "sri":
                lines.append("    SRI/ " + " ".join(g(x) for x in r.sri) + " /")
        if r.rtype == "plog":
            for (P, pA, pb, pE) in r.plog:
                lines.append(f"    PLOG/ {g(P / _ATM_CGS)} {g(pA)} {g(pb)} {g(pE)} /")
        if r.third_body:
            lines.append("    " + " ".join(f"{s}/{g(e + 1.0)}/" for s, e in r.third_body.items()))
        if r.duplicate:
            lines.append("    DUPLICATE")
    lines.append("END")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def _thermo_text(mech: Mechanism) -> List[str]:
    out = ["THERMO ALL", "   300.000  1000.000  5000.000"]
    for s in mech.species:
        th = s.thermo
        comp = list(s.composition.items())
        cs = ""
        for el, n in comp[:4]:
            cs += f"{el.upper():<2}{int(round(n)):>3d}"
        cs = cs.ljust(20)
        l1 = f"{s.name:<18}{'':6}{cs}G{th.t_min:10.3f}{th.t_max:10.3f}{th.t_mid:8.2f}"
        l1 = l1.ljust(79) + "1"
        co = list(th.high) + list(th.low)
        fmt = lambda x: f"{x:15.8E}"
        l2 = "".join(fmt(x) for x in co[0:5]).ljust(79) + "2"
        l3 = "".join(fmt(x) for x in co[5:10]).ljust(79) + "3"
        l4 = "".join(fmt(x) for x in co[10:14]).ljust(79) + "4"
        out.extend([l1, l2, l3, l4])
    out.append("END")
    return out
