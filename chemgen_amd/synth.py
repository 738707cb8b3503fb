"""Seeded synthetic mechanisms and grid-point states for benchmarks and tests.

A synthetic mechanism keeps real species (composition, NASA-7 thermo perturbed by up to
+/-5% per coefficient) and draws element-balanced reactions among them:
log10 A ~ U(6, 16), beta ~ U(-2, 3), Ea ~ U(0, 60) kcal/mol, a fixed number of Troe /
Lindemann / SRI falloff and third-body reactions with random efficiencies in
[0.5, 12].  Most bimolecular reactions pair an arbitrary species with one of the small
radicals/molecules of the set, the shape real combustion mechanisms have, and every
reaction is written in its exergonic direction at 1000 K, as mechanisms are.  The result is
written as Chemkin text and parsed back, so the mechanism used downstream is exactly
what the Chemkin reader produces from the file.

States follow one recipe everywhere: ``rng = default_rng(seed)``; ``T ~ U(Tlo, Thi)``;
``P = 1 atm * exp(U(ln Plo, ln Phi))`` in dyn/cm^2; ``Y ~ Dirichlet(0.5)`` renormalised.
"""
from __future__ import annotations

import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .mechanism import Mechanism, Nasa7, Reaction, Species, parse_chemkin, write_chemkin

KCAL = 1000.0


def make_states(n_species: int, ng: int, seed: int = 0, T_range=(300.0, 2500.0),
                P_atm=(1.0, 40.0)):
    """Seeded states: returns P[ng], T[ng], Y[Ns, ng] (species-major), all float64."""
    rng = np.random.default_rng(seed)
    T = rng.uniform(T_range[0], T_range[1], ng)
    P = 1013250.0 * np.exp(rng.uniform(math.log(P_atm[0]), math.log(P_atm[1]), ng))
    Y = rng.dirichlet(0.5 * np.ones(n_species), ng)
    Y /= Y.sum(axis=1, keepdims=True)
    return P, T, np.ascontiguousarray(Y.T)


def ragged_patches(n_patches: int = 64, lo: int = 1 << 8, hi: int = 1 << 17, total: int = 1 << 22,
                   seed: int = 0) -> np.ndarray:
    """Patch offsets [n_patches + 1] of a ragged batch: sizes drawn log-uniform in [lo, hi],
    then rescaled and rounded so that they sum to exactly ``total`` (the remainder of the
    rounding goes to the largest patches, one point each)."""
    rng = np.random.default_rng(seed)
    raw = np.exp(rng.uniform(math.log(lo), math.log(hi), n_patches))
    sizes = np.floor(raw * (total / raw.sum())).astype(np.int64)
    rest = int(total - sizes.sum())
    order = np.argsort(-sizes, kind="stable")
    sizes[order[:rest]] += 1
    return np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)


def _perturb_thermo(th: Nasa7, rng, frac: float) -> Nasa7:
    f = lambda c: tuple(float(x * (1.0 + rng.uniform(-frac, frac))) for x in c)
    return Nasa7(th.t_min, th.t_mid, th.t_max, f(th.low), f(th.high))


def synthetic_mechanism(base_species: Sequence[Species], elements: Sequence[str], n_reactions: int,
                        n_troe: int, n_lindemann: int, n_sri: int, n_third_body: int,
                        seed: int, name: str, frac_irreversible: float = 0.05,
                        thermo_perturb: float = 0.05, n_small: int = 12, n_plog: int = 0) -> Mechanism:
    rng = np.random.default_rng(seed)
    els = list(elements)
    species = [Species(s.name, dict(s.composition), _perturb_thermo(s.thermo, rng, thermo_perturb))
               for s in base_species]
    ns = len(species)
    comp = np.array([[s.composition.get(e, 0.0) for e in els] for s in species], dtype=np.int64)
    key = lambda v: tuple(int(x) for x in v)
    by_comp: Dict[tuple, List[int]] = {}
    for k in range(ns):
        by_comp.setdefault(key(comp[k]), []).append(k)
    pair_sum: Dict[tuple, List[Tuple[int, int]]] = {}
    for i in range(ns):
        for j in range(i, ns):
            pair_sum.setdefault(key(comp[i] + comp[j]), []).append((i, j))
    natoms = comp.sum(axis=1)
    small = [int(k) for k in np.argsort(natoms, kind="stable")[:n_small] if natoms[k] > 0]
    splittable = [k for k in range(ns) if len(pair_sum.get(key(comp[k]), [])) > 0]
    names = [s.name for s in species]

    def arrh():
        return (10.0 ** rng.uniform(6.0, 16.0), float(rng.uniform(-2.0, 3.0)),
                float(rng.uniform(0.0, 60.0)) * KCAL)

    def effs():
        n = int(rng.integers(3, 9))
        pick = rng.choice(ns, size=min(n, ns), replace=False)
        return {names[int(k)]: float(rng.uniform(0.5, 12.0)) - 1.0 for k in pick}

    def st(idx):
        d: Dict[str, float] = {}
        for k in idx:
            d[names[k]] = d.get(names[k], 0.0) + 1.0
        return d

    def assoc():
        """A + B <=> AB (randomly oriented)."""
        for _ in range(10000):
            k = splittable[int(rng.integers(len(splittable)))]
            pairs = pair_sum[key(comp[k])]
            i, j = pairs[int(rng.integers(len(pairs)))]
            if k in (i, j):
                continue
            if rng.uniform() < 0.5:
                return [i, j], [k]
            return [k], [i, j]
        raise RuntimeError("no association reaction found")

    def bimol():
        for _ in range(100000):
            i = int(rng.integers(ns))
            j = small[int(rng.integers(len(small)))] if rng.uniform() < 0.7 else int(rng.integers(ns))
            cands = [p for p in pair_sum.get(key(comp[i] + comp[j]), []) if sorted(p) != sorted((i, j))]
            if not cands:
                continue
            p = cands[int(rng.integers(len(cands)))]
            return [i, j], list(p)
        raise RuntimeError("no bimolecular reaction found")

    def smh1000(k):
        a = species[k].thermo.high if 1000.0 >= species[k].thermo.t_mid else species[k].thermo.low
        T = 1000.0
        return (a[6] - a[0] + a[0] * math.log(T) + a[1] / 2 * T + a[2] / 6 * T ** 2
                + a[3] / 12 * T ** 3 + a[4] / 20 * T ** 4 - a[5] / T)

    def orient(r_idx, p_idx):
        """Write the reaction in its exergonic direction at 1000 K, as mechanisms are."""
        d = sum(smh1000(k) for k in p_idx) - sum(smh1000(k) for k in r_idx)
        return (r_idx, p_idx) if d >= 0.0 else (p_idx, r_idx)

    kinds = (["troe"] * n_troe + ["lindemann"] * n_lindemann + ["sri"] * n_sri
             + ["third_body"] * n_third_body + ["plog"] * n_plog)
    kinds += ["standard"] * (n_reactions - len(kinds))
    order = rng.permutation(len(kinds))
    reactions: List[Reaction] = []
    for o in order:
        kind = kinds[int(o)]
        rev = bool(rng.uniform() >= frac_irreversible)
        if kind == "standard":
            r_idx, p_idx = orient(*(bimol() if rng.uniform() < 0.85 else assoc()))
            r = Reaction("", st(r_idx), st(p_idx), "standard", rev, arrh())
        elif kind == "plog":
            # 3-6 pressure points log-uniform in 0.05-100 atm, one or two Arrhenius terms each
            # (a PLOG sum), drawn around one base rate so that k(P) varies smoothly
            r_idx, p_idx = orient(*bimol())
            A, b, Ea = arrh()
            pts = np.sort(10.0 ** rng.uniform(math.log10(0.05), 2.0, int(rng.integers(3, 7)))) * 1013250.0
            plog = []
            for pp in pts:
                nr = 1 if rng.uniform() < 0.7 else 2
                for _ in range(nr):
                    plog.append((float(pp), A * 10.0 ** rng.uniform(-1.0, 1.0) / nr,
                                 b + float(rng.uniform(-0.3, 0.3)), Ea * float(rng.uniform(0.8, 1.2))))
            r = Reaction("", st(r_idx), st(p_idx), "plog", rev, plog[0][1:], plog=plog)
        elif kind == "third_body":
            r_idx, p_idx = orient(*assoc())
            r = Reaction("", st(r_idx), st(p_idx), "third_body", rev, arrh())
            r.third_body = effs()
        else:
            r_idx, p_idx = orient(*assoc())
            A, b, Ea = arrh()
            low = (A * 10.0 ** rng.uniform(3.0, 8.0), float(rng.uniform(-3.0, 0.0)),
                   Ea * float(rng.uniform(0.0, 1.0)))
            r = Reaction("", st(r_idx), st(p_idx), kind, rev, (A, b, Ea), low=low)
            r.third_body = effs()
            if kind == "troe":
                tp = (float(rng.uniform(0.2, 0.95)), 10.0 ** rng.uniform(1.5, 3.0),
                      10.0 ** rng.uniform(2.5, 4.0))
                if rng.uniform() < 0.5:
                    tp = tp + (10.0 ** rng.uniform(3.0, 4.0),)
                r.troe = tp
            elif kind == "sri":
                r.sri = (float(rng.uniform(0.1, 1.0)), float(rng.uniform(50.0, 500.0)),
                         10.0 ** rng.uniform(2.5, 3.5), float(rng.uniform(0.5, 2.0)),
                         float(rng.uniform(-0.5, 0.5)))
        reactions.append(r)
    return Mechanism(name, els, species, reactions)


def with_extended_reactions(mech: Mechanism, n: int, seed: int, name: str) -> Mechanism:
    """``mech`` plus ``n`` element-balanced reactions of the shapes real mechanisms hold beyond
    three species per side and integer coefficients: the sum of two of its reactions (sides
    of four or more species; repeated species merge into one coefficient) or half of one
    (coefficients 0.5), cycling over the kinds of the reaction taken first (elementary,
    third-body, falloff, PLOG) so that each kind appears with such sides."""
    import copy
    rng = np.random.default_rng(seed)
    base = list(mech.reactions)
    by_kind: Dict[str, List[Reaction]] = {}
    for r in base:
        by_kind.setdefault(r.rtype, []).append(r)
    kinds = [k for k in ("standard", "third_body", "troe", "plog", "lindemann", "sri") if k in by_kind]
    simple = by_kind["standard"]

    def add(a: Dict[str, float], b: Dict[str, float], f: float = 1.0) -> Dict[str, float]:
        out = dict(a)
        for s2, nu in b.items():
            out[s2] = out.get(s2, 0.0) + f * nu
        return out

    extra: List[Reaction] = []
    for i in range(n):
        r = copy.deepcopy(by_kind[kinds[i % len(kinds)]][int(rng.integers(len(by_kind[kinds[i % len(kinds)]])))])
        if i % 2 == 0:
            o = simple[int(rng.integers(len(simple)))]
            r.reactants, r.products = add(r.reactants, o.reactants), add(r.products, o.products)
        else:
            r.reactants = {k: 0.5 * v for k, v in r.reactants.items()}
            r.products = {k: 0.5 * v for k, v in r.products.items()}
        r.reversible = bool(i % 3 != 2)
        r.duplicate = True
        extra.append(r)
    return Mechanism(name, list(mech.elements), list(mech.species), base + extra)


def with_long_efficiency_lists(mech: Mechanism, lo: int, hi: int, seed: int, name: str) -> Mechanism:
    """``mech`` with every third-body and falloff reaction given between ``lo`` and ``hi``
    collision efficiencies (eff - 1 in [-0.7, 5], as the parser stores them) on species drawn
    at random: efficiency lists longer than a lane group of 8 lanes, which the broadcast table
    kernel of 8-point tiles loads in two parts."""
    import copy
    rng = np.random.default_rng(seed)
    names = [sp.name for sp in mech.species]
    out: List[Reaction] = []
    for r in mech.reactions:
        r = copy.deepcopy(r)
        if r.rtype in ("third_body", "troe", "lindemann", "sri"):
            k = int(rng.integers(lo, hi + 1))
            pick = rng.choice(len(names), size=k, replace=False)
            r.third_body = {names[i]: float(np.round(rng.uniform(-0.7, 5.0), 3)) for i in sorted(pick)}
        out.append(r)
    return Mechanism(name, list(mech.elements), list(mech.species), out)


def write_and_parse(mech: Mechanism, path: str) -> Mechanism:
    write_chemkin(mech, path)
    return parse_chemkin(path, name=mech.name)


# sizes of the benchmark mechanisms: (species source, reactions, troe, lindemann, sri,
# third-body, seed)
SPECS = {
    "synth_gri53": dict(n_reactions=325, n_troe=25, n_lindemann=4, n_sri=0, n_third_body=40, seed=53),
    "synth_nhept88": dict(n_reactions=387, n_troe=33, n_lindemann=6, n_sri=0, n_third_body=20, seed=88),
    "synth_jetsurf348": dict(n_reactions=2163, n_troe=150, n_lindemann=0, n_sri=23, n_third_body=324,
                             seed=348),
    "synth_plog53": dict(n_reactions=120, n_troe=8, n_lindemann=0, n_sri=0, n_third_body=12, n_plog=30,
                         seed=531),
}
