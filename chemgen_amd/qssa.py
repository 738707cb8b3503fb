"""Quasi-steady-state (QSS) species: getrates over the transported species only.

The reference carries the scaffolding of QSS species -- ``chemistry(qssa_species=[...])``,
the reduced species list and index (chemgen/chemistry.py:30-52, :98-120), NSP_RED / NSP_SK /
NSP_QSSA in its constants header (chemgen/constants.py:485-489) -- but computes nothing with
them.  Here a mechanism with a QSS set is a :class:`QssMechanism`; its generated kernel
(``codegen._emit_qss``) takes Y and returns wdot for the transported species R (the species
not in the QSS set, in mechanism order) and, per grid point,

1. C_k for k in R (the mixture of the transported species), ctot and the third-body sums
   over R only (QSS species are present in small amounts by assumption);
2. EG for every species, h = C / EG for k in R;
3. for each reaction j, F_j = k_f prod_{R reactants} C and B_j = k_b prod_{R products} C
   (times M for three-body reactions), so that q_f = F_j C_a and q_b = B_j C_b with a and b
   the QSS species of each side (or factor 1);
4. the QSS balance sum_j nu_ij (q_f,j - q_b,j) = 0 for every QSS species i: a linear system
   A C_Q = -c, assembled in registers and solved by Gaussian elimination without pivoting
   (-A is a diagonally dominant M-matrix when the QSS species are consumed faster than they
   are made from each other, the regime the approximation is for);
5. wdot_k for k in R with those C_Q.

Linearity requires every QSS species to enter its reactions with coefficient 1 and no
reaction side to hold two QSS species; :func:`qss_errors` lists violations.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Iterable, List, Sequence, Tuple

from .mechanism import Mechanism


# QSS sets of the bundled mechanisms that the build compiles kernels for
BUNDLED_QSS = {
    "h2_burke": ("H2O2",),
    "gri30": ("CH2(S)", "CH2OH", "CH3O", "C2H3", "CH2CHO", "C", "CH", "HCCOH", "NNH", "HNO", "H2CN", "HOCN"),
}


def bundled_qss(name: str) -> "QssMechanism":
    from .kernels import load_bundled
    return with_qss(load_bundled(name), BUNDLED_QSS[name])


def qss_errors(mech: Mechanism, qss: Sequence[str]) -> List[str]:
    """Why ``qss`` is not a linear QSS set of ``mech`` (empty when it is)."""
    names = set(mech.species_names)
    out = [f"{s}: not a species" for s in qss if s not in names]
    q = set(qss)
    if len(q) != len(list(qss)):
        out.append("a QSS species is listed twice")
    consumed = set()
    for j, r in enumerate(mech.reactions):
        for side, label in ((r.reactants, "reactants"), (r.products, "products")):
            on = [(s, n) for s, n in side.items() if s in q]
            if len(on) > 1:
                out.append(f"R{j}: two QSS species among the {label}")
            for s, n in on:
                if n != 1.0:
                    out.append(f"R{j}: QSS species {s} with coefficient {n}")
        a = [s for s in r.reactants if s in q]
        b = [s for s in r.products if s in q]
        if a != b:                       # a species on both sides has no net rate here
            consumed.update(a)
            if r.reversible:
                consumed.update(b)
    # a QSS species that no reaction consumes has no diagonal in its balance row
    out += [f"{s}: never consumed" for s in qss if s in names and s not in consumed]
    return out


def pick_qss(mech: Mechanism, candidates: Iterable[str], limit: int = 16) -> List[str]:
    """Greedily take candidates (in order) that keep the set linear, up to ``limit``."""
    out: List[str] = []
    for s in candidates:
        if len(out) >= limit:
            break
        if s in mech.species_names and not qss_errors(mech, out + [s]):
            out.append(s)
    return out


@dataclass
class QssMechanism(Mechanism):
    """A mechanism with a linear QSS set (species names, in the order of the QSS unknowns)."""
    qss: Tuple[str, ...] = field(default_factory=tuple)

    @property
    def transported(self) -> List[int]:
        q = set(self.qss)
        return [k for k, s in enumerate(self.species) if s.name not in q]

    @property
    def n_transported(self) -> int:
        return self.n_species - len(self.qss)


def with_qss(mech: Mechanism, qss: Sequence[str], name: str = None) -> QssMechanism:
    """``mech`` with the QSS set ``qss`` (checked for linearity)."""
    errs = qss_errors(mech, qss)
    if errs:
        raise ValueError("not a linear QSS set: " + "; ".join(errs[:5]))
    return QssMechanism(name or f"{mech.name}_qss{len(qss)}", list(mech.elements), list(mech.species),
                        list(mech.reactions), tuple(qss))
