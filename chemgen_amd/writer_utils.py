"""Small emission helpers for the HIP target (the reference keeps its Fortran-array
formatting helpers in chemgen/writer_utils.py:5-41; these serve the C++/HIP text).
"""
from __future__ import annotations

from typing import Iterable, List, Sequence


def dlit(x: float) -> str:
    """An exactly round-tripping C double literal."""
    x = float(x)
    if x != x:
        return "__builtin_nan(\"\")"
    if x in (float("inf"), float("-inf")):
        return "__builtin_inf()" if x > 0 else "(-__builtin_inf())"
    r = repr(x)
    if "e" not in r and "." not in r and "inf" not in r:
        r += ".0"
    return r


def tree_sum(terms: Sequence[str]) -> str:
    """Balanced-tree sum of expressions (shorter dependency chains than a left fold)."""
    terms = list(terms)
    if not terms:
        return "0.0"
    while len(terms) > 1:
        nxt = []
        for i in range(0, len(terms) - 1, 2):
            nxt.append(f"({terms[i]} + {terms[i + 1]})")
        if len(terms) % 2:
            nxt.append(terms[-1])
        terms = nxt
    return terms[0]


def power_expr(var: str, nu: float) -> str:
    """var**nu for a stoichiometric coefficient: products for small integers."""
    if nu == 1.0:
        return var
    if float(nu).is_integer() and 2 <= nu <= 4:
        return "*".join([var] * int(nu))
    return f"cg_pow({var}, {dlit(nu)})"


def stoich_product(terms: Iterable[str]) -> str:
    t = list(terms)
    return " * ".join(t) if t else "1.0"


def chunk_lines(items: List[str], per_line: int, indent: str = "    ") -> List[str]:
    out = []
    for i in range(0, len(items), per_line):
        out.append(indent + ", ".join(items[i:i + per_line]) + ",")
    return out
