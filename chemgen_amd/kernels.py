"""Bundled mechanisms and the kernel variants compiled for them."""
from __future__ import annotations

import functools
import os
from typing import List

from .codegen import Variant
from .mechanism import Mechanism, parse_chemkin

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


@functools.lru_cache(maxsize=None)
def load_bundled(name: str) -> Mechanism:
    """A mechanism shipped in chemgen_amd/data (<name>.inp, thermo inline)."""
    return parse_chemkin(os.path.join(DATA_DIR, name + ".inp"), name=name)


TUNED = ("synth_gri53", "synth_nhept88", "gri30", "synth_jetsurf348", "h2_burke", "fuller2021", "ic8_llnl")

# table-kernel shapes compiled for small mechanisms whose tuned kernel is another family, so
# that the table format's PLOG, extended and negative-prefactor records are exercised: 16-point tiles of 4
# wavefronts (compiled class loops and the generic loop) and 4-point tiles of 8
TABLE_CHECKED = {
    "synth_plog53": (Variant(64, 1, "lean", "gh", "horner", False, 4, 4, 16),
                     Variant(64, 2, "lean", "gh", "horner", False, 4, 8, 4)),
    "synth_ext53": (Variant(64, 1, "lean", "gh", "horner", False, 4, 4, 16),
                    Variant(64, 2, "lean", "gh", "horner", False, 4, 8, 4)),
    "hashemi2016": (Variant(64, 1, "lean", "gh", "horner", False, 4, 4, 16),
                    Variant(64, 2, "lean", "gh", "horner", False, 4, 8, 4)),
    # compiled third-body and Troe rounds with 9-15 efficiencies: two loads per 8-lane group
    "synth_eff53": (Variant(64, 1, "lean", "gh", "horner", False, 4, 4, 8, fdiv=True, bcast=True),
                    Variant(64, 2, "lean", "gh", "horner", False, 4, 8, 16, fdiv=True, bcast=True)),
}


def default_variants(mech: Mechanism) -> List[Variant]:
    """Variants compiled by the build: the tuned (or default) shape first, then the
    autotuner's candidates for the mechanisms it sweeps."""
    from .autotune import best_variant, candidate_variants
    from .autotune import STRAIGHT_LINE_MAX_SPECIES
    out = [best_variant(mech)]
    if mech.name in TUNED:
        out += [v for v in candidate_variants(mech) if v not in out]
    live = Variant(64, 1, "lean", "gh", "horner", False, 4, 1, 0, False, True)
    if mech.n_species <= STRAIGHT_LINE_MAX_SPECIES and live not in out:
        out.append(live)   # the live-range-ordered lane kernel, for its parity tests
    out += [v for v in TABLE_CHECKED.get(mech.name, ()) if v not in out]
    return out
