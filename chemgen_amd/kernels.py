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


def default_variants(mech: Mechanism) -> List[Variant]:
    """Variants compiled by the build: the default shape first."""
    return [Variant(64, 1)]
