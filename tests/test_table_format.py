"""The table kernel's reaction table decodes back to the mechanism (host-side check of the
format the kernel reads)."""
from collections import Counter

import pytest

from chemgen_amd.autotune import table_supported
from chemgen_amd.codegen import build_coef, table_records
from chemgen_amd.kernels import load_bundled


@pytest.mark.parametrize("name", ["h2_burke", "gri30", "synth_gri53", "synth_jetsurf348"])
@pytest.mark.parametrize("npt", [16, 64])
def test_table_decodes_to_mechanism(name, npt):
    m = load_bundled(name)
    rows, effsp, order = table_records(m, npt)
    _, _, roffs = build_coef(m)
    pad = m.n_species * npt
    idx = {s.name: i for i, s in enumerate(m.species)}
    assert sorted(order) == list(range(m.n_reactions))
    kinds = {"standard": {0}, "third_body": {1}, "lindemann": {2}, "troe": {3, 4}, "sri": {5, 6}}
    for n, j in enumerate(order):
        r = m.reactions[j]
        (x, y, z, w), (x2, y2, z2, _) = rows[2 * n], rows[2 * n + 1]
        assert (x & 7) in kinds[r.rtype]
        assert ((x >> 3) & 1) == int(r.reversible)
        nrs, nps, ne, dnu = (x >> 4) & 3, (x >> 6) & 3, (x >> 8) & 15, ((x >> 24) & 15) - 8
        assert y == roffs[j]
        slots_r = [z & 0xFFFF, z >> 16, w & 0xFFFF]
        slots_p = [w >> 16, x2 & 0xFFFF, x2 >> 16]
        for slots, count, side in ((slots_r, nrs, r.reactants), (slots_p, nps, r.products)):
            assert all(s % npt == 0 for s in slots)
            assert slots[count:] == [pad] * (3 - count)
            assert Counter(s // npt for s in slots[:count]) == Counter({idx[k]: int(v) for k, v in side.items()})
        assert dnu == int(r.coeff_sum)
        assert ne == len(r.third_body)
        assert effsp[y2:y2 + ne] == [idx[k] for k in r.third_body]


def test_plog_mechanism_is_not_tabled():
    assert not table_supported(load_bundled("synth_plog53"))
    assert table_supported(load_bundled("synth_jetsurf348"))
