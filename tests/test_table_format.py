"""The table kernel's reaction table decodes back to the mechanism (host-side check of the
format the kernel reads)."""
from collections import Counter

import pytest

from chemgen_amd.autotune import table_supported
from chemgen_amd.codegen import build_coef, table_class, table_records
from chemgen_amd.kernels import load_bundled


@pytest.mark.parametrize("name", ["h2_burke", "gri30", "synth_gri53", "synth_jetsurf348"])
@pytest.mark.parametrize("npt,min_class", [(16, 64), (64, 0), (16, 1000)])
def test_table_decodes_to_mechanism(name, npt, min_class):
    m = load_bundled(name)
    rows, effsp, order, segs, gen_start = table_records(m, npt, min_class)
    _, _, roffs = build_coef(m)
    pad = m.n_species * npt
    idx = {s.name: i for i, s in enumerate(m.species)}
    assert sorted(order) == list(range(m.n_reactions))
    kinds = {"standard": {0}, "third_body": {1}, "lindemann": {2}, "troe": {3, 4}, "sri": {5, 6}}
    for n, j in enumerate(order):
        r = m.reactions[j]
        (y, z, w, x2), (x, e0, e1, _) = rows[2 * n], rows[2 * n + 1]
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
        assert effsp[e0:e0 + ne] == [idx[k] for k in r.third_body]
    # class segments: contiguous, one shape each, at least min_class long, then the generic rest
    pos = 0
    for a, b, key in segs:
        assert a == pos and b - a >= max(1, min_class)
        assert all(table_class(m, order[n]) == key for n in range(a, b))
        pos = b
    assert gen_start == pos
    seg_keys = {key for _, _, key in segs}
    assert all(table_class(m, order[n]) not in seg_keys for n in range(gen_start, m.n_reactions))


def test_plog_mechanism_is_not_tabled():
    assert not table_supported(load_bundled("synth_plog53"))
    assert table_supported(load_bundled("synth_jetsurf348"))


@pytest.mark.parametrize("name", ["h2_burke", "gri30", "synth_gri53", "synth_plog53"])
def test_live_order_is_a_permutation_with_its_peak(name):
    """The live-range order visits every reaction once, and the peak it reports is the largest
    set of species between their first and last reaction, recounted here from first/last
    positions; for GRI-Mech 3.0 it keeps fewer species live than the file order."""
    from chemgen_amd.codegen import live_order
    m = load_bundled(name)
    order, peak = live_order(m)
    assert sorted(order) == list(range(m.n_reactions))
    idx = {s.name: i for i, s in enumerate(m.species)}
    first, last = {}, {}
    for pos, j in enumerate(order):
        r = m.reactions[j]
        for s in list(r.reactants) + list(r.products) + list(r.third_body):
            first.setdefault(idx[s], pos)
            last[idx[s]] = pos
    recount = max(sum(1 for k in first if first[k] <= pos <= last[k]) for pos in range(len(order)))
    assert recount == peak
    if name == "gri30":
        fo = {}
        lo = {}
        for pos, r in enumerate(m.reactions):
            for s in list(r.reactants) + list(r.products) + list(r.third_body):
                fo.setdefault(idx[s], pos)
                lo[idx[s]] = pos
        file_peak = max(sum(1 for k in fo if fo[k] <= pos <= lo[k]) for pos in range(m.n_reactions))
        assert peak < file_peak
