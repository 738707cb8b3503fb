"""The table kernel's reaction table decodes back to the mechanism (host-side check of the
format the kernel reads)."""
from collections import Counter

import pytest

from chemgen_amd.autotune import table_supported
from chemgen_amd.codegen import TABLE_LONG, TABLE_NEG, build_coef, negative_a, table_class, table_layout, table_records
from chemgen_amd.kernels import load_bundled


@pytest.mark.parametrize("name", ["h2_burke", "gri30", "synth_gri53", "synth_jetsurf348", "synth_plog53",
                                  "synth_ext53", "fuller2021", "hashemi2016"])
@pytest.mark.parametrize("npt,min_class", [(16, 64), (64, 0), (16, 1000), (2, 8)])
def test_table_decodes_to_mechanism(name, npt, min_class):
    m = load_bundled(name)
    if m.n_species * npt * 8 >= 1 << 16:
        pytest.skip("LDS row byte offsets beyond 16 bits")
    lay = table_layout(m, npt, min_class)
    rows, effsp, order, segs, gen_start = lay.rows, lay.effsp, lay.order, lay.segs, lay.gen_start
    assert table_records(m, npt, min_class) == (rows, effsp, order, segs, gen_start)
    _, _, roffs = build_coef(m)
    pad = m.n_species * npt * 8       # slots: LDS byte offsets of the species rows
    rowb = npt * 8
    idx = {s.name: i for i, s in enumerate(m.species)}
    assert sorted(order) == list(range(m.n_reactions))
    kinds = {"standard": {0}, "third_body": {1}, "lindemann": {2}, "troe": {3, 4}, "sri": {5, 6}, "plog": {7}}
    n_long = 0
    for n, j in enumerate(order):
        r = m.reactions[j]
        (y, z, w, x2), (x, e0, e1, xo) = rows[2 * n], rows[2 * n + 1]
        assert (x & 7) in kinds[r.rtype]
        assert ((x >> 3) & 1) == int(r.reversible)
        nrs, nps, ne, dnu = (x >> 4) & 3, (x >> 6) & 3, (x >> 8) & 15, ((x >> 24) & 15) - 8
        lng = bool(x & TABLE_LONG)
        assert bool(x & TABLE_NEG) == negative_a(r)
        assert y == 8 * roffs[j]
        slots_r = [z & 0xFFFF, z >> 16, w & 0xFFFF]
        slots_p = [w >> 16, x2 & 0xFFFF, x2 >> 16]
        if lng:
            # extended record: entries with their coefficients, dnu in xnu
            n_long += 1
            assert nrs == nps == 0 and slots_r + slots_p == [pad] * 6 and xo > 0
            a, b = lay.xsl[xo], lay.xsl[xo + 1]
            ent = list(zip(lay.xsl[xo + 2:xo + 2 + a + b], lay.xnu[xo + 2:xo + 2 + a + b]))
            assert all(s % npt == 0 for s, _ in ent)
            assert {s // npt: nu for s, nu in ent[:a]} == {idx[k]: v for k, v in r.reactants.items()}
            assert {s // npt: nu for s, nu in ent[a:]} == {idx[k]: v for k, v in r.products.items()}
            assert lay.xnu[xo] == r.coeff_sum
            assert any(not float(v).is_integer() for v in list(r.reactants.values()) + list(r.products.values())) \
                or max(sum(r.reactants.values()), sum(r.products.values())) > 3
        else:
            assert xo == 0
            for slots, count, side in ((slots_r, nrs, r.reactants), (slots_p, nps, r.products)):
                assert all(s % rowb == 0 for s in slots)
                assert slots[count:] == [pad] * (3 - count)
                assert Counter(s // rowb for s in slots[:count]) == Counter({idx[k]: int(v) for k, v in side.items()})
            assert dnu == int(r.coeff_sum)
        assert ne == len(r.third_body)
        if r.rtype == "plog":
            # pressure points: blob offsets walk the record, rates per point as in the mechanism
            pts = sorted(set(c[0] for c in r.plog))
            assert e1 == len(pts)
            o = roffs[j]
            for q, p in enumerate(pts):
                nrates = len([c for c in r.plog if c[0] == p])
                assert lay.plog[e0 + q] == (o, nrates)
                o += 3 + 4 * nrates
        else:
            assert effsp[e0:e0 + ne] == [idx[k] for k in r.third_body]
    assert n_long == sum(1 for r in m.reactions
                         if any(not float(v).is_integer() for v in list(r.reactants.values()) + list(r.products.values()))
                         or max(sum(r.reactants.values()), sum(r.products.values())) > 3)
    if name in ("synth_ext53", "fuller2021"):
        assert n_long >= 12
    # class segments: contiguous, one shape each, at least min_class long, then the generic rest;
    # extended records never get a loop of their own
    pos = 0
    for a, b, key in segs:
        assert a == pos and b - a >= max(1, min_class) and key[2] > 0 and not key[5]
        assert all(table_class(m, order[n]) == key for n in range(a, b))
        pos = b
    assert gen_start == pos
    seg_keys = {key for _, _, key in segs}
    assert all(table_class(m, order[n]) not in seg_keys for n in range(gen_start, m.n_reactions))
    longs = [bool(rows[2 * n + 1][0] & TABLE_LONG) for n in range(gen_start, m.n_reactions)]
    assert longs == sorted(longs)     # extended records last in the generic segment


def test_every_bundled_mechanism_is_tabled():
    for name in ("synth_plog53", "synth_ext53", "synth_jetsurf348", "fuller2021"):
        assert table_supported(load_bundled(name)), name


def test_default_variant_fits_the_lds():
    """Large mechanisms get the longest tile of at most 16 points whose state fits 160 KB."""
    from chemgen_amd.autotune import LDS_BYTES, default_variant, table_lds_bytes
    for name, npt in (("synth_jetsurf348", 16), ("fuller2021", 8)):
        v = default_variant(load_bundled(name))
        assert v.table == npt, (name, v)
        assert table_lds_bytes(load_bundled(name).n_species, v.table, v.split) <= LDS_BYTES


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


def test_negative_prefactor_compiled_in():
    """Negative prefactors (Hashemi 2016: 22 duplicates) keep log2 |A| in the blob and their sign
    in the code: the straight-line kernel negates exactly those rates, the table kernel marks
    them with a class bit and keeps them out of the compiled class loops."""
    import numpy as np
    from chemgen_amd.codegen import Variant, emit_hip
    m = load_bundled("hashemi2016")
    neg = [j for j, r in enumerate(m.reactions) if negative_a(r)]
    assert len(neg) == 22
    coef, _, roffs = build_coef(m)
    assert np.isfinite(coef).all()
    for j in neg:
        assert coef[roffs[j]] == pytest.approx(np.log2(abs(m.reactions[j].arrh[0])), rel=1e-15)
    src = emit_hip(m, Variant())[0]
    assert src.count("const double kf = -cg_exp2_fin(lk);") == 22
    lay = table_layout(m, 16, 8)
    pos = {j: n for n, j in enumerate(lay.order)}
    assert all(lay.rows[2 * pos[j] + 1][0] & TABLE_NEG for j in neg)
    assert all(pos[j] >= lay.gen_start for j in neg)


def test_generator_level3_short_tiles():
    """--parallel-level 3 takes tiles down to one point (what 500+-species mechanisms need) and
    refuses a tile whose species state does not fit the LDS."""
    from argparse import Namespace
    from chemgen_amd.generate_chemistry import variant_for
    m = load_bundled("fuller2021")
    v = variant_for(Namespace(parallel_level=3, veclen=8, nreact_per_block=16), m)
    assert (v.table, v.split) == (8, 16)
    assert variant_for(Namespace(parallel_level=3, veclen=1, nreact_per_block=4), m).table == 1
    with pytest.raises(SystemExit):
        variant_for(Namespace(parallel_level=3, veclen=16, nreact_per_block=16), m)


@pytest.mark.parametrize("name,npt,W", [("synth_jetsurf348", 16, 16), ("fuller2021", 8, 16), ("hashemi2016", 16, 16),
                                        ("synth_ext53", 16, 4), ("gri30", 64, 8), ("synth_plog53", 2, 16)])
def test_chunks_and_incidence_give_the_stoichiometry(name, npt, W):
    """The table kernel's chunks cover the table in order, each within the LDS buffer, as
    few as the buffer allows (ceil(reactions / rc)); and its per-chunk species lists, applied to rates of
    progress, give wdot_k = sum_j (nu''_kj - nu'_kj) q_j of the mechanism's own coefficients
    (a dense product over the reactions in mechanism order)."""
    import numpy as np
    from chemgen_amd.codegen import table_gather_plan
    m = load_bundled(name)
    plan = table_gather_plan(m, npt, W)
    lay, rc, slots, spans, chunks = plan["layout"], plan["rc"], plan["slots"], plan["spans"], plan["chunks"]
    assert chunks[0][0] == 0 and chunks[-1][1] == m.n_reactions
    assert len(chunks) == -(-m.n_reactions // rc)
    for (c0, c1, pieces), nxt in zip(chunks, chunks[1:] + [(m.n_reactions, None, None)]):
        assert c1 == nxt[0] and 0 < c1 - c0 <= rc
        assert pieces[0][0] == c0 and pieces[-1][1] == c1
        for (a, b, si), after in zip(pieces, pieces[1:] + [(c1, None, None)]):
            sa, sb = spans[si]
            assert sa <= a < b <= sb and b == after[0]
    assert plan["lds"] <= 160 * 1024
    # the lists applied to random rates of progress
    rng = np.random.default_rng(0)
    q = rng.normal(size=(m.n_reactions, npt))             # in table order
    w = np.zeros((m.n_species, npt))
    goff, grow, gnu = plan["goff"], plan["grow"], plan["gnu"]
    for c, (c0, c1, _) in enumerate(chunks):
        buf = q[c0:c1].reshape(-1)                          # the chunk's LDS buffer [row][pt]
        for k in range(m.n_species):
            for e in range(goff[c * (m.n_species + 1) + k], goff[c * (m.n_species + 1) + k + 1]):
                assert grow[e] % npt == 0 and grow[e] < (c1 - c0) * npt
                w[k] += gnu[e] * buf[grow[e]:grow[e] + npt]
    N = _net_matrix(m)
    q_mech = np.empty_like(q)
    q_mech[lay.order] = q                                   # table position n holds reaction order[n]
    assert np.allclose(w, N @ q_mech, rtol=1e-12, atol=1e-12)
    # the kernel's gather loop, run on the host-built schedule
    wk = _emulate_gather(plan, m.n_species, npt, W, q)
    assert np.allclose(wk, N @ q_mech, rtol=1e-12, atol=1e-12)


def _net_matrix(m):
    import numpy as np
    idx = {s.name: i for i, s in enumerate(m.species)}
    N = np.zeros((m.n_species, m.n_reactions))
    for j, r in enumerate(m.reactions):
        for sp, nu in r.reactants.items():
            N[idx[sp], j] -= nu
        for sp, nu in r.products.items():
            N[idx[sp], j] += nu
    return N


def _emulate_gather(plan, ns, npt, W, q):
    """The table kernel's gather on the CPU, as its generated code runs it: in chunk c every
    lane group (w, g) sums its stream slot by slot (common slot lengths), a split slot's parts
    are summed over the wavefront's groups, and each slot's writer stores (first chunk) or
    adds.  Every species must be stored exactly once, before any add."""
    import struct
    import numpy as np
    from chemgen_amd.codegen import _gather_order
    st = plan["streams"]
    G, rc = st["G"], plan["rc"]
    assert G * npt == 64
    out = np.full((ns, npt), np.nan)
    stored = set()
    for c, (c0, c1, _) in enumerate(plan["chunks"]):
        rec = st["chunks"][c]
        lens, steps = rec["lens"], rec["steps"]
        assert steps % 4 == 0 and steps >= sum(lens)
        buf = np.zeros((rc + 1) * npt)
        buf[:(c1 - c0) * npt] = q[c0:c1].reshape(-1)
        for w in range(W):
            cm = st["comb"][rec["comb_base"] + w]
            a = np.zeros((len(lens), G, npt))
            order = _gather_order(lens)
            for g in range(G):
                base = rec["str_base"] + (w * G + g) * steps
                for t, (s_, _) in enumerate(order):
                    if True:
                        e = st["stream"][base + t]
                        o = (e & 0x3ffff) // 8
                        assert (e & 0x3ffff) % (8 * npt) == 0 and o <= rc * npt
                        if st["fnu"] is not None:
                            nu = st["fnu"][base + t]
                        else:                                    # the double {0, bits 18-31}
                            nu = struct.unpack("<d", struct.pack("<II", 0, e & 0xfffc0000))[0]
                        a[s_, g] = nu * buf[o:o + npt] + a[s_, g]
                for e in st["stream"][base + len(order):base + steps]:    # padding after the last slot
                    assert e == rc * npt * 8
            for s_ in range(len(lens)):
                if cm & (1 << s_):
                    a[s_, :] = a[s_].sum(axis=0)
                for g in range(G):
                    k = st["spk"][rec["spk_base"] + (w * G + g) * len(lens) + s_]
                    if k < 0:
                        continue
                    sp = k & 0xffffff
                    if k & (1 << 30):
                        assert sp not in stored
                        stored.add(sp)
                        out[sp] = a[s_, g]
                    else:
                        assert sp in stored
                        out[sp] = out[sp] + a[s_, g]
    assert stored == set(range(ns))
    return out
