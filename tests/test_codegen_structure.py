"""What a compiled code object depends on: the mechanism's structure, not its coefficient values.

One code object serves every coefficient set of its structure (cg_mech_set_coef, the
generated module's blob), so nothing value-dependent may be compiled into the code unless
the structure hash records it.  SRI's T^e factor is compiled whenever the SRI record has five
parameters (e = 0 included), so e = 0 and e = 0.2 give the same hash and the same source.
The one-shot API caches loaded mechanisms per (name, structure, blob).
"""
import dataclasses

from chemgen_amd.api import cache_key
from chemgen_amd.codegen import Variant, build_coef, emit_hip, structure_hash
from chemgen_amd.kernels import load_bundled
from chemgen_amd.mechanism import Mechanism, Reaction


def _sri_mech(e, A=2.0e12):
    base = load_bundled("h2_burke")
    sp = [s for s in base.species if s.name in ("H", "O2", "HO2", "N2")]
    r = Reaction("", {"H": 1.0, "O2": 1.0}, {"HO2": 1.0}, "sri", True, (A, 0.5, 300.0),
                 low=(5.0e17, -1.0, 0.0), sri=(0.6, 300.0, 900.0, 1.3, e))
    return Mechanism("sri_t", ["H", "O", "N"], sp, [r])


def test_sri_e_is_a_coefficient():
    m0, m1 = _sri_mech(0.0), _sri_mech(0.2)
    assert structure_hash(m0) == structure_hash(m1)
    for v in (Variant(), Variant(64, 1, "lean", "gh", "horner", False, 4, 1, 0, False, True),
              Variant(64, 2, "lean", "gh", "horner", False, 4, 2, 0),
              Variant(64, 1, "lean", "gh", "horner", False, 4, 2, 4)):
        s0, s1 = emit_hip(m0, v)[0], emit_hip(m1, v)[0]
        assert s0 == s1, v.tag
    c0, c1 = build_coef(m0)[0], build_coef(m1)[0]
    assert len(c0) == len(c1) and (c0 != c1).sum() == 1


def test_cache_key_separates_coefficient_values():
    a, b = _sri_mech(0.2), _sri_mech(0.2, A=3.0e12)
    assert structure_hash(a) == structure_hash(b)
    assert cache_key(a) != cache_key(b)
    assert cache_key(a) == cache_key(_sri_mech(0.2))


def test_fdiv_variant_divides_with_one_newton_step():
    """The ``fdiv`` option (every tuner candidate) divides with the reciprocal estimate and one
    Newton step in log2, h_k = C_k / EG_k and the falloff blending, in every kernel family; the
    default variant keeps the corrected quotient; the option is part of the code-object tag."""
    from chemgen_amd.autotune import candidate_variants
    from chemgen_amd.codegen import Variant, emit_hip
    from chemgen_amd.kernels import load_bundled
    m = load_bundled("gri30")   # Troe (with and without T2) and Lindemann reactions
    assert all(v.fdiv for v in candidate_variants(m))
    for v in (Variant(), Variant(64, 2, "lean", "gh", "horner", False, 28, 4),
              Variant(64, 1, "lean", "gh", "horner", False, 4, 1, 0, False, True),
              Variant(64, 4, "lean", "gh", "horner", False, 4, 16, 16)):
        slow, _, _ = emit_hip(m, v)
        fast, _, _ = emit_hip(m, dataclasses.replace(v, fdiv=True))
        assert "cg_div_fast(m - 1.0, m + 1.0)" in fast and "cg_div(m - 1.0, m + 1.0)" in slow
        assert fast.count("cg_div_fast(") > slow.count("cg_div_fast(") + 3, v.tag
        assert dataclasses.replace(v, fdiv=True).tag == v.tag + "d"


def test_bcast_table_kernel_reads_each_word_from_its_lane():
    """The broadcast table kernel (16- and 8-point tiles): in the gather, lane pt of a group
    loads word B b + pt of block b and step t takes lane t mod B of block t div B, so step t
    reads stream word t; in the class rounds, lane pt loads the record of round n0 + pt and
    round n takes lane n - n0.  Checked on the generated source of both tile lengths."""
    import re
    from chemgen_amd.autotune import candidate_variants
    m = load_bundled("synth_jetsurf348")
    for npt in (16, 8):
        v = Variant(64, 4, "lean", "gh", "horner", False, 4, 16, npt, fdiv=True, bcast=True)
        assert v in candidate_variants(m) or npt == 8
        src = emit_hip(m, v)[0]
        assert "cg_bci<" in src and "0x158 + N" in src if npt == 8 else "0x150 + N" in src
        gathers = src.split("// gather of chunk ")[1:]
        assert gathers
        for g in gathers:
            g = g.split("\n  }\n")[0]
            for b, off in re.findall(r"const unsigned u(\d+) = st\[(\d+)\];", g):
                assert int(off) == npt * int(b)
            steps = [(int(t), int(u), int(n)) for t, u, n in
                     re.findall(r"const unsigned x(\d+) = CG_BCW\(u(\d+), (\d+)\);", g)]
            assert steps and [t for t, _, _ in steps] == list(range(len(steps)))
            assert all(u == t // npt and n == t % npt for t, u, n in steps)
        rounds = src.split("round(s)\n")[1:]
        assert rounds
        for r in rounds:
            blocks = re.findall(r"const int jl = min\(j0 \+ \((\d+) \+ pt\) \* CG_SLOTS", r.split("  }\n  {")[0])
            for n0 in blocks:
                body = r.split(f"const int jl = min(j0 + ({n0} + pt) * CG_SLOTS")[1]
                body = body.split("const int jl = min(")[0]
                for n, m_ in re.findall(r"const int j = j0 \+ (\d+) \* CG_SLOTS;\n\s+const uint32_t ry = CG_BCI\(lt\.y, (\d+)\)", body):
                    assert int(n) - int(n0) == int(m_) and int(m_) < npt
    # 32-point tiles (a group wider than a DPP row) keep the per-lane loads
    v32 = [v for v in candidate_variants(m) if v.table == 32]
    assert all(not v.bcast for v in v32)


def test_short_split_tiles_repeat_lanes_and_fit_two_per_cu():
    """Reaction-split tiles of tp < 64 points: lanes tp..63 work on columns 0..63-tp (same
    values, same LDS writes), the LDS rows are tp wide, the tile length in the metadata is tp,
    and two tiles of the n-heptane-sized mechanism's state fit the 160 KB of a CU at 56 points
    (not at 64)."""
    from chemgen_amd.autotune import LDS_BYTES, candidate_variants, two_tile_points
    m = load_bundled("synth_nhept88")
    assert two_tile_points(m.n_species) == 56
    assert two_tile_points(53) == 0                     # 64-point tiles already fit twice
    gran = lambda b: -(-b // 512) * 512
    assert 2 * gran(8 * 2 * 88 * 64) > LDS_BYTES >= 2 * gran(8 * 2 * 88 * 56)
    v = next(v for v in candidate_variants(m) if v.tp == 56 and v.baked and v.group == 2)
    assert v.tag == "s8w4lghhg2kdn56" and v.ppb == 56
    src, _, meta = emit_hip(m, v)
    assert "#define CG_TP 56" in src and "lane < CG_TP ? lane : lane - CG_TP" in src
    assert "* 64]" not in src and meta["ppb"] == 56 and meta["compact_lds"]
    assert "#define CG_SMEM_DOUBLES (2 * CG_NS * CG_TP)" in src
