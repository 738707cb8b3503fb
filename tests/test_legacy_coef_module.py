"""The reference's Fortran coefficient module written by this build (chemgen_amd/legacy_coef.py):
at parallel level 1 (no padding) it holds, under the reference's names and sizes, the arrays that
the oracle's restatement of write_constants.py gives (oracle/legacy_coef.py), for the bundled
mechanisms whose reaction types the reference's HIP path holds; the padded layouts of levels 2-4
are pinned by the hand-derived arrays of tests/test_legacy_levels.py."""
import re

import numpy as np
import pytest

from chemgen_amd.kernels import load_bundled
from chemgen_amd.legacy_coef import emit_coef_module
from oracle.legacy_coef import legacy_arrays


def parse_module(src):
    """{name: (fortran type, values)} from ``type :: name(n)=(/& ... /)`` blocks."""
    out = {}
    for m in re.finditer(r"^\s*(real\(kind=8\)|integer) :: (\w+)\((\d+)\)=\(/&\n(.*?)^\s*/\)", src,
                         re.S | re.M):
        ftype, name, n, body = m.group(1), m.group(2), int(m.group(3)), m.group(4)
        toks = [t.strip() for t in body.replace("&", "").replace("\n", " ").split(",") if t.strip()]
        assert len(toks) == n, name
        vals = [int(t) for t in toks] if ftype == "integer" else [float(t.replace("D", "E")) for t in toks]
        out[name] = (ftype, vals)
    return out


@pytest.mark.parametrize("name", ["h2_burke", "h2_burke2012"])
@pytest.mark.parametrize("level", [1])
def test_coef_module_matches_reference_arrays(name, level):
    m = load_bundled(name)
    src = emit_coef_module(m, level)
    assert src.startswith(f"module coef_m_v{level}\n") and f"end module coef_m_v{level}" in src
    assert max(len(l) for l in src.splitlines()) <= 132       # free-form line limit
    got = parse_module(src)
    ref = legacy_arrays(m)
    assert set(got) == set(ref)
    for k, v in ref.items():
        ftype, vals = got[k]
        assert ftype == ("integer" if v.dtype.kind == "i" else "real(kind=8)"), k
        assert np.array_equal(np.asarray(vals, dtype=v.dtype), v), k


def test_coef_module_refuses_what_the_reference_cannot_hold():
    with pytest.raises(ValueError):
        emit_coef_module(load_bundled("gri30"))          # Lindemann reactions
    with pytest.raises(ValueError):
        emit_coef_module(load_bundled("h2_burke"), 3)    # levels 2-4 need reactions per block
    with pytest.raises(ValueError):
        emit_coef_module(load_bundled("h2_burke"), 5)


def test_generator_writes_the_module(tmp_path):
    import os
    from chemgen_amd.generate_chemistry import main
    from chemgen_amd.kernels import DATA_DIR
    # with --compat reference the defaults are the reference's (generate_chemistry.py:21-26):
    # level 1, and each reaction type padded to a multiple of 8 -- at level 1 too, as the
    # reference's gpu_coef mode passes nreact_per_block to write_coef_module at every level
    main(["--mode", "gpu_coef", "--mech", os.path.join(DATA_DIR, "h2_burke.inp"), "--output", str(tmp_path),
          "--compat", "reference"])
    src = (tmp_path / "coef_m_v1.f90").read_text()
    assert parse_module(src) == parse_module(emit_coef_module(load_bundled("h2_burke"), 1, 8))
    main(["--mode", "gpu_coef", "--mech", os.path.join(DATA_DIR, "h2_burke.inp"), "--output", str(tmp_path),
          "--compat", "reference", "--parallel-level", "2"])
    src = (tmp_path / "coef_m_v2.f90").read_text()
    assert parse_module(src) == parse_module(emit_coef_module(load_bundled("h2_burke"), 2, 8))
    for level in (1, 3, 4):
        main(["--mode", "gpu_coef", "--mech", os.path.join(DATA_DIR, "h2_burke.inp"), "--output", str(tmp_path),
              "--compat", "reference", "--parallel-level", str(level), "--nreact-per-block", "8"])
        got = (tmp_path / f"coef_m_v{level}.f90").read_text()
        assert got == emit_coef_module(load_bundled("h2_burke"), level, 8)
