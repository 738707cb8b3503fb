"""Regenerate the mechanism files in this folder (run in a checkout that has the
reference's ``ck_files`` next to it; the files it writes are committed, so nothing reads
the reference at run time).

    python -m chemgen_amd.data.make_synthetic /path/to/reference/ck_files

Outputs (Chemkin text, thermo inline):
  h2_burke.inp          ck_files/H2_burke/chem.yaml re-serialised (9 species, 23 reactions)
  h2_burke2012.inp      ck_files/Hydrogen_H2/Burke_et_al_2012/chem.yaml (13 species, 27 reactions)
  gri30.inp             ck_files/GRI30/chem.yaml (53 species, 325 reactions)
  fuller2021.inp        ck_files/Nitrogen/Fuller-2021/chem.yaml (835 species, 4289 reactions:
                        523 PLOG, 73 falloff, 35 third-body; 43 reaction sides of 4-5 species)
  hashemi2016.inp       ck_files/CH4_Methane/Hashemi2016/chem.yaml (68 species, 631 reactions:
                        114 PLOG, 34 falloff, 11 third-body; 22 with a negative prefactor)
  ic8_llnl.inp          ck_files/i-Octane_C8H18/llnl_version3/chem.yaml (874 species, 6864
                        reactions: the reference's largest species count)
  synth_gri53.inp       GRI30 species; 325 reactions: 25 Troe + 4 Lindemann, 40 third-body
  synth_nhept88.inp     n-heptane sk88 species; 387 reactions: 33 Troe + 6 Lindemann, 20 third-body
  synth_jetsurf348.inp  JetSurf 2.0 species; 2163 reactions: 150 Troe + 23 SRI, 324 third-body
  synth_plog53.inp      GRI30 species; 120 reactions: 30 PLOG, 8 Troe, 12 third-body
  synth_ext53.inp       synth_plog53 plus 12 reactions with sides of 4+ species slots or
                        coefficients 0.5 (sums and halves of its own reactions, every kind)
  synth_eff53.inp       synth_gri53 with 9-15 collision efficiencies on every third-body and
                        falloff reaction (from synth_gri53.inp alone: no ck_files needed)

    python -m chemgen_amd.data.make_synthetic /path/to/reference/ck_files synth_plog53

writes only the named files.
"""
import os
import sys

from ..mechanism import load_yaml, write_chemkin
from ..synth import SPECS, synthetic_mechanism, write_and_parse

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ck_files: str, only=()) -> None:
    real = {
        "h2_burke": "H2_burke/chem.yaml",
        "h2_burke2012": "Hydrogen_H2/Burke_et_al_2012/chem.yaml",
        "gri30": "GRI30/chem.yaml",
        "fuller2021": "Nitrogen/Fuller-2021/chem.yaml",
        "hashemi2016": "CH4_Methane/Hashemi2016/chem.yaml",
        "ic8_llnl": "i-Octane_C8H18/llnl_version3/chem.yaml",
    }
    for name, rel in real.items():
        if only and name not in only:
            continue
        m = load_yaml(os.path.join(ck_files, rel), name=name)
        write_chemkin(m, os.path.join(HERE, name + ".inp"))
    bases = {
        "synth_gri53": "GRI30/chem.yaml",
        "synth_nhept88": "n-Heptane_C7H16/llnl_v3.1/nheptane_sk88/chem.yaml",
        "synth_jetsurf348": "JetSurf/2.0/chem.yaml",
        "synth_plog53": "GRI30/chem.yaml",
    }
    for name, rel in bases.items():
        if only and name not in only:
            continue
        base = load_yaml(os.path.join(ck_files, rel))
        m = synthetic_mechanism(base.species, base.elements, name=name, **SPECS[name])
        m2 = write_and_parse(m, os.path.join(HERE, name + ".inp"))
        print(name, m2.n_species, m2.n_reactions, m2.counts())
    if not only or "synth_eff53" in only:
        from ..synth import with_long_efficiency_lists
        from ..mechanism import parse_chemkin
        base = parse_chemkin(os.path.join(HERE, "synth_gri53.inp"), name="synth_gri53")
        m = with_long_efficiency_lists(base, 9, 15, seed=5354, name="synth_eff53")
        m2 = write_and_parse(m, os.path.join(HERE, "synth_eff53.inp"))
        print("synth_eff53", m2.n_species, m2.n_reactions, m2.counts())
    if not only or "synth_ext53" in only:
        from ..synth import with_extended_reactions
        from ..mechanism import parse_chemkin
        base = parse_chemkin(os.path.join(HERE, "synth_plog53.inp"), name="synth_plog53")
        m = with_extended_reactions(base, 12, seed=5353, name="synth_ext53")
        m2 = write_and_parse(m, os.path.join(HERE, "synth_ext53.inp"))
        print("synth_ext53", m2.n_species, m2.n_reactions, m2.counts())


if __name__ == "__main__":
    main(sys.argv[1], tuple(sys.argv[2:]))
