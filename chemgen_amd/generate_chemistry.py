"""Command line of the generator, with the reference's flags (generate_chemistry.py:7-54).

    python -m chemgen_amd.generate_chemistry --mode getrates --language hip \
        --mech chem.inp --therm therm.dat --output getrates_hip.py [--parallel-level N]
    python -m chemgen_amd.generate_chemistry --mode gpu_coef --mech chem.inp --output DIR

``--mode getrates --language hip`` writes, next to ``--output``:
  <stem>.py          Python module exposing ``wdot = getrates(P, T, Y)``
  <stem>.hip         the generated gfx950 kernel source
  <stem>.hsaco       its code object (hipcc --genco)
  <stem>.coef.npy    the gpu_coef blob
  <stem>.json, .inp  mechanism metadata and the mechanism as parsed
``--mode gpu_coef`` writes DIR/gpu_coef_cdna4.npy (+ .json): the coefficient blob the
kernels read (the layout is documented in chemgen_amd/codegen.py); with ``--compat
reference`` also DIR/coef_m_vN.f90 (N = --parallel-level), the reference's Fortran
coefficient module that its ``chemgen_m`` uses (chemgen_amd/legacy_coef.py): species maps at
levels 1-3 (MAX_SP a power of two at 3), dense rocBLAS matrices at 4, each reaction type padded
to a multiple of --nreact-per-block with the reference's dummy reactions (at every level, as the
reference's gpu_coef mode does; with --compat reference the defaults of --parallel-level,
--veclen and --nreact-per-block are the reference's: 1, 16, 8).

Launch shape, the analogue of the reference's parallel levels (--parallel-level,
--veclen, --nreact-per-block):
  1  one lane per grid point, VECLEN points per workgroup (rounded up to 64)
  2  reaction-split tiles: NREACT_PER_BLOCK wavefronts share a 64-point tile in LDS
  3  table-driven tiles of VECLEN points (1, 2, 4, 8, 16, 32 or 64: the nearest) and
     NREACT_PER_BLOCK wavefronts, each lane one (point, reaction slot) pair -- the shape for
     mechanisms too large for 2 (500+ species need tiles of 8 points or fewer)
  4  the autotuned shape for this mechanism (data/tuning.json), else the default

``--legacy-abi`` (with ``--compat reference``) also writes ``<stem>_legacy.so``: the
reference's C entry points for this mechanism (copyConstantsToDevice taking the reference's
coefficient arrays, getrates_hip_wrapper_v1..v4, allocate_intermediate_DeviceMemory) over the
generated kernel, so that the reference's Fortran module links it unchanged.

The reference's python/fortran/OpenMP targets, its constants_vN.h (read only by its own HIP
kernels) and the mini-app are not produced here (those targets exist in the reference).
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np


def parse_arguments(argv=None):
    p = argparse.ArgumentParser(description="Generate GPU coefficients and getrates functions")
    p.add_argument("--mode", choices=["gpu_coef", "getrates", "mini_app"], required=True)
    p.add_argument("--mech")
    p.add_argument("--therm")
    p.add_argument("--yaml_file")
    p.add_argument("--output", required=True)
    # defaults: this build's (the autotuned shape) unless --compat reference, which takes the
    # reference's (generate_chemistry.py:21-26: level 1, VECLEN 16, 8 reactions per block), so
    # that a coefficient module written without these flags has the layout the reference's own
    # would have
    p.add_argument("--parallel-level", type=int, choices=[1, 2, 3, 4], default=None)
    p.add_argument("--veclen", type=int, default=None)
    p.add_argument("--nreact-per-block", type=int, default=None)
    p.add_argument("--language", choices=["python", "fortran", "hip"], default="hip")
    p.add_argument("--omp", action="store_true")
    p.add_argument("--vector", action="store_true")
    p.add_argument("--module", action="store_true")
    p.add_argument("--input-mw", action="store_true")
    p.add_argument("--rtypes-together", action="store_true")
    p.add_argument("--ng", type=int, default=1)
    p.add_argument("--ncpu", type=int, default=64)
    p.add_argument("--ngpu", type=int, default=8)
    p.add_argument("--nt", type=int, default=1)
    p.add_argument("--time-cpu", action="store_true")
    p.add_argument("--compat", choices=["full", "reference"], default="full",
                   help="reference: keep only the reaction types the reference handles")
    p.add_argument("--qssa", default="",
                   help="comma-separated quasi-steady-state species (the reference's "
                        "chemistry(qssa_species=...)): Y and wdot then hold the other species "
                        "only, the QSS concentrations come from the linear QSS balance")
    p.add_argument("--legacy-abi", action="store_true",
                   help="also build lib<stem>_legacy.so exporting the reference's C entry points "
                        "(copyConstantsToDevice, getrates_hip_wrapper_v1..v4) over this kernel")
    a = p.parse_args(argv)
    ref = a.compat == "reference"
    for name, ours, theirs in (("parallel_level", 4, 1), ("veclen", 64, 16), ("nreact_per_block", 4, 8)):
        if getattr(a, name) is None:
            setattr(a, name, theirs if ref else ours)
    return a


def variant_for(args, mech):
    from .autotune import best_variant
    from .codegen import Variant
    if getattr(mech, "qss", ()):
        return Variant()          # QSS kernels: lane per point
    lvl = args.parallel_level
    if lvl == 1:
        return Variant(block=max(64, (args.veclen + 63) // 64 * 64), waves_per_eu=1)
    if lvl == 2:
        return Variant(64, 2, "lean", "gh", "horner", False, 8, max(2, args.nreact_per_block))
    if lvl == 3:
        from .autotune import table_fits
        npt = min((1, 2, 4, 8, 16, 32, 64), key=lambda n: abs(n - args.veclen))
        waves = max(1, args.nreact_per_block)
        if not table_fits(mech.n_species, mech.n_reactions, npt, waves):
            raise SystemExit(f"--parallel-level 3: {npt} points x {mech.n_species} species exceed the LDS; "
                             "use a smaller --veclen")
        return Variant(64, max(1, min(4, waves // 4)), "lean", "gh", "horner", False, 4, waves, npt)
    return best_variant(mech)


def _load(args):
    from .mechanism import load_mechanism
    mech = load_mechanism(args.mech, args.therm, args.yaml_file, compat=args.compat)
    qss = [s.strip() for s in getattr(args, "qssa", "").split(",") if s.strip()]
    if qss:
        from .qssa import with_qss
        mech = with_qss(mech, qss)
    return mech


def generate_gpu_coefficients(args) -> None:
    from .codegen import build_coef, structure_hash
    mech = _load(args)
    os.makedirs(args.output, exist_ok=True)
    coef, off_smh, roffs = build_coef(mech)
    np.save(os.path.join(args.output, "gpu_coef_cdna4.npy"), coef)
    with open(os.path.join(args.output, "gpu_coef_cdna4.json"), "w") as fh:
        json.dump({"mechanism": mech.name, "species": mech.species_names, "n_coef": len(coef),
                   "smh_offset": off_smh, "reaction_offsets": roffs,
                   "structure_hash": f"{structure_hash(mech):016x}"}, fh, indent=1)
    if args.compat == "reference":
        # the reference's Fortran coefficient module, for chemgen_m over the legacy library
        from .legacy_coef import emit_coef_module
        level = args.parallel_level
        path = os.path.join(args.output, f"coef_m_v{level}.f90")
        with open(path, "w") as fh:
            # the reference's CLI always pads each reaction type to a multiple of
            # --nreact-per-block (generate_chemistry.py:56-70)
            fh.write(emit_coef_module(mech, level, args.nreact_per_block))
        print(f"reference coefficient module written to {path}")
    print(f"gpu_coef written to {args.output}")


def generate_getrates(args) -> None:
    if args.language != "hip":
        raise SystemExit("this generator emits --language hip; the python/fortran targets are the "
                         "reference's own")
    from .api import save_generated
    from .codegen import compile_code_object, emit_hip, variant_flags
    mech = _load(args)
    v = variant_for(args, mech)
    stem = os.path.splitext(args.output)[0]
    os.makedirs(os.path.dirname(os.path.abspath(stem)), exist_ok=True)
    src, _, _ = emit_hip(mech, v)
    with open(stem + ".hip", "w") as fh:
        fh.write(src)
    compile_code_object(stem + ".hip", stem + ".hsaco", variant_flags(v))
    save_generated(stem, mech, v, stem + ".hsaco")
    base = os.path.basename(stem)
    with open(stem + ".py", "w") as fh:
        fh.write(f'''"""Generated by chemgen_amd: getrates for mechanism '{mech.name}'
({mech.n_species} species, {mech.n_reactions} reactions), FP64 HIP kernel for gfx950."""
import os

from chemgen_amd.api import load_generated

_impl = load_generated(os.path.join(os.path.dirname(os.path.abspath(__file__)), "{base}"))
species = _impl.species


def getrates(P, T, Y):
    """wdot = getrates(P, T, Y): P [dyn/cm^2], T [K], Y mass fractions -> wdot [mol/cm^3/s]."""
    return _impl(P, T, Y)
''')
    if args.legacy_abi:
        from .legacy_abi import build_legacy_library
        build_legacy_library(mech, os.path.abspath(stem + ".hsaco"), stem + "_legacy.so", args.parallel_level)
        print(f"reference C ABI library written to {stem}_legacy.so")
    print(f"Expressions and getrates function written to {args.output}")


def main(argv=None) -> None:
    args = parse_arguments(argv)
    if args.mode == "gpu_coef":
        generate_gpu_coefficients(args)
    elif args.mode == "getrates":
        generate_getrates(args)
    else:
        raise SystemExit("--mode mini_app: the Fortran mini-app is the reference's own harness; "
                         "use bench.py for timing")


if __name__ == "__main__":
    main()
