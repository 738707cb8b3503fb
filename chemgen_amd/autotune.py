"""Launch-shape / code-shape autotuner for the generated getrates kernels.

The reference picks its HIP launch shape by hand through ``--parallel-level``, ``--veclen``
and ``--nreact-per-block`` (generate_chemistry.py:20-26) and a sweep of mini-app builds
(compile_mini_apps.sh, compare_mini_apps_write_csv.sh).  Here each candidate is a
:class:`Variant` (points per block, register budget via waves/EU, exp/log
implementation, reverse-rate formulation, polynomial scheme, EG placement), compiled
ahead of time; the tuner times every candidate on the GPU with HIP events, checks that
all candidates agree with the first one, and records the winner per mechanism structure
in ``data/tuning.json`` so later loads pick it without re-measuring.
"""
from __future__ import annotations

import dataclasses
import json
import os
from typing import Dict, List, Optional

from .codegen import Variant, structure_hash
from .mechanism import Mechanism

TUNING_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "tuning.json")


LDS_BYTES = 160 * 1024          # per workgroup on gfx950
STRAIGHT_LINE_MAX_SPECIES = 150  # above this the lane-per-point and split kernels do not fit


def table_lds_bytes(ns: int, npt: int, waves: int) -> int:
    """Least LDS of one table-kernel tile: EG, h [ns + 1][npt] (one row of ones for absent
    species slots), the per-wavefront partial sums, and the rates of progress of one round of
    reaction slots (waves x 64 lanes); the kernel gives its chunk buffer the rest of the LDS."""
    return 8 * (2 * (ns + 1) * npt + waves * npt + waves * 64)


def table_chunk_reactions(ns: int, nr: int, npt: int, waves: int) -> int:
    """Reactions whose rates of progress the table kernel holds at once (its chunk): the LDS
    left after EG, h, the partial sums and the zero row, in rows of npt doubles."""
    return min(nr, (LDS_BYTES // 8 - (2 * (ns + 1) + waves + 1) * npt) // npt)


def table_fits(ns: int, nr: int, npt: int, waves: int, min_chunk: int = 0) -> bool:
    """Whether the table kernel of npt-point tiles and ``waves`` wavefronts holds the mechanism:
    its LDS, the 16-bit LDS row offsets of the reaction table (ns npt 8 < 2^16), and a chunk
    of at least min(nr, min_chunk) reactions."""
    return (table_lds_bytes(ns, npt, waves) <= LDS_BYTES and ns * npt * 8 < 1 << 16
            and table_chunk_reactions(ns, nr, npt, waves) >= min(nr, min_chunk))


def two_tile_points(ns: int) -> int:
    """Points per reaction-split tile (a multiple of 8, 48 or more) that lets two tiles' species
    state (2 x ns x points doubles each, in 512-byte LDS granules) share a CU where 64-point
    tiles cannot; 0 when 64 points already fit twice or 48 do not."""
    gran = lambda b: -(-b // 512) * 512
    if 2 * gran(8 * 2 * ns * 64) <= LDS_BYTES:
        return 0
    for tp in range(56, 47, -8):
        if 2 * gran(8 * 2 * ns * tp) <= LDS_BYTES:
            return tp
    return 0


def candidate_variants(mech: Mechanism) -> List[Variant]:
    """Candidates swept by default (all compiled by the build for the bundled mechanisms):
    straight-line lane-per-point and reaction-split kernels while the species state of a
    64-point tile fits (registers / LDS), table-driven kernels with every tile length whose
    LDS fits.  Every candidate divides with one Newton step after the reciprocal estimate
    (``fdiv``: 1.9% faster on the GRI-sized split kernel, the same parity statistics); table
    kernels of 16- and 8-point tiles broadcast per-reaction data across each lane group
    (``bcast``: 6% faster on the JetSurf-sized mechanism and on Hashemi 2016, 16% on Fuller 2021,
    the same sums in the same order).  Table workgroups of 12 wavefronts (3 per SIMD, 168
    registers) sit between 8 and 16: the JetSurf-sized kernel spills at 16 (128 registers) and
    runs 14% faster at 12."""
    return [dataclasses.replace(v, fdiv=True) for v in _candidates(mech)]


def _candidates(mech: Mechanism) -> List[Variant]:
    out = []
    ns = mech.n_species
    if ns <= STRAIGHT_LINE_MAX_SPECIES:
        out += [Variant(64, 1, "lean", "gh", "horner", False),   # lane per point, state in registers
                Variant(64, 2, "lean", "gh", "horner", True, 4)]  # lane per point, EG in LDS
        # reaction-split tiles: (wavefronts per tile, waves/SIMD the registers are sized for),
        # LDS re-read distances
        for split, waves, groups in ((4, 2, (8, 16, 20, 28)), (4, 3, (2, 4)), (6, 3, (4, 8)), (8, 2, (4, 8)),
                                     (8, 4, (2, 4, 8)), (16, 4, (2, 4))):
            if 8 * (2 * ns * 64 + split * 64) <= LDS_BYTES and split <= mech.n_reactions:
                for group in groups:
                    out.append(Variant(64, waves, "lean", "gh", "horner", False, group, split))
                # the same with the coefficients compiled in
                out.append(Variant(64, waves, "lean", "gh", "horner", False, groups[-1], split, 0, True))
        # 8-wavefront tiles shortened so that two tiles share a CU where one 64-point tile alone
        # fills more than half the LDS (n-heptane-sized, 88 species: 56 points, 4 waves/SIMD)
        tp2 = two_tile_points(ns)
        if tp2 and mech.n_reactions >= 8:
            for group in (2, 4):
                for baked in (False, True):
                    out.append(Variant(64, 4, "lean", "gh", "horner", False, group, 8, 0, baked, tp=tp2))
        out.append(Variant(64, 1, "lean", "gh", "horner", False, 4, 1, 0, True))
        # lane per point with reactions in live-range order, coefficients read or compiled in
        out += [Variant(64, 1, "lean", "gh", "horner", False, 4, 1, 0, False, True),
                Variant(64, 1, "lean", "gh", "horner", False, 4, 1, 0, True, True)]
        if ns <= 24:   # small state: lane per point at higher occupancy
            for wpe in (2, 4, 8):
                out += [Variant(64, wpe, "lean", "gh", "horner", False, 4, 1, 0, False, True),
                        Variant(64, wpe, "lean", "gh", "horner", False, 4, 1, 0, True, True)]
    if table_supported(mech):
        # tiles shorter than 16 points only where a 16-point tile does not fit
        nr = mech.n_reactions
        short = (1, 2, 4, 8) if not table_fits(ns, nr, 16, 4, 512) else ()
        for npt in short + (16, 32, 64):
            for waves, wpe in ((4, 1), (8, 2), (12, 3), (16, 4)):
                if table_fits(ns, nr, npt, waves):
                    out.append(Variant(64, wpe, "lean", "gh", "horner", False, 4, waves, npt,
                                       bcast=npt in (8, 16)))
    return out


def table_supported(mech: Mechanism) -> bool:
    """Whether the table kernel's reaction format holds this mechanism (reaction types up to
    PLOG, at most 15 collision efficiencies per reaction, ...)."""
    from .codegen import table_records
    try:
        table_records(mech, npt=1)
    except ValueError:
        return False
    return True


def straight_line_size(mech: Mechanism) -> int:
    """Reactions a straight-line kernel spells out: one per reaction plus one per extra
    Arrhenius term of a PLOG reaction (every term of every pressure point is compiled in)."""
    return mech.n_reactions + sum(len(r.plog) - 1 for r in mech.reactions if r.rtype == "plog")


def default_variant(mech: Mechanism) -> Variant:
    """Shape used before any tuning, from the tuner's measurements on MI355X: lane per point for
    small mechanisms (<= 24 species; QSS mechanisms always); reaction-split tiles up to 150
    species while the straight-line code stays small (<= 800 reactions spelled out: GRI-sized
    and n-heptane-sized tuned there, 2x the lane kernel), 4 wavefronts per tile up to 64
    species (two tiles per CU), 8 above; otherwise the table kernel (a mechanism such as Hashemi
    2016, 68 species with 114 PLOG reactions of 908 terms, spills 2-9 KB per lane in the
    straight-line kernels and runs 1.4-3.6x faster in the table kernel), with the longest tile of
    at most 16 points whose species state fits the LDS with a chunk of at least 512 reactions
    (or all of them), and the most wavefronts that fit with it (16, 4 per SIMD, won at every
    tile length on the JetSurf-sized mechanism and Fuller 2021); 16- and 8-point tiles
    broadcast per-reaction data across each lane group (``bcast``)."""
    ns = mech.n_species
    if ns <= 24 or getattr(mech, "qss", ()):
        return Variant()
    if ns <= STRAIGHT_LINE_MAX_SPECIES and straight_line_size(mech) <= 800:
        if ns <= 64:
            return Variant(64, 2, "lean", "gh", "horner", False, 28, 4, fdiv=True)
        return Variant(64, 2, "lean", "gh", "horner", False, 8, 8, fdiv=True)
    for npt in (16, 8, 4, 2, 1):
        for waves, wpe in ((16, 4), (8, 2), (4, 1)):
            if table_fits(ns, mech.n_reactions, npt, waves, 512):
                return Variant(64, wpe, "lean", "gh", "horner", False, 4, waves, npt, fdiv=True,
                               bcast=npt in (8, 16))
    for npt in (16, 8, 4, 2, 1):
        for waves, wpe in ((16, 4), (8, 2), (4, 1)):
            if table_fits(ns, mech.n_reactions, npt, waves):
                return Variant(64, wpe, "lean", "gh", "horner", False, 4, waves, npt, fdiv=True,
                               bcast=npt in (8, 16))
    raise ValueError(f"{ns} species: the table kernel's one-point tile does not fit the LDS")


def _load_db() -> Dict:
    if os.path.exists(TUNING_PATH):
        with open(TUNING_PATH) as fh:
            return json.load(fh)
    return {}


def best_variant(mech: Mechanism) -> Variant:
    """The tuned variant for this mechanism structure, or the default shape."""
    db = _load_db()
    e = db.get(mech.name)
    if e and e.get("structure_hash") == f"{structure_hash(mech):016x}":
        return Variant(**e["best"])
    return default_variant(mech)


def _time_ms(launch, reps: int, blocks: int = 3) -> float:
    """Fastest of ``blocks`` timed blocks of ``reps`` back-to-back launches (HIP events on
    torch's current stream, which the launches use), after two untimed launches."""
    import torch
    for _ in range(2):
        launch()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(blocks):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            launch()
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1) / reps)
    return best


def _warm_gpu(seconds: float = 0.5) -> None:
    """Busy the GPU for a moment so that the first candidate is not timed at idle clocks."""
    import time
    import torch
    x = torch.ones(1 << 24, dtype=torch.float64, device="cuda:0")
    t0 = time.time()
    while time.time() - t0 < seconds:
        for _ in range(20):
            x.mul_(1.0000001).add_(1e-9)
        torch.cuda.synchronize()


def tune(mech: Mechanism, ng: int = 1 << 20, variants: Optional[List[Variant]] = None,
         reps: int = 10, seed: int = 0, record: bool = True, build: bool = False,
         out_path: Optional[str] = None) -> Dict:
    """Time each variant on cuda:0 over ng seeded states (fastest of three blocks of
    ``reps`` launches, after a warm-up of the GPU); returns (and records) the results."""
    import numpy as np
    import torch
    from .runtime import DeviceMechanism
    from .synth import make_states
    variants = variants or candidate_variants(mech)
    P, T, Y = make_states(mech.n_species, ng, seed=seed)
    dev = torch.device("cuda:0")
    Pd, Td, Yd = (torch.from_numpy(x).to(dev) for x in (P, T, Y))
    W = torch.empty_like(Yd)
    ref = None
    results = []
    _warm_gpu()
    for v in variants:
        print(f"[autotune] {mech.name} {v.tag}", flush=True)
        dm = DeviceMechanism(mech, v, build=build)
        ms = _time_ms(lambda: dm.getrates_(Pd, Td, Yd, W), reps)
        w = W[:, :4096].cpu().numpy()
        if ref is None:
            ref = w
            agree = 0.0
        else:
            # difference relative to the largest |wdot| of the same grid point: species near
            # steady state have |wdot| far below the terms summed into it, so a per-species
            # scale would flag harmless reordering; a wrong kernel is off by O(1) here
            scale = np.maximum(np.abs(ref).max(axis=0, keepdims=True), 1e-300)
            agree = float(np.max(np.abs(w - ref) / scale))
        ok = bool(np.isfinite(w).all()) and agree <= 1e-9
        results.append({"variant": dataclasses.asdict(v), "tag": v.tag, "ms": ms,
                        "points_per_s": ng / (ms / 1e3), "agree": agree, "ok": ok})
        print(json.dumps(results[-1]), flush=True)
        dm.close()
    good = [r for r in results if r["ok"]]
    best = min(good, key=lambda r: r["ms"]) if good else None
    entry = {"structure_hash": f"{structure_hash(mech):016x}", "ng": ng,
             "best": best["variant"] if best else dataclasses.asdict(default_variant(mech)), "results": results}
    if record and best:
        db = _load_db()
        if out_path and os.path.exists(out_path):
            with open(out_path) as fh:
                db = json.load(fh)
        old = db.get(mech.name, {})
        if "batched" in old and old.get("structure_hash") == entry["structure_hash"]:
            entry["batched"] = old["batched"]
        db[mech.name] = entry
        with open(out_path or TUNING_PATH, "w") as fh:
            json.dump(db, fh, indent=1, sort_keys=True)
    return entry


def best_batched(mech: Mechanism):
    """(variant, grid) tuned for ragged batches of this mechanism; grid 0 = the runtime's
    default of 8 persistent workgroups per CU."""
    e = _load_db().get(mech.name, {})
    b = e.get("batched")
    if b and e.get("structure_hash") == f"{structure_hash(mech):016x}":
        return Variant(**b["best"]), int(b["grid"])
    return best_variant(mech), 0


def tune_batched(mech: Mechanism, offsets, P, T, Y, variants: Optional[List[Variant]] = None,
                 grids=(-1, 0, 1, 2, 4), reps: int = 10, record: bool = True,
                 out_path: Optional[str] = None) -> Dict:
    """Time the prepared ragged-batch launch for every (variant, persistent grid) pair on
    device-resident states P, T, Y (patches laid out back to back, ``offsets`` [n+1]).
    ``grids`` are workgroups per CU (0 = runtime default of 8 per CU, -1 = one workgroup
    per tile, the hardware dispatcher then balances the tiles).  All candidates must agree with
    the first to 1e-9 of each point's largest |wdot|; the fastest is recorded under the
    mechanism's ``batched`` entry."""
    import numpy as np
    import torch
    from .runtime import DeviceMechanism
    variants = variants or candidate_variants(mech)
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    W = torch.empty_like(Y)
    ref = None
    _warm_gpu()
    chk = np.unique(np.clip(np.asarray(offsets[:-1]), 0, Y.shape[1] - 1))[:64]
    results = []
    for v in variants:
        dm = DeviceMechanism(mech, v, build=False)
        dm.batched_prepare(offsets)
        for gpc in grids:
            grid = gpc * n_cu if gpc >= 0 else -1
            print(f"[autotune-batched] {mech.name} {v.tag} grid={grid}", flush=True)
            ms = _time_ms(lambda: dm.getrates_batched_prepared_(P, T, Y, W, grid=grid), reps)
            w = W[:, torch.from_numpy(chk).to(W.device)].cpu().numpy()
            if ref is None:
                ref, agree = w, 0.0
            else:
                scale = np.maximum(np.abs(ref).max(axis=0, keepdims=True), 1e-300)
                agree = float(np.max(np.abs(w - ref) / scale))
            ok = bool(np.isfinite(w).all()) and agree <= 1e-9
            results.append({"variant": dataclasses.asdict(v), "tag": v.tag, "grid_per_cu": gpc, "ms": ms,
                            "points_per_s": int(offsets[-1]) / (ms / 1e3), "agree": agree, "ok": ok})
            print(json.dumps(results[-1]), flush=True)
        dm.close()
    good = [r for r in results if r["ok"]]
    best = min(good, key=lambda r: r["ms"]) if good else None
    entry = {"n_points": int(offsets[-1]), "n_patches": len(offsets) - 1, "results": results,
             "best": best["variant"] if best else None,
             "grid": (best["grid_per_cu"] * n_cu if best["grid_per_cu"] >= 0 else -1) if best else 0}
    if record and best:
        path = out_path or TUNING_PATH
        db = _load_db()
        if out_path and os.path.exists(out_path):
            with open(out_path) as fh:
                db = json.load(fh)
        db.setdefault(mech.name, {"structure_hash": f"{structure_hash(mech):016x}",
                                  "best": dataclasses.asdict(best_variant(mech))})
        db[mech.name]["batched"] = entry
        with open(path, "w") as fh:
            json.dump(db, fh, indent=1, sort_keys=True)
    return entry


if __name__ == "__main__":
    import argparse
    from .kernels import load_bundled
    ap = argparse.ArgumentParser()
    ap.add_argument("mechs", nargs="*", default=["synth_gri53"],
                    help="bundled mechanism names or mechanism files (Chemkin with thermo inline, or YAML)")
    ap.add_argument("--ng", type=int, default=1 << 20)
    ap.add_argument("--out", default=None, help="where to write the tuning database")
    ap.add_argument("--build", action="store_true", help="compile the candidates that are not built yet")
    ap.add_argument("--batched", action="store_true",
                    help="tune the ragged-batch launch: 64 patches, log-uniform 2^8-2^17 points, 2^22 in total")
    a = ap.parse_args()
    for name in a.mechs:
        if os.path.isfile(name):
            from .mechanism import load_mechanism
            m = load_mechanism(name, name=os.path.splitext(os.path.basename(name))[0])
        else:
            m = load_bundled(name)
        if a.batched:
            import torch
            from .synth import make_states, ragged_patches
            offs = ragged_patches()
            P, T, Y = make_states(m.n_species, int(offs[-1]), seed=0, T_range=(600.0, 2800.0))
            d = torch.device("cuda:0")
            Pd, Td, Yd = (torch.from_numpy(x).to(d) for x in (P, T, Y))
            del P, T, Y
            tune_batched(m, offs, Pd, Td, Yd, out_path=a.out)
        else:
            tune(m, ng=a.ng, out_path=a.out, build=a.build)
