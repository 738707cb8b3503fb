"""Small drivers to run one kernel repeatedly under a profiler (rocprofv3 --kernel-trace /
--pmc): the FP64 getrates kernel of a bundled mechanism, the fused surrogate inference, or
the ragged-batch launch.

    python -m chemgen_amd.profile_run getrates --mech synth_gri53 --ng 1048576 --reps 5
    python -m chemgen_amd.profile_run infer --mech synth_gri53 --ng 1048576 --reps 5
    python -m chemgen_amd.profile_run batched --mech synth_nhept88 --reps 5

States are drawn on the GPU (T U(800, 2800) K unless --tlo/--thi, P 1-40 atm, Y Dirichlet(0.5)).
"""
from __future__ import annotations

import argparse
import os


def _warm(launch, seconds):
    import time
    import torch
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(4):
            launch()
        torch.cuda.synchronize()


def main(argv=None):
    import numpy as np
    import torch
    from .autotune import best_batched, best_variant
    from .kernels import load_bundled
    from .runtime import DeviceMechanism
    from .surrogate import Surrogate, random_states
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["getrates", "infer", "batched"])
    ap.add_argument("--mech", default="synth_gri53")
    ap.add_argument("--ng", type=int, default=1 << 20)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--tlo", type=float, default=800.0)
    ap.add_argument("--thi", type=float, default=2800.0)
    ap.add_argument("--weights", default=None, help="surrogate weights (.npz); default: the bundled ones")
    ap.add_argument("--warm", type=float, default=0.0,
                    help="seconds of untimed launches before the --reps ones (clocks settle under load)")
    a = ap.parse_args(argv)
    m = load_bundled(a.mech)
    if a.what == "batched":
        from .synth import ragged_patches
        offs = ragged_patches()
        v, grid = best_batched(m)
        dm = DeviceMechanism(m, v, build=False)
        P, T, Y = random_states(m.n_species, int(offs[-1]), seed=3, T_range=(a.tlo, a.thi))
        W = torch.empty_like(Y)
        dm.batched_prepare(offs)
        _warm(lambda: dm.getrates_batched_prepared_(P, T, Y, W, grid=grid), a.warm)
        for _ in range(a.reps):
            dm.getrates_batched_prepared_(P, T, Y, W, grid=grid)
        tag = f"{v.tag} grid={grid}"
    elif a.what == "getrates":
        dm = DeviceMechanism(m, best_variant(m), build=False)
        P, T, Y = random_states(m.n_species, a.ng, seed=3, T_range=(a.tlo, a.thi))
        W = torch.empty_like(Y)
        _warm(lambda: dm.getrates_(P, T, Y, W), a.warm)
        for _ in range(a.reps):
            dm.getrates_(P, T, Y, W)
        tag = dm.variant.tag
    else:
        path = a.weights or os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                                         f"surrogate_{a.mech}.npz")
        d = dict(np.load(path))
        dims = [int(x) for x in d["dims"]]
        sur = Surrogate(m, hidden=dims[1], n_hidden=len(dims) - 2, seed=0)
        sur.load_state_dict(d)
        P, T, Y = random_states(m.n_species, a.ng, seed=7, T_range=(a.tlo, a.thi))
        W = torch.empty_like(Y)
        _warm(lambda: sur.predict_(P, T, Y, W, fused=True), a.warm)
        for _ in range(a.reps):
            sur.predict_(P, T, Y, W, fused=True)
        tag = f"MLP {dims}"
    torch.cuda.synchronize()
    print(f"done {a.what} {a.mech} {tag}", flush=True)


if __name__ == "__main__":
    main()
