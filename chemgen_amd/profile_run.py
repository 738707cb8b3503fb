"""Small drivers to run one kernel repeatedly under a profiler (rocprofv3 --kernel-trace /
--pmc): the FP64 getrates kernel of a bundled mechanism, the fused surrogate inference, the
whole surrogate path (audit, network, FP64 fallback), or the ragged-batch launch.

    python -m chemgen_amd.profile_run getrates --mech synth_gri53 --ng 1048576 --reps 5
    python -m chemgen_amd.profile_run infer --mech synth_gri53 --ng 1048576 --reps 5
    python -m chemgen_amd.profile_run surpath --mech synth_gri53 --ng 1048576 --reps 5
    python -m chemgen_amd.profile_run batched --mech synth_nhept88 --reps 5
    python -m chemgen_amd.profile_run gemm --mech fuller2021 --ng 65536 --reps 5
    python -m chemgen_amd.profile_run gemm --shape 3761 16532 --ng 16384 --reps 5
    python -m chemgen_amd.profile_run host --mech synth_gri53 --ng 1048576 --reps 5

``gemm`` times the GEMM formulation of a mechanism's stoichiometric products -- the logs of
the forward and reverse rate-of-progress products, [2 Nr x Ns] x log C [Ns x ng], and the
production rates, nu [Ns x Nr] x q [Nr x ng] -- as dense FP64 GEMMs (torch.matmul, rocBLAS
on the FP64 matrix cores), and, for a bundled mechanism, the tuned getrates kernel at the
same ng; it prints one JSON line with both times per point.  ``--shape NS NR`` takes a
mechanism of that size with two reactants and two products per reaction.

``host`` times ``GetRates(mech)(P, T, Y)`` called with NumPy arrays (the generated module's
host-buffer form: copies to the GPU, the kernel, the copy back) in both Y layouts, against the
kernel alone on device-resident data, and prints one JSON line.

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
    ap.add_argument("what", choices=["getrates", "infer", "surpath", "batched", "gemm", "host"])
    ap.add_argument("--mech", default="synth_gri53")
    ap.add_argument("--ng", type=int, default=1 << 20)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--tlo", type=float, default=800.0)
    ap.add_argument("--thi", type=float, default=2800.0)
    ap.add_argument("--weights", default=None, help="surrogate weights (.npz); default: the bundled ones")
    ap.add_argument("--warm", type=float, default=0.0,
                    help="seconds of untimed launches before the --reps ones (clocks settle under load)")
    ap.add_argument("--shape", type=int, nargs=2, default=None, metavar=("NS", "NR"),
                    help="gemm: mechanism size without a mechanism (two reactants, two products each)")
    a = ap.parse_args(argv)
    if a.what == "gemm":
        return _gemm(a)
    if a.what == "host":
        return _host(a)
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
        if a.what == "surpath":     # the whole surrogate path: audit, network, FP64 fallback
            dm = DeviceMechanism(m, best_variant(m), build=False)
            _warm(lambda: sur.getrates_(P, T, Y, W, dm), a.warm)
            for _ in range(a.reps):
                info = sur.getrates_(P, T, Y, W, dm)
            tag = f"MLP {dims} path {Surrogate.fallback_info(info)}"
        else:
            _warm(lambda: sur.predict_(P, T, Y, W, fused=True), a.warm)
            for _ in range(a.reps):
                sur.predict_(P, T, Y, W, fused=True)
            tag = f"MLP {dims}"
    torch.cuda.synchronize()
    print(f"done {a.what} {a.mech} {tag}", flush=True)


def _gemm(a):
    """The stoichiometric GEMMs of the GEMM formulation, and for a bundled mechanism the complete
    GEMM-formulation path (chemgen_amd.gemm), against the getrates kernel."""
    import json
    import numpy as np
    import torch
    from .autotune import _time_ms, best_variant
    from .kernels import load_bundled
    from .runtime import DeviceMechanism
    from .surrogate import random_states
    dev = torch.device("cuda:0")
    if a.shape:
        ns, nr = a.shape
        rng = np.random.default_rng(0)
        nu_r = np.zeros((nr, ns))
        nu_p = np.zeros((nr, ns))
        for j in range(nr):
            nu_r[j, rng.choice(ns, 2, replace=False)] = 1.0
            nu_p[j, rng.choice(ns, 2, replace=False)] = 1.0
        name, m = f"{ns}x{nr}", None
    else:
        m = load_bundled(a.mech)
        ns, nr, name = m.n_species, m.n_reactions, a.mech
        idx = {s.name: i for i, s in enumerate(m.species)}
        nu_r = np.zeros((nr, ns))
        nu_p = np.zeros((nr, ns))
        for j, r in enumerate(m.reactions):
            for sp, nu in r.reactants.items():
                nu_r[j, idx[sp]] += nu
            for sp, nu in r.products.items():
                nu_p[j, idx[sp]] += nu
    A1 = torch.from_numpy(np.concatenate([nu_r, nu_p])).to(dev)            # [2 Nr, Ns]
    A2 = torch.from_numpy(np.ascontiguousarray((nu_p - nu_r).T)).to(dev)   # [Ns, Nr]
    g = torch.Generator(device=dev).manual_seed(1)
    lnC = torch.randn(ns, a.ng, dtype=torch.float64, device=dev, generator=g)
    q = torch.randn(nr, a.ng, dtype=torch.float64, device=dev, generator=g)
    L = torch.empty(2 * nr, a.ng, dtype=torch.float64, device=dev)
    W = torch.empty(ns, a.ng, dtype=torch.float64, device=dev)

    def both():
        torch.matmul(A1, lnC, out=L)
        torch.matmul(A2, q, out=W)
    _warm(both, a.warm)
    ms_gemm = _time_ms(both, a.reps)
    flops = 2.0 * (2 * nr * ns + ns * nr) * a.ng
    out = dict(what="gemm", mech=name, ns=ns, nr=nr, ng=a.ng, gemm_ms=ms_gemm,
               gemm_tflops=flops / ms_gemm * 1e-9, gemm_ns_per_point=ms_gemm * 1e6 / a.ng)
    if m is not None:
        from .gemm import GemmGetRates
        v = best_variant(m)
        dm = DeviceMechanism(m, v, build=False)
        P, T, Y = random_states(ns, a.ng, seed=3, T_range=(a.tlo, a.thi))
        Wk = torch.empty_like(Y)
        _warm(lambda: dm.getrates_(P, T, Y, Wk), a.warm)
        ms_k = _time_ms(lambda: dm.getrates_(P, T, Y, Wk), a.reps)
        # the complete GEMM-formulation path (cg_getrates_gemm: element-wise kernels + DGEMMs)
        gp = GemmGetRates(m, slab=min(a.ng, 1 << 16))
        Wg = torch.empty_like(Y)
        _warm(lambda: gp.getrates_(P, T, Y, Wg), a.warm)
        ms_p = _time_ms(lambda: gp.getrates_(P, T, Y, Wg), a.reps)
        scale = torch.clamp(Wk.abs().amax(dim=0, keepdim=True), min=1e-300)
        out.update(kernel=v.tag, getrates_ms=ms_k, getrates_ns_per_point=ms_k * 1e6 / a.ng,
                   gemm_over_getrates=ms_gemm / ms_k, gemm_path_ms=ms_p,
                   gemm_path_ns_per_point=ms_p * 1e6 / a.ng, gemm_path_over_getrates=ms_p / ms_k,
                   gemm_path_vs_kernel_max_diff=float(((Wg - Wk).abs() / scale).max()))
        gp.close()
    print(json.dumps(out), flush=True)


def _host(a):
    """The host-buffer call (PCIe copies included) against the device-resident kernel."""
    import json
    import time
    import numpy as np
    import torch
    from .api import GetRates
    from .autotune import _time_ms
    from .kernels import load_bundled
    from .synth import make_states
    m = load_bundled(a.mech)
    g = GetRates(m, build=False)
    P, T, Y = make_states(m.n_species, a.ng, seed=3, T_range=(a.tlo, a.thi))   # Y [Ns, ng]
    Yp = np.ascontiguousarray(Y.T)                                           # Y [ng, Ns]
    out = dict(what="host", mech=a.mech, ng=a.ng, kernel=g._dm.variant.tag)
    for layout, Yh in (("species", Y), ("point", Yp)):
        g(P, T, Yh, layout=layout)                       # first call: allocations, clocks
        best = float("inf")
        for _ in range(a.reps):
            t0 = time.perf_counter()
            g(P, T, Yh, layout=layout)
            best = min(best, time.perf_counter() - t0)
        out[f"{layout}_ms"] = best * 1e3
        out[f"{layout}_points_per_s"] = a.ng / best
    dev = torch.device("cuda:0")
    Pd, Td, Yd = (torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in (P, T, Y))
    Wd = torch.empty_like(Yd)
    _warm(lambda: g._dm.getrates_(Pd, Td, Yd, Wd), 1.0)
    ms = _time_ms(lambda: g._dm.getrates_(Pd, Td, Yd, Wd), a.reps)
    out.update(device_ms=ms, device_points_per_s=a.ng / (ms / 1e3),
               host_bytes=int((m.n_species * 2 + 2) * 8 * a.ng))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
