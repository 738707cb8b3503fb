"""getrates throughput on one MI355X: FP64 HIP kernel vs the host-core C/OpenMP oracle.

One step = one getrates call over ng grid points of the named mechanism (inputs resident
in HBM, species-major layout).  Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the CPU baseline's OpenMP threads: one per core, packed (set before any OpenMP runtime loads)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")


WARM_S = 1.0   # untimed warm-up time before each timed region, seconds


def warm_up(launch, warmup, seconds=WARM_S):
    """At least ``warmup`` untimed launches, more until ``seconds`` have passed: the timed
    launches then run at the clocks a sustained load settles at, not at idle clocks.
    Returns the number of launches made."""
    import torch
    n = 0
    t0 = time.perf_counter()
    while n < warmup or time.perf_counter() - t0 < seconds:
        launch()
        n += 1
        if n % 8 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
    return n


def measured_fp64_peak(iters=20000):
    """FP64 FMA ceiling of this GPU (cg_probe_fp64_fma: eight independent FMA chains per lane,
    8 waves per SIMD), TFLOP/s, best of five timed launches after a warm-up."""
    import ctypes
    import torch
    from chemgen_amd.runtime import lib
    L = lib()
    L.cg_probe_fp64_fma.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]
    L.cg_probe_fp64_fma.restype = ctypes.c_int
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    blocks = n_cu * 8
    out = torch.empty(blocks * 256, dtype=torch.float64, device="cuda:0")
    s = torch.cuda.current_stream()
    launch = lambda: L.cg_probe_fp64_fma(out.data_ptr(), blocks, iters, ctypes.c_void_p(s.cuda_stream))
    warm_up(launch, 2, 0.5)
    best = 0.0
    for _ in range(5):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        launch()
        e1.record(s)
        torch.cuda.synchronize()
        best = max(best, 2.0 * 8 * iters * 256 * blocks / (e0.elapsed_time(e1) / 1e3))
    return best / 1e12


def host_info():
    """The host the CPU baseline runs on: logical CPUs, CPUs this process may use, CPU model,
    and the OpenMP settings."""
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)), "cpu_model": model,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"),
            "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"), "OMP_PLACES": os.environ.get("OMP_PLACES")}


def _avx512():
    try:
        with open("/proc/cpuinfo") as fh:
            flags = next((l for l in fh if l.startswith("flags")), "")
        return all(f in flags.split() for f in ("avx512f", "avx512dq", "avx512vl", "avx512bw"))
    except OSError:
        return False


def _median_rate(fn, npts, runs=5, min_s=1.0):
    """Median over ``runs`` timed runs of points/s; each run repeats ``fn`` for at least
    ``min_s`` seconds."""
    import numpy as np
    fn()
    rates = []
    for _ in range(runs):
        reps, t0 = 0, time.perf_counter()
        while reps < 1 or time.perf_counter() - t0 < min_s:
            fn()
            reps += 1
        rates.append(npts * reps / (time.perf_counter() - t0))
    return float(np.median(rates)), rates


def cpu_baseline(mech_name, mech, P, T, Y, target_s=15.0, sample_ng=1 << 14, host=None):
    """The C/OpenMP restatement of the reference's getrates_i on the host cores, on the first
    ``sample_ng`` of the GPU's own (seed 0) states: all OpenMP threads (OMP_PROC_BIND=close,
    OMP_PLACES=cores) and one thread, median of 5 runs each, for the strict build (libm) and
    the vectorised build (libmvec; AVX-512 where the CPU has it); ``value`` is the faster
    all-thread figure.  Also the reference's Python-language getrates as restated in NumPy by
    the oracle, on the H2 Burke mechanisms at ng = 4096, one thread."""
    import numpy as np
    from chemgen_amd.kernels import load_bundled
    from chemgen_amd.synth import make_states
    from oracle.build import COracle
    from oracle.getrates_np import getrates as np_oracle
    from oracle.metrics import parity_stats
    n = sample_ng
    Pc, Tc, Yc = P[:n].copy(), T[:n].copy(), np.ascontiguousarray(Y[:, :n])
    ref, qf, qb = np_oracle(mech, Pc[:256], Tc[:256], Yc[:, :256], return_parts=True)
    vec = "vec512" if _avx512() else "vec"
    run_s = target_s / 5.0 / 4.0  # 4 timed configurations x 5 runs ~ target_s
    builds = {}
    for flavor in ("", vec):
        co = COracle(mech_name, flavor)
        par = parity_stats(mech, co.getrates(Pc[:256], Tc[:256], Yc[:, :256]), ref, qf, qb)
        threads = co.threads
        all_t, runs_all = _median_rate(lambda: co.getrates(Pc, Tc, Yc), n, 5, run_s)
        co.set_threads(1)
        one_t, runs_one = _median_rate(lambda: co.getrates(Pc[:4096], Tc[:4096], Yc[:, :4096]), 4096, 5, run_s)
        co.set_threads(threads)
        builds[flavor or "strict"] = {"threads": threads, "points_per_s": all_t, "runs": runs_all,
                                      "one_thread_points_per_s": one_t, "one_thread_runs": runs_one,
                                      "parity_max_gross": par["max_gross"]}
    best = max(builds, key=lambda k: builds[k]["points_per_s"])
    # the reference's Python-language getrates as the NumPy restatement of the oracle times it,
    # one thread, 4096 points (the reference's own Python is not run)
    py = {}
    for h2 in ("h2_burke2012", "h2_burke"):
        m0 = load_bundled(h2)
        P0, T0, Y0 = make_states(m0.n_species, 4096, seed=0)
        rate, _ = _median_rate(lambda: np_oracle(m0, P0, T0, Y0), 4096, 5, 0.3)
        py[h2] = {"points_per_s": rate, "ng": 4096}
    return {"value": builds[best]["points_per_s"], "unit": "points/s", "cores": builds[best]["threads"],
            "kind": "port",
            "sample": f"median of 5 runs of >= {run_s:.2f} s, each over the first {n} of the benchmark's "
                      f"seed-0 states of {mech_name}; C/OpenMP restatement of the reference's getrates_i "
                      f"({best} build)",
            "build": best, "builds": builds, "host": host or host_info(),
            "numpy_restatement_ng4096": py}


def bench_batched(steps, warmup):
    """Ragged batch on the n-heptane-sized mechanism: 64 patches with log-uniform sizes in
    2^8..2^17 points, 2^22 points in total, T U(600, 2800) K; the prepared persistent launch
    is captured in a HIP graph and the graph is replayed and timed."""
    import numpy as np
    import torch
    from chemgen_amd.autotune import best_batched
    from chemgen_amd.kernels import load_bundled
    from chemgen_amd.roofline import PEAK_FP64_VECTOR, work_per_point
    from chemgen_amd.runtime import DeviceMechanism
    from chemgen_amd.synth import make_states, ragged_patches
    from oracle.getrates_np import getrates as np_oracle
    from oracle.metrics import parity_stats
    name = "synth_nhept88"
    mech = load_bundled(name)
    offs = ragged_patches()
    ng = int(offs[-1])
    variant, grid = best_batched(mech)
    P, T, Y = make_states(mech.n_species, ng, seed=0, T_range=(600.0, 2800.0))
    dev = torch.device("cuda:0")
    Pd, Td, Yd = (torch.from_numpy(x).to(dev) for x in (P, T, Y))
    Wd = torch.empty_like(Yd)
    dm = DeviceMechanism(mech, variant, build=False)
    dm.batched_prepare(offs)
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            dm.getrates_batched_prepared_(Pd, Td, Yd, Wd, grid=grid, stream=s)
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        warm_up(g.replay, warmup)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(steps):
            g.replay()
        e1.record(s)
    torch.cuda.synchronize()
    t_graph = e0.elapsed_time(e1) / steps / 1e3
    # the same points as one contiguous plain launch, for the cost of raggedness
    e0.record(s)
    for _ in range(steps):
        dm.getrates_(Pd, Td, Yd, Wd, s)
    e1.record(s)
    torch.cuda.synchronize()
    t_plain = e0.elapsed_time(e1) / steps / 1e3
    g.replay()
    torch.cuda.synchronize()
    pick = np.concatenate([np.arange(a, min(a + 4, b)) for a, b in zip(offs[:-1], offs[1:])])
    Wc = Wd[:, torch.from_numpy(pick).to(dev)].cpu().numpy()
    ref, qf, qb = np_oracle(mech, P[pick], T[pick], Y[:, pick], return_parts=True)
    par = parity_stats(mech, Wc, ref, qf, qb)
    w = work_per_point(mech)
    sizes = np.diff(offs)
    return {"workload": f"{name} ({mech.n_species} species, {mech.n_reactions} reactions): ragged batch of "
                        f"{len(sizes)} patches, log-uniform 2^8-2^17 points, {ng} points in total, "
                        f"T U(600,2800) K, seed 0; one launch over the tile list replayed from a HIP graph",
            "points_per_s": ng / t_graph, "kernel_avg_ms": t_graph * 1e3,
            "contiguous_points_per_s": ng / t_plain, "kernel_variant": variant.tag,
            "grid": grid if grid > 0 else ("8 x CUs" if grid == 0 else "one workgroup per tile"), "patch_min": int(sizes.min()), "patch_max": int(sizes.max()),
            "roofline_frac": w["flops"] * ng / t_graph / PEAK_FP64_VECTOR,
            "parity": {"sample_points": int(len(pick)), **par}}


def bench_large(steps, warmup, ng=1 << 19, name="synth_jetsurf348", nchk=256):
    """Plain FP64 getrates on a large mechanism, T U(600, 2800) K, with its tuned or default
    (table-driven) kernel: the JetSurf-sized synthetic mechanism (348 species, 2163 reactions)
    over 2^19 points, Fuller et al. 2021 (835 species, 4289 reactions with 523 PLOG) over 2^18,
    or Hashemi et al. 2016 (68 species, 631 reactions with 114 PLOG of 908 terms) over 2^20."""
    import numpy as np
    import torch
    from chemgen_amd.autotune import best_variant
    from chemgen_amd.kernels import load_bundled
    from chemgen_amd.roofline import PEAK_FP64_VECTOR, work_per_point
    from chemgen_amd.runtime import DeviceMechanism
    from chemgen_amd.synth import make_states
    from oracle.getrates_np import getrates as np_oracle
    from oracle.metrics import parity_stats
    mech = load_bundled(name)
    variant = best_variant(mech)
    P, T, Y = make_states(mech.n_species, ng, seed=0, T_range=(600.0, 2800.0))
    dev = torch.device("cuda:0")
    Pd, Td, Yd = (torch.from_numpy(x).to(dev) for x in (P, T, Y))
    Wd = torch.empty_like(Yd)
    dm = DeviceMechanism(mech, variant, build=False)
    s = torch.cuda.current_stream()
    warm_up(lambda: dm.getrates_(Pd, Td, Yd, Wd, s), warmup)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(steps):
        dm.getrates_(Pd, Td, Yd, Wd, s)
    e1.record(s)
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / steps / 1e3
    Wc = Wd[:, :nchk].cpu().numpy()
    ref, qf, qb = np_oracle(mech, P[:nchk], T[:nchk], Y[:, :nchk], return_parts=True)
    par = parity_stats(mech, Wc, ref, qf, qb)
    w = work_per_point(mech)
    return {"workload": f"{name} ({mech.n_species} species, {mech.n_reactions} reactions), ng={ng}, "
                        f"T U(600,2800) K, P 1-40 atm log-uniform, Y Dirichlet(0.5), seed 0",
            "points_per_s": ng / t, "kernel_avg_ms": t * 1e3, "kernel_variant": variant.tag,
            "flops_per_point": w["flops"], "roofline_frac": w["flops"] * ng / t / PEAK_FP64_VECTOR,
            "parity": {"sample_points": nchk, **par}}


# the reference's only published getrates numbers: H2 Burke mini-app on one gfx90a GCD
# (performance_table_mini_apps_H2_burke.csv:213, ngg = 2,097,152 points, best HIP variant;
# csv:222, one CPU core, getrates_i, ng = 262,144)
H2_PUBLISHED_GPU = 7.69e8
H2_PUBLISHED_CPU_CORE = 5.06e6


def bench_h2_miniapp(steps, warmup, ngg=1 << 21, ng_cpu=1 << 18):
    """The published H2 Burke mini-app measurement redone here: getrates over ngg = 2^21
    points on the GPU (tuned variant, HIP events, inputs resident) and one host core running
    the C restatement of getrates_i over 2^18 points; T U(300, 2500) K, P 1-40 atm, Y
    Dirichlet(0.5), seed 0.  The published figures are from other hardware (gfx90a GCD, an
    unstated CPU) and are shown for reference only."""
    import numpy as np
    import torch
    from chemgen_amd.autotune import best_variant
    from chemgen_amd.kernels import load_bundled
    from chemgen_amd.roofline import PEAK_FP64_VECTOR, PEAK_HBM, work_per_point
    from chemgen_amd.runtime import DeviceMechanism
    from chemgen_amd.synth import make_states
    from oracle.build import COracle
    from oracle.getrates_np import getrates as np_oracle
    from oracle.metrics import parity_stats
    name = "h2_burke"
    mech = load_bundled(name)
    variant = best_variant(mech)
    P, T, Y = make_states(mech.n_species, ngg, seed=0)
    dev = torch.device("cuda:0")
    Pd, Td, Yd = (torch.from_numpy(x).to(dev) for x in (P, T, Y))
    Wd = torch.empty_like(Yd)
    dm = DeviceMechanism(mech, variant, build=False)
    s = torch.cuda.current_stream()
    warm_up(lambda: dm.getrates_(Pd, Td, Yd, Wd, s), warmup)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(steps):
        dm.getrates_(Pd, Td, Yd, Wd, s)
    e1.record(s)
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / steps / 1e3
    nchk = 4096
    ref, qf, qb = np_oracle(mech, P[:nchk], T[:nchk], Y[:, :nchk], return_parts=True)
    par = parity_stats(mech, Wd[:, :nchk].cpu().numpy(), ref, qf, qb)
    # one host core, C restatement of getrates_i (vectorised build), on the full mechanism and
    # on its 17 elementary reactions only (what the reference's CPU mini-app computes, see
    # oracle.build.h2_burke_nofalloff)
    Pc, Tc, Yc = P[:ng_cpu].copy(), T[:ng_cpu].copy(), np.ascontiguousarray(Y[:, :ng_cpu])
    vec = "vec512" if _avx512() else "vec"
    one_core = {}
    for nm in (name, "h2_burke_nofalloff"):
        co = COracle(nm, vec)
        threads0 = co.threads
        co.set_threads(1)
        one_core[nm], _ = _median_rate(lambda: co.getrates(Pc, Tc, Yc), ng_cpu, 5, 0.5)
        co.set_threads(threads0)
    w = work_per_point(mech)
    achieved_bw = w["bytes"] * ngg / t
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tr = json.load(fh).get(f"{name}/{ngg}/{variant.tag}")
        if tr:
            traffic = tr["read_bytes"] + tr["write_bytes"]
    return {"workload": f"{name} ({mech.n_species} species, {mech.n_reactions} reactions), ngg={ngg}, "
                        f"T U(300,2500) K, P 1-40 atm log-uniform, Y Dirichlet(0.5), seed 0 -- the published "
                        f"mini-app size",
            "points_per_s": ngg / t, "kernel_avg_ms": t * 1e3, "kernel_variant": variant.tag,
            "published_gfx90a_points_per_s": H2_PUBLISHED_GPU,
            "hbm_frac": achieved_bw / PEAK_HBM, "fp64_frac": w["flops"] * ngg / t / PEAK_FP64_VECTOR,
            "traffic": traffic,
            "cpu_one_core_points_per_s": one_core[name],
            "cpu_one_core_elementary_only_points_per_s": one_core["h2_burke_nofalloff"],
            "cpu_one_core_sample": f"median of 5 runs over {ng_cpu} points, C restatement of getrates_i "
                                   f"({vec} build), 1 OpenMP thread; elementary_only: without the 4 "
                                   f"third-body and 2 Troe reactions",
            "published_cpu_one_core_points_per_s": H2_PUBLISHED_CPU_CORE,
            "parity": {"sample_points": nchk, **par}}


def bench_qss(steps, warmup, ng=1 << 20):
    """GRI-Mech 3.0 with 12 quasi-steady-state species (chemgen_amd/qssa.py): getrates over
    the 41 transported species of 2^20 points, T U(800, 2500) K, seed 0, with the QSS kernel
    whose coefficients are compiled in; parity against the QSS oracle on a sample."""
    import numpy as np
    import torch
    from chemgen_amd.codegen import Variant
    from chemgen_amd.qssa import bundled_qss
    from chemgen_amd.runtime import DeviceMechanism
    from chemgen_amd.synth import make_states
    from oracle.qssa_np import qss_gross_scale, qss_getrates
    qm = bundled_qss("gri30")
    v = Variant(64, 1, "lean", "gh", "horner", False, 4, 1, 0, True)
    P, T, Y = make_states(qm.n_transported, ng, seed=0, T_range=(800.0, 2500.0))
    dev = torch.device("cuda:0")
    Pd, Td, Yd = (torch.from_numpy(x).to(dev) for x in (P, T, Y))
    Wd = torch.empty_like(Yd)
    dm = DeviceMechanism(qm, v, build=False)
    s = torch.cuda.current_stream()
    warm_up(lambda: dm.getrates_(Pd, Td, Yd, Wd, s), warmup)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(steps):
        dm.getrates_(Pd, Td, Yd, Wd, s)
    e1.record(s)
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / steps / 1e3
    nchk = 1024
    ref, _, qf, qb, _ = qss_getrates(qm, qm.qss, P[:nchk], T[:nchk], Y[:, :nchk], return_parts=True)
    G = qss_gross_scale(qm, qm.qss, qf, qb)
    d = np.abs(Wd[:, :nchk].cpu().numpy() - ref)
    live = G > 0
    return {"workload": f"gri30 with {len(qm.qss)} QSS species ({', '.join(qm.qss)}): {qm.n_transported} "
                        f"transported species, ng={ng}, T U(800,2500) K, P 1-40 atm, Y Dirichlet(0.5), seed 0",
            "points_per_s": ng / t, "kernel_avg_ms": t * 1e3, "kernel_variant": v.tag,
            "parity": {"sample_points": nchk, "max_gross": float(np.max(d[live] / (np.abs(ref) + G)[live])),
                       "inert_exact": bool(np.all(d[~live] == 0.0))}}


def bench_surrogate(steps, warmup, ng=1 << 20):
    """The learned surrogate for the GRI-sized mechanism (4 x 256 SiLU MLP, BF16 on the matrix
    cores, weights in chemgen_amd/data) on 2^20 held-out states (T U(800, 2800) K, seed 7):
    its error against the FP64 kernel per point, max_k |w_sur - w| / (|w| + s_k) with s_k = 1e-3
    of the 99th percentile of |w_k| over the training labels, before and after the per-point
    FP64 fallback, and the throughput of the fused inference kernel and of surrogate +
    fallback (timed on the current stream with HIP events)."""
    import numpy as np
    import torch
    from chemgen_amd.autotune import best_variant
    from chemgen_amd.kernels import load_bundled
    from chemgen_amd.roofline import PEAK_BF16_DENSE
    from chemgen_amd.runtime import DeviceMechanism
    from chemgen_amd.surrogate import Surrogate, random_states
    name = "synth_gri53"
    mech = load_bundled(name)
    dm = DeviceMechanism(mech, best_variant(mech), build=False)
    sur = Surrogate(mech, hidden=256, n_hidden=4, seed=0)
    sur.load_state_dict(dict(np.load(os.path.join(ROOT, "chemgen_amd", "data", "surrogate_synth_gri53.npz"))))
    P, T, Y = random_states(mech.n_species, ng, seed=7)
    ref = torch.empty_like(Y)
    dm.getrates_(P, T, Y, ref)
    W = torch.empty_like(Y)
    flag, _ = sur.predict_(P, T, Y, W)
    err_sur = sur.error(W, ref)
    info = sur.fallback_info(sur.getrates_(P, T, Y, W, dm))
    err = sur.error(W, ref)
    s = torch.cuda.current_stream()

    def timed(fn):
        warm_up(fn, warmup)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(steps):
            fn()
        e1.record(s)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps / 1e3

    t_fp64 = timed(lambda: dm.getrates_(P, T, Y, ref, s))
    t_inf = timed(lambda: sur.predict_(P, T, Y, W))
    t_path = timed(lambda: sur.getrates_(P, T, Y, W, dm))
    flops = 2.0 * sum(sur.dims[i] * sur.dims[i + 1] for i in range(len(sur.dims) - 1))
    q = lambda e, p: float(torch.quantile(e.float(), p).item())
    return {"workload": f"{name}: MLP {sur.dims}, SiLU, BF16 MFMA / FP32 accumulation, {ng} held-out states "
                        f"(T U(800,2800) K, P 1-40 atm, Y Dirichlet(0.5), seed 7); fallback: element residual "
                        f"> 1e-3 or audit (about 4096 points) > 5% off by > 1%",
            "fp64_points_per_s": ng / t_fp64,
            "surrogate_only": {"frac_err_le_1pct": float((err_sur <= 1e-2).double().mean()),
                               "median_err": q(err_sur, 0.5), "p99_err": q(err_sur, 0.99),
                               "residual_flag_frac": float(flag.double().mean())},
            "with_fallback": {"frac_err_le_1pct": float((err <= 1e-2).double().mean()),
                              "fallback_frac": info["fallback_points"] / ng,
                              "global_fallback": info["global_fallback"]},
            "inference_points_per_s": ng / t_inf,
            "inference_roofline": {"bound": "mfma", "achieved": flops * ng / t_inf / 1e12,
                                   "peak": PEAK_BF16_DENSE / 1e12, "unit": "TFLOP/s",
                                   "frac": flops * ng / t_inf / PEAK_BF16_DENSE, "flops_per_point": flops},
            "path_points_per_s": ng / t_path, "path_speedup_vs_fp64": t_fp64 / t_path}


def _finite(o):
    """JSON has no inf/nan: report them as strings."""
    if isinstance(o, dict):
        return {k: _finite(v) for k, v in o.items()}
    if isinstance(o, list):
        return [_finite(v) for v in o]
    if isinstance(o, float) and not math.isfinite(o):
        return str(o)
    return o


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mech", default="synth_gri53")
    ap.add_argument("--ng", type=int, default=1 << 20)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the ragged-batch workload")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the JetSurf-sized, Fuller 2021 and Hashemi 2016 workloads")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the learned-surrogate workload")
    ap.add_argument("--no-h2", action="store_true", help="skip the H2 Burke mini-app workload")
    ap.add_argument("--no-qss", action="store_true", help="skip the QSS-species workload")
    ap.add_argument("--mode", choices=["getrates", "copy"], default="getrates",
                    help="copy: a plain 1 GiB device copy, the calibration for HBM byte counters")
    args = ap.parse_args()
    host = host_info()   # before any OpenMP runtime binds this thread to its first place
    if args.gpus != 1:
        print("bench.py measures one GPU only: the getrates path is per-point independent and "
              "this build runs on a single MI355X; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    if args.mode == "copy":
        n = 1 << 27  # 2^27 doubles = 1 GiB read + 1 GiB written per copy
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(args.steps):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"mode": "copy", "bytes_read_per_launch": 8 * n, "bytes_written_per_launch": 8 * n}))
        return
    from chemgen_amd.kernels import load_bundled
    from chemgen_amd.roofline import PEAK_FP64_VECTOR, work_per_point
    from chemgen_amd.runtime import DeviceMechanism
    from chemgen_amd.synth import make_states

    from chemgen_amd.autotune import best_variant
    mech = load_bundled(args.mech)
    dm = DeviceMechanism(mech, best_variant(mech), build=False)
    ng = args.ng
    P, T, Y = make_states(mech.n_species, ng, seed=0)
    dev = torch.device("cuda:0")
    Pd, Td, Yd = (torch.from_numpy(x).to(dev) for x in (P, T, Y))
    Wd = torch.empty_like(Yd)
    stream = torch.cuda.current_stream()
    n_warm = warm_up(lambda: dm.getrates_(Pd, Td, Yd, Wd, stream), args.warmup)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    t0 = time.perf_counter()
    evs[0].record(stream)
    for s in range(args.steps):
        dm.getrates_(Pd, Td, Yd, Wd, stream)
        evs[s + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = [evs[s].elapsed_time(evs[s + 1]) for s in range(args.steps)]
    avg_launch_s = float(np.mean(launch_ms)) / 1e3
    ms_per_step = wall / args.steps * 1e3

    # parity of the benchmarked output on a sample, against the C oracle
    from oracle.build import COracle
    from oracle.getrates_np import getrates as np_oracle
    from oracle.metrics import parity_stats
    nchk = 2048
    Wc = Wd[:, :nchk].cpu().numpy()
    ref, qf, qb = np_oracle(mech, P[:nchk], T[:nchk], Y[:, :nchk], return_parts=True)
    par = parity_stats(mech, Wc, ref, qf, qb)

    # the 40-digit arbiter on the first 64 points: raw and gross errors of the GPU and of the
    # FP64 oracle against the exact value of the reference formula
    from oracle.arbiter_mp import arbiter_stats, getrates_mp
    na = 64
    wm, G = getrates_mp(mech, P[:na], T[:na], Y[:, :na])
    par["arbiter"] = {"sample_points": na, "gpu": arbiter_stats(Wc[:, :na], wm, G),
                      "oracle": arbiter_stats(ref[:, :na], wm, G)}

    w = work_per_point(mech)
    # the other sizes of the GRI-sized configuration: the first n points of the same resident
    # arrays (leading dimension 2^20), timed the same way
    sizes = {}
    for n in (1 << 16, 1 << 18):
        if n >= ng:
            continue
        a = (Pd[:n], Td[:n], Yd[:, :n], Wd[:, :n])
        warm_up(lambda: dm.getrates_(*a, stream), args.warmup)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = max(args.steps, (args.steps * ng) // n)
        e0.record(stream)
        for _ in range(reps):
            dm.getrates_(*a, stream)
        e1.record(stream)
        torch.cuda.synchronize()
        tn = e0.elapsed_time(e1) / reps / 1e3
        sizes[str(n)] = {"points_per_s": n / tn, "kernel_avg_ms": tn * 1e3,
                         "roofline_frac": w["flops"] * n / tn / PEAK_FP64_VECTOR}
    sizes[str(ng)] = {"points_per_s": ng / avg_launch_s, "kernel_avg_ms": avg_launch_s * 1e3,
                      "roofline_frac": w["flops"] * ng / avg_launch_s / PEAK_FP64_VECTOR}
    achieved = w["flops"] * ng / avg_launch_s
    peak_meas = measured_fp64_peak()
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            t = json.load(fh).get(f"{args.mech}/{ng}/{dm.variant.tag}")
        if t:
            traffic = t["read_bytes"] + t["write_bytes"]
    out = {
        "metric": "getrates points/s: FP64 HIP vs MLP surrogate @<=1% wdot err; % roofline",
        "value": ng / (ms_per_step / 1e3),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_launches": n_warm,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.mech} ({mech.n_species} species, {mech.n_reactions} reactions), "
                               f"ng={ng}, T U(300,2500) K, P 1-40 atm log-uniform, Y Dirichlet(0.5), seed 0; "
                               f"at least {WARM_S:g} s of untimed launches before the timed ones",
                   "mechanism": args.mech, "ng": ng, "kernel_variant": dm.variant.tag},
        "roofline": {"bound": "fp64-valu", "achieved": achieved / 1e12, "peak": PEAK_FP64_VECTOR / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_VECTOR, "traffic": traffic,
                     "flops_per_point": w["flops"], "flops_arith_per_point": w["arith"],
                     "transcendentals_per_point": w["transc"], "bytes_per_point": w["bytes"],
                     "kernel_avg_ms": avg_launch_s * 1e3,
                     "measured_peak": peak_meas, "frac_of_measured": achieved / 1e12 / peak_meas},
        "parity": {"sample_points": nchk, **par},
        "sizes": sizes,
    }
    if not args.no_batched:
        del Pd, Td, Yd, Wd
        torch.cuda.empty_cache()
        out["batched"] = bench_batched(args.steps, args.warmup)
    if not args.no_large:
        torch.cuda.empty_cache()
        out["large"] = bench_large(max(3, args.steps // 4), min(args.warmup, 2))
        torch.cuda.empty_cache()
        out["large_real"] = bench_large(max(3, args.steps // 4), min(args.warmup, 2), ng=1 << 18,
                                        name="fuller2021", nchk=128)
        torch.cuda.empty_cache()
        out["plog_real"] = bench_large(max(3, args.steps // 4), min(args.warmup, 2), ng=1 << 20,
                                       name="hashemi2016", nchk=256)
    if not args.no_surrogate:
        torch.cuda.empty_cache()
        out["surrogate"] = bench_surrogate(args.steps, args.warmup)
    if not args.no_h2:
        torch.cuda.empty_cache()
        out["h2_miniapp"] = bench_h2_miniapp(args.steps, args.warmup)
    if not args.no_qss:
        torch.cuda.empty_cache()
        out["qss"] = bench_qss(args.steps, args.warmup)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.mech, mech, P, T, Y, args.cpu_seconds, host=host)
    print(json.dumps(_finite(out)))


if __name__ == "__main__":
    main()
