"""Train and evaluate the getrates surrogate on the GPU.

    python -m chemgen_amd.surrogate_train --mech synth_gri53 --n-train 8388608 --epochs 12 \
        --out chemgen_amd/data/surrogate_synth_gri53.npz --report report.json
    python -m chemgen_amd.surrogate_train --mech synth_gri53 --sweep 256x2,256x4,512x4 \
        --n-train 8388608 --epochs 8 --report sweep.json

Training states: T U(800, 2800) K, P 1-40 atm log-uniform, Y Dirichlet(0.5) (seeded, drawn on
the GPU), labelled by the FP64 getrates kernel.  Evaluation: 2^20 held-out states (another
seed): per-point error max_k |w_sur - w| / (|w| + s_k) with s_k = 1e-3 x p99 |w_k|, before
and after the fallback, the fraction of points that fell back, and the throughput of
surrogate + fallback against the FP64 kernel on the same points.
"""
from __future__ import annotations

import argparse
import json
import time


def bf16_floor(sur, P, T, Y, W, n: int = 4096) -> dict:
    """How far BF16 arithmetic alone moves this network's outputs: the network's fused BF16
    evaluation W against the same weights (FP32 masters) evaluated in FP64 on the same BF16
    features, per point max_k |dw| / (|w| + s_k).  A network whose FP64 evaluation were exact
    could still be within 1% only on the points this leaves within 1%."""
    import numpy as np
    n = min(n, P.shape[0])
    X = sur.features(P[:n], T[:n], Y[:, :n].contiguous()).double().cpu().numpy()
    h = X
    for l, L in enumerate(sur.layers):
        h = h @ L.w.double().cpu().numpy().T + L.b.double().cpu().numpy()
        if l < len(sur.layers) - 1:
            h = h / (1.0 + np.exp(-h))
    s = sur.s.cpu().numpy()
    w64 = s[None, :] * np.sinh(h)
    wq = W[:, :n].double().cpu().numpy().T
    e = np.max(np.abs(wq - w64) / (np.abs(w64) + s[None, :]), axis=1)
    return {"points": n, "frac_le_1pct": float(np.mean(e <= 1e-2)), "median": float(np.median(e)),
            "p90": float(np.percentile(e, 90))}


def evaluate(sur, dm, n_eval: int, seed: int, tol: float = 1e-3, audit_every: int = 64, reps: int = 5):
    """Accuracy (before / after the fallback) and throughput on n_eval held-out states."""
    import torch
    from .surrogate import random_states
    P, T, Y = random_states(sur.ns, n_eval, seed=seed)
    ref = torch.empty_like(Y)
    dm.getrates_(P, T, Y, ref)
    # the FP64 kernel against itself: zero for the deterministic kernels; the table kernel's
    # LDS atomics reorder additions between runs, which shows where a species' net rate is a
    # near-cancellation of much larger terms
    rep = torch.empty_like(Y)
    dm.getrates_(P, T, Y, rep)
    err_rep = sur.error(rep, ref)
    del rep
    W = torch.empty_like(Y)
    flag, resid = sur.predict_(P, T, Y, W, tol)
    err_raw = sur.error(W, ref)
    floor = bf16_floor(sur, P, T, Y, W)
    info = sur.fallback_info(sur.getrates_(P, T, Y, W, dm, tol=tol, audit_every=audit_every))
    err = sur.error(W, ref)
    torch.cuda.synchronize()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps / 1e3, (time.perf_counter() - t0) / reps

    t_fp64, _ = timed(lambda: dm.getrates_(P, T, Y, ref))
    t_sur, wall_sur = timed(lambda: sur.getrates_(P, T, Y, W, dm, tol=tol, audit_every=audit_every))
    X = sur.features(P, T, Y)
    t_mlp, _ = timed(lambda: sur.forward(X))
    t_fused, _ = timed(lambda: sur.predict_(P, T, Y, W, tol)) if sur.fused_ok else (float("nan"), 0.0)
    flops = 2.0 * sum(sur.dims[i] * sur.dims[i + 1] for i in range(len(sur.dims) - 1)) * n_eval
    q = lambda e, p: float(torch.quantile(e[: 1 << 20].float(), p).item())
    return {
        "n_eval": n_eval,
        "surrogate_only": {"frac_le_1pct": float((err_raw <= 1e-2).double().mean()),
                           "median_err": q(err_raw, 0.5), "p99_err": q(err_raw, 0.99)},
        "with_fallback": {"frac_le_1pct": float((err <= 1e-2).double().mean()), "median_err": q(err, 0.5),
                          "p99_err": q(err, 0.99), **info,
                          "fallback_frac": info["fallback_points"] / n_eval},
        "residual_flag_frac": float(flag.double().mean()),
        "bf16_vs_fp64_same_weights": floor,
        "fp64_repeat": {"frac_le_1pct": float((err_rep <= 1e-2).double().mean()),
                        "median_err": q(err_rep, 0.5), "max_err": float(err_rep.max().item())},
        "fp64_kernel_points_per_s": n_eval / t_fp64,
        "surrogate_path_points_per_s": n_eval / t_sur,
        "surrogate_path_wall_points_per_s": n_eval / wall_sur,
        "speedup_vs_fp64": t_fp64 / t_sur,
        "mlp_forward_ms": t_mlp * 1e3,
        "mlp_forward_tflops": flops / t_mlp / 1e12,
        "fused_infer_ms": t_fused * 1e3,
        "fused_infer_tflops": flops / t_fused / 1e12,
        "fused_infer_points_per_s": n_eval / t_fused,
    }


def sweep(mech_name: str, shapes, n_train: int, epochs: int, n_eval: int = 1 << 20, batch: int = 8192,
          lr: float = 2e-3, log=print) -> list:
    """Accuracy against speed over network shapes ((hidden, layers) pairs), each trained with the
    same data and budget: per shape the fraction of held-out points within 1% and the median
    error of the fused BF16 network, the same for its FP32 master weights evaluated in FP64 (what
    any wider-precision evaluation of this network could reach at most), and the points/s of the
    fused inference kernel against the FP64 kernel on the same points."""
    import numpy as np
    import torch
    from .autotune import best_variant
    from .kernels import load_bundled
    from .runtime import DeviceMechanism
    from .surrogate import Surrogate, random_states
    mech = load_bundled(mech_name)
    dm = DeviceMechanism(mech, best_variant(mech), build=False)
    P, T, Y = random_states(mech.n_species, n_train, seed=100)
    W = torch.empty_like(Y)
    dm.getrates_(P, T, Y, W)
    Pe, Te, Ye = random_states(mech.n_species, n_eval, seed=7)
    ref = torch.empty_like(Ye)
    dm.getrates_(Pe, Te, Ye, ref)
    torch.cuda.synchronize()
    out = []

    def timed(fn, reps=10):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps / 1e3

    t64 = timed(lambda: dm.getrates_(Pe, Te, Ye, ref))
    for hidden, layers in shapes:
        sur = Surrogate(mech, hidden=hidden, n_hidden=layers, seed=0)
        sur.fit_scales(W)
        t0 = time.perf_counter()
        sur.train(P, T, Y, W, epochs=epochs, batch=batch, lr=lr,
                  log=lambda m: log(f"[sweep {hidden}x{layers}] {m}"))
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        Ws = torch.empty_like(Ye)
        sur.predict_(Pe, Te, Ye, Ws)
        err = sur.error(Ws, ref)
        # the same weights in FP64 on a sample: features -> FP32 masters in FP64 -> decode
        n = 8192
        X = sur.features(Pe[:n], Te[:n], Ye[:, :n].contiguous()).double().cpu().numpy()
        h = X
        for l, L in enumerate(sur.layers):
            h = h @ L.w.double().cpu().numpy().T + L.b.double().cpu().numpy()
            if l < len(sur.layers) - 1:
                h = h / (1.0 + np.exp(-h))
        s = sur.s.cpu().numpy()
        w64 = (s[None, :] * np.sinh(h)).T
        r = ref[:, :n].cpu().numpy()
        e64 = np.max(np.abs(w64 - r) / (np.abs(r) + s[:, None]), axis=0)
        t_inf = timed(lambda: sur.predict_(Pe, Te, Ye, Ws))
        flops = 2.0 * sum(sur.dims[i] * sur.dims[i + 1] for i in range(len(sur.dims) - 1))
        row = {"hidden": hidden, "layers": layers, "params": int(sum(L.w.numel() + L.b.numel() for L in sur.layers)),
               "flops_per_point": flops, "train_s": t_train, "epochs": epochs, "n_train": n_train,
               "bf16": {"frac_le_1pct": float((err <= 1e-2).double().mean()), "median_err": float(err.median()),
                        "frac_le_10pct": float((err <= 1e-1).double().mean())},
               "fp64_eval": {"points": n, "frac_le_1pct": float(np.mean(e64 <= 1e-2)),
                             "median_err": float(np.median(e64)), "frac_le_10pct": float(np.mean(e64 <= 1e-1))},
               "inference_points_per_s": n_eval / t_inf, "fp64_points_per_s": n_eval / t64,
               "inference_vs_fp64": t64 / t_inf, "inference_tflops": flops * n_eval / t_inf / 1e12}
        log("[sweep] " + json.dumps(row))
        out.append(row)
        del sur, Ws
        torch.cuda.empty_cache()
    return out


def main(argv=None):
    import numpy as np
    import torch
    from .autotune import best_variant
    from .kernels import load_bundled
    from .runtime import DeviceMechanism
    from .surrogate import Surrogate, random_states
    ap = argparse.ArgumentParser()
    ap.add_argument("--mech", default="synth_gri53")
    ap.add_argument("--n-train", type=int, default=1 << 23)
    ap.add_argument("--n-eval", type=int, default=1 << 20)
    ap.add_argument("--hidden", type=int, default=256)
    ap.add_argument("--layers", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=None, help="weights (.npz)")
    ap.add_argument("--report", default=None, help="JSON report")
    ap.add_argument("--load", default=None, help="skip training, load weights")
    ap.add_argument("--sweep", default=None,
                    help="accuracy against speed over network shapes, e.g. 256x2,256x4,512x4 (same data and epochs)")
    a = ap.parse_args(argv)
    if a.sweep:
        shapes = [tuple(int(v) for v in t.split("x")) for t in a.sweep.split(",")]
        rows = sweep(a.mech, shapes, a.n_train, a.epochs, a.n_eval, a.batch, a.lr,
                     log=lambda m: print(m, flush=True))
        if a.report:
            with open(a.report, "w") as fh:
                json.dump({"mechanism": a.mech, "sweep": rows}, fh, indent=1)
        return
    mech = load_bundled(a.mech)
    dm = DeviceMechanism(mech, best_variant(mech), build=False)
    sur = Surrogate(mech, hidden=a.hidden, n_hidden=a.layers, seed=0)
    report = {"mechanism": a.mech, "dims": sur.dims, "n_train": a.n_train, "epochs": a.epochs, "batch": a.batch}
    if a.load:
        sur.load_state_dict(dict(np.load(a.load)))
    else:
        t0 = time.perf_counter()
        P, T, Y = random_states(mech.n_species, a.n_train, seed=100)
        W = torch.empty_like(Y)
        dm.getrates_(P, T, Y, W)
        torch.cuda.synchronize()
        t_data = time.perf_counter() - t0
        sur.fit_scales(W)
        t0 = time.perf_counter()
        hist = sur.train(P, T, Y, W, epochs=a.epochs, batch=a.batch, lr=a.lr,
                         log=lambda s: print("[train]", s, flush=True))
        torch.cuda.synchronize()
        report.update({"data_s": t_data, "train_s": time.perf_counter() - t0, "loss_history": hist})
        del P, T, Y, W
        torch.cuda.empty_cache()
        if a.out:
            np.savez(a.out, **sur.state_dict())
    report["eval"] = evaluate(sur, dm, a.n_eval, seed=7)
    print(json.dumps(report["eval"]), flush=True)
    if a.report:
        with open(a.report, "w") as fh:
            json.dump(report, fh, indent=1)


if __name__ == "__main__":
    main()
