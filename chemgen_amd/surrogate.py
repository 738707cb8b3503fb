"""Learned getrates surrogate with a per-point FP64 fallback.

An MLP maps (T, log P, log Y_1..log Y_Ns) to z_k = asinh(wdot_k / s_k) (s_k: a per-species scale,
1e-3 of the 99th percentile of |wdot_k| over the training labels), hidden layers of width
256 with SiLU.  Weights and activations are BF16 on the matrix cores with FP32
accumulation (cg_mlp_gemm); master weights, pre-activations and Adam state are FP32.  The
network is trained on the GPU by this module's own forward and backward kernels on states
labelled by the FP64 getrates kernel.

Inference (:meth:`Surrogate.getrates_`) decodes wdot = s sinh(z) in FP64 and computes each
point's element-conservation residual max_e |sum_k a_ek wdot_k| / sum_k |a_ek wdot_k|; points
above ``tol`` are recomputed by the FP64 kernel.  An audit, run before the network touches
the rest of the batch, evaluates every ``audit_every``-th point both ways; if more than
``audit_max_bad`` of them are off by more than 1 % the whole batch goes to FP64, so a poor
network costs about one FP64 launch plus the audit, not the network plus FP64.

No counterpart exists in the reference (chemgen/chemistry_nn.py:15-85 is a non-functional
prototype of log-space Arrhenius layers).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .mechanism import Mechanism
from .runtime import ChemgenError, DeviceMechanism, check, lib

_BOUND = False


def _bind():
    global _BOUND
    if _BOUND:
        return
    L = lib()
    p, i64, i32, f32, f64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_float, ctypes.c_double
    sig = {
        "cg_mlp_gemm": [p, i64, i64, p, i64, i64, i32, i32, i32, i32, p, i32, p, i64, p, i64, i32, p],
        "cg_mlp_features": [p, p, p, i64, i64, i32, i32, p, p, p],
        "cg_mlp_targets": [p, i64, i64, i32, p, p, p],
        "cg_mlp_mse_grad": [p, p, i64, f32, p, p, p],
        "cg_mlp_silu_back": [p, p, i64, p, p],
        "cg_mlp_colsum": [p, i64, i32, p, p],
        "cg_mlp_adam": [p, p, p, p, i64, f32, f32, f32, f32, i32, p, p],
        "cg_mlp_to_bf16": [p, i64, p, p],
        "cg_mlp_decode": [p, i64, i32, p, p, i32, f64, p, i64, p, p, p],
        "cg_mlp_compact": [p, i64, i32, p, p, p, p],
        "cg_mlp_audit": [p, i64, p, i64, i64, i32, p, f64, f64, p, p, p, p, p],
        "cg_mlp_fallback_plan": [p, p, i64, f64, i32, p, p, p, p, p],
        "cg_count_tiles": [p, i64, i32, p, p, p],
        "cg_mlp_gather": [p, p, p, i64, p, p, i64, i32, p, p, p, i64, p],
        "cg_mlp_scatter": [p, i64, p, p, i64, i32, p, i64, p],
        "cg_mlp_pack": [p, i32, i32, i32, i32, p, p],
        "cg_mlp_infer": [p, p, p, i64, i64, i32, p, p, p, i32, i32, i32, p, p, i32, f64, p, i64, p, p, p, p],
    }
    for name, a in sig.items():
        f = getattr(L, name)
        f.argtypes = a
        f.restype = ctypes.c_int
    _BOUND = True


def _s(stream):
    import torch
    return int((stream or torch.cuda.current_stream()).cuda_stream)


def _ptr(t):
    return None if t is None else t.data_ptr()


@dataclass
class Layer:
    w: "object"          # FP32 master [N, K]
    b: "object"          # FP32 [N]
    w16: "object"        # BF16 copy [N, K]
    mw: "object" = None  # Adam moments
    vw: "object" = None
    mb: "object" = None
    vb: "object" = None
    gw: "object" = None  # gradients
    gb: "object" = None


class Surrogate:
    """MLP surrogate for one mechanism (see the module docstring)."""

    def __init__(self, mech: Mechanism, hidden: int = 256, n_hidden: int = 4, seed: int = 0,
                 device: str = "cuda:0"):
        import torch
        _bind()
        self.mech = mech
        self.ns = mech.n_species
        self.kin = ((self.ns + 3 + 7) // 8) * 8   # T, ln P, ln Y_k, 1000/T, zero padding
        self.dims = [self.kin] + [hidden] * n_hidden + [self.ns]
        self.dev = torch.device(device)
        g = torch.Generator(device="cpu").manual_seed(seed)
        self.layers: List[Layer] = []
        for l in range(len(self.dims) - 1):
            K, N = self.dims[l], self.dims[l + 1]
            w = (torch.randn(N, K, generator=g, dtype=torch.float32) * math.sqrt(2.0 / K)).to(self.dev)
            b = torch.zeros(N, dtype=torch.float32, device=self.dev)
            self.layers.append(Layer(w, b, w.to(torch.bfloat16)))
        # features: T, ln P, ln(Y + 1e-12) and 1000/T, centred and scaled to O(1)
        self.fscale = torch.tensor([1800.0, 1.0 / 600.0, math.log(1013250.0 * 6.3), 1.0 / 1.1, -5.0, 0.2,
                                    0.8, 2.0], dtype=torch.float32, device=self.dev)
        self.s = torch.ones(self.ns, dtype=torch.float64, device=self.dev)
        a = mech.element_matrix()
        a = a[np.abs(a).sum(axis=1) > 0]
        self.aek = torch.tensor(a, dtype=torch.float32, device=self.dev).contiguous()
        self.step_count = 0
        self._packed = None   # (weights in MFMA operand order, biases) for the fused inference kernel

    # ---------------- building blocks ----------------
    def features(self, P, T, Y, stream=None):
        import torch
        ng = P.shape[0]
        ld = Y.stride(0) if Y.shape[0] > 1 else Y.shape[1]
        X = torch.empty(ng, self.kin, dtype=torch.bfloat16, device=self.dev)
        check(lib().cg_mlp_features(P.data_ptr(), T.data_ptr(), Y.data_ptr(), ld, ng, self.ns, self.kin,
                                    self.fscale.data_ptr(), X.data_ptr(), _s(stream)), "cg_mlp_features")
        return X

    def targets(self, W, stream=None):
        import torch
        ng = W.shape[1]
        Z = torch.empty(ng, self.ns, dtype=torch.float32, device=self.dev)
        check(lib().cg_mlp_targets(W.data_ptr(), W.stride(0), ng, self.ns, self.s.data_ptr(), Z.data_ptr(),
                                   _s(stream)), "cg_mlp_targets")
        return Z

    def _gemm(self, A, sAi, sAk, B, sBk, sBj, M, N, K, bias=None, act=0, C32=None, ldc=0, Cb=None,
              ldcb=0, ksplit=1, accumulate=0, stream=None):
        check(lib().cg_mlp_gemm(A.data_ptr(), sAi, sAk, B.data_ptr(), sBk, sBj, M, N, K, ksplit, _ptr(bias),
                                act, _ptr(C32), ldc, _ptr(Cb), ldcb, accumulate, _s(stream)), "cg_mlp_gemm")

    def forward(self, X, train: bool = False, stream=None):
        """Zh [M, Ns] FP32; with ``train`` the pre-activations and inputs are kept."""
        import torch
        M = X.shape[0]
        H = X
        self._acts = [X] if train else None
        self._pre = [] if train else None
        nl = len(self.layers)
        for l, L in enumerate(self.layers):
            K, N = self.dims[l], self.dims[l + 1]
            last = l == nl - 1
            Zp = torch.empty(M, N, dtype=torch.float32, device=self.dev) if (train or last) else None
            Hn = None if last else torch.empty(M, N, dtype=torch.bfloat16, device=self.dev)
            self._gemm(H, K, 1, L.w16, 1, K, M, N, K, bias=L.b, act=0 if last else 1, C32=Zp, ldc=N,
                       Cb=Hn, ldcb=N, stream=stream)
            if train:
                self._pre.append(Zp)
                if not last:
                    self._acts.append(Hn)
            H = Hn
            if last:
                return Zp

    def backward(self, G, stream=None):
        """Gradients of all weights for the output gradient G [M, Ns] (BF16)."""
        import torch
        M = G.shape[0]
        ks = max(1, min(64, M // 4096))
        for l in range(len(self.layers) - 1, -1, -1):
            L = self.layers[l]
            K, N = self.dims[l], self.dims[l + 1]
            Hin = self._acts[l]
            if L.gw is None:
                L.gw = torch.zeros_like(L.w)
                L.gb = torch.zeros_like(L.b)
            else:
                L.gw.zero_()
                L.gb.zero_()
            # dW[N, K] = G^T H_in: A(n, m) = G[m, n], B(m, k) = H_in[m, k]
            self._gemm(G, 1, N, Hin, K, 1, N, K, M, C32=L.gw, ldc=K, ksplit=ks, accumulate=1, stream=stream)
            check(lib().cg_mlp_colsum(G.data_ptr(), M, N, L.gb.data_ptr(), _s(stream)), "cg_mlp_colsum")
            if l > 0:
                # dH[M, K] = G W: A(m, n) = G[m, n], B(n, k) = W[n, k]
                dH = torch.empty(M, K, dtype=torch.float32, device=self.dev)
                self._gemm(G, N, 1, L.w16, K, 1, M, K, N, C32=dH, ldc=K, stream=stream)
                Gn = torch.empty(M, K, dtype=torch.bfloat16, device=self.dev)
                check(lib().cg_mlp_silu_back(dH.data_ptr(), self._pre[l - 1].data_ptr(), M * K, Gn.data_ptr(),
                                             _s(stream)), "cg_mlp_silu_back")
                G = Gn

    def loss_grad(self, Zh, Z, stream=None):
        import torch
        n = Zh.numel()
        G = torch.empty(Zh.shape, dtype=torch.bfloat16, device=self.dev)
        loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        check(lib().cg_mlp_mse_grad(Zh.data_ptr(), Z.data_ptr(), n, 1.0 / n, G.data_ptr(), loss.data_ptr(),
                                    _s(stream)), "cg_mlp_mse_grad")
        return G, loss

    def adam(self, lr: float, b1=0.9, b2=0.999, eps=1e-8, stream=None):
        import torch
        self.step_count += 1
        self._packed = None
        for L in self.layers:
            if L.mw is None:
                L.mw, L.vw = torch.zeros_like(L.w), torch.zeros_like(L.w)
                L.mb, L.vb = torch.zeros_like(L.b), torch.zeros_like(L.b)
            check(lib().cg_mlp_adam(L.w.data_ptr(), L.gw.data_ptr(), L.mw.data_ptr(), L.vw.data_ptr(), L.w.numel(),
                                    lr, b1, b2, eps, self.step_count, L.w16.data_ptr(), _s(stream)), "cg_mlp_adam")
            check(lib().cg_mlp_adam(L.b.data_ptr(), L.gb.data_ptr(), L.mb.data_ptr(), L.vb.data_ptr(), L.b.numel(),
                                    lr, b1, b2, eps, self.step_count, None, _s(stream)), "cg_mlp_adam")

    # ---------------- training ----------------
    def fit_scales(self, W):
        """s_k = 1e-3 x 99th percentile of |wdot_k| over labels W [Ns, n]."""
        import torch
        q = torch.quantile(W.abs()[:, : min(W.shape[1], 1 << 20)].float(), 0.99, dim=1).double()
        self.s = torch.clamp(q * 1e-3, min=1e-30).contiguous()

    def train(self, P, T, Y, W, epochs: int = 4, batch: int = 8192, lr: float = 2e-3, log=None):
        """Adam on (features(P,T,Y), asinh(W/s)); W: FP64 labels [Ns, n] from the FP64 kernel."""
        import torch
        n = P.shape[0]
        X = self.features(P, T, Y)
        Z = self.targets(W)
        steps = epochs * (n // batch)
        hist = []
        g = torch.Generator(device=self.dev).manual_seed(1)
        s = 0
        for ep in range(epochs):
            perm = torch.randperm(n, device=self.dev, generator=g)
            tot = 0.0
            for b0 in range(0, n - batch + 1, batch):
                idx = perm[b0:b0 + batch]
                Xb, Zb = X[idx].contiguous(), Z[idx].contiguous()
                Zh = self.forward(Xb, train=True)
                G, loss = self.loss_grad(Zh, Zb)
                self.backward(G)
                cur = lr * 0.5 * (1.0 + math.cos(math.pi * s / max(1, steps)))
                self.adam(cur)
                s += 1
                if b0 % (batch * 64) == 0:
                    tot = float(loss.item())
                    hist.append((ep, s, tot))
                    if log:
                        log(f"epoch {ep} step {s} loss {tot:.4e} lr {cur:.2e}")
        return hist

    # ---------------- inference with fallback ----------------
    @property
    def fused_ok(self) -> bool:
        """Whether the one-launch inference kernel covers this network's shape."""
        hid = self.dims[1:-1]
        return (len(hid) >= 1 and len(set(hid)) == 1 and hid[0] in (256, 512) and self.kin <= hid[0]
                and self.ns <= hid[0] and self.aek.shape[0] <= 8)

    def pack(self, stream=None):
        """Weights in the fused kernel's operand order and the biases back to back.  BF16
        weights changed directly (not through adam/load_state_dict) need ``_packed = None``."""
        import torch
        if self._packed is not None:
            return self._packed
        H = self.dims[1]
        kch, nth = H // 32, H // 16
        shapes = [(nth, (self.kin + 31) // 32)] + [(nth, kch)] * (len(self.dims) - 3) + [((self.ns + 15) // 16, kch)]
        Wf = torch.empty(sum(nt * kc * 512 for nt, kc in shapes), dtype=torch.bfloat16, device=self.dev)
        off = 0
        for (nt, kc), L, K, N in zip(shapes, self.layers, self.dims[:-1], self.dims[1:]):
            check(lib().cg_mlp_pack(L.w16.data_ptr(), N, K, nt, kc, Wf[off:].data_ptr(), _s(stream)), "cg_mlp_pack")
            off += nt * kc * 512
        bias = torch.cat([L.b for L in self.layers]).contiguous()
        self._packed = (Wf, bias)
        return self._packed

    def predict_(self, P, T, Y, W, tol: float = 1e-3, stream=None, fused: Optional[bool] = None,
                 flag=None, resid=None, skip=None):
        """Surrogate wdot into W [Ns, ld] (FP64) plus per-point flags/residuals (into ``flag`` /
        ``resid`` when given).  ``fused`` (default: whenever the shape allows) runs the
        one-launch inference kernel, which ``skip`` (a device int32, fused only) can cancel on
        the device; otherwise the layer-by-layer products of the training path are used."""
        import torch
        ng = P.shape[0]
        flag = torch.empty(ng, dtype=torch.int32, device=self.dev) if flag is None else flag
        resid = torch.empty(ng, dtype=torch.float64, device=self.dev) if resid is None else resid
        ld = W.stride(0) if W.shape[0] > 1 else W.shape[1]
        if fused is None:
            fused = self.fused_ok
        if fused:
            if not self.fused_ok:
                raise ChemgenError(f"fused inference does not cover network {self.dims}")
            Wf, bias = self.pack(stream)
            ldy = Y.stride(0) if Y.shape[0] > 1 else Y.shape[1]
            check(lib().cg_mlp_infer(P.data_ptr(), T.data_ptr(), Y.data_ptr(), ldy, ng, self.ns,
                                     self.fscale.data_ptr(), Wf.data_ptr(), bias.data_ptr(), self.dims[1], self.kin,
                                     len(self.dims) - 2, self.s.data_ptr(), self.aek.data_ptr(), self.aek.shape[0],
                                     tol, W.data_ptr(), ld, flag.data_ptr(), resid.data_ptr(), _ptr(skip),
                                     _s(stream)),
                  "cg_mlp_infer")
            return flag, resid
        X = self.features(P, T, Y, stream)
        Zh = self.forward(X, stream=stream)
        check(lib().cg_mlp_decode(Zh.data_ptr(), ng, self.ns, self.s.data_ptr(), self.aek.data_ptr(),
                                  self.aek.shape[0], tol, W.data_ptr(), ld, flag.data_ptr(), resid.data_ptr(),
                                  _s(stream)), "cg_mlp_decode")
        return flag, resid

    def _gather(self, P, T, Y, idx, count, n, out=None, stream=None):
        """States of the points idx[:*count] (at most n) packed into [n] / [Ns, n] tensors (new,
        or the three of ``out``)."""
        import torch
        if out is None:
            Pp = torch.empty(n, dtype=torch.float64, device=self.dev)
            Tp = torch.empty_like(Pp)
            Yp = torch.empty(self.ns, n, dtype=torch.float64, device=self.dev)
        else:
            Pp, Tp, Yp = out
        ldy = Y.stride(0) if Y.shape[0] > 1 else Y.shape[1]
        check(lib().cg_mlp_gather(P.data_ptr(), T.data_ptr(), Y.data_ptr(), ldy, idx.data_ptr(), count.data_ptr(),
                                  n, self.ns, Pp.data_ptr(), Tp.data_ptr(), Yp.data_ptr(), Yp.shape[1], _s(stream)),
              "cg_mlp_gather")
        return Pp, Tp, Yp

    def _scatter(self, Wp, idx, count, n, W, stream=None):
        ldw = W.stride(0) if W.shape[0] > 1 else W.shape[1]
        check(lib().cg_mlp_scatter(Wp.data_ptr(), Wp.shape[1], idx.data_ptr(), count.data_ptr(), n, self.ns,
                                   W.data_ptr(), ldw, _s(stream)), "cg_mlp_scatter")

    def error(self, Wsur, Wref):
        """Per-point surrogate error max_k |w_sur - w| / (|w| + s_k) (the accuracy measure)."""
        return ((Wsur - Wref).abs() / (Wref.abs() + self.s[:, None])).max(dim=0).values

    def _workspace(self, ng: int, audit_every: int, tile: int):
        """Device buffers of getrates_ for ng points, allocated once per (ng, audit_every, tile)."""
        import torch
        key = (ng, audit_every, tile)
        ws = getattr(self, "_ws", None)
        if ws is not None and ws["key"] == key:
            return ws
        d = self.dev
        i32 = lambda n: torch.zeros(n, dtype=torch.int32, device=d)
        nt = (ng + tile - 1) // tile
        na = (ng + audit_every - 1) // audit_every if audit_every > 0 else 0
        ws = {"key": key, "na": na, "flag": i32(ng), "resid": torch.empty(ng, dtype=torch.float64, device=d),
              "reject": i32(1), "n_bad": i32(1), "count": i32(1), "full": i32(1), "pcount": i32(1),
              "acc": i32(2),      # the audit's ticket counters, left zero by each call
              "idx": torch.zeros(ng, dtype=torch.int64, device=d),
              "packed": (torch.empty(ng, dtype=torch.float64, device=d), torch.empty(ng, dtype=torch.float64, device=d),
                         torch.empty(self.ns, ng, dtype=torch.float64, device=d)),
              "Wp": torch.empty(self.ns, ng, dtype=torch.float64, device=d),
              "ts": torch.arange(0, nt * tile, tile, dtype=torch.int64, device=d), "tl_p": i32(nt), "tl_f": i32(nt),
              "side": torch.cuda.Stream(device=d)}
        if na:
            ws["aidx"] = torch.arange(0, ng, audit_every, dtype=torch.int64, device=d)
            ws["acount"] = torch.full((1,), na, dtype=torch.int32, device=d)
            ws["apacked"] = (torch.empty(na, dtype=torch.float64, device=d), torch.empty(na, dtype=torch.float64, device=d),
                             torch.empty(self.ns, na, dtype=torch.float64, device=d))
            ws["Wa"] = torch.empty(self.ns, na, dtype=torch.float64, device=d)
            ws["Wsa"] = torch.empty(self.ns, na, dtype=torch.float64, device=d)
            ws["aflag"] = i32(na)
            ws["aresid"] = torch.empty(na, dtype=torch.float64, device=d)
        self._ws = ws
        return ws

    def getrates_(self, P, T, Y, W, fallback: DeviceMechanism, tol: float = 1e-3, audit_every: Optional[int] = None,
                  audit_max_bad: float = 0.05, dense_frac: float = 0.5, stream=None) -> dict:
        """wdot into W via surrogate + per-point FP64 fallback, with no host synchronisation.

        1. Audit: every ``audit_every``-th point (default max(16, ng // 4096), about 4096
           points) is evaluated by the network and, on a side stream at the same time, by the
           FP64 kernel; the device counts the audited points off by more than 1 % and sets
           ``reject`` if they are more than ``audit_max_bad`` of them (cg_mlp_audit).
        2. The network runs over the batch unless ``reject`` (cg_mlp_infer reads the flag and
           returns at once); the audited points get their FP64 values.
        3. The points whose element-conservation residual exceeds ``tol`` (or is not finite) are
           compacted on the device; if the audit rejected the network or more than
           ``dense_frac`` of the points are flagged, the whole batch goes to one FP64 launch,
           else the flagged points are recomputed packed (cg_mlp_compact,
           cg_mlp_fallback_plan).  Both FP64 launches run over tile lists whose lengths follow
           those device counts
           (cg_mlp_fallback_plan, cg_getrates_tiles), so the launch sequence is fixed and the
           whole path can be captured into a HIP graph (after one eager call with the same ng,
           which allocates the workspace).
        Only the fused inference kernel can be cancelled on the device: networks it does not
        cover run layer by layer even when rejected (the FP64 launch then overwrites them).
        Runs on ``stream`` (default: the current stream).  Returns the device counters;
        ``fallback_info`` turns them into numbers (that synchronises)."""
        import contextlib
        import torch
        ng = P.shape[0]
        if ng == 0:
            return {"ng": 0}
        if ng > 0x7fffffff:
            raise ChemgenError("Surrogate.getrates_: at most 2^31 - 1 points per call")
        if audit_every is None:
            audit_every = max(16, ng // 4096)
        # the torch work below runs on the current stream: make it ``stream``
        with torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext():
            return self._getrates_dev(P, T, Y, W, fallback, tol, audit_every, audit_max_bad, dense_frac)

    def _getrates_dev(self, P, T, Y, W, fallback, tol, audit_every, audit_max_bad, dense_frac):
        import torch
        ng = P.shape[0]
        # only the fused kernel can be cancelled on the device; the layer-by-layer path (networks
        # it does not cover) runs even when the audit rejects, and the whole-batch FP64 launch then
        # overwrites its output
        fused = self.fused_ok
        tile = fallback.block
        ws = self._workspace(ng, audit_every, tile)
        cur = torch.cuda.current_stream()
        st = int(cur.cuda_stream)
        na = ws["na"]
        if na:
            Pa, Ta, Ya = self._gather(P, T, Y, ws["aidx"], ws["acount"], na, out=ws["apacked"], stream=cur)
            side = ws["side"]
            side.wait_stream(cur)
            fallback.getrates_(Pa, Ta, Ya, ws["Wa"], stream=side)
            self.predict_(Pa, Ta, Ya, ws["Wsa"], tol, stream=cur, fused=fused, flag=ws["aflag"], resid=ws["aresid"])
            cur.wait_stream(side)
            check(lib().cg_mlp_audit(ws["Wsa"].data_ptr(), na, ws["Wa"].data_ptr(), na, na, self.ns,
                                     self.s.data_ptr(), 1e-2, audit_max_bad, ws["reject"].data_ptr(),
                                     ws["n_bad"].data_ptr(), ws["acc"].data_ptr(), ws["count"].data_ptr(), st),
                  "cg_mlp_audit")
        else:
            ws["reject"].zero_()
            ws["count"].zero_()
        self.predict_(P, T, Y, W, tol, stream=cur, fused=fused, flag=ws["flag"], resid=ws["resid"],
                      skip=ws["reject"] if fused else None)
        if na:
            self._scatter(ws["Wa"], ws["aidx"], ws["acount"], na, W, stream=cur)
        check(lib().cg_mlp_compact(ws["flag"].data_ptr(), ng, audit_every, ws["count"].data_ptr(),
                                   ws["idx"].data_ptr(), ws["reject"].data_ptr(), st), "cg_mlp_compact")
        check(lib().cg_mlp_fallback_plan(ws["count"].data_ptr(), ws["reject"].data_ptr(), ng, dense_frac, tile,
                                         ws["full"].data_ptr(), ws["pcount"].data_ptr(), ws["tl_p"].data_ptr(),
                                         ws["tl_f"].data_ptr(), st), "cg_mlp_fallback_plan")
        # packed recompute of the flagged points, on a bounded grid (the tile kernel strides over
        # the list, so an empty list costs one short launch)
        Pp, Tp, Yp = self._gather(P, T, Y, ws["idx"], ws["pcount"], ng, out=ws["packed"], stream=cur)
        fallback.getrates_tiles_(Pp, Tp, Yp, ws["Wp"], ws["ts"], ws["tl_p"], ng, grid=512, stream=cur)
        self._scatter(ws["Wp"], ws["idx"], ws["pcount"], ng, W, stream=cur)
        # the whole batch in FP64 (zero-length tiles unless the plan says so)
        fallback.getrates_tiles_(P, T, Y, W, ws["ts"], ws["tl_f"], ng, grid=-1, stream=cur)
        return {"ng": ng, "na": na, "reject": ws["reject"], "n_bad": ws["n_bad"], "count": ws["pcount"],
                "full": ws["full"]}

    @staticmethod
    def fallback_info(d: dict) -> dict:
        """Host numbers of a getrates_ result (synchronises with the device)."""
        if d.get("ng", 0) == 0:
            return {"fallback_points": 0, "audit_points": 0, "audit_bad_frac": 0.0, "global_fallback": False}
        full = int(d["full"].item())
        na = d["na"]
        return {"fallback_points": full if full else int(d["count"].item()),
                "audit_points": na, "audit_bad_frac": int(d["n_bad"].item()) / max(1, na),
                "global_fallback": bool(full), "audit_rejected": bool(int(d["reject"].item()))}

    def state_dict(self) -> dict:
        return {"dims": self.dims, "fscale": self.fscale.cpu().numpy(), "s": self.s.cpu().numpy(),
                **{f"w{l}": L.w.cpu().numpy() for l, L in enumerate(self.layers)},
                **{f"b{l}": L.b.cpu().numpy() for l, L in enumerate(self.layers)}}

    def load_state_dict(self, d: dict) -> None:
        import torch
        for l, L in enumerate(self.layers):
            L.w.copy_(torch.from_numpy(np.asarray(d[f"w{l}"])))
            L.b.copy_(torch.from_numpy(np.asarray(d[f"b{l}"])))
            L.w16.copy_(L.w.to(torch.bfloat16))
        self._packed = None
        self.s = torch.from_numpy(np.asarray(d["s"])).to(self.dev).contiguous()
        self.fscale = torch.from_numpy(np.asarray(d["fscale"])).to(self.dev).contiguous()


def random_states(ns: int, n: int, seed: int, T_range=(800.0, 2800.0), P_atm=(1.0, 40.0), device="cuda:0",
                  chunk: int = 1 << 21):
    """Seeded states drawn on the GPU (T uniform, P log-uniform, Y Dirichlet(0.5)); Y is drawn
    ``chunk`` points at a time, so the temporaries stay small next to Y itself."""
    import torch
    g = torch.Generator(device=device).manual_seed(seed)
    T = torch.rand(n, generator=g, device=device, dtype=torch.float64) * (T_range[1] - T_range[0]) + T_range[0]
    lp = torch.rand(n, generator=g, device=device, dtype=torch.float64) * math.log(P_atm[1] / P_atm[0])
    P = 1013250.0 * P_atm[0] * torch.exp(lp)
    Y = torch.empty(ns, n, dtype=torch.float64, device=device)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        gam = torch._standard_gamma(torch.full((ns, b - a), 0.5, dtype=torch.float64, device=device), generator=g)
        Y[:, a:b] = gam / gam.sum(dim=0, keepdim=True)
        del gam
    return P.contiguous(), T.contiguous(), Y
