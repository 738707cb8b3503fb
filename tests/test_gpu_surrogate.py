"""Surrogate kernels on the GPU: MFMA products vs NumPy, gradients vs the CPU oracle,
fallback behaviour, and what training gains."""
import numpy as np
import pytest

from chemgen_amd.kernels import load_bundled
from oracle.mlp_np import bf16, loss_and_grads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


def _bf(torch, a):
    return torch.from_numpy(np.asarray(a, dtype=np.float32)).to("cuda:0").to(torch.bfloat16)


@pytest.mark.parametrize("M,N,K,mode", [
    (64, 64, 32, "nt"), (200, 96, 72, "nt"), (37, 53, 264, "nt"),
    (130, 70, 50, "nn"), (256, 48, 1000, "tn_split"),
])
def test_mfma_gemm_matches_numpy(torch, M, N, K, mode):
    from chemgen_amd.surrogate import Surrogate, _bind
    from chemgen_amd.runtime import lib
    _bind()
    rng = np.random.default_rng(M + N + K)
    A = bf16(rng.normal(size=(M, K)))
    B = bf16(rng.normal(size=(K, N)))
    bias = rng.normal(size=N).astype(np.float32)
    ref = A @ B + bias.astype(np.float64)
    C32 = torch.zeros(M, N, dtype=torch.float32, device="cuda:0")
    Cb = torch.zeros(M, N, dtype=torch.bfloat16, device="cuda:0")
    bt = torch.from_numpy(bias).to("cuda:0")
    s = int(torch.cuda.current_stream().cuda_stream)
    print(f"[gemm] {M}x{N}x{K} {mode}", flush=True)
    if mode == "nt":    # A row-major [M,K], B given as W[N,K] (B(k,j) = W[j,k])
        Ad, Wd = _bf(torch, A), _bf(torch, B.T.copy())
        rc = lib().cg_mlp_gemm(Ad.data_ptr(), K, 1, Wd.data_ptr(), 1, K, M, N, K, 1, bt.data_ptr(), 1,
                               C32.data_ptr(), N, Cb.data_ptr(), N, 0, s)
    elif mode == "nn":  # B row-major [K,N]
        Ad, Bd = _bf(torch, A), _bf(torch, B)
        rc = lib().cg_mlp_gemm(Ad.data_ptr(), K, 1, Bd.data_ptr(), N, 1, M, N, K, 1, bt.data_ptr(), 1,
                               C32.data_ptr(), N, Cb.data_ptr(), N, 0, s)
    else:               # A given transposed (At[K,M]), split K with atomic accumulation
        Atd, Bd = _bf(torch, A.T.copy()), _bf(torch, B)
        rc = lib().cg_mlp_gemm(Atd.data_ptr(), 1, M, Bd.data_ptr(), N, 1, M, N, K, 4, bt.data_ptr(), 0,
                               C32.data_ptr(), N, None, 0, 1, s)
    assert rc == 0, lib().cg_last_error()
    torch.cuda.synchronize()
    c = C32.cpu().numpy().astype(np.float64)
    scale = np.abs(A) @ np.abs(B) + np.abs(bias)
    assert np.max(np.abs(c - ref) / scale) < 1e-5
    if mode != "tn_split":
        silu = ref / (1 + np.exp(-ref))
        cb = Cb.float().cpu().numpy()
        assert np.max(np.abs(cb - silu) / (np.abs(silu) + 1e-2)) < 2 ** -7


def test_gradients_match_cpu_oracle(torch):
    from chemgen_amd.surrogate import Surrogate, random_states
    from chemgen_amd.runtime import DeviceMechanism
    m = load_bundled("h2_burke")
    sur = Surrogate(m, hidden=32, n_hidden=2, seed=3)
    P, T, Y = random_states(m.n_species, 300, seed=5)
    W = DeviceMechanism(m, build=False).getrates(P, T, Y)
    sur.fit_scales(W)
    X = sur.features(P, T, Y)
    Z = sur.targets(W)
    Zh = sur.forward(X, train=True)
    G, loss = sur.loss_grad(Zh, Z)
    sur.backward(G)
    torch.cuda.synchronize()
    Xn = X.float().cpu().numpy().astype(np.float64)
    Ws = [L.w.cpu().numpy().astype(np.float64) for L in sur.layers]
    bs = [L.b.cpu().numpy().astype(np.float64) for L in sur.layers]
    ref_loss, gW, gb = loss_and_grads(Xn, Z.cpu().numpy().astype(np.float64), Ws, bs, quantize=True)
    assert float(loss.item()) == pytest.approx(ref_loss, rel=1e-3)
    for l, L in enumerate(sur.layers):
        g = L.gw.cpu().numpy()
        assert np.linalg.norm(g - gW[l]) <= 1e-2 * np.linalg.norm(gW[l]), l
        assert np.linalg.norm(L.gb.cpu().numpy() - gb[l]) <= 1e-2 * np.linalg.norm(gb[l]) + 1e-12, l


def test_fallback_paths(torch):
    """tol = 0 sends every point to the FP64 kernel, through the whole-batch launch and
    through the packed sub-batch (bit-identical output both ways); weights forced to
    nonsense from the host are caught by the conservation residual, and by the audit
    alone when the residual test is switched off."""
    from chemgen_amd.surrogate import Surrogate, random_states
    from chemgen_amd.runtime import DeviceMechanism
    m = load_bundled("h2_burke")
    dm = DeviceMechanism(m, build=False)
    sur = Surrogate(m, hidden=32, n_hidden=2, seed=1)
    P, T, Y = random_states(m.n_species, 5000, seed=9)
    ref = dm.getrates(P, T, Y)
    sur.fit_scales(ref)
    for dense in (0.5, 1.0):
        W = torch.full_like(Y, float("nan"))
        info = sur.fallback_info(sur.getrates_(P, T, Y, W, dm, tol=0.0, audit_every=0, dense_frac=dense))
        torch.cuda.synchronize()
        assert info["fallback_points"] == 5000
        assert torch.equal(W, ref), dense
    for L in sur.layers:       # garbage network: outputs violate element conservation
        L.w16.copy_((L.w * 40.0).to(torch.bfloat16))
    W = torch.empty_like(Y)
    info = sur.fallback_info(sur.getrates_(P, T, Y, W, dm, tol=1e-3, audit_every=0, dense_frac=1.0))
    torch.cuda.synchronize()
    assert info["fallback_points"] > 4900 and not info["global_fallback"]
    err = sur.error(W, ref)
    assert float((err <= 1e-12).double().mean()) >= info["fallback_points"] / 5000 - 1e-9
    W = torch.empty_like(Y)
    info = sur.fallback_info(sur.getrates_(P, T, Y, W, dm, tol=float("inf"), audit_every=16))
    torch.cuda.synchronize()
    print(info)
    assert info["global_fallback"] and info["audit_points"] == 313 and info["audit_bad_frac"] > 0.5
    assert torch.equal(W, ref)


def test_audit_points_take_fp64_values(torch):
    """With the residual test off and the audit passing (threshold 1), the audited points
    hold the FP64 kernel's values and every other point the network's."""
    from chemgen_amd.surrogate import Surrogate, random_states
    from chemgen_amd.runtime import DeviceMechanism
    m = load_bundled("h2_burke")
    dm = DeviceMechanism(m, build=False)
    sur = Surrogate(m, hidden=32, n_hidden=2, seed=4)
    P, T, Y = random_states(m.n_species, 3000, seed=13)
    ref = dm.getrates(P, T, Y)
    sur.fit_scales(ref)
    Wsur = torch.empty_like(Y)
    sur.predict_(P, T, Y, Wsur, float("inf"))
    W = torch.empty_like(Y)
    info = sur.fallback_info(sur.getrates_(P, T, Y, W, dm, tol=float("inf"), audit_every=7, audit_max_bad=1.0))
    torch.cuda.synchronize()
    assert info["fallback_points"] == 0 and info["audit_points"] == 429
    mask = torch.zeros(3000, dtype=torch.bool, device=W.device)
    mask[::7] = True
    assert torch.equal(W[:, mask], ref[:, mask])
    assert torch.equal(W[:, ~mask], Wsur[:, ~mask])


def test_surrogate_path_against_fp64(torch):
    """The trained GRI-sized network (4 x 256, chemgen_amd/data) against the plain FP64 path on
    2^20 held-out states: the surrogate path's output is within 1 % of the FP64 kernel's on
    >= 99 % of points (the fallback's guarantee); training gains a network error far below the
    untrained network's; and the audit-first fallback keeps the whole path within 15 % of the
    FP64 kernel's time when the network is rejected (the audit costs a fixed ~0.1 ms: one
    network tile and one FP64 tile side by side, then a handful of small launches), where
    running the network and then the FP64 kernel would cost ~2.5 x.  Numbers are printed."""
    import os
    from chemgen_amd.autotune import best_variant
    from chemgen_amd.runtime import DeviceMechanism
    from chemgen_amd.surrogate import Surrogate, random_states
    m = load_bundled("synth_gri53")
    dm = DeviceMechanism(m, best_variant(m), build=False)
    sur = Surrogate(m, hidden=256, n_hidden=4, seed=0)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sur.load_state_dict(dict(np.load(os.path.join(root, "chemgen_amd", "data", "surrogate_synth_gri53.npz"))))
    untrained = Surrogate(m, hidden=256, n_hidden=4, seed=0)
    untrained.s = sur.s.clone()
    ng = 1 << 20
    P, T, Y = random_states(m.n_species, ng, seed=7)
    ref = torch.empty_like(Y)
    dm.getrates_(P, T, Y, ref)
    Wn = torch.empty_like(Y)
    sur.predict_(P, T, Y, Wn)
    Wu = torch.empty_like(Y)
    untrained.predict_(P, T, Y, Wu)
    err_t, err_u = sur.error(Wn, ref), sur.error(Wu, ref)
    W = torch.empty_like(Y)
    info = sur.fallback_info(sur.getrates_(P, T, Y, W, dm))
    err = sur.error(W, ref)

    def timed(fn, reps=20):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    t64 = timed(lambda: dm.getrates_(P, T, Y, ref))
    tpath = timed(lambda: sur.getrates_(P, T, Y, W, dm))
    tnet = timed(lambda: sur.predict_(P, T, Y, Wn))
    print(f"[surrogate vs fp64] median error trained {float(err_t.median()):.3g} untrained "
          f"{float(err_u.median()):.3g}; network within 1%: {float((err_t <= 1e-2).double().mean()):.4f}; "
          f"path within 1%: {float((err <= 1e-2).double().mean()):.4f}; {info}; ms fp64 {t64:.4f} "
          f"path {tpath:.4f} network {tnet:.4f}", flush=True)
    assert float((err <= 1e-2).double().mean()) >= 0.99
    assert float(err_t.median()) < 0.5 * float(err_u.median())
    if info["global_fallback"]:
        assert tpath <= 1.15 * t64


@pytest.mark.parametrize("mode", ["reject", "packed"])
def test_surrogate_path_graph_replay(torch, mode):
    """The whole surrogate path (audit, guarded inference, device-side fallback plan, packed and
    whole-batch FP64 launches) has no host synchronisation: captured into a HIP graph and
    replayed three times, with eager work between, every replay equals the eager result bit for
    bit.  "reject": an untrained network that the audit rejects (output = the FP64 kernel's);
    "packed": the audit passes (threshold 1) and the residual test flags points, which are
    recomputed packed (those points = FP64, the others = the network)."""
    from chemgen_amd.runtime import DeviceMechanism
    from chemgen_amd.surrogate import Surrogate, random_states
    m = load_bundled("h2_burke")
    dm = DeviceMechanism(m, build=False)
    sur = Surrogate(m, hidden=256, n_hidden=2, seed=3)
    ng = 20000
    P, T, Y = random_states(m.n_species, ng, seed=17)
    ref = dm.getrates(P, T, Y)
    sur.fit_scales(ref)
    kw = {} if mode == "reject" else {"audit_max_bad": 1.0, "tol": 1e-3, "dense_frac": 1.0}
    eager = torch.full_like(Y, -1.0)
    d = sur.getrates_(P, T, Y, eager, dm, **kw)
    info = sur.fallback_info(d)
    print(f"[surrogate graph] {mode} {info}", flush=True)
    # the FP64 values come from the tile kernels, compiled apart from the plain kernel that made
    # ref (its FMA contractions may differ in the last bit): compared at 1e-10 of the error measure
    close = lambda a, b: float(sur.error(a, b).max()) <= 1e-10
    if mode == "reject":
        assert info["global_fallback"] and close(eager, ref)
    else:
        assert not info["global_fallback"] and 0 < info["fallback_points"] < ng
        Wsur = torch.empty_like(Y)
        flag, _ = sur.predict_(P, T, Y, Wsur)
        flag[::64] = 0
        fb = flag.bool()
        mask = torch.zeros(ng, dtype=torch.bool, device=Y.device)
        mask[::64] = True
        assert close(eager[:, fb | mask], ref[:, fb | mask])
        assert torch.equal(eager[:, ~(fb | mask)], Wsur[:, ~(fb | mask)])
    W = torch.full_like(Y, -1.0)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            sur.getrates_(P, T, Y, W, dm, stream=s, **kw)
    torch.cuda.synchronize()
    for rep in range(3):
        W.fill_(-1.0)
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(W, eager), f"replay {rep}"
        _ = (Y * 2.0).sum().item()   # eager work between the replays


def test_features_match_cpu_oracle(torch):
    """Network inputs computed on the GPU equal the NumPy restatement after BF16 rounding
    (a last-bit difference of the FP64 log may move a value by one BF16 step, rarely)."""
    from chemgen_amd.surrogate import Surrogate, random_states
    from oracle.mlp_np import features as np_features
    m = load_bundled("synth_gri53")
    sur = Surrogate(m, hidden=256, n_hidden=1, seed=0)
    P, T, Y = random_states(m.n_species, 3000, seed=21)
    X = sur.features(P, T, Y).float().cpu().numpy().astype(np.float64)
    ref = bf16(np_features(P.cpu().numpy(), T.cpu().numpy(), Y.cpu().numpy(), sur.fscale.cpu().numpy(), sur.kin))
    assert X.shape == ref.shape == (3000, 56)
    diff = np.abs(X - ref)
    assert float((diff == 0).mean()) >= 0.999
    assert np.all(diff <= 2.0 ** -7 * np.maximum(np.abs(ref), 1e-30))


@pytest.mark.parametrize("name,hidden,layers,ng", [
    # H = 256 picks its tile shape from the batch: at most 64 points per CU (4 wavefronts x 16
    # points), 128 (8 x 16), more (8 x 32) -- ng 5000, 20000, 40000 on a 256-CU part
    ("h2_burke", 256, 2, 1000), ("synth_gri53", 256, 4, 5000), ("synth_gri53", 256, 4, 20000),
    ("synth_gri53", 256, 4, 40000), ("h2_burke2012", 512, 2, 777),
    ("synth_nhept88", 512, 3, 300),
    ("synth_jetsurf348", 512, 5, 600),   # the JetSurf-sized network: 352 -> 5 x 512 -> 348
])
def test_fused_inference_matches_oracle(torch, name, hidden, layers, ng):
    """The one-launch inference kernel against the CPU oracle (features -> BF16 MLP -> decode)
    on a strided Y, with a point count that leaves a partial tile; and against the
    layer-by-layer GPU path."""
    from chemgen_amd.surrogate import Surrogate, random_states
    from chemgen_amd.runtime import DeviceMechanism
    from oracle.mlp_np import decode as np_decode, features as np_features, forward as np_forward
    m = load_bundled(name)
    dm = DeviceMechanism(m, build=False)
    sur = Surrogate(m, hidden=hidden, n_hidden=layers, seed=7)
    g = torch.Generator().manual_seed(5)
    for L in sur.layers:          # biases away from zero, so the bias path is exercised
        L.b.copy_((torch.randn(L.b.shape, generator=g) * 0.3).to(L.b.device))
    P, T, Y0 = random_states(m.n_species, ng, seed=31)
    sur.fit_scales(dm.getrates(P, T, Y0))
    Y = torch.zeros(m.n_species, ng + 37, dtype=torch.float64, device="cuda:0")[:, :ng]
    Y.copy_(Y0)
    W = torch.zeros(m.n_species, ng + 5, dtype=torch.float64, device="cuda:0")[:, :ng]
    print(f"[fused] {name} H={hidden} L={layers} ng={ng}", flush=True)
    flag, resid = sur.predict_(P, T, Y, W, tol=1e-3, fused=True)
    Wl = torch.empty(m.n_species, ng, dtype=torch.float64, device="cuda:0")
    sur.predict_(P, T, Y, Wl, tol=1e-3, fused=False)
    torch.cuda.synchronize()
    s = sur.s.cpu().numpy()
    Xn = bf16(np_features(P.cpu().numpy(), T.cpu().numpy(), Y0.cpu().numpy(), sur.fscale.cpu().numpy(), sur.kin))
    Ws = [L.w16.float().cpu().numpy().astype(np.float64) for L in sur.layers]
    bs = [L.b.cpu().numpy().astype(np.float64) for L in sur.layers]
    z_ref = np_forward(Xn, Ws, bs, quantize=True)[0]
    Wn = W.cpu().numpy()
    z = np.arcsinh(Wn / s[:, None]).T
    assert np.all(np.isfinite(z))
    # BF16 activations: where the kernel's FP32 sums and the oracle's FP64 sums round an
    # activation to neighbouring BF16 values, an output moves by up to ~2^-8 of its size
    # per layer; a wrong fragment layout or permutation moves it by O(1)
    tol = 2.0 ** -8 * (layers + 1)
    assert np.max(np.abs(z - z_ref) / (1.0 + np.abs(z_ref))) <= tol
    zl = np.arcsinh(Wl.cpu().numpy() / s[:, None]).T
    assert np.max(np.abs(z - zl) / (1.0 + np.abs(zl))) <= tol
    # residual and flag of the kernel's own wdot as the restated decode defines them (the
    # kernel's sinh and residual sums are single precision: ~1e-7 relative)
    w_chk, r_chk = np_decode(z, s, sur.aek.cpu().numpy())
    assert np.max(np.abs(w_chk - Wn) / (np.abs(w_chk) + 1e-7 * s[:, None])) <= 1e-5
    r = resid.cpu().numpy()
    assert np.max(np.abs(r - r_chk)) <= 1e-4
    fl = flag.cpu().numpy()
    far = np.abs(r_chk - 1e-3) > 2e-4
    assert np.array_equal(fl[far], (r_chk > 1e-3)[far])
