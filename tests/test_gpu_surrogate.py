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
    """tol = 0 sends every point to the FP64 kernel (bit-identical output); weights forced
    to nonsense from the host are caught by the conservation residual."""
    from chemgen_amd.surrogate import Surrogate, random_states
    from chemgen_amd.runtime import DeviceMechanism
    m = load_bundled("h2_burke")
    dm = DeviceMechanism(m, build=False)
    sur = Surrogate(m, hidden=32, n_hidden=2, seed=1)
    P, T, Y = random_states(m.n_species, 5000, seed=9)
    ref = dm.getrates(P, T, Y)
    sur.fit_scales(ref)
    W = torch.empty_like(Y)
    info = sur.getrates_(P, T, Y, W, dm, tol=0.0, audit_every=0)
    torch.cuda.synchronize()
    assert info["fallback_points"] == 5000
    assert torch.equal(W, ref)
    for L in sur.layers:       # garbage network: outputs violate element conservation
        L.w16.copy_((L.w * 40.0).to(torch.bfloat16))
    info = sur.getrates_(P, T, Y, W, dm, tol=1e-3, audit_every=0)
    torch.cuda.synchronize()
    assert info["fallback_points"] > 4900
    err = ((W - ref).abs() / (ref.abs() + sur.s[:, None])).max(dim=0).values
    flagged_ok = float((err <= 1e-12).double().mean())
    assert flagged_ok >= info["fallback_points"] / 5000 - 1e-9


def test_training_gains_over_untrained(torch):
    """A short training run reduces the surrogate's error on held-out states well below the
    untrained network's (the method without learning)."""
    from chemgen_amd.surrogate import Surrogate, random_states
    from chemgen_amd.runtime import DeviceMechanism
    m = load_bundled("h2_burke")
    dm = DeviceMechanism(m, build=False)
    sur = Surrogate(m, hidden=64, n_hidden=3, seed=2)
    P, T, Y = random_states(m.n_species, 1 << 16, seed=11)
    W = dm.getrates(P, T, Y)
    sur.fit_scales(W)
    Ph, Th, Yh = random_states(m.n_species, 8192, seed=12)
    Wh = dm.getrates(Ph, Th, Yh)
    Zh_ref = sur.targets(Wh)

    def zerr():
        return float((sur.forward(sur.features(Ph, Th, Yh)) - Zh_ref).pow(2).mean().item())

    before = zerr()
    print("[train] start", flush=True)
    sur.train(P, T, Y, W, epochs=4, batch=2048, lr=3e-3)
    after = zerr()
    print("z mse before", before, "after", after)
    assert after < 0.25 * before
