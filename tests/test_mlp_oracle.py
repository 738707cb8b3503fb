"""The surrogate's CPU oracle (oracle/mlp_np.py) pinned by finite differences and BF16 facts."""
import numpy as np

from oracle.mlp_np import bf16, loss_and_grads


def test_bf16_rounding_known_values():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties to even -> 1; 1 + 3*2^-9 rounds up
    assert bf16(1.0 + 2.0 ** -8) == 1.0
    assert bf16(1.0 + 3 * 2.0 ** -9) == 1.0 + 2.0 ** -7
    assert bf16(3.140625) == 3.140625          # exactly representable
    assert bf16(-2.0 ** -130) == -(2.0 ** -130)  # subnormal bf16 keeps 7 bits


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    dims = [10, 16, 16, 7]
    Ws = [rng.normal(size=(dims[i + 1], dims[i])) * 0.5 for i in range(3)]
    bs = [rng.normal(size=dims[i + 1]) * 0.1 for i in range(3)]
    X = rng.normal(size=(13, 10))
    Z = rng.normal(size=(13, 7))
    loss, gW, gb = loss_and_grads(X, Z, Ws, bs, quantize=False)
    h = 1e-6
    for l in range(3):
        for (i, j) in [(0, 0), (3, 5), (dims[l + 1] - 1, dims[l] - 1)]:
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[l][i, j] += h
            Wm[l][i, j] -= h
            fd = (loss_and_grads(X, Z, Wp, bs, False)[0] - loss_and_grads(X, Z, Wm, bs, False)[0]) / (2 * h)
            assert abs(fd - gW[l][i, j]) <= 1e-6 * max(1.0, abs(fd)), (l, i, j, fd, gW[l][i, j])
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[l][2] += h
        bm[l][2] -= h
        fd = (loss_and_grads(X, Z, Ws, bp, False)[0] - loss_and_grads(X, Z, Ws, bm, False)[0]) / (2 * h)
        assert abs(fd - gb[l][2]) <= 1e-6 * max(1.0, abs(fd))
