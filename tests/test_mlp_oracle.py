"""Pins the CPU oracle of the learned guide's network: BF16 rounding on known
values, and its backward pass against central finite differences of the
loss (float64, no rounding)."""
import numpy as np

import mlp_oracle as M


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1 + 2**-8, 1 + 3 * 2**-9, -2.5, 3.0e-39, 65504.0, 1 + 2**-7], dtype=np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0  # tie -> even
    assert r[2] == 1 + 2**-7  # above the tie -> up
    assert r[3] == -2.5
    assert r[6] == 1 + 2**-7
    assert abs(r[5] - 65536.0) == 0.0  # 65504 rounds to 2^16 in bf16


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    Ws, bs = M.init_params(3)
    bs = [rng.standard_normal(b.shape) * 0.1 for b in bs]
    d = 7
    n = 20
    X = rng.standard_normal((n, M.IN))
    axis = rng.integers(0, d, n)
    fin = (rng.random(n) < 0.4).astype(np.float64)
    L, gW, gb, Y = M.grads(Ws, bs, X, axis, fin, d, round_bf16=False)
    h = 1e-6
    for l in range(4):
        for _ in range(12):
            i = rng.integers(0, Ws[l].shape[0]); j = rng.integers(0, Ws[l].shape[1])
            Wp = [w.copy() for w in Ws]; Wm = [w.copy() for w in Ws]
            Wp[l][i, j] += h; Wm[l][i, j] -= h
            fd = (M.loss_only(Wp, bs, X, axis, fin, d) - M.loss_only(Wm, bs, X, axis, fin, d)) / (2 * h)
            assert abs(fd - gW[l][i, j]) <= 1e-6 * max(1.0, abs(fd)), (l, i, j, fd, gW[l][i, j])
        for _ in range(6):
            j = rng.integers(0, bs[l].shape[0])
            bp = [b.copy() for b in bs]; bm = [b.copy() for b in bs]
            bp[l][j] += h; bm[l][j] -= h
            fd = (M.loss_only(Ws, bp, X, axis, fin, d) - M.loss_only(Ws, bm, X, axis, fin, d)) / (2 * h)
            assert abs(fd - gb[l][j]) <= 1e-6 * max(1.0, abs(fd)), (l, j, fd, gb[l][j])


def test_loss_is_cross_entropy_plus_bce():
    Y = np.zeros((1, 16))
    Y[0, :3] = [1.0, 2.0, 0.5]
    Y[0, 15] = 0.3
    L, dY = M.loss_and_dY(Y, np.array([1]), np.array([1.0]), 3)
    ce = -np.log(np.exp(2.0) / np.exp([1.0, 2.0, 0.5]).sum())
    bce = -np.log(1 / (1 + np.exp(-0.3)))
    assert abs(L - (ce + bce)) < 1e-12
