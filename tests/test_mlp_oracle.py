"""Pins the CPU oracle of the learned guide's network: BF16 rounding on known
values, and its backward pass against central finite differences of the
loss (float64, no rounding); the training order (a permutation of the
sample log per epoch) and the axis relabelling of training rows."""
import numpy as np
import pytest

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


def test_mix64_is_splitmix64():
    import workloads as W
    for s in (0, 1, 7, 0xA0761D6478BD642F):
        assert M.mix64(s) == W.SplitMix64(s).next_u64()


@pytest.mark.parametrize("n", [1, 2, 3, 64, 1000, 8192, 65537, 300_000])
def test_shuffle_is_a_permutation(n):
    P = M.make_shuffle(n, M.mix64(n))
    v = M.shuffle_index(P, np.arange(n))
    assert np.array_equal(np.sort(v), np.arange(n))


def test_shuffle_mixes_large_logs():
    # neighbours in the sample log land in the same 8192-row batch about as
    # often as under a uniform permutation (8192 / n)
    n = 600_000
    v = M.shuffle_index(M.make_shuffle(n, 11), np.arange(n))
    inv = np.empty(n, np.int64)
    inv[v] = np.arange(n)
    b = inv // 8192
    assert abs(np.mean(b[1:] == b[:-1]) - 8192 / n) <= 0.2 * 8192 / n
    assert abs(np.corrcoef(np.arange(n), v)[0, 1]) < 0.01


@pytest.mark.parametrize("dim", [2, 5, 8])
def test_relabel_moves_axis_blocks_with_the_label(dim):
    x = np.arange(32, dtype=np.float64)
    for smp in range(50):
        for lab in range(dim):
            y, nl = M.relabel(x, lab, 0xC0FFEE, smp, dim)
            # the label's features follow it, in all three blocks
            for blk in range(3):
                assert y[8 * blk + nl] == x[8 * blk + lab]
            # axes beyond dim and the scalar features stay put
            assert np.array_equal(y[24:], x[24:])
            for q in range(24):
                if (q & 7) >= dim:
                    assert y[q] == x[q]
            assert sorted(y[:8][:dim]) == list(x[:dim])
    y, nl = M.relabel(x, 3, 0, 5, 8)
    assert nl == 3 and np.array_equal(y, x)
