"""CPU oracle of the learned split guide's network (test infrastructure only).

Network: 32 -> 64 -> 64 -> 64 -> 16, ReLU between layers.  Weights are BF16,
activations enter each product as BF16, products accumulate in FP32 (here in
float64, which bounds the difference by FP32 rounding).  Output columns
0..d-1 are split-axis logits, column 15 is the finish logit.

Loss per sample: softmax cross-entropy over the d axis logits against the
axis label, plus binary cross-entropy of the finish logit against the finish
label; the batch loss is the mean.  Backward products (and the bias
gradients) take BF16 copies of the upstream gradients, as the MFMA kernels
do.

The functions mirror csrc/learn.hip step by step:
  forward(Wb, X)            -> (H0, H1, H2, H3, Y)
  backward(Wb, cache, ...)  -> gradients of the four weight matrices and biases
and the training loop's sample order and axis relabelling:
  make_shuffle / shuffle_index -> the rows of an epoch (common.hpp Shuffle)
  relabel                      -> a row's features and label under a random
                                  permutation of its axes
"""
import numpy as np

IN, HID, OUT = 32, 64, 16
FINISH = 15
SHAPES = [(IN, HID), (HID, HID), (HID, HID), (HID, OUT)]


def bf16(x):
    """Round-to-nearest-even to bfloat16, returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isnan(a), np.nan, out)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    Ws, bs = [], []
    for (k, n) in SHAPES:
        Ws.append(rng.standard_normal((k, n)) * np.sqrt(2.0 / k))
        bs.append(np.zeros(n))
    return Ws, bs


def forward(Ws, bs, X, round_bf16=True):
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    H = [r(X)]
    for l in range(4):
        Z = H[-1] @ r(Ws[l]) + bs[l]
        if l < 3:
            H.append(r(np.maximum(Z, 0.0)))
        else:
            Y = Z
    return H, Y


def loss_and_dY(Y, axis, finish, d):
    n = Y.shape[0]
    logits = Y[:, :d]
    m = logits.max(axis=1, keepdims=True)
    e = np.exp(logits - m)
    p = e / e.sum(axis=1, keepdims=True)
    ce = -np.log(p[np.arange(n), axis])
    z = Y[:, FINISH]
    q = 1.0 / (1.0 + np.exp(-z))
    bce = np.maximum(z, 0) - z * finish + np.log1p(np.exp(-np.abs(z)))
    dY = np.zeros_like(Y)
    dY[:, :d] = p
    dY[np.arange(n), axis] -= 1.0
    dY[:, FINISH] = q - finish
    dY /= n
    return float((ce + bce).mean()), dY


def backward(Ws, bs, H, dY, round_bf16=True):
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    gW = [None] * 4
    gb = [None] * 4
    dZ = dY
    for l in (3, 2, 1, 0):
        dZr = r(dZ)
        gW[l] = H[l].T @ dZr
        gb[l] = dZr.sum(axis=0)
        if l > 0:
            dH = dZr @ r(Ws[l]).T
            dZ = dH * (H[l] > 0)
    return gW, gb


def grads(Ws, bs, X, axis, finish, d, round_bf16=True):
    H, Y = forward(Ws, bs, X, round_bf16)
    L, dY = loss_and_dY(Y, axis, finish, d)
    gW, gb = backward(Ws, bs, H, dY, round_bf16)
    return L, gW, gb, Y


def loss_only(Ws, bs, X, axis, finish, d):
    H, Y = forward(Ws, bs, X, round_bf16=False)
    return loss_and_dY(Y, axis, finish, d)[0]


# ------------------------------------------------- training order, relabelling

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def mix64(z):
    """SplitMix64 step: the output for state z (z advanced by the golden gamma
    first), as learn.hip's mix64."""
    z = (z + GOLDEN) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def make_shuffle(n, seed):
    """The permutation of [0, n) drawn from seed: four rounds of (add key,
    multiply by an odd constant, xor-shift) on [0, 2^bits), cycle-walked
    below n."""
    bits = 1
    while bits < 63 and (1 << bits) < n:
        bits += 1
    keys = [mix64((seed + i * GOLDEN) & M64) for i in range(4)]
    return n, (1 << bits) - 1, keys, (bits + 1) // 2


def shuffle_index(P, x):
    n, mask, keys, shift = P
    x = np.array(x, dtype=np.uint64, ndmin=1)
    mult = np.uint64(0xD1342543DE82EF95)
    todo = np.ones(x.shape, dtype=bool)
    while todo.any():
        y = x[todo]
        for k in keys:
            y = (y + np.uint64(k)) & np.uint64(mask)
            y = (y * mult) & np.uint64(mask)
            y ^= y >> np.uint64(shift)
        x[todo] = y
        todo[todo] = y >= np.uint64(n)
    return x.astype(np.int64)


def relabel(x, label, perm_seed, sample, dim):
    """Features and label of one row after the random relabelling of its
    axes: new axis j takes position, width and fourth-difference features
    (blocks 0-7, 8-15, 16-23) of logged axis pm[j]; pm is a Fisher-Yates
    shuffle of the first dim axes from a hash of (perm_seed, sample)."""
    pm = list(range(8))
    if perm_seed != 0:
        h = mix64(perm_seed ^ ((sample * GOLDEN) & M64))
        for j in range(dim - 1, 0, -1):
            k = ((h >> 32) * (j + 1)) >> 32
            h = mix64(h)
            pm[j], pm[k] = pm[k], pm[j]
    dst = [0] * 8
    for j in range(8):
        dst[pm[j]] = j
    y = np.array(x, copy=True)
    for q in range(24):
        if (q & 7) < dim:
            y[(q & ~7) | dst[q & 7]] = x[q]
    return y, dst[label]
