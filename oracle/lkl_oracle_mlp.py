"""f64 numpy restatement of the surrogate network and its training step.

TEST INFRASTRUCTURE ONLY: the checker of the HIP surrogate kernels.  Imported
by tests/ and bench.py's checks; never by the product.

The network has no counterpart in the reference (it is the learned screening
step added on top of the exact scorer):
  MLP 6144 -> 512 -> 512 -> 512 -> 1, ReLU, weights stored input-major,
  flat fp32 layout W1, b1, W2, b2, W3, b3, W4, b4 (3,672,065 values).
With round=True the forward mirrors the kernels' operand precision: weights
and input features are bf16 (round to nearest even), each hidden activation is
relu(acc + bias) rounded to f32 then to bf16; sums are exact f64 here where the
kernels accumulate in fp32.  With round=False everything stays f64, which is
the function the central finite differences in the tests differentiate.

Pinned by tests: central finite differences of the f64 loss (every layer),
torch-CPU float64 autograd of the same network, and the closed form of the
first Adam step (m_hat = g, v_hat = g^2).
"""
import numpy as np

NF, D = 6144, 512
SIZES = [("W1", NF * D), ("b1", D), ("W2", D * D), ("b2", D), ("W3", D * D), ("b3", D),
         ("W4", D), ("b4", 1)]
NPARAM = sum(n for _, n in SIZES)


def offsets():
    out, o = {}, 0
    for name, n in SIZES:
        out[name] = (o, n)
        o += n
    return out


def split(p):
    """Flat parameter vector -> dict of f64 arrays in their matrix shapes."""
    p = np.asarray(p)
    shapes = {"W1": (NF, D), "W2": (D, D), "W3": (D, D), "W4": (D,), "b1": (D,), "b2": (D,),
              "b3": (D,), "b4": (1,)}
    return {k: p[o:o + n].astype(np.float64).reshape(shapes[k]) for k, (o, n) in offsets().items()}


def bf16_bits_f32(a):
    """f32 -> bf16 bits, round to nearest even (finite values)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_bits_f64(a):
    """Non-negative f64 -> bf16 bits with ONE rounding (nearest even); 0 -> 0."""
    a = np.ascontiguousarray(a, np.float64)
    u = a.view(np.uint64)
    e = (u >> np.uint64(52)).astype(np.int64) - 1023 + 127
    keep = ((u >> np.uint64(45)) & np.uint64(0x7F)).astype(np.int64)
    rest = u & np.uint64((1 << 45) - 1)
    half = np.uint64(1 << 44)
    up = (rest > half) | ((rest == half) & ((keep & 1) == 1))
    keep = keep + up
    e = e + (keep == 128)
    keep = np.where(keep == 128, 0, keep)
    bits = ((e << 7) | keep).astype(np.uint16)
    return np.where(a == 0.0, np.uint16(0), bits)


def bf16_to_f64(bits):
    b = np.ascontiguousarray(bits, np.uint16).astype(np.uint32) << 16
    return b.view(np.float32).astype(np.float64)


def round_bf16(a):
    """Values of a (any float) after f32 then bf16 rounding."""
    return bf16_to_f64(bf16_bits_f32(np.asarray(a, np.float32)))


def features_bits(counts):
    """log1p(count) in f64, rounded once to bf16 (the feature values)."""
    return bf16_bits_f64(np.log1p(np.asarray(counts, np.float64)))


def forward(p, x, round=True):
    """x: bf16 bits (B x 6144) if round else f64 values.  Returns y, cache."""
    P = split(p)
    if round:
        X = bf16_to_f64(x)
        W = {k: round_bf16(P[k]) for k in ("W1", "W2", "W3", "W4")}
        act = lambda z: round_bf16(np.maximum(z, 0.0))
    else:
        X = np.asarray(x, np.float64)
        W = {k: P[k] for k in ("W1", "W2", "W3", "W4")}
        act = lambda z: np.maximum(z, 0.0)
    h1 = act(X @ W["W1"] + P["b1"])
    h2 = act(h1 @ W["W2"] + P["b2"])
    h3 = act(h2 @ W["W3"] + P["b3"])
    y = h3 @ W["W4"] + P["b4"][0]
    return y, {"X": X, "W": W, "h1": h1, "h2": h2, "h3": h3}


def loss_grad(p, x, t, round=True):
    """Loss 0.5 * mean((y - t)^2) and its gradient (flat f64, parameter layout)."""
    y, c = forward(p, x, round)
    t = np.asarray(t, np.float64)
    B = y.shape[0]
    dy = (y - t) / B
    loss = 0.5 * np.mean((y - t) ** 2)
    W = c["W"]
    g = {}
    g["W4"] = c["h3"].T @ dy
    g["b4"] = np.array([dy.sum()])
    dz3 = np.outer(dy, W["W4"]) * (c["h3"] > 0)
    g["W3"] = c["h2"].T @ dz3
    g["b3"] = dz3.sum(0)
    dz2 = (dz3 @ W["W3"].T) * (c["h2"] > 0)
    g["W2"] = c["h1"].T @ dz2
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ W["W2"].T) * (c["h1"] > 0)
    g["W1"] = c["X"].T @ dz1
    g["b1"] = dz1.sum(0)
    flat = np.zeros(NPARAM, np.float64)
    for k, (o, n) in offsets().items():
        flat[o:o + n] = g[k].ravel()
    return loss, flat, y


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """Kingma & Ba bias-corrected Adam (step counts from 1); f64."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


# ── spread of valid fp32 evaluations (tolerances of the device checks) ────────
# The kernels take the same bf16 operands as forward(round=True) but
# accumulate in fp32 (MFMA tiles, split-K partials, fp32 epilogues), in an
# order of their own; a hidden activation then rounds to a neighbouring bf16
# value now and then, and the gradients inherit those differences through the
# bf16 dZ operands.  A worst-case bound of that chain grows by the layers'
# absolute row sums (tens per layer) and is vacuous, so the checks compare the
# device with the spread of other valid fp32 evaluations of the same network:
# the same operands and bf16 roundings (activations, dZ operands) with fp32
# matrix products whose summation order is permuted (a random permutation of
# the reduced dimension per evaluation).


def _mm32(a, b, perm):
    """fp32 product with the reduced dimension permuted."""
    return np.asarray(a, np.float32)[:, perm] @ np.asarray(b, np.float32)[perm, :]


def forward_f32(p, x, seed):
    """forward(round=True) with fp32 products in a permuted order: y, cache."""
    rng = np.random.default_rng(seed)
    P = split(p)
    X = bf16_to_f64(x).astype(np.float32)
    W = {k: round_bf16(P[k]).astype(np.float32) for k in ("W1", "W2", "W3", "W4")}
    act = lambda z: round_bf16(np.maximum(z, 0.0)).astype(np.float32)
    h1 = act(_mm32(X, W["W1"], rng.permutation(NF)) + P["b1"].astype(np.float32))
    h2 = act(_mm32(h1, W["W2"], rng.permutation(D)) + P["b2"].astype(np.float32))
    h3 = act(_mm32(h2, W["W3"], rng.permutation(D)) + P["b3"].astype(np.float32))
    pk = rng.permutation(D)
    y = (h3[:, pk] @ W["W4"][pk] + np.float32(P["b4"][0])).astype(np.float64)
    return y, {"X": X, "W": W, "h1": h1, "h2": h2, "h3": h3}


def loss_grad_f32(p, x, t, seed):
    """loss_grad(round=True) evaluated like the kernels, fp32 products in a
    permuted order, dZ rounded to bf16 where the kernels use it as a GEMM
    operand.  Returns (loss, flat gradient f64, y)."""
    y, c = forward_f32(p, x, seed)
    rng = np.random.default_rng(seed + 1)
    t = np.asarray(t, np.float64)
    B = y.shape[0]
    dy = ((y.astype(np.float32) - t.astype(np.float32)) / np.float32(B)).astype(np.float32)
    loss = float(0.5 * np.mean((y - t) ** 2))
    W = c["W"]
    bf = lambda a: round_bf16(a).astype(np.float32)
    g = {}
    g["W4"] = (c["h3"].T @ dy).astype(np.float64)
    g["b4"] = np.array([float(dy.sum(dtype=np.float32))])
    dz = np.outer(dy, W["W4"]).astype(np.float32) * (c["h3"] > 0)
    hs = {3: c["h2"], 2: c["h1"], 1: c["X"]}
    for l in (3, 2, 1):
        dzb = bf(dz)
        g["W%d" % l] = _mm32(hs[l].T, dzb, rng.permutation(B)).astype(np.float64)
        g["b%d" % l] = dz.sum(0, dtype=np.float32).astype(np.float64)
        if l > 1:
            dz = _mm32(dzb, W["W%d" % l].T, rng.permutation(D)) * (hs[l] > 0)
    flat = np.zeros(NPARAM, np.float64)
    for k, (o, n) in offsets().items():
        flat[o:o + n] = np.asarray(g[k]).ravel()
    return loss, flat, y
