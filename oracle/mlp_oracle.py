"""CPU oracle of the rule-selection network (test infrastructure only: imported by tests/ and
nothing else; the product trains and runs the network with its own HIP kernels).

The network has no counterpart in the reference, whose regular quadrature uses one fixed rule per
run (src/lib_Stokes.f90:2341-2446, src/lib_sdBem.f90:859-905). What is restated here is the
network's own definition, in float64:

    6 inputs -> 64 -> 64 -> 64 -> 8 logits, ReLU between layers, softmax cross-entropy,
    loss = mean over samples of -log softmax(z)[label].

Parameter layout (one flat float vector of 9288 entries, weights input-major so that
W[i, o] = p[off + i * n_out + o]):
    W1 6x64, b1 64, W2 64x64, b2 64, W3 64x64, b3 64, W4 64x8, b4 8.

`loss_grad` is checked against central finite differences of `loss` in the tests, which pins the
backward pass by its definition rather than by a second copy of the same formulas.
"""
import numpy as np

NIN, NH, NOUT = 6, 64, 8
SHAPES = [(NIN, NH), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NH, NOUT), (NOUT,)]
NPARAM = sum(int(np.prod(s)) for s in SHAPES)


def unpack(p):
    """Split the flat parameter vector into [W1, b1, W2, b2, W3, b3, W4, b4] (float64 views)."""
    p = np.asarray(p, dtype=np.float64)
    assert p.size == NPARAM
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(p[o:o + n].reshape(s))
        o += n
    return out


def pack(parts):
    return np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x in parts])


def init_params(seed=0):
    """He-style random parameters for the tests."""
    rng = np.random.default_rng(seed)
    parts = []
    for s in SHAPES:
        if len(s) == 2:
            parts.append(rng.standard_normal(s) * np.sqrt(2.0 / s[0]))
        else:
            parts.append(np.zeros(s))
    return pack(parts).astype(np.float32)


def forward(p, X):
    """Logits (M x 8) and the cached activations."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    Z1 = X @ W1 + b1
    H1 = np.maximum(Z1, 0.0)
    Z2 = H1 @ W2 + b2
    H2 = np.maximum(Z2, 0.0)
    Z3 = H2 @ W3 + b3
    H3 = np.maximum(Z3, 0.0)
    Z4 = H3 @ W4 + b4
    return Z4, (X, Z1, H1, Z2, H2, Z3, H3)


def softmax(Z):
    Z = Z - Z.max(axis=1, keepdims=True)
    E = np.exp(Z)
    return E / E.sum(axis=1, keepdims=True)


def loss(p, X, y):
    Z, _ = forward(p, X)
    Zs = Z - Z.max(axis=1, keepdims=True)
    lse = np.log(np.exp(Zs).sum(axis=1))
    return float(np.mean(lse - Zs[np.arange(len(y)), y]))


def loss_grad(p, X, y):
    """Mean cross-entropy and its gradient (flat, float64) by reverse-mode differentiation."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    Z4, (X, Z1, H1, Z2, H2, Z3, H3) = forward(p, X)
    M = X.shape[0]
    P = softmax(Z4)
    L = float(np.mean(-np.log(P[np.arange(M), y])))
    dZ4 = P.copy()
    dZ4[np.arange(M), y] -= 1.0
    dZ4 /= M
    gW4 = H3.T @ dZ4
    gb4 = dZ4.sum(0)
    dZ3 = (dZ4 @ W4.T) * (Z3 > 0)
    gW3 = H2.T @ dZ3
    gb3 = dZ3.sum(0)
    dZ2 = (dZ3 @ W3.T) * (Z2 > 0)
    gW2 = H1.T @ dZ2
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2.T) * (Z1 > 0)
    gW1 = X.T @ dZ1
    gb1 = dZ1.sum(0)
    return L, pack([gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4])


def bf16_round(p):
    """Round float32 values to bfloat16 (nearest, ties to even) and back to float32."""
    u = np.asarray(p, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def forward_bf16(p, X):
    """Logits as the in-assembly classifier computes them: inputs, weights and the activations
    between layers rounded to bfloat16 (nearest even), biases and accumulation in float (here
    float64, which differs from the device's float32 sums only in rounding)."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    r = lambda a: bf16_round(np.asarray(a, dtype=np.float32)).astype(np.float64)  # noqa: E731
    H = r(X)
    for W, b in ((W1, b1), (W2, b2), (W3, b3)):
        H = r(np.maximum(H @ r(W) + b, 0.0))
    return H @ r(W4) + b4
