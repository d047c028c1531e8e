"""ORACLE -- test infrastructure only.

CPU restatement (numpy) of the temperature surrogate: a 5-128-128-128-128-1
MLP with SiLU hidden layers.  The reference has no such network; this file is
the checker for the device forward and backward kernels.  Only tests/ load it.

Two forms:
  * `forward` / `gradients` round every matrix-product operand to BF16 (round
    to nearest even) where the device stages it to the matrix cores and
    accumulate in float64, so they differ from the device only by FP32
    accumulation order;
  * `forward64` / `gradients64` are exact float64 maths, pinned by central
    finite differences in the tests.
Parameter vector: W1[5][128] b1[128] W2..W4[128][128] b2..b4[128] W5[128] b5,
weights stored [fan_in][fan_out].
"""
import numpy as np

FAN_IN = (5, 128, 128, 128, 128)
FAN_OUT = (128, 128, 128, 128, 1)


def layout():
    off, o = [], 0
    for fi, fo in zip(FAN_IN, FAN_OUT):
        off.append((o, o + fi * fo))
        o += fi * fo + fo
    return off, o


def unpack(p):
    off, _ = layout()
    W = [np.asarray(p[w:b], np.float64).reshape(fi, fo)
         for (w, b), fi, fo in zip(off, FAN_IN, FAN_OUT)]
    B = [np.asarray(p[b:b + fo], np.float64) for (w, b), fo in zip(off, FAN_OUT)]
    return W, B


def pack(W, B):
    return np.concatenate([np.concatenate([w.ravel(), b.ravel()]) for w, b in zip(W, B)])


def bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def f32(x):
    """Round to FP32, returned as float64."""
    return np.asarray(x, np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X):
    """Device-rounded forward: returns (T, cache)."""
    W, B = unpack(p)
    a = np.asarray(X, np.float64)
    Z, A = [], [a]
    for l in range(4):
        z = (bf16(a) @ bf16(W[l]) + B[l]).astype(np.float32).astype(np.float64)
        a = bf16(silu(z))
        Z.append(z)
        A.append(a)
    T = (bf16(a) @ bf16(W[4]) + B[4]).astype(np.float32).astype(np.float64)[:, 0]
    return T, (Z, A)


def forward_scaled(p, X):
    """Forward at the fused inference kernel's rounding points: the hidden
    layers carry y = -log2(e) z (FP32), sigmoid(z) = 1 / (1 + 2^y), and the
    BF16-rounded activation is y * sigmoid(z) = -log2(e) silu(z); the biases
    are scaled by -log2(e) in FP32, and the output layer accumulates against
    the BF16 W5 in FP32, then takes acc * (-1/log2(e)) + b5 in one FMA.
    The scalings cancel: without rounding this is `forward64`."""
    W, B = unpack(p)
    c, ci = f32(-1.4426950408889634), f32(-0.6931471805599453)
    a = bf16(np.asarray(X, np.float64))
    for l in range(4):
        b = f32(c * f32(B[l]))
        acc = f32(a @ bf16(W[l]))
        y = f32(f32(acc * c) + b) if l == 0 else f32(acc + b)
        a = bf16(f32(y * f32(1.0 / (1.0 + np.exp2(y)))))
    acc = f32(a @ bf16(W[4][:, 0]))
    return f32(acc * ci + f32(B[4][0]))


def gradients(p, X, Y):
    """Device-rounded gradients of mean((T - Y)^2): returns (grads, loss)."""
    W, B = unpack(p)
    T, (Z, A) = forward(p, X)
    n = len(T)
    r = T - np.asarray(Y, np.float64)
    dT = (2.0 * r / n).astype(np.float32).astype(np.float64)[:, None]
    gW, gB = [None] * 5, [None] * 5
    gW[4] = bf16(A[4]).T @ bf16(dT)
    gB[4] = dT.sum(0)
    dz = (bf16(dT) @ bf16(W[4]).T) * dsilu(Z[3])
    for l in range(3, -1, -1):
        dz = dz.astype(np.float32).astype(np.float64)
        gW[l] = bf16(A[l]).T @ bf16(dz)
        gB[l] = dz.sum(0)
        if l > 0:
            dz = (bf16(dz) @ bf16(W[l]).T) * dsilu(Z[l - 1])
    return pack(gW, gB), float((r * r).mean())


def forward64(p, X):
    W, B = unpack(p)
    a = np.asarray(X, np.float64)
    for l in range(4):
        a = silu(a @ W[l] + B[l])
    return (a @ W[4] + B[4])[:, 0]


def gradients64(p, X, Y):
    W, B = unpack(p)
    a = np.asarray(X, np.float64)
    Z, A = [], [a]
    for l in range(4):
        z = a @ W[l] + B[l]
        a = silu(z)
        Z.append(z)
        A.append(a)
    T = (a @ W[4] + B[4])[:, 0]
    n = len(T)
    r = T - np.asarray(Y, np.float64)
    dT = (2.0 * r / n)[:, None]
    gW, gB = [None] * 5, [None] * 5
    gW[4] = A[4].T @ dT
    gB[4] = dT.sum(0)
    dz = (dT @ W[4].T) * dsilu(Z[3])
    for l in range(3, -1, -1):
        gW[l] = A[l].T @ dz
        gB[l] = dz.sum(0)
        if l > 0:
            dz = (dz @ W[l].T) * dsilu(Z[l - 1])
    return pack(gW, gB), float((r * r).mean())
