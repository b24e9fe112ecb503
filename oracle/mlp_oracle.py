"""CPU restatement of the warm-start network (cpfft_amd/csrc/mlp.hip):
forward pass, mean-squared-error loss and its gradient.

TEST INFRASTRUCTURE ONLY.  Only tests/ may import this module; the product
never loads it.  The network has no counterpart in the reference: it only
sets the starting point of the local Newton solve (mm10_a.f:2945-3023).

Network: nin -> 128 (tanh) -> 128 (tanh) -> nout (42 -> 12 for fcc, 66 -> 24
for bcc24), parameters in one flat FP32 vector [W1 (128 x nin) | b1 |
W2 (128x128) | b2 | W3 (nout x 128) | b3] with rows of W the output units.
Loss: mean over rows x nout of (Y - T)^2.

bf16=True rounds every GEMM operand to BF16 (round to nearest even) at the
points where the GPU kernels round them (inputs, weights, hidden
activations, back-propagated errors) and accumulates in FP64, so it agrees
with the GPU's FP32 accumulation to ~1e-6 relative apart from rare rounding
ties.  bf16=False is the exact FP64 model; its gradient is pinned by finite
differences (tests/test_oracle_mlp.py).
"""
import numpy as np

IN, HID, OUT = 42, 128, 12
O_W1 = 0
O_B1 = O_W1 + HID * IN
O_W2 = O_B1 + HID
O_B2 = O_W2 + HID * HID
O_W3 = O_B2 + HID
O_B3 = O_W3 + OUT * HID
NPARAM = O_B3 + OUT


def to_bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def offsets(nin=IN, nout=OUT):
    """(W1, b1, W2, b2, W3, b3, end) offsets of the flat parameter vector"""
    w1 = 0
    b1 = w1 + HID * nin
    w2 = b1 + HID
    b2 = w2 + HID * HID
    w3 = b2 + HID
    b3 = w3 + nout * HID
    return w1, b1, w2, b2, w3, b3, b3 + nout


def unpack(p, nin=IN, nout=OUT):
    p = np.asarray(p, dtype=np.float64)
    w1, b1, w2, b2, w3, b3, end = offsets(nin, nout)
    return (p[w1:b1].reshape(HID, nin), p[b1:w2], p[w2:b2].reshape(HID, HID),
            p[b2:w3], p[w3:b3].reshape(nout, HID), p[b3:end])


def _r(x, bf16):
    return to_bf16(x).astype(np.float64) if bf16 else np.asarray(x, dtype=np.float64)


def forward(p, X, bf16=True, nout=OUT):
    """Network output for inputs X[n][nin]; returns (Y, cache)."""
    X = np.asarray(X)
    W1, b1, W2, b2, W3, b3 = unpack(p, X.shape[1], nout)
    f32 = np.float32 if bf16 else np.float64
    Xb = _r(X, bf16)
    H1 = np.tanh((Xb @ _r(W1, bf16).T + b1).astype(f32)).astype(np.float64)
    H1b = _r(H1, bf16)
    H2 = np.tanh((H1b @ _r(W2, bf16).T + b2).astype(f32)).astype(np.float64)
    H2b = _r(H2, bf16)
    Y = (H2b @ _r(W3, bf16).T + b3).astype(f32).astype(np.float64)
    return Y, (Xb, H1, H1b, H2, H2b)


def loss_and_grad(p, X, T, bf16=True):
    """(loss, gradient[nparam]) of mean((Y - T)^2); the sizes follow X[n][nin]
    and T[n][nout]."""
    X, T = np.asarray(X), np.asarray(T, dtype=np.float64)
    nin, nout = X.shape[1], T.shape[1]
    w1, b1o, w2, b2o, w3, b3o, end = offsets(nin, nout)
    W1, b1, W2, b2, W3, b3 = unpack(p, nin, nout)
    f32 = np.float32 if bf16 else np.float64
    Y, (Xb, H1, H1b, H2, H2b) = forward(p, X, bf16, nout)
    n = Y.size
    d = Y - T
    loss = float((d * d).sum() / n)
    dY = (2.0 / n * d).astype(f32).astype(np.float64)
    dYb = _r(dY, bf16)
    g = np.zeros(end)
    g[w3:b3o] = (dYb.T @ H2b).ravel()
    g[b3o:end] = dY.sum(0)
    dZ2 = ((dYb @ _r(W3, bf16)) * (1.0 - H2 * H2)).astype(f32).astype(np.float64)
    dZ2b = _r(dZ2, bf16)
    g[w2:b2o] = (dZ2b.T @ H1b).ravel()
    g[b2o:w3] = dZ2.sum(0)
    dZ1 = ((dZ2b @ _r(W2, bf16)) * (1.0 - H1 * H1)).astype(f32).astype(np.float64)
    dZ1b = _r(dZ1, bf16)
    g[w1:b1o] = (dZ1b.T @ Xb).ravel()
    g[b1o:w2] = dZ1.sum(0)
    return loss, g
