"""CPU FP64 restatement of the warm-start network (test infrastructure only).

Per-node MLP 8 -> 128 -> 128 -> 3 with tanh hidden layers, flat weight layout
W1[8][H] b1[H] W2[H][H] b2[H] W3[H][3] b3[3], mean-squared-error loss over all
rows*3 outputs.  The reference has no learned component (its Newmark Newton starts
from the predictor, CBSolverNewmarkBeta.cpp:845-855); this module is the checker of
the library's forward/backward kernels, and is itself pinned by central finite
differences of `loss` (tests/test_oracle_pins.py).
"""
import numpy as np

IN, OUT = 8, 3


def unpack(w, H=128):
    w = np.asarray(w, np.float64)
    o = 0
    W1 = w[o:o + IN * H].reshape(IN, H); o += IN * H
    b1 = w[o:o + H]; o += H
    W2 = w[o:o + H * H].reshape(H, H); o += H * H
    b2 = w[o:o + H]; o += H
    W3 = w[o:o + H * OUT].reshape(H, OUT); o += H * OUT
    b3 = w[o:o + OUT]; o += OUT
    assert o == w.size
    return W1, b1, W2, b2, W3, b3


def n_weights(H=128):
    return IN * H + H + H * H + H + H * OUT + OUT


def forward(w, X, H=128):
    W1, b1, W2, b2, W3, b3 = unpack(w, H)
    A1 = np.tanh(X @ W1 + b1)
    A2 = np.tanh(A1 @ W2 + b2)
    return A2 @ W3 + b3, A1, A2


def loss(w, X, T, H=128):
    Y, _, _ = forward(w, X, H)
    return float(np.mean((Y - T) ** 2))


def loss_grad(w, X, T, H=128):
    """(loss, dloss/dw) by the chain rule, same flat layout as w."""
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(w, H)
    Y, A1, A2 = forward(w, X, H)
    dY = 2.0 * (Y - T) / Y.size
    dW3 = A2.T @ dY
    db3 = dY.sum(0)
    dZ2 = (dY @ W3.T) * (1.0 - A2 ** 2)
    dW2 = A1.T @ dZ2
    db2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2.T) * (1.0 - A1 ** 2)
    dW1 = X.T @ dZ1
    db1 = dZ1.sum(0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3])
    return float(np.mean((Y - T) ** 2)), g


def adam(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return w, m, v
