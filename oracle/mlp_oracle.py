"""FP64 numpy restatement of the drag-surrogate network (test infrastructure
only: the checker of the GPU training and inference kernels).

Network 5 -> 64 -> 64 -> 3, GELU in its tanh form
    gelu(x) = x/2 (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))
(Hendrycks & Gimpel, arXiv:1606.08415, eq. for the tanh approximation).
Parameters are one flat vector W1[5][64] b1[64] W2[64][64] b2[64] W3[64][3]
b3[3] (row = input unit), IMOGEN_MLP_NPARAMS = 4739 floats.
"""
import numpy as np

N_IN, N_HID, N_OUT = 5, 64, 3
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_IN * N_HID
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_HID * N_OUT
NPARAMS = OFF_B3 + N_OUT
A = np.sqrt(2.0 / np.pi)
B = 0.044715


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(N_IN, N_HID), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(N_HID, N_HID),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(N_HID, N_OUT), p[OFF_B3:NPARAMS])


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(A * (x + B * x ** 3)))


def gelu_grad(x):
    t = np.tanh(A * (x + B * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * A * (1.0 + 3.0 * B * x * x)


def forward(p, x):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    z1 = x @ W1 + b1
    h1 = gelu(z1)
    z2 = h1 @ W2 + b2
    h2 = gelu(z2)
    y = h2 @ W3 + b3
    return y, (z1, h1, z2, h2)


def backward(p, x, cache, dy):
    """Gradient of sum(dy * y) with respect to the flat parameter vector."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    z1, h1, z2, h2 = cache
    g = np.zeros(NPARAMS)
    g[OFF_W3:OFF_B3] = (h2.T @ dy).ravel()
    g[OFF_B3:NPARAMS] = dy.sum(0)
    dz2 = (dy @ W3.T) * gelu_grad(z2)
    g[OFF_W2:OFF_B2] = (h1.T @ dz2).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * gelu_grad(z1)
    g[OFF_W1:OFF_B1] = (x.T @ dz1).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    return g


def adam(p, g, m, v, lr, b1, b2, eps, step):
    """One Adam step (Kingma & Ba 2015, Algorithm 1); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mhat = m / (1 - b1 ** step)
    vhat = v / (1 - b2 ** step)
    return p - lr * mhat / (np.sqrt(vhat) + eps), m, v


def to_bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def forward_bf16(p, x):
    """The inference path's rounding: inputs, weights and hidden activations in
    bfloat16, sums and GELU in (here: float64 instead of float32)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    h1 = to_bf16(gelu(to_bf16(x) @ to_bf16(W1) + b1.astype(np.float32)))
    h2 = to_bf16(gelu(h1 @ to_bf16(W2) + b2.astype(np.float32)))
    return h2 @ to_bf16(W3) + b3.astype(np.float32)
