"""Float64 NumPy reference of the frequency-map network (4-128-128-128-3,
GELU in its tanh form): forward pass and manual back-propagation of the mean
squared error.  Test infrastructure only.

With bf16=True it rounds to BF16 (round to nearest even) at the same points
as the GPU kernels (emri_kludge_suite_amd/csrc/aak_mlp.hip): the weight
matrices, the activations A1..A3 entering the matrix products, the output
gradient dY and the back-propagated dZ3, dZ2 entering theirs.  Everything
else stays float64, so the GPU result differs from it by FP32 rounding and
summation order only.
"""
import numpy as np

H = 128
C_GELU = np.sqrt(2.0 / np.pi)


def offsets():
    o = {}
    k = 0
    for name, size in (("W1", 4 * H), ("b1", H), ("W2", H * H), ("b2", H), ("W3", H * H), ("b3", H),
                       ("W4", H * 3), ("b4", 3)):
        o[name] = (k, size)
        k += size
    return o, k


OFF, NP = offsets()


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    shapes = {"W1": (4, H), "b1": (H,), "W2": (H, H), "b2": (H,), "W3": (H, H), "b3": (H,), "W4": (H, 3),
              "b4": (3,)}
    return {k: p[o:o + n].reshape(shapes[k]) for k, (o, n) in OFF.items()}


def pack(d):
    out = np.zeros(NP)
    for k, (o, n) in OFF.items():
        out[o:o + n] = np.asarray(d[k]).ravel()
    return out


def bf16(x):
    """Round to BF16 (nearest even) and back."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C_GELU * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(C_GELU * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 3 * 0.044715 * z * z)


def forward(p, X, bf=False):
    """Outputs Y and the cache of the backward pass."""
    q = unpack(p)
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    W = {k: r(q[k]) for k in ("W1", "W2", "W3", "W4")}
    X = np.asarray(X, dtype=np.float64)
    Z1 = X @ W["W1"] + q["b1"]
    A1 = r(gelu(Z1))
    Z2 = A1 @ W["W2"] + q["b2"]
    A2 = r(gelu(Z2))
    Z3 = A2 @ W["W3"] + q["b3"]
    A3 = r(gelu(Z3))
    Y = A3 @ W["W4"] + q["b4"]
    return Y, (X, W, Z1, A1, Z2, A2, Z3, A3)


def loss_and_grad(p, X, T, bf=False):
    """Mean squared error over all n x 3 outputs and its gradient."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    Y, (X, W, Z1, A1, Z2, A2, Z3, A3) = forward(p, X, bf)
    n = len(X)
    D = Y - T
    loss = np.mean(D * D)
    dY = 2.0 * D / (3.0 * n)
    g = {"W4": A3.T @ dY, "b4": dY.sum(0)}
    dZ3 = (r(dY) @ W["W4"].T) * gelu_grad(Z3)
    g["W3"], g["b3"] = A2.T @ r(dZ3), dZ3.sum(0)
    dZ2 = (r(dZ3) @ W["W3"].T) * gelu_grad(Z2)
    g["W2"], g["b2"] = A1.T @ r(dZ2), dZ2.sum(0)
    dZ1 = (r(dZ2) @ W["W2"].T) * gelu_grad(Z1)
    g["W1"], g["b1"] = X.T @ dZ1, dZ1.sum(0)
    return loss, pack(g)
