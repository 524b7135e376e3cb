"""FP64 numpy restatement of the learned-proposal MLP and of its one-ring
features (test infrastructure only).

The network is 96-128-128-3 with GELU (erf form) and mean-squared-error
loss; parameters are packed W1[128x96], b1, W2[128x128], b2, W3[3x128], b3.
With ``emulate=True`` the inputs, weights and hidden activations are rounded
to BF16 (round to nearest even) where the GPU kernels round them, and the
backward pass treats those roundings as the identity (the kernels'
convention); with ``emulate=False`` it is the plain FP64 function, whose
gradient the finite-difference test pins.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

N_IN, N_HID, N_OUT = 96, 128, 3
LAYOUT = [("W1", (N_HID, N_IN)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
          ("W3", (N_OUT, N_HID)), ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in LAYOUT)


def unpack(p):
    out, o = {}, 0
    for name, shape in LAYOUT:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in LAYOUT])


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), as float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(p, X, emulate=True):
    w = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    W1, W2, W3 = r(w["W1"]), r(w["W2"]), r(w["W3"])
    Xb = r(X)
    Z1 = Xb @ W1.T + w["b1"]
    H1 = r(gelu(Z1))
    Z2 = H1 @ W2.T + w["b2"]
    H2 = r(gelu(Z2))
    Y = H2 @ W3.T + w["b3"]
    return Y, dict(X=Xb, Z1=Z1, H1=H1, Z2=Z2, H2=H2, W1=W1, W2=W2, W3=W3)


def loss_and_grad(p, X, T, emulate=True):
    Y, c = forward(p, X, emulate)
    n = X.shape[0]
    D = Y - np.asarray(T, dtype=np.float64)
    loss = float((D * D).sum() / n)
    dY = 2.0 * D / n
    g = {}
    g["W3"] = dY.T @ c["H2"]
    g["b3"] = dY.sum(axis=0)
    dZ2 = (dY @ c["W3"]) * gelu_d(c["Z2"])
    g["W2"] = dZ2.T @ c["H1"]
    g["b2"] = dZ2.sum(axis=0)
    dZ1 = (dZ2 @ c["W2"]) * gelu_d(c["Z1"])
    g["W1"] = dZ1.T @ c["X"]
    g["b1"] = dZ1.sum(axis=0)
    return loss, pack(g)


def block_slices():
    out, o = {}, 0
    for name, shape in LAYOUT:
        k = int(np.prod(shape))
        out[name] = slice(o, o + k)
        o += k
    return out


# ----------------------------------------------------------- one-ring features
def canonical_neighbours(ring_row, ring_val, stride, v, k=32):
    """Distinct vertices of v's ring elements in order of first appearance."""
    ring_val = np.asarray(ring_val).reshape(-1)
    out = []
    for e in range(ring_row[v], ring_row[v + 1]):
        for t in range(stride):
            u = int(ring_val[e * stride + t])
            if u not in out:
                out.append(u)
            if len(out) == k:
                return out
    return out


def features(pos, nbrs, v):
    """float32 features (bf16-rounded, as float32) and L of vertex v: the
    kernel's sequential float32 arithmetic."""
    p = np.asarray(pos, dtype=np.float32)
    vx, vy, vz = p[v]
    s = np.float32(0.0)
    for u in nbrs:
        dx, dy, dz = p[u, 0] - vx, p[u, 1] - vy, p[u, 2] - vz
        s = np.float32(s + np.sqrt(np.float32(np.float32(dx * dx + dy * dy) + dz * dz)))
    L = np.float32(s / np.float32(len(nbrs))) if nbrs else np.float32(1.0)
    f = np.zeros(96, dtype=np.float32)
    for k, u in enumerate(nbrs):
        f[3 * k] = (p[u, 0] - vx) / L
        f[3 * k + 1] = (p[u, 1] - vy) / L
        f[3 * k + 2] = (p[u, 2] - vz) / L
    return bf16(f).astype(np.float32), L
