"""CPU reference of the learned warm start's network and features (numpy).

TEST INFRASTRUCTURE ONLY (checker for conductorrigid_amd/csrc/k_mlp.hip), like the rest
of oracle/.  The network has no counterpart in the reference; this module states it
independently of the GPU kernels:

    x = feature row; trunk inputs x[:TRUNK], gate g = (E0 from x[GATE:GATE+6] as hi + lo, 1)
    z1 = x W1 + b1, h1 = silu(z1); z2 = h1 W2 + b2, h2 = silu(z2); u = h2 W3 + b3 (4)
    y = g . u;  loss = mean_m (y_m - t_m)^2

With ``bf16_points=True`` it rounds at the same points as the GPU (features, W1, W2, W3,
h1, h2 where they feed a product, dz1, dz2 where they feed a product); with
``bf16_points=False`` it is exact float64 arithmetic, which the finite-difference checks
pin.
"""
from __future__ import annotations

import numpy as np

IN, TRUNK, GATE, H, OUT, NEIGH, RADIUS = 176, 168, 168, 128, 4, 32, 3.0
OFF_W1 = 0
OFF_B1 = OFF_W1 + IN * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * OUT
PARAMS = OFF_B3 + OUT


def bf16_bits(a) -> np.ndarray:
    """float32 -> bf16 bits, round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_to_f32(bits) -> np.ndarray:
    return (np.asarray(bits, dtype=np.uint32) << 16).view(np.float32)


def bf16(a) -> np.ndarray:
    return bf16_to_f32(bf16_bits(a))


def unpack(theta):
    t = np.asarray(theta, dtype=np.float64)
    return (t[OFF_W1:OFF_B1].reshape(IN, H), t[OFF_B1:OFF_W2], t[OFF_W2:OFF_B2].reshape(H, H), t[OFF_B2:OFF_W3],
            t[OFF_W3:OFF_B3].reshape(H, OUT), t[OFF_B3:PARAMS])


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def gates(X):
    """[M, 4]: (E0x, E0y, E0z, 1) per row, E0 = hi + lo from the gate columns."""
    X = np.asarray(X, dtype=np.float64)
    g = np.ones((X.shape[0], OUT))
    g[:, :3] = X[:, GATE:GATE + 3] + X[:, GATE + 3:GATE + 6]
    return g


def forward(theta, X, bf16_points=True):
    """X: float array [M, IN] (already bf16-exact when bf16_points).  Returns y and the
    cached activations."""
    W1, b1, W2, b2, W3, b3 = unpack(theta)
    X = np.asarray(X, dtype=np.float64)
    Xt = X.copy()
    Xt[:, TRUNK:] = 0.0
    r = (lambda a: bf16(a).astype(np.float64)) if bf16_points else (lambda a: a)
    z1 = Xt @ r(W1) + b1
    h1 = r(_silu(z1))
    z2 = h1 @ r(W2) + b2
    h2 = r(_silu(z2))
    u = h2 @ r(W3) + b3
    g = gates(X)
    y = (g * u).sum(1)
    return y, dict(Xt=Xt, z1=z1, h1=h1, z2=z2, h2=h2, g=g)


def gradient(theta, X, t, bf16_points=True):
    """Flat gradient (length PARAMS) of mean (y - t)^2, and y."""
    W1, b1, W2, b2, W3, b3 = unpack(theta)
    M = np.asarray(X).shape[0]
    y, c = forward(theta, X, bf16_points)
    r = (lambda a: bf16(a).astype(np.float64)) if bf16_points else (lambda a: a)
    dy = 2.0 * (y - t) / M
    du = dy[:, None] * c["g"]
    dz2 = r((du @ r(W3).T) * _dsilu(c["z2"]))
    dh1 = dz2 @ r(W2).T
    dz1 = r(dh1 * _dsilu(c["z1"]))
    g = np.zeros(PARAMS)
    g[OFF_W1:OFF_B1] = (c["Xt"].T @ dz1).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W2:OFF_B2] = (c["h1"].T @ dz2).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    g[OFF_W3:OFF_B3] = (c["h2"].T @ du).ravel()
    g[OFF_B3:PARAMS] = du.sum(0)
    return g, y


def loss(theta, X, t, bf16_points=False):
    y, _ = forward(theta, X, bf16_points)
    return float(np.mean((y - t) ** 2))


def features(pos, L, E0, iso, body_charge, rigid_membership, rigid_com, head, nn, nlist):
    """Feature rows [N, IN] as floats that are exactly the kernel's bf16 values' sources:
    neighbours from the given list in its order, the NEIGH nearest with r < RADIUS by
    increasing float32 distance (ties keep list order), each as d = (dx, dy, dz)/RADIUS,
    a +1/-1 same-body flag and |d|; then Q_b / n_b, the offset o from the rigid body's
    centre of mass / 4 and |o|; the gate E0 as bf16 hi and lo = E0 - hi.  float32
    arithmetic where the kernel uses it.  Round with bf16_bits() to compare."""
    N = pos.shape[0]
    L = np.asarray(L, dtype=np.float64)
    f32 = np.float32
    e = np.asarray(E0, dtype=np.float64).astype(f32)
    counts = np.bincount(iso, minlength=body_charge.shape[0])
    X = np.zeros((N, IN), dtype=np.float64)
    inv = f32(1.0 / RADIUS)

    def mi(d):
        return d - L * np.rint(d / L)

    for i in range(N):
        js = nlist[head[i]:head[i] + nn[i]].astype(np.int64)
        d = mi(pos[i, :3] - pos[js, :3])
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]).astype(np.float32)
        keep = d2 < np.float32(RADIUS * RADIUS)
        js, d2 = js[keep], d2[keep]
        order = np.argsort(d2, kind="stable")[:NEIGH]
        for k, o in enumerate(order):
            j = js[o]
            X[i, 5 * k:5 * k + 3] = mi(pos[j, :3] - pos[i, :3]).astype(f32) * inv
            X[i, 5 * k + 3] = 1.0 if iso[j] == iso[i] else -1.0
            X[i, 5 * k + 4] = np.sqrt(d2[o]) * inv
        o = 5 * NEIGH
        X[i, o] = body_charge[iso[i]] / counts[iso[i]]
        r = rigid_membership[i]
        if r != 0xFFFFFFFF:
            c = mi(pos[i, :3] - rigid_com[r, :3]).astype(f32) * f32(0.25)
            X[i, o + 1:o + 4] = c
            X[i, o + 4] = np.sqrt(f32(f32(f32(c[0] * c[0]) + f32(c[1] * c[1])) + f32(c[2] * c[2])))
        hi = bf16(e)
        X[i, GATE:GATE + 3] = hi
        X[i, GATE + 3:GATE + 6] = (e - hi).astype(f32)
    return X
