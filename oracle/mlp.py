"""CPU restatement of the learned warm start's network (numpy).

TEST INFRASTRUCTURE ONLY: the checker used by tests/.  The product never
imports this module.

The network has no counterpart in the reference (it stands in for the
impulse initialisation at src/World/PhysicsWorld.cpp:514-530 and learns the
impulses stored at :568-581).  Two modes:

* ``emulate=True`` mirrors the device arithmetic: normalised inputs,
  weights, hidden activations and backpropagated errors are rounded to BF16
  (round to nearest even) where they enter a product; products accumulate
  exactly here (float64) and are rounded to FP32, as the matrix cores
  accumulate in FP32; biases, bias gradients and the Adam update are FP32.
  The device's FP32 accumulation order differs, so device and oracle agree
  to FP32 rounding of the sums (and, rarely, one BF16 step of an activation).
* ``emulate=False`` is the same network in float64 without rounding: the
  mathematical definition that finite differences check.

Parameter layout (flat, FP32): W1[64][16], b1[64], W2[64][64], b2[64],
W3[3][64], b3[3].  Sample rows: 16 feature floats (all inputs; 14-15 are the
optional body-load features, else 0) + 3 targets + 0.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 16, 64, 3
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_HID * N_IN
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_OUT * N_HID
N_PARAMS = OFF_B3 + N_OUT
assert N_PARAMS == 5443


def bf16(x) -> np.ndarray:
    """Round FP32 values to BF16 (nearest, ties to even); returned as float32."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def unpack(params):
    p = np.asarray(params)
    W1 = p[OFF_W1:OFF_B1].reshape(N_HID, N_IN)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(N_HID, N_HID)
    b2 = p[OFF_B2:OFF_W3]
    W3 = p[OFF_W3:OFF_B3].reshape(N_OUT, N_HID)
    b3 = p[OFF_B3:N_PARAMS]
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3, dtype=np.float32) -> np.ndarray:
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, np.ravel(W3), b3]).astype(dtype)


def _f32(x):
    return np.asarray(x, np.float64).astype(np.float32)


def forward(params, mean, inv_std, features, emulate=True):
    """Outputs y (n x 3) and the intermediates the backward pass needs."""
    feats = np.asarray(features)[:, :N_IN]
    W1, b1, W2, b2, W3, b3 = unpack(params)
    if emulate:
        x = bf16((feats.astype(np.float32) - np.asarray(mean, np.float32)) * np.asarray(inv_std, np.float32))
        W1b, W2b, W3b = (bf16(w).astype(np.float64) for w in (W1, W2, W3))
        xb = x.astype(np.float64)
        z1 = _f32(_f32(xb @ W1b.T) + np.asarray(b1, np.float32))
        h1 = bf16(np.where(z1 > 0, z1, np.float32(0))).astype(np.float64)
        z2 = _f32(_f32(h1 @ W2b.T) + np.asarray(b2, np.float32))
        h2 = bf16(np.where(z2 > 0, z2, np.float32(0))).astype(np.float64)
        y = _f32(_f32(h2 @ W3b.T) + np.asarray(b3, np.float32))
    else:
        xb = (feats.astype(np.float64) - np.asarray(mean, np.float64)) * np.asarray(inv_std, np.float64)
        W1b, W2b, W3b = (np.asarray(w, np.float64) for w in (W1, W2, W3))
        z1 = xb @ W1b.T + b1
        h1 = np.maximum(z1, 0.0)
        z2 = h1 @ W2b.T + b2
        h2 = np.maximum(z2, 0.0)
        y = h2 @ W3b.T + b3
    return y, dict(x=xb, z1=z1, h1=h1, z2=z2, h2=h2, W1=W1b, W2=W2b, W3=W3b)


def loss_and_grad(params, mean, inv_std, samples, emulate=True):
    """Summed squared error and the gradient of the mean squared error
    (over samples and the 3 outputs) with respect to the flat parameters."""
    s = np.asarray(samples)
    n = len(s)
    y, c = forward(params, mean, inv_std, s[:, :16], emulate)
    t = s[:, 16:19]
    if emulate:
        d = _f32(y.astype(np.float32) - t.astype(np.float32))
        scale = np.float32(2.0) / np.float32(3.0 * n)
        dy = (d * scale).astype(np.float32)
        loss = float(np.sum(d.astype(np.float64) ** 2))
        dyb = bf16(dy).astype(np.float64)
        dz2 = _f32(dyb @ c["W3"]) * (c["z2"] > 0)
        dz2b = bf16(dz2).astype(np.float64)
        dz1 = _f32(dz2b @ c["W2"]) * (c["z1"] > 0)
        dz1b = bf16(dz1).astype(np.float64)
        gW3 = dyb.T @ c["h2"]
        gW2 = dz2b.T @ c["h1"]
        gW1 = dz1b.T @ c["x"]
        gb3, gb2, gb1 = dy.astype(np.float64).sum(0), dz2.astype(np.float64).sum(0), dz1.astype(np.float64).sum(0)
    else:
        d = y - t
        loss = float(np.sum(d ** 2))
        dy = d * (2.0 / (3.0 * n))
        dz2 = (dy @ c["W3"]) * (c["z2"] > 0)
        dz1 = (dz2 @ c["W2"]) * (c["z1"] > 0)
        gW3, gW2, gW1 = dy.T @ c["h2"], dz2.T @ c["h1"], dz1.T @ c["x"]
        gb3, gb2, gb1 = dy.sum(0), dz2.sum(0), dz1.sum(0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3, np.float64)


def mean_loss(params, mean, inv_std, samples):
    """Mean squared error in float64 (the function finite differences see)."""
    s = np.asarray(samples)
    y, _ = forward(params, mean, inv_std, s[:, :16], emulate=False)
    return float(np.mean((y - s[:, 16:19]) ** 2))


def adam(params, m, v, grad, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step in FP32 in the device's operation order; returns (params, m, v)."""
    f = np.float32
    g = np.asarray(grad, f)
    b1, b2 = f(beta1), f(beta2)
    m = b1 * np.asarray(m, f) + (f(1) - b1) * g
    v = b2 * np.asarray(v, f) + (f(1) - b2) * g * g
    # bias corrections from the FP32 betas, in double: 1 - double(beta)^t
    c1 = f(1.0 - float(b1) ** t)
    c2 = f(1.0 - float(b2) ** t)
    w = np.asarray(params, f) - f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps))
    return w.astype(f), m.astype(f), v.astype(f)


def sample_hash(seed: int, q: np.ndarray) -> np.ndarray:
    """Minibatch row selection of the training entry point: row = hash(seed, q) % held."""
    M = 0xFFFFFFFF
    h = (np.uint64(seed) ^ ((q.astype(np.uint64) * np.uint64(0x9E3779B1)) & np.uint64(M))) & np.uint64(M)
    h ^= h >> np.uint64(16)
    h = (h * np.uint64(0x85EBCA6B)) & np.uint64(M)
    h ^= h >> np.uint64(13)
    h = (h * np.uint64(0xC2B2AE35)) & np.uint64(M)
    h ^= h >> np.uint64(16)
    return h


def body_loads(body_a, body_b, cached_normal, inv_mass, dynamic, gdt):
    """Load features 14-15 of every point (float64): for each dynamic body,
    the sum over all its points of the cached normal impulse clamped at 0,
    times its inverse mass over |gravity| dt (the weight impulse); 0 for a
    static or kinematic body."""
    lam = np.maximum(np.asarray(cached_normal, np.float64), 0.0)
    nb = len(inv_mass)
    acc = np.zeros(nb)
    for bodies in (body_a, body_b):
        w = np.where(dynamic[bodies], lam * inv_mass[bodies] / gdt, 0.0)
        np.add.at(acc, bodies, w)
    acc = np.where(dynamic, acc, 0.0)
    return acc[body_a], acc[body_b]


def contact_features(ra, rb, n, t1, t2, sep, inv_mass_a, inv_mass_b, inv_inertia_a, inv_inertia_b,
                     va, wa, vb, wb, hit, load_a=0.0, load_b=0.0):
    """Feature rows (n x 16, float64) from contact-point and body data, as
    defined for the learned warm start: velocities are those after velocity
    integration, inverse inertias are world-space 3x3 matrices; load_a /
    load_b are the body-load features (body_loads) or 0 without them."""
    def proj(I, r, d):
        x = np.cross(r, d)
        return np.einsum("ni,nij,nj->n", x, I, x)
    vrel = (va + np.cross(wa, ra)) - (vb + np.cross(wb, rb))
    out = np.zeros((len(ra), 16))
    out[:, 0] = np.einsum("ni,ni->n", vrel, n)
    out[:, 1] = np.einsum("ni,ni->n", vrel, t1)
    out[:, 2] = np.einsum("ni,ni->n", vrel, t2)
    out[:, 3:6] = n
    out[:, 6] = -sep
    out[:, 7] = inv_mass_a
    out[:, 8] = inv_mass_b
    out[:, 9] = proj(inv_inertia_a, ra, n)
    out[:, 10] = proj(inv_inertia_b, rb, n)
    out[:, 11] = 0.5 * (proj(inv_inertia_a, ra, t1) + proj(inv_inertia_a, ra, t2))
    out[:, 12] = 0.5 * (proj(inv_inertia_b, rb, t1) + proj(inv_inertia_b, rb, t2))
    out[:, 13] = hit
    out[:, 14] = load_a
    out[:, 15] = load_b
    return out
