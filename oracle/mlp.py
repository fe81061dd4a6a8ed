"""CPU restatement of the c2p surrogate MLP (TEST INFRASTRUCTURE ONLY).

The network has no counterpart in the reference; this numpy version is the
checker for the device kernels' forward pass and gradients.  Parameter layout
and features follow include/rhllc_hip.h (RHLLC_MLP_*):

    W1[64][3] b1[64] W2[64][64] b2[64] W3[64] b3 | mu_x[3] sd_x[3] mu_y sd_y | gamma

Inputs (log D, log1p(|m|/D), log(p0/D)), output log(p/p0), p0 the c2p root
with the Lorentz factor frozen at W0 = E/sqrt(E^2 - m^2) (cold_pressure).

With bf16=True the inputs of the 64x64 products (H1, W2, dZ2, W2^T) are
rounded to bfloat16 (round to nearest even) as the MFMA kernels do; all other
arithmetic is float64.
"""
from __future__ import annotations

import numpy as np

H, IN = 64, 3
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 192, 256, 4352, 4416, 4480
N_PARAMS, N_MODEL = 4481, 4490


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    return (P[OFF_W1:OFF_B1].reshape(H, IN), P[OFF_B1:OFF_W2], P[OFF_W2:OFF_B2].reshape(H, H), P[OFF_B2:OFF_W3],
            P[OFF_W3:OFF_B3], P[OFF_B3])


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def cold_pressure(Gg, D, m, E):
    """Root of E + p = D W0 + Gg p W0^2 with W0 = E/sqrt(E^2 - m^2) fixed, floored at 1e-12 D."""
    with np.errstate(invalid="ignore", divide="ignore"):
        q = (E - m) * (E + m)
        W0 = E / np.sqrt(q)
        p0 = (E - D * W0) / (Gg * (E * E / q) - 1.0)
        return np.where(q > 0, np.fmax(p0, 1e-12 * D), np.nan)


def raw_features(U, gamma):
    """Un-normalised features (log D, log1p(|m|/D), log(p0/D)) as float32, and p0."""
    U = np.asarray(U, dtype=np.float64)
    D, m, E = U[:, 0], np.linalg.norm(U[:, 1:4], axis=1), U[:, 4]
    p0 = cold_pressure(gamma / (gamma - 1.0), D, m, E)
    f = np.float32
    with np.errstate(invalid="ignore", divide="ignore"):
        raw = np.column_stack([np.log(D.astype(f)), np.log1p((m / D).astype(f)), np.log((p0 / D).astype(f))])
    return raw.astype(np.float64), p0


def features(U, model):
    """Normalised features of conserved states U[n,5], and p0."""
    norm = np.asarray(model, dtype=np.float64)[N_PARAMS:]
    raw, p0 = raw_features(U, norm[8])
    return (raw - norm[0:3]) / norm[3:6], p0


def forward(P, X, bf=True):
    W1, b1, W2, b2, W3, b3 = unpack(P)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    z1 = X @ W1.T + b1
    h1 = silu(z1)
    z2 = r(h1) @ r(W2).T + b2
    h2 = silu(z2)
    out = h2 @ W3 + b3
    return out, (z1, h1, z2, h2)


def predict_p(model, U, bf=True):
    norm = np.asarray(model, dtype=np.float64)[N_PARAMS:]
    X, p0 = features(U, model)
    out, _ = forward(model, X, bf)
    return p0 * np.exp(out * norm[7] + norm[6])


def loss_and_grad(P, X, y, ok, bf=True):
    """sum_i ok_i (f(x_i) - y_i)^2 and its gradient / n (the device kernels' convention)."""
    n = X.shape[0]
    W1, b1, W2, b2, W3, b3 = unpack(P)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    out, (z1, h1, z2, h2) = forward(P, X, bf)
    d = (out - y) * ok
    loss = float((d * d).sum())
    dout = 2.0 * d / n
    g = np.zeros(N_PARAMS)
    g[OFF_W3:OFF_B3] = h2.T @ dout
    g[OFF_B3] = dout.sum()
    dz2 = (dout[:, None] * W3[None, :]) * dsilu(z2)
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    g[OFF_W2:OFF_B2] = (r(dz2).T @ r(h1)).reshape(-1)
    dh1 = r(dz2) @ r(W2)
    dz1 = dh1 * dsilu(z1)
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W1:OFF_B1] = (dz1.T @ X).reshape(-1)
    return loss, g
