"""ORACLE - test infrastructure only: the neural radiance cache in numpy float64.

The reference has no NRC code (docs/implementation/M7-nrc-adaptive-integrator/README.md: rows 2-5
"not started"), so this module states the network the workload specifies and is the checker of the
GPU kernels (verion_amd/csrc/vh_nrc.hip):
  input   12-frequency position encoding (sin, cos of 2^k pi p for k = 0..11 and the three axes of the
          position normalised to the scene box), real spherical harmonics of the outgoing direction up
          to degree 3 (16 terms, "degree-4" counts the bands), the shading normal, the roughness,
          4 zeros: 96 values;
  MLP     96 -> 64 -> 64 -> 64 -> 64 -> 64 -> 16 (the first 3 outputs are RGB), ReLU on the five
          hidden layers, linear output;
  loss    relative L2, sum over RGB of (y - t)^2 / (y^2 + 0.01) with y held constant in the
          denominator, averaged over the batch;
  Adam    beta1 0.9, beta2 0.99, eps 1e-8, lr 1e-2, bias-corrected.
Parameters are one flat float32 vector: W0 (96x64, row-major [k][n]), b0 (64), W1..W4 (64x64) with
b1..b4, W5 (64x16), b5 (16).

`forward(..., bf16=True)` rounds the encoding, the weights and every hidden activation to bfloat16
(round to nearest even) as the GPU does, accumulating in float64; the GPU accumulates in float32,
so the two agree to bfloat16 rounding. Pinned by tests/test_nrc_oracle.py: central finite
differences of the loss against `backward`, Adam's closed form, the SH basis orthonormality.
"""
from __future__ import annotations

import math

import numpy as np

IN = 96
HID = 64
OUT = 16
LAYERS = [(IN, HID), (HID, HID), (HID, HID), (HID, HID), (HID, HID), (HID, OUT)]
EPS_LOSS = 0.01
BETA1, BETA2, EPS_ADAM = 0.9, 0.99, 1.0e-8
FREQS = 12


def param_count() -> int:
    return sum(k * n + n for k, n in LAYERS)


def unpack(theta: np.ndarray):
    """Flat parameters -> [(W [k][n], b [n])]."""
    out, o = [], 0
    for k, n in LAYERS:
        W = theta[o:o + k * n].reshape(k, n)
        o += k * n
        b = theta[o:o + n]
        o += n
        out.append((W, b))
    return out


def init(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases (the output layer scaled by 0.1), float32."""
    rng = np.random.default_rng(seed)
    parts = []
    for li, (k, n) in enumerate(LAYERS):
        lim = math.sqrt(6.0 / k) * (0.1 if li == len(LAYERS) - 1 else 1.0)
        W = rng.uniform(-lim, lim, size=(k, n))
        if li == len(LAYERS) - 1:
            W[:, 3:] = 0.0
        parts += [W.reshape(-1), np.zeros(n)]
    return np.concatenate(parts).astype(np.float32)


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    nan = np.isnan(f)
    out = rounded.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


SH_C = [0.28209479177387814, 0.4886025119029199, 1.0925484305920792, 0.31539156525252005, 0.5462742152960396,
        0.5900435899266435, 2.890611442640554, 0.4570457994644658, 0.3731763325901154, 1.445305721320277]


def sh16(d: np.ndarray) -> np.ndarray:
    """Real spherical harmonics up to degree 3 of unit directions d (n, 3)."""
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    c = SH_C
    return np.stack([
        np.full_like(x, c[0]),
        -c[1] * y, c[1] * z, -c[1] * x,
        c[2] * x * y, -c[2] * y * z, c[3] * (3 * z * z - 1), -c[2] * x * z, c[4] * (x * x - y * y),
        -c[5] * y * (3 * x * x - y * y), c[6] * x * y * z, -c[7] * y * (5 * z * z - 1), c[8] * z * (5 * z * z - 3),
        -c[7] * x * (5 * z * z - 1), c[9] * z * (x * x - y * y), -c[5] * x * (x * x - 3 * y * y),
    ], axis=1)


def encode(raw: np.ndarray) -> np.ndarray:
    """raw (n, 12): position in [0,1]^3, roughness, direction xyz, -, normal xyz, - -> (n, 96) float64."""
    raw = np.asarray(raw, np.float64).reshape(-1, 12)
    p = np.clip(raw[:, 0:3], 0.0, 1.0)
    rough = raw[:, 3]
    wo = raw[:, 4:7]
    n = raw[:, 8:11]
    feats = np.zeros((len(raw), IN))
    for d in range(3):
        for k in range(FREQS):
            a = (2.0 ** k) * math.pi * p[:, d]
            feats[:, d * 24 + 2 * k] = np.sin(a)
            feats[:, d * 24 + 2 * k + 1] = np.cos(a)
    feats[:, 72:88] = sh16(wo)
    feats[:, 88:91] = n
    feats[:, 91] = rough
    return feats


def forward(theta, raw, bf16_round: bool = False):
    """Returns the output (n, 16) and the per-layer inputs and pre-activations for backward."""
    layers = unpack(np.asarray(theta, np.float64))
    x = encode(raw)
    if bf16_round:
        x = bf16(x)
    acts, pres = [x], []
    for li, (W, b) in enumerate(layers):
        Wl = bf16(W) if bf16_round else W
        z = acts[-1] @ Wl + b
        pres.append(z)
        if li < len(layers) - 1:
            h = np.maximum(z, 0.0)
            if bf16_round:
                h = bf16(h)
            acts.append(h)
    return pres[-1], acts, pres


def loss(y, target):
    y3 = y[:, :3]
    return float(np.mean(np.sum((y3 - target) ** 2 / (y3 * y3 + EPS_LOSS), axis=1)))


def loss_grad(y, target):
    """dL/dy (n, 16) with the denominator held constant."""
    g = np.zeros_like(y)
    y3 = y[:, :3]
    g[:, :3] = 2.0 * (y3 - target) / (y3 * y3 + EPS_LOSS) / len(y)
    return g


def backward(theta, raw, target, bf16_round: bool = False) -> np.ndarray:
    """Gradient of `loss` with respect to the flat parameters (float64)."""
    layers = unpack(np.asarray(theta, np.float64))
    y, acts, pres = forward(theta, raw, bf16_round)
    dz = loss_grad(y, target)
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        grads[li] = (acts[li].T @ dz, dz.sum(axis=0))
        if li > 0:
            dh = dz @ (bf16(W) if bf16_round else W).T
            dz = dh * (pres[li - 1] > 0.0)
    return np.concatenate([np.concatenate([g.reshape(-1), gb]) for g, gb in grads])


def adam(theta, m, v, g, step: int, lr: float = 1.0e-2):
    """One bias-corrected Adam update (step counts from 1); returns the new theta, m, v."""
    m = BETA1 * m + (1 - BETA1) * g
    v = BETA2 * v + (1 - BETA2) * g * g
    mh = m / (1 - BETA1 ** step)
    vh = v / (1 - BETA2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + EPS_ADAM), m, v
