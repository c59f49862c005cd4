"""Restatement of the radiance cache's network for its tests: input encoding (scene-box
position in 6 frequencies per axis, outgoing direction in real spherical harmonics of
bands 0-3, normal, roughness, base colour, constant 1) and the forward pass with BF16
rounding of inputs, weights and hidden activations, accumulated in float64."""
import math

import numpy as np

# real spherical harmonics, bands 0-3 (closed forms)
SH = (0.28209479177387814, 0.4886025119029199, 1.0925484305920792, 0.9461746957575601, 0.31539156525252005,
      0.5462742152960396, 0.5900435899266435, 2.890611442640554, 0.4570457994644658, 0.3731763325901154,
      1.445305721320277)


def sh16(d):
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    xx, yy, zz = x * x, y * y, z * z
    return np.stack([np.full_like(x, SH[0]), -SH[1] * y, SH[1] * z, -SH[1] * x,
                     SH[2] * x * y, -SH[2] * y * z, SH[3] * zz - SH[4], -SH[2] * x * z, SH[5] * (xx - yy),
                     -SH[6] * y * (3 * xx - yy), SH[7] * x * y * z, -SH[8] * y * (5 * zz - 1),
                     SH[9] * z * (5 * zz - 3), -SH[8] * x * (5 * zz - 1), SH[10] * z * (xx - yy),
                     -SH[6] * x * (xx - 3 * yy)], axis=1)


def encode(rec, lo, hi):
    """64 network inputs per record (float64 from the float32 record)."""
    rec = rec.astype(np.float64)
    n = len(rec)
    x = np.zeros((n, 64))
    lo = np.asarray(lo, np.float32).astype(np.float64)
    ext = np.asarray(hi, np.float32).astype(np.float64) - lo
    u = ((rec[:, 0:3] - lo) / ext).astype(np.float32).astype(np.float64)
    for a in range(3):
        for f in range(6):
            v = u[:, a] * (1 << f)
            x[:, a * 12 + 2 * f] = np.sin(math.pi * v)
            x[:, a * 12 + 2 * f + 1] = np.cos(math.pi * v)
    x[:, 36:52] = sh16(rec[:, 3:6])
    x[:, 52:55] = rec[:, 6:9]
    x[:, 55] = rec[:, 9]
    x[:, 56:59] = rec[:, 10:13]
    x[:, 59] = 1.0
    return x


def bf16(a):
    import torch
    return torch.as_tensor(np.asarray(a, np.float32)).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)


def unpack(p):
    return [p[l * 4096:(l + 1) * 4096].reshape(64, 64) for l in range(4)] + [p[16384:].reshape(64, 16)]


def forward_ref(p, X):
    W = unpack(p)
    h = bf16(X)
    for l in range(4):
        h = bf16(np.maximum(h @ bf16(W[l]), 0.0))
    return (h @ bf16(W[4]))[:, :3]
