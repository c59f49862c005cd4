"""The radiance cache's input encoding, pinned by mathematics (CPU only).

The 16 direction features must be the real spherical harmonics of bands 0-3: an
orthonormal set on the sphere (checked by a product quadrature exact for these degrees)
that spans each band's polynomials (so the band-l functions are harmonic: Laplacian 0 on
the polynomial extension of degree l).  The frequency features are (sin, cos) pairs at
doubling frequencies.  The GPU encoding is compared with this restatement through the
network in test_gpu_nrc.py.
"""
import math

import numpy as np

from nrc_ref import encode, sh16


def sphere_quadrature(n_theta=24, n_phi=48):
    """Gauss-Legendre in cos(theta) x uniform phi: exact for spherical polynomials of degree < min(2 n_theta, n_phi)."""
    x, w = np.polynomial.legendre.leggauss(n_theta)
    phi = (np.arange(n_phi) + 0.5) * (2 * math.pi / n_phi)
    ct = np.repeat(x, n_phi)
    st = np.sqrt(1 - ct * ct)
    ph = np.tile(phi, n_theta)
    d = np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=1)
    wt = np.repeat(w, n_phi) * (2 * math.pi / n_phi)
    return d, wt


def test_sh_orthonormal():
    d, w = sphere_quadrature()
    Y = sh16(d)
    gram = (Y * w[:, None]).T @ Y
    np.testing.assert_allclose(gram, np.eye(16), atol=1e-12)


def test_sh_bands_are_harmonic():
    # band l (indices l^2 .. l^2 + 2l) as polynomials in (x, y, z): Laplacian zero,
    # checked by central differences of the polynomial extension
    rng = np.random.default_rng(0)
    p = rng.normal(size=(64, 3))
    h = 1e-3
    lap = np.zeros((64, 16))
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        lap += (sh16(p + e) - 2 * sh16(p) + sh16(p - e)) / (h * h)
    # band 0 is constant; bands 1-3 are homogeneous harmonic polynomials except for the
    # z^2 term of (l=2, m=0), whose constant -0.315... makes it 0.315 (3 z^2 - r^2) on the sphere
    on_sphere = p / np.linalg.norm(p, axis=1, keepdims=True)
    Y = sh16(on_sphere)
    r2 = np.sum(on_sphere ** 2, axis=1)
    np.testing.assert_allclose(r2, 1.0)
    for idx in (0, 1, 2, 3, 4, 5, 7, 8):
        np.testing.assert_allclose(lap[:, idx], 0.0, atol=1e-5)
    # (2, 0): 0.946 z^2 - 0.315 equals 0.315 (3 z^2 - (x^2 + y^2 + z^2)) on the sphere
    x, y, z = on_sphere.T
    np.testing.assert_allclose(Y[:, 6], 0.31539156525252005 * (3 * z * z - (x * x + y * y + z * z)), atol=1e-12)
    # band 3 terms with a (5 z^2 - 1) or (5 z^2 - 3) factor are harmonic once 1 -> r^2
    np.testing.assert_allclose(Y[:, 12], 0.3731763325901154 * z * (5 * z * z - 3 * (x * x + y * y + z * z)),
                               atol=1e-12)
    for idx in (9, 10, 14, 15):
        np.testing.assert_allclose(lap[:, idx], 0.0, atol=1e-5)


def test_frequency_features():
    rng = np.random.default_rng(1)
    n = 256
    rec = np.zeros((n, 16), np.float32)
    lo, hi = (-1.0, 0.0, 2.0), (3.0, 2.0, 4.0)
    rec[:, 0:3] = rng.uniform(lo, hi, (n, 3))
    rec[:, 3:6] = [0.0, 0.0, 1.0]
    x = encode(rec, lo, hi)
    s, c = x[:, 0:36:2], x[:, 1:36:2]
    np.testing.assert_allclose(s * s + c * c, 1.0, atol=1e-12)
    # doubling: sin(2a) = 2 sin(a) cos(a) between consecutive frequencies of an axis
    for a in range(3):
        for f in range(5):
            i = a * 6 + f
            np.testing.assert_allclose(s[:, i + 1], 2 * s[:, i] * c[:, i], atol=1e-9)
    # the lowest frequency is sin(pi u) with u the position in the box
    u = (rec[:, 0].astype(np.float64) - lo[0]) / (hi[0] - lo[0])
    np.testing.assert_allclose(s[:, 0], np.sin(math.pi * u), atol=1e-6)
    assert np.all(x[:, 59] == 1.0) and np.all(x[:, 60:] == 0.0)
