"""Pins of the oracle's integer RNG and float transcendentals (CPU only)."""
import json
import os

import numpy as np
import pytest

import scenes

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_pcg32_published_demo_vector(oracle_lib):
    # pcg32-demo of the PCG reference distribution: initstate 42, initseq 54
    kat = json.load(open(os.path.join(GOLDEN, "pcg32_kat.json")))
    got = oracle_lib.pcg32(kat["initstate"], kat["initseq"], len(kat["outputs"]))
    assert [f"{x:08x}" for x in got] == kat["outputs"]


def test_pcg32_keyed_stream_layout(oracle_lib):
    # Key layout: initstate = pixel << 32 | sample, initseq = seed << 8 | bounce;
    # floats are (next >> 8) * 2^-24.  Cross-checked against Python-integer PCG32.
    rng = np.random.default_rng(1)
    for _ in range(50):
        seed, pixel, sample, bounce = (int(rng.integers(0, 2**32)), int(rng.integers(0, 2**26)),
                                       int(rng.integers(0, 2**20)), int(rng.integers(0, 255)))
        ref = scenes.Pcg32((pixel << 32) | sample, (seed << 8) | bounce)
        want = np.array([ref.uniform() for _ in range(6)], np.float32)
        got = oracle_lib.pcg32_keyed(seed, pixel, sample, bounce, 6)
        assert np.array_equal(got, want)


def test_pcg32_keyed_streams_are_distinct(oracle_lib):
    a = oracle_lib.pcg32_keyed(7, 100, 3, 2, 4)
    b = oracle_lib.pcg32_keyed(7, 100, 3, 3, 4)   # next bounce
    c = oracle_lib.pcg32_keyed(7, 100, 4, 2, 4)   # next sample
    d = oracle_lib.pcg32_keyed(7, 101, 3, 2, 4)   # next pixel
    e = oracle_lib.pcg32_keyed(8, 100, 3, 2, 4)   # next seed
    rows = [a, b, c, d, e]
    for i in range(5):
        for j in range(i + 1, 5):
            assert not np.array_equal(rows[i], rows[j])
    assert np.all((a >= 0) & (a < 1))


def test_pcg32_uniformity(oracle_lib):
    u = np.concatenate([oracle_lib.pcg32_keyed(3, p, 0, 0, 4) for p in range(4000)])
    hist, _ = np.histogram(u, bins=16, range=(0, 1))
    exp = len(u) / 16
    chi2 = ((hist - exp) ** 2 / exp).sum()
    assert chi2 < 40.0  # 15 dof, p ~ 5e-4


@pytest.mark.parametrize("fn,lo,hi", [(0, 0.0, 2 * np.pi), (1, 0.0, 2 * np.pi), (0, -7.0, 7.0), (1, -7.0, 7.0)])
def test_sin_cos_accuracy(oracle_lib, fn, lo, hi):
    x = np.linspace(lo, hi, 200001).astype(np.float32)
    got = oracle_lib.math32(fn, x).astype(np.float64)
    ref = (np.sin if fn == 0 else np.cos)(x.astype(np.float64))
    assert np.max(np.abs(got - ref)) < 2.5e-7
    # exact special points of the octant reduction
    assert oracle_lib.math32(0, np.array([0.0], np.float32))[0] == 0.0
    assert oracle_lib.math32(1, np.array([0.0], np.float32))[0] == 1.0


def test_acos_accuracy(oracle_lib):
    x = np.linspace(-1, 1, 200001).astype(np.float32)
    got = oracle_lib.math32(2, x).astype(np.float64)
    ref = np.arccos(x.astype(np.float64))
    assert np.max(np.abs(got - ref)) < 5e-7
    assert abs(oracle_lib.math32(2, np.array([1.0], np.float32))[0]) == 0.0


def test_atan2_accuracy_and_quadrants(oracle_lib):
    rng = np.random.default_rng(0)
    y = rng.normal(size=100000).astype(np.float32)
    x = rng.normal(size=100000).astype(np.float32)
    got = oracle_lib.math32(3, y, x).astype(np.float64)
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert np.max(np.abs(got - ref)) < 5e-7
    axes = np.array([[1, 0], [0, 1], [-1, 0], [0, -1]], np.float32)
    got = oracle_lib.math32(3, axes[:, 0], axes[:, 1])
    np.testing.assert_allclose(got, [np.pi / 2, 0.0, -np.pi / 2, np.pi], atol=1e-7)
