"""Neural radiance cache (bi_pt_nrc_*): the network and its place in the wave engine.

Network numerics (BF16 operands on the matrix cores, FP32 accumulation) are checked
against a torch restatement with the same BF16 rounding points (inputs, weights, every
hidden activation) accumulated in float64; the training step's gradient against torch
autograd of the same loss; the Adam step against its update formula; and learning on a
known smooth function.  In the engine, the CPU restatement's cache modes
(oracle/ref32.cpp CacheMode) pin the wiring bit for bit: a cache out of the paths' reach
leaves the image unchanged, paths ending in a constant cache give the restatement's
image, and the training records and targets equal the restatement's.
"""
import math

import numpy as np
import pytest

import bi_pt
import scenes
from nrc_ref import bf16, encode, forward_ref, unpack

pytestmark = pytest.mark.gpu


def random_params(rng, out_scale=0.3):
    p = np.zeros(bi_pt.NRC_PARAMS, np.float32)
    p[:16384] = rng.normal(0.0, math.sqrt(2.0 / 64), 16384)
    w4 = np.zeros((64, 16), np.float32)
    w4[:, :3] = rng.normal(0.0, out_scale, (64, 3))
    p[16384:] = w4.reshape(-1)
    return p


def random_records(rng, n, lo, hi, valid_frac=1.0):
    rec = np.zeros((n, 16), np.float32)
    rec[:, 0:3] = rng.uniform(lo, hi, (n, 3))
    for k in (3, 6):
        d = rng.normal(size=(n, 3))
        rec[:, k:k + 3] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rec[:, 9] = rng.uniform(0.05, 1.0, n)
    rec[:, 10:13] = rng.uniform(0.1, 0.9, (n, 3))
    rec[:, 13].view(np.uint32)[:] = (rng.uniform(size=n) < valid_frac).astype(np.uint32)
    return rec


@pytest.fixture(scope="module")
def nrc_pair(c0_scene):
    t = bi_pt.PathTracer()
    t.set_scene(c0_scene)
    n = bi_pt.Nrc(t, train_samples=4096, train_steps=4, learning_rate=1e-2)   # the Adam tests' step size
    yield t, n
    n.close()
    t.close()


LO, HI = (-2.0, 0.0, -3.0), (4.0, 5.0, 1.0)


def test_forward_matches_torch(nrc_pair):
    import torch
    print("case: cache forward, 1000 records, random weights", flush=True)
    _, nrc = nrc_pair
    rng = np.random.default_rng(1)
    p = random_params(rng)
    nrc.set_params(p)
    nrc.set_bounds(LO, HI)
    assert np.array_equal(nrc.params(), p)
    rec = random_records(rng, 1000, LO, HI)
    got = nrc.infer(torch.as_tensor(rec).cuda()).cpu().numpy().astype(np.float64)
    want = forward_ref(p, encode(rec, LO, HI))
    scale = np.abs(want).max()
    err = np.abs(got - want)
    print(f"forward: max err {err.max() / scale:.2e} of max |out|, mean {err.mean() / np.abs(want).mean():.2e}")
    assert err.max() <= 1e-2 * scale
    assert err.mean() <= 2e-3 * np.abs(want).mean()


def test_gradient_and_adam_match(nrc_pair):
    import torch
    print("case: cache training step, 200 samples (40 invalid)", flush=True)
    _, nrc = nrc_pair
    rng = np.random.default_rng(2)
    p = random_params(rng)
    nrc.set_params(p)
    nrc.set_bounds(LO, HI)
    n = 200
    rec = random_records(rng, n, LO, HI, valid_frac=0.8)
    tgt = rng.uniform(0.0, 2.0, (n, 3)).astype(np.float32)
    nrc.train(torch.as_tensor(rec).cuda(), torch.as_tensor(tgt).cuda())
    g = nrc.gradient()
    p1 = nrc.params()
    # reference gradient: float64 autograd through the BF16-rounded forward (rounding
    # of the hidden activations passed straight through)
    valid = torch.as_tensor((rec[:, 13].view(np.uint32) == 1).astype(np.float64))
    W = [torch.tensor(bf16(w), requires_grad=True) for w in unpack(p)]
    h = torch.as_tensor(bf16(encode(rec, LO, HI)))
    for l in range(4):
        r = torch.relu(h @ W[l])
        h = r + (torch.as_tensor(bf16(r.detach().numpy())) - r.detach())
    y = (h @ W[4])[:, :3]
    t = torch.as_tensor(tgt.astype(np.float64))
    loss = ((((y - t) ** 2) / (y.detach() ** 2 + 0.01)).sum(1) * valid).sum() / (3 * n)
    loss.backward()
    gr = np.concatenate([w.grad.numpy().reshape(-1) for w in W])
    for l, (a, b) in enumerate([(0, 4096), (4096, 8192), (8192, 12288), (12288, 16384), (16384, 17408)]):
        rel = np.linalg.norm(g[a:b] - gr[a:b]) / np.linalg.norm(gr[a:b])
        print(f"layer {l}: gradient relative error {rel:.3e}")
        assert rel < 3e-2, l
    assert np.all(g[16384:].reshape(64, 16)[:, 3:] == 0)
    # Adam, first step, in float32 in the kernel's operation order
    f = np.float32
    b1, b2, lr, eps = f(0.9), f(0.99), f(0.01), f(1e-15)
    m = b1 * f(0) + (f(1) - b1) * g
    v = b2 * f(0) + (f(1) - b2) * (g * g)
    c1 = f(1.0 / (1.0 - float(b1)))
    c2 = f(1.0 / (1.0 - float(b2)))
    want = (p - lr * (m * c1) / (np.sqrt(v * c2) + eps)).astype(np.float32)
    np.testing.assert_allclose(p1, want, rtol=2e-6, atol=1e-9)


def test_learns_a_smooth_function(nrc_pair):
    import torch
    print("case: cache learns t = albedo (0.2 + max(n.y, 0)) (1 + 0.5 sin 3x), 300 steps of 2048", flush=True)
    _, nrc = nrc_pair
    nrc.set_params(random_params(np.random.default_rng(3), out_scale=0.05))
    nrc.set_bounds(LO, HI)
    rng = np.random.default_rng(4)

    def batch(n):
        rec = random_records(rng, n, LO, HI)
        t = rec[:, 10:13] * (0.2 + np.maximum(rec[:, 7:8], 0.0)) * (1.0 + 0.5 * np.sin(3.0 * rec[:, 0:1]))
        return rec, t.astype(np.float32)

    losses = []
    for k in range(300):
        rec, t = batch(2048)
        nrc.train(torch.as_tensor(rec).cuda(), torch.as_tensor(t).cuda())
        if k < 5 or k >= 290:
            losses.append(nrc.stats()["last_loss"])
    first, last = np.mean(losses[:5]), np.mean(losses[5:])
    rec, t = batch(4096)
    pred = nrc.infer(torch.as_tensor(rec).cuda()).cpu().numpy()
    rel = np.abs(pred - t).mean() / np.abs(t).mean()
    print(f"loss {first:.4f} -> {last:.4f}; held-out mean relative error {rel:.3f}")
    assert last < 0.1 * first
    assert rel < 0.1


def test_cache_out_of_reach_leaves_image(c0_scene):
    print("case: c0 64x48 4 spp, cache at segment 12 of 4: image equals bi_pt_render", flush=True)
    cam = scenes.camera_c0(64, 48)
    plain = bi_pt.PathTracer()
    plain.set_scene(c0_scene)
    plain.render(cam, spp=4, max_bounces=5, seed=31)
    want = plain.image().copy()
    t = bi_pt.PathTracer()
    t.set_scene(c0_scene)
    nrc = bi_pt.Nrc(t, train_samples=4096, train_steps=4, cache_bounce=12)
    nrc.render(cam, spp=4, max_bounces=5, seed=31)
    assert np.array_equal(t.image().view(np.uint32), want.view(np.uint32))
    assert nrc.stats()["last_valid_samples"] == 0
    nrc.close()
    t.close()
    plain.close()


def constant_params(k):
    """A network whose output is the constant k (the constant input carried through unit 0)."""
    p = np.zeros(bi_pt.NRC_PARAMS, np.float32)
    W = unpack(p)
    W[0][59, 0] = 1.0
    for l in (1, 2, 3):
        W[l][0, 0] = 1.0
    W[4][0, :3] = k
    return p


def test_constant_cache_matches_oracle(c0_scene, c0_oracle):
    print("case: c0 96x64 2 spp depth 4, paths end at segment 2 in a constant cache", flush=True)
    cam = scenes.camera_c0(96, 64)
    k = np.array([0.5, 0.25, 1.0], np.float32)
    t = bi_pt.PathTracer()
    t.set_scene(c0_scene)
    nrc = bi_pt.Nrc(t, train_samples=4096, train_steps=4, cache_bounce=2)
    nrc.set_params(constant_params(k))
    nrc.render(cam, spp=2, max_bounces=5, seed=77)   # trains after the frame, renders with k
    got = t.sums(cam.width * cam.height)
    want = c0_oracle.render32_cached(cam, 5, 77, 2, 2, k)
    plain = c0_oracle.render32(cam, 5, 77, spp=2)["sums"]
    assert not np.array_equal(want, plain)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    nrc.close()
    t.close()


def test_training_samples_match_oracle(c0_scene, c0_oracle):
    print("case: c0 80x56, 4096 training paths recording at segment 2: records and targets", flush=True)
    cam = scenes.camera_c0(80, 56)
    t = bi_pt.PathTracer()
    t.set_scene(c0_scene)
    nrc = bi_pt.Nrc(t, train_samples=4096, train_steps=4, cache_bounce=2)
    nrc.render(cam, spp=1, max_bounces=5, seed=5)   # frame 0
    nrc.render(cam, spp=1, max_bounces=5, seed=5)   # frame 1: another set of training pixels
    rec, tgt = nrc.samples(4096)
    orec, otgt = c0_oracle.train_samples(cam, 5, bi_pt.nrc_train_seed(5, 1), 2, 4096)
    valid = orec[:, 13].view(np.uint32) == 1
    print(f"training samples: {valid.sum()} of 4096 reach segment 2")
    assert valid.sum() > 500
    assert np.array_equal(rec.view(np.uint32), orec.view(np.uint32))
    assert np.array_equal(tgt.view(np.uint32), otgt.view(np.uint32))
    st = nrc.stats()
    assert st["frames"] == 2 and st["steps"] == 8 and st["last_valid_samples"] == valid.sum()
    nrc.close()
    t.close()


def test_online_training_learns_suffix_radiance(c0_scene, c0_oracle):
    """128 frames of online training on C0 (65,536 training paths, 4 Adam steps per frame,
    targets clamped to 16x their mean luminance).  Held out: 8,192 vertices of other paths,
    each target the mean of 1,024 continuations (the restatement's), so the targets' own
    noise stays well under the spread the cache has to explain.  The trained cache's
    relative squared error must be at most 0.7x that of the best constant."""
    import torch
    print("case: c0 160x120, 128 frames of online training; held-out vertices, 1024-suffix targets", flush=True)
    cam = scenes.camera_c0(160, 120)
    t = bi_pt.PathTracer()
    t.set_scene(c0_scene)
    nrc = bi_pt.Nrc(t, train_samples=65536, train_steps=4, cache_bounce=2, seed=7)
    fresh = bi_pt.Nrc(t, train_samples=4096, train_steps=4, cache_bounce=2, seed=7)   # the initial weights
    orec, otgt = c0_oracle.train_samples(cam, 5, 0xABCD, 2, 8192, suffixes=1024)
    v = orec[:, 13].view(np.uint32) == 1
    rec = torch.as_tensor(orec[v]).cuda()
    tgt = otgt[v].astype(np.float64)
    losses = []
    for f in range(128):
        nrc.render(cam, spp=1, max_bounces=5, seed=100 + f)
        if f % 16 == 15:
            losses.append(nrc.stats()["last_loss"])
    cb = t.bvh()["centroid_bounds"]
    fresh.set_bounds(cb[:3], cb[3:])

    def rel_err(n):
        p = np.maximum(n.infer(rec).cpu().numpy().astype(np.float64), 0.0)
        return ((p - tgt) ** 2).sum() / (tgt ** 2).sum()

    before, after = rel_err(fresh), rel_err(nrc)
    const = ((tgt - tgt.mean(0)) ** 2).sum() / (tgt ** 2).sum()
    img = t.image()
    print(f"held-out relative squared error: initial {before:.4f}, trained {after:.4f}, best constant {const:.4f} "
          f"(ratio {after / const:.3f}, {int(v.sum())} vertices); training loss every 16 frames "
          f"{' '.join(f'{x:.2f}' for x in losses)}", flush=True)
    assert np.all(np.isfinite(img)) and np.all(img[:, :3] >= 0)
    assert after < 0.9 * before
    assert after <= 0.7 * const
    fresh.close()
    nrc.close()
    t.close()
