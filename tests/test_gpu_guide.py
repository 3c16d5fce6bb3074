"""Path-guiding network on the GPU (BF16 MFMA kernels) against the float64
oracle in oracle/guide_ref.py.

Tolerances: the oracle mirrors the device's bfloat16 rounding points
(weights, activations, output gradients), so the remaining difference is
FP32 accumulation order plus an occasional one-ulp bfloat16 flip of an
activation: raw outputs agree to 2e-2 of their scale at worst and 2e-3 on
average; the gradient agrees to 2% in norm with the bf16-mirroring oracle
and has cosine > 0.99 with the exact float64 gradient.
"""
import numpy as np
import pytest

import guide_ref as G
from test_oracle_guide import BOX, rand_params, rand_samples

pytestmark = pytest.mark.gpu


def _guide(ctx, params=None, seed=1):
    from ajisai_render_amd.core import PathGuide
    g = PathGuide(ctx, BOX, seed=seed)
    if params is not None:
        g.set_params(params.astype(np.float32))
    return g


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


@pytest.mark.parametrize("n", [1, 15, 16, 1037, 65536])
def test_forward_matches_oracle(gpu_ctx, n):
    print(f"case: guide forward n={n}", flush=True)
    rng = np.random.default_rng(100 + n)
    params = rand_params(rng).astype(np.float32)
    g = _guide(gpu_ctx, params)
    smp = rand_samples(rng, n).astype(np.float32)
    mix, raw = g.forward(_dev(smp[:, 0:3]), _dev(smp[:, 3:6]))
    raw = raw.cpu().numpy().astype(np.float64)
    x = G.encode(smp[:, 0:3], smp[:, 3:6], BOX, f32=True)
    ref, _ = G.forward(params.astype(np.float64), x, bf16_mode=True)
    scale = max(1.0, np.abs(ref).max())
    err = np.abs(raw - ref)
    assert err.max() <= 2e-2 * scale, err.max()
    assert err.mean() <= 2e-3 * scale, err.mean()
    # decoded mixture = decode of the device's own raw outputs
    mu, kappa, w = G.decode(raw)
    m = mix.cpu().numpy().astype(np.float64).reshape(n, 8, 5)
    assert np.allclose(m[..., :3], mu, atol=1e-5)
    assert np.allclose(m[..., 3], kappa, rtol=1e-5)
    assert np.allclose(m[..., 4], w, rtol=1e-4, atol=1e-7)
    g.close()


def test_initial_guide_is_broad(gpu_ctx):
    g = _guide(gpu_ctx, seed=5)
    rng = np.random.default_rng(11)
    smp = rand_samples(rng, 4096).astype(np.float32)
    mix, _ = g.forward(_dev(smp[:, 0:3]), _dev(smp[:, 3:6]), want_raw=False)
    m = mix.cpu().numpy().reshape(-1, 8, 5)
    # small output layer: kappa near 1, weights near 1/8
    assert np.all(m[..., 3] < 3) and np.all(np.abs(m[..., 4] - 0.125) < 0.1)
    p = g.get_params()
    assert p.shape == (G.N_PARAMS,)
    # unused inputs / outputs carry zero weights
    W1 = p[:112 * 64].reshape(112, 64)
    assert np.all(W1[99:] == 0)
    W4 = p[G.offsets()[3][0]:G.offsets()[3][1]].reshape(64, 48)
    assert np.all(W4[:, 40:] == 0)
    g.close()


@pytest.mark.parametrize("n", [300, 4099])
def test_gradient_matches_oracle(gpu_ctx, n):
    print(f"case: guide gradient n={n}", flush=True)
    rng = np.random.default_rng(200 + n)
    params = rand_params(rng).astype(np.float32)
    g = _guide(gpu_ctx, params)
    smp = rand_samples(rng, n).astype(np.float32)
    loss, grad = g.train_step(_dev(smp), lr=0.0, want_grad=True)
    p64 = params.astype(np.float64)
    loss_q, grad_q = G.loss_and_grad(p64, smp.astype(np.float64), BOX, bf16_mode=True)
    loss_x, grad_x = G.loss_and_grad(p64, smp.astype(np.float64), BOX, bf16_mode=False)
    assert abs(loss - loss_q) <= 1e-3 * (1 + abs(loss_q)), (loss, loss_q)
    assert abs(loss - loss_x) <= 2e-2 * (1 + abs(loss_x)), (loss, loss_x)
    rel = np.linalg.norm(grad - grad_q) / np.linalg.norm(grad_q)
    assert rel < 2e-2, rel
    cos = grad @ grad_x / (np.linalg.norm(grad) * np.linalg.norm(grad_x))
    assert cos > 0.99, cos
    # lr = 0 leaves the parameters untouched
    assert np.array_equal(g.get_params(), params)
    g.close()


def test_adam_step_matches_oracle(gpu_ctx):
    rng = np.random.default_rng(300)
    params = rand_params(rng).astype(np.float32)
    g = _guide(gpu_ctx, params)
    smp = rand_samples(rng, 2048).astype(np.float32)
    _, grad1 = g.train_step(_dev(smp), lr=1e-3, want_grad=True)
    p1 = g.get_params().astype(np.float64)
    ref1, m, v = G.adam(params.astype(np.float64), np.zeros(G.N_PARAMS), np.zeros(G.N_PARAMS),
                        grad1.astype(np.float64), 1, 1e-3)
    assert np.allclose(p1, ref1, rtol=0, atol=1e-6)
    _, grad2 = g.train_step(_dev(smp), lr=1e-3, want_grad=True)
    p2 = g.get_params().astype(np.float64)
    ref2, _, _ = G.adam(p1, m, v, grad2.astype(np.float64), 2, 1e-3)
    assert np.allclose(p2, ref2, rtol=0, atol=1e-6)
    g.close()


def test_training_learns_a_position_dependent_lobe(gpu_ctx):
    """Directions drawn from a vMF (kappa 30) whose mean depends on the
    position: the trained mixture must put its mass there."""
    print("case: guide training 400 steps x 16384", flush=True)
    rng = np.random.default_rng(400)
    g = _guide(gpu_ctx, seed=9)
    n = 16384

    def batch():
        pos = rng.uniform(BOX[:3], BOX[3:], (n, 3))
        nrm = np.tile([0.0, 0.0, 1.0], (n, 1))
        mu = np.where(pos[:, :1] > 1.0, [0.0, 0.6, 0.8], [0.8, 0.0, -0.6])
        d = G.vmf_sample(mu, np.full(n, 30.0), rng.random(n), rng.random(n))
        return np.concatenate([pos, nrm, d, np.ones((n, 1))], axis=1), mu

    losses = []
    for _ in range(400):
        smp, _ = batch()
        loss, _ = g.train_step(_dev(smp), lr=1e-3)
        losses.append(loss)
    smp, mu = batch()
    # entropy of vMF(30) = -E[log q] ~ -log(30 / 2pi) + 1 = -0.563
    assert losses[0] > 1.5 and np.mean(losses[-20:]) < 0.0, (losses[0], losses[-20:])
    mix, raw = g.forward(_dev(smp[:, 0:3]), _dev(smp[:, 3:6]))
    q_mean = G.mixture_pdf(raw.cpu().numpy(), mu)
    assert np.median(q_mean) > 2.0  # uniform density is 1/(4 pi) = 0.08
    g.close()


@pytest.mark.parametrize("n", [4096, 32768])
def test_training_step_is_bitwise_reproducible(gpu_ctx, n):
    """Identical steps on identical inputs give bit-identical gradients
    (fixed-order reductions throughout; the network's MFMAs are 16x16x16)."""
    print(f"case: guide gradient reproducibility n={n}", flush=True)
    rng = np.random.default_rng(7 + n)
    g = _guide(gpu_ctx, rand_params(rng).astype(np.float32))
    smp = _dev(rand_samples(rng, n))
    grads = [g.train_step(smp, lr=0.0, want_grad=True)[1] for _ in range(4)]
    for k in range(1, 4):
        assert np.array_equal(grads[0], grads[k]), np.abs(grads[0] - grads[k]).max()
    g.close()
