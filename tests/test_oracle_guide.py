"""Pins the CPU oracle of the path-guiding network (oracle/guide_ref.py) by
mathematics: finite differences, closed-form vMF moments and normalisation,
bfloat16 rounding known answers and the first Adam step's closed form."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import guide_ref as G  # noqa: E402

BOX = np.array([-1.0, -0.5, -2.0, 3.0, 2.5, 2.0])


def rand_params(rng, scale=1.0):
    layers = []
    for ni, no in [(G.N_IN, G.N_HID), (G.N_HID, G.N_HID), (G.N_HID, G.N_HID), (G.N_HID, G.N_OUT)]:
        W = rng.uniform(-1, 1, (ni, no)) * np.sqrt(6.0 / ni) * scale
        b = rng.uniform(-0.1, 0.1, no)
        layers.append((W, b))
    return G.pack(layers)


def rand_samples(rng, n):
    pos = rng.uniform(BOX[:3], BOX[3:], (n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.exponential(size=n)
    return np.concatenate([pos, nrm, d, w[:, None]], axis=1)


def test_layout():
    offs = G.offsets()
    assert offs[0] == (0, 112 * 64, 112, 64)
    assert G.N_PARAMS == 112 * 64 + 64 + 2 * (64 * 64 + 64) + 64 * 48 + 48 == 18672
    p = np.arange(G.N_PARAMS, dtype=np.float64)
    assert np.array_equal(G.pack(G.unpack(p)), p)


def test_encode_known_values():
    pos = np.array([[-1.0, -0.5, -2.0], [3.0, 2.5, 2.0], [1.0, 1.0, 0.0]])
    nrm = np.array([[0.0, 0.0, 1.0], [0.5, -0.5, 0.0], [1.0, 0.0, 0.0]])
    x = G.encode(pos, nrm, BOX, f32=False)
    assert np.allclose(x[0, 96:99], -1) and np.allclose(x[1, 96:99], 1)
    assert np.allclose(x[2, 96:99], 0)
    # n_z = 1 (features 64..95): sin(pi m) = 0, cos(pi m) = (-1)^m for harmonic m = h + 1
    for j in (4, 5):
        for g in range(4):
            for i in range(4):
                m = 8 * (j % 2) + 2 * g + i // 2 + 1
                want = 0.0 if i % 2 == 0 else (-1.0) ** m
                assert abs(x[0, 16 * j + 4 * g + i] - want) < 1e-12
    # n_x = 0.5, harmonic 1: sin(pi/2) = 1, cos = 0; harmonic 2: sin(pi) = 0, cos = -1
    assert np.allclose(x[1, 0:4], [1, 0, 0, -1], atol=1e-12)
    # every (component, harmonic, sin/cos) appears exactly once
    rng = np.random.default_rng(2)
    nn = rng.normal(size=(5, 3))
    xs = G.encode(np.zeros((5, 3)), nn, BOX, f32=False)
    ref = [np.sin(np.pi * k * nn[:, c]) for c in range(3) for k in range(1, 17)] + \
          [np.cos(np.pi * k * nn[:, c]) for c in range(3) for k in range(1, 17)]
    ref = np.sort(np.stack(ref, 1), axis=1)
    assert np.allclose(np.sort(xs[:, :96], axis=1), ref)
    assert np.all(x[:, 99:] == 0)


def test_bf16_known_answers():
    torch = pytest.importorskip("torch")
    one = 1.0
    cases = np.array([one, one + 2 ** -8, one + 3 * 2 ** -8, one + 2 ** -8 + 2 ** -20,
                      -2.5e-3, 65504.0, 3.0e38, 1e-40, 0.0, -0.0], np.float32)
    got = G.bf16(cases)
    assert got[0] == 1.0 and got[1] == 1.0  # tie to even
    assert got[2] == 1.0 + 2 ** -6          # tie to even (upward)
    assert got[3] == 1.0 + 2 ** -7          # above the tie
    ref = torch.from_numpy(cases).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(got, ref)
    rng = np.random.default_rng(3)
    r = (rng.normal(size=10000) * 10.0 ** rng.uniform(-20, 20, 10000)).astype(np.float32)
    ref = torch.from_numpy(r).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(G.bf16(r), ref)


def test_decode_identities():
    rng = np.random.default_rng(4)
    raw = rng.normal(size=(50, 48)) * 4
    mu, kappa, pi = G.decode(raw)
    assert np.allclose(np.linalg.norm(mu, axis=-1), 1)
    assert np.allclose(pi.sum(axis=1), 1) and np.all(pi > 0)
    assert np.all(kappa >= np.exp(-3) * (1 - 1e-12)) and np.all(kappa <= np.exp(7) * (1 + 1e-12))
    # softmax is shift invariant; outputs 40..47 are unused
    raw2 = raw.copy()
    raw2[:, 4:40:5] += 17.0
    raw2[:, 40:] = 1e9
    for a, b in zip(G.decode(raw), G.decode(raw2)):
        assert np.allclose(a, b)


@pytest.mark.parametrize("kappa", [np.exp(-3), 0.5, 3.0, 40.0, np.exp(7)])
def test_vmf_normalisation(kappa):
    # integrate over the polar angle around mu: int_{-1}^{1} 2 pi q(c) dc = 1
    mu = np.array([0.0, 0.0, 1.0])
    # fine near c = 1 where a sharp lobe lives
    c = 1 - np.concatenate([np.geomspace(1e-9, 2.0, 400001)[::-1], [0.0]])
    d = np.stack([np.sqrt(1 - c * c), np.zeros_like(c), c], axis=1)
    q = G.vmf_pdf(mu[None], np.full(len(c), kappa), d)
    integral = np.trapezoid(2 * np.pi * q, c)
    assert abs(integral - 1) < 2e-4


@pytest.mark.parametrize("kappa", [0.3, 5.0, 200.0])
def test_vmf_sampling_moments(kappa):
    rng = np.random.default_rng(5)
    n = 200000
    mu = rng.normal(size=3)
    mu /= np.linalg.norm(mu)
    d = G.vmf_sample(np.broadcast_to(mu, (n, 3)), np.full(n, kappa), rng.random(n), rng.random(n))
    assert np.allclose(np.linalg.norm(d, axis=1), 1)
    c = d @ mu
    # E[mu.d] = coth(kappa) - 1/kappa
    mean = 1 / np.tanh(kappa) - 1 / kappa
    assert abs(c.mean() - mean) < 5 * c.std() / np.sqrt(n) + 1e-9
    # importance check: E_q[1/q] = area of the sphere
    q = G.vmf_pdf(np.broadcast_to(mu, (n, 3)), np.full(n, kappa), d)
    if kappa < 10:
        est = np.mean(1 / q)
        assert abs(est - 4 * np.pi) < 6 * np.std(1 / q) / np.sqrt(n)


def test_mixture_pdf_integrates_to_one():
    rng = np.random.default_rng(6)
    raw = rng.normal(size=(3, 48))
    raw[:, 3:40:5] = rng.uniform(-3, 2.5, (3, 8))  # kappa up to ~12
    # Fibonacci sphere quadrature
    m = 400000
    i = np.arange(m) + 0.5
    z = 1 - 2 * i / m
    phi = np.pi * (1 + 5 ** 0.5) * i
    r = np.sqrt(1 - z * z)
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    for k in range(3):
        q = G.mixture_pdf(np.broadcast_to(raw[k], (m, 48)), d)
        assert abs(q.mean() * 4 * np.pi - 1) < 1e-3


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    params = rand_params(rng)
    smp = rand_samples(rng, 64)
    loss, g = G.loss_and_grad(params, smp, BOX)
    assert np.isfinite(loss)
    idx = np.concatenate([rng.choice(np.arange(o, o + ni * no), 40, replace=False)
                          for o, _, ni, no in G.offsets()] +
                         [rng.choice(np.arange(b, b + no), 8, replace=False)
                          for _, b, _, no in G.offsets()])
    h = 1e-6
    for i in idx:
        p1 = params.copy()
        p1[i] += h
        p2 = params.copy()
        p2[i] -= h
        fd = (G.loss_and_grad(p1, smp, BOX)[0] - G.loss_and_grad(p2, smp, BOX)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_raw_gradient_matches_finite_differences():
    # direct check of d loss / d raw, including the kappa clamp and the
    # unnormalised mean directions
    rng = np.random.default_rng(8)
    n = 16
    raw = rng.normal(size=(n, 48))
    raw[0, 3] = -5.0   # clamped below: zero kappa gradient
    raw[1, 8] = 9.0    # clamped above
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.exponential(size=n)
    loss, g = G.loss_and_draw(raw, d, w)
    assert g[0, 3] == 0 and g[1, 8] == 0 and np.all(g[:, 40:] == 0)
    h = 1e-6
    for r in range(n):
        for c in range(40):
            a = raw.copy()
            a[r, c] += h
            b = raw.copy()
            b[r, c] -= h
            fd = (G.loss_and_draw(a, d, w)[0] - G.loss_and_draw(b, d, w)[0]) / (2 * h)
            assert abs(fd - g[r, c]) <= 1e-7 + 1e-5 * abs(fd), (r, c, fd, g[r, c])


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(9)
    p = rng.normal(size=100)
    g = rng.normal(size=100)
    p1, m, v = G.adam(p, np.zeros(100), np.zeros(100), g, 1, 1e-3)
    # m_hat = g, v_hat = g^2: step = lr g / (|g| + eps)
    assert np.allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8))
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)
