"""CPU oracle of the path-guiding network (test infrastructure only).

Restates, in numpy float64, the network of ajisai_render_amd's path guide:
the input encoding, the MLP x(112) -> 64 -> 64 -> 64 -> 48 with ReLU, the
decode of the outputs into an 8-lobe von Mises-Fisher mixture, the weighted
negative log-likelihood loss and its gradient by hand-written backprop.
The reference renderer has no guiding (it samples continuation directions
from the BSDF only: renderer/pt.cpp:215-222), so there is no reference
call site to follow; the oracle is pinned by mathematics instead:
  * the analytic gradient against central finite differences (float64),
  * the vMF density against its closed-form normalisation (integral = 1)
    and against samples drawn by the published inversion of Jakob (2012),
  * the decode against softmax / normalisation identities.
``bf16=True`` mirrors the rounding points of the device kernels (weights,
activations and output gradients rounded to bfloat16, round-to-nearest-even)
so that device results can be compared at accumulation-order tolerance.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline use this.
"""
from __future__ import annotations

import numpy as np

N_IN, N_USED_IN, N_HID, N_OUT, N_USED_OUT, N_LOBES, N_FREQ = 112, 99, 64, 48, 40, 8, 16
KAPPA_LO, KAPPA_HI = -3.0, 7.0

_LAYERS = [(N_IN, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_HID, N_OUT)]


def offsets():
    """[(w_off, b_off, n_in, n_out)] of the flat FP32 parameter vector:
    W1[112][64] b1 W2[64][64] b2 W3[64][64] b3 W4[64][48] b4."""
    out, o = [], 0
    for ni, no in _LAYERS:
        out.append((o, o + ni * no, ni, no))
        o += ni * no + no
    return out


N_PARAMS = offsets()[-1][1] + N_OUT  # 18672


def unpack(params):
    p = np.asarray(params, np.float64)
    return [(p[w:w + ni * no].reshape(ni, no), p[b:b + no]) for w, b, ni, no in offsets()]


def pack(layers):
    return np.concatenate([np.concatenate([W.reshape(-1), b]) for W, b in layers])


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even); returns float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, ((u >> 16) | 0x40) << 16, r)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def encode(pos, nrm, box, f32=True):
    """Input features x[16j + 4g + i] (j = 0..6, g = 0..3, i = 0..3): for
    j < 6, component c = j // 2 and harmonic h = 8 (j % 2) + 2g + i // 2 give
    sin(pi (h+1) n_c) (even i) or cos(pi (h+1) n_c) (odd i); features 96..98
    are p_hat = 2 (p - min) / (max - min) - 1, features 99..111 are zero."""
    box = np.asarray(box, np.float64)
    dt = np.float32 if f32 else np.float64
    pos = np.asarray(pos, dt)
    nrm = np.asarray(nrm, dt)
    n = pos.shape[0]
    x = np.zeros((n, N_IN), np.float64)
    lo = box[:3].astype(dt)
    hi = box[3:].astype(dt)
    x[:, 96:99] = dt(2) * (pos - lo) / (hi - lo) - dt(1)
    pi = dt(3.14159265) if f32 else np.pi
    for j in range(6):
        c = j // 2
        for g in range(4):
            for i in range(4):
                h = 8 * (j % 2) + 2 * g + i // 2
                a = dt(pi * dt(h + 1)) * nrm[:, c]
                x[:, 16 * j + 4 * g + i] = np.sin(a) if i % 2 == 0 else np.cos(a)
    return x


def forward(params, x, bf16_mode=False):
    """Returns (raw outputs n x 48, cache for backward)."""
    layers = unpack(params)
    acts = []
    h = bf16(x) if bf16_mode else np.asarray(x, np.float64)
    for li, (W, b) in enumerate(layers):
        acts.append(h)
        Wq = bf16(W) if bf16_mode else W
        y = h @ Wq + b
        if li < 3:
            y = np.maximum(y, 0.0)
            h = bf16(y) if bf16_mode else y
        else:
            h = y
    return h, acts


def decode(raw):
    """mu (n,8,3), kappa (n,8), pi (n,8) from raw outputs."""
    o = np.asarray(raw, np.float64)[:, :N_USED_OUT].reshape(-1, N_LOBES, 5)
    v = o[..., :3]
    ln = np.maximum(np.linalg.norm(v, axis=-1), 1e-6)
    mu = v / ln[..., None]
    kappa = np.exp(np.clip(o[..., 3], KAPPA_LO, KAPPA_HI))
    a = o[..., 4]
    e = np.exp(a - a.max(axis=1, keepdims=True))
    return mu, kappa, e / e.sum(axis=1, keepdims=True)


def vmf_pdf(mu, kappa, d):
    """kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.d - 1))."""
    md = np.sum(mu * d, axis=-1)
    return kappa / (2 * np.pi * -np.expm1(-2 * kappa)) * np.exp(kappa * (md - 1.0))


def mixture_pdf(raw, d):
    mu, kappa, w = decode(raw)
    return np.sum(w * vmf_pdf(mu, kappa, np.asarray(d, np.float64)[:, None, :]), axis=1)


def vmf_sample(mu, kappa, u1, u2):
    """Jakob (2012), "Numerically stable sampling of the von Mises-Fisher
    distribution on S^2": w = 1 + log(u1 + (1 - u1) e^{-2 kappa}) / kappa."""
    w = 1.0 + np.log(u1 + (1.0 - u1) * np.exp(-2.0 * kappa)) / kappa
    w = np.clip(w, -1.0, 1.0)
    s = np.sqrt(np.maximum(0.0, 1.0 - w * w))
    phi = 2 * np.pi * u2
    local = np.stack([s * np.cos(phi), s * np.sin(phi), w], axis=-1)
    # any orthonormal frame with third axis mu
    a = np.where(np.abs(mu[..., :1]) > 0.9, np.array([0.0, 1.0, 0.0]), np.array([1.0, 0.0, 0.0]))
    t1 = np.cross(a, mu)
    t1 /= np.linalg.norm(t1, axis=-1, keepdims=True)
    t2 = np.cross(mu, t1)
    return local[..., :1] * t1 + local[..., 1:2] * t2 + local[..., 2:3] * mu


def loss_and_draw(raw, dirs, weights):
    """Loss = -sum_k w_k log q(d_k) / sum_k w_k and d loss / d raw (n x 48)."""
    raw = np.asarray(raw, np.float64)
    d = np.asarray(dirs, np.float64)
    wgt = np.asarray(weights, np.float64) / np.sum(weights)
    o = raw[:, :N_USED_OUT].reshape(-1, N_LOBES, 5)
    mu, kappa, pi = decode(raw)
    qj = vmf_pdf(mu, kappa, d[:, None, :])
    q = np.maximum(np.sum(pi * qj, axis=1), 1e-300)
    loss = -np.sum(wgt * np.log(q))
    r = pi * qj / q[:, None]
    g = np.zeros((raw.shape[0], N_LOBES, 5))
    g[..., 4] = r - pi
    md = np.sum(mu * d[:, None, :], axis=-1)
    with np.errstate(over="ignore"):
        dk = 1.0 / kappa - 2.0 / np.expm1(2.0 * kappa) + (md - 1.0)
    inside = (o[..., 3] > KAPPA_LO) & (o[..., 3] < KAPPA_HI)
    g[..., 3] = r * np.where(inside, kappa * dk, 0.0)
    ln = np.maximum(np.linalg.norm(o[..., :3], axis=-1), 1e-6)
    g[..., :3] = (r * kappa / ln)[..., None] * (d[:, None, :] - md[..., None] * mu)
    out = np.zeros_like(raw)
    out[:, :N_USED_OUT] = (-wgt[:, None, None] * g).reshape(-1, N_USED_OUT)
    return loss, out


def backward(params, acts, draw, bf16_mode=False):
    """Gradient of the loss w.r.t. the flat parameter vector."""
    layers = unpack(params)
    grads = [None] * 4
    dy = bf16(draw) if bf16_mode else draw
    for li in range(3, -1, -1):
        W, _ = layers[li]
        X = acts[li]
        grads[li] = (X.T @ dy, dy.sum(axis=0))
        if li > 0:
            Wq = bf16(W) if bf16_mode else W
            dx = (dy @ Wq.T) * (acts[li] > 0)
            dy = bf16(dx) if bf16_mode else dx
    return pack(grads)


def loss_and_grad(params, samples, box, bf16_mode=False):
    """samples: n x 10 (pos, nrm, dir, weight)."""
    s = np.asarray(samples)
    x = encode(s[:, 0:3], s[:, 3:6], box)
    raw, acts = forward(params, x, bf16_mode)
    loss, draw = loss_and_draw(raw, s[:, 6:9], s[:, 9])
    return loss, backward(params, acts, draw, bf16_mode)


def adam(params, m, v, grad, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
