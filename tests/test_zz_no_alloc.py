"""The allocation-free training step (ajg_guide_train_step_noalloc) against
the allocating one (ajg_guide_train_step): two guides from the same seed
take the same steps, one through each entry point; after every step their
FP32 parameters are bit-identical. Two steps in a row, the second on fewer
samples, so the scratch holds the first step's larger contents when the
second runs."""
import numpy as np
import pytest

from ajisai_render_amd import _capi as capi

pytestmark = pytest.mark.gpu


def _samples(n, seed):
    import torch
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, 4, (n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.exponential(1.0, (n, 1))
    return torch.from_numpy(np.concatenate([pos, nrm, d, w], 1).astype(np.float32)).cuda()


def test_noalloc_train_step_matches_allocating_step(gpu_ctx):
    import torch
    from ajisai_render_amd.core import PathGuide
    box = np.array([0, 0, 0, 4, 4, 4], np.float32)
    ga = PathGuide(gpu_ctx, box, seed=17)
    gb = PathGuide(gpu_ctx, box, seed=17)
    assert np.array_equal(ga.get_params(), gb.get_params())
    for k, n in enumerate((PathGuide.NOALLOC_MAX_SAMPLES, 3000)):
        print(f"case: no-alloc train step {k}, {n} samples", flush=True)
        smp = _samples(n, 100 + k)
        ga.train_step(smp, lr=1e-3)
        gb.train_step_noalloc(smp, lr=1e-3)
        torch.cuda.synchronize()
        pa, pb = ga.get_params(), gb.get_params()
        assert np.isfinite(pa).all()
        assert np.array_equal(pa, pb), np.abs(pa - pb).max()
    ga.close()
    gb.close()


def test_noalloc_train_step_rejects_oversized_batch(gpu_ctx):
    import torch
    from ajisai_render_amd.core import PathGuide
    g = PathGuide(gpu_ctx, np.array([0, 0, 0, 1, 1, 1], np.float32), seed=3)
    big = torch.zeros((PathGuide.NOALLOC_MAX_SAMPLES + 1, 10), dtype=torch.float32, device="cuda")
    with pytest.raises(capi.AjgError, match="at most 32768"):
        g.train_step_noalloc(big)
    g.close()
