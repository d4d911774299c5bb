"""GPU checks of the learned pre-sizer: the MFMA forward and backward
kernels against the float64 restatement (oracle/mlp_ref.py), training on the
device, and the learned mode drawing exactly what the exact mode draws.

Tolerances: operands are FP16 (weights, activations and back-propagated
deltas), accumulation FP32.  The oracle rounds weights and activations to
FP16 as the device does (not the deltas), so gradients agree to ~1e-3
relative; the bar is 2e-2 relative L2 error on the gradient vector and
2e-2 absolute (normalised units) on predictions.
"""
import numpy as np
import pytest

from oracle import mlp_ref as R
from skity_amd import abi, device, presizer

from helpers import workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = device.Context(0)
    yield c
    c.close()


def _random_problem(n, seed):
    rng = np.random.default_rng(seed)
    feats = rng.normal(2.0, 1.5, (n, 16)).astype(np.float32)
    feats[:, 14:] = 1.0
    targets = np.stack([np.log1p(np.abs(feats[:, :4]).sum(1) * 40),
                        np.log1p(np.abs(feats[:, 5] * feats[:, 6]) * 30)], 1).astype(np.float32)
    prm = presizer.init_params(feats, targets, seed=seed)
    return prm, feats, targets


def test_forward_matches_float64(ctx):
    prm, feats, targets = _random_problem(777, 1)
    got = presizer.predict(ctx, prm, feats)
    want = R.predict(prm, feats, half=True)
    ysig = prm[R.YSIG:R.TOTAL]
    err = np.abs(got - want) / ysig
    print(f"forward: max normalised error {err.max():.2e}", flush=True)
    assert err.max() < 2e-2


def test_gradients_match_float64(ctx):
    prm, feats, targets = _random_problem(1000, 2)
    loss, g = presizer.gradients(ctx, prm, feats, targets)
    want_loss, want_g = R.loss_and_grad(prm, feats, targets, half=True)
    rel = np.linalg.norm(g - want_g) / np.linalg.norm(want_g)
    print(f"gradients: loss {loss:.5f} vs {want_loss:.5f}, relative L2 error {rel:.2e}",
          flush=True)
    assert abs(loss - want_loss) <= 1e-2 * want_loss
    assert rel < 2e-2
    # per block, so a wrong layout in one layer cannot hide behind the others
    for lo, hi in ((R.W1, R.B1), (R.W2, R.B2), (R.W3, R.B3)):
        rb = np.linalg.norm(g[lo:hi] - want_g[lo:hi]) / np.linalg.norm(want_g[lo:hi])
        assert rb < 3e-2, (lo, rb)


def test_training_on_device_reduces_loss(ctx):
    prm, feats, targets = _random_problem(8192, 3)
    loss0, _ = presizer.gradients(ctx, prm, feats, targets)
    trained, curve = presizer.train(ctx, prm, feats, targets, steps=600, batch=1024, lr=3e-3)
    loss1, _ = presizer.gradients(ctx, trained, feats, targets)
    print(f"training: loss {loss0:.4f} -> {loss1:.4f}", flush=True)
    assert np.isfinite(curve).all()
    assert loss1 < 0.2 * loss0


def test_learned_mode_draws_what_exact_mode_draws(ctx):
    """Train on two small scenes, install with a 0.99-quantile margin and draw
    a third scene: pixels equal the exact mode's, whatever the fallbacks."""
    train_sc = [workload(2, 96, seed=31), workload(2, 96, seed=32)]
    feats, targets = presizer.collect(ctx, train_sc)
    prm = presizer.init_params(feats, targets)
    prm, _ = presizer.train(ctx, prm, feats, targets, steps=400, batch=256, lr=3e-3)
    m = presizer.margins(presizer.predict(ctx, prm, feats), targets, 0.99)
    presizer.install(ctx, prm, m)
    sc = workload(2, 64, seed=33)
    surf = device.Surface(ctx, sc.width, sc.height)
    surf.draw(sc.scene, mode=abi.MODE_EXACT)
    ref = surf.read()
    surf.clear(0)
    surf.draw(sc.scene, mode=abi.MODE_LEARNED)
    got = surf.read()
    st = surf.stats()
    surf.close()
    print(f"learned mode: {st['fallbacks']} fallbacks of {st['raster_paths']} jobs", flush=True)
    assert np.array_equal(got, ref)
