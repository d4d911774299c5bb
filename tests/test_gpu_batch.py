"""GPU: the batched form (many canvases, one pipeline run) against
per-canvas draws and the oracle."""
import numpy as np
import pytest

from oracle import oracle_py as O
from skity_amd import abi, device, presizer, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = device.Context(0)
    yield c
    c.close()


def uneven_scenes(n, seed0, max_paths=None):
    """Canvases of the batched uneven family (scaled): side 256-4096,
    10-1,000 paths of 4-316 line/quad/cubic segments, 30% gradients."""
    out = []
    for i in range(n):
        sc = scenes.generate(5, n_paths=0, seed=seed0 + i)
        if max_paths and sc.n_paths > max_paths:
            sc = sc.slice(0, max_paths)
        out.append(sc)
    return out


def draw_single(ctx, sc, **kw):
    s = device.Surface(ctx, sc.width, sc.height)
    s.clear(0)
    s.draw(sc.scene, **kw)
    img = s.read()
    s.close()
    return img


def draw_batch(ctx, scs, repeat=1, **kw):
    preps = [device.Prepared(ctx, sc.scene) for sc in scs]
    surfs = [device.Surface(ctx, sc.width, sc.height) for sc in scs]
    b = device.Batch(ctx)
    for s, p in zip(surfs, preps):
        b.add(s, p)
    for _ in range(repeat):
        for s in surfs:
            s.clear(0)
        b.draw(**kw)
    imgs = [s.read() for s in surfs]
    st = b.stats()
    b.close()
    for s in surfs:
        s.close()
    for p in preps:
        p.close()
    return imgs, st


def test_batch_equals_individual_draws(ctx):
    scs = uneven_scenes(10, 400, max_paths=120)
    # a mask-filter scene exercises the layer canvases of a batch
    scs.append(scenes.generate(3, n_paths=6, seed=77))
    scs.append(scenes.generate(1, seed=78))
    print("canvases", [(sc.width, sc.n_paths) for sc in scs], flush=True)
    imgs, st = draw_batch(ctx, scs, repeat=2)
    assert st["paths"] == sum(sc.n_paths for sc in scs)
    for i, sc in enumerate(scs):
        want = draw_single(ctx, sc)
        assert np.array_equal(imgs[i], want), f"canvas {i}"


def test_batch_small_canvases_match_oracle(ctx):
    scs = [sc for sc in uneven_scenes(24, 500, max_paths=60) if sc.width <= 512][:4]
    assert scs, "the seeds give small canvases"
    imgs, _ = draw_batch(ctx, scs)
    for i, sc in enumerate(scs):
        want = O.render(sc.scene, sc.width, sc.height)
        d = np.abs(imgs[i].astype(np.int32) - want.astype(np.int32))
        print(f"canvas {i} {sc.width}^2 {sc.n_paths} paths: max diff {d.max()}", flush=True)
        assert d.max() <= 1


def test_batch_learned_and_autotuned_equal_exact(ctx):
    if not __import__("os").path.exists(presizer.WEIGHTS):
        pytest.skip("no trained pre-sizer weights")
    prm, margin = presizer.load()
    presizer.install(ctx, prm, margin)
    scs = uneven_scenes(8, 600, max_paths=150)
    ref, _ = draw_batch(ctx, scs, mode=abi.MODE_EXACT)
    got, st = draw_batch(ctx, scs, mode=abi.MODE_LEARNED)
    print("learned fallbacks", st["fallbacks"], flush=True)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    preps = [device.Prepared(ctx, sc.scene) for sc in scs]
    surfs = [device.Surface(ctx, sc.width, sc.height) for sc in scs]
    b = device.Batch(ctx)
    for s, p in zip(surfs, preps):
        b.add(s, p)
    split = b.autotune(abi.MODE_LEARNED, repeats=1)
    print("autotuned split", split, flush=True)
    for s in surfs:
        s.clear(0)
    b.draw(mode=abi.MODE_LEARNED)
    for s, want in zip(surfs, ref):
        assert np.array_equal(s.read(), want)
    b.close()
    for s in surfs:
        s.close()
    for p in preps:
        p.close()


def test_batch_rejects_a_surface_twice(ctx):
    sc = scenes.generate(1, seed=3)
    p = device.Prepared(ctx, sc.scene)
    s = device.Surface(ctx, sc.width, sc.height)
    b = device.Batch(ctx)
    b.add(s, p)
    with pytest.raises(device.HiprError):
        b.add(s, p)
    b.close()
    s.close()
    p.close()
