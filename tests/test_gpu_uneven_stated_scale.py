"""The batched uneven family at its stated scale (workload 4: canvases of side
256-4096 with 100-50,000 closed outlines of 4-10,000 line/quad/cubic
segments, 30% gradient paints) through the batched form, against the oracle.

The eight canvases are the first eight of the benchmark's batched leg
(seeds SEED_BASE + 400 + i), drawn whole: one batch in the learned form
against the oracle's full frames, and against per-canvas exact draws.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from skity_amd import abi, device, presizer, scenes

from helpers import diff_report

pytestmark = pytest.mark.gpu

SEEDS = [scenes.SEED_BASE + 400 + i for i in range(8)]


@pytest.fixture(scope="module")
def ctx():
    c = device.Context(0)
    presizer.install(c, *presizer.load())
    yield c
    c.close()


@pytest.fixture(scope="module")
def canvases():
    return [scenes.generate(4, n_paths=0, seed=s) for s in SEEDS]


def test_stated_scale_batch_matches_oracle_and_exact(ctx, canvases):
    preps = [device.Prepared(ctx, sc.scene) for sc in canvases]
    surfs = [device.Surface(ctx, sc.width, sc.height) for sc in canvases]
    batch = device.Batch(ctx)
    for su, pr in zip(surfs, preps):
        batch.add(su, pr)
    batch.draw(mode=abi.MODE_LEARNED)
    st = batch.stats()
    got = [su.read() for su in surfs]
    batch.close()
    n_paths = sum(sc.n_paths for sc in canvases)
    print(f"{len(canvases)} canvases, {n_paths} paths, {st['edges']} edges, "
          f"max active {st['max_active']}, fallbacks {st['fallbacks']}", flush=True)
    worst = 0
    for i, (sc, pr, su) in enumerate(zip(canvases, preps, surfs)):
        su.clear(0)
        su.draw(pr, mode=abi.MODE_EXACT)
        exact = su.read()
        assert np.array_equal(got[i], exact), f"canvas {i}: batched learned != exact"
        want = O.render_mt(sc.scene, sc.width, sc.height)
        mx, cnt = diff_report(got[i], want)
        print(f"canvas {i}: {sc.width}^2, {sc.n_paths} paths: max |diff| {mx}, "
              f"{cnt} pixels differ", flush=True)
        # solid fills are bit-exact; gradients within 1/255
        assert mx <= 1, f"canvas {i}"
        worst = max(worst, mx)
        su.close()
        pr.close()
    assert n_paths > 40000
