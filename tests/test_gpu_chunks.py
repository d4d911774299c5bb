"""Draws larger than one pipeline run's memory budget run as consecutive
sub-plans of whole draws; the targets must see exactly the same blends.

The budget is lowered here (chunk_points) so that small scenes split into
many sub-plans: curves + strokes in the exact and learned forms, and a batch
of outline canvases of the uneven family (workload 4), each against the same
draw in one run.
"""
import numpy as np
import pytest

from skity_amd import abi, device, presizer, scenes

from helpers import workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = device.Context(0)
    presizer.install(c, *presizer.load())
    yield c
    c.close()


@pytest.mark.parametrize("mode", [abi.MODE_EXACT, abi.MODE_LEARNED])
def test_chunked_scene_equals_one_run(ctx, mode):
    sc = workload(2, 200)
    prep = device.Prepared(ctx, sc.scene)
    out = []
    for chunk in (0, 3000):
        surf = device.Surface(ctx, sc.width, sc.height)
        surf.draw(prep, mode=mode, chunk_points=chunk)
        out.append((surf.read(), surf.stats(), surf.stage_times()))
        surf.close()
    prep.close()
    (a, sa, _), (b, sb, tb) = out
    print(f"mode {mode}: one run {sa['edges']} edges; chunked {sb['edges']} edges, "
          f"{tb['walk'][1]} walk launches", flush=True)
    assert np.array_equal(a, b)
    assert sa["edges"] == sb["edges"] and sa["blends"] == sb["blends"]


def test_chunked_batch_equals_per_canvas_draws(ctx):
    scs = [scenes.generate(4, n_paths=40 + 30 * i, seed=scenes.SEED_BASE + 900 + i)
           for i in range(4)]
    preps = [device.Prepared(ctx, sc.scene) for sc in scs]
    want = []
    for sc, pr in zip(scs, preps):
        s = device.Surface(ctx, sc.width, sc.height)
        s.draw(pr)
        want.append(s.read())
        s.close()
    surfs = [device.Surface(ctx, sc.width, sc.height) for sc in scs]
    batch = device.Batch(ctx)
    for su, pr in zip(surfs, preps):
        batch.add(su, pr)
    batch.draw(mode=abi.MODE_LEARNED, chunk_points=20000)
    for i, su in enumerate(surfs):
        assert np.array_equal(su.read(), want[i]), f"canvas {i}"
    print(f"batch of {len(scs)} outline canvases in sub-plans: equal", flush=True)
    batch.close()
    for su in surfs:
        su.close()
    for pr in preps:
        pr.close()
