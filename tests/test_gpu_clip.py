"""GPU parity of rect clips (f2: Canvas::ClipRect narrowing the global clip
bounds, canvas.cc:130-135,463-476, the scan clip of RastePath,
sw_raster.cc:731-786, edges outside it dropped by SWEdge::CanBeIgnored,
sw_edge.cc:72-88): nested Save / ClipRect / Restore with integer,
fractional, off-canvas and empty clips under scale + translate matrices,
over fills, strokes, gradients, blurred draws, pixmap shaders and DrawPaint,
in the exact and learned forms and batched, against the oracle (whose clip
tests/test_oracle_clip.py pins)."""
import numpy as np
import pytest

from oracle import oracle_py as O
from skity_amd import abi, device, presizer

from helpers import diff_report
from scene_kit import clip_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = device.Context(0)
    presizer.install(c, *presizer.load())
    yield c
    c.close()


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
@pytest.mark.parametrize("mode", [abi.MODE_EXACT, abi.MODE_LEARNED])
def test_clipped_scenes_match_oracle(ctx, seed, mode):
    sc = clip_scene(seed)
    assert sc.scene.contents.clip_rects
    surf = device.Surface(ctx, sc.width, sc.height)
    surf.clear(0)
    surf.draw(sc.scene, mode=mode)
    got = surf.read()
    surf.close()
    want = O.render(sc.scene, sc.width, sc.height)
    mx, cnt = diff_report(got, want)
    print(f"clip scene {seed}, mode {mode}: max diff {mx}, {cnt} pixels differ", flush=True)
    # blurred layers: within 1/255
    assert mx <= 1


def test_clipped_batch_matches_single(ctx):
    scenes = [clip_scene(s, w=128 + 32 * s, h=96) for s in (5, 6, 7)]
    singles = []
    for sc in scenes:
        surf = device.Surface(ctx, sc.width, sc.height)
        surf.clear(0)
        surf.draw(sc.scene)
        singles.append(surf.read())
        surf.close()
    preps = [device.Prepared(ctx, sc.scene) for sc in scenes]
    surfs = [device.Surface(ctx, sc.width, sc.height) for sc in scenes]
    b = device.Batch(ctx)
    for s, p in zip(surfs, preps):
        s.clear(0)
        b.add(s, p)
    b.draw(mode=abi.MODE_LEARNED)
    for k, (s, want) in enumerate(zip(surfs, singles)):
        mx, _ = diff_report(s.read(), want)
        print(f"batched clip canvas {k}: max diff {mx}", flush=True)
        assert mx == 0
        s.close()
    b.close()
    for p in preps:
        p.close()


def test_clipped_scene_in_sub_plans(ctx):
    """A clipped frame split into sub-plans (a small chunk budget) equals the
    frame drawn in one run: the jobs carry their clips into every chunk."""
    sc = clip_scene(9, w=320, h=240, n=80)
    imgs = []
    for chunk in (0, 2048):
        surf = device.Surface(ctx, sc.width, sc.height)
        surf.clear(0)
        surf.draw(sc.scene, mode=abi.MODE_LEARNED, chunk_points=chunk)
        imgs.append(surf.read())
        surf.close()
    want = O.render(sc.scene, sc.width, sc.height)
    for k, got in enumerate(imgs):
        mx, _ = diff_report(got, want)
        print(f"clipped scene, chunk budget {(0, 2048)[k]}: max diff {mx}", flush=True)
        assert mx <= 1
    assert np.array_equal(imgs[0], imgs[1])
