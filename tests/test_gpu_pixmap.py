"""GPU parity of pixmap shaders (PixmapBrush + BitmapSampler,
src/render/sw/sw_span_brush.cc:555-579, src/graphic/bitmap_sampler.cc:
26-108) on fills, strokes, blurred draws and DrawPaint: nearest and linear
sampling, every pair of tile modes, local matrices that scale and rotate,
canvas matrices, premultiplied and unpremultiplied images, colour filters
and other blend modes, on premultiplied and unpremultiplied, RGBA and BGRA
targets, against the oracle (whose sampler tests/test_oracle_pixmap.py
pins)."""
import numpy as np
import pytest

from oracle import oracle_py as O
from skity_amd import abi, device, presizer
from skity_amd.canvas import Canvas, Paint, Path

from helpers import diff_report
from scene_kit import pixmap_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = device.Context(0)
    presizer.install(c, *presizer.load())
    yield c
    c.close()


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("alpha,color", [(abi.ALPHA_PREMUL, abi.COLOR_RGBA),
                                         (abi.ALPHA_UNPREMUL, abi.COLOR_BGRA)])
@pytest.mark.parametrize("mode", [abi.MODE_EXACT, abi.MODE_LEARNED])
def test_pixmap_shaders_match_oracle(ctx, seed, alpha, color, mode):
    sc = pixmap_scene(seed)
    surf = device.Surface(ctx, sc.width, sc.height, alpha_type=alpha, color_type=color)
    surf.draw(sc.scene, mode=mode)
    got = surf.read()
    surf.close()
    want = O.render_ex(sc.scene, sc.width, sc.height, alpha, color)
    mx, cnt = diff_report(got, want)
    print(f"pixmap scene {seed}, target {alpha}/{color}, mode {mode}: max diff {mx}, "
          f"{cnt} pixels differ", flush=True)
    # blurred layers: within 1/255; everything else is exact (below)
    assert mx <= 1


def test_unblurred_pixmap_draws_are_exact(ctx):
    """Without blurs the device and the oracle evaluate the same float
    expressions in the same order: bit-exact."""
    rng = np.random.default_rng(7)
    cv = Canvas(160, 120)
    for i in range(24):
        img = rng.integers(0, 256, (int(rng.integers(1, 30)), int(rng.integers(1, 30)), 4))
        img = img.astype(np.uint8)
        at = abi.ALPHA_UNPREMUL if i % 2 else abi.ALPHA_PREMUL
        if at == abi.ALPHA_PREMUL:
            img[..., :3] = np.minimum(img[..., :3], img[..., 3:])
        cx, cy = float(rng.uniform(10, 150)), float(rng.uniform(10, 110))
        th = float(rng.uniform(-3, 3))
        s = float(rng.uniform(0.3, 3))
        local = (s * np.cos(th), -s * np.sin(th), cx, s * np.sin(th), s * np.cos(th), cy)
        p = Paint((1, 1, 1, float(rng.uniform(0.4, 1)))).set_image_shader(
            img, at, i % 2, i % 4, (i // 4) % 4, local)
        if i % 3 == 2:
            p.set_style(abi.STYLE_STROKE)
            p.set_stroke_width(float(rng.uniform(1, 6)))
        r = float(rng.uniform(8, 40))
        cv.draw_path(Path().add_rect(cx - r, cy - r, cx + r, cy + 0.7 * r), p)
    sc = cv.scene()
    surf = device.Surface(ctx, sc.width, sc.height)
    surf.draw(sc.scene)
    got = surf.read()
    surf.close()
    want = O.render(sc.scene, sc.width, sc.height)
    mx, cnt = diff_report(got, want)
    print(f"unblurred pixmap draws: max diff {mx}, {cnt} pixels differ", flush=True)
    assert mx == 0


def test_pixmap_batched_matches_single(ctx):
    """Canvases with pixmap shaders drawn as one batch (images concatenated
    across canvases) equal the same canvases drawn one at a time."""
    scenes = [pixmap_scene(s, w=96 + 16 * s, h=80) for s in (4, 5, 6)]
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
    b.draw()
    for k, (s, want) in enumerate(zip(surfs, singles)):
        got = s.read()
        s.close()
        mx, cnt = diff_report(got, want)
        print(f"batched pixmap canvas {k}: max diff {mx}", flush=True)
        assert mx == 0
    b.close()
    for p in preps:
        p.close()
