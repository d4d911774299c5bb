"""GPU parity of colour filters (blend, matrix, the two sRGB filters,
compose; ColorFilter::FilterColor, src/effect/color_filter.cc:113-185)
applied after the cover and before the blend (SWSpanBrush::BrushH,
src/render/sw/sw_span_brush.cc:106-137), on fills, strokes, gradients,
blurred draws and DrawPaint, with several blend modes, on premultiplied and
unpremultiplied, RGBA and BGRA targets, against the oracle (whose filters
tests/test_oracle_color_filter.py pins)."""
import numpy as np
import pytest

from oracle import oracle_py as O
from skity_amd import abi, device, presizer

from helpers import diff_report
from scene_kit import filter_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = device.Context(0)
    presizer.install(c, *presizer.load())
    yield c
    c.close()


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("alpha,color", [(abi.ALPHA_PREMUL, abi.COLOR_RGBA),
                                         (abi.ALPHA_UNPREMUL, abi.COLOR_BGRA)])
@pytest.mark.parametrize("mode", [abi.MODE_EXACT, abi.MODE_LEARNED])
def test_color_filters_match_oracle(ctx, seed, alpha, color, mode):
    sc = filter_scene(seed)
    surf = device.Surface(ctx, sc.width, sc.height, alpha_type=alpha, color_type=color)
    surf.draw(sc.scene, mode=mode)
    got = surf.read()
    surf.close()
    want = O.render_ex(sc.scene, sc.width, sc.height, alpha, color)
    mx, cnt = diff_report(got, want)
    print(f"filter scene {seed}, target {alpha}/{color}, mode {mode}: max diff {mx}, "
          f"{cnt} pixels differ", flush=True)
    # gradients and blurred layers: within 1/255
    assert mx <= 1
