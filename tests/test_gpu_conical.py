"""GPU parity of two-point conical gradients (ConicalGradientColorBrush,
src/render/sw/sw_span_brush.cc:385-506) against the oracle, whose conical
brush tests/test_oracle_conical.py pins to the gradient's definition:
nested circles, cones, end radius 0, equal radii, concentric circles, every
tile mode, rotated local matrices and canvas matrices; fills and strokes.
Gradient colours are float code: within 1/255."""
import numpy as np
import pytest

from oracle import oracle_py as O
from skity_amd import abi, device, presizer

from helpers import diff_report
from scene_kit import conical_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = device.Context(0)
    presizer.install(c, *presizer.load())
    yield c
    c.close()


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("mode", [abi.MODE_EXACT, abi.MODE_LEARNED])
def test_conical_gradients_match_oracle(ctx, seed, mode):
    sc = conical_scene(seed)
    surf = device.Surface(ctx, sc.width, sc.height)
    surf.draw(sc.scene, mode=mode)
    got = surf.read()
    surf.close()
    want = O.render(sc.scene, sc.width, sc.height)
    mx, cnt = diff_report(got, want)
    print(f"conical scene {seed}, mode {mode}: {int((want[..., 3] > 0).sum())} pixels drawn, "
          f"max diff {mx}, {cnt} differ", flush=True)
    assert mx <= 1
