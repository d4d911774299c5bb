"""The frames the benchmark times, checked whole against the oracle.

Curves + strokes (4,096 paths, 2048^2) in the exact form and in the learned
form, and gradients + blur (1,024 paths, 4096^2): every pixel of the device
frame against the oracle's frame (rendered on host threads with
orc_render_scene_mt, which is the serial oracle's image).  Curves + strokes
is integer code after the float stroker and flattener, so the expected
difference is 0; gradients carry the 1/255 tolerance of the shading.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from skity_amd import abi, device, presizer

from helpers import diff_report, workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = device.Context(0)
    prm, margin = presizer.load()
    presizer.install(c, prm, margin)
    yield c
    c.close()


@pytest.fixture(scope="module")
def curves():
    sc = workload(2)
    print(f"curves + strokes: oracle over {sc.n_paths} paths", flush=True)
    return sc, O.render_mt(sc.scene, sc.width, sc.height)


@pytest.mark.parametrize("mode", [abi.MODE_EXACT, abi.MODE_LEARNED])
def test_curves_strokes_full_frame(ctx, curves, mode):
    sc, want = curves
    prep = device.Prepared(ctx, sc.scene)
    surf = device.Surface(ctx, sc.width, sc.height)
    surf.draw(prep, mode=mode)
    got = surf.read()
    st = surf.stats()
    surf.close()
    prep.close()
    mx, cnt = diff_report(got, want)
    print(f"curves + strokes, mode {mode}: {sc.n_paths} paths, {st['blends']} blends, "
          f"max |diff| {mx}, differing pixels {cnt}, fallbacks {st['fallbacks']}", flush=True)
    assert cnt == 0


def test_gradients_blur_full_frame(ctx):
    sc = workload(3)
    want = O.render_mt(sc.scene, sc.width, sc.height)
    surf = device.Surface(ctx, sc.width, sc.height)
    surf.draw(sc.scene, mode=abi.MODE_EXACT)
    got = surf.read()
    surf.close()
    mx, cnt = diff_report(got, want)
    print(f"gradients + blur: {sc.n_paths} paths 4096^2, max |diff| {mx}, "
          f"pixels differing {cnt}", flush=True)
    assert mx <= 1
