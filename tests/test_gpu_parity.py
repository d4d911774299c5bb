"""GPU parity: the HIP pipeline against the CPU oracle on seeded scenes.

Bars (from the task statement): bit-exact edge records and segment counts;
pixels within 1/255 per channel (the walk and the brushes are integer code,
so the expected difference is 0 for solid fills); StackBlur bit-exact.
"""
import sys

import numpy as np
import pytest

from oracle import oracle_py as O
from skity_amd import abi, device

from helpers import diff_report, fill_only, no_blur, workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = device.Context(0)
    yield c
    c.close()


def render_gpu(ctx, sc, mode=0):
    surf = device.Surface(ctx, sc.width, sc.height)
    surf.draw(sc.scene, mode=mode)
    img = surf.read()
    st = surf.stats()
    surf.close()
    return img, st


@pytest.mark.parametrize("wl,n", [(1, 0), (2, 48)])
def test_edges_bit_exact(ctx, wl, n):
    sc = fill_only(workload(wl, n))
    offs, edges = ctx.build_edges(sc.scene, sc.n_paths)
    for p in range(sc.n_paths):
        want = O.path_edges(sc.scene, p)
        got = edges[offs[p]:offs[p + 1]]
        assert got.shape == want.shape, f"path {p}: {got.shape[0]} edges vs {want.shape[0]}"
        bad = np.nonzero((got != want).any(1))[0]
        assert bad.size == 0, f"path {p}: first differing edge {bad[0]}: {got[bad[0]]} vs {want[bad[0]]}"


def test_plain_fill_800x600(ctx):
    sc = workload(1)
    got, st = render_gpu(ctx, sc)
    want = O.render(sc.scene, sc.width, sc.height)
    mx, cnt = diff_report(got, want)
    print(f"workload 1: max diff {mx}, differing pixels {cnt}, stats {st}", flush=True)
    assert mx <= 1
    assert cnt == 0, "solid fills are integer code end to end: expected bit-exact"


@pytest.mark.parametrize("seed", [11, 12])
def test_curve_fills_2048(ctx, seed):
    sc = fill_only(workload(2, 24, seed=seed))
    got, st = render_gpu(ctx, sc)
    want = O.render(sc.scene, sc.width, sc.height)
    mx, cnt = diff_report(got, want)
    print(f"curves seed {seed}: max diff {mx}, differing pixels {cnt}, stats {st}", flush=True)
    assert mx <= 1
    assert cnt == 0


def test_stackblur_bit_exact(ctx):
    rng = np.random.default_rng(5)
    for (h, w, r) in [(37, 53, 2), (64, 80, 7), (120, 33, 25), (50, 300, 55)]:
        img = rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8)
        img[..., :3] = np.minimum(img[..., :3], img[..., 3:4])
        img[0] = 0  # a mask-filter layer's top row is transparent
        want = O.stackblur(img, r)
        got = ctx.stackblur(img, r)
        assert np.array_equal(got, want), f"{h}x{w} r={r}"


def test_gradients_no_blur(ctx):
    sc = no_blur(workload(3, 24))
    got, st = render_gpu(ctx, sc)
    want = O.render(sc.scene, sc.width, sc.height)
    mx, cnt = diff_report(got, want)
    print(f"gradients: max diff {mx}, differing pixels {cnt}", flush=True)
    assert mx <= 1


def test_gradients_with_blur(ctx):
    sc = workload(3, 12)
    got, st = render_gpu(ctx, sc)
    want = O.render(sc.scene, sc.width, sc.height)
    mx, cnt = diff_report(got, want)
    print(f"gradients+blur: max diff {mx}, differing pixels {cnt}", flush=True)
    assert mx <= 1


@pytest.mark.parametrize("fill", [abi.FILL_WINDING, abi.FILL_EVEN_ODD])
def test_wide_active_edge_list(ctx, fill):
    """A comb whose rows cross 600 edges: more than the 4-slot (256-edge)
    register list holds, so the 16-slot walk takes the job."""
    from skity_amd.canvas import Canvas, Paint, Path
    c = Canvas(1220, 240)
    p = Path().move_to(0.5, 220.25)
    for i in range(300):
        x = 2.3 + i * 4
        p.line_to(x, 10.6 + (i % 7))
        p.line_to(x + 2.1, 220.25 - (i % 5))
    p.close()
    p.fill_type = fill
    c.draw_path(p, Paint((0.2, 0.5, 0.9, 0.9)))
    sc = c.scene()
    got, st = render_gpu(ctx, sc)
    want = O.render(sc.scene, sc.width, sc.height)
    mx, cnt = diff_report(got, want)
    print(f"comb: max active {st['max_active']}, max diff {mx}, differing pixels {cnt}", flush=True)
    assert st["max_active"] > 256
    assert mx <= 1 and cnt == 0
