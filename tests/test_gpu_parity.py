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


@pytest.mark.parametrize("h,w,r", [(37, 53, 2), (64, 80, 7), (120, 33, 25), (50, 300, 55),
                                   (1, 1, 3), (3, 700, 16), (700, 5, 17), (130, 129, 128),
                                   (300, 260, 254), (64, 64, 254)])
def test_stackblur_mfma_bit_exact(ctx, h, w, r):
    """The MFMA banded-product StackBlur equals the oracle byte for byte:
    fp16 holds every pixel and weight exactly and every fp32 partial sum is
    an integer below 2^24 (255 (r+1)^2 at r = 254), so no rounding occurs."""
    rng = np.random.default_rng(1000 + r)
    img = rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8)
    img[..., 3] = np.maximum(img[..., 3], 250) if r % 2 else img[..., 3]
    want = O.stackblur(img, r)
    got = ctx.stackblur(img, r, impl=abi.BLUR_IMPL_MFMA)
    bad = int(np.count_nonzero(np.any(got != want, axis=-1)))
    assert bad == 0, f"{h}x{w} r={r}: {bad} pixels differ"


def test_stackblur_mfma_saturated(ctx):
    """All-255 input at the largest radius: the largest sums the kernel sees."""
    img = np.full((90, 600, 4), 255, np.uint8)
    want = O.stackblur(img, 254)
    got = ctx.stackblur(img, 254, impl=abi.BLUR_IMPL_MFMA)
    assert np.array_equal(got, want)


def test_blur_impls_render_identically(ctx):
    sc = workload(3, 12)
    surfs = []
    for impl in (abi.BLUR_IMPL_MFMA, abi.BLUR_IMPL_LINES):
        s = device.Surface(ctx, sc.width, sc.height)
        s.clear(0)
        s.draw(sc.scene, blur_impl=impl)
        surfs.append(s.read())
        s.close()
    assert np.array_equal(surfs[0], surfs[1])


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
