"""GPU parity of paths whose rows keep more than 1024 edges active.

The reference's WalkEdges keeps a linked active list of any length
(src/render/sw/sw_raster.cc:546-677).  Here such paths take the 32-slot
(2048 edges) and 64-slot (4096 edges) walks; a longer list is refused with
HIPR_ERROR_CAPACITY, and the context stays usable afterwards.
"""
import numpy as np
import pytest

from oracle import oracle_py as O
from skity_amd import abi, device, presizer
from skity_amd.canvas import Canvas, Paint, Path

from helpers import diff_report

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = device.Context(0)
    prm, margin = presizer.load()
    presizer.install(c, prm, margin)
    yield c
    c.close()


def comb_scene(teeth, fill, extra=True):
    """A closed comb: each tooth is a rising and a falling edge, so the rows
    between y = 17 and 183 cross about 2 * teeth edges."""
    pitch = 2.6
    w = int(teeth * pitch) + 24
    c = Canvas(w, 200)
    p = Path().move_to(0.5, 190.25)
    for i in range(teeth):
        x = 2.3 + i * pitch
        p.line_to(x, 10.6 + (i % 7))
        p.line_to(x + 1.3, 190.25 - (i % 5))
    p.close()
    p.fill_type = fill
    c.draw_path(p, Paint((0.2, 0.5, 0.9, 0.9)))
    if extra:
        # narrow paths in the same frame take the 4-slot walk beside it
        for k in range(6):
            c.draw_path(Path().add_circle(40 + 150 * k, 100, 30 + 3 * k),
                        Paint((0.9, 0.3 + 0.1 * k, 0.1, 0.6)))
    return c.scene()


def draw(ctx, sc, mode):
    surf = device.Surface(ctx, sc.width, sc.height)
    try:
        surf.draw(sc.scene, mode=mode)
        return surf.read(), surf.stats()
    finally:
        surf.close()


@pytest.mark.parametrize("mode", [abi.MODE_EXACT, abi.MODE_LEARNED])
@pytest.mark.parametrize("fill", [abi.FILL_WINDING, abi.FILL_EVEN_ODD])
@pytest.mark.parametrize("teeth,tier_min", [(700, 1024), (1500, 2048)])
def test_wide_active_lists_match_oracle(ctx, teeth, tier_min, fill, mode):
    sc = comb_scene(teeth, fill)
    got, st = draw(ctx, sc, mode)
    want = O.render(sc.scene, sc.width, sc.height)
    mx, cnt = diff_report(got, want)
    print(f"comb of {teeth} teeth, fill {fill}, mode {mode}: max active {st['max_active']}, "
          f"max diff {mx}, differing pixels {cnt}", flush=True)
    assert st["max_active"] > tier_min
    assert mx == 0 and cnt == 0


def test_active_list_beyond_limit_is_refused(ctx):
    sc = comb_scene(2100, abi.FILL_WINDING, extra=False)
    surf = device.Surface(ctx, sc.width, sc.height)
    with pytest.raises(device.HiprError) as e:
        surf.draw(sc.scene)
    surf.close()
    assert e.value.code == abi.ERROR_CAPACITY
    # the context still draws
    ok = comb_scene(100, abi.FILL_WINDING, extra=False)
    got, st = draw(ctx, ok, abi.MODE_EXACT)
    want = O.render(ok.scene, ok.width, ok.height)
    assert np.array_equal(got, want)
