"""The estimate modes' partial flatten redo: jobs whose edges outgrew their
estimated range are flattened again into exact space appended after the
others, while the other jobs' edges stay where they were written, copied
across the reallocation when the edge buffer has to grow.  On a fresh
context (no edge buffer yet) with exact estimates for half the paths and 1
for the rest, the buffer must grow for the redo; the pixels must equal the
exact form's and the oracle's, and the flatten fallbacks must be exactly the
under-estimated paths."""
import numpy as np
import pytest

from oracle import oracle_py as O
from skity_amd import abi, device

from helpers import fill_only, workload

pytestmark = pytest.mark.gpu


def test_partial_flatten_redo_keeps_other_edges():
    sc = fill_only(workload(2, 96, seed=31))
    n = sc.n_paths
    probe = device.Context(0)
    surf = device.Surface(probe, sc.width, sc.height)
    surf.draw(sc.scene, mode=abi.MODE_EXACT)
    ref = surf.read()
    path, kind, _ = surf.jobs()
    assert np.array_equal(path, np.arange(n)) and not kind.any()
    exact = np.array([surf.job_edges(j).shape[0] for j in range(n)], np.int32)
    surf.close()
    probe.close()

    est = exact.copy()
    short = np.arange(n) % 2 == 1
    est[short] = 1
    expect_redo = int((exact[short] > 1).sum())
    # the redo appends sum(exact[short]) edges after sum(est): more than the
    # first allocation (the estimated total + 25%) holds
    assert est.sum() + exact[short].sum() > 1.25 * est.sum() + 8

    ctx = device.Context(0)
    try:
        surf = device.Surface(ctx, sc.width, sc.height)
        surf.draw(sc.scene, mode=abi.MODE_ESTIMATES, edge_estimates=est,
                  tile_estimates=np.full(n, 2 * 16384, np.int32))
        got = surf.read()
        st = surf.stats()
        edges_after = np.array([surf.job_edges(j).shape[0] for j in range(n)], np.int32)
        surf.close()
    finally:
        ctx.close()
    want = O.render(sc.scene, sc.width, sc.height)
    print(f"{n} paths, {expect_redo} under-estimated: fallbacks {st['fallbacks']}, "
          f"under-estimates {st['underestimates']}", flush=True)
    assert np.array_equal(got, ref)
    assert np.array_equal(got, want)
    assert np.array_equal(edges_after, exact)
    assert st["fallbacks"] == expect_redo
    assert st["underestimates"] == 0
