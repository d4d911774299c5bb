"""CPU checks of the C-ABI libraries: they load and export every entry point
declared in include/skity_hip.h and include/skity_hip_picture.h (no device
calls)."""
import ctypes
import os
import re

from skity_amd import abi

HEADER = os.path.join(abi.ROOT, "include", "skity_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"HIPR_API\s+[\w\s\*]+?\b(hipr_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "hipr_surface_draw_scene" in syms
    assert len(syms) >= 10


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(abi.HERE, "libskity_hip.so"))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"missing exports: {missing}"


def test_picture_binding_exports_every_declared_symbol():
    text = open(os.path.join(abi.ROOT, "include", "skity_hip_picture.h")).read()
    syms = sorted(set(re.findall(r"\b(skity_hip_\w+)\s*\(", text)))
    assert len(syms) == 8
    lib = ctypes.CDLL(os.path.join(abi.HERE, "libskity_hip_canvas.so"))
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, f"missing exports: {missing}"


def test_struct_sizes_match_header_layout():
    # hipr_paint: 8 floats + 9 x 4-byte fields
    assert ctypes.sizeof(abi.Paint) == 68
    assert ctypes.sizeof(abi.Gradient) == 4 * (4 + 4 + 2 + 6 + 64 + 16)
    assert ctypes.sizeof(abi.Stats) == 7 * 8 + 8 + 3 * 8


def test_scene_generator_matches_header_and_workload_spec():
    """The generator library is refused when built from another header; the
    curves + strokes workload has 4096 paths, about 20% stroked at 1-16 px."""
    import numpy as np
    from skity_amd import scenes
    sc = scenes.generate(2)
    s = sc.scene.contents
    assert s.n_paths == 4096 and (sc.width, sc.height) == (2048, 2048)
    style = np.array([s.paints[p].style for p in range(s.n_paths)])
    width = np.array([s.paints[p].stroke_width for p in range(s.n_paths)])
    assert set(np.unique(style)) <= {abi.STYLE_FILL, abi.STYLE_STROKE}
    frac = (style == abi.STYLE_STROKE).mean()
    assert 0.17 < frac < 0.23
    assert width[style == abi.STYLE_STROKE].min() >= 1.0
    assert width[style == abi.STYLE_STROKE].max() <= 16.0


def test_uneven_generator_spans_the_family():
    """Workload 5 (scaled batched uneven family): square canvases of side
    2^8..2^12, 10-1,000 paths, segment counts 4-316 with lines, quads and
    cubics at about 40/30/30 %, about 30% gradient paints."""
    import numpy as np
    from skity_amd import scenes
    sides, paths, grads, verbs = set(), [], 0, np.zeros(8, np.int64)
    for seed in range(40):
        sc = scenes.generate(5, seed=1000 + seed)
        s = sc.scene.contents
        assert sc.width == sc.height and sc.width in (256, 512, 1024, 2048, 4096)
        sides.add(sc.width)
        paths.append(s.n_paths)
        assert 10 <= s.n_paths <= 1000
        n_v = s.verb_offsets[s.n_paths]
        v = np.ctypeslib.as_array(s.verbs, shape=(n_v,))
        verbs += np.bincount(v, minlength=8)[:8]
        grads += sum(1 for p in range(s.n_paths) if s.paints[p].gradient >= 0)
    assert len(sides) >= 4
    seg = verbs[abi.VERB_LINE] + verbs[abi.VERB_QUAD] + verbs[abi.VERB_CUBIC]
    assert abs(verbs[abi.VERB_LINE] / seg - 0.4) < 0.02
    assert abs(verbs[abi.VERB_CUBIC] / seg - 0.3) < 0.02
    assert abs(grads / sum(paths) - 0.3) < 0.03
