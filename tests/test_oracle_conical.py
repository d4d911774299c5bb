"""Pins of the oracle's two-point conical gradient (ConicalGradientColorBrush,
src/render/sw/sw_span_brush.cc:385-506) against the definition of the
gradient: at a point p, t is the largest root of

    |p - c(t)| = r(t),   c(t) = (1 - t) c0 + t c1,   r(t) = (1 - t) r0 + t r1,

with r(t) >= 0, and the pixel is transparent where there is none.  The
definition is solved here in float64 at pixel centres; pixels where the
discriminant or r(t) is within a small margin of 0 (where float rounding
decides) are left out.  Red -> blue stops, clamp tiling: a channel is
trunc(255 * lerp), so the expected value is known to 1 LSB.

The focal-on-circle case (the start centre on the end circle) is not pinned:
there the reference evaluates x_t = (x^2 + y^2) / 2, which is not the
definition's root."""
import math

import pytest

from oracle import oracle_py as O
from skity_amd import abi


def _grad(c0, r0, c1, r1, tile=abi.TILE_CLAMP):
    g = abi.Gradient()
    g.type = abi.GRADIENT_CONICAL
    g.tile_mode = tile
    g.color_count = 2
    for j, v in enumerate((1, 0, 0, 1)):
        g.colors[0][j] = v
    for j, v in enumerate((0, 0, 1, 1)):
        g.colors[1][j] = v
    g.points[0], g.points[1], g.points[2], g.points[3] = c0[0], c0[1], c1[0], c1[1]
    g.radius[0], g.radius[1] = r0, r1
    for i, v in enumerate((1, 0, 0, 0, 1, 0)):
        g.local_matrix[i] = v
    return g


def conical_t(p, c0, r0, c1, r1, margin=1e-3):
    """(t, robust) from the definition; t None when no root has r(t) >= 0."""
    dx, dy, dr = c1[0] - c0[0], c1[1] - c0[1], r1 - r0
    qx, qy = p[0] - c0[0], p[1] - c0[1]
    a = dx * dx + dy * dy - dr * dr
    b = -2 * (qx * dx + qy * dy + r0 * dr)
    c = qx * qx + qy * qy - r0 * r0
    roots = []
    robust = True
    if abs(a) < 1e-12:
        if abs(b) > 1e-12:
            roots = [-c / b]
    else:
        disc = b * b - 4 * a * c
        if abs(disc) < margin * max(1.0, b * b):
            robust = False
        if disc >= 0:
            s = math.sqrt(disc)
            roots = [(-b + s) / (2 * a), (-b - s) / (2 * a)]
    good = [t for t in roots if (1 - t) * r0 + t * r1 >= 0]
    for t in roots:
        if abs((1 - t) * r0 + t * r1) < margin * max(r0, r1, 1):
            robust = False
    return (max(good) if good else None), robust


def _argb(v):
    return ((v >> 16) & 255, (v >> 8) & 255, v & 255, v >> 24)


CASES = {
    # start circle inside the end circle: every point has a t
    "nested": ((40, 50), 10, (60, 50), 40),
    # circles apart: a cone, transparent outside it
    "cone": ((20, 50), 5, (80, 50), 12),
    # end radius 0: the reference swaps the circles
    "to_point": ((40, 50), 30, (62, 55), 0),
    # equal radii: the strip case
    "strip": ((30, 50), 15, (70, 55), 15),
    # equal centres: the radial case
    "concentric": ((50, 50), 10, (50, 50), 40),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_conical_against_definition(case):
    c0, r0, c1, r1 = CASES[case]
    g = _grad(c0, r0, c1, r1)
    checked = transparent = 0
    for y in range(0, 100, 3):
        for x in range(0, 100, 3):
            t, robust = conical_t((x + 0.5, y + 0.5), c0, r0, c1, r1)
            if not robust:
                continue
            r, gg, b, a = _argb(O.gradient_color(g, x, y))
            if t is None:
                assert (r, gg, b, a) == (0, 0, 0, 0), (case, x, y, (r, gg, b, a))
                transparent += 1
                continue
            tc = min(max(t, 0.0), 1.0)
            assert a == 255 and gg == 0, (case, x, y)
            assert abs(r - (1 - tc) * 255) <= 1.5 and abs(b - tc * 255) <= 1.5, \
                (case, x, y, t, r, b)
            checked += 1
    assert checked > 150
    if case == "cone":
        assert transparent > 100


def test_concentric_conical_equals_radial_ring():
    """c0 = c1, r0 = 0: t = |p - c| / r1, a radial gradient of radius r1."""
    g = _grad((50, 50), 0, (50, 50), 40)
    rad = abi.Gradient()
    rad.type = abi.GRADIENT_RADIAL
    rad.tile_mode = abi.TILE_CLAMP
    rad.color_count = 2
    for j, v in enumerate((1, 0, 0, 1)):
        rad.colors[0][j] = v
    for j, v in enumerate((0, 0, 1, 1)):
        rad.colors[1][j] = v
    rad.points[0], rad.points[1] = 50, 50
    rad.radius[0] = 40
    for i, v in enumerate((1, 0, 0, 0, 1, 0)):
        rad.local_matrix[i] = v
    diff = max(abs(a - b) for y in range(0, 100, 2) for x in range(0, 100, 2)
               for a, b in zip(_argb(O.gradient_color(g, x, y)), _argb(O.gradient_color(rad, x, y))))
    assert diff <= 1


def test_negative_radius_is_transparent():
    g = _grad((40, 50), -1, (60, 50), 20)
    assert all(O.gradient_color(g, x, 50) == 0 for x in range(0, 100, 5))
