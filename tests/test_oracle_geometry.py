"""Pins of the shape-rasterisation oracle (oracle/geometry.py) by lattice
counts that integer arithmetic fixes, independent of the FP64 expressions
the oracle restates."""

import numpy as np
import pytest

from oracle import geometry as G


def _lattice(lo, hi):
    return np.arange(lo, hi + 1, dtype=np.float64)


def test_box_count_is_product_of_intervals():
    x, y, z = _lattice(-7, 9), _lattice(-5, 6), _lattice(-4, 8)
    # closed box [cx - hx, cx + hx] per axis, half sizes from integer sizes
    shape = {"kind": "box", "center": (1.0, -0.5, 2.0), "size": (6.0, 5.0, 7.0), "material": 3}
    m = G.paint([shape], x, y, z)

    def count(v, c, s):  # integers v with |v - c| <= s/2, done in halves
        return sum(1 for t in v if abs(2 * int(t) - int(2 * c)) <= int(s))

    want = count(x, 1.0, 6) * count(y, -0.5, 5) * count(z, 2.0, 7)
    assert int((m == 3).sum()) == want
    assert set(np.unique(m)) <= {0, 3}


def test_sphere_integer_radius_lattice_count():
    r = 5
    v = _lattice(-8, 8)
    m = G.paint([{"kind": "sphere", "center": (0.0, 0.0, 0.0), "radius": float(r),
                  "material": 1}], v, v, v)
    want = sum(1 for i in range(-8, 9) for j in range(-8, 9) for k in range(-8, 9)
               if i * i + j * j + k * k <= r * r)
    assert int(m.sum()) == want


@pytest.mark.parametrize("axis", ["x", "y", "z"])
def test_cylinder_integer_lattice_count(axis):
    r, h = 4, 6
    v = _lattice(-6, 6)
    m = G.paint([{"kind": "cylinder", "center": (0.0, 0.0, 0.0), "radius": float(r),
                  "height": float(h), "axis": axis, "material": 2}], v, v, v)
    disc = sum(1 for i in range(-6, 7) for j in range(-6, 7) if i * i + j * j <= r * r)
    along = sum(1 for k in range(-6, 7) if 2 * abs(k) <= h)
    assert int((m == 2).sum()) == disc * along


def test_polygon_rectangle_is_half_open():
    x, y = _lattice(-3, 12), _lattice(-2, 9)
    rect = [(1.0, 2.0), (7.0, 2.0), (7.0, 6.0), (1.0, 6.0)]
    m = G.paint([{"kind": "polygon", "vertices": rect, "material": 5}], x, y)
    inside = {(i, j) for i in range(1, 7) for j in range(2, 6)}  # [1, 7) x [2, 6)
    got = {(int(x[a]), int(y[b])) for a, b in zip(*np.nonzero(m[:, :, 0] == 5))}
    assert got == inside


def test_polygon_l_shape_decomposes():
    x, y = _lattice(-1, 10), _lattice(-1, 10)
    ell = [(0.0, 0.0), (8.0, 0.0), (8.0, 3.0), (3.0, 3.0), (3.0, 7.0), (0.0, 7.0)]
    m = G.paint([{"kind": "polygon", "vertices": ell, "material": 1}], x, y)
    want = 8 * 3 + 3 * 4  # [0, 8) x [0, 3) and [0, 3) x [3, 7)
    assert int(m.sum()) == want


def test_polygon_extrusion_bounds():
    x, y, z = _lattice(0, 5), _lattice(0, 5), _lattice(-3, 3)
    sq = [(0.0, 0.0), (4.0, 0.0), (4.0, 4.0), (0.0, 4.0)]
    m = G.paint([{"kind": "polygon", "vertices": sq, "z_min": -1.0, "z_max": 2.0,
                  "material": 1}], x, y, z)
    assert int(m.sum()) == 16 * 4  # z in {-1, 0, 1, 2}


def test_painter_order_last_wins():
    v = _lattice(-4, 4)
    big = {"kind": "box", "center": (0.0, 0.0, 0.0), "size": (8.0, 8.0, 8.0), "material": 1}
    small = {"kind": "box", "center": (0.0, 0.0, 0.0), "size": (2.0, 2.0, 2.0), "material": 2}
    m = G.paint([big, small], v, v, v)
    assert int((m == 2).sum()) == 27 and int((m == 1).sum()) == 9 ** 3 - 27
    m2 = G.paint([small, big], v, v, v)
    assert int((m2 == 2).sum()) == 0
