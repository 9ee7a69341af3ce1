"""Tests that pin the oracle of the renderer's proxy-plane slicing geometry
(project/boundingBox.cpp:133-361): per view, `steps` planes perpendicular to
the view direction, each cut with the volume's bounding box, its corner
points ordered around their mean and fanned into triangles.

Expected values come from independent constructions: axis-aligned views whose
cuts are known rectangles, a float64 Sutherland-Hodgman clip of the plane
against the box's six half-spaces for the cut's area and corner count, exact
plane and edge membership, and the plane spacing the reference defines
(t = (k + 1/2) / steps between the nearest and farthest box corner)."""
import numpy as np
import pytest

import oracle as orc


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def clip_area(box_pos, box_dims, plane_pos, n):
    """Area and vertex count of plane ∩ box by Sutherland-Hodgman (float64)."""
    n = unit(n)
    a = unit(np.cross(n, [1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.cross(n, [0.0, 1.0, 0.0]))
    b = np.cross(n, a)
    big = 10.0 * float(np.max(box_dims)) + 10.0 * float(np.linalg.norm(plane_pos))
    poly = [plane_pos + big * (sa * a + sb * b) for sa, sb in ((1, 1), (-1, 1), (-1, -1), (1, -1))]
    lo = np.asarray(box_pos, np.float64) - np.asarray(box_dims, np.float64) / 2
    hi = np.asarray(box_pos, np.float64) + np.asarray(box_dims, np.float64) / 2
    for ax in range(3):
        for sign, bound in ((1, hi[ax]), (-1, lo[ax])):
            out = []
            for i in range(len(poly)):
                p, q = poly[i], poly[(i + 1) % len(poly)]
                fp, fq = sign * (p[ax] - bound), sign * (q[ax] - bound)
                if fp <= 0:
                    out.append(p)
                if (fp < 0) != (fq < 0) and fp != fq:
                    out.append(p + (q - p) * (fp / (fp - fq)))
            poly = out
            if not poly:
                return 0.0, 0
    c = np.mean(poly, axis=0)
    area = sum(0.5 * abs(np.dot(np.cross(poly[i] - c, poly[(i + 1) % len(poly)] - c), n))
               for i in range(len(poly)))
    return area, len(poly)


def fan_area(points, n):
    c = points[-1].astype(np.float64)
    m = len(points) - 1
    cr = [np.dot(np.cross(points[i] - c, points[(i + 1) % m] - c), n) for i in range(m)]
    return 0.5 * sum(abs(x) for x in cr), cr


def test_axis_aligned_view_cuts_rectangles():
    dims = (4.0, 6.0, 8.0)
    v, idx, planes = orc.proxy_geometry((0, 0, 0), dims, (0, 0, 0), (0, 0, -1), 4)
    assert v.shape == (20, 3) and idx.shape == (48,)
    for k in range(4):
        z = -4.0 + 8.0 * (k + 0.5) / 4
        pts = v[5 * k:5 * k + 5]
        assert sorted(map(tuple, pts[:4])) == sorted((x, y, z) for x in (-2.0, 2.0) for y in (-3.0, 3.0))
        assert tuple(pts[4]) == (0.0, 0.0, z)
        assert tuple(planes[k]) == (4, 4 * k)
        want = []
        for i in range(3):
            want += [5 * k + 4, 5 * k + i, 5 * k + i + 1]
        want += [5 * k + 4, 5 * k + 3, 5 * k]
        assert list(idx[12 * k:12 * k + 12]) == want


DIRS = [(0.3, -0.8, 0.5), (1.0, 1.0, 1.0), (-0.2, 0.1, -1.0), (0.7, 0.0, -0.7), (0.01, 1.0, 0.02),
        (-1.0, -0.5, 0.25)]


@pytest.mark.parametrize("view_dir", DIRS)
@pytest.mark.parametrize("steps", [1, 7, 100])
def test_every_cut_is_the_ordered_plane_box_polygon(view_dir, steps):
    dims = np.array([3.0, 5.0, 2.5], np.float32)
    origin = np.array([0.25, -0.5, 4.0], np.float32)
    v, idx, planes = orc.proxy_geometry((0, 0, 0), dims, origin, view_dir, steps)
    n = unit(-np.asarray(view_dir, np.float64))
    corners = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]) * dims / 2
    dist = (corners - origin) @ n
    lo, hi = dist.min(), dist.max()
    off = 0
    for k in range(steps):
        m, toff = planes[k]
        pts = v[off:off + m + 1]
        assert toff * 3 == sum(3 * planes[j][0] for j in range(k))
        # On the plane the reference places: depth lo + (hi - lo) (k + 1/2) / steps.
        depth = (pts.astype(np.float64) - origin) @ n
        want = lo + (hi - lo) * (k + 0.5) / steps
        assert np.allclose(depth, want, atol=1e-4 * (hi - lo + 1)), (k, depth, want)
        # Every corner on a box edge: two coordinates exactly on box faces.
        for p in pts[:-1]:
            on = np.isclose(np.abs(p), dims / 2, rtol=0, atol=0)
            assert on.sum() >= 2 and np.all(np.abs(p) <= dims / 2 * (1 + 1e-6)), p
        # Mean last, polygon ordered (one winding), area of the true cut.
        assert np.allclose(pts[-1], pts[:-1].mean(axis=0), atol=1e-5)
        area, ncut = clip_area(np.zeros(3), dims.astype(np.float64), n * want + origin, n)
        got, cr = fan_area(pts, n)
        assert m == ncut and 3 <= m <= 6
        assert all(c > 0 for c in cr) or all(c < 0 for c in cr)
        assert got == pytest.approx(area, rel=1e-4)
        # Fan triangles (centre, i, i + 1), closed by (centre, m - 1, 0).
        tri = idx[3 * toff:3 * (toff + m)].reshape(m, 3) - off
        assert np.all(tri[:, 0] == m)
        assert list(tri[:, 1]) == list(range(m - 1)) + [m - 1]
        assert list(tri[:, 2]) == list(range(1, m)) + [0]
        off += m + 1
    assert off == len(v) and len(idx) == 3 * planes[:, 0].sum()


def test_corner_hits_are_deduplicated():
    # The plane x + y + z = 3 touches the box [-1, 1]^3 only at (1, 1, 1),
    # which three edges reach (t = max_t exactly): one point survives.
    pts = orc.plane_box((0, 0, 0), (2, 2, 2), (1, 1, 1), (1, 1, 1))
    assert pts.tolist() == [[1.0, 1.0, 1.0], [1.0, 1.0, 1.0]]


def test_plane_containing_edges_keeps_the_references_nan_points():
    # The face plane x = 1: the four x-edges end on it (t = max_t), the four
    # y/z-edges on it give 0/0 (NaN points, never equal, kept), the rest miss.
    pts = orc.plane_box((0, 0, 0), (2, 2, 2), (1, 0, 0), (1, 0, 0))
    assert len(pts) == 9
    assert np.isfinite(pts[:4]).all() and np.all(pts[:4, 0] == 1.0)
    assert np.isnan(pts[4:]).all()


def test_off_centre_box_uses_origin_centred_depths():
    # getMinMaxPoints measures the corners of a box centred at the origin,
    # not at the box position, so planes can miss a shifted box: those give
    # only their (NaN) mean and no triangles.
    v, idx, planes = orc.proxy_geometry((10, 0, 0), (2, 2, 2), (0, 0, 0), (-1, 0, 0), 5)
    assert np.all(planes[:, 0] == 0) and len(idx) == 0 and len(v) == 5
    assert np.isnan(v).all()
    v, idx, planes = orc.proxy_geometry((0.5, 0, 0), (2, 2, 2), (0, 0, 0), (-1, 0, 0), 4)
    # Depths -1 + 2 (k + 1/2) / 4 = -0.75, -0.25, 0.25, 0.75: the box spans
    # x in [-0.5, 1.5], so the first plane misses and the other three hit.
    assert list(planes[:, 0]) == [0, 4, 4, 4]


def test_sort_is_descending_pseudoangle():
    # A hexagonal cut (normal (1,1,1) through the centre) sorted by the
    # oracle: the pseudoangles about the mean, measured in the frame the
    # reference builds from the first point found, decrease along the list.
    raw = orc.plane_box((0, 0, 0), (2, 2, 2), (0, 0, 0), (1, 1, 1))
    srt = orc.plane_box((0, 0, 0), (2, 2, 2), (0, 0, 0), (1, 1, 1), sort=True)
    assert len(raw) == 7 and sorted(map(tuple, raw[:-1])) == sorted(map(tuple, srt[:-1]))
    c = raw[-1].astype(np.float64)
    n = unit((1, 1, 1))
    ax = unit(raw[0] - c)
    ay = np.cross(n, ax)

    def pa(p):
        x, y = np.dot(ax, p - c), np.dot(ay, p - c)
        r = x / (abs(x) + abs(y))
        return 1 + r if y < 0 else 3 - r

    keys = [pa(p) for p in srt[:-1].astype(np.float64)]
    assert all(keys[i] > keys[i + 1] for i in range(len(keys) - 1))
