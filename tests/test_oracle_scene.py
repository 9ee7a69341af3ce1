"""Known-answer tests that pin the oracle's scene handling: the solid mask,
the cell flags and the density sources (project/smokeSimulation.cu:193-213,
:251-273, :289-313).  Every expected value is computed here by brute force
from the definitions, independently of the oracle's code."""
import numpy as np
import pytest

import oracle as orc


def brute_flags(mask):
    """Flags from a mask by direct neighbour inspection (numpy padding)."""
    nz, ny, nx = mask.shape
    s = mask.astype(bool)
    pad = np.zeros((nz + 2, ny + 2, nx + 2), dtype=bool)
    pad[1:-1, 1:-1, 1:-1] = s
    c = (slice(1, -1),) * 3
    nb = {
        orc.F_XM: pad[1:-1, 1:-1, :-2], orc.F_XP: pad[1:-1, 1:-1, 2:],
        orc.F_YM: pad[1:-1, :-2, 1:-1], orc.F_YP: pad[1:-1, 2:, 1:-1],
        orc.F_ZM: pad[:-2, 1:-1, 1:-1], orc.F_ZP: pad[2:, 1:-1, 1:-1],
    }
    fl = np.where(pad[c], orc.F_SELF, 0).astype(np.uint8)
    anyn = np.zeros_like(s)
    for bit, a in nb.items():
        fl |= np.where(a, bit, 0).astype(np.uint8)
        anyn |= a
    interior = np.zeros_like(s)
    interior[1:-1, 1:-1, 1:-1] = True
    fl |= np.where(interior & s & anyn, orc.F_ACTIVE, 0).astype(np.uint8)
    return fl


def brute_sphere(dims, cx, cy, cz, r):
    """Interior cells with (x-cx)^2+(y-cy)^2+(z-cz)^2 < r^2, in float32 as the
    reference computes it (integer index, float centre)."""
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    f = np.float32
    dx = x.astype(f) - f(cx)
    dy = y.astype(f) - f(cy)
    dz = z.astype(f) - f(cz)
    d = (dx * dx + dy * dy) + dz * dz
    inside = d < f(r) * f(r)
    interior = (x >= 1) & (x <= nx - 2) & (y >= 1) & (y <= ny - 2) & (z >= 1) & (z <= nz - 2)
    return inside & interior


@pytest.mark.parametrize("bits", [64, 32])
def test_initial_mask_is_floor(bits):
    o = orc.Oracle(7, 9, 11, bits)
    m = o.get_mask()
    assert m.shape == (11, 9, 7)
    assert (m[:, 0, :] == 0).all()
    assert (m[:, 1:, :] == 1).all()
    assert np.array_equal(o.get_flags(), brute_flags(m))


@pytest.mark.parametrize("centre", [(8, 8, 8), (7.5, 8.5, 6.5), (3, 12.5, 9)])
@pytest.mark.parametrize("r", [1.0, 2.5, 4.0, 6.0])
def test_sphere_rasterisation(centre, r):
    dims = (16, 16, 16)
    o = orc.Oracle(*dims, 64)
    o.add_obstacle(*centre, r)
    o.stage(orc.STAGE_SCENE, 0.05)
    expect = np.ones((16, 16, 16), dtype=np.uint8)
    expect[:, 0, :] = 0
    expect[brute_sphere(dims, *centre, r)] = 0
    assert np.array_equal(o.get_mask(), expect)
    assert np.array_equal(o.get_flags(), brute_flags(expect))


def test_last_obstacle_wins_and_union():
    dims = (16, 16, 16)
    a = (6.0, 8.0, 8.0, 4.0)
    b = (9.0, 8.0, 8.0, 3.0)
    o = orc.Oracle(*dims, 64)
    o.add_obstacle(*a)
    o.add_obstacle(*b)
    o.stage(orc.STAGE_SCENE, 0.05)
    m = o.get_mask()
    in_a = brute_sphere(dims, *a)
    in_b = brute_sphere(dims, *b)
    # Reference: the second obstacle overwrites the first everywhere inside
    # the interior (smokeSimulation.cu:304-310).
    assert (m[in_b] == 0).all()
    assert (m[in_a & ~in_b] == 1).all()
    o.set_params(obstacle_mode=orc.OBST_UNION)
    o.rebuild_mask()
    m = o.get_mask()
    assert (m[in_a | in_b] == 0).all()
    rest = ~(in_a | in_b)
    rest[:, 0, :] = False
    assert (m[rest] == 1).all()


def test_moving_obstacle_restores_fluid():
    dims = (16, 16, 16)
    o = orc.Oracle(*dims, 64)
    oid = o.add_obstacle(5.0, 8.0, 8.0, 3.0)
    o.stage(orc.STAGE_SCENE, 0.05)
    o.move_object(oid, 11.0, 8.0, 8.0)
    o.stage(orc.STAGE_SCENE, 0.05)
    m = o.get_mask()
    assert (m[brute_sphere(dims, 11.0, 8.0, 8.0, 3.0)] == 0).all()
    old = brute_sphere(dims, 5.0, 8.0, 8.0, 3.0) & ~brute_sphere(dims, 11.0, 8.0, 8.0, 3.0)
    assert (m[old] == 1).all()
    with pytest.raises(IndexError):
        o.move_object(7, 0, 0, 0)


def test_obstacles_never_touch_the_ring_and_base_mask_is_kept():
    dims = (12, 10, 14)
    solid = np.zeros((14, 10, 12), dtype=np.uint8)
    solid[:, 0, :] = 1
    solid[3:6, 2:4, 5:9] = 1          # a box
    o = orc.Oracle(*dims, 64)
    o.set_solids(solid)
    o.add_obstacle(0.0, 5.0, 7.0, 5.0)  # overlaps the x = 0 wall
    o.stage(orc.STAGE_SCENE, 0.05)
    m = o.get_mask()
    assert (m[:, :, 0][solid[:, :, 0] == 0] == 1).all()        # ring x=0 untouched
    assert (m[3:6, 2:4, 5:9] == 0).all()                         # base box kept
    assert (m[brute_sphere(dims, 0.0, 5.0, 7.0, 5.0)] == 0).all()
    assert np.array_equal(o.get_flags(), brute_flags(m))


def test_flags_random_masks():
    rng = np.random.default_rng(1)
    for dims in [(5, 6, 7), (9, 4, 3), (10, 10, 10)]:
        solid = (rng.random(dims[::-1]) < 0.3).astype(np.uint8)
        o = orc.Oracle(*dims, 64)
        o.set_solids(solid)
        m = o.get_mask()
        assert np.array_equal(m, (solid == 0).astype(np.uint8))
        assert np.array_equal(o.get_flags(), brute_flags(m))


@pytest.mark.parametrize("bits", [64, 32])
def test_sources_set_both_halves(bits):
    dims = (14, 12, 10)
    o = orc.Oracle(*dims, bits)
    rng = np.random.default_rng(2)
    d0 = rng.random((10, 12, 14)) * 0.5
    d1 = rng.random((10, 12, 14)) * 0.5
    o.set_density(0, d0)
    o.set_density(1, d1)
    src = [(3.0, 2.0, 4.0, 2.5), (12.5, 9.0, 8.0, 3.0)]   # second one straddles the ring
    for s in src:
        o.add_source(*s)
    o.stage(orc.STAGE_SCENE, 0.05)
    inside = np.zeros((10, 12, 14), dtype=bool)
    for s in src:
        inside |= brute_sphere(dims, *s)
    for h, d in ((0, d0), (1, d1)):
        got = o.get_density(h)
        assert (got[inside] == 1.0).all()
        assert np.array_equal(got[~inside], d[~inside].astype(np.float32 if bits == 32 else np.float64))


def test_source_inside_solid_still_writes_density():
    # The reference fills sources regardless of the mask (smokeSimulation.cu:256-269).
    dims = (12, 12, 12)
    o = orc.Oracle(*dims, 64)
    o.add_obstacle(6.0, 6.0, 6.0, 3.0)
    o.add_source(6.0, 6.0, 6.0, 2.0)
    o.stage(orc.STAGE_SCENE, 0.05)
    assert (o.get_density(0)[brute_sphere(dims, 6.0, 6.0, 6.0, 2.0)] == 1.0).all()
    assert (o.get_mask()[brute_sphere(dims, 6.0, 6.0, 6.0, 2.0)] == 0).all()


def test_colour_classes_share_no_face():
    """Red-black parity rule (smokeSimulation.cu:362): no two cells of one
    colour touch the same face, so the in-place sweep is race-free."""
    nx, ny, nz = 9, 8, 7
    for o in (0, 1):
        seen = set()
        for z in range(1, nz - 1):
            for y in range(1, ny - 1):
                for x in range(1, nx - 1):
                    if (x + y + z + o) % 2:
                        continue
                    faces = [("u", x, y, z), ("u", x + 1, y, z), ("v", x, y, z),
                             ("v", x, y + 1, z), ("w", x, y, z), ("w", x, y, z + 1)]
                    for f in faces:
                        assert f not in seen
                        seen.add(f)
