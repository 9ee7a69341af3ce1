"""The oracle's step norms (oracle.step_norms, the reduction behind
smoke_get_step_norms) pinned by closed forms: a uniform velocity has zero
divergence and its components as maxima; the face field u = x has divergence
exactly 1 on every relaxed cell; a field whose components do not vary along
their own axis (u(y, z), v(x, z), w(x, y)) has divergence exactly 0; injected
NaN/Inf values are counted and do not poison the maxima; the divergence
operand order equals the oracle's own right-hand side bit for bit; and the
projection lowers the divergence more with more Jacobi sweeps."""
import numpy as np

import oracle as orc
from smokesim import scenes

DIMS = (13, 11, 9)


def fresh(dims=DIMS, seed=3, boxes=0.05):
    sc = scenes.test_scene(dims, seed=seed, box_fraction=boxes, projection=2, iterations=10)
    o = orc.Oracle(*dims, 32)
    scenes.apply(o, sc)
    return o, sc


def set_both(o, comp, arr):
    for h in (0, 1):
        o.set_velocity(comp, h, arr)


def test_uniform_velocity_has_zero_divergence_and_known_maxima():
    o, _ = fresh()
    shape = o.face_shape
    vals = (0.75, -1.5, 0.125)
    for k, a in enumerate(vals):
        set_both(o, k, np.full(shape, a))
    n = orc.step_norms(o, 0.05)
    assert n["div_l2"] == 0.0
    assert (n["max_u"], n["max_v"], n["max_w"]) == (0.75, 1.5, 0.125)
    assert n["nonfinite"] == 0
    assert n["cfl"] == float(np.float64(np.float32(0.05)) * 1.5)


def test_linear_field_has_unit_divergence_on_relaxed_cells():
    o, _ = fresh()
    nz, ny, nx = o.cell_shape
    u = np.broadcast_to(np.arange(nx + 1, dtype=np.float64), o.face_shape).copy()
    set_both(o, 0, u)
    set_both(o, 1, np.zeros(o.face_shape))
    set_both(o, 2, np.zeros(o.face_shape))
    active = int(np.count_nonzero(o.get_flags() & orc.F_ACTIVE))
    assert active > 0
    n = orc.step_norms(o, 0.05)
    assert n["div_l2"] == float(np.sqrt(active))
    # faces at x = 0 and x = nx are never written and are left out
    assert n["max_u"] == float(nx - 1)


def test_fields_constant_along_their_own_axis_are_divergence_free():
    o, _ = fresh(seed=5)
    rng = np.random.default_rng(7)
    nz1, ny1, nx1 = o.face_shape
    u = np.broadcast_to(rng.standard_normal((nz1, ny1, 1)), o.face_shape).copy()
    v = np.broadcast_to(rng.standard_normal((nz1, 1, nx1)), o.face_shape).copy()
    w = np.broadcast_to(rng.standard_normal((1, ny1, nx1)), o.face_shape).copy()
    for k, a in enumerate((u, v, w)):
        set_both(o, k, a.astype(np.float32))
    n = orc.step_norms(o, 0.05)
    assert n["div_l2"] == 0.0
    assert n["max_u"] > 0 and n["max_v"] > 0 and n["max_w"] > 0


def test_nonfinite_values_are_counted_not_maximised():
    o, _ = fresh()
    h = 1 - o.index_now
    u = np.full(o.face_shape, 0.5)
    u[3, 4, 5] = np.nan
    o.set_velocity(0, h, u)
    rho = np.zeros(o.cell_shape)
    rho[2, 2, 2] = np.inf
    rho[1, 1, 1] = 0.25
    o.set_density(h, rho)
    n = orc.step_norms(o, 0.05)
    assert n["nonfinite"] == 2
    assert n["max_u"] == 0.5
    assert n["max_density"] == np.inf
    # a NaN on a face the stages never write (u at x = 0) is not counted
    u[3, 4, 5] = 0.5
    u[2, 2, 0] = np.nan
    o.set_velocity(0, h, u)
    rho[2, 2, 2] = 0.0
    o.set_density(h, rho)
    assert orc.step_norms(o, 0.05)["nonfinite"] == 0


def test_divergence_order_is_the_right_hand_side_of_the_oracle():
    o, sc = fresh(seed=9)
    o.stage(orc.STAGE_SCENE, sc["dt"])
    o.stage(orc.STAGE_FORCE, sc["dt"])
    now = o.index_now
    div = orc.divergence_f32(*(o.get_velocity(k, now) for k in range(3)))
    o.stage(orc.STAGE_PROJECT, sc["dt"])
    act = (o.get_flags() & orc.F_ACTIVE) != 0
    rhs = o.get_rhs()
    assert np.array_equal(np.where(act, div, 0).astype(np.float32).view(np.uint32),
                          np.where(act, rhs, 0).astype(np.float32).view(np.uint32))


def test_more_sweeps_leave_less_divergence():
    rel = []
    for k in (0, 5, 40, 400):
        o, sc = fresh(dims=(16, 14, 12), seed=11)
        o.set_params(iterations=k)
        o.step(sc["dt"])
        n = orc.step_norms(o, sc["dt"], rhs=o.get_rhs())
        assert n["rhs_l2"] > 0
        rel.append(n["div_l2"] / n["rhs_l2"])
    assert rel[0] == 1.0  # no sweeps: the projection leaves the velocity as it was
    assert rel[0] > rel[1] > rel[2] > rel[3]
    assert rel[3] < 1e-2
