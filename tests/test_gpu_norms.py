"""Step norms on the GPU (smoke_get_step_norms): the device reduction equals
the oracle's restatement of it (oracle.step_norms) on the same fields -- maxima
and the non-finite count exactly, FP64 sums to 1e-12 relative (only their
summation order differs) -- after real steps of every projection and of BF16
fields, on the closed-form fields of the oracle tests, with NaN/Inf injected
(status SMOKE_ERR_NONFINITE), taken in-step or afterwards, and repeatably (a
fixed reduction order gives the same bits every time)."""
import numpy as np
import pytest

import oracle as orc
import smokesim
from helpers import gpu_for, oracle_for
from smokesim import scenes

pytestmark = pytest.mark.gpu

SUM_RTOL = 1e-12


def check(gpu_norms, ref, what):
    for k in ("max_u", "max_v", "max_w", "max_density", "nonfinite"):
        assert gpu_norms[k] == ref[k], (what, k, gpu_norms[k], ref[k])
    for k in ("div_l2", "rhs_l2", "density_sum", "cfl"):
        a, b = gpu_norms[k], ref[k]
        assert abs(a - b) <= SUM_RTOL * max(abs(b), 1e-300), (what, k, a, b)


def oracle_steps_with_rhs(sc, steps):
    """The FP32 oracle after `steps` steps, and the right-hand side of the last
    step's projection: the divergence of [now] after the force stage, over
    relaxed cells (what every projection form solves against)."""
    o = oracle_for(sc, 32)
    o.step(sc["dt"], steps - 1)
    o.stage(orc.STAGE_SCENE, sc["dt"])
    o.stage(orc.STAGE_FORCE, sc["dt"])
    now = o.index_now
    rhs = orc.divergence_f32(*(o.get_velocity(k, now) for k in range(3)))
    for st in (orc.STAGE_PROJECT, orc.STAGE_ADVECT_VEL, orc.STAGE_ADVECT_DENS):
        o.stage(st, sc["dt"])
    return o, rhs


@pytest.mark.parametrize("proj", [smokesim.PROJ_JACOBI, smokesim.PROJ_RBSOR,
                                  smokesim.PROJ_REFERENCE])
def test_norms_after_steps_equal_the_restatement(proj):
    sc = scenes.test_scene((37, 29, 23), seed=21, projection=proj, iterations=12, noise=0.3)
    g = gpu_for(sc)
    g.norms_enable(True)
    for _ in range(3):
        g.step(sc["dt"], readback=False)
    n = g.step_norms()
    assert n["status"] == smokesim.SMOKE_OK and n["step"] == 3
    o, rhs = oracle_steps_with_rhs(sc, 3)
    check(n, orc.step_norms(g, sc["dt"], rhs=rhs), f"gpu fields p{proj}")
    check(n, orc.step_norms(o, sc["dt"], rhs=rhs), f"oracle fields p{proj}")
    assert 0.0 <= n["div_l2"] < n["rhs_l2"]
    # taken again after the step (not in-step): identical bits
    g.norms_enable(False)
    m = g.step_norms()
    assert all(m[k] == n[k] for k in n), (m, n)
    g.close()


@pytest.mark.parametrize("storage", [smokesim.STORAGE_BF16, smokesim.STORAGE_FP16,
                                     smokesim.STORAGE_FP16_VELOCITY])
def test_norms_of_half_fields(storage):
    sc = scenes.test_scene((33, 21, 26), seed=22, projection=2, iterations=10, noise=0.3)
    g = gpu_for(sc, storage=storage)
    for _ in range(2):
        g.step_async(sc["dt"])
    n = g.step_norms()
    ref = orc.step_norms(g, sc["dt"])   # the 16-bit values widen exactly
    ref["rhs_l2"] = n["rhs_l2"]
    check(n, ref, f"storage {storage}")
    assert n["rhs_l2"] > n["div_l2"] >= 0.0
    g.close()


def test_closed_form_fields():
    sc = scenes.test_scene((21, 17, 13), seed=4, projection=2, iterations=8)
    g = gpu_for(sc)
    o = oracle_for(sc, 32)
    shape = g.face_shape
    nz1, ny1, nx1 = shape
    rng = np.random.default_rng(3)
    fields = {
        "uniform": [np.full(shape, a, np.float32) for a in (0.75, -1.5, 0.125)],
        "linear": [np.broadcast_to(np.arange(nx1, dtype=np.float32), shape).copy(),
                   np.zeros(shape, np.float32), np.zeros(shape, np.float32)],
        "axis-free": [np.broadcast_to(rng.standard_normal((nz1, ny1, 1)), shape).astype(np.float32),
                      np.broadcast_to(rng.standard_normal((nz1, 1, nx1)), shape).astype(np.float32),
                      np.broadcast_to(rng.standard_normal((1, ny1, nx1)), shape).astype(np.float32)],
    }
    active = int(np.count_nonzero(g.get_flags() & smokesim.F_ACTIVE))
    for name, (u, v, w) in fields.items():
        for sim in (g, o):
            for h in (0, 1):
                for k, a in enumerate((u, v, w)):
                    sim.set_velocity(k, h, a)
        n = g.step_norms()
        check(n, orc.step_norms(o, 0.0), name)
        if name == "linear":
            assert n["div_l2"] == float(np.sqrt(active))
        else:
            assert n["div_l2"] == 0.0
    g.close()


def test_nonfinite_values_are_reported():
    sc = scenes.test_scene((19, 15, 12), seed=6, projection=2, iterations=6)
    g = gpu_for(sc)
    g.step(sc["dt"], readback=False)
    n = g.step_norms()
    assert n["status"] == smokesim.SMOKE_OK and n["nonfinite"] == 0
    h = 1 - g.index_now
    u = g.get_velocity(0, h)
    u[4, 5, 6] = np.nan
    g.set_velocity(0, h, u)
    rho = g.get_density(h)
    rho[3, 3, 3] = np.inf
    rho[2, 2, 2] = -np.inf
    g.set_density(h, rho)
    n = g.step_norms()
    assert n["status"] == smokesim.ERR_NONFINITE and n["nonfinite"] == 3
    assert n["max_density"] == np.inf
    assert np.isfinite(n["max_u"])
    with pytest.raises(smokesim.SmokeError) as e:
        smokesim._check(g.lib.smoke_get_step_norms(g.h, None))
    assert e.value.status == smokesim.ERR_INVALID
    g.close()


def test_in_step_norms_track_every_step_and_repeat_bitwise():
    sc = scenes.test_scene((40, 33, 27), seed=8, projection=2, iterations=20)
    a, b = gpu_for(sc), gpu_for(sc)
    a.norms_enable(True)
    for k in range(4):
        a.step_async(sc["dt"])
        b.step_async(sc["dt"])
        na, nb = a.step_norms(), b.step_norms()
        assert na["step"] == nb["step"] == k + 1
        assert {x: na[x] for x in na} == {x: nb[x] for x in nb}
    a.close()
    b.close()


def test_full_size_step_norms_256():
    """At the 256^3 benchmark size: norms equal the restatement on the GPU's own
    fields and the projection lowers the divergence."""
    sc = scenes.grid_256()
    g = gpu_for(sc)
    g.norms_enable(True)
    for _ in range(2):
        g.step_async(sc["dt"])
    n = g.step_norms()
    assert n["status"] == smokesim.SMOKE_OK and n["nonfinite"] == 0
    assert 0 < n["div_l2"] < n["rhs_l2"]
    ref = orc.step_norms(g, sc["dt"])
    for k in ("max_u", "max_v", "max_w", "max_density", "nonfinite"):
        assert n[k] == ref[k], k
    for k in ("div_l2", "density_sum"):
        assert abs(n[k] - ref[k]) <= SUM_RTOL * abs(ref[k]), k
    g.close()
