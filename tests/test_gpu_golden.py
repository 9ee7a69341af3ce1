"""The HIP step against the committed oracle fixtures (tests/golden), without
running the oracle: each fixture's scene is rebuilt from its stored inputs,
stepped five times on the GPU, and held bit for bit to the stored FP32
outputs (mask, flags, density, u, v, w of the output half, and the pressure
of the pressure forms) and within the stated tolerance of the stored FP64
outputs (relative L2: density <= 1e-4, each velocity component and the
pressure <= 1e-3; density max |diff| <= 1e-4)."""
import numpy as np
import pytest

import smokesim
from golden import make_fixtures as mf
from helpers import rel_l2

pytestmark = pytest.mark.gpu

IDS = [c[0] for c in mf.CASES]


def gpu_from_fixture(fx, storage=smokesim.STORAGE_FP32):
    nx, ny, nz = (int(v) for v in fx["dims"])
    return mf.load_inputs(smokesim.Simulation(nx, ny, nz, storage=storage), fx)


def outputs(sim, now):
    out = {"mask": sim.get_mask(), "flags": sim.get_flags(), "rho": sim.get_density(now),
           "pressure": sim.get_pressure()}
    for k, name in enumerate("uvw"):
        out[name] = sim.get_velocity(k, now)
    return out


def check_fp32_bitwise(out, fx, context):
    assert np.array_equal(out["mask"], fx["mask"]), context
    assert np.array_equal(out["flags"], fx["flags"]), context
    keys = ["rho", "u", "v", "w"] + (["pressure"] if json_params(fx)["projection"] != 0 else [])
    for k in keys:
        n = int(np.count_nonzero(out[k].view(np.uint32) != fx[f"{k}32"].view(np.uint32)))
        assert n == 0, f"{context}: {k} differs from the FP32 fixture in {n} values"


def json_params(fx):
    return mf.json.loads(str(fx["params"]))


@pytest.mark.parametrize("name", IDS)
def test_step_equals_fp32_fixture_and_meets_fp64_tolerance(name):
    fx = mf.load_fixture(name)
    with gpu_from_fixture(fx) as g:
        dt = float(fx["dt"])
        for _ in range(int(fx["steps"])):
            g.step(dt, readback=False)
        g.synchronize()
        now = g.index_now
        assert now == int(fx["now"])
        out = outputs(g, now)
    check_fp32_bitwise(out, fx, name)
    assert rel_l2(out["rho"], fx["rho64"]) <= 1e-4
    assert np.abs(out["rho"] - fx["rho64"]).max() <= 1e-4
    for k in ("u", "v", "w"):
        assert rel_l2(out[k], fx[f"{k}64"]) <= 1e-3, k
    if json_params(fx)["projection"] != 0:
        assert rel_l2(out["pressure"], fx["pressure64"]) <= 1e-3


@pytest.mark.parametrize("use_graph", [False, True], ids=["launches", "graph"])
def test_batched_step_equals_fp32_fixture(use_graph):
    """The Jacobi fixture stepped as a batch member (persistent kernels over
    the batch's work list; with use_graph, replays of the captured step)."""
    fx = mf.load_fixture("p2_20x16x12")
    with gpu_from_fixture(fx) as g:
        with smokesim.Batch([g], [float(fx["dt"])]) as b:
            for _ in range(int(fx["steps"])):
                b.step(use_graph=use_graph)
            b.synchronize()
            assert g.index_now == int(fx["now"])
            out = outputs(g, g.index_now)
    check_fp32_bitwise(out, fx, f"batch graph={use_graph}")


@pytest.mark.parametrize("name,storage,refined", [
    ("p1_16x12x20", smokesim.STORAGE_FP16, False),
    ("p2_20x16x12", smokesim.STORAGE_FP16, False),
    ("p2_20x16x12", smokesim.STORAGE_FP32, True),
    ("p2_20x16x12", smokesim.STORAGE_FP16, True)],
    ids=["fp16-rbsor", "fp16-jacobi", "fp32-refined", "fp16-refined"])
def test_reduced_precision_forms_meet_tolerance_of_fp64_fixture(name, storage, refined):
    """The 16-bit storage forms and the refined solve against the stored FP64
    outputs: mask and flags bit-exact; relative L2 of density <= 1e-3 and of
    velocity and pressure <= 3e-3 (the refined solve runs 2 refinements of
    half the Jacobi sweeps, the same sweep count)."""
    fx = mf.load_fixture(name)
    with gpu_from_fixture(fx, storage=storage) as g:
        if refined:
            its = json_params(fx)["iterations"]
            g.set_params(projection=smokesim.PROJ_REFINED, inner_sweeps=its // 2, iterations=its)
        for _ in range(int(fx["steps"])):
            g.step(float(fx["dt"]), readback=False)
        g.synchronize()
        out = outputs(g, g.index_now)
    assert np.array_equal(out["mask"], fx["mask"]) and np.array_equal(out["flags"], fx["flags"])
    errs = {k: rel_l2(out[k], fx[f"{k}64"]) for k in ("rho", "u", "v", "w", "pressure")}
    print(name, storage, refined, errs)
    assert errs["rho"] <= 1e-3, errs
    assert max(errs["u"], errs["v"], errs["w"], errs["pressure"]) <= 3e-3, errs
