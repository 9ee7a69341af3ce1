"""The oracle against its committed regression fixtures (tests/golden).

Each fixture holds a small scene (one per projection form), its inputs and
the oracle's outputs after five steps in FP64 and in FP32.  Rebuilding the
scene from the stored inputs and stepping it again must give every stored
value bit for bit: a change to the oracle's arithmetic, to its parameter
handling or to the scene stage shows here before it can shift the bar the
HIP step is held to.  The fixtures also pin the FP32/FP64 gap of each form
(the tolerance the FP32 step is held to against FP64)."""
import numpy as np
import pytest

from golden import make_fixtures as mf
from helpers import rel_l2

FIELDS = ("rho", "u", "v", "w", "pressure")


@pytest.mark.parametrize("case", mf.CASES, ids=[c[0] for c in mf.CASES])
@pytest.mark.parametrize("bits", [64, 32])
def test_oracle_reproduces_fixture(case, bits):
    fx = mf.load_fixture(case[0])
    out = mf.run_oracle(fx, bits)
    assert np.array_equal(out["mask"], fx["mask"])
    assert np.array_equal(out["flags"], fx["flags"])
    assert int(out["now"]) == int(fx["now"])
    for k in FIELDS:
        want = fx[f"{k}{bits}"]
        assert out[k].dtype == want.dtype and out[k].shape == want.shape, k
        assert mf.same(out[k], want), f"{k}: {np.count_nonzero(out[k] != want)} values differ"


@pytest.mark.parametrize("case", mf.CASES, ids=[c[0] for c in mf.CASES])
def test_fixture_inputs_match_their_case(case):
    name, dims, seed, projection, iterations, steps = case
    fx = mf.load_fixture(name)
    assert tuple(int(v) for v in fx["dims"]) == dims
    assert int(fx["steps"]) == steps
    nx, ny, nz = dims
    assert fx["rho_in"].shape == fx["rho64"].shape == fx["mask"].shape == (nz, ny, nx)
    for k in "uvw":
        assert fx[f"{k}_in"].shape == fx[f"{k}64"].shape == (nz + 1, ny + 1, nx + 1)
    params = mf.json.loads(str(fx["params"]))
    assert (params["projection"], params["iterations"]) == (projection, iterations)


@pytest.mark.parametrize("case", mf.CASES, ids=[c[0] for c in mf.CASES])
def test_fixture_precisions_agree_to_the_stated_tolerance(case):
    """FP32 against FP64 after five steps: the bar the HIP step is held to
    (density <= 1e-4, each velocity component and the pressure <= 1e-3
    relative L2), with real content in every field."""
    fx = mf.load_fixture(case[0])
    for k, tol in (("rho", 1e-4), ("u", 1e-3), ("v", 1e-3), ("w", 1e-3), ("pressure", 1e-3)):
        assert rel_l2(fx[f"{k}32"], fx[f"{k}64"]) <= tol, k
    for k in ("rho", "u", "v", "w"):
        assert np.abs(fx[f"{k}64"]).max() > 0.1, k
    if case[3] != 0:  # the pressure forms leave their last pressure
        assert np.abs(fx["pressure64"]).max() > 0.1
