"""Known-answer and invariant tests that pin the oracle's physics stages:
buoyancy and clamp (project/smokeSimulation.cu:315-352), the sampler and the
advection kernels (:409-638), and the projection (:356-394) in its literal
velocity form and in the two pressure forms.  Expected values come from
closed forms, exact integer constructions or a dense direct solve, not from
re-evaluating the oracle's own expressions."""
import numpy as np
import pytest

import oracle as orc
from helpers import random_density


def fresh(dims, bits=64, **params):
    o = orc.Oracle(*dims, bits)
    if params:
        o.set_params(**params)
    return o


def cell_div(u, v, w):
    """Divergence of every cell from face arrays shaped (nz+1, ny+1, nx+1)."""
    return ((u[:-1, :-1, 1:] - u[:-1, :-1, :-1]) + (v[:-1, 1:, :-1] - v[:-1, :-1, :-1])
            + (w[1:, :-1, :-1] - w[:-1, :-1, :-1]))


def curl_field(dims, seed, scale=1.0 / 64):
    """Exactly divergence-free staggered field: the discrete curl of an
    integer vector potential, scaled by a power of two (exact in FP32)."""
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    A = [rng.integers(-8, 9, size=(nz + 2, ny + 2, nx + 2)).astype(np.float64) for _ in range(3)]
    Ax, Ay, Az = A
    sl = (slice(0, nz + 1), slice(0, ny + 1), slice(0, nx + 1))

    def sh(a, dz=0, dy=0, dx=0):
        return a[dz:dz + nz + 1, dy:dy + ny + 1, dx:dx + nx + 1]

    u = (sh(Az, dy=1) - Az[sl]) - (sh(Ay, dz=1) - Ay[sl])
    v = (sh(Ax, dz=1) - Ax[sl]) - (sh(Az, dx=1) - Az[sl])
    w = (sh(Ay, dx=1) - Ay[sl]) - (sh(Ax, dy=1) - Ax[sl])
    return [a * scale for a in (u, v, w)]


# ---------------------------------------------------------------- buoyancy

@pytest.mark.parametrize("gate", [0.0, 0.5])
def test_buoyancy_closed_form(gate):
    dims = (10, 9, 8)
    o = fresh(dims, gravity=1.5, buoyancy=2.25, max_velocity=1e6, density_gate=gate)
    solid = np.zeros((8, 9, 10), dtype=np.uint8)
    solid[:, 0, :] = 1
    solid[3, 4, 5] = solid[6, 2, 2] = 1
    o.set_solids(solid)
    rho = random_density(dims, 3).astype(np.float64)
    o.set_density(0, rho)
    rng = np.random.default_rng(4)
    vel = [rng.uniform(-1, 1, (9, 10, 11)) for _ in range(3)]
    for k in range(3):
        o.set_velocity(k, 0, vel[k])
    o.stage(orc.STAGE_SCENE, 0.125)
    o.stage(orc.STAGE_FORCE, 0.125)
    fluid = solid == 0
    expect = vel[1].copy()
    for z in range(8):
        for y in range(1, 9):
            for x in range(10):
                if fluid[z, y, x] and fluid[z, y - 1, x] and (gate <= 0 or rho[z, y, x] > gate):
                    expect[z, y, x] += rho[z, y, x] * (1.5 + 2.25) * 0.125
    np.testing.assert_allclose(o.get_velocity(1, 0), expect, rtol=0, atol=1e-14)
    np.testing.assert_array_equal(o.get_velocity(0, 0), vel[0])
    np.testing.assert_array_equal(o.get_velocity(2, 0), vel[2])


# -------------------------------------------------------------------- clamp

@pytest.mark.parametrize("w0,scaled", [(0.0, False), (0.25, True)])
def test_clamp_closed_form(w0, scaled):
    """velocityConfinement: scale by M^2/(L dt) when L dt > M^2, at the same
    staggered index, for x,y,z in [1, n-1] only."""
    dims = (6, 5, 7)
    dt = 0.5
    o = fresh(dims, gravity=0.0, buoyancy=0.0, max_velocity=3.0)
    shape = (8, 6, 7)
    vals = (3.0, 3.0, w0)      # L = 18 (+ w0^2); L*dt = 9 = M^2 exactly when w0 = 0
    for k in range(3):
        o.set_velocity(k, 0, np.full(shape, vals[k]))
    o.stage(orc.STAGE_SCENE, dt)
    o.stage(orc.STAGE_FORCE, dt)
    L = 18.0 + w0 * w0
    factor = 9.0 / (L * dt) if scaled else 1.0
    inside = np.zeros(shape, dtype=bool)
    inside[1:7, 1:5, 1:6] = True   # z,y,x in [1, n-1]
    for k in range(3):
        got = o.get_velocity(k, 0)
        np.testing.assert_allclose(got[inside], vals[k] * factor, rtol=1e-15, atol=0)
        np.testing.assert_array_equal(got[~inside], vals[k])
    if scaled:
        # Not a magnitude clamp: the new squared speed is M^4 / (L dt^2).
        got = [o.get_velocity(k, 0)[2, 2, 2] for k in range(3)]
        np.testing.assert_allclose(sum(g * g for g in got), 81.0 / (L * dt * dt), rtol=1e-14)


# ---------------------------------------------------------------- advection

def prep_density_advection(dims, bits, uvw, rho, solid=None):
    """Run the scene stage, then set the post-advection velocity half [past]
    to the given uniform values and the density half [now] to rho."""
    o = fresh(dims, bits)
    if solid is not None:
        o.set_solids(solid)
    o.stage(orc.STAGE_SCENE, 1.0)
    now = o.index_now
    past = 1 - now
    fshape = (dims[2] + 1, dims[1] + 1, dims[0] + 1)
    for k in range(3):
        o.set_velocity(k, past, np.full(fshape, uvw[k]))
    o.set_density(now, rho)
    return o, now, past


def test_sampler_reproduces_affine_density():
    dims = (12, 11, 10)
    z, y, x = np.meshgrid(np.arange(10), np.arange(11), np.arange(12), indexing="ij")
    rho = 0.25 + 0.01 * x - 0.02 * y + 0.015 * z
    o, now, past = prep_density_advection(dims, 64, (0.3, -0.2, 0.1), rho)
    o.stage(orc.STAGE_ADVECT_DENS, 1.0)
    got = o.get_density(past)
    expect = 0.25 + 0.01 * (x - 0.3) - 0.02 * (y + 0.2) + 0.015 * (z - 0.1)
    sel = np.zeros(rho.shape, dtype=bool)
    sel[1:-1, 1:-1, 1:-1] = True     # interior, fluid (floor is y = 0)
    np.testing.assert_allclose(got[sel], expect[sel], rtol=0, atol=1e-13)


@pytest.mark.parametrize("bits", [64, 32])
def test_integer_shift_is_exact(bits):
    dims = (12, 9, 8)
    rho = random_density(dims, 5)
    o, now, past = prep_density_advection(dims, bits, (1.0, 0.0, 0.0), rho)
    stale = o.get_density(past)
    o.stage(orc.STAGE_ADVECT_DENS, 1.0)
    got = o.get_density(past)
    # cells x >= 2: exact shift by one; x = 1 is clamped to position 1.
    np.testing.assert_array_equal(got[1:-1, 1:-1, 2:-1], rho[1:-1, 1:-1, 1:-2])
    np.testing.assert_allclose(got[1:-1, 1:-1, 1], 0.5 * (rho[1:-1, 1:-1, 0].astype(np.float64)
                                                           + rho[1:-1, 1:-1, 1]), atol=1e-7)
    # ring cells are never written: they keep the stale value
    np.testing.assert_array_equal(got[:, :, 0], stale[:, :, 0])
    np.testing.assert_array_equal(got[:, -1, :], stale[:, -1, :])


def test_wall_clamp_of_back_trace():
    dims = (14, 8, 8)
    rho = random_density(dims, 6).astype(np.float64)
    o, now, past = prep_density_advection(dims, 64, (10.0, 0.0, 0.0), rho)
    o.stage(orc.STAGE_ADVECT_DENS, 1.0)
    got = o.get_density(past)
    # x + 0.5 - 10 < 1 for x <= 9: position clamped to 1, i.e. halfway
    # between cells 0 and 1.
    np.testing.assert_allclose(got[1:-1, 1:-1, 1:10],
                               0.5 * (rho[1:-1, 1:-1, 0:1] + rho[1:-1, 1:-1, 1:2]).repeat(9, axis=2),
                               atol=1e-15)


def test_solid_cells_keep_stale_density():
    dims = (10, 10, 10)
    solid = np.zeros((10, 10, 10), dtype=np.uint8)
    solid[:, 0, :] = 1
    solid[4:7, 4:7, 4:7] = 1
    rho = random_density(dims, 7)
    o, now, past = prep_density_advection(dims, 64, (0.2, 0.1, -0.3), rho, solid)
    o.set_density(past, np.full((10, 10, 10), 7.0))
    o.stage(orc.STAGE_ADVECT_DENS, 1.0)
    got = o.get_density(past)
    assert (got[4:7, 4:7, 4:7] == 7.0).all()
    assert (got[1:-1, 1:-1, 1:-1][solid[1:-1, 1:-1, 1:-1] == 0] != 7.0).all()


@pytest.mark.parametrize("bits", [64, 32])
def test_velocity_advection_dt0_identity_and_stale_faces(bits):
    dims = (9, 8, 7)
    solid = np.zeros((7, 8, 9), dtype=np.uint8)
    solid[:, 0, :] = 1
    solid[3, 3:5, 4] = 1
    o = fresh(dims, bits)
    o.set_solids(solid)
    rng = np.random.default_rng(8)
    vel = [rng.uniform(-1, 1, (8, 9, 10)).astype(np.float32) for _ in range(3)]
    for k in range(3):
        o.set_velocity(k, 0, vel[k])
        o.set_velocity(k, 1, np.full((8, 9, 10), 7.0))
    o.stage(orc.STAGE_SCENE, 0.0)
    assert o.index_now == 0
    o.stage(orc.STAGE_ADVECT_VEL, 0.0)
    fl = solid == 0
    nx, ny, nz = dims
    for k, (lo, hi) in enumerate([((1, 1, 1), (nx - 1, ny - 2, nz - 2)),
                                  ((1, 1, 1), (nx - 2, ny - 1, nz - 2)),
                                  ((1, 1, 1), (nx - 2, ny - 2, nz - 1))]):
        got = o.get_velocity(k, 1)
        written = np.zeros(got.shape, dtype=bool)
        for z in range(lo[2], hi[2] + 1):
            for y in range(lo[1], hi[1] + 1):
                for x in range(lo[0], hi[0] + 1):
                    nb = [(z, y, x - 1), (z, y - 1, x), (z - 1, y, x)][k]
                    written[z, y, x] = fl[z, y, x] and fl[nb]
        np.testing.assert_array_equal(got[written], vel[k][written])
        assert (got[~written] == 7.0).all()


def test_velocity_advection_uses_8_point_averages():
    """V is traced back in x with avgU, an 8-point mean over z-1 and z
    (smokeSimulation.cu:409-420, :569): with u = z/2 the x-offset is
    (z - 1/2)/2, not z/2."""
    dims = (16, 8, 8)
    shape = (9, 9, 17)
    z, y, x = np.meshgrid(np.arange(9), np.arange(9), np.arange(17), indexing="ij")
    o = fresh(dims, 64)
    o.set_velocity(0, 0, 0.5 * z)
    o.set_velocity(1, 0, 0.25 * x)
    o.set_velocity(2, 0, np.zeros(shape))
    o.stage(orc.STAGE_SCENE, 1.0)
    o.stage(orc.STAGE_ADVECT_VEL, 1.0)
    got = o.get_velocity(1, 1)
    # v-faces at y = 1 touch the solid floor and are not written.
    zz, yy, xx = z[1:7, 2:8, 6:15], y[1:7, 2:8, 6:15], x[1:7, 2:8, 6:15]
    qx = xx + 0.5 - (0.5 * zz - 0.25) - 0.5
    np.testing.assert_allclose(got[1:7, 2:8, 6:15], 0.25 * qx, atol=1e-13)


def test_advection_selects_the_right_average_per_axis():
    dims = (10, 10, 10)
    shape = (11, 11, 11)
    z, y, x = np.meshgrid(np.arange(11), np.arange(11), np.arange(11), indexing="ij")
    # U traced in y by avgV: u depends on y only, v = 1.
    o = fresh(dims, 64)
    o.set_velocity(0, 0, 0.25 * y)
    o.set_velocity(1, 0, np.ones(shape))
    o.set_velocity(2, 0, np.zeros(shape))
    o.stage(orc.STAGE_SCENE, 1.0)
    o.stage(orc.STAGE_ADVECT_VEL, 1.0)
    got = o.get_velocity(0, 1)
    np.testing.assert_allclose(got[1:9, 2:9, 1:10], 0.25 * (y[1:9, 2:9, 1:10] - 1), atol=1e-13)
    # W traced in x by avgU: w depends on x only, u = 1.
    o = fresh(dims, 64)
    o.set_velocity(0, 0, np.ones(shape))
    o.set_velocity(1, 0, np.zeros(shape))
    o.set_velocity(2, 0, 0.25 * x)
    o.stage(orc.STAGE_SCENE, 1.0)
    o.stage(orc.STAGE_ADVECT_VEL, 1.0)
    got = o.get_velocity(2, 1)
    np.testing.assert_allclose(got[1:10, 1:9, 2:9], 0.25 * (x[1:10, 1:9, 2:9] - 1), atol=1e-13)


# --------------------------------------------------------------- projection

PROJ = [orc.PROJ_REFERENCE, orc.PROJ_RBSOR, orc.PROJ_JACOBI]


@pytest.mark.parametrize("bits", [64, 32])
@pytest.mark.parametrize("proj", PROJ)
def test_divergence_free_field_is_a_fixed_point(bits, proj):
    dims = (11, 9, 10)
    vel = curl_field(dims, 9)
    assert np.abs(cell_div(*vel)).max() == 0.0
    o = fresh(dims, bits, projection=proj, iterations=7)
    o.add_obstacle(5.0, 4.0, 5.0, 2.5)
    for k in range(3):
        o.set_velocity(k, 0, vel[k])
    o.stage(orc.STAGE_SCENE, 0.05)
    o.stage(orc.STAGE_PROJECT, 0.05)
    for k in range(3):
        np.testing.assert_array_equal(o.get_velocity(k, 0), vel[k])


def scene_after_force(dims, seed, bits=64, **params):
    o = fresh(dims, bits, **params)
    solid = np.zeros(dims[::-1], dtype=np.uint8)
    solid[:, 0, :] = 1
    rng = np.random.default_rng(seed)
    for _ in range(3):
        x, y, z = (int(rng.integers(1, n - 2)) for n in dims)
        solid[z:z + 2, y:y + 2, x:x + 2] = 1
    o.set_solids(solid)
    o.add_obstacle(dims[0] / 2, dims[1] / 2, dims[2] / 2, min(dims) / 5)
    o.set_density(0, random_density(dims, seed))
    for k in range(3):
        o.set_velocity(k, 0, rng.uniform(-1, 1, (dims[2] + 1, dims[1] + 1, dims[0] + 1)))
    o.stage(orc.STAGE_SCENE, 0.05)
    o.stage(orc.STAGE_FORCE, 0.05)
    return o


def test_velocity_form_equals_pressure_form():
    """The reference's in-place SOR on faces is SOR on acc P - sum s P = -div
    followed by u -= grad P: the two agree to rounding in FP64."""
    dims = (13, 12, 11)
    a = scene_after_force(dims, 10, projection=orc.PROJ_REFERENCE, iterations=12)
    b = scene_after_force(dims, 10, projection=orc.PROJ_RBSOR, iterations=12)
    a.stage(orc.STAGE_PROJECT, 0.05)
    b.stage(orc.STAGE_PROJECT, 0.05)
    for k in range(3):
        ua, ub = a.get_velocity(k, 0), b.get_velocity(k, 0)
        np.testing.assert_allclose(ua, ub, rtol=0, atol=1e-12 * np.abs(ua).max())


def direct_solution(o):
    """Dense Gaussian-elimination solve of acc_i P_i - sum_n s_n P_n = -div_i
    over the active cells, and the corrected velocities."""
    flags = o.get_flags()
    mask = o.get_mask()
    now = o.index_now
    u, v, w = (o.get_velocity(k, now) for k in range(3))
    div = cell_div(u, v, w)
    active = (flags & orc.F_ACTIVE) != 0
    idx = -np.ones(mask.shape, dtype=np.int64)
    cells = np.argwhere(active)
    idx[tuple(cells.T)] = np.arange(len(cells))
    n = len(cells)
    A = np.zeros((n, n))
    b = np.zeros(n)
    for r, (z, y, x) in enumerate(cells):
        acc = 0
        for dz, dy, dx in ((0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)):
            zz, yy, xx = z + dz, y + dy, x + dx
            if mask[zz, yy, xx]:
                acc += 1
                if idx[zz, yy, xx] >= 0:
                    A[r, idx[zz, yy, xx]] -= 1.0
        A[r, r] = acc
        b[r] = -div[z, y, x]
    P = np.zeros(mask.shape)
    P[tuple(cells.T)] = np.linalg.solve(A, b)
    s = mask.astype(np.float64)
    u2, v2, w2 = u.copy(), v.copy(), w.copy()
    nz, ny, nx = mask.shape
    u2[:nz, :ny, 1:nx] -= s[:, :, 1:] * s[:, :, :-1] * (P[:, :, 1:] - P[:, :, :-1])
    v2[:nz, 1:ny, :nx] -= s[:, 1:, :] * s[:, :-1, :] * (P[:, 1:, :] - P[:, :-1, :])
    w2[1:nz, :ny, :nx] -= s[1:, :, :] * s[:-1, :, :] * (P[1:, :, :] - P[:-1, :, :])
    return P, (u2, v2, w2), active


@pytest.mark.parametrize("proj,iters", [(orc.PROJ_REFERENCE, 600), (orc.PROJ_RBSOR, 600),
                                        (orc.PROJ_JACOBI, 6000)])
def test_projection_converges_to_direct_solve(proj, iters):
    dims = (8, 7, 6)
    o = scene_after_force(dims, 11, projection=proj, iterations=iters)
    P, vel, active = direct_solution(o)
    o.stage(orc.STAGE_PROJECT, 0.05)
    now = o.index_now
    got = [o.get_velocity(k, now) for k in range(3)]
    for k in range(3):
        np.testing.assert_allclose(got[k], vel[k], rtol=0, atol=1e-9)
    div = cell_div(*got)
    assert np.abs(div[active]).max() < 1e-9
    if proj != orc.PROJ_REFERENCE:
        np.testing.assert_allclose(o.get_pressure(), P, rtol=0, atol=1e-9)


def test_damped_jacobi_converges_to_the_same_solution():
    dims = (7, 6, 6)
    o = scene_after_force(dims, 12, projection=orc.PROJ_JACOBI, iterations=8000, jacobi_omega=0.8)
    P, vel, _ = direct_solution(o)
    o.stage(orc.STAGE_PROJECT, 0.05)
    np.testing.assert_allclose(o.get_pressure(), P, rtol=0, atol=1e-9)


def test_jacobi_matches_one_hand_sweep():
    """One Jacobi sweep from P = 0 is P_i = -d_i / acc_i on active cells."""
    dims = (9, 8, 7)
    o = scene_after_force(dims, 13, projection=orc.PROJ_JACOBI, iterations=1)
    flags = o.get_flags()
    now = o.index_now
    d = cell_div(*(o.get_velocity(k, now) for k in range(3)))
    o.stage(orc.STAGE_PROJECT, 0.05)
    active = (flags & orc.F_ACTIVE) != 0
    acc = np.zeros(flags.shape)
    for bit in (orc.F_XM, orc.F_XP, orc.F_YM, orc.F_YP, orc.F_ZM, orc.F_ZP):
        acc += (flags & bit) != 0
    expect = np.where(active, -d / np.maximum(acc, 1), 0.0)
    np.testing.assert_allclose(o.get_pressure(), expect, rtol=1e-15, atol=1e-16)


# ------------------------------------------------------------- whole steps

def full_scene(dims, bits, proj=orc.PROJ_REFERENCE, iters=20):
    o = fresh(dims, bits, projection=proj, iterations=iters, gravity=7.8, buoyancy=2.0,
              density_gate=0.001, max_velocity=4.0)
    o.add_source(dims[0] / 2, 3.0, dims[2] / 2, 3.0)
    o.add_obstacle(dims[0] / 2, dims[1] / 2, dims[2] / 2, min(dims) / 6)
    rng = np.random.default_rng(14)
    for k in range(3):
        o.set_velocity(k, 0, rng.uniform(-0.01, 0.01, (dims[2] + 1, dims[1] + 1, dims[0] + 1)))
    return o


@pytest.mark.parametrize("proj", PROJ)
def test_step_is_deterministic_across_thread_counts(proj):
    dims = (16, 18, 14)
    out = []
    for threads in (1, 4):
        orc.set_threads(threads)
        o = full_scene(dims, 32, proj)
        o.step(0.05, 8)
        out.append(o.state())
    orc.set_threads(0)
    for k in ("rho0", "rho1", "u0", "u1", "v0", "v1", "w0", "w1"):
        assert np.array_equal(out[0][k], out[1][k]), k


@pytest.mark.parametrize("bits", [64, 32])
def test_density_stays_in_unit_interval(bits):
    o = full_scene((16, 20, 16), bits)
    for _ in range(6):
        o.step(0.05, 5)
        d = o.output_density()
        assert d.min() >= 0.0
        assert d.max() <= 1.0 + 1e-6
    assert d.sum() > 1.0


@pytest.mark.parametrize("proj", PROJ)
def test_fp32_tracks_fp64(proj):
    dims = (16, 20, 16)
    a = full_scene(dims, 64, proj)
    b = full_scene(dims, 32, proj)
    a.step(0.05, 10)
    b.step(0.05, 10)
    now = a.index_now
    da, db = a.output_density(), b.output_density()
    assert np.linalg.norm(da - db) / np.linalg.norm(da) < 1e-4
    for k in range(3):
        va, vb = a.get_velocity(k, 1 - now), b.get_velocity(k, 1 - now)
        assert np.linalg.norm(va - vb) / np.linalg.norm(va) < 1e-3


# ------------------------------------------------------ vorticity confinement
# The oracle's vorticity confinement (an addition: the reference lists it as
# future work, README.md:33) is pinned by closed forms: a velocity whose
# vorticity grows linearly in x gets the force eps * (N x omega) with
# N = +x exactly, and a solid-body rotation (uniform vorticity) gets none.

def _force_stage(o, dt):
    o.stage(orc.STAGE_SCENE, dt)
    o.stage(orc.STAGE_FORCE, dt)


def test_vorticity_confinement_linear_vorticity_closed_form():
    nx, ny, nz = 14, 11, 10
    eps, c, dt = 0.5, 0.01, 0.25
    o = fresh((nx, ny, nz), 64, vorticity=eps, gravity=0.0, buoyancy=0.0, max_velocity=1e6)
    xs = np.arange(nx + 1, dtype=np.float64)
    V = np.broadcast_to(c * (xs + 0.5) ** 2, (nz + 1, ny + 1, nx + 1)).copy()
    o.set_velocity(1, 0, V)
    _force_stage(o, dt)
    got = o.get_velocity(1, 0)
    # omega_z(x) = 2c(x + 1/2) on interior cells, grad|omega| = (2c, 0, 0) away
    # from the walls, so F_y = -eps * omega_z and each interior v-face moves by
    # dt * F_y of its (identical) two cells.
    for x in range(2, nx - 2):
        want = V[3:nz - 3, 3:ny - 3, x] - dt * eps * 2 * c * (x + 0.5)
        np.testing.assert_allclose(got[3:nz - 3, 3:ny - 3, x], want, rtol=0, atol=1e-14)
    for k in (0, 2):  # F_x = F_z = 0 in the interior: u and w untouched there
        assert np.array_equal(o.get_velocity(k, 0)[3:nz - 3, 3:ny - 3, 3:nx - 3], 0 * V[3:nz - 3, 3:ny - 3, 3:nx - 3])


def test_vorticity_confinement_leaves_uniform_vorticity_alone():
    nx, ny, nz = 13, 12, 9
    om = 0.02
    o = fresh((nx, ny, nz), 64, vorticity=0.7, gravity=0.0, buoyancy=0.0, max_velocity=1e6)
    yc, xc = ny / 2, nx / 2
    U = np.zeros((nz + 1, ny + 1, nx + 1))
    V = np.zeros_like(U)
    U[:, :, :] = (-om * (np.arange(ny + 1) + 0.5 - yc))[None, :, None]   # u at (x, y+.5, z+.5)
    V[:, :, :] = (om * (np.arange(nx + 1) + 0.5 - xc))[None, None, :]    # v at (x+.5, y, z+.5)
    o.set_velocity(0, 0, U)
    o.set_velocity(1, 0, V)
    _force_stage(o, 0.5)
    # omega = (0, 0, 2*om) everywhere inside: grad|omega| = 0 away from the
    # walls, N = 0 there, so interior faces keep their values.
    sl = (slice(3, nz - 3), slice(3, ny - 3), slice(3, nx - 3))
    assert np.array_equal(o.get_velocity(0, 0)[sl], U[sl])
    assert np.array_equal(o.get_velocity(1, 0)[sl], V[sl])


def test_vorticity_off_is_the_reference_force_stage():
    dims = (11, 10, 9)
    a = fresh(dims, 32, gravity=7.8, buoyancy=2.0)
    b = fresh(dims, 32, gravity=7.8, buoyancy=2.0, vorticity=0.0)
    rng = np.random.default_rng(5)
    for o in (a, b):
        for k in range(3):
            o.set_velocity(k, 0, np.random.default_rng(k).uniform(-0.1, 0.1, (10, 11, 12)))
        o.set_density(0, rng.uniform(0, 1, (9, 10, 11)))
        rng = np.random.default_rng(5)
        _force_stage(o, 0.05)
    for k in range(3):
        assert np.array_equal(a.get_velocity(k, 0), b.get_velocity(k, 0))


# Moving-obstacle velocity boundary condition (an extension: the reference
# stores obstacle velocities, smokeSimulation.cu:45-52, and never reads them,
# :299-301).  Pinned by a brute-force restatement of which faces an obstacle
# owns (sphere membership evaluated with numpy float32 scalars) and by the
# flow it induces after a converged projection: divergence-free fluid cells
# and the dipole pattern of a sphere pushed through still fluid.

def _sphere_cells(dims, sph):
    nx, ny, nz = dims
    cx, cy, cz, r = (np.float32(t) for t in sph)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    dx = x.astype(np.float32) - cx
    dy = y.astype(np.float32) - cy
    dz = z.astype(np.float32) - cz
    dist = (dx * dx + dy * dy) + dz * dz
    inner = np.zeros(dims[::-1], bool)
    inner[1:-1, 1:-1, 1:-1] = True
    return (dist < r * r) & inner


@pytest.mark.parametrize("mode", [orc.OBST_LAST_WINS, orc.OBST_UNION])
def test_obstacle_velocity_bc_owns_exactly_the_obstacle_faces(mode):
    dims = (13, 11, 10)
    nx, ny, nz = dims
    obst = [((4.0, 5.0, 5.0, 3.0), (0.25, -0.5, 0.75)), ((7.5, 4.5, 4.0, 2.6), (-1.0, 0.5, 2.0))]
    o = fresh(dims, 64, obstacle_velocity=1, obstacle_mode=mode)
    ids = [o.add_obstacle(*sph[:3], sph[3], vel=vel) for sph, vel in obst]
    init = [np.random.default_rng(10 + k).uniform(-1, 1, (nz + 1, ny + 1, nx + 1)) for k in range(3)]
    for k in range(3):
        o.set_velocity(k, 0, init[k])
    o.stage(orc.STAGE_SCENE, 0.05)  # swaps to half 0 and applies the condition there
    now = o.index_now
    assert now == 0
    owner = -np.ones(dims[::-1], int)
    used = obst[-1:] if mode == orc.OBST_LAST_WINS else obst
    base = len(obst) - len(used)
    for j, (sph, _) in enumerate(used):
        owner[_sphere_cells(dims, sph)] = ids[base + j]
    vel = {ids[j]: obst[j][1] for j in range(len(obst))}
    for k in range(3):
        want = init[k].copy()
        for z in range(nz):
            for y in range(ny):
                for x in range(nx):
                    p = (x, y, z)
                    if p[k] < 1:
                        continue
                    lo = list(p)
                    lo[k] -= 1
                    who = owner[z, y, x] if owner[z, y, x] >= 0 else owner[lo[2], lo[1], lo[0]]
                    if who >= 0:
                        want[z, y, x] = np.float32(vel[who][k])
        assert np.array_equal(o.get_velocity(k, now), want), k
    assert np.array_equal(o.get_mask() == 0, (owner >= 0) | (o.get_mask() == 0))


def test_obstacle_velocity_bc_off_leaves_faces_alone():
    dims = (10, 9, 8)
    o = fresh(dims, 64)
    o.add_obstacle(5.0, 4.0, 4.0, 2.5, vel=(1.0, 2.0, 3.0))
    init = np.random.default_rng(3).uniform(-1, 1, (9, 10, 11))
    o.set_velocity(0, 0, init)
    o.stage(orc.STAGE_SCENE, 0.05)
    o.stage(orc.STAGE_SCENE, 0.05)
    assert np.array_equal(o.get_velocity(0, 0), init)


def test_obstacle_velocity_bc_pushes_still_fluid_divergence_free():
    n = 16
    o = fresh((n, n, n), 64, projection=orc.PROJ_REFERENCE, iterations=1500, obstacle_velocity=1,
              gravity=0.0, buoyancy=0.0)
    o.add_obstacle(8.0, 8.0, 8.0, 3.2, vel=(1.0, 0.0, 0.0))
    for st in (orc.STAGE_SCENE, orc.STAGE_FORCE, orc.STAGE_PROJECT):
        o.stage(st, 0.05)
    now = o.index_now
    u, v, w = (o.get_velocity(k, now) for k in range(3))
    act = (o.get_flags() & orc.F_ACTIVE) != 0
    assert np.abs(cell_div(u, v, w)[act]).max() < 1e-10
    # A sphere moving +x through still fluid: fluid ahead and behind moves
    # with it, fluid beside it flows back.
    # Obstacle cells on the centre row are x = 5..11.
    assert u[8, 8, 12] == 1.0 and u[8, 8, 5] == 1.0           # obstacle faces
    assert u[8, 8, 13] > 0.2 and u[8, 8, 4] > 0.2             # ahead, behind
    assert u[8, 12, 8] < -0.05 and u[12, 8, 8] < -0.05 and u[4, 8, 8] < -0.05  # beside
