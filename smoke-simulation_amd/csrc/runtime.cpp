// Host runtime of libsmokesim.so: the C ABI of include/smoke.h over HIP.
//
// One smoke_sim owns one device arena (every field of both ping-pong halves,
// pressure, right-hand side, masks and a staging buffer), one non-blocking
// stream at default priority, the object list and the step parameters.  A
// step is a fixed sequence of kernel launches on that stream; nothing in it
// allocates or synchronises unless the caller asks for the density back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <vector>

#include "sim_state.h"

using smk::Grid;
using smk::fail;
using smk::Object;

namespace {
thread_local std::string g_last_error;
}  // namespace

smoke_status smk::fail(smoke_status st, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return st;
}

namespace {




size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

Grid make_grid(unsigned nx, unsigned ny, unsigned nz)
{
    Grid g;
    g.nx = (int)nx;
    g.ny = (int)ny;
    g.nz = (int)nz;
    g.px = (int)((nx + 1 + 31) / 32 * 32);
    g.py = (int)ny + 1;
    g.plane = g.px * g.py;
    g.elems = g.plane * ((int)nz + 1);
    return g;
}

constexpr int kStoredFields = 8;  // rho0 rho1 u0 u1 v0 v1 w0 w1 (FP32 or 16-bit)
constexpr int kFloatFields = 3;   // P0 P1 d (always FP32)
constexpr int kByteFields = 3;    // base s flags

bool storage_ok(int storage)
{
    return storage >= SMOKE_STORAGE_FP32 && storage <= SMOKE_STORAGE_FP16_VELOCITY;
}
// Bytes per element of the stored density and velocity fields.
size_t density_bytes(int storage)
{
    return storage == SMOKE_STORAGE_BF16 || storage == SMOKE_STORAGE_FP16 ? 2 : 4;
}
size_t velocity_bytes(int storage) { return storage == SMOKE_STORAGE_FP32 ? 4 : 2; }

// Arena: stored fields, P0 P1 d, staging (reference-layout faces, FP32), then
// the byte fields.
size_t arena_bytes(const Grid& g, int storage)
{
    size_t e = (size_t)g.elems;
    size_t staging = (size_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1) * sizeof(float);
    return 2 * align256(e * density_bytes(storage)) + 6 * align256(e * velocity_bytes(storage)) +
           kFloatFields * align256(e * sizeof(float)) + align256(staging) +
           kByteFields * align256(e);
}

// REFINED buffers: q (FP32), e0 e1 qi inv (FP16), control block, partials.
size_t refine_bytes(const Grid& g)
{
    size_t e = (size_t)g.elems;
    size_t parts = (size_t)smk::refine_parts(g) + smk::kRefineStage2;
    return align256(e * sizeof(float)) + 3 * align256(e * 2) + align256(sizeof(smk::RefineCtl)) +
           align256(parts * sizeof(double)) + align256(parts * sizeof(unsigned));
}

bool dims_ok(unsigned nx, unsigned ny, unsigned nz)
{
    if (nx < 3 || ny < 3 || nz < 3) return false;
    // 32-bit element indices in the kernels.
    double elems = (double)((nx + 1 + 31) / 32 * 32) * (ny + 1) * (nz + 1);
    return elems < 2147483647.0;
}

smoke_status check_params(const smoke_params* p)
{
    if (!p) return fail(SMOKE_ERR_INVALID, "params: NULL");
    if (p->projection < SMOKE_PROJ_REFERENCE || p->projection > SMOKE_PROJ_REFINED)
        return fail(SMOKE_ERR_INVALID, "params: unknown projection %d", p->projection);
    if (p->projection == SMOKE_PROJ_REFINED) {
        if (p->inner_sweeps < 2 || p->inner_sweeps > 1000 || (p->inner_sweeps & 1))
            return fail(SMOKE_ERR_INVALID, "params: inner_sweeps %d not an even number in 2..1000",
                        p->inner_sweeps);
        if (!std::isfinite(p->tolerance) || p->tolerance < 0.0)
            return fail(SMOKE_ERR_INVALID, "params: tolerance must be finite and >= 0");
        if (p->warm_start)
            return fail(SMOKE_ERR_INVALID, "params: the REFINED projection starts cold (warm_start 0)");
        if (p->iterations < 1)
            return fail(SMOKE_ERR_INVALID, "params: REFINED needs iterations >= 1");
    }
    if (p->obstacle_velocity != 0 && p->obstacle_velocity != 1)
        return fail(SMOKE_ERR_INVALID, "params: obstacle_velocity must be 0 or 1");
    if (!std::isfinite(p->vorticity) || p->vorticity < 0.0f)
        return fail(SMOKE_ERR_INVALID, "params: vorticity must be finite and >= 0");
    if (p->obstacle_mode != SMOKE_OBSTACLES_LAST_WINS && p->obstacle_mode != SMOKE_OBSTACLES_UNION)
        return fail(SMOKE_ERR_INVALID, "params: unknown obstacle_mode %d", p->obstacle_mode);
    if (p->iterations < 0 || p->iterations > 1000000)
        return fail(SMOKE_ERR_INVALID, "params: iterations %d out of range", p->iterations);
    if (!std::isfinite(p->sor_omega) || !std::isfinite(p->jacobi_omega))
        return fail(SMOKE_ERR_INVALID, "params: omega must be finite");
    if (!std::isfinite(p->gravity) || !std::isfinite(p->buoyancy) || !std::isfinite(p->max_velocity) ||
        !std::isfinite(p->density_gate))
        return fail(SMOKE_ERR_INVALID, "params: physics parameters must be finite");
    return SMOKE_OK;
}

}  // namespace


namespace {

constexpr size_t kMaxStepHistory = 1 << 20;

smk::SphereList sphere_list(const smoke_sim* s, int type)
{
    smk::SphereList l;
    l.count = 0;
    for (const Object& o : s->objects)
        if (o.type == type) l.s[l.count++] = make_float4(o.pos[0], o.pos[1], o.pos[2], o.r);
    return l;
}

int count_type(const smoke_sim* s, int type)
{
    int n = 0;
    for (const Object& o : s->objects) n += (o.type == type);
    return n;
}

// Rebuilds mask and flags from the base mask and the obstacles.  The mask is a
// pure function of those two, so it is recomputed only when one changed; the
// reference recomputes it every step (drawObjects, smokeSimulation.cu:766).

}  // namespace

hipError_t smk::refresh_mask(smoke_sim* s, long long* launches)
{
    if (!s->mask_dirty) return hipSuccess;
    smk::SphereList obst = sphere_list(s, 0);
    hipError_t e = smk::build_mask(s->g, s->base, s->s, obst, s->prm.obstacle_mode, s->stream);
    if (e != hipSuccess) return e;
    e = smk::build_flags(s->g, s->s, s->flags, s->stream);
    if (e != hipSuccess) return e;
    if (launches) *launches += 2;
    if (s->owner) {
        e = smk::build_owner(s->g, s->owner, obst, s->prm.obstacle_mode, s->stream);
        if (e != hipSuccess) return e;
        if (launches) *launches += 1;
    }
    s->mask_dirty = false;
    return hipSuccess;
}

namespace {

// Allocates and clears the REFINED buffers (once per simulation).
smoke_status ensure_refine(smoke_sim* s)
{
    // The FP16 sweeps address a block's planes with 32-bit buffer offsets
    // (as the FP32 fused sweeps do) and have no one-sweep fallback.
    if (smk::fused_max_zc(s->g.plane) < 1)
        return fail(SMOKE_ERR_INVALID, "refined solver: planes of %d cells are too wide (at most %lld)",
                    s->g.plane, 2147483647LL / (4LL * (smk::kFusedSlackPlanes + 1)));
    if (s->refine_mem) return SMOKE_OK;
    const Grid& g = s->g;
    const size_t bytes = refine_bytes(g);
    char* m = nullptr;
    if (hipMalloc((void**)&m, bytes) != hipSuccess)
        return fail(SMOKE_ERR_NOMEM, "refined solver: cannot allocate %.2f GiB", bytes / 1073741824.0);
    const size_t e = (size_t)g.elems;
    char* p = m;
    smk::RefineBufs& b = s->rb;
    b.q = reinterpret_cast<float*>(p);
    p += align256(e * sizeof(float));
    b.e[0] = reinterpret_cast<uint32_t*>(p);
    p += align256(e * 2);
    b.e[1] = reinterpret_cast<uint32_t*>(p);
    p += align256(e * 2);
    b.qi = reinterpret_cast<uint32_t*>(p);
    p += align256(e * 2);
    b.ctl = reinterpret_cast<smk::RefineCtl*>(p);
    p += align256(sizeof(smk::RefineCtl));
    b.part = reinterpret_cast<double*>(p);
    p += align256(((size_t)smk::refine_parts(g) + smk::kRefineStage2) * sizeof(double));
    b.pmax = reinterpret_cast<unsigned*>(p);
    s->refine_mem = m;
    hipError_t err = smk::fill_u8(reinterpret_cast<uint8_t*>(m), (long long)bytes, 0, s->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(s->stream);
    if (err != hipSuccess) return fail(SMOKE_ERR_HIP, "refined solver: init failed: %s", hipGetErrorString(err));
    return SMOKE_OK;
}

// Allocates the owner bytes of the moving-obstacle boundary condition (once
// per simulation); the next mask refresh builds them.
smoke_status ensure_owner(smoke_sim* s)
{
    if (s->owner) return SMOKE_OK;
    if (hipMalloc((void**)&s->owner, (size_t)s->g.elems) != hipSuccess) {
        s->owner = nullptr;
        return fail(SMOKE_ERR_NOMEM, "obstacle_velocity: cannot allocate %d bytes", s->g.elems);
    }
    s->mask_dirty = true;
    return SMOKE_OK;
}

// Allocates the vorticity-confinement scratch (once per simulation): four
// FP32 padded fields, every cell of which the curl pass writes before it is
// read, so nothing needs clearing.
smoke_status ensure_vort(smoke_sim* s)
{
    if (s->vort_mem) return SMOKE_OK;
    const size_t field = align256((size_t)s->g.elems * sizeof(float));
    char* m = nullptr;
    if (hipMalloc((void**)&m, 4 * field) != hipSuccess)
        return fail(SMOKE_ERR_NOMEM, "vorticity: cannot allocate %.2f GiB", 4 * field / 1073741824.0);
    s->vort_mem = m;
    s->vb.mag = reinterpret_cast<float*>(m);
    for (int k = 0; k < 3; ++k) s->vb.f[k] = reinterpret_cast<float*>(m + (k + 1) * field);
    return SMOKE_OK;
}

// Allocates the step-norms records (once per simulation): the reduction's
// scratch, then the public record it publishes.
size_t norms_bytes() { return align256(smk::norms_scratch_bytes()) + sizeof(smoke_step_norms); }
smoke_step_norms* norms_record(smoke_sim* s)
{
    return reinterpret_cast<smoke_step_norms*>(static_cast<char*>(s->norms_mem) +
                                               align256(smk::norms_scratch_bytes()));
}
smoke_status ensure_norms(smoke_sim* s)
{
    if (s->norms_mem) return SMOKE_OK;
    if (hipMalloc(&s->norms_mem, norms_bytes()) != hipSuccess) {
        s->norms_mem = nullptr;
        return fail(SMOKE_ERR_NOMEM, "norms: cannot allocate %zu bytes", norms_bytes());
    }
    return SMOKE_OK;
}

}  // namespace

// Enqueues the norms reduction of the current state -- divergence of half
// [now], output half [1 - now], right-hand side d if the last projection
// left one -- with `scratch` for the partial records and the public record
// written to `out` (device memory).
hipError_t smk::enqueue_norms(smoke_sim* s, void* scratch, smoke_step_norms* out)
{
    const int n = s->now, p = 1 - s->now;
    return smk::with_fields(s, [&](auto rho, auto u, auto v, auto w) {
        return smk::step_norms(s->g, rho[p], u[p], v[p], w[p], u[n], v[n], w[n], s->d,
                               s->rhs_valid ? 1 : 0, s->flags, scratch, out, s->last_dt,
                               s->steps_enqueued, s->stream);
    });
}

namespace {

hipError_t enqueue_own_norms(smoke_sim* s)
{
    return smk::enqueue_norms(s, s->norms_mem, norms_record(s));
}

}  // namespace

namespace {

using smk::refresh_mask;

hipError_t stage_scene(smoke_sim* s, long long* launches)
{
    s->now = 1 - s->now;  // smokeSimulation.cu:777-779
    for (const Object& o : s->objects) {
        if (o.type != 1) continue;
        int launched = 0;
        hipError_t e = smk::with_fields(s, [&](auto rho, auto, auto, auto) {
            return smk::fill_source(s->g, rho[0], rho[1],
                                    make_float4(o.pos[0], o.pos[1], o.pos[2], o.r), s->stream,
                                    &launched);
        });
        if (e != hipSuccess) return e;
        *launches += launched;
    }
    hipError_t e = refresh_mask(s, launches);
    if (e != hipSuccess || !s->prm.obstacle_velocity) return e;
    if (!s->owner) return hipErrorNotInitialized;  // set_params allocates it
    smk::SphereList vel;
    vel.count = 0;
    for (const Object& o : s->objects)
        if (o.type == 0) vel.s[vel.count++] = make_float4(o.vel[0], o.vel[1], o.vel[2], 0.0f);
    *launches += 1;
    const int n = s->now;
    return smk::with_fields(s, [&](auto, auto u, auto v, auto w) {
        return smk::obstacle_velocity_bc(s->g, u[n], v[n], w[n], s->owner, vel, s->stream);
    });
}

hipError_t stage_force(smoke_sim* s, float dt, long long* launches)
{
    smk::ForceArgs a;
    a.gravity = s->prm.gravity;
    a.buoyancy = s->prm.buoyancy;
    a.dt = dt;
    a.max_v2 = s->prm.max_velocity * s->prm.max_velocity;
    a.gate = s->prm.density_gate;
    a.gate_on = s->prm.density_gate > 0.0f;
    const int n = s->now;
    const float eps = s->prm.vorticity;
    if (eps > 0.0f && !s->vort_mem) return hipErrorNotInitialized;  // set_params allocates it
    *launches += eps > 0.0f ? 4 : 1;
    return smk::with_fields(s, [&](auto rho, auto u, auto v, auto w) {
        hipError_t e = smk::force(s->g, u[n], v[n], w[n], rho[n], s->flags, a, s->stream);
        if (e != hipSuccess || !(eps > 0.0f)) return e;
        return smk::vorticity(s->g, u[n], v[n], w[n], s->flags, s->vb, eps, dt, s->stream);
    });
}

// REFINED: R refinements of inner_sweeps FP16 sweeps (launches of an even
// number of fused sweeps), each closed by the FP32 update/residual pass; the
// result ends in P[pcur].
hipError_t enqueue_refined(smoke_sim* s, long long* launches, long long* sweep_launches)
{
    const Grid& g = s->g;
    const smoke_params& p = s->prm;
    const smk::RefineBufs& b = s->rb;
    const int R = (p.iterations + p.inner_sweeps - 1) / p.inner_sweeps;
    int T = s->plan.sweeps_per_launch;
    if (T < 2 || T > 6 || (T & 1)) T = 4;
    const int two_rows = (s->plan.tile == smk::JacobiPlan::TILE_64x32 ||
                          s->plan.tile == smk::JacobiPlan::TILE_32x64) ? 1 : 0;
    hipError_t e = smk::refine_norm0(g, s->d, s->flags, b, p.tolerance, s->stream);
    if (e != hipSuccess) return e;
    *launches += 2;
    int a = s->pcur;
    for (int r = 0; r < R; ++r) {
        int ein = 0;
        bool first = true;
        for (int left = p.inner_sweeps; left > 0;) {
            const int t = std::min(left, T);
            const int eout = first ? 0 : 1 - ein;
            e = smk::refine_sweeps(g, b, r == 0 ? s->d : b.q, s->flags, first ? 1 : 0, ein, eout, t,
                                   s->plan.zc, two_rows, s->stream);
            if (e != hipSuccess) return e;
            ein = eout;
            first = false;
            left -= t;
            *launches += 1;
            *sweep_launches += 1;
        }
        e = smk::refine_resid(g, s->P[a], s->P[a ^ 1], s->d, s->flags, b, ein, r == 0 ? 1 : 0,
                              r == R - 1 ? 1 : 0, r + 1, p.tolerance, s->stream);
        if (e != hipSuccess) return e;
        *launches += 2;
        a ^= 1;
    }
    s->pcur = a;
    return hipSuccess;
}

// K Jacobi sweeps as launches of plan.sweeps_per_launch fused sweeps each, the
// remainder in one shorter launch; ping-pongs the two P buffers.
hipError_t enqueue_jacobi(smoke_sim* s, const smk::JacobiPlan& plan, int sweeps, bool zero_in,
                          long long* launches)
{
    const bool plain = s->prm.jacobi_omega == 1.0;
    // A plane too large for one fused block's 32-bit buffer span: one sweep
    // per launch (same bits).
    const int per_launch = smk::fused_max_zc(s->g.plane) < 1 ? 1 : plan.sweeps_per_launch;
    for (int left = sweeps; left > 0;) {
        const int t = std::min(left, per_launch);
        const int out = 1 - s->pcur;
        hipError_t e = smk::jacobi_fused(s->g, s->P[s->pcur], s->P[out], s->d, s->flags, plan, t,
                                         (float)s->prm.jacobi_omega, plain ? 1 : 0, zero_in ? 1 : 0,
                                         s->stream);
        if (e != hipSuccess) return e;
        *launches += 1;
        s->pcur = out;
        zero_in = false;
        left -= t;
    }
    return hipSuccess;
}

// Records the begin (0) or end (1) event of the sweep loop of a timed step.
hipError_t sweep_mark(smoke_sim* s, int which, long long sweeps)
{
    if (!s->cur) return hipSuccess;
    if (which == 1) s->cur->sweeps = sweeps;
    return hipEventRecord(s->cur->sweep_ev[which], s->stream);
}

hipError_t stage_project(smoke_sim* s, long long* launches)
{
    const Grid& g = s->g;
    const int n = s->now;
    const smoke_params& p = s->prm;
    hipError_t e = hipSuccess;
    if (p.projection == SMOKE_PROJ_REFERENCE) {
        // The in-place form has no right-hand side of its own; with norms on,
        // the divergence before the sweeps is kept for rhs_l2.
        s->rhs_valid = s->norms_on;
        if (s->norms_on) {
            e = smk::divergence(g, s->u[n], s->v[n], s->w[n], s->flags, s->d, s->stream);
            if (e != hipSuccess) return e;
            *launches += 1;
        }
        if ((e = sweep_mark(s, 0, 0)) != hipSuccess) return e;
        for (int it = 0; it < p.iterations; ++it)
            for (int o = 0; o < 2; ++o) {
                e = smk::project_reference_half(g, s->u[n], s->v[n], s->w[n], s->flags, o,
                                                p.sor_omega, s->stream);
                if (e != hipSuccess) return e;
                *launches += 1;
            }
        return sweep_mark(s, 1, 2LL * p.iterations);
    }
    e = smk::with_fields(s, [&](auto, auto u, auto v, auto w) {
        return smk::divergence(g, u[n], v[n], w[n], s->flags, s->d, s->stream);
    });
    if (e != hipSuccess) return e;
    *launches += 1;
    s->rhs_valid = true;
    if (p.projection == SMOKE_PROJ_REFINED) {
        if ((e = sweep_mark(s, 0, 0)) != hipSuccess) return e;
        long long sweep_launches = 0;
        if ((e = enqueue_refined(s, launches, &sweep_launches)) != hipSuccess) return e;
        if ((e = sweep_mark(s, 1, sweep_launches)) != hipSuccess) return e;
    } else if (p.projection == SMOKE_PROJ_JACOBI) {
        const bool zero_in = !p.warm_start;
        if (p.iterations == 0 && zero_in) {  // P = 0: no correction
            if ((e = sweep_mark(s, 0, 0)) != hipSuccess) return e;
            return sweep_mark(s, 1, 0);
        }
        if (!zero_in) {
            e = smk::mask_pressure(g, s->P[s->pcur], s->flags, s->stream);
            if (e != hipSuccess) return e;
            *launches += 1;
        }
        if ((e = sweep_mark(s, 0, 0)) != hipSuccess) return e;
        long long sweep_launches = 0;
        e = enqueue_jacobi(s, s->plan, p.iterations, zero_in, &sweep_launches);
        if (e != hipSuccess) return e;
        *launches += sweep_launches;
        if ((e = sweep_mark(s, 1, sweep_launches)) != hipSuccess) return e;
    } else {
        if (p.warm_start) e = smk::mask_pressure(g, s->P[s->pcur], s->flags, s->stream);
        else e = smk::fill_f32(s->P[s->pcur], g.elems, 0.0f, s->stream);
        if (e != hipSuccess) return e;
        *launches += 1;
        if ((e = sweep_mark(s, 0, 0)) != hipSuccess) return e;
        for (int it = 0; it < p.iterations; ++it)
            for (int o = 0; o < 2; ++o) {
                e = smk::rbsor_half(g, s->P[s->pcur], s->d, s->flags, o, (float)p.sor_omega, s->stream);
                if (e != hipSuccess) return e;
                *launches += 1;
            }
        if ((e = sweep_mark(s, 1, 2LL * p.iterations)) != hipSuccess) return e;
    }
    e = smk::with_fields(s, [&](auto, auto u, auto v, auto w) {
        return smk::gradient(g, u[n], v[n], w[n], s->P[s->pcur], s->flags, s->stream);
    });
    *launches += 1;
    return e;
}

hipError_t stage_advect_velocity(smoke_sim* s, float dt, long long* launches)
{
    const int n = s->now, p = 1 - s->now;
    *launches += 1;
    return smk::with_fields(s, [&](auto, auto u, auto v, auto w) {
        return smk::advect_velocity(s->g, u[n], v[n], w[n], u[p], v[p], w[p], s->flags, dt,
                                    s->stream);
    });
}

hipError_t stage_advect_density(smoke_sim* s, float dt, long long* launches)
{
    const int n = s->now, p = 1 - s->now;
    *launches += 1;
    hipError_t e = smk::with_fields(s, [&](auto rho, auto u, auto v, auto w) {
        return smk::advect_density(s->g, rho[n], rho[p], u[p], v[p], w[p], s->flags, dt, s->stream);
    });
    if (e != hipSuccess || !s->norms_on) return e;
    // The step's norms, reduced in-stream after its last stage (the record's
    // step number is the one this step will have).
    if (!s->norms_mem) return hipErrorNotInitialized;  // smoke_norms_enable allocates it
    *launches += 3;
    s->steps_enqueued += 1;
    e = enqueue_own_norms(s);
    s->steps_enqueued -= 1;
    return e;
}

hipError_t run_stage(smoke_sim* s, int stage, float dt, long long* launches)
{
    switch (stage) {
    case SMOKE_STAGE_SCENE: return stage_scene(s, launches);
    case SMOKE_STAGE_FORCE: return stage_force(s, dt, launches);
    case SMOKE_STAGE_PROJECT: return stage_project(s, launches);
    case SMOKE_STAGE_ADVECT_VEL: return stage_advect_velocity(s, dt, launches);
    case SMOKE_STAGE_ADVECT_DENS: return stage_advect_density(s, dt, launches);
    default: return hipErrorInvalidValue;
    }
}

hipError_t get_event(smoke_sim* s, hipEvent_t* ev)
{
    if (!s->event_pool.empty()) {
        *ev = s->event_pool.back();
        s->event_pool.pop_back();
        return hipSuccess;
    }
    return hipEventCreate(ev);
}

// Folds finished timed steps into the totals (all of them if `wait`).
hipError_t collect_timing(smoke_sim* s, bool wait)
{
    size_t done = 0;
    for (; done < s->pending.size(); ++done) {
        smoke_sim::StepEvents& se = s->pending[done];
        if (wait) {
            hipError_t e = hipEventSynchronize(se.ev[SMOKE_STAGE_COUNT]);
            if (e != hipSuccess) return e;
        } else if (hipEventQuery(se.ev[SMOKE_STAGE_COUNT]) != hipSuccess) {
            break;
        }
        double total = 0.0;
        for (int k = 0; k < SMOKE_STAGE_COUNT; ++k) {
            float ms = 0.0f;
            hipError_t e = hipEventElapsedTime(&ms, se.ev[k], se.ev[k + 1]);
            if (e != hipSuccess) return e;
            s->stage_ms[k] += ms;
            s->stage_launches[k] += se.launches[k];
            total += ms;
        }
        float sweep = 0.0f;
        hipError_t e = hipEventElapsedTime(&sweep, se.sweep_ev[0], se.sweep_ev[1]);
        if (e != hipSuccess) return e;
        s->sweep_ms += sweep;
        s->sweep_launches += se.sweeps;
        s->timed_steps += 1;
        s->step_ms.push_back(total);
        if (s->step_ms.size() > kMaxStepHistory) s->step_ms.pop_front();
        for (int k = 0; k <= SMOKE_STAGE_COUNT; ++k) s->event_pool.push_back(se.ev[k]);
        s->event_pool.push_back(se.sweep_ev[0]);
        s->event_pool.push_back(se.sweep_ev[1]);
    }
    s->pending.erase(s->pending.begin(), s->pending.begin() + (long)done);
    return hipSuccess;
}

void drop_step_graphs(smoke_sim* s)
{
    for (smoke_sim::StepGraph& sg : s->graphs) {
        if (sg.exec) (void)hipGraphExecDestroy(sg.exec);
        if (sg.graph) (void)hipGraphDestroy(sg.graph);
        sg = smoke_sim::StepGraph{};
    }
}

// One step as a replay of the graph captured for the current parity (captured
// first if there is none for this version, dt and pressure parity).  The mask
// refresh, when due, runs before it, outside the graph.
smoke_status step_graph(smoke_sim* s, float dt)
{
    HIP_TRY(refresh_mask(s, nullptr));
    smoke_sim::StepGraph& sg = s->graphs[s->now];
    if (!sg.exec || sg.version != s->version || sg.dt != dt || sg.pcur_in != s->pcur) {
        if (sg.exec) {
            HIP_TRY(hipStreamSynchronize(s->stream));  // no replay of it still running
            (void)hipGraphExecDestroy(sg.exec);
        }
        if (sg.graph) (void)hipGraphDestroy(sg.graph);
        sg = smoke_sim::StepGraph{};
        const int now0 = s->now, pcur0 = s->pcur;
        const bool rhs0 = s->rhs_valid;
        HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        for (int st = 0; st < SMOKE_STAGE_COUNT && e == hipSuccess; ++st) {
            long long launches = 0;
            e = run_stage(s, st, dt, &launches);
        }
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(s->stream, &g);
        sg.now_out = s->now;
        sg.pcur_out = s->pcur;
        sg.rhs_out = s->rhs_valid;
        s->now = now0;  // the replay below applies the step's parities
        s->pcur = pcur0;
        s->rhs_valid = rhs0;
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            sg = smoke_sim::StepGraph{};
            return fail(SMOKE_ERR_HIP, "step: graph capture failed: %s",
                        hipGetErrorString(e != hipSuccess ? e : e2));
        }
        sg.graph = g;
        const hipError_t e3 = hipGraphInstantiate(&sg.exec, g, nullptr, nullptr, 0);
        if (e3 != hipSuccess) {
            (void)hipGraphDestroy(g);
            sg = smoke_sim::StepGraph{};
            return fail(SMOKE_ERR_HIP, "step: graph instantiation failed: %s", hipGetErrorString(e3));
        }
        sg.version = s->version;
        sg.dt = dt;
        sg.pcur_in = pcur0;
    }
    HIP_TRY(hipGraphLaunch(sg.exec, s->stream));
    s->now = sg.now_out;
    s->pcur = sg.pcur_out;
    s->rhs_valid = sg.rhs_out;
    s->last_dt = dt;
    s->steps_enqueued += 1;
    return SMOKE_OK;
}

smoke_status step_locked(smoke_sim* s, float dt)
{
    SMOKE_REFUSE_IN_BATCH(s, "step");
    if (!std::isfinite(dt)) return fail(SMOKE_ERR_INVALID, "step: dt must be finite");
    if (s->graph_on && !s->timing && !s->norms_on) return step_graph(s, dt);
    smoke_sim::StepEvents se{};
    if (s->timing) {
        if (s->pending.size() >= 512) HIP_TRY(collect_timing(s, false));
        for (int k = 0; k <= SMOKE_STAGE_COUNT; ++k) HIP_TRY(get_event(s, &se.ev[k]));
        HIP_TRY(get_event(s, &se.sweep_ev[0]));
        HIP_TRY(get_event(s, &se.sweep_ev[1]));
        HIP_TRY(hipEventRecord(se.ev[0], s->stream));
        s->cur = &se;
    }
    s->last_dt = dt;
    for (int st = 0; st < SMOKE_STAGE_COUNT; ++st) {
        long long launches = 0;
        hipError_t e = run_stage(s, st, dt, &launches);
        if (e != hipSuccess) {
            s->cur = nullptr;
            return fail(SMOKE_ERR_HIP, "step: stage %d launch failed: %s", st, hipGetErrorString(e));
        }
        if (s->timing) {
            se.launches[st] = launches;
            HIP_TRY(hipEventRecord(se.ev[st + 1], s->stream));
        }
    }
    if (s->timing) s->pending.push_back(se);
    s->steps_enqueued += 1;
    if (s->norms_on) s->norms_step = s->steps_enqueued;
    s->cur = nullptr;
    return SMOKE_OK;
}

// Blocking device->host copy of a padded cell field through the staging
// buffer.
template <typename T>
smoke_status read_cells_f32(smoke_sim* s, const T* src, float* dst)
{
    SMOKE_WAIT_BATCH(s);
    HIP_TRY(smk::pack_cells_f32(s->g, src, s->staging, s->stream));
    size_t bytes = (size_t)s->g.nx * s->g.ny * s->g.nz * sizeof(float);
    HIP_TRY(hipMemcpyAsync(dst, s->staging, bytes, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SMOKE_OK;
}

}  // namespace

// =================================================================== C ABI

extern "C" {

const char* smoke_last_error(void) { return g_last_error.c_str(); }

const char* smoke_version(void) { return "smokesim 0.2 (gfx950, FP32 and BF16/FP16 mixed)"; }

smoke_status smoke_device_info(char* buf, size_t len)
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) return fail(SMOKE_ERR_HIP, "no HIP device visible");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    if (buf && len)
        snprintf(buf, len, "device %d: %s (%s), %d CUs, %.1f GiB, LDS/block %zu B, %d devices", dev,
                 prop.name, prop.gcnArchName, prop.multiProcessorCount,
                 prop.totalGlobalMem / 1073741824.0, prop.sharedMemPerBlock, count);
    return SMOKE_OK;
}

smoke_status smoke_params_default(smoke_params* out)
{
    if (!out) return fail(SMOKE_ERR_INVALID, "params_default: NULL");
    out->gravity = -9.82f;
    out->buoyancy = 2.0f;
    out->density_gate = 0.0f;
    out->max_velocity = 3.0f;
    out->sor_omega = 1.9;
    out->jacobi_omega = 1.0;
    out->projection = SMOKE_PROJ_REFERENCE;
    out->iterations = 30;
    out->obstacle_mode = SMOKE_OBSTACLES_LAST_WINS;
    out->warm_start = 0;
    out->inner_sweeps = 20;
    out->vorticity = 0.0f;
    out->obstacle_velocity = 0;
    out->tolerance = 0.0;
    return SMOKE_OK;
}

size_t smoke_footprint_bytes(unsigned nx, unsigned ny, unsigned nz)
{
    return smoke_footprint_bytes_ex(nx, ny, nz, SMOKE_STORAGE_FP32, 0);
}

size_t smoke_footprint_bytes_ex(unsigned nx, unsigned ny, unsigned nz, int storage, int refined)
{
    if (!dims_ok(nx, ny, nz) || !storage_ok(storage)) return 0;
    const Grid g = make_grid(nx, ny, nz);
    return arena_bytes(g, storage) + (refined ? refine_bytes(g) : 0);
}

int smoke_get_storage(const smoke_sim* s) { return s ? s->storage : -1; }

smoke_status smoke_create(unsigned nx, unsigned ny, unsigned nz, const float* density,
                          smoke_sim** out)
{
    return smoke_create_ex(nx, ny, nz, density, SMOKE_STORAGE_FP32, out);
}

smoke_status smoke_create_ex(unsigned nx, unsigned ny, unsigned nz, const float* density,
                             int storage, smoke_sim** out)
{
    if (!out) return fail(SMOKE_ERR_INVALID, "create: out is NULL");
    *out = nullptr;
    if (!dims_ok(nx, ny, nz))
        return fail(SMOKE_ERR_INVALID, "create: dims %ux%ux%u invalid (each >= 3, padded size < 2^31)",
                    nx, ny, nz);
    if (!storage_ok(storage)) return fail(SMOKE_ERR_INVALID, "create: unknown storage %d", storage);
    smoke_sim* s = new smoke_sim();
    s->g = make_grid(nx, ny, nz);
    s->storage = storage;
    smoke_params_default(&s->prm);
    // 16-bit velocity cannot run the face-based reference projection.
    if (storage != SMOKE_STORAGE_FP32) s->prm.projection = SMOKE_PROJ_JACOBI;
    auto bail = [&](smoke_status st) {
        smoke_destroy(s);
        return st;
    };
    hipError_t e = hipGetDevice(&s->device);
    if (e != hipSuccess) return bail(fail(SMOKE_ERR_HIP, "create: no device: %s", hipGetErrorString(e)));
    e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(fail(SMOKE_ERR_HIP, "create: stream: %s", hipGetErrorString(e)));
    size_t bytes = arena_bytes(s->g, storage);
    e = hipMalloc((void**)&s->arena, bytes);
    if (e != hipSuccess) {
        s->arena = nullptr;
        return bail(fail(SMOKE_ERR_NOMEM, "create: cannot allocate %.2f GiB: %s", bytes / 1073741824.0,
                         hipGetErrorString(e)));
    }
    const size_t fbytes = align256((size_t)s->g.elems * sizeof(float));
    char* p = s->arena;
    for (int k = 0; k < kStoredFields; ++k) {
        const int h = k & 1, field = k >> 1;  // rho, u, v, w
        const bool half = field == 0 ? density_bytes(storage) == 2 : velocity_bytes(storage) == 2;
        const bool fp16 = storage == SMOKE_STORAGE_FP16 || storage == SMOKE_STORAGE_FP16_VELOCITY;
        if (half && fp16) {
            smk::f16s** f[4] = {s->frho, s->fu, s->fv, s->fw};
            f[field][h] = reinterpret_cast<smk::f16s*>(p);
        } else if (half) {
            smk::bf16s** f[4] = {s->hrho, s->hu, s->hv, s->hw};
            f[field][h] = reinterpret_cast<smk::bf16s*>(p);
        } else {
            float** f[4] = {s->rho, s->u, s->v, s->w};
            f[field][h] = reinterpret_cast<float*>(p);
        }
        p += align256((size_t)s->g.elems * (half ? 2 : 4));
    }
    float** fields[kFloatFields] = {&s->P[0], &s->P[1], &s->d};
    for (int k = 0; k < kFloatFields; ++k) {
        *fields[k] = reinterpret_cast<float*>(p);
        p += fbytes;
    }
    s->staging = reinterpret_cast<float*>(p);
    p += align256((size_t)(nx + 1) * (ny + 1) * (nz + 1) * sizeof(float));
    uint8_t** bfields[kByteFields] = {&s->base, &s->s, &s->flags};
    for (int k = 0; k < kByteFields; ++k) {
        *bfields[k] = reinterpret_cast<uint8_t*>(p);
        p += align256((size_t)s->g.elems);
    }
    // Zero everything (the reference leaves half [1] uninitialised; zero is
    // the defined value here; all-zero bits are +0 in FP32, BF16 and FP16), then
    // the floor mask and the initial density.
    e = smk::fill_u8(reinterpret_cast<uint8_t*>(s->arena), (long long)bytes, 0, s->stream);
    if (e == hipSuccess) e = smk::floor_mask(s->g, s->base, s->stream);
    if (e == hipSuccess && density) {
        size_t db = (size_t)nx * ny * nz * sizeof(float);
        e = hipMemcpyAsync(s->staging, density, db, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess)
            e = smk::with_fields(s, [&](auto rho, auto, auto, auto) {
                return smk::unpack_cells_f32(s->g, s->staging, rho[0], s->stream);
            });
    }
    if (e == hipSuccess) e = refresh_mask(s, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return bail(fail(SMOKE_ERR_HIP, "create: init failed: %s", hipGetErrorString(e)));
    *out = s;
    return SMOKE_OK;
}

void smoke_destroy(smoke_sim* s)
{
    if (!s) return;
    if (s->batch) smk::batch_detach(s->batch, s);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto& se : s->pending) {
        for (int k = 0; k <= SMOKE_STAGE_COUNT; ++k) (void)hipEventDestroy(se.ev[k]);
        (void)hipEventDestroy(se.sweep_ev[0]);
        (void)hipEventDestroy(se.sweep_ev[1]);
    }
    for (hipEvent_t ev : s->event_pool) (void)hipEventDestroy(ev);
    drop_step_graphs(s);
    if (s->copy_stream) (void)hipStreamSynchronize(s->copy_stream);
    hipEvent_t rb_events[5] = {s->rb_step_done, s->rb_packed[0], s->rb_packed[1], s->rb_copied[0],
                               s->rb_copied[1]};
    for (hipEvent_t ev : rb_events)
        if (ev) (void)hipEventDestroy(ev);
    for (void* h : s->rb_host)
        if (h) (void)hipHostFree(h);
    if (s->rb_dev) (void)hipFree(s->rb_dev);
    if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
    if (s->refine_mem) (void)hipFree(s->refine_mem);
    if (s->vort_mem) (void)hipFree(s->vort_mem);
    if (s->owner) (void)hipFree(s->owner);
    if (s->norms_mem) (void)hipFree(s->norms_mem);
    if (s->arena) (void)hipFree(s->arena);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

smoke_status smoke_readback_begin(smoke_sim* s, int format)
{
    if (!s) return fail(SMOKE_ERR_INVALID, "readback_begin: NULL sim");
    if (format < SMOKE_READBACK_F32 || format > SMOKE_READBACK_BF16)
        return fail(SMOKE_ERR_INVALID, "readback_begin: unknown format %d", format);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->rb_pending.size() >= 2)
        return fail(SMOKE_ERR_STATE, "readback_begin: two read-backs pending; call smoke_readback_wait");
    const size_t cells = (size_t)s->g.nx * s->g.ny * s->g.nz;
    const size_t slot_bytes = cells * sizeof(float);
    if (!s->copy_stream) {
        HIP_TRY(hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&s->rb_step_done, hipEventDisableTiming));
        for (int k = 0; k < 2; ++k) {
            HIP_TRY(hipEventCreateWithFlags(&s->rb_packed[k], hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&s->rb_copied[k], hipEventDisableTiming));
        }
    }
    if (!s->rb_dev) {
        if (hipMalloc(&s->rb_dev, 2 * slot_bytes) != hipSuccess) {
            s->rb_dev = nullptr;
            return fail(SMOKE_ERR_NOMEM, "readback_begin: cannot allocate device slots");
        }
        for (int k = 0; k < 2; ++k)
            if (hipHostMalloc(&s->rb_host[k], slot_bytes, hipHostMallocDefault) != hipSuccess) {
                s->rb_host[k] = nullptr;
                return fail(SMOKE_ERR_NOMEM, "readback_begin: cannot allocate pinned host buffers");
            }
    }
    const int slot = s->rb_next;
    char* dev = static_cast<char*>(s->rb_dev) + slot * slot_bytes;
    const int src_half = 1 - s->now;
    // The copy stream packs after the last enqueued step; the next step's
    // scene stage (which writes sources into both density halves) waits for
    // the pack, not for the host copy.  A simulation in a batch is stepped on
    // the batch's stream.
    hipStream_t compute = s->batch_stream ? s->batch_stream : s->stream;
    HIP_TRY(hipEventRecord(s->rb_step_done, compute));
    HIP_TRY(hipStreamWaitEvent(s->copy_stream, s->rb_step_done, 0));
    size_t bytes = slot_bytes;
    if (format == SMOKE_READBACK_F32) {
        HIP_TRY(smk::with_fields(s, [&](auto rho, auto, auto, auto) {
            return smk::pack_cells_f32(s->g, rho[src_half], reinterpret_cast<float*>(dev),
                                       s->copy_stream);
        }));
    } else {
        HIP_TRY(smk::with_fields(s, [&](auto rho, auto, auto, auto) {
            return smk::pack_cells_half(s->g, rho[src_half], reinterpret_cast<uint16_t*>(dev), format,
                                        s->copy_stream);
        }));
        bytes = cells * sizeof(uint16_t);
    }
    HIP_TRY(hipEventRecord(s->rb_packed[slot], s->copy_stream));
    HIP_TRY(hipStreamWaitEvent(compute, s->rb_packed[slot], 0));
    HIP_TRY(hipMemcpyAsync(s->rb_host[slot], dev, bytes, hipMemcpyDeviceToHost, s->copy_stream));
    HIP_TRY(hipEventRecord(s->rb_copied[slot], s->copy_stream));
    s->rb_pending.push_back({slot, s->steps_enqueued});
    s->rb_next = 1 - slot;
    return SMOKE_OK;
}

smoke_status smoke_readback_wait(smoke_sim* s, const void** host, long long* step)
{
    if (!s) return fail(SMOKE_ERR_INVALID, "readback_wait: NULL sim");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->rb_pending.empty()) return fail(SMOKE_ERR_STATE, "readback_wait: nothing pending");
    const smoke_sim::Readback r = s->rb_pending.front();
    HIP_TRY(hipEventSynchronize(s->rb_copied[r.slot]));
    s->rb_pending.pop_front();
    if (host) *host = s->rb_host[r.slot];
    if (step) *step = r.step;
    return SMOKE_OK;
}

smoke_status smoke_get_dims(const smoke_sim* s, unsigned* nx, unsigned* ny, unsigned* nz)
{
    if (!s) return fail(SMOKE_ERR_INVALID, "get_dims: NULL sim");
    if (nx) *nx = (unsigned)s->g.nx;
    if (ny) *ny = (unsigned)s->g.ny;
    if (nz) *nz = (unsigned)s->g.nz;
    return SMOKE_OK;
}

smoke_status smoke_set_params(smoke_sim* s, const smoke_params* p)
{
    if (!s) return fail(SMOKE_ERR_INVALID, "set_params: NULL sim");
    smoke_status st = check_params(p);
    if (st != SMOKE_OK) return st;
    std::lock_guard<std::mutex> lk(s->mu);
    SMOKE_REFUSE_IN_BATCH(s, "set_params");
    if (s->storage != SMOKE_STORAGE_FP32 && p->projection == SMOKE_PROJ_REFERENCE)
        return fail(SMOKE_ERR_INVALID,
                    "set_params: 16-bit storage runs the pressure forms (RBSOR, JACOBI, REFINED), not REFERENCE");
    if (p->obstacle_mode != s->prm.obstacle_mode) s->mask_dirty = true;
    if (p->projection == SMOKE_PROJ_REFINED) {
        smoke_status rs = ensure_refine(s);
        if (rs != SMOKE_OK) return rs;
    }
    if (p->vorticity > 0.0f) {
        smoke_status vs = ensure_vort(s);
        if (vs != SMOKE_OK) return vs;
    }
    if (p->obstacle_velocity) {
        smoke_status os = ensure_owner(s);
        if (os != SMOKE_OK) return os;
    }
    s->prm = *p;
    s->version += 1;
    return SMOKE_OK;
}

namespace {

// Snapshot blob: this header, then the object list, then the raw arena (its
// layout is a function of dims and storage only).
struct SnapshotHeader {
    char magic[8];  // "SMKSNAP1"
    uint32_t nx, ny, nz;
    int32_t storage;
    int32_t now, pcur;
    int32_t n_objects;
    int32_t plan[3];
    uint64_t arena_bytes;
    smoke_params prm;
};

size_t snapshot_size(const smoke_sim* s, size_t n_objects)
{
    return sizeof(SnapshotHeader) + n_objects * sizeof(Object) + arena_bytes(s->g, s->storage);
}

}  // namespace

size_t smoke_snapshot_bytes(const smoke_sim* s)
{
    return s ? snapshot_size(s, s->objects.size()) : 0;
}

smoke_status smoke_snapshot(smoke_sim* s, void* host, size_t bytes)
{
    if (!s || !host) return fail(SMOKE_ERR_INVALID, "snapshot: NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    const size_t need = snapshot_size(s, s->objects.size());
    if (bytes < need) return fail(SMOKE_ERR_INVALID, "snapshot: buffer holds %zu bytes, needs %zu", bytes, need);
    SMOKE_WAIT_BATCH(s);
    HIP_TRY(hipStreamSynchronize(s->stream));
    SnapshotHeader h;
    std::memset(&h, 0, sizeof h);
    std::memcpy(h.magic, "SMKSNAP1", 8);
    h.nx = (uint32_t)s->g.nx;
    h.ny = (uint32_t)s->g.ny;
    h.nz = (uint32_t)s->g.nz;
    h.storage = s->storage;
    h.now = s->now;
    h.pcur = s->pcur;
    h.n_objects = (int32_t)s->objects.size();
    h.plan[0] = s->plan.sweeps_per_launch;
    h.plan[1] = s->plan.tile;
    h.plan[2] = s->plan.zc;
    h.arena_bytes = arena_bytes(s->g, s->storage);
    h.prm = s->prm;
    char* p = static_cast<char*>(host);
    std::memcpy(p, &h, sizeof h);
    p += sizeof h;
    if (!s->objects.empty()) std::memcpy(p, s->objects.data(), s->objects.size() * sizeof(Object));
    p += s->objects.size() * sizeof(Object);
    HIP_TRY(hipMemcpy(p, s->arena, h.arena_bytes, hipMemcpyDeviceToHost));
    return SMOKE_OK;
}

smoke_status smoke_restore(smoke_sim* s, const void* host, size_t bytes)
{
    if (!s || !host) return fail(SMOKE_ERR_INVALID, "restore: NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    s->norms_step = -1;  // an in-step norms result no longer describes the state
    SMOKE_REFUSE_IN_BATCH(s, "restore");
    SnapshotHeader h;
    if (bytes < sizeof h) return fail(SMOKE_ERR_INVALID, "restore: blob too short");
    std::memcpy(&h, host, sizeof h);
    if (std::memcmp(h.magic, "SMKSNAP1", 8) != 0) return fail(SMOKE_ERR_INVALID, "restore: not a snapshot");
    if ((int)h.nx != s->g.nx || (int)h.ny != s->g.ny || (int)h.nz != s->g.nz || h.storage != s->storage)
        return fail(SMOKE_ERR_INVALID, "restore: snapshot of a %ux%ux%u storage-%d simulation, this one is %dx%dx%d storage-%d",
                    h.nx, h.ny, h.nz, h.storage, s->g.nx, s->g.ny, s->g.nz, s->storage);
    if (h.n_objects < 0 || h.n_objects > 2 * SMOKE_MAX_OBJECTS || h.arena_bytes != arena_bytes(s->g, s->storage) ||
        bytes < snapshot_size(s, (size_t)h.n_objects) || (h.now != 0 && h.now != 1) ||
        (h.pcur != 0 && h.pcur != 1) || h.plan[0] < 1 || h.plan[0] > 6 ||
        h.plan[1] < SMOKE_TILE_32x32 || h.plan[1] > SMOKE_TILE_32x64 || h.plan[2] < 1)
        return fail(SMOKE_ERR_INVALID, "restore: inconsistent snapshot");
    smoke_status st = check_params(&h.prm);
    if (st != SMOKE_OK) return st;
    if (h.prm.projection == SMOKE_PROJ_REFINED && (st = ensure_refine(s)) != SMOKE_OK) return st;
    if (h.prm.vorticity > 0.0f && (st = ensure_vort(s)) != SMOKE_OK) return st;
    if (h.prm.obstacle_velocity && (st = ensure_owner(s)) != SMOKE_OK) return st;
    const char* p = static_cast<const char*>(host) + sizeof h;
    // The object list must satisfy what add_object enforces: type 0 or 1 and
    // at most SMOKE_MAX_OBJECTS of each (the mask and boundary-condition
    // sphere lists hold that many).
    {
        int per_type[2] = {0, 0};
        for (int32_t k = 0; k < h.n_objects; ++k) {
            Object o;
            std::memcpy(&o, p + (size_t)k * sizeof(Object), sizeof o);
            if (o.type != 0 && o.type != 1)
                return fail(SMOKE_ERR_INVALID, "restore: object %d has unknown type %d", k, o.type);
            if (++per_type[o.type] > SMOKE_MAX_OBJECTS)
                return fail(SMOKE_ERR_INVALID, "restore: more than %d %s", SMOKE_MAX_OBJECTS,
                            o.type == 0 ? "obstacles" : "sources");
        }
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->objects.assign(reinterpret_cast<const Object*>(p), reinterpret_cast<const Object*>(p) + h.n_objects);
    p += (size_t)h.n_objects * sizeof(Object);
    HIP_TRY(hipMemcpy(s->arena, p, h.arena_bytes, hipMemcpyHostToDevice));
    s->now = h.now;
    s->pcur = h.pcur;
    s->prm = h.prm;
    s->plan.sweeps_per_launch = h.plan[0];
    s->plan.tile = h.plan[1];
    s->plan.zc = h.plan[2];
    s->version += 1;
    s->mask_dirty = true;  // flags (and the refined solver's inverse) follow the restored masks
    HIP_TRY(refresh_mask(s, nullptr));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SMOKE_OK;
}

smoke_status smoke_get_solve_info(smoke_sim* s, double* rel_residual, int* refinements)
{
    if (!s) return fail(SMOKE_ERR_INVALID, "get_solve_info: NULL sim");
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->refine_mem) return fail(SMOKE_ERR_STATE, "get_solve_info: the REFINED solver never ran");
    HIP_TRY(hipStreamSynchronize(s->stream));
    smk::RefineCtl c;
    HIP_TRY(hipMemcpy(&c, s->rb.ctl, sizeof c, hipMemcpyDeviceToHost));
    if (rel_residual) *rel_residual = c.rel;
    if (refinements) *refinements = c.refinements;
    return SMOKE_OK;
}

smoke_status smoke_set_jacobi_plan(smoke_sim* s, const smoke_jacobi_plan* plan)
{
    if (!s || !plan) return fail(SMOKE_ERR_INVALID, "set_jacobi_plan: NULL");
    if (plan->sweeps_per_launch < 1 || plan->sweeps_per_launch > 6)
        return fail(SMOKE_ERR_INVALID, "set_jacobi_plan: sweeps_per_launch %d not in 1..6",
                    plan->sweeps_per_launch);
    if (plan->tile < SMOKE_TILE_32x32 || plan->tile > SMOKE_TILE_32x64)
        return fail(SMOKE_ERR_INVALID, "set_jacobi_plan: unknown tile %d", plan->tile);
    if (plan->zc < 1) return fail(SMOKE_ERR_INVALID, "set_jacobi_plan: zc %d < 1", plan->zc);
    std::lock_guard<std::mutex> lk(s->mu);
    s->plan.sweeps_per_launch = plan->sweeps_per_launch;
    s->plan.tile = plan->tile;
    s->plan.zc = plan->zc;
    s->version += 1;
    return SMOKE_OK;
}

smoke_status smoke_get_jacobi_plan(const smoke_sim* s, smoke_jacobi_plan* out)
{
    if (!s || !out) return fail(SMOKE_ERR_INVALID, "get_jacobi_plan: NULL");
    out->sweeps_per_launch = s->plan.sweeps_per_launch;
    out->tile = s->plan.tile;
    out->zc = s->plan.zc;
    return SMOKE_OK;
}

smoke_status smoke_autotune(smoke_sim* s, smoke_jacobi_plan* chosen, double* ms_per_sweep)
{
    if (!s) return fail(SMOKE_ERR_INVALID, "autotune: NULL sim");
    std::lock_guard<std::mutex> lk(s->mu);
    SMOKE_REFUSE_IN_BATCH(s, "autotune");
    const int K = s->prm.iterations > 0 ? s->prm.iterations : 100;
    HIP_TRY(refresh_mask(s, nullptr));
    hipEvent_t ev[2] = {nullptr, nullptr};
    auto cleanup = [&]() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    };
    if (hipEventCreate(&ev[0]) != hipSuccess || hipEventCreate(&ev[1]) != hipSuccess) {
        cleanup();
        return fail(SMOKE_ERR_HIP, "autotune: cannot create events");
    }
    smk::JacobiPlan best = s->plan;
    double best_ms = -1.0;
    const int tiles[4] = {smk::JacobiPlan::TILE_32x32, smk::JacobiPlan::TILE_64x16,
                          smk::JacobiPlan::TILE_64x32, smk::JacobiPlan::TILE_32x64};
    // z-chunks: fixed lengths, and the interior split into 1..4 equal chunks
    // (one chunk per block gives the fewest blocks and no z halo; when the
    // tiles of a plane about fill the GPU, that is fastest).
    std::vector<int> zcs = {32, 64, 128};
    for (int c = 1; c <= 4; ++c) {
        const int zc = (s->g.nz - 2 + c - 1) / c;
        if (zc >= 1 && std::find(zcs.begin(), zcs.end(), zc) == zcs.end()) zcs.push_back(zc);
    }
    constexpr int kReps = 2;
    hipError_t e = hipSuccess;
    // REFINED: the refined solve's sequence (FP16 sweep launches of an even
    // T, one- or two-row tiles, and the FP32 residual passes), per FP16 sweep.
    const bool refined = s->prm.projection == SMOKE_PROJ_REFINED;
    const int pcur0 = s->pcur;
    const smk::JacobiPlan plan0 = s->plan;
    const long long fp16_sweeps =
        (long long)((s->prm.iterations + s->prm.inner_sweeps - 1) / s->prm.inner_sweeps) *
        s->prm.inner_sweeps;
    if (refined) {
        // A right-hand side to solve: the divergence of the last output
        // velocity, half 1 - now (a zero one would end every refined solve at
        // its first norm).  d then no longer holds the last projection's
        // right-hand side.
        const int h = 1 - s->now;
        e = smk::with_fields(s, [&](auto, auto u, auto v, auto w) {
            return smk::divergence(s->g, u[h], v[h], w[h], s->flags, s->d, s->stream);
        });
        s->rhs_valid = false;
    }
    for (int t = 2; t <= 6 && e == hipSuccess; ++t) {
        if (refined && (t & 1)) continue;
        for (int tile : tiles) {
            if (e != hipSuccess) break;
            if (refined && tile != smk::JacobiPlan::TILE_32x32 && tile != smk::JacobiPlan::TILE_64x32)
                continue;  // the refined sweeps have one tile per row count
            for (int zc : zcs) {
                smk::JacobiPlan cand;
                cand.sweeps_per_launch = t;
                cand.tile = tile;
                cand.zc = zc;
                long long n = 0, n2 = 0;
                auto run = [&]() {
                    if (!refined) return enqueue_jacobi(s, cand, K, false, &n);
                    s->plan = cand;
                    const hipError_t r = enqueue_refined(s, &n, &n2);
                    s->pcur = pcur0;
                    return r;
                };
                e = run();  // warm-up
                if (e == hipSuccess) e = hipEventRecord(ev[0], s->stream);
                for (int r = 0; r < kReps && e == hipSuccess; ++r) e = run();
                if (e == hipSuccess) e = hipEventRecord(ev[1], s->stream);
                if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
                float ms = 0.0f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
                if (e != hipSuccess) break;
                const double per_sweep = ms / (double)(kReps * (refined ? fp16_sweeps : K));
                if (best_ms < 0.0 || per_sweep < best_ms) {
                    best_ms = per_sweep;
                    best = cand;
                }
            }
        }
    }
    s->plan = plan0;
    s->pcur = pcur0;
    // The sweeps ran on whatever right-hand side the buffer held: start the
    // next warm start from zero.
    if (e == hipSuccess) e = smk::fill_f32(s->P[0], s->g.elems, 0.0f, s->stream);
    if (e == hipSuccess) e = smk::fill_f32(s->P[1], s->g.elems, 0.0f, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    cleanup();
    if (e != hipSuccess) return fail(SMOKE_ERR_HIP, "autotune: %s", hipGetErrorString(e));
    s->plan = best;
    s->version += 1;
    if (chosen) {
        chosen->sweeps_per_launch = best.sweeps_per_launch;
        chosen->tile = best.tile;
        chosen->zc = best.zc;
    }
    if (ms_per_sweep) *ms_per_sweep = best_ms;
    return SMOKE_OK;
}

smoke_status smoke_get_params(const smoke_sim* s, smoke_params* out)
{
    if (!s || !out) return fail(SMOKE_ERR_INVALID, "get_params: NULL");
    *out = s->prm;
    return SMOKE_OK;
}

static smoke_status add_object(smoke_sim* s, int type, float x, float y, float z, float vx, float vy,
                               float vz, float r, int* id)
{
    if (!s) return fail(SMOKE_ERR_INVALID, "add object: NULL sim");
    std::lock_guard<std::mutex> lk(s->mu);
    SMOKE_REFUSE_IN_BATCH(s, "add object");
    if (count_type(s, type) >= SMOKE_MAX_OBJECTS)
        return fail(SMOKE_ERR_INVALID, "add object: more than %d %s", SMOKE_MAX_OBJECTS,
                    type == 0 ? "obstacles" : "sources");
    Object o;
    o.pos[0] = x; o.pos[1] = y; o.pos[2] = z;
    o.vel[0] = vx; o.vel[1] = vy; o.vel[2] = vz;
    o.r = r;
    o.type = type;
    s->objects.push_back(o);
    s->version += 1;
    if (type == 0) s->mask_dirty = true;
    if (id) *id = (int)s->objects.size() - 1;
    return SMOKE_OK;
}

smoke_status smoke_add_obstacle(smoke_sim* s, float x, float y, float z, float vx, float vy, float vz,
                                float r, int* id)
{
    return add_object(s, 0, x, y, z, vx, vy, vz, r, id);
}

smoke_status smoke_add_source(smoke_sim* s, float x, float y, float z, float r, int* id)
{
    return add_object(s, 1, x, y, z, 0.0f, 0.0f, 0.0f, r, id);
}

smoke_status smoke_move_object(smoke_sim* s, int id, float x, float y, float z)
{
    if (!s) return fail(SMOKE_ERR_INVALID, "move_object: NULL sim");
    std::lock_guard<std::mutex> lk(s->mu);
    SMOKE_REFUSE_IN_BATCH(s, "move_object");
    if (id < 0 || id >= (int)s->objects.size())
        return fail(SMOKE_ERR_INVALID, "move_object: no object with id %d", id);
    Object& o = s->objects[id];
    o.pos[0] = x; o.pos[1] = y; o.pos[2] = z;
    s->version += 1;
    if (o.type == 0) s->mask_dirty = true;
    return SMOKE_OK;
}

smoke_status smoke_set_object_velocity(smoke_sim* s, int id, float vx, float vy, float vz)
{
    if (!s) return fail(SMOKE_ERR_INVALID, "set_object_velocity: NULL sim");
    if (!std::isfinite(vx) || !std::isfinite(vy) || !std::isfinite(vz))
        return fail(SMOKE_ERR_INVALID, "set_object_velocity: velocity must be finite");
    std::lock_guard<std::mutex> lk(s->mu);
    SMOKE_REFUSE_IN_BATCH(s, "set_object_velocity");
    if (id < 0 || id >= (int)s->objects.size())
        return fail(SMOKE_ERR_INVALID, "set_object_velocity: no object with id %d", id);
    Object& o = s->objects[id];
    o.vel[0] = vx; o.vel[1] = vy; o.vel[2] = vz;
    s->version += 1;
    return SMOKE_OK;
}

smoke_status smoke_set_solids(smoke_sim* s, const uint8_t* solid)
{
    if (!s || !solid) return fail(SMOKE_ERR_INVALID, "set_solids: NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    s->norms_step = -1;  // an in-step norms result no longer describes the state
    SMOKE_REFUSE_IN_BATCH(s, "set_solids");
    size_t n = (size_t)s->g.nx * s->g.ny * s->g.nz;
    HIP_TRY(hipMemcpyAsync(s->staging, solid, n, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(smk::unpack_cells_solid(s->g, reinterpret_cast<const uint8_t*>(s->staging), s->base, s->stream));
    s->mask_dirty = true;
    HIP_TRY(refresh_mask(s, nullptr));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SMOKE_OK;
}

smoke_status smoke_set_density(smoke_sim* s, int half, const float* src)
{
    if (!s || !src || (half != 0 && half != 1)) return fail(SMOKE_ERR_INVALID, "set_density: bad argument");
    std::lock_guard<std::mutex> lk(s->mu);
    s->norms_step = -1;  // an in-step norms result no longer describes the state
    SMOKE_REFUSE_IN_BATCH(s, "set_density");
    size_t bytes = (size_t)s->g.nx * s->g.ny * s->g.nz * sizeof(float);
    HIP_TRY(hipMemcpyAsync(s->staging, src, bytes, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(smk::with_fields(s, [&](auto rho, auto, auto, auto) {
        return smk::unpack_cells_f32(s->g, s->staging, rho[half], s->stream);
    }));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SMOKE_OK;
}

smoke_status smoke_get_density(smoke_sim* s, int half, float* dst)
{
    if (!s || !dst || (half != 0 && half != 1)) return fail(SMOKE_ERR_INVALID, "get_density: bad argument");
    std::lock_guard<std::mutex> lk(s->mu);
    return smk::with_fields(s, [&](auto rho, auto, auto, auto) { return read_cells_f32(s, rho[half], dst); });
}

smoke_status smoke_set_velocity(smoke_sim* s, int comp, int half, const float* src)
{
    if (!s || !src) return fail(SMOKE_ERR_INVALID, "set_velocity: NULL");
    if ((half != 0 && half != 1) || comp < 0 || comp > 2)
        return fail(SMOKE_ERR_INVALID, "set_velocity: bad component %d or half %d", comp, half);
    std::lock_guard<std::mutex> lk(s->mu);
    s->norms_step = -1;  // an in-step norms result no longer describes the state
    SMOKE_REFUSE_IN_BATCH(s, "set_velocity");
    size_t bytes = (size_t)(s->g.nx + 1) * (s->g.ny + 1) * (s->g.nz + 1) * sizeof(float);
    HIP_TRY(hipMemcpyAsync(s->staging, src, bytes, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(smk::with_fields(s, [&](auto, auto u, auto v, auto w) {
        auto dst = comp == 0 ? u[half] : (comp == 1 ? v[half] : w[half]);
        return smk::unpack_faces_f32(s->g, s->staging, dst, s->stream);
    }));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SMOKE_OK;
}

smoke_status smoke_get_velocity(smoke_sim* s, int comp, int half, float* dst)
{
    if (!s || !dst) return fail(SMOKE_ERR_INVALID, "get_velocity: NULL");
    if ((half != 0 && half != 1) || comp < 0 || comp > 2)
        return fail(SMOKE_ERR_INVALID, "get_velocity: bad component %d or half %d", comp, half);
    std::lock_guard<std::mutex> lk(s->mu);
    SMOKE_WAIT_BATCH(s);
    HIP_TRY(smk::with_fields(s, [&](auto, auto u, auto v, auto w) {
        auto src = comp == 0 ? u[half] : (comp == 1 ? v[half] : w[half]);
        return smk::pack_faces_f32(s->g, src, s->staging, s->stream);
    }));
    size_t bytes = (size_t)(s->g.nx + 1) * (s->g.ny + 1) * (s->g.nz + 1) * sizeof(float);
    HIP_TRY(hipMemcpyAsync(dst, s->staging, bytes, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SMOKE_OK;
}

static smoke_status read_bytes(smoke_sim* s, const uint8_t* src, uint8_t* dst)
{
    SMOKE_WAIT_BATCH(s);
    uint8_t* st = reinterpret_cast<uint8_t*>(s->staging);
    HIP_TRY(smk::pack_cells_u8(s->g, src, st, s->stream));
    HIP_TRY(hipMemcpyAsync(dst, st, (size_t)s->g.nx * s->g.ny * s->g.nz, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SMOKE_OK;
}

smoke_status smoke_get_mask(smoke_sim* s, uint8_t* dst)
{
    if (!s || !dst) return fail(SMOKE_ERR_INVALID, "get_mask: NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    return read_bytes(s, s->s, dst);
}

smoke_status smoke_get_flags(smoke_sim* s, uint8_t* dst)
{
    if (!s || !dst) return fail(SMOKE_ERR_INVALID, "get_flags: NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    return read_bytes(s, s->flags, dst);
}

smoke_status smoke_get_pressure(smoke_sim* s, float* dst)
{
    if (!s || !dst) return fail(SMOKE_ERR_INVALID, "get_pressure: NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    return read_cells_f32(s, s->P[s->pcur], dst);
}

int smoke_index_now(const smoke_sim* s) { return s ? s->now : -1; }

smoke_status smoke_set_step_graph(smoke_sim* s, int enable)
{
    if (!s) return fail(SMOKE_ERR_INVALID, "set_step_graph: NULL sim");
    std::lock_guard<std::mutex> lk(s->mu);
    if (!enable && s->graph_on) {
        HIP_TRY(hipStreamSynchronize(s->stream));
        drop_step_graphs(s);
    }
    s->graph_on = enable != 0;
    return SMOKE_OK;
}

smoke_status smoke_step_async(smoke_sim* s, float dt)
{
    if (!s) return fail(SMOKE_ERR_INVALID, "step: NULL sim");
    std::lock_guard<std::mutex> lk(s->mu);
    return step_locked(s, dt);
}

smoke_status smoke_step(smoke_sim* s, float dt, float* density_out)
{
    if (!s) return fail(SMOKE_ERR_INVALID, "step: NULL sim");
    std::lock_guard<std::mutex> lk(s->mu);
    smoke_status st = step_locked(s, dt);
    if (st != SMOKE_OK) return st;
    if (density_out)
        return smk::with_fields(s, [&](auto rho, auto, auto, auto) {
            return read_cells_f32(s, rho[1 - s->now], density_out);
        });
    return SMOKE_OK;
}

smoke_status smoke_run_stage(smoke_sim* s, int stage, float dt)
{
    if (!s) return fail(SMOKE_ERR_INVALID, "run_stage: NULL sim");
    if (stage < 0 || stage >= SMOKE_STAGE_COUNT) return fail(SMOKE_ERR_INVALID, "run_stage: bad stage %d", stage);
    if (!std::isfinite(dt)) return fail(SMOKE_ERR_INVALID, "run_stage: dt must be finite");
    std::lock_guard<std::mutex> lk(s->mu);
    SMOKE_REFUSE_IN_BATCH(s, "run_stage");
    s->norms_step = -1;
    s->last_dt = dt;
    long long launches = 0;
    hipError_t e = run_stage(s, stage, dt, &launches);
    if (e != hipSuccess) return fail(SMOKE_ERR_HIP, "run_stage %d: %s", stage, hipGetErrorString(e));
    return SMOKE_OK;
}

smoke_status smoke_read_density(smoke_sim* s, float* dst)
{
    if (!s || !dst) return fail(SMOKE_ERR_INVALID, "read_density: NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    return smk::with_fields(s, [&](auto rho, auto, auto, auto) {
        return read_cells_f32(s, rho[1 - s->now], dst);
    });
}

smoke_status smoke_synchronize(smoke_sim* s)
{
    if (!s) return fail(SMOKE_ERR_INVALID, "synchronize: NULL sim");
    SMOKE_WAIT_BATCH(s);
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SMOKE_OK;
}

smoke_status smoke_norms_enable(smoke_sim* s, int enable)
{
    if (!s) return fail(SMOKE_ERR_INVALID, "norms_enable: NULL sim");
    std::lock_guard<std::mutex> lk(s->mu);
    if (enable) {
        smoke_status st = ensure_norms(s);
        if (st != SMOKE_OK) return st;
    }
    s->norms_on = enable != 0;
    return SMOKE_OK;
}

smoke_status smoke_get_step_norms(smoke_sim* s, smoke_step_norms* out)
{
    if (!s || !out) return fail(SMOKE_ERR_INVALID, "get_step_norms: NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    SMOKE_WAIT_BATCH(s);
    smoke_status st = ensure_norms(s);
    if (st != SMOKE_OK) return st;
    // A batch step or a stage-level call leaves no in-step result for the
    // current state: reduce now.
    if (!(s->norms_on && s->norms_step == s->steps_enqueued)) HIP_TRY(enqueue_own_norms(s));
    HIP_TRY(hipMemcpyAsync(out, norms_record(s), sizeof *out, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (out->nonfinite)
        return fail(SMOKE_ERR_NONFINITE, "step %lld: %lld non-finite value(s) in density/velocity",
                    out->step, out->nonfinite);
    return SMOKE_OK;
}

smoke_status smoke_get_step_norms_noalloc(smoke_sim* s, smoke_step_norms* out_device)
{
    if (!s || !out_device) return fail(SMOKE_ERR_INVALID, "get_step_norms_noalloc: NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->batch)
        return fail(SMOKE_ERR_STATE, "get_step_norms_noalloc: the simulation belongs to a batch");
    if (smk::noalloc_norms_scratch_bytes() < smk::norms_scratch_bytes())
        return fail(SMOKE_ERR_INVALID, "get_step_norms_noalloc: the reduction needs %zu bytes of "
                    "scratch, the static scratch holds %zu", smk::norms_scratch_bytes(),
                    smk::noalloc_norms_scratch_bytes());
    void* scratch = nullptr;
    HIP_TRY(smk::noalloc_norms_scratch(s->device, &scratch));
    HIP_TRY(smk::enqueue_norms(s, scratch, out_device));
    return SMOKE_OK;
}

smoke_status smoke_timing_enable(smoke_sim* s, int enable)
{
    if (!s) return fail(SMOKE_ERR_INVALID, "timing_enable: NULL sim");
    std::lock_guard<std::mutex> lk(s->mu);
    s->timing = enable != 0;
    return SMOKE_OK;
}

smoke_status smoke_timing_read(smoke_sim* s, double* ms, long long* launches, long long* steps)
{
    if (!s) return fail(SMOKE_ERR_INVALID, "timing_read: NULL sim");
    std::lock_guard<std::mutex> lk(s->mu);
    HIP_TRY(collect_timing(s, true));
    for (int k = 0; k < SMOKE_STAGE_COUNT; ++k) {
        if (ms) ms[k] = s->stage_ms[k];
        if (launches) launches[k] = s->stage_launches[k];
    }
    if (steps) *steps = s->timed_steps;
    return SMOKE_OK;
}

smoke_status smoke_timing_reset(smoke_sim* s)
{
    if (!s) return fail(SMOKE_ERR_INVALID, "timing_reset: NULL sim");
    std::lock_guard<std::mutex> lk(s->mu);
    HIP_TRY(collect_timing(s, true));
    for (int k = 0; k < SMOKE_STAGE_COUNT; ++k) {
        s->stage_ms[k] = 0.0;
        s->stage_launches[k] = 0;
    }
    s->sweep_ms = 0.0;
    s->sweep_launches = 0;
    s->timed_steps = 0;
    s->step_ms.clear();
    return SMOKE_OK;
}

smoke_status smoke_timing_sweeps(smoke_sim* s, double* ms, long long* launches)
{
    if (!s) return fail(SMOKE_ERR_INVALID, "timing_sweeps: NULL sim");
    std::lock_guard<std::mutex> lk(s->mu);
    HIP_TRY(collect_timing(s, true));
    if (ms) *ms = s->sweep_ms;
    if (launches) *launches = s->sweep_launches;
    return SMOKE_OK;
}

smoke_status smoke_timing_step_ms(smoke_sim* s, double* ms, int cap, int* count)
{
    if (!s || (!ms && cap > 0)) return fail(SMOKE_ERR_INVALID, "timing_step_ms: NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    HIP_TRY(collect_timing(s, true));
    int n = (int)std::min<size_t>((size_t)std::max(cap, 0), s->step_ms.size());
    size_t first = s->step_ms.size() - (size_t)n;
    for (int i = 0; i < n; ++i) ms[i] = s->step_ms[first + (size_t)i];
    if (count) *count = n;
    return SMOKE_OK;
}

}  // extern "C"
