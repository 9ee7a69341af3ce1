// Request loop with continuous batching (smoke_server_* of include/smoke.h).
//
// Requests wait in arrival order.  Every tick: (1) admit from the head of the
// queue while the request's device footprint fits the HBM budget (its
// simulation is created and loaded on its own stream, overlapping the batch
// still running on the GPU); (2) if the set of resident requests changed,
// rebuild the batch, otherwise reuse it (its step is replayed from a HIP graph
// after the first step); (3) advance every resident request by one step;
// (4) retire requests that ran all their steps: copy out the final density,
// free the simulation and its budget.  Built on the public C ABI only.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "sim_state.h"

using smk::fail;

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
    long long id = 0;
    unsigned nx = 0, ny = 0, nz = 0;
    int steps = 0;
    float dt = 0.0f;
    smoke_params params{};
    std::vector<uint8_t> solid;
    std::vector<float> sources, obstacles;
    std::vector<float> velocity[3];
    size_t footprint = 0;
    int state = SMOKE_REQUEST_QUEUED;
    smoke_sim* sim = nullptr;
    int steps_done = 0;
    Clock::time_point t_submit, t_admit, t_done;
    std::vector<float> result;
};

}  // namespace

struct smoke_server {
    std::mutex mu;
    size_t budget = 0, used = 0, peak = 0;
    int iterations = 0;
    long long next_id = 0;
    std::map<long long, std::unique_ptr<Request>> requests;
    std::deque<Request*> queue;
    std::vector<Request*> active;
    smoke_batch* batch = nullptr;
    int steps_since_rebuild = 0;
    bool dirty = false;
    long long admitted = 0, completed = 0, ticks = 0, cell_updates = 0;
    double gpu_ms = 0.0;
    std::vector<int> batch_sizes;         // requests advanced by each step of the current batch
    std::vector<double> tick_ms;          // GPU ms per tick (collected)
    std::vector<int> tick_requests;
};

namespace {

// GPU step times of the current batch are collected when it is replaced.
smoke_status close_batch(smoke_server* sv)
{
    if (!sv->batch) return SMOKE_OK;
    smoke_status st = smoke_batch_synchronize(sv->batch);
    if (st != SMOKE_OK) return st;
    std::vector<double> ms(sv->batch_sizes.size());
    int n = 0;
    st = smoke_batch_step_ms(sv->batch, ms.data(), (int)ms.size(), &n);
    if (st != SMOKE_OK) return st;
    for (int i = 0; i < n; ++i) {
        sv->tick_ms.push_back(ms[i]);
        sv->tick_requests.push_back(sv->batch_sizes[(size_t)i]);
        sv->gpu_ms += ms[i];
    }
    sv->batch_sizes.clear();
    smoke_batch_destroy(sv->batch);
    sv->batch = nullptr;
    return SMOKE_OK;
}

smoke_status load_request(smoke_server* sv, Request* r)
{
    smoke_sim* s = nullptr;
    smoke_status st = smoke_create(r->nx, r->ny, r->nz, nullptr, &s);
    if (st != SMOKE_OK) return st;
    auto bail = [&](smoke_status e) {
        smoke_destroy(s);
        return e;
    };
    if (!r->solid.empty() && (st = smoke_set_solids(s, r->solid.data())) != SMOKE_OK) return bail(st);
    for (size_t k = 0; k + 3 < r->obstacles.size(); k += 4) {
        const float* o = &r->obstacles[k];
        if ((st = smoke_add_obstacle(s, o[0], o[1], o[2], 0.0f, 0.0f, 0.0f, o[3], nullptr)) != SMOKE_OK)
            return bail(st);
    }
    for (size_t k = 0; k + 3 < r->sources.size(); k += 4) {
        const float* o = &r->sources[k];
        if ((st = smoke_add_source(s, o[0], o[1], o[2], o[3], nullptr)) != SMOKE_OK) return bail(st);
    }
    for (int c = 0; c < 3; ++c)
        if (!r->velocity[c].empty() &&
            (st = smoke_set_velocity(s, c, 0, r->velocity[c].data())) != SMOKE_OK)
            return bail(st);
    if ((st = smoke_set_params(s, &r->params)) != SMOKE_OK) return bail(st);
    r->sim = s;
    return SMOKE_OK;
}

}  // namespace

extern "C" {

smoke_status smoke_server_create(size_t hbm_budget_bytes, int iterations, smoke_server** out)
{
    if (!out) return fail(SMOKE_ERR_INVALID, "server_create: out is NULL");
    *out = nullptr;
    if (iterations < 1) return fail(SMOKE_ERR_INVALID, "server_create: iterations must be >= 1");
    if (hbm_budget_bytes == 0) return fail(SMOKE_ERR_INVALID, "server_create: zero budget");
    smoke_server* sv = new smoke_server();
    sv->budget = hbm_budget_bytes;
    sv->iterations = iterations;
    *out = sv;
    return SMOKE_OK;
}

void smoke_server_destroy(smoke_server* sv)
{
    if (!sv) return;
    std::lock_guard<std::mutex> lk(sv->mu);
    (void)close_batch(sv);
    for (Request* r : sv->active) smoke_destroy(r->sim);
    delete sv;
}

smoke_status smoke_server_submit(smoke_server* sv, const smoke_request* q, long long* id)
{
    if (!sv || !q) return fail(SMOKE_ERR_INVALID, "server_submit: NULL");
    if (q->steps < 1) return fail(SMOKE_ERR_INVALID, "server_submit: steps must be >= 1");
    if (q->n_sources < 0 || q->n_sources > 8 || (q->n_sources && !q->sources))
        return fail(SMOKE_ERR_INVALID, "server_submit: 0..8 sources");
    if (q->n_obstacles < 0 || q->n_obstacles > SMOKE_MAX_OBJECTS || (q->n_obstacles && !q->obstacles))
        return fail(SMOKE_ERR_INVALID, "server_submit: bad obstacle list");
    const size_t fp = smoke_footprint_bytes(q->nx, q->ny, q->nz);
    if (fp == 0) return fail(SMOKE_ERR_INVALID, "server_submit: invalid dims");
    std::lock_guard<std::mutex> lk(sv->mu);
    if (fp > sv->budget) return fail(SMOKE_ERR_NOMEM, "server_submit: request needs %zu bytes, budget %zu", fp, sv->budget);
    auto r = std::make_unique<Request>();
    r->id = sv->next_id++;
    r->nx = q->nx;
    r->ny = q->ny;
    r->nz = q->nz;
    r->steps = q->steps;
    r->dt = q->dt;
    smoke_params_default(&r->params);
    r->params.gravity = q->gravity;
    r->params.buoyancy = q->buoyancy;
    r->params.density_gate = q->density_gate;
    r->params.max_velocity = q->max_velocity;
    r->params.projection = SMOKE_PROJ_JACOBI;
    r->params.iterations = sv->iterations;
    const size_t cells = (size_t)q->nx * q->ny * q->nz;
    const size_t faces = (size_t)(q->nx + 1) * (q->ny + 1) * (q->nz + 1);
    if (q->solid) r->solid.assign(q->solid, q->solid + cells);
    if (q->n_sources) r->sources.assign(q->sources, q->sources + 4 * q->n_sources);
    if (q->n_obstacles) r->obstacles.assign(q->obstacles, q->obstacles + 4 * q->n_obstacles);
    for (int c = 0; c < 3; ++c)
        if (q->velocity[c]) r->velocity[c].assign(q->velocity[c], q->velocity[c] + faces);
    r->footprint = fp;
    r->t_submit = Clock::now();
    if (id) *id = r->id;
    sv->queue.push_back(r.get());
    sv->requests[r->id] = std::move(r);
    return SMOKE_OK;
}

smoke_status smoke_server_tick(smoke_server* sv, int* active, int* queued)
{
    if (!sv) return fail(SMOKE_ERR_INVALID, "server_tick: NULL");
    std::lock_guard<std::mutex> lk(sv->mu);
    smoke_status st = SMOKE_OK;
    // (1) admission, in arrival order, while the head fits.
    while (!sv->queue.empty() && (int)sv->active.size() < SMOKE_MAX_BATCH) {
        Request* r = sv->queue.front();
        if (sv->used + r->footprint > sv->budget) break;
        if ((st = load_request(sv, r)) != SMOKE_OK) return st;
        sv->queue.pop_front();
        sv->used += r->footprint;
        sv->peak = std::max(sv->peak, sv->used);
        r->state = SMOKE_REQUEST_RUNNING;
        r->t_admit = Clock::now();
        sv->active.push_back(r);
        sv->admitted += 1;
        sv->dirty = true;
    }
    if (sv->active.empty()) {
        if (active) *active = 0;
        if (queued) *queued = (int)sv->queue.size();
        return SMOKE_OK;
    }
    // (2) batch of the resident requests.
    if (sv->dirty || !sv->batch) {
        if ((st = close_batch(sv)) != SMOKE_OK) return st;
        std::vector<smoke_sim*> sims;
        std::vector<float> dts;
        for (Request* r : sv->active) {
            sims.push_back(r->sim);
            dts.push_back(r->dt);
        }
        if ((st = smoke_batch_create(sims.data(), (int)sims.size(), dts.data(), &sv->batch)) != SMOKE_OK)
            return st;
        sv->dirty = false;
        sv->steps_since_rebuild = 0;
    }
    // (3) one step of every resident request.
    if ((st = smoke_batch_step(sv->batch, sv->steps_since_rebuild > 0)) != SMOKE_OK) return st;
    sv->steps_since_rebuild += 1;
    sv->batch_sizes.push_back((int)sv->active.size());
    sv->ticks += 1;
    bool finishing = false;
    for (Request* r : sv->active) {
        r->steps_done += 1;
        sv->cell_updates += (long long)r->nx * r->ny * r->nz;
        finishing = finishing || r->steps_done >= r->steps;
    }
    // (4) retire finished requests.
    if (finishing) {
        if ((st = close_batch(sv)) != SMOKE_OK) return st;
        std::vector<Request*> keep;
        for (Request* r : sv->active) {
            if (r->steps_done < r->steps) {
                keep.push_back(r);
                continue;
            }
            r->result.resize((size_t)r->nx * r->ny * r->nz);
            if ((st = smoke_read_density(r->sim, r->result.data())) != SMOKE_OK) return st;
            smoke_destroy(r->sim);
            r->sim = nullptr;
            sv->used -= r->footprint;
            r->state = SMOKE_REQUEST_DONE;
            r->t_done = Clock::now();
            sv->completed += 1;
        }
        sv->active.swap(keep);
        sv->dirty = true;
    }
    if (active) *active = (int)sv->active.size();
    if (queued) *queued = (int)sv->queue.size();
    return SMOKE_OK;
}

smoke_status smoke_server_result(smoke_server* sv, long long id, int* state, float* density,
                                 double* queue_ms, double* run_ms)
{
    if (!sv) return fail(SMOKE_ERR_INVALID, "server_result: NULL");
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->requests.find(id);
    if (it == sv->requests.end()) return fail(SMOKE_ERR_INVALID, "server_result: no request %lld", id);
    Request* r = it->second.get();
    if (state) *state = r->state;
    if (r->state == SMOKE_REQUEST_DONE) {
        if (density) std::memcpy(density, r->result.data(), r->result.size() * sizeof(float));
        if (queue_ms) *queue_ms = ms_between(r->t_submit, r->t_admit);
        if (run_ms) *run_ms = ms_between(r->t_admit, r->t_done);
    }
    return SMOKE_OK;
}

smoke_status smoke_server_stats_get(smoke_server* sv, smoke_server_stats* out)
{
    if (!sv || !out) return fail(SMOKE_ERR_INVALID, "server_stats: NULL");
    std::lock_guard<std::mutex> lk(sv->mu);
    out->submitted = sv->next_id;
    out->admitted = sv->admitted;
    out->completed = sv->completed;
    out->ticks = sv->ticks;
    out->cell_updates = sv->cell_updates;
    out->gpu_ms = sv->gpu_ms;
    out->hbm_used = sv->used;
    out->hbm_budget = sv->budget;
    out->hbm_peak = sv->peak;
    out->active = (int)sv->active.size();
    out->queued = (int)sv->queue.size();
    return SMOKE_OK;
}

smoke_status smoke_server_tick_ms(smoke_server* sv, double* ms, int* requests, int cap, int* count)
{
    if (!sv || ((!ms || !requests) && cap > 0)) return fail(SMOKE_ERR_INVALID, "server_tick_ms: NULL");
    std::lock_guard<std::mutex> lk(sv->mu);
    // Fold in the steps of the batch still running.
    if (sv->batch) {
        smoke_status st = smoke_batch_synchronize(sv->batch);
        if (st != SMOKE_OK) return st;
    }
    const int n = (int)std::min<size_t>((size_t)std::max(cap, 0), sv->tick_ms.size());
    const size_t first = sv->tick_ms.size() - (size_t)n;
    for (int i = 0; i < n; ++i) {
        ms[i] = sv->tick_ms[first + (size_t)i];
        requests[i] = sv->tick_requests[first + (size_t)i];
    }
    if (count) *count = n;
    return SMOKE_OK;
}

}  // extern "C"
