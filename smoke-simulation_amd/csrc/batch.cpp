// Batches of simulations (smoke_batch_* of include/smoke.h): one launch per
// stage for all member simulations, persistent kernels over (simulation,
// block) work lists (batch_kernels.hip), optional HIP-graph replay.
//
// A batch keeps, in one device allocation, a descriptor per simulation
// (smk::SimDesc: grid, field pointers, per-simulation physics and dt, and
// the ping-pong parities that the step flips on the device), the work tables
// of every stage and the source-fill table.  The host mirrors the parity
// flips in each smoke_sim so that the per-simulation read calls find the
// right halves.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sim_state.h"

using smk::fail;

struct smoke_batch {
    std::vector<smoke_sim*> sims;
    std::vector<float> dts;
    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    char* dmem = nullptr;
    smk::SimDesc* d_sims = nullptr;
    // [0] force, [1] faces 1..n-1, [2] interior, [3] jacobi, [4] velocity
    // advection columns, [5] density advection columns
    smk::StageTable* d_tab = nullptr;
    smk::SourceTable* d_src = nullptr;
    std::vector<smk::SimDesc> h_sims;
    smk::StageTable h_tab[6];
    int totals[6] = {0, 0, 0, 0, 0, 0};
    int src_total = 0;
    int iterations = 0;
    float omega = 1.0f;
    int plain = 1;
    smk::JacobiPlan plan;
    long long cells = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    std::vector<hipEvent_t> pool;
    std::vector<double> step_ms;
    bool broken = false;  // a member simulation was destroyed
    // Largest z-chunk of the fused sweeps every member's planes allow (the
    // 32-bit buffer span of a block, fused_max_zc); plans are capped to it.
    int zc_cap = 0;
    // Resident workgroups per CU (queried outside graph capture).
    int occ_cells = 1;
    int occ_adv_v = 1;
    int occ_adv_d = 1;
    int occ_jacobi[7] = {1, 1, 1, 1, 1, 1, 1};  // by sweeps per launch
};

namespace {
void refresh_occupancy(smoke_batch* b)
{
    b->occ_cells = smk::batch_cells_blocks_per_cu();
    b->occ_adv_v = smk::batch_advect_velocity_blocks_per_cu();
    b->occ_adv_d = smk::batch_advect_density_blocks_per_cu();
    for (int t = 1; t <= 6; ++t) b->occ_jacobi[t] = smk::batch_jacobi_blocks_per_cu(b->plan, t);
}
}  // namespace

void smk::batch_detach(smoke_batch* b, smoke_sim* s)
{
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (smoke_sim*& m : b->sims)
        if (m == s) m = nullptr;
    s->batch = nullptr;
    s->batch_stream = nullptr;
    b->broken = true;
}

namespace {

constexpr int kTabForce = 0, kTabFaces = 1, kTabInterior = 2, kTabJacobi = 3, kTabAdvV = 4,
              kTabAdvD = 5;

int ceil_div(int a, int b) { return (a + b - 1) / b; }

void fill_table(smk::StageTable& t, const std::vector<smoke_sim*>& sims, int kind,
                const smk::JacobiPlan& plan, int* total)
{
    int acc = 0;
    for (size_t s = 0; s < sims.size(); ++s) {
        const smk::Grid& g = sims[s]->g;
        int gx, gy, gz;
        if (kind == kTabJacobi) {
            const int T = plan.sweeps_per_launch;
            const int tx = smk::tile_width(plan.tile), ty = smk::tile_height(plan.tile);
            gx = ceil_div(g.nx - 2, tx - 2 * T);
            gy = ceil_div(g.ny - 2, ty - 2 * T);
            gz = ceil_div(g.nz - 2, plan.zc);
        } else if (kind == kTabAdvV) {
            gx = ceil_div(g.nx - 1, smk::kBlockX);
            gy = ceil_div(g.ny - 1, smk::kBlockY);
            gz = ceil_div(g.nz - 1, smk::kAdvChunkZ);
        } else if (kind == kTabAdvD) {
            gx = ceil_div(g.nx - 1, smk::kBlockX);  // x from 0
            gy = ceil_div(g.ny - 2, smk::kBlockY);
            gz = ceil_div(g.nz - 2, smk::kAdvChunkZ);
        } else {
            // Cells from x = 0 (the rows' line-aligned start; a stage starting
            // at x = 1 idles lane 0), y and z from 1 except for the force stage.
            const int sub = kind == kTabForce ? 0 : (kind == kTabFaces ? 1 : 2);
            const int wx = g.nx - sub + (kind == kTabForce ? 0 : 1);
            const int wy = g.ny - (kind == kTabForce ? 1 : sub), wz = g.nz - sub;
            gx = ceil_div(wx, smk::kBlockX);
            gy = ceil_div(wy, smk::kBlockY);
            gz = ceil_div(wz, smk::kColZ);
        }
        t.start[s] = acc;
        t.gx[s] = gx;
        t.gy[s] = gy;
        acc += gx * gy * gz;
    }
    t.start[sims.size()] = acc;
    *total = acc;
}

// Persistent grid: enough workgroups to fill every CU, a multiple of 8 (the
// kernels split the work list into one contiguous share per XCD).
int persistent_grid(int blocks_per_cu, int cus, int total)
{
    long long g = (long long)blocks_per_cu * cus;
    if (g > total) g = total;
    g = (g + 7) / 8 * 8;
    if (g < 8) g = 8;
    return (int)g;
}

hipError_t upload_tables(smoke_batch* b)
{
    hipError_t e = hipMemcpyAsync(b->d_sims, b->h_sims.data(), b->h_sims.size() * sizeof(smk::SimDesc),
                                  hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_tab, b->h_tab, sizeof b->h_tab, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    return e;
}

hipError_t rebuild_jacobi_table(smoke_batch* b)
{
    refresh_occupancy(b);
    fill_table(b->h_tab[kTabJacobi], b->sims, kTabJacobi, b->plan, &b->totals[kTabJacobi]);
    return hipMemcpy(&b->d_tab[kTabJacobi], &b->h_tab[kTabJacobi], sizeof(smk::StageTable),
                     hipMemcpyHostToDevice);
}

// K sweeps of every simulation: launches of plan.sweeps_per_launch fused
// sweeps, the remainder in one shorter launch.  Returns the launch count.
hipError_t enqueue_sweeps(smoke_batch* b, const smk::JacobiPlan& plan, int* launches)
{
    const int n = (int)b->sims.size();
    int l = 0;
    for (int left = b->iterations; left > 0; ++l) {
        const int t = std::min(left, plan.sweeps_per_launch);
        const smk::StageTable* tab = t == 1 ? &b->d_tab[kTabInterior] : &b->d_tab[kTabJacobi];
        const int total = t == 1 ? b->totals[kTabInterior] : b->totals[kTabJacobi];
        const int grid = persistent_grid(b->occ_jacobi[t], b->cus, total);
        hipError_t e = smk::batch_jacobi(b->d_sims, tab, n, total, grid, plan, t, l & 1, l == 0,
                                         b->omega, b->plain, b->stream);
        if (e != hipSuccess) return e;
        left -= t;
    }
    *launches = l;
    return hipSuccess;
}

hipError_t enqueue_step(smoke_batch* b, int* sweep_launches)
{
    const int n = (int)b->sims.size();
    const int gc = persistent_grid(b->occ_cells, b->cus, 1 << 30);
    hipError_t e = smk::batch_sources(b->d_sims, b->d_src, b->src_total,
                                      std::min(gc, std::max(8, (b->src_total + 7) / 8 * 8)), b->stream);
    auto cells = [&](int stage, int tab, int lp) {
        return smk::batch_cells(stage, b->d_sims, &b->d_tab[tab], n, b->totals[tab],
                                std::min(gc, std::max(8, (b->totals[tab] + 7) / 8 * 8)), lp, b->stream);
    };
    if (e == hipSuccess) e = cells(smk::BS_FORCE, kTabForce, 0);
    if (e == hipSuccess) e = cells(smk::BS_DIV, kTabInterior, 0);
    int L = 0;
    if (e == hipSuccess) e = enqueue_sweeps(b, b->plan, &L);
    // K = 0: P[pcur] still holds the previous step's pressure; the batch
    // requires K >= 1 (checked at creation), so the gradient always reads
    // this step's result.
    if (e == hipSuccess) e = cells(smk::BS_GRAD, kTabFaces, L & 1);
    if (e == hipSuccess)
        e = smk::batch_advect_velocity(
            b->d_sims, &b->d_tab[kTabAdvV], n, b->totals[kTabAdvV],
            persistent_grid(b->occ_adv_v, b->cus, b->totals[kTabAdvV]), b->stream);
    if (e == hipSuccess)
        e = smk::batch_advect_density(
            b->d_sims, &b->d_tab[kTabAdvD], n, b->totals[kTabAdvD],
            persistent_grid(b->occ_adv_d, b->cus, b->totals[kTabAdvD]), b->stream);
    if (e == hipSuccess) e = smk::batch_flip(b->d_sims, n, L & 1, b->stream);
    *sweep_launches = L;
    return e;
}

void drop_graph(smoke_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

hipError_t collect(smoke_batch* b)
{
    for (auto& pr : b->pending) {
        hipError_t e = hipEventSynchronize(pr.second);
        if (e != hipSuccess) return e;
        float ms = 0.0f;
        e = hipEventElapsedTime(&ms, pr.first, pr.second);
        if (e != hipSuccess) return e;
        b->step_ms.push_back(ms);
        b->pool.push_back(pr.first);
        b->pool.push_back(pr.second);
    }
    b->pending.clear();
    return hipSuccess;
}

hipError_t pool_event(smoke_batch* b, hipEvent_t* ev)
{
    if (!b->pool.empty()) {
        *ev = b->pool.back();
        b->pool.pop_back();
        return hipSuccess;
    }
    return hipEventCreate(ev);
}

}  // namespace

extern "C" {

smoke_status smoke_batch_create(smoke_sim* const* sims, int count, const float* dts,
                                smoke_batch** out)
{
    if (!out || !sims || !dts) return fail(SMOKE_ERR_INVALID, "batch_create: NULL argument");
    *out = nullptr;
    if (count < 1 || count > SMOKE_MAX_BATCH)
        return fail(SMOKE_ERR_INVALID, "batch_create: count %d not in 1..%d", count, SMOKE_MAX_BATCH);
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    int sources = 0;
    for (int i = 0; i < count; ++i) {
        smoke_sim* s = sims[i];
        if (!s) return fail(SMOKE_ERR_INVALID, "batch_create: simulation %d is NULL", i);
        if (s->batch) return fail(SMOKE_ERR_STATE, "batch_create: simulation %d is already in a batch", i);
        for (int j = 0; j < i; ++j)
            if (sims[j] == s) return fail(SMOKE_ERR_INVALID, "batch_create: simulation %d listed twice", i);
        if (s->device != dev) return fail(SMOKE_ERR_INVALID, "batch_create: simulation %d on another device", i);
        const smoke_params& p = s->prm;
        const smoke_params& p0 = sims[0]->prm;
        if (s->storage != SMOKE_STORAGE_FP32)
            return fail(SMOKE_ERR_INVALID, "batch_create: simulation %d stores 16-bit fields (batches are FP32)", i);
        if (p.obstacle_velocity)
            return fail(SMOKE_ERR_INVALID, "batch_create: simulation %d uses the moving-obstacle condition (batches do not)", i);
        if (p.vorticity != 0.0f)
            return fail(SMOKE_ERR_INVALID, "batch_create: simulation %d uses vorticity confinement (batches do not)", i);
        if (p.projection != SMOKE_PROJ_JACOBI || p.warm_start)
            return fail(SMOKE_ERR_INVALID, "batch_create: simulation %d is not a cold-start Jacobi projection", i);
        if (p.iterations < 1 || p.iterations != p0.iterations || p.jacobi_omega != p0.jacobi_omega)
            return fail(SMOKE_ERR_INVALID, "batch_create: simulations differ in iterations/jacobi_omega (or K < 1)");
        if (!std::isfinite(dts[i])) return fail(SMOKE_ERR_INVALID, "batch_create: dt %d not finite", i);
        int ns = 0;
        for (const smk::Object& o : s->objects) ns += (o.type == 1);
        if (ns > 8) return fail(SMOKE_ERR_INVALID, "batch_create: simulation %d has more than 8 sources", i);
        if (smk::fused_max_zc(s->g.plane) < 1)
            return fail(SMOKE_ERR_INVALID, "batch_create: simulation %d's planes are too large to batch", i);
        sources += ns;
    }
    smoke_batch* b = new smoke_batch();
    b->sims.assign(sims, sims + count);
    b->dts.assign(dts, dts + count);
    b->device = dev;
    b->iterations = sims[0]->prm.iterations;
    b->omega = (float)sims[0]->prm.jacobi_omega;
    b->plain = sims[0]->prm.jacobi_omega == 1.0 ? 1 : 0;
    b->zc_cap = 1 << 30;
    for (int i = 0; i < count; ++i) b->zc_cap = std::min(b->zc_cap, smk::fused_max_zc(sims[i]->g.plane));
    b->plan = sims[0]->plan;
    b->plan.zc = std::min(b->plan.zc, b->zc_cap);
    auto bail = [&](smoke_status st) {
        smoke_batch_destroy(b);
        return st;
    };
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return bail(fail(SMOKE_ERR_HIP, "batch_create: device query"));
    b->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(SMOKE_ERR_HIP, "batch_create: stream"));
    // Freeze the scene: masks up to date, all earlier work on the
    // simulations' own streams finished.
    for (smoke_sim* s : b->sims) {
        std::lock_guard<std::mutex> lk(s->mu);
        if (smk::refresh_mask(s, nullptr) != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess)
            return bail(fail(SMOKE_ERR_HIP, "batch_create: preparing a simulation failed"));
    }
    // Descriptors.
    b->h_sims.resize(count);
    for (int i = 0; i < count; ++i) {
        smoke_sim* s = b->sims[i];
        smk::SimDesc& d = b->h_sims[i];
        d.g = s->g;
        for (int h = 0; h < 2; ++h) {
            d.rho[h] = s->rho[h];
            d.u[h] = s->u[h];
            d.v[h] = s->v[h];
            d.w[h] = s->w[h];
            d.P[h] = s->P[h];
        }
        d.d = s->d;
        d.flags = s->flags;
        d.force.gravity = s->prm.gravity;
        d.force.buoyancy = s->prm.buoyancy;
        d.force.dt = b->dts[i];
        d.force.max_v2 = s->prm.max_velocity * s->prm.max_velocity;
        d.force.gate = s->prm.density_gate;
        d.force.gate_on = s->prm.density_gate > 0.0f;
        d.now = 1 - s->now;  // the parity the next step works on
        d.pcur = s->pcur;
        b->cells += (long long)s->g.nx * s->g.ny * s->g.nz;
    }
    fill_table(b->h_tab[kTabForce], b->sims, kTabForce, b->plan, &b->totals[kTabForce]);
    fill_table(b->h_tab[kTabFaces], b->sims, kTabFaces, b->plan, &b->totals[kTabFaces]);
    fill_table(b->h_tab[kTabInterior], b->sims, kTabInterior, b->plan, &b->totals[kTabInterior]);
    fill_table(b->h_tab[kTabJacobi], b->sims, kTabJacobi, b->plan, &b->totals[kTabJacobi]);
    fill_table(b->h_tab[kTabAdvV], b->sims, kTabAdvV, b->plan, &b->totals[kTabAdvV]);
    fill_table(b->h_tab[kTabAdvD], b->sims, kTabAdvD, b->plan, &b->totals[kTabAdvD]);
    refresh_occupancy(b);
    // Source table: each source's bounding box, as the single-simulation
    // source fill computes it.
    std::vector<smk::SourceTable> hsrc(1);
    smk::SourceTable& st = hsrc[0];
    st.count = 0;
    int acc = 0;
    for (int i = 0; i < count; ++i) {
        const smk::Grid& g = b->sims[i]->g;
        for (const smk::Object& o : b->sims[i]->objects) {
            if (o.type != 1) continue;
            int lo[3], ext[3];
            if (!smk::source_box(g, make_float4(o.pos[0], o.pos[1], o.pos[2], o.r), lo, ext)) continue;
            const int k = st.count++;
            st.start[k] = acc;
            st.sim[k] = i;
            st.sph[k] = make_float4(o.pos[0], o.pos[1], o.pos[2], o.r);
            st.lo[k] = make_int4(lo[0], lo[1], lo[2], 0);
            const int gxb = ceil_div(ext[0], smk::kBlockX);
            st.ext[k] = make_int4(ext[0], ext[1], ext[2], gxb);
            acc += gxb * ceil_div(ext[1], smk::kBlockY) * ext[2];
        }
    }
    st.start[st.count] = acc;
    b->src_total = acc;
    (void)sources;
    const size_t bytes = count * sizeof(smk::SimDesc) + 256 + sizeof b->h_tab + 256 + sizeof(smk::SourceTable);
    if (hipMalloc((void**)&b->dmem, bytes) != hipSuccess) {
        b->dmem = nullptr;
        return bail(fail(SMOKE_ERR_NOMEM, "batch_create: cannot allocate batch tables"));
    }
    char* p = b->dmem;
    b->d_sims = reinterpret_cast<smk::SimDesc*>(p);
    p += (count * sizeof(smk::SimDesc) + 255) / 256 * 256;
    b->d_tab = reinterpret_cast<smk::StageTable*>(p);
    p += (sizeof b->h_tab + 255) / 256 * 256;
    b->d_src = reinterpret_cast<smk::SourceTable*>(p);
    hipError_t e = hipMemcpyAsync(b->d_src, &st, sizeof st, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = upload_tables(b);
    if (e != hipSuccess) return bail(fail(SMOKE_ERR_HIP, "batch_create: upload: %s", hipGetErrorString(e)));
    for (smoke_sim* s : b->sims) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->batch = b;
        s->batch_stream = b->stream;
    }
    *out = b;
    return SMOKE_OK;
}

void smoke_batch_destroy(smoke_batch* b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (smoke_sim* s : b->sims) {
        if (s && s->batch == b) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->batch = nullptr;
            s->batch_stream = nullptr;
        }
    }
    drop_graph(b);
    for (auto& pr : b->pending) {
        (void)hipEventDestroy(pr.first);
        (void)hipEventDestroy(pr.second);
    }
    for (hipEvent_t ev : b->pool) (void)hipEventDestroy(ev);
    if (b->dmem) (void)hipFree(b->dmem);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

smoke_status smoke_batch_step(smoke_batch* b, int use_graph)
{
    if (!b) return fail(SMOKE_ERR_INVALID, "batch_step: NULL batch");
    if (b->broken) return fail(SMOKE_ERR_STATE, "batch_step: a simulation of the batch was destroyed");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(pool_event(b, &e0));
    HIP_TRY(pool_event(b, &e1));
    HIP_TRY(hipEventRecord(e0, b->stream));
    int L = 0;
    if (use_graph) {
        if (!b->exec) {
            HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = enqueue_step(b, &L);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(b->stream, &g);
            if (e != hipSuccess || e2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return fail(SMOKE_ERR_HIP, "batch_step: graph capture failed: %s",
                            hipGetErrorString(e != hipSuccess ? e : e2));
            }
            b->graph = g;
            HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        } else {
            L = (b->iterations + b->plan.sweeps_per_launch - 1) / b->plan.sweeps_per_launch;
        }
        HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    } else {
        hipError_t e = enqueue_step(b, &L);
        if (e != hipSuccess) return fail(SMOKE_ERR_HIP, "batch_step: launch failed: %s", hipGetErrorString(e));
    }
    HIP_TRY(hipEventRecord(e1, b->stream));
    b->pending.emplace_back(e0, e1);
    if (b->pending.size() > 256) HIP_TRY(collect(b));
    // Mirror the device-side parity flips.
    for (size_t i = 0; i < b->sims.size(); ++i) {
        smoke_sim* s = b->sims[i];
        std::lock_guard<std::mutex> lk(s->mu);
        s->now = 1 - s->now;
        s->pcur ^= (L & 1);
        s->steps_enqueued += 1;
        s->rhs_valid = true;  // the batched divergence stage fills d
        s->last_dt = b->dts[i];
    }
    return SMOKE_OK;
}

smoke_status smoke_batch_synchronize(smoke_batch* b)
{
    if (!b) return fail(SMOKE_ERR_INVALID, "batch_synchronize: NULL batch");
    HIP_TRY(hipStreamSynchronize(b->stream));
    return SMOKE_OK;
}

smoke_status smoke_batch_set_jacobi_plan(smoke_batch* b, const smoke_jacobi_plan* plan)
{
    if (!b || !plan) return fail(SMOKE_ERR_INVALID, "batch_set_jacobi_plan: NULL");
    if (b->broken) return fail(SMOKE_ERR_STATE, "batch_set_jacobi_plan: a simulation was destroyed");
    if (plan->sweeps_per_launch < 1 || plan->sweeps_per_launch > 6 || plan->zc < 1 ||
        plan->tile < SMOKE_TILE_32x32 || plan->tile > SMOKE_TILE_32x64)
        return fail(SMOKE_ERR_INVALID, "batch_set_jacobi_plan: invalid plan");
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->plan.sweeps_per_launch = plan->sweeps_per_launch;
    b->plan.tile = plan->tile;
    b->plan.zc = std::min(plan->zc, b->zc_cap);
    drop_graph(b);
    HIP_TRY(rebuild_jacobi_table(b));
    return SMOKE_OK;
}

smoke_status smoke_batch_get_jacobi_plan(const smoke_batch* b, smoke_jacobi_plan* out)
{
    if (!b || !out) return fail(SMOKE_ERR_INVALID, "batch_get_jacobi_plan: NULL");
    out->sweeps_per_launch = b->plan.sweeps_per_launch;
    out->tile = b->plan.tile;
    out->zc = b->plan.zc;
    return SMOKE_OK;
}

smoke_status smoke_batch_autotune(smoke_batch* b, smoke_jacobi_plan* chosen, double* ms_per_sweep)
{
    if (!b) return fail(SMOKE_ERR_INVALID, "batch_autotune: NULL batch");
    if (b->broken) return fail(SMOKE_ERR_STATE, "batch_autotune: a simulation was destroyed");
    HIP_TRY(hipStreamSynchronize(b->stream));
    drop_graph(b);
    hipEvent_t ev[2];
    HIP_TRY(hipEventCreate(&ev[0]));
    HIP_TRY(hipEventCreate(&ev[1]));
    smk::JacobiPlan best = b->plan;
    double best_ms = -1.0;
    const int tiles[4] = {smk::JacobiPlan::TILE_32x32, smk::JacobiPlan::TILE_64x16,
                          smk::JacobiPlan::TILE_64x32, smk::JacobiPlan::TILE_32x64};
    const int zcs[4] = {32, 64, 128, 256};
    hipError_t e = hipSuccess;
    for (int t = 2; t <= 6 && e == hipSuccess; ++t)
        for (int tile : tiles)
            for (int zc : zcs) {
                if (e != hipSuccess) break;  // a failed candidate ends the search
                b->plan.sweeps_per_launch = t;
                b->plan.tile = tile;
                b->plan.zc = std::min(zc, b->zc_cap);
                e = rebuild_jacobi_table(b);
                int L = 0;
                if (e == hipSuccess) e = enqueue_sweeps(b, b->plan, &L);  // warm-up
                if (e == hipSuccess) e = hipEventRecord(ev[0], b->stream);
                for (int r = 0; r < 2 && e == hipSuccess; ++r) e = enqueue_sweeps(b, b->plan, &L);
                if (e == hipSuccess) e = hipEventRecord(ev[1], b->stream);
                if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
                float ms = 0.0f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
                if (e != hipSuccess) break;
                const double per = ms / (2.0 * b->iterations);
                if (best_ms < 0.0 || per < best_ms) {
                    best_ms = per;
                    best = b->plan;
                }
            }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    b->plan = best;
    if (e == hipSuccess) e = rebuild_jacobi_table(b);
    if (e != hipSuccess) return fail(SMOKE_ERR_HIP, "batch_autotune: %s", hipGetErrorString(e));
    if (chosen) {
        chosen->sweeps_per_launch = best.sweeps_per_launch;
        chosen->tile = best.tile;
        chosen->zc = best.zc;
    }
    if (ms_per_sweep) *ms_per_sweep = best_ms;
    return SMOKE_OK;
}

smoke_status smoke_batch_time_sweeps(smoke_batch* b, int reps, double* ms_per_launch, int* launches)
{
    if (!b || reps < 1) return fail(SMOKE_ERR_INVALID, "batch_time_sweeps: NULL batch or reps < 1");
    if (b->broken) return fail(SMOKE_ERR_STATE, "batch_time_sweeps: a simulation was destroyed");
    HIP_TRY(hipStreamSynchronize(b->stream));
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    int L = 0;
    if (e == hipSuccess) e = enqueue_sweeps(b, b->plan, &L);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(ev[0], b->stream);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = enqueue_sweeps(b, b->plan, &L);
    if (e == hipSuccess) e = hipEventRecord(ev[1], b->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    if (e != hipSuccess) return fail(SMOKE_ERR_HIP, "batch_time_sweeps: %s", hipGetErrorString(e));
    if (ms_per_launch) *ms_per_launch = ms / (double)(reps * L);
    if (launches) *launches = L;
    return SMOKE_OK;
}

long long smoke_batch_interior_cells(const smoke_batch* b)
{
    if (!b) return 0;
    long long n = 0;
    for (const smk::SimDesc& d : b->h_sims) n += (long long)(d.g.nx - 2) * (d.g.ny - 2) * (d.g.nz - 2);
    return n;
}

smoke_status smoke_batch_step_ms(smoke_batch* b, double* ms, int cap, int* count)
{
    if (!b || (!ms && cap > 0)) return fail(SMOKE_ERR_INVALID, "batch_step_ms: NULL");
    HIP_TRY(collect(b));
    const int n = (int)std::min<size_t>((size_t)std::max(cap, 0), b->step_ms.size());
    const size_t first = b->step_ms.size() - (size_t)n;
    for (int i = 0; i < n; ++i) ms[i] = b->step_ms[first + (size_t)i];
    if (count) *count = n;
    return SMOKE_OK;
}

long long smoke_batch_cells(const smoke_batch* b) { return b ? b->cells : 0; }

}  // extern "C"
