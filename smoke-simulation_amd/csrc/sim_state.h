// Private host-side state of a simulation, shared by the runtime
// (runtime.cpp) and the batch/serving layer (batch.cpp).  Not part of the
// public ABI (include/smoke.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <deque>
#include <mutex>
#include <vector>

#include "smoke_internal.h"

namespace smk {

// Records the message for smoke_last_error() (per thread) and returns st.
smoke_status fail(smoke_status st, const char* fmt, ...);

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return smk::fail(SMOKE_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

struct Object {
    float pos[3];
    float vel[3];
    float r;
    int type;  // 0 obstacle, 1 source (smokeSimulation.cu:50)
};

}  // namespace smk

struct smoke_sim {
    std::mutex mu;
    smk::Grid g{};
    int device = 0;
    hipStream_t stream = nullptr;
    char* arena = nullptr;
    float* rho[2] = {nullptr, nullptr};
    float* u[2] = {nullptr, nullptr};
    float* v[2] = {nullptr, nullptr};
    float* w[2] = {nullptr, nullptr};
    float* P[2] = {nullptr, nullptr};
    float* d = nullptr;
    // SMOKE_STORAGE_BF16 (FP16): density and u, v, w live in hrho, hu, hv, hw
    // (frho, fu, fv, fw) instead (the float pointers above are then NULL);
    // the *_VELOCITY forms: u, v, w live there and the density in rho.
    int storage = SMOKE_STORAGE_FP32;
    smk::bf16s* hrho[2] = {nullptr, nullptr};
    smk::bf16s* hu[2] = {nullptr, nullptr};
    smk::bf16s* hv[2] = {nullptr, nullptr};
    smk::bf16s* hw[2] = {nullptr, nullptr};
    smk::f16s* frho[2] = {nullptr, nullptr};
    smk::f16s* fu[2] = {nullptr, nullptr};
    smk::f16s* fv[2] = {nullptr, nullptr};
    smk::f16s* fw[2] = {nullptr, nullptr};
    // REFINED solver buffers (allocated when the projection is first set to
    // REFINED).
    char* refine_mem = nullptr;
    smk::RefineBufs rb;
    // Vorticity-confinement scratch (allocated when vorticity is first set
    // above 0).
    char* vort_mem = nullptr;
    smk::VortBufs vb{};
    // Moving-obstacle boundary condition: per-cell owner bytes (allocated
    // when obstacle_velocity is first set; built with the mask).
    uint8_t* owner = nullptr;
    // Step norms: partial records and the result (allocated on first use).
    // With norms_on every step ends with the reduction; norms_step is the
    // value of steps_enqueued the stored result describes (-1: none).
    // rhs_valid: d holds the right-hand side of the last projection stage.
    void* norms_mem = nullptr;
    bool norms_on = false;
    long long norms_step = -1;
    bool rhs_valid = false;
    float last_dt = 0.0f;
    float* staging = nullptr;
    uint8_t* base = nullptr;
    uint8_t* s = nullptr;
    uint8_t* flags = nullptr;

    smoke_params prm{};
    std::vector<smk::Object> objects;
    int now = 1;   // indexNow (smokeSimulation.cu:707)
    int pcur = 0;  // P buffer holding the latest pressure
    bool mask_dirty = true;
    smk::JacobiPlan plan;  // launch plan of the fused Jacobi sweeps
    smoke_batch* batch = nullptr;         // batch this simulation belongs to
    hipStream_t batch_stream = nullptr;   // that batch's stream
    long long steps_enqueued = 0;

    // Asynchronous read-back: a copy stream, two device pack slots and two
    // pinned host buffers (allocated on first use).
    struct Readback {
        int slot;
        long long step;
    };
    hipStream_t copy_stream = nullptr;
    void* rb_dev = nullptr;          // 2 slots of nx*ny*nz*4 bytes
    void* rb_host[2] = {nullptr, nullptr};
    hipEvent_t rb_step_done = nullptr, rb_packed[2] = {nullptr, nullptr},
               rb_copied[2] = {nullptr, nullptr};
    std::deque<Readback> rb_pending;
    int rb_next = 0;

    // Step graphs (smoke_set_step_graph): a step captured into a HIP graph
    // per ping-pong parity and replayed.  `version` counts the changes that
    // alter what a step enqueues (parameters, plan, objects, solids); a graph
    // captured at another version, dt or pressure parity is captured again.
    // The host-side parities a captured step leaves are kept with its graph.
    struct StepGraph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        unsigned long long version = 0;
        float dt = 0.0f;
        int pcur_in = -1, now_out = 0, pcur_out = 0;
        bool rhs_out = false;
    };
    bool graph_on = false;
    unsigned long long version = 1;
    StepGraph graphs[2];

    // Timing: events at every stage boundary of a step, plus two around the
    // projection's sweep launches (the step's dominant kernel).
    bool timing = false;
    struct StepEvents {
        hipEvent_t ev[SMOKE_STAGE_COUNT + 1];
        long long launches[SMOKE_STAGE_COUNT];
        hipEvent_t sweep_ev[2];
        long long sweeps;
    };
    StepEvents* cur = nullptr;  // events of the step being enqueued
    std::vector<StepEvents> pending;
    std::vector<hipEvent_t> event_pool;
    double stage_ms[SMOKE_STAGE_COUNT] = {0};
    long long stage_launches[SMOKE_STAGE_COUNT] = {0};
    double sweep_ms = 0.0;
    long long sweep_launches = 0;
    long long timed_steps = 0;
    std::deque<double> step_ms;
};

namespace smk {
// Calls fn(rho, u, v, w) with the simulation's field pointer pairs typed by
// its storage (float* [2], bf16s* [2] or f16s* [2], density and velocity
// each).
template <typename Fn>
auto with_fields(smoke_sim* s, Fn&& fn)
{
    switch (s->storage) {
    case SMOKE_STORAGE_BF16: return fn(s->hrho, s->hu, s->hv, s->hw);
    case SMOKE_STORAGE_BF16_VELOCITY: return fn(s->rho, s->hu, s->hv, s->hw);
    case SMOKE_STORAGE_FP16: return fn(s->frho, s->fu, s->fv, s->fw);
    case SMOKE_STORAGE_FP16_VELOCITY: return fn(s->rho, s->fu, s->fv, s->fw);
    default: return fn(s->rho, s->u, s->v, s->w);
    }
}
// Enqueues the step-norms reduction of the simulation's current state on its
// stream, with `scratch` (norms_scratch_bytes()) for the partial records and
// the public record written to `out` (device memory).
hipError_t enqueue_norms(smoke_sim* s, void* scratch, smoke_step_norms* out);
// Rebuilds mask and flags if the base mask or an obstacle changed.
hipError_t refresh_mask(smoke_sim* s, long long* launches);
// Called by smoke_destroy for a simulation that is still in a batch: waits
// for the batch and marks it unusable (its steps would touch freed memory).
void batch_detach(smoke_batch* b, smoke_sim* s);
}  // namespace smk

// Rejects calls that would change or step a simulation owned by a batch.
#define SMOKE_REFUSE_IN_BATCH(s, what)                                                         \
    do {                                                                                   \
        if ((s)->batch)                                                                    \
            return smk::fail(SMOKE_ERR_STATE, "%s: the simulation belongs to a batch", what); \
    } while (0)

// Work of the batch owning the simulation is finished before a read.
#define SMOKE_WAIT_BATCH(s)                                                                \
    do {                                                                                   \
        if ((s)->batch_stream) HIP_TRY(hipStreamSynchronize((s)->batch_stream));           \
    } while (0)
