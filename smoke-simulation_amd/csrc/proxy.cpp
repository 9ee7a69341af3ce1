// C ABI of the proxy-plane slicing geometry (include/smoke.h) over the
// kernel of proxy_kernels.hip.
//
// smoke_proxy_geometry_device enqueues the kernel on the caller's stream into
// the caller's device arrays.  smoke_proxy_geometry serves host arrays: it
// keeps, per device, one non-blocking stream and one pinned, host-coherent
// workspace holding the largest view asked for so far; the kernel writes its
// arrays straight into that memory (a few KB per view) and, last, a sequence
// number the host waits for by polling the workspace (a stream
// synchronisation wakes the caller tens of microseconds after the kernel
// ends; the poll sees the number about a microsecond after it is written).
// The poll gives up after kPollLimit and falls back to synchronising the
// stream, which also reports a failed kernel.  Then the vertices and indices
// produced are copied into the caller's arrays.
#include <hip/hip_runtime.h>

#include <immintrin.h>

#include <chrono>
#include <cstring>
#include <mutex>
#include <vector>

#include "sim_state.h"

using smk::fail;

namespace {

struct ProxyWorkspace {
    int device = -1;
    hipStream_t stream = nullptr;
    char* mem = nullptr;  // pinned host memory the kernel writes
    int steps = 0;        // capacity in planes
    unsigned seq = 0;     // number of the last view asked for
};

constexpr std::chrono::milliseconds kPollLimit{20};

std::mutex g_proxy_mu;
std::vector<ProxyWorkspace> g_proxy_ws;

size_t vert_bytes(int steps) { return (size_t)steps * SMOKE_PROXY_VERTICES_PER_PLANE * 3 * sizeof(float); }
size_t index_bytes(int steps) { return (size_t)steps * SMOKE_PROXY_INDICES_PER_PLANE * sizeof(int16_t); }
size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

smoke_status check_view(const float* a, const float* b, const float* c, const float* d, int steps,
                        smk::ProxyView& v)
{
    if (!a || !b || !c || !d) return fail(SMOKE_ERR_INVALID, "proxy_geometry: NULL vector");
    if (steps < 1 || steps > SMOKE_PROXY_MAX_STEPS)
        return fail(SMOKE_ERR_INVALID, "proxy_geometry: steps %d not in 1..%d", steps, SMOKE_PROXY_MAX_STEPS);
    for (int k = 0; k < 3; ++k) {
        v.box_pos[k] = a[k];
        v.box_dims[k] = b[k];
        v.origin[k] = c[k];
        v.dir[k] = d[k];
    }
    return SMOKE_OK;
}

// Workspace of the current device holding `steps` planes (caller holds the
// lock).
smoke_status workspace(int steps, ProxyWorkspace** out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    ProxyWorkspace* ws = nullptr;
    for (ProxyWorkspace& w : g_proxy_ws)
        if (w.device == dev) ws = &w;
    if (!ws) {
        g_proxy_ws.emplace_back();
        ws = &g_proxy_ws.back();
        ws->device = dev;
        HIP_TRY(hipStreamCreateWithFlags(&ws->stream, hipStreamNonBlocking));
    }
    if (ws->steps < steps) {
        if (ws->mem) {
            HIP_TRY(hipStreamSynchronize(ws->stream));  // no kernel still writes it
            HIP_TRY(hipHostFree(ws->mem));
        }
        ws->mem = nullptr;
        ws->steps = 0;
        const size_t bytes = align256(vert_bytes(steps)) + align256(index_bytes(steps)) +
                             align256((size_t)steps * sizeof(smoke_proxy_plane)) + 256;
        HIP_TRY(hipHostMalloc((void**)&ws->mem, bytes, hipHostMallocCoherent));
        std::memset(ws->mem, 0, bytes);  // no stale completion mark
        ws->steps = steps;
    }
    *out = ws;
    return SMOKE_OK;
}

}  // namespace

smoke_status smoke_proxy_geometry_device(const float box_position[3], const float box_dimensions[3],
                                         const float view_origin[3], const float view_direction[3],
                                         int steps, float* vertices, int16_t* indices,
                                         smoke_proxy_plane* planes, int* counts, void* stream)
{
    smk::ProxyView v;
    smoke_status st = check_view(box_position, box_dimensions, view_origin, view_direction, steps, v);
    if (st != SMOKE_OK) return st;
    if (!vertices || !indices || !planes || !counts)
        return fail(SMOKE_ERR_INVALID, "proxy_geometry_device: NULL output");
    HIP_TRY(smk::proxy_geometry(v, steps, vertices, indices, planes, counts,
                                static_cast<hipStream_t>(stream)));
    return SMOKE_OK;
}

smoke_status smoke_proxy_geometry(const float box_position[3], const float box_dimensions[3],
                                  const float view_origin[3], const float view_direction[3],
                                  int steps, float* vertices, int16_t* indices,
                                  smoke_proxy_plane* planes, int* n_vertices, int* n_indices)
{
    smk::ProxyView v;
    smoke_status st = check_view(box_position, box_dimensions, view_origin, view_direction, steps, v);
    if (st != SMOKE_OK) return st;
    if (!vertices || !indices || !planes || !n_vertices || !n_indices)
        return fail(SMOKE_ERR_INVALID, "proxy_geometry: NULL output");
    std::lock_guard<std::mutex> lk(g_proxy_mu);
    ProxyWorkspace* ws = nullptr;
    if ((st = workspace(steps, &ws)) != SMOKE_OK) return st;
    char* p = ws->mem;
    float* dv = reinterpret_cast<float*>(p);
    p += align256(vert_bytes(ws->steps));
    int16_t* di = reinterpret_cast<int16_t*>(p);
    p += align256(index_bytes(ws->steps));
    smoke_proxy_plane* dp = reinterpret_cast<smoke_proxy_plane*>(p);
    p += align256((size_t)ws->steps * sizeof(smoke_proxy_plane));
    int* dc = reinterpret_cast<int*>(p);
    unsigned* done = reinterpret_cast<unsigned*>(p + 64);
    const unsigned seq = ++ws->seq == 0 ? ++ws->seq : ws->seq;  // 0 never marks a view
    HIP_TRY(smk::proxy_geometry(v, steps, dv, di, dp, dc, ws->stream, done, seq));
    const auto t0 = std::chrono::steady_clock::now();
    while (__atomic_load_n(done, __ATOMIC_ACQUIRE) != seq) {
        if (std::chrono::steady_clock::now() - t0 > kPollLimit) {
            HIP_TRY(hipStreamSynchronize(ws->stream));
            if (__atomic_load_n(done, __ATOMIC_ACQUIRE) != seq)
                return fail(SMOKE_ERR_HIP, "proxy_geometry: the kernel ended without its completion mark");
            break;
        }
        _mm_pause();
    }
    const int counts[2] = {dc[0], dc[1]};
    std::memcpy(vertices, dv, (size_t)counts[0] * 3 * sizeof(float));
    std::memcpy(indices, di, (size_t)counts[1] * sizeof(int16_t));
    std::memcpy(planes, dp, (size_t)steps * sizeof(smoke_proxy_plane));
    *n_vertices = counts[0];
    *n_indices = counts[1];
    return SMOKE_OK;
}
