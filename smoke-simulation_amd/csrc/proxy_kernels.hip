// HIP kernel of the renderer's proxy-plane slicing geometry for gfx950:
// BoundingBox::generateProxyGeometry (project/boundingBox.cpp:133-183) with
// getMinMaxPoints (:231-241), planeBoxIntersections (:285-354), duplicate
// (:356-361), sortPoints (:259-272), pseudoangle (:87-92) and teselate
// (:243-255), computed on the device instead of sequentially on the CPU.
//
// One workgroup of nt = min(1024, steps rounded up to 64) lanes; lane j of
// chunk c builds plane k = c * nt + j: its edge points (at most 12, plus their mean) in
// registers, sorted with the same insertion sort, then a workgroup prefix sum
// over the per-plane vertex and index counts places every plane's vertices,
// fan triangles and (num_triangles, triangle_offset) entry exactly where the
// reference's sequential appends put them.  Vector arithmetic is FP32 in
// glm 0.9.9's operation order (dot = (x*x' + y*y') + z*z', normalize = v * (1
// / sqrt(dot(v, v))), cross, component-wise vec3 / scalar); the file is
// compiled with -ffp-contract=off and HIP's correctly rounded division and
// square root, so the vertices are bit-identical to the CPU oracle's.
#include <algorithm>

#include "smoke_internal.h"

namespace smk {

namespace {

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 mul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 muls(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 divs(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(V3 a, V3 b)
{
    const V3 t = mul(a, b);
    return t.x + t.y + t.z;
}
__device__ __forceinline__ V3 normalize(V3 v) { return muls(v, 1.0f / sqrtf(dot(v, v))); }
__device__ __forceinline__ V3 cross(V3 a, V3 b)
{
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}

// The ratio in FP32; the reference's offsets are double literals and the
// result returns as float.
__device__ __forceinline__ float pseudoangle(float y, float x)
{
    const float r = x / (fabsf(x) + fabsf(y));
    if (y < 0) return (float)(1.0 + (double)r);
    return (float)(3.0 - (double)r);
}

// Edge points of one plane (12 at most: points of 0/0 edges are NaN and never
// equal, so none of them is dropped) and their mean, last.
constexpr int kEdges = 12;
constexpr int kMaxPoints = kEdges + 1;
struct Poly {
    V3 p[kMaxPoints];
    int n;
};

// All 12 edge tests of planeBoxIntersections in the reference's order (x
// edges over y then z, y edges over x then z, z edges over y then x), kept in
// registers: q[e] and keep[e] (hit, and no earlier kept point equal to it).
// The mean sums the kept points in that order, as the sequential appends do.
__device__ __forceinline__ void plane_box(V3 bp, V3 bs, V3 po, V3 pn, V3 (&q)[kEdges],
                                          bool (&keep)[kEdges], int& n, V3& mean)
{
    const V3 h = divs(bs, 2.0f);
#pragma unroll
    for (int e = 0; e < kEdges; ++e) {
        const int axis = e >> 2;
        const float a = (e & 2) ? 1.0f : -1.0f, b = (e & 1) ? 1.0f : -1.0f;
        V3 off, rd;
        float max_t;
        if (axis == 0) {
            off = V3{h.x, h.y * a, h.z * b};
            rd = V3{1, 0, 0};
            max_t = bs.x;
        } else if (axis == 1) {
            off = V3{h.x * a, h.y, h.z * b};
            rd = V3{0, 1, 0};
            max_t = bs.y;
        } else {
            off = V3{h.x * b, h.y * a, h.z};
            rd = V3{0, 0, 1};
            max_t = bs.z;
        }
        const V3 ro = sub(bp, off);
        const float t = dot(sub(po, ro), pn) / dot(rd, pn);
        keep[e] = !(t > max_t || t < 0);
        q[e] = add(ro, muls(rd, t));
    }
    V3 avg{0.0f, 0.0f, 0.0f};
    n = 0;
#pragma unroll
    for (int j = 0; j < kEdges; ++j) {
        bool dup = false;
#pragma unroll
        for (int i = 0; i < j; ++i)
            dup = dup || (keep[i] && q[j].x == q[i].x && q[j].y == q[i].y && q[j].z == q[i].z);
        keep[j] = keep[j] && !dup;
        if (keep[j]) {
            avg = add(avg, q[j]);
            ++n;
        }
    }
    mean = divs(avg, (float)n);
}

// Pseudoangle sort keys about the mean in the frame (ax, normal x ax), ax
// toward the first kept point.  Returns false if a key is NaN.
__device__ __forceinline__ bool sort_keys(const V3 (&q)[kEdges], const bool (&keep)[kEdges], V3 o,
                                          V3 normal, float (&key)[kEdges])
{
    normal = normalize(normal);
    V3 first = o;
    bool found = false;
#pragma unroll
    for (int j = kEdges - 1; j >= 0; --j)
        if (keep[j]) {
            first = q[j];
            found = true;
        }
    const V3 ax = normalize(sub(first, o));
    const V3 ay = cross(normal, ax);
    bool ok = true;
#pragma unroll
    for (int j = 0; j < kEdges; ++j) {
        const V3 v = sub(q[j], o);
        key[j] = pseudoangle(dot(ay, v), dot(ax, v));
        ok = ok && (!keep[j] || key[j] == key[j]);
    }
    return ok || !found;
}

// The reference sorts with std::sort, which for ranges this short is
// libstdc++'s insertion sort: a stable sort, here by descending key.  With
// keys that are all numbers (the regular case) each kept point's place is its
// rank: the kept points with a larger key plus the earlier kept points with
// an equal one.  NaN keys (a plane through an edge lying in it, or a plane
// that meets the box in a single point) take the insertion sort itself.
__device__ void insertion_sort(Poly& P, float (&key)[kMaxPoints])
{
    const int m = P.n - 1;
    for (int i = 1; i < m; ++i) {
        const V3 pv = P.p[i];
        const float kv = key[i];
        int j = i;
        if (kv > key[0]) {
            for (; j > 0; --j) {
                P.p[j] = P.p[j - 1];
                key[j] = key[j - 1];
            }
        } else {
            while (kv > key[j - 1]) {
                P.p[j] = P.p[j - 1];
                key[j] = key[j - 1];
                --j;
            }
        }
        P.p[j] = pv;
        key[j] = kv;
    }
}

constexpr int kProxyThreads = 1024;

// `done` (may be NULL): after every output is written and made visible to
// the host (system-scope fences), lane 0 stores `seq` there with a
// system-scope release, so a host waiting on pinned memory may read the
// outputs as soon as it sees the value.
__global__ __launch_bounds__(kProxyThreads) void k_proxy_geometry(
    ProxyView view, int steps, float* __restrict__ verts, int16_t* __restrict__ indices,
    smoke_proxy_plane* __restrict__ planes, int* __restrict__ counts, unsigned* done, unsigned seq)
{
    __shared__ int scan_v[kProxyThreads];
    __shared__ int scan_i[kProxyThreads];
    const int tid = threadIdx.x, nt = blockDim.x;
    const V3 bp{view.box_pos[0], view.box_pos[1], view.box_pos[2]};
    const V3 dims{view.box_dims[0], view.box_dims[1], view.box_dims[2]};
    const V3 origin{view.origin[0], view.origin[1], view.origin[2]};
    const V3 dir = normalize(V3{-view.dir[0], -view.dir[1], -view.dir[2]});
    // getMinMaxPoints: corners of a box centred at the origin (the
    // reference's choice, not the box position), direction normalised again.
    const V3 d2 = normalize(dir);
    float lo = 0.0f, hi = 0.0f;
    for (int c = 0; c < 8; ++c) {
        const V3 corner{(c & 1) ? 1.0f : -1.0f, (c & 2) ? 1.0f : -1.0f, (c & 4) ? -1.0f : 1.0f};
        const float dist = dot(sub(divs(mul(corner, dims), 2.0f), origin), d2);
        if (c == 0) lo = hi = dist;
        if (dist < lo) lo = dist;
        if (dist > hi) hi = dist;
    }
    const V3 start = add(origin, muls(dir, lo));
    const V3 end = add(origin, muls(dir, hi));

    int vbase = 0, ibase = 0;
    for (int c0 = 0; c0 < steps; c0 += nt) {
        const int k = c0 + tid;
        V3 q[kEdges];
        bool keep[kEdges];
        float key[kEdges];
        V3 mean{0.0f, 0.0f, 0.0f};
        int npts = 0;
        bool ranked = true;
        if (k < steps) {
            const float t = (0.5f + (float)k) / (float)steps;
            const V3 pos = add(start, muls(sub(end, start), t));
            plane_box(bp, dims, pos, dir, q, keep, npts, mean);
            ranked = sort_keys(q, keep, mean, dir, key);
        } else {
#pragma unroll
            for (int e = 0; e < kEdges; ++e) keep[e] = false;
        }
        const int nv = k < steps ? npts + 1 : 0;
        const int ni = k < steps && npts >= 1 ? 3 * npts : 0;
        // Inclusive prefix sums of the counts over the workgroup.
        scan_v[tid] = nv;
        scan_i[tid] = ni;
        __syncthreads();
        for (int off = 1; off < nt; off <<= 1) {
            const int av = tid >= off ? scan_v[tid - off] : 0;
            const int ai = tid >= off ? scan_i[tid - off] : 0;
            __syncthreads();
            scan_v[tid] += av;
            scan_i[tid] += ai;
            __syncthreads();
        }
        if (k < steps) {
            const int vo = vbase + scan_v[tid] - nv;
            const int io = ibase + scan_i[tid] - ni;
            if (ranked) {
#pragma unroll
                for (int j = 0; j < kEdges; ++j) {
                    if (!keep[j]) continue;
                    int slot = 0;
#pragma unroll
                    for (int i = 0; i < kEdges; ++i)
                        slot += keep[i] && (key[i] > key[j] || (i < j && key[i] == key[j]));
                    verts[3 * (vo + slot) + 0] = q[j].x;
                    verts[3 * (vo + slot) + 1] = q[j].y;
                    verts[3 * (vo + slot) + 2] = q[j].z;
                }
            } else {
                Poly P;
                float pk[kMaxPoints];
                P.n = 0;
                for (int j = 0; j < kEdges; ++j)
                    if (keep[j]) {
                        pk[P.n] = key[j];
                        P.p[P.n++] = q[j];
                    }
                P.p[P.n++] = mean;
                insertion_sort(P, pk);
                for (int i = 0; i + 1 < P.n; ++i) {
                    verts[3 * (vo + i) + 0] = P.p[i].x;
                    verts[3 * (vo + i) + 1] = P.p[i].y;
                    verts[3 * (vo + i) + 2] = P.p[i].z;
                }
            }
            verts[3 * (vo + npts) + 0] = mean.x;
            verts[3 * (vo + npts) + 1] = mean.y;
            verts[3 * (vo + npts) + 2] = mean.z;
            // teselate: fan about the mean, closed by (mean, last, first);
            // the reference stores 16-bit indices (values wrap past 32767).
            const int centre = vo + npts;
            int w = io;
            if (npts >= 1) {
                for (int i = 0; i < npts - 1; ++i) {
                    indices[w++] = (int16_t)(uint16_t)centre;
                    indices[w++] = (int16_t)(uint16_t)(vo + i);
                    indices[w++] = (int16_t)(uint16_t)(vo + i + 1);
                }
                indices[w++] = (int16_t)(uint16_t)centre;
                indices[w++] = (int16_t)(uint16_t)(vo + npts - 1);
                indices[w++] = (int16_t)(uint16_t)vo;
            }
            planes[k].num_triangles = npts;
            planes[k].triangle_offset = io / 3;
        }
        vbase += scan_v[nt - 1];
        ibase += scan_i[nt - 1];
        __syncthreads();  // every lane has read the totals before they are overwritten
    }
    if (tid == 0) {
        counts[0] = vbase;
        counts[1] = ibase;
    }
    if (done) {
        __threadfence_system();
        __syncthreads();
        if (tid == 0) __hip_atomic_store(done, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

}  // namespace

hipError_t proxy_geometry(const ProxyView& view, int steps, float* verts, int16_t* indices,
                          smoke_proxy_plane* planes, int* counts, hipStream_t st, unsigned* done,
                          unsigned seq)
{
    if (steps < 1) return hipErrorInvalidValue;
    const int nt = std::min(kProxyThreads, (steps + 63) / 64 * 64);
    hipLaunchKernelGGL(k_proxy_geometry, dim3(1), dim3(nt), 0, st, view, steps, verts, indices,
                       planes, counts, done, seq);
    return hipGetLastError();
}

}  // namespace smk
