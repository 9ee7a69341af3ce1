// HIP kernel of the renderer's proxy-plane slicing geometry for gfx950:
// BoundingBox::generateProxyGeometry (project/boundingBox.cpp:133-183) with
// getMinMaxPoints (:231-241), planeBoxIntersections (:285-354), duplicate
// (:356-361), sortPoints (:259-272), pseudoangle (:87-92) and teselate
// (:243-255), computed on the device instead of sequentially on the CPU.
//
// One workgroup of kProxyThreads lanes; lane j of chunk c builds plane
// k = c * kProxyThreads + j: its edge points (at most 12, plus their mean) in
// registers, sorted with the same insertion sort, then a workgroup prefix sum
// over the per-plane vertex and index counts places every plane's vertices,
// fan triangles and (num_triangles, triangle_offset) entry exactly where the
// reference's sequential appends put them.  Vector arithmetic is FP32 in
// glm 0.9.9's operation order (dot = (x*x' + y*y') + z*z', normalize = v * (1
// / sqrt(dot(v, v))), cross, component-wise vec3 / scalar); the file is
// compiled with -ffp-contract=off and HIP's correctly rounded division and
// square root, so the vertices are bit-identical to the CPU oracle's.
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
constexpr int kMaxPoints = 13;
struct Poly {
    V3 p[kMaxPoints];
    int n;
};

__device__ __forceinline__ void visit(Poly& P, V3& avg, V3 ro, V3 rd, float max_t, V3 po, V3 pn)
{
    const float t = dot(sub(po, ro), pn) / dot(rd, pn);
    if (t > max_t || t < 0) return;
    const V3 q = add(ro, muls(rd, t));
    for (int i = 0; i < P.n; ++i)
        if (q.x == P.p[i].x && q.y == P.p[i].y && q.z == P.p[i].z) return;
    P.p[P.n++] = q;
    avg = add(avg, q);
}

__device__ void plane_box(Poly& P, V3 bp, V3 bs, V3 po, V3 pn)
{
    P.n = 0;
    const V3 h = divs(bs, 2.0f);
    V3 avg{0.0f, 0.0f, 0.0f};
    for (int y = -1; y < 2; y += 2)
        for (int z = -1; z < 2; z += 2)
            visit(P, avg, sub(bp, V3{h.x, h.y * (float)y, h.z * (float)z}), V3{1, 0, 0}, bs.x, po, pn);
    for (int x = -1; x < 2; x += 2)
        for (int z = -1; z < 2; z += 2)
            visit(P, avg, sub(bp, V3{h.x * (float)x, h.y, h.z * (float)z}), V3{0, 1, 0}, bs.y, po, pn);
    for (int y = -1; y < 2; y += 2)
        for (int x = -1; x < 2; x += 2)
            visit(P, avg, sub(bp, V3{h.x * (float)x, h.y * (float)y, h.z}), V3{0, 0, 1}, bs.z, po, pn);
    P.p[P.n] = divs(avg, (float)P.n);
    P.n += 1;
}

// Descending pseudoangle about the mean in the frame (ax, normal x ax), ax
// toward the first point found; the insertion sort keeps equal keys in
// discovery order (it is the algorithm libstdc++'s std::sort applies to
// ranges this short).
__device__ void sort_points(Poly& P, V3 normal)
{
    const int m = P.n - 1;
    const V3 o = P.p[m];
    normal = normalize(normal);
    const V3 ax = normalize(sub(P.p[0], o));
    const V3 ay = cross(normal, ax);
    float key[kMaxPoints];
    for (int i = 0; i < m; ++i) {
        const V3 v = sub(P.p[i], o);
        key[i] = pseudoangle(dot(ay, v), dot(ax, v));
    }
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

constexpr int kProxyThreads = 256;

__global__ __launch_bounds__(kProxyThreads) void k_proxy_geometry(
    ProxyView view, int steps, float* __restrict__ verts, int16_t* __restrict__ indices,
    smoke_proxy_plane* __restrict__ planes, int* __restrict__ counts)
{
    __shared__ int scan_v[kProxyThreads];
    __shared__ int scan_i[kProxyThreads];
    const int tid = threadIdx.x;
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
    for (int c0 = 0; c0 < steps; c0 += kProxyThreads) {
        const int k = c0 + tid;
        Poly P;
        P.n = 0;
        int nv = 0, ni = 0;
        if (k < steps) {
            const float t = (0.5f + (float)k) / (float)steps;
            const V3 pos = add(start, muls(sub(end, start), t));
            plane_box(P, bp, dims, pos, dir);
            sort_points(P, dir);
            nv = P.n;
            ni = P.n >= 2 ? 3 * (P.n - 1) : 0;
        }
        // Inclusive prefix sums of the counts over the workgroup.
        scan_v[tid] = nv;
        scan_i[tid] = ni;
        __syncthreads();
        for (int off = 1; off < kProxyThreads; off <<= 1) {
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
            for (int i = 0; i < P.n; ++i) {
                verts[3 * (vo + i) + 0] = P.p[i].x;
                verts[3 * (vo + i) + 1] = P.p[i].y;
                verts[3 * (vo + i) + 2] = P.p[i].z;
            }
            // teselate: fan about the mean, closed by (mean, last, first);
            // the reference stores 16-bit indices (values wrap past 32767).
            const int centre = vo + P.n - 1;
            int w = io;
            if (P.n >= 2) {
                for (int i = 0; i < P.n - 2; ++i) {
                    indices[w++] = (int16_t)(uint16_t)centre;
                    indices[w++] = (int16_t)(uint16_t)(vo + i);
                    indices[w++] = (int16_t)(uint16_t)(vo + i + 1);
                }
                indices[w++] = (int16_t)(uint16_t)centre;
                indices[w++] = (int16_t)(uint16_t)(vo + P.n - 2);
                indices[w++] = (int16_t)(uint16_t)vo;
            }
            planes[k].num_triangles = P.n - 1;
            planes[k].triangle_offset = io / 3;
        }
        vbase += scan_v[kProxyThreads - 1];
        ibase += scan_i[kProxyThreads - 1];
        __syncthreads();  // every lane has read the totals before they are overwritten
    }
    if (tid == 0) {
        counts[0] = vbase;
        counts[1] = ibase;
    }
}

}  // namespace

hipError_t proxy_geometry(const ProxyView& view, int steps, float* verts, int16_t* indices,
                          smoke_proxy_plane* planes, int* counts, hipStream_t st)
{
    hipLaunchKernelGGL(k_proxy_geometry, dim3(1), dim3(kProxyThreads), 0, st, view, steps, verts,
                       indices, planes, counts);
    return hipGetLastError();
}

}  // namespace smk
