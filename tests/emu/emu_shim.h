// Test harness: force-included (-include) when the library's HIP kernel
// sources are compiled as plain host C++ for the CPU emulator.  Kernels then
// run on host threads (emu_runtime.cpp) and "device" memory is malloc'ed, so
// AddressSanitizer checks every kernel access against its allocation and the
// results can be compared with the oracle on small grids.  Never part of the
// shipped library: smokesim only loads the gfx950 build.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>

#undef __shared__
#define __shared__ static  // one block runs at a time, so a static is per-block LDS
#undef __launch_bounds__
#define __launch_bounds__(...)
#ifndef __align__
#define __align__(n) __attribute__((aligned(n)))
#endif
// A __device__ variable is a host global here; its "device address" is its
// host address (emu_runtime.cpp's hipGetSymbolAddress).
#undef HIP_SYMBOL
#define HIP_SYMBOL(x) x

struct EmuIdx {
    unsigned x, y, z;
};
extern thread_local EmuIdx threadIdx, blockIdx, blockDim, gridDim;

void emu_syncthreads();
#define __syncthreads() emu_syncthreads()

// System-scope fence and release store: host threads share one coherent
// memory, so a sequentially consistent fence and an atomic store suffice.
inline void __threadfence_system() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }
#ifndef __HIP_MEMORY_SCOPE_SYSTEM
#define __HIP_MEMORY_SCOPE_SYSTEM 5
#endif
#define __hip_atomic_store(p, v, order, scope) __atomic_store_n((p), (v), (order))

// Host arithmetic rounds every FP32 operation before a conversion, so the
// device's opaque-value barrier is not needed (and its "v" register
// constraint does not exist on the host).
#define SMK_OPAQUE_F32(x) ((void)(x))

// Wave-uniform values only (the kernels apply it to values that already are):
// the value itself.
#define __builtin_amdgcn_readfirstlane(x) (x)

inline int __popc(unsigned v) { return __builtin_popcount(v); }
inline unsigned __float_as_uint(float f)
{
    unsigned u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float __uint_as_float(unsigned u)
{
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Lane exchange of a block run on host threads: every thread of the block
// publishes its value and reads its partner's (both sides of a barrier).
unsigned long long emu_exchange(unsigned long long bits, int partner);

// DPP row moves used by the kernels: wave_shr:1 (0x138) and wave_shl:1 (0x130)
// within each wave64; a lane without a source gets 0 (bound_ctrl) or `old`.
int emu_dpp(int old, int src, int ctrl, bool bound_ctrl);
#define __builtin_amdgcn_update_dpp(old, src, ctrl, row_mask, bank_mask, bound_ctrl) \
    emu_dpp((old), (src), (ctrl), (bound_ctrl))

// Raw buffer access: a base and a size in bytes; loads at or past the size
// return 0 and such stores are dropped (the hardware's range check for
// untyped buffers with stride 0).
struct EmuRsrc {
    char* base;
    uint32_t bytes;
};
typedef EmuRsrc __amdgpu_buffer_rsrc_t;
inline EmuRsrc emu_make_rsrc(void* p, int bytes) { return EmuRsrc{static_cast<char*>(p), (uint32_t)bytes}; }
inline bool emu_in_range(const EmuRsrc& r, int off, int soff, uint32_t size)
{
    const uint64_t o = (uint64_t)(uint32_t)off + (uint64_t)(uint32_t)soff;
    return o + size <= r.bytes;
}
inline uint32_t emu_buf_ld32(EmuRsrc r, int off, int soff)
{
    uint32_t v = 0;
    if (emu_in_range(r, off, soff, 4)) std::memcpy(&v, r.base + (uint32_t)off + (uint32_t)soff, 4);
    return v;
}
inline uint8_t emu_buf_ld8(EmuRsrc r, int off, int soff)
{
    return emu_in_range(r, off, soff, 1) ? (uint8_t)r.base[(uint32_t)off + (uint32_t)soff] : (uint8_t)0;
}
inline uint16_t emu_buf_ld16(EmuRsrc r, int off, int soff)
{
    uint16_t v = 0;
    if (emu_in_range(r, off, soff, 2)) std::memcpy(&v, r.base + (uint32_t)off + (uint32_t)soff, 2);
    return v;
}
// Two dwords, each range-checked on its own.
typedef uint32_t emu_u32x2 __attribute__((vector_size(8)));
inline emu_u32x2 emu_buf_ld64(EmuRsrc r, int off, int soff)
{
    emu_u32x2 v;
    v[0] = emu_buf_ld32(r, off, soff);
    v[1] = emu_buf_ld32(r, (int)((uint32_t)off + 4u), soff);
    return v;
}
inline void emu_buf_st32(uint32_t v, EmuRsrc r, int off, int soff)
{
    if (emu_in_range(r, off, soff, 4)) std::memcpy(r.base + (uint32_t)off + (uint32_t)soff, &v, 4);
}
#define __builtin_amdgcn_make_buffer_rsrc(p, stride, bytes, flags) emu_make_rsrc((p), (bytes))
#define __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, aux) emu_buf_ld32((r), (off), (soff))
#define __builtin_amdgcn_raw_buffer_load_b8(r, off, soff, aux) emu_buf_ld8((r), (off), (soff))
#define __builtin_amdgcn_raw_buffer_load_b16(r, off, soff, aux) emu_buf_ld16((r), (off), (soff))
#define __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, aux) emu_buf_ld64((r), (off), (soff))
#define __builtin_amdgcn_raw_buffer_store_b32(v, r, off, soff, aux) emu_buf_st32((v), (r), (off), (soff))
template <typename T>
T __shfl_xor(T v, int mask, int width = 64)
{
    static_assert(sizeof(T) <= 8, "shuffle of at most 8 bytes");
    const int t = (int)(threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z));
    const int lane = t % width;
    const int partner = (lane ^ mask) < width ? t - lane + (lane ^ mask) : t;
    unsigned long long b = 0;
    std::memcpy(&b, &v, sizeof(T));
    b = emu_exchange(b, partner);
    T r;
    std::memcpy(&r, &b, sizeof(T));
    return r;
}
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }

hipError_t emu_last_error();
#define hipGetLastError() emu_last_error()

// Runs (or, on a capturing stream, records) one launch.
void emu_launch(const char* name, std::function<void()> fn, dim3 grid, dim3 block, hipStream_t st);
#undef hipLaunchKernelGGL
#define hipLaunchKernelGGL(K, G, B, SH, ST, ...) \
    emu_launch(#K, [=]() { K(__VA_ARGS__); }, dim3(G), dim3(B), ST)

// Resident workgroups per CU; EMU_OCCUPANCY overrides it, so the persistent
// grids of the batched kernels take other sizes.
template <class F>
hipError_t emu_occupancy(int* n, F, int threads, size_t)
{
    const char* env = getenv("EMU_OCCUPANCY");
    *n = env ? atoi(env) : (threads >= 1024 ? 2 : 2048 / threads);
    return hipSuccess;
}
#define hipOccupancyMaxActiveBlocksPerMultiprocessor emu_occupancy
