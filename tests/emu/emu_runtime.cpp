// Test harness: the subset of the HIP runtime API that libsmokesim's host code
// calls, implemented on the CPU, plus the kernel emulator behind
// hipLaunchKernelGGL (emu_shim.h).
//
//  * "Device" allocations are plain heap blocks (AddressSanitizer bounds).
//  * Work runs synchronously in submission order; a stream under capture
//    records launches instead, and a graph launch replays them.
//  * With EMU_DEFERRED set, stream work (launches, asynchronous copies,
//    graph launches, event waits) is queued per stream and runs only when a
//    synchronisation needs it -- stream or event synchronisation, an event
//    wait reached by another stream, stream destruction, and hipFree /
//    hipHostFree (which wait for the whole device, as HIP's do); the
//    synchronous hipMemcpy runs at once without waiting for any stream (the
//    library's streams are non-blocking).  Pending work of different streams
//    runs newest stream first.  A host buffer changed or freed before the
//    copy that reads it ran, a result read before its stream was
//    synchronised, or a missing event wait between streams then gives wrong
//    results or an AddressSanitizer report here.  With EMU_REVERSE_BLOCKS set,
//    kernels that run their blocks one after another do so in reverse order,
//    so results that depend on block order show up.
//  * Kernels without barriers run their threads one after another; the
//    fused-Jacobi kernels (the only ones with __syncthreads) run every thread
//    of a block on its own host thread with a real barrier, one block at a
//    time, which is what makes `static` stand in for LDS.
//  * Handles carry magic numbers and are checked on every use, so a call on
//    a destroyed stream, event or graph stops the run.
#include <hip/hip_runtime_api.h>

#include <barrier>
#include <condition_variable>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <deque>
#include <functional>
#include <memory>
#include <set>
#include <thread>
#include <vector>

namespace {
constexpr unsigned kStream = 0x5354524dU, kEvent = 0x4556454eU, kGraph = 0x47524150U,
                   kExec = 0x45584543U;
using Op = std::function<void()>;
}  // namespace

struct ihipStream_t {
    unsigned magic;
    bool capturing;
    std::vector<Op> rec;
    std::deque<Op> queue;  // deferred work not yet run
    unsigned long long enqueued = 0, done = 0;
};
struct ihipEvent_t {
    unsigned magic;
    ihipStream_t* stream = nullptr;  // where it was last recorded
    unsigned long long target = 0;   // that stream's work count at the record
};
struct ihipGraph {
    unsigned magic;
    std::vector<Op> ops;
};
struct hipGraphExec {
    unsigned magic;
    std::vector<Op> ops;
};

struct EmuIdx {
    unsigned x, y, z;
};
thread_local EmuIdx threadIdx, blockIdx, blockDim, gridDim;

namespace {

std::set<void*> g_device;
const bool g_deferred = getenv("EMU_DEFERRED") != nullptr;
const bool g_reverse_blocks = getenv("EMU_REVERSE_BLOCKS") != nullptr;
ihipStream_t g_null_stream{0x5354524dU, false, {}, {}, 0, 0};
std::vector<ihipStream_t*> g_streams;  // live streams, in creation order (null stream first)
std::barrier<>* g_barrier = nullptr;
// One barrier per wave of 64 threads for lane exchanges (DPP, shuffles): a
// wave may skip work that other waves of the block do (wave-uniform
// branches), so exchanges synchronise the wave, not the block.
std::vector<std::barrier<>*> g_wave_barrier;
bool g_threaded = false;

[[noreturn]] void die(const char* what)
{
    fprintf(stderr, "emu: %s\n", what);
    abort();
}
void check(hipStream_t s)
{
    if (s && s->magic != kStream) die("use of an invalid stream");
}
void check(hipEvent_t e)
{
    if (!e || e->magic != kEvent) die("use of an invalid event");
}
ihipStream_t* S(hipStream_t s) { return s ? s : &g_null_stream; }

// Runs queued work of `s` until `upto` items of it have run.  Work may drain
// other streams (event waits).
void drain(ihipStream_t* s, unsigned long long upto)
{
    while (s->done < upto) {
        if (s->queue.empty()) die("drain past the queued work");
        Op op = std::move(s->queue.front());
        s->queue.pop_front();
        op();
        s->done += 1;
    }
}
void drain_all()
{
    if (g_null_stream.done < g_null_stream.enqueued) drain(&g_null_stream, g_null_stream.enqueued);
    for (;;) {
        bool any = false;
        for (size_t i = g_streams.size(); i-- > 0;) {
            ihipStream_t* s = g_streams[i];
            if (s->done < s->enqueued) {
                drain(s, s->enqueued);
                any = true;
                break;  // the list may have changed
            }
        }
        if (!any) break;
    }
    if (g_null_stream.done < g_null_stream.enqueued) drain(&g_null_stream, g_null_stream.enqueued);
}
void enqueue(hipStream_t st, Op op)
{
    ihipStream_t* s = S(st);
    if (!g_deferred) {
        op();
        s->enqueued += 1;
        s->done += 1;
        return;
    }
    s->queue.push_back(std::move(op));
    s->enqueued += 1;
}

// Lane exchanges use two buffers in turn, so one wave barrier per exchange
// suffices: a thread can only overwrite a buffer after every thread of its
// wave (the only readers of its slot) has passed the next exchange's
// barrier, i.e. finished reading it.
unsigned long long g_xbuf[2][1024];
thread_local unsigned g_xcount = 0;

// Persistent host threads for the threaded kernels (created once, reused by
// every launch; never joined: the process exits with them parked).
class Pool {
public:
    void run(unsigned n, std::function<void(unsigned)> job)
    {
        std::unique_lock<std::mutex> lk(m_);
        while (threads_.size() < n) {
            const unsigned id = (unsigned)threads_.size();
            threads_.emplace_back([this, id]() { worker(id); });
            threads_.back().detach();
        }
        job_ = std::move(job);
        njob_ = n;
        active_ = n;
        ++gen_;
        cv_.notify_all();
        done_.wait(lk, [&]() { return active_ == 0; });
    }

private:
    void worker(unsigned id)
    {
        unsigned long long seen = 0;
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            cv_.wait(lk, [&]() { return gen_ != seen; });
            seen = gen_;
            if (id >= njob_) continue;
            std::function<void(unsigned)> job = job_;
            lk.unlock();
            job(id);
            lk.lock();
            if (--active_ == 0) done_.notify_all();
        }
    }
    std::mutex m_;
    std::condition_variable cv_, done_;
    std::vector<std::thread> threads_;
    std::function<void(unsigned)> job_;
    unsigned njob_ = 0, active_ = 0;
    unsigned long long gen_ = 0;
};
Pool& pool()
{
    static Pool* p = new Pool();  // intentionally never destroyed
    return *p;
}

void run(const char* name, const std::function<void()>& fn, dim3 grid, dim3 block)
{
    // Kernels with barriers or lane exchanges run their threads concurrently.
    const bool threaded =
        strstr(name, "jacobi_fused") || strstr(name, "kb_jacobi<") || strstr(name, "k_ref_") ||
        strstr(name, "k_proxy") || strstr(name, "k_advect_velocity") || strstr(name, "kb_adv_v") ||
        strstr(name, "k_advect_density") || strstr(name, "kb_adv_d") || strstr(name, "k_norms");
    const unsigned nt = block.x * block.y * block.z;
    const EmuIdx gd{grid.x, grid.y, grid.z}, bd{block.x, block.y, block.z};
    auto set_thread = [&](unsigned t) {
        threadIdx = {t % block.x, (t / block.x) % block.y, t / (block.x * block.y)};
    };
    if (!threaded) {
        blockDim = bd;
        gridDim = gd;
        const unsigned long long nb = (unsigned long long)grid.x * grid.y * grid.z;
        for (unsigned long long k = 0; k < nb; ++k) {
            const unsigned long long b = g_reverse_blocks ? nb - 1 - k : k;
            blockIdx = {(unsigned)(b % grid.x), (unsigned)((b / grid.x) % grid.y),
                        (unsigned)(b / ((unsigned long long)grid.x * grid.y))};
            for (unsigned t = 0; t < nt; ++t) {
                set_thread(t);
                fn();
            }
        }
        return;
    }
    std::barrier<> bar((std::ptrdiff_t)nt);
    g_barrier = &bar;
    std::vector<std::unique_ptr<std::barrier<>>> waves;
    g_wave_barrier.clear();
    for (unsigned w = 0; w * 64 < nt; ++w) {
        waves.emplace_back(new std::barrier<>((std::ptrdiff_t)std::min(64u, nt - 64 * w)));
        g_wave_barrier.push_back(waves.back().get());
    }
    g_threaded = true;
    pool().run(nt, [&](unsigned t) {
        blockDim = bd;
        gridDim = gd;
        g_xcount = 0;
        set_thread(t);
        for (unsigned bz = 0; bz < grid.z; ++bz)
            for (unsigned by = 0; by < grid.y; ++by)
                for (unsigned bx = 0; bx < grid.x; ++bx) {
                    blockIdx = {bx, by, bz};
                    fn();
                    bar.arrive_and_wait();  // the block ends before the next starts
                }
    });
    g_threaded = false;
    g_barrier = nullptr;
}

}  // namespace

void emu_syncthreads()
{
    if (!g_threaded) die("__syncthreads in a kernel run without threads");
    g_barrier->arrive_and_wait();
}

static unsigned long long exchange(unsigned long long bits, int partner_t)
{
    const int t = (int)(threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z));
    unsigned long long* buf = g_xbuf[g_xcount++ & 1];
    buf[t] = bits;
    g_wave_barrier[t / 64]->arrive_and_wait();
    return buf[partner_t];
}

int emu_dpp(int old, int src, int ctrl, bool bound_ctrl)
{
    if (!g_threaded) die("DPP in a kernel run without threads");
    const int t = (int)(threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z));
    const int lane = t & 63;
    int from;
    if (ctrl == 0x138) from = lane - 1;       // wave_shr:1
    else if (ctrl == 0x130) from = lane + 1;  // wave_shl:1
    else die("unsupported DPP control");
    const bool ok = from >= 0 && from < 64;
    const int r = (int)(unsigned)exchange((unsigned)src, ok ? t - lane + from : t);
    return ok ? r : (bound_ctrl ? 0 : old);
}

unsigned long long emu_exchange(unsigned long long bits, int partner)
{
    if (!g_threaded) die("lane exchange in a kernel run without threads");
    return exchange(bits, partner);
}

hipError_t emu_last_error() { return hipSuccess; }

void emu_launch(const char* name, std::function<void()> fn, dim3 grid, dim3 block, hipStream_t st)
{
    check(st);
    if (st && st->capturing) {
        st->rec.push_back([=]() { run(name, fn, grid, block); });
        return;
    }
    enqueue(st, [=]() { run(name, fn, grid, block); });
}

extern "C" {

hipError_t hipMalloc(void** p, size_t n)
{
    *p = malloc(n ? n : 1);
    g_device.insert(*p);
    return hipSuccess;
}
hipError_t hipFree(void* p)
{
    if (!p) return hipSuccess;
    drain_all();
    if (!g_device.erase(p)) die("hipFree of a pointer hipMalloc did not return");
    free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned)
{
    *p = malloc(n ? n : 1);
    return hipSuccess;
}
hipError_t hipHostFree(void* p)
{
    drain_all();
    free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind)
{
    memcpy(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t st)
{
    check(st);
    if (st && st->capturing) die("copy recorded into a graph");
    enqueue(st, [=]() { memcpy(d, s, n); });
    return hipSuccess;
}
hipError_t hipGetDevice(int* d)
{
    *d = 0;
    return hipSuccess;
}
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidDevice; }
hipError_t hipGetSymbolAddress(void** p, const void* symbol)
{
    *p = const_cast<void*>(symbol);
    return hipSuccess;
}
hipError_t hipGetDeviceCount(int* c)
{
    *c = 1;
    return hipSuccess;
}
hipError_t hipGetDevicePropertiesR0600(hipDeviceProp_tR0600* p, int)
{
    memset(p, 0, sizeof *p);
    strcpy(p->name, "cpu emulator");
    strcpy(p->gcnArchName, "emu");
    p->multiProcessorCount = 256;
    p->totalGlobalMem = 288ull << 30;
    return hipSuccess;
}
const char* hipGetErrorString(hipError_t) { return "emulator error"; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned)
{
    *s = new ihipStream_t{kStream, false, {}, {}, 0, 0};
    g_streams.push_back(*s);
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s)
{
    check(s);
    if (!s) return hipErrorInvalidValue;
    drain(s, s->enqueued);
    g_streams.erase(std::remove(g_streams.begin(), g_streams.end(), s), g_streams.end());
    s->magic = 0;
    // Deferred event waits of other streams may still name it (drained, so
    // they return at once): kept allocated in that mode.
    if (!g_deferred) delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s)
{
    check(s);
    if (s && s->capturing) die("synchronize on a capturing stream");
    drain(S(s), S(s)->enqueued);
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned)
{
    check(s);
    check(e);
    if (s && s->capturing) die("event wait recorded into a graph");
    ihipStream_t* es = e->stream;
    const unsigned long long t = e->target;
    if (es) enqueue(s, [es, t]() { drain(es, t); });
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode)
{
    check(s);
    s->capturing = true;
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g)
{
    check(s);
    s->capturing = false;
    *g = new ihipGraph{kGraph, std::move(s->rec)};
    s->rec.clear();
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t)
{
    if (!g || g->magic != kGraph) die("instantiate of an invalid graph");
    *e = new hipGraphExec{kExec, g->ops};
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s)
{
    check(s);
    if (!e || e->magic != kExec) die("launch of an invalid graph");
    for (Op& op : e->ops) enqueue(s, op);
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e)
{
    if (!e || e->magic != kExec) die("destroy of an invalid graph exec");
    e->magic = 0;
    delete e;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g)
{
    if (!g || g->magic != kGraph) die("destroy of an invalid graph");
    g->magic = 0;
    delete g;
    return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t* e)
{
    *e = new ihipEvent_t{kEvent};
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned)
{
    *e = new ihipEvent_t{kEvent};
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e)
{
    check(e);
    e->magic = 0;
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s)
{
    check(e);
    check(s);
    if (s && s->capturing) return hipSuccess;  // not part of the recorded work
    e->stream = S(s);
    e->target = S(s)->enqueued;
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e)
{
    check(e);
    if (e->stream) drain(e->stream, e->target);
    return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t e)
{
    check(e);
    return !e->stream || e->stream->done >= e->target ? hipSuccess : hipErrorNotReady;
}
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b)
{
    check(a);
    check(b);
    if (hipEventQuery(a) != hipSuccess || hipEventQuery(b) != hipSuccess) return hipErrorNotReady;
    *ms = 1.0f;
    return hipSuccess;
}

}  // extern "C"
