// Host model of the HIP runtime for the served-path replay (test infrastructure; see
// hip/hip_runtime.h).
//
// What is modelled:
//   * "Device" and pinned host memory come from one reserved address range with a bump pointer,
//     so no address is ever handed out twice: a freed block stays in the registry as dead and any
//     later access to it is reported (a real pool could reuse it and hide the bug).
//   * Streams are queues of operations that run later, not at the call: a synchronising call
//     (stream / event synchronize, blocking copies, hipFree) runs what it needs, and every
//     hipEventQuery first runs a random number of ready operations of random streams, so kernels
//     of different streams interleave with each other and with the host in varying orders.
//     hipStreamWaitEvent blocks its stream until the recording it names has run.
//   * Stream-ordered allocations are live from the position of hipMallocAsync in their stream to
//     the position of hipFreeAsync; a kernel that runs outside that window is reported.
//   * Copies from pageable host memory take their bytes at the call (as the runtime stages
//     them); copies from or into pinned memory touch it when they run, and check it is alive.
//   * Events used after hipEventDestroy are reported.
// Kernels are address models (kernel_model.cpp) that run as stream operations and check every
// buffer they would touch against the live allocations at that moment.
#include "model.h"

#include <sys/mman.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <vector>

struct ModelEvent {
    uint64_t recorded = 0;    // generation of the latest hipEventRecord
    uint64_t completed = 0;   // highest generation whose record operation has run
    bool destroyed = false;
};

struct ModelOp {
    enum Kind { kRun, kWait } kind = kRun;
    ModelEvent* ev = nullptr;   // kWait
    uint64_t gen = 0;
    std::function<void()> run;
};

struct ModelGraph {
    std::vector<ModelOp> ops;
};

struct ModelStream {
    std::deque<ModelOp> q;
    bool destroyed = false;
    ModelGraph* capture = nullptr;
};

namespace model {

namespace {

std::recursive_mutex g_mu;
struct Alloc {
    uintptr_t base;
    size_t size;
    int kind;   // 0 device, 1 device stream-ordered, 2 pinned host
    bool live;
    uint64_t epoch;   // bumped by every copy / fill into the block
    uint64_t id;
};
std::map<uintptr_t, Alloc> g_allocs;
uint64_t g_next_id = 1;
char* g_arena = nullptr;
size_t g_arena_size = 0, g_arena_used = 0;
std::vector<ModelStream*> g_streams;
std::mt19937_64 g_rng(12345);
std::vector<std::string> g_violations;
size_t g_violation_count = 0;
Stats g_stats;
hipError_t g_last = hipSuccess;

void reserve_arena() {
    if (g_arena) return;
    for (size_t gb : {4096ull, 1024ull, 256ull, 64ull, 4ull}) {
        void* p = mmap(nullptr, gb << 30, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
        if (p != MAP_FAILED) {
            g_arena = static_cast<char*>(p);
            g_arena_size = gb << 30;
            return;
        }
    }
    fprintf(stderr, "hip model: cannot reserve an address range\n");
    abort();
}

void* take(size_t bytes, int kind) {
    reserve_arena();
    g_stats.allocations++;
    const size_t need = (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
    const size_t off = (g_arena_used + 255) & ~size_t(255);
    if (off + need + 4096 > g_arena_size) return nullptr;
    g_arena_used = off + need + 4096;   // a gap that belongs to no allocation after every block
    const uintptr_t base = reinterpret_cast<uintptr_t>(g_arena + off);
    g_allocs[base] = Alloc{base, bytes, kind, true, 0, g_next_id++};
    return g_arena + off;
}

// The allocation containing address p (live or dead), or null.
Alloc* find(uintptr_t p) {
    auto it = g_allocs.upper_bound(p);
    if (it == g_allocs.begin()) return nullptr;
    --it;
    Alloc& a = it->second;
    if (p >= a.base && p < a.base + std::max<size_t>(a.size, 1)) return &a;
    return nullptr;
}

void release(Alloc& a) {
    a.live = false;
    // return the block's whole pages to the system; the addresses stay reserved
    const uintptr_t lo = (a.base + 4095) & ~uintptr_t(4095), hi = (a.base + a.size) & ~uintptr_t(4095);
    if (hi > lo) madvise(reinterpret_cast<void*>(lo), hi - lo, MADV_DONTNEED);
}

bool in_arena(const void* p) {
    return g_arena && p >= g_arena && p < g_arena + g_arena_size;
}

bool ready(ModelStream* s) {
    if (s->q.empty()) return false;
    const ModelOp& op = s->q.front();
    return op.kind != ModelOp::kWait || op.ev->completed >= op.gen;
}

void step(ModelStream* s) {
    ModelOp op = std::move(s->q.front());
    s->q.pop_front();
    if (op.kind == ModelOp::kRun && op.run) op.run();
    g_stats.ops_run++;
}

// Runs ready operations of random streams until pred() holds; reports a deadlock if nothing is
// ready before that.
template <typename P>
void run_until(P pred, const char* what) {
    while (!pred()) {
        std::vector<ModelStream*> r;
        for (ModelStream* s : g_streams)
            if (ready(s)) r.push_back(s);
        if (r.empty()) {
            violation("deadlock: %s waits for stream work that can never run", what);
            return;
        }
        step(r[g_rng() % r.size()]);
    }
}

void run_random(unsigned max_steps) {
    const unsigned n = (unsigned)(g_rng() % (max_steps + 1));
    for (unsigned i = 0; i < n; ++i) {
        std::vector<ModelStream*> r;
        for (ModelStream* s : g_streams)
            if (ready(s)) r.push_back(s);
        if (r.empty()) return;
        step(r[g_rng() % r.size()]);
    }
}

void run_all(const char* what) {
    run_until([] {
        for (ModelStream* s : g_streams)
            if (!s->q.empty()) return false;
        return true;
    }, what);
}

void push(ModelStream* s, ModelOp op) {
    if (s->capture) {
        s->capture->ops.push_back(std::move(op));
        return;
    }
    s->q.push_back(std::move(op));
}

hipError_t ret(hipError_t e) {
    if (e != hipSuccess) g_last = e;
    return e;
}

}  // namespace

std::recursive_mutex& mutex() { return g_mu; }

void violation(const char* fmt, ...) {
    g_violation_count++;
    if (g_violations.size() >= 200) return;
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_violations.emplace_back(buf);
}

bool check(const void* p, size_t bytes, const char* what) {
    g_stats.checks++;
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    Alloc* al = find(a);
    if (!al) {
        violation("%s: %zu bytes at %p lie in no allocation", what, bytes, p);
        return false;
    }
    if (!al->live) {
        violation("%s: %zu bytes at %p belong to allocation #%llu, already freed", what, bytes, p, (unsigned long long)al->id);
        return false;
    }
    if (a + bytes > al->base + al->size) {
        violation("%s: %zu bytes at offset %zu overrun allocation #%llu of %zu bytes", what, bytes, (size_t)(a - al->base),
                  (unsigned long long)al->id, al->size);
        return false;
    }
    return true;
}

uint64_t content_key(const void* p) {
    Alloc* al = find(reinterpret_cast<uintptr_t>(p));
    if (!al) return 0;
    return (al->id << 24) ^ al->epoch;
}

void enqueue(hipStream_t s, const char* name, std::function<void()> body) {
    ModelOp op;
    op.run = [name, body = std::move(body)] {
        g_stats.kernels++;
        (void)name;
        body();
    };
    push(s, std::move(op));
}

Stats& stats() { return g_stats; }

const std::vector<std::string>& violations() { return g_violations; }
size_t violation_count() { return g_violation_count; }

}  // namespace model

using namespace model;

#define LOCK std::lock_guard<std::recursive_mutex> lock_(model::mutex())

hipError_t hipGetDeviceCount(int* count) {
    *count = 1;
    return hipSuccess;
}
hipError_t hipSetDevice(int device) { return device == 0 ? hipSuccess : ret(hipErrorInvalidValue); }
hipError_t hipGetDeviceProperties(hipDeviceProp_t* prop, int) {
    memset(prop, 0, sizeof *prop);
    strcpy(prop->name, "host model of a 256-CU device");
    prop->multiProcessorCount = 256;
    prop->totalGlobalMem = size_t(288) << 30;
    return hipSuccess;
}
const char* hipGetErrorString(hipError_t e) {
    switch (e) {
    case hipSuccess: return "success";
    case hipErrorInvalidValue: return "invalid value (model)";
    case hipErrorMemoryAllocation: return "out of memory (model)";
    case hipErrorNotReady: return "not ready";
    default: return "error (model)";
    }
}
hipError_t hipGetLastError() {
    LOCK;
    hipError_t e = g_last;
    g_last = hipSuccess;
    return e;
}

hipError_t hipMalloc(void** p, size_t bytes) {
    LOCK;
    run_all("hipMalloc");
    *p = take(bytes, 0);
    return *p ? hipSuccess : ret(hipErrorMemoryAllocation);
}
hipError_t hipFree(void* p) {
    LOCK;
    if (!p) return hipSuccess;
    run_all("hipFree");   // hipFree waits for the device
    Alloc* a = find(reinterpret_cast<uintptr_t>(p));
    if (!a || a->base != reinterpret_cast<uintptr_t>(p) || !a->live || a->kind == 2) {
        violation("hipFree(%p): not the start of a live device allocation", p);
        return ret(hipErrorInvalidValue);
    }
    release(*a);
    return hipSuccess;
}
hipError_t hipMallocAsync(void** p, size_t bytes, hipStream_t s) {
    LOCK;
    void* q = take(bytes, 1);
    if (!q) return ret(hipErrorMemoryAllocation);
    // the block may be used only by stream work that runs after this point of the stream
    Alloc* a = find(reinterpret_cast<uintptr_t>(q));
    a->live = false;
    const uintptr_t base = a->base;
    ModelOp op;
    op.run = [base] {
        auto it = g_allocs.find(base);
        it->second.live = true;
    };
    push(s, std::move(op));
    *p = q;
    return hipSuccess;
}
hipError_t hipFreeAsync(void* p, hipStream_t s) {
    LOCK;
    if (!p) return hipSuccess;
    const uintptr_t base = reinterpret_cast<uintptr_t>(p);
    ModelOp op;
    op.run = [base] {
        auto it = g_allocs.find(base);
        if (it == g_allocs.end() || !it->second.live || it->second.kind != 1) {
            violation("hipFreeAsync(%p): not a live stream-ordered allocation when the free runs", reinterpret_cast<void*>(base));
            return;
        }
        release(it->second);
    };
    push(s, std::move(op));
    return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) {
    LOCK;
    *p = take(bytes, 2);
    return *p ? hipSuccess : ret(hipErrorMemoryAllocation);
}
hipError_t hipHostFree(void* p) {
    LOCK;
    if (!p) return hipSuccess;
    Alloc* a = find(reinterpret_cast<uintptr_t>(p));
    if (!a || a->base != reinterpret_cast<uintptr_t>(p) || !a->live || a->kind != 2) {
        violation("hipHostFree(%p): not a live pinned allocation", p);
        return ret(hipErrorInvalidValue);
    }
    release(*a);
    return hipSuccess;
}

namespace {
// One copy as a stream operation. Pageable host sources are read now; everything else when it runs.
ModelOp copy_op(void* dst, const void* src, size_t bytes) {
    ModelOp op;
    if (!in_arena(src)) {
        auto snap = std::make_shared<std::vector<char>>((const char*)src, (const char*)src + bytes);
        op.run = [dst, snap, bytes] {
            if (!check(dst, bytes, "copy destination")) return;
            memcpy(dst, snap->data(), bytes);
            find(reinterpret_cast<uintptr_t>(dst))->epoch++;
        };
    } else {
        op.run = [dst, src, bytes] {
            if (!check(src, bytes, "copy source")) return;
            if (in_arena(dst)) {
                if (!check(dst, bytes, "copy destination")) return;
                memcpy(dst, src, bytes);
                find(reinterpret_cast<uintptr_t>(dst))->epoch++;
            } else {
                memcpy(dst, src, bytes);
            }
        };
    }
    return op;
}
}  // namespace

hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind) {
    LOCK;
    run_all("hipMemcpy");
    ModelOp op = copy_op(dst, src, bytes);
    op.run();
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind, hipStream_t s) {
    LOCK;
    const bool to_pageable = !in_arena(dst);
    push(s, copy_op(dst, src, bytes));
    if (to_pageable && !s->capture) {   // a copy into pageable memory returns when it is done
        run_until([s] { return s->q.empty(); }, "hipMemcpyAsync to pageable memory");
    }
    return hipSuccess;
}
hipError_t hipMemsetAsync(void* dst, int value, size_t bytes, hipStream_t s) {
    LOCK;
    ModelOp op;
    op.run = [dst, value, bytes] {
        if (!check(dst, bytes, "memset destination")) return;
        memset(dst, value, bytes);
        find(reinterpret_cast<uintptr_t>(dst))->epoch++;
    };
    push(s, std::move(op));
    return hipSuccess;
}
hipError_t hipGetSymbolAddress(void** out, const void* symbol) {
    LOCK;
    // A symbol's device copy: one block of the model's address range per symbol address.
    static std::map<const void*, void*> copies;
    auto it = copies.find(symbol);
    // (the model cannot see a symbol's size; 256 MiB covers every array of the library and,
    // untouched, costs no memory)
    if (it == copies.end()) it = copies.emplace(symbol, take(size_t(256) << 20, 0)).first;
    *out = it->second;
    return *out ? hipSuccess : ret(hipErrorMemoryAllocation);
}

// The stream-ordered pool's settings do not change what the model checks.
hipError_t hipDeviceGetDefaultMemPool(hipMemPool_t* pool, int) {
    *pool = reinterpret_cast<hipMemPool_t>(uintptr_t(1));
    return hipSuccess;
}
hipError_t hipMemPoolSetAttribute(hipMemPool_t, hipMemPoolAttr, void*) { return hipSuccess; }

hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
    LOCK;
    *s = new ModelStream();
    g_streams.push_back(*s);
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
    LOCK;
    if (!s || s->destroyed) return ret(hipErrorInvalidResourceHandle);
    run_until([s] { return s->q.empty(); }, "hipStreamDestroy");
    s->destroyed = true;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
    LOCK;
    run_until([s] { return s->q.empty(); }, "hipStreamSynchronize");
    return hipSuccess;
}
hipError_t hipDeviceSynchronize() {
    LOCK;
    run_all("hipDeviceSynchronize");
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned) {
    LOCK;
    if (e->destroyed) violation("hipStreamWaitEvent on a destroyed event");
    if (e->recorded == 0) return hipSuccess;
    ModelOp op;
    op.kind = ModelOp::kWait;
    op.ev = e;
    op.gen = e->recorded;
    push(s, std::move(op));
    return hipSuccess;
}

hipError_t hipEventCreate(hipEvent_t* e) {
    LOCK;
    g_stats.events++;
    *e = new ModelEvent();   // never freed: a use after destroy stays detectable
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
    LOCK;
    if (!e || e->destroyed) {
        violation("hipEventDestroy of a destroyed event");
        return ret(hipErrorInvalidResourceHandle);
    }
    e->destroyed = true;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
    LOCK;
    if (e->destroyed) violation("hipEventRecord on a destroyed event");
    const uint64_t gen = ++e->recorded;
    ModelOp op;
    op.run = [e, gen] { e->completed = std::max(e->completed, gen); };
    push(s, std::move(op));
    return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t e) {
    LOCK;
    if (e->destroyed) violation("hipEventQuery on a destroyed event");
    run_random(6);
    return e->completed >= e->recorded ? hipSuccess : hipErrorNotReady;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
    LOCK;
    if (e->destroyed) violation("hipEventSynchronize on a destroyed event");
    run_until([e] { return e->completed >= e->recorded; }, "hipEventSynchronize");
    return hipSuccess;
}
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    LOCK;
    if (a->completed < a->recorded || b->completed < b->recorded) return ret(hipErrorNotReady);
    *ms = 0.05f;
    return hipSuccess;
}

hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    LOCK;
    if (s->capture) return ret(hipErrorIllegalState);
    s->capture = new ModelGraph();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
    LOCK;
    if (!s->capture) return ret(hipErrorIllegalState);
    *g = s->capture;
    s->capture = nullptr;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* exec, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
    LOCK;
    *exec = new ModelGraph(*g);
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t exec, hipStream_t s) {
    LOCK;
    for (const ModelOp& op : exec->ops) push(s, op);
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t exec) {
    LOCK;
    delete exec;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
    LOCK;
    delete g;
    return hipSuccess;
}

// Model queries for the replay driver.
extern "C" __attribute__((visibility("default"))) size_t mmr_model_violations(char* buffer, size_t capacity) {
    LOCK;
    std::string all;
    for (const std::string& v : g_violations) all += v + "\n";
    if (buffer && capacity) {
        const size_t n = std::min(capacity - 1, all.size());
        memcpy(buffer, all.data(), n);
        buffer[n] = 0;
    }
    return g_violation_count;
}

// kernels run, checks, stream operations run, jobs checked, device / pinned allocations, events created
extern "C" __attribute__((visibility("default"))) void mmr_model_stats(uint64_t* out6) {
    LOCK;
    out6[0] = g_stats.kernels;
    out6[1] = g_stats.checks;
    out6[2] = g_stats.ops_run;
    out6[3] = g_stats.jobs;
    out6[4] = g_stats.allocations;
    out6[5] = g_stats.events;
}

extern "C" __attribute__((visibility("default"))) void mmr_model_seed(uint64_t seed) {
    LOCK;
    g_rng.seed(seed);
}
