// Host side of libmmr_recon: the C ABI of include/mmr_recon.h.
//
// A context owns one device and one HIP stream. Every submit call validates its descriptors,
// packs them into device job records (recon_internal.h), uploads them with the work-queue
// counters, launches the kernels stream-ordered and returns a ticket; completion is observed
// through an event per ticket. Errors follow crates/capi: status codes plus a thread-local
// message (crates/capi/src/lib.rs:46-48, :305-321). Kernels report unsupported or invalid records
// found on the device through a per-ticket status word that mmr_recon_wait / mmr_recon_poll
// return. No C++ exception leaves the library: every export runs its body inside `guarded`, which
// turns an exception (std::bad_alloc from a container, std::system_error from a thread) into
// MMR_STATUS_INTERNAL_ERROR with the message set, as crates/capi wraps every export in
// catch_unwind (crates/capi/src/lib.rs:572-588).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <atomic>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/mmr_recon.h"
#include "recon_internal.h"

namespace {

// The thread's last error message: a fixed buffer, so that recording an error never allocates
// (and never throws, also while an out-of-memory exception is being turned into a status).
thread_local char g_last_error[512];

mmr_status fail(mmr_status status, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof g_last_error, fmt, ap);
    va_end(ap);
    return status;
}

mmr_status hip_fail(hipError_t e, const char* what) {
    return fail(MMR_STATUS_INTERNAL_ERROR, "%s: %s", what, hipGetErrorString(e));
}

// Runs an export's body; an exception becomes MMR_STATUS_INTERNAL_ERROR with the message set.
template <typename F>
mmr_status guarded(const char* what, F&& body) noexcept {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return fail(MMR_STATUS_INTERNAL_ERROR, "%s: out of host memory", what);
    } catch (const std::exception& e) {
        return fail(MMR_STATUS_INTERNAL_ERROR, "%s: internal error: %s", what, e.what());
    } catch (...) {
        return fail(MMR_STATUS_INTERNAL_ERROR, "%s: internal error (unknown exception)", what);
    }
}

// Same for the exports without a status (destroy calls): the error is only recorded.
template <typename F>
void guarded_void(const char* what, F&& body) noexcept {
    (void)guarded(what, [&]() -> mmr_status {
        body();
        return MMR_STATUS_OK;
    });
}

struct TicketRecord {
    hipEvent_t start = nullptr;       // before the first kernel (after the job upload)
    hipEvent_t kernels_done = nullptr;  // after the last kernel
    hipEvent_t done = nullptr;        // after the status word is copied back
    hipEvent_t marks[MMR_RECON_MAX_PHASES - 1] = {};   // between launch phases
    uint32_t mark_count = 0;
    std::vector<double> phase_ms;     // start -> marks[0] -> ... -> kernels_done
    std::vector<char> host_jobs;      // host copy of the job records (prepared / allocation-free modes)
    char* pinned = nullptr;           // pinned staging: status word at 0, job records at kPinnedJobs
    size_t pinned_bytes = 0;
    uint32_t* host_status = nullptr;  // pinned; device status word copied here
    bool finished = false;
    bool reported = false;            // status returned to the caller (wait / poll / server completion)
    mmr_status status = MMR_STATUS_OK;
    double kernel_ms = -1.0;
};

// Offset of the job records in a ticket's pinned staging block (the status word sits at 0).
constexpr size_t kPinnedJobs = 256;

constexpr uint32_t kStatusUnsupported = mmr::kDevStatusUnsupported;
constexpr uint32_t kStatusInvalidData = mmr::kDevStatusInvalidData;

}  // namespace

// Resampling taps of one pass (filter, source length n, scaled length m) for every output
// coordinate: left tap, tap count, and weights padded to `max` per coordinate.
struct ScaleTaps {
    std::vector<int32_t> left;
    std::vector<uint32_t> count;
    std::vector<float> w;
    uint32_t max = 0;
};

// Device scratch of the allocation-free entry point (mmr_recon_mpeg2_reconstruct_noalloc): the
// submission's job records, tile table, counters, status word and block order live here instead of
// a stream-ordered allocation. Every GPU the library runs on holds its own copy, from the load of
// the code object until the process ends. On the host the name is only a stand-in for the symbol,
// not a device address: host code uses the per-device pointer that noalloc_scratch() obtains once
// per device with hipGetSymbolAddress, never the name itself.
static __device__ __align__(256) char scratch[MMR_RECON_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
std::map<int, char*> g_scratch_ptr;   // device ordinal -> that device's copy of `scratch`

// The current device's copy of `scratch`, resolved on the first call for the device.
hipError_t noalloc_scratch(int device, char** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it == g_scratch_ptr.end()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        it = g_scratch_ptr.emplace(device, static_cast<char*>(p)).first;
    }
    *out = it->second;
    return hipSuccess;
}
}  // namespace

struct mmr_recon_context {
    int device = 0;
    char* noalloc_scratch = nullptr;  // this device's copy of `scratch`
    hipStream_t stream = nullptr;
    uint32_t queue_depth = 256;
    uint32_t cu_count = 256;
    uint32_t persistent_groups = 0;
    std::mutex mu;
    uint64_t next_ticket = 1;
    std::map<uint64_t, TicketRecord> tickets;
    uint8_t* flat_scaling = nullptr;  // device Flat_16 scaling lists ([6][16] + [2][64])
    uint8_t* linear_tables = nullptr; // device srgb_to_linear_u16 + linear_u16_to_srgb, on first use
    mmr_recon_tuning tuning{};        // launch shapes used by the next submissions
    // Scaler taps by (filter, n, m); they depend on the geometry only (bounded: cleared when full).
    std::map<std::tuple<uint32_t, uint32_t, uint32_t>, std::shared_ptr<const ScaleTaps>> scale_taps;
    // Ticket resources kept for reuse once a ticket has completed: timing events and pinned
    // staging blocks by size (powers of two), so a submission neither creates events nor pins
    // memory in the steady state. Bounded; the surplus is released.
    std::vector<hipEvent_t> free_events;
    std::multimap<size_t, char*> free_pinned;
    size_t free_pinned_bytes = 0;
    // Pinned staging ring: one pinned allocation made on the first submission, carved in
    // submission order and released as tickets complete (in any order; the tail advances over
    // released spans). A submission that does not fit takes a pooled block instead. Allocating
    // pinned memory per submission would stall the stream pipeline.
    // Device block of the submissions on the context's own stream: the stream runs them in order,
    // so each submission's upload into it starts after the previous submission's kernels have
    // finished and the one block serves all of them (grown, stream-ordered, when one needs more).
    char* stream_block = nullptr;
    size_t stream_block_cap = 0;
    char* ring = nullptr;
    size_t ring_cap = 0, ring_head = 0, ring_tail = 0;
    struct RingSpan {
        size_t off, bytes;
        bool released;
    };
    std::deque<RingSpan> ring_live;
};

namespace {

constexpr size_t kMaxPooledEvents = 4096;
constexpr size_t kMaxPooledPinnedBytes = size_t(256) << 20;

hipError_t take_event(mmr_recon_context* ctx, hipEvent_t* ev) {
    if (!ctx->free_events.empty()) {
        *ev = ctx->free_events.back();
        ctx->free_events.pop_back();
        return hipSuccess;
    }
    return hipEventCreate(ev);
}

void give_event(mmr_recon_context* ctx, hipEvent_t& ev) {
    if (!ev) return;
    if (ctx->free_events.size() < kMaxPooledEvents) ctx->free_events.push_back(ev);
    else (void)hipEventDestroy(ev);
    ev = nullptr;
}

constexpr size_t kPinnedRingBytes = size_t(32) << 20;

// `bytes` (a multiple of 256) from the ring, or nullptr when it does not fit now.
char* ring_take(mmr_recon_context* ctx, size_t bytes) {
    if (!ctx->ring) {
        if (hipHostMalloc((void**)&ctx->ring, kPinnedRingBytes, hipHostMallocDefault) != hipSuccess) {
            ctx->ring = nullptr;
            return nullptr;
        }
        ctx->ring_cap = kPinnedRingBytes;
    }
    if (ctx->ring_live.empty()) ctx->ring_head = ctx->ring_tail = 0;
    size_t off;
    if (ctx->ring_live.empty() || ctx->ring_head > ctx->ring_tail) {   // free: [head, cap) and [0, tail)
        if (ctx->ring_cap - ctx->ring_head >= bytes) off = ctx->ring_head;
        else if (!ctx->ring_live.empty() && ctx->ring_tail > bytes) off = 0;
        else if (ctx->ring_live.empty() && ctx->ring_cap >= bytes) off = 0;
        else return nullptr;
    } else {                                                           // free: [head, tail)
        if (ctx->ring_tail - ctx->ring_head > bytes) off = ctx->ring_head;
        else return nullptr;
    }
    ctx->ring_head = off + bytes;
    ctx->ring_live.push_back({off, bytes, false});
    return ctx->ring + off;
}

bool ring_give(mmr_recon_context* ctx, char* p) {
    if (!ctx->ring || p < ctx->ring || p >= ctx->ring + ctx->ring_cap) return false;
    const size_t off = (size_t)(p - ctx->ring);
    for (auto& sp : ctx->ring_live)
        if (sp.off == off && !sp.released) {
            sp.released = true;
            break;
        }
    while (!ctx->ring_live.empty() && ctx->ring_live.front().released) {
        ctx->ring_tail = ctx->ring_live.front().off + ctx->ring_live.front().bytes;
        ctx->ring_live.pop_front();
    }
    if (ctx->ring_live.empty()) ctx->ring_head = ctx->ring_tail = 0;
    else ctx->ring_tail = ctx->ring_live.front().off;
    return true;
}

// Pinned staging of at least `bytes`: from the ring, else a pooled block (a power of two, at least
// 4 KiB).
hipError_t take_pinned(mmr_recon_context* ctx, size_t bytes, char** out, size_t* out_bytes) {
    const size_t rounded = (bytes + 255) & ~size_t(255);
    if (char* r = ring_take(ctx, rounded)) {
        *out = r;
        *out_bytes = rounded;
        return hipSuccess;
    }
    size_t cap = 4096;
    while (cap < bytes) cap <<= 1;
    auto it = ctx->free_pinned.find(cap);
    if (it != ctx->free_pinned.end()) {
        *out = it->second;
        ctx->free_pinned_bytes -= cap;
        ctx->free_pinned.erase(it);
    } else {
        hipError_t e = hipHostMalloc((void**)out, cap, hipHostMallocDefault);
        if (e != hipSuccess) return e;
    }
    *out_bytes = cap;
    return hipSuccess;
}

void give_pinned(mmr_recon_context* ctx, char*& p, size_t bytes) {
    if (!p) return;
    if (ring_give(ctx, p)) {
        p = nullptr;
        return;
    }
    if (ctx->free_pinned_bytes + bytes <= kMaxPooledPinnedBytes) {
        ctx->free_pinned.emplace(bytes, p);
        ctx->free_pinned_bytes += bytes;
    } else {
        (void)hipHostFree(p);
    }
    p = nullptr;
}

// Returns a ticket's events and pinned block to the pools. Only for tickets whose work has
// completed (or never reached the stream).
void release_ticket(mmr_recon_context* ctx, TicketRecord& r) {
    give_event(ctx, r.start);
    give_event(ctx, r.kernels_done);
    give_event(ctx, r.done);
    for (uint32_t m = 0; m < r.mark_count; ++m) give_event(ctx, r.marks[m]);
    r.mark_count = 0;
    give_pinned(ctx, r.pinned, r.pinned_bytes);
    r.host_status = nullptr;
}

// Retire finished tickets from the front so the map stays bounded.
void reap(mmr_recon_context* ctx) {
    while (!ctx->tickets.empty()) {
        auto it = ctx->tickets.begin();
        if (!it->second.finished) {
            if (hipEventQuery(it->second.done) != hipSuccess) break;
        }
        release_ticket(ctx, it->second);
        ctx->tickets.erase(it);
    }
}

// The status a submission's device status word stands for (message set when not OK).
mmr_status status_from_word(uint32_t st) {
    if (st & mmr::kDevStatusTimeout)
        return fail(MMR_STATUS_INTERNAL_ERROR, "a device-side wait timed out (internal error); outputs are invalid");
    if (st & kStatusUnsupported)
        return fail(MMR_STATUS_UNSUPPORTED, "a record of this submission uses a feature the device path does not reconstruct "
                                            "(MPEG-2 field prediction or H.264 intra macroblock); fall back to CPU reconstruction");
    if (st & kStatusInvalidData)
        return fail(MMR_STATUS_INVALID_DATA, "a record of this submission is invalid (H.264 reference index or MPEG-2 level record index out of range)");
    return MMR_STATUS_OK;
}

mmr_status finish(mmr_recon_context* ctx, TicketRecord& r) {
    if (!r.finished) {
        hipError_t e = hipEventSynchronize(r.done);
        if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.start, r.kernels_done) == hipSuccess) r.kernel_ms = ms;
        hipEvent_t from = r.start;
        for (uint32_t m = 0; m <= r.mark_count; ++m) {
            hipEvent_t to = m < r.mark_count ? r.marks[m] : r.kernels_done;
            float seg = 0.f;
            r.phase_ms.push_back(hipEventElapsedTime(&seg, from, to) == hipSuccess ? seg : -1.0);
            from = to;
        }
        r.status = status_from_word(r.host_status ? *r.host_status : 0u);
        r.finished = true;
        release_ticket(ctx, r);   // the ticket's work is complete: its events and staging are free
    }
    return r.status;
}

// A submission in flight: records the start event, owns the device allocation of job records +
// counters + status word, and finishes by recording the completion event.
// A submission prepared for replay: the device block is a persistent allocation and the kernel
// launches (counter fill included) are captured into a graph on `stream`.
struct PreparedCapture {
    hipStream_t stream = nullptr;
    char* d_block = nullptr;
    uint32_t* status = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

// Device memory a submission may use instead of allocating its block: the allocation-free entry
// point passes the device's copy of the file-scope `scratch` array (see below).
struct Workspace {
    char* base = nullptr;
    size_t capacity = 0;
};

struct Submission {
    mmr_recon_context* ctx;
    hipStream_t stream;
    PreparedCapture* cap;
    const Workspace* ws;      // non-null: synchronous submission inside this memory, no allocation
    bool capturing = false;
    bool enqueued = false;    // stream work of this submission may be in flight
    bool owns_block = true;   // d_block is this submission's own allocation (not the context's stream block)
    TicketRecord rec;
    char* d_block = nullptr;
    size_t jobs_bytes = 0;
    size_t counter_words = 0;
    size_t counters_bytes = 0;

    Submission(mmr_recon_context* c, hipStream_t s, PreparedCapture* capture = nullptr, const Workspace* workspace = nullptr)
        : ctx(c), stream(s), cap(capture), ws(workspace) {}

    // Workspace mode: the block is the workspace (checked against its capacity), the job records
    // are copied from pageable host memory, the counters and status word cleared by the fill
    // kernel; no device, pinned or event allocation.
    mmr_status begin_workspace(const void* jobs, size_t bytes, size_t scratch_bytes,
                               const std::function<void(char* host_jobs, char* scratch)>& patch) {
        const size_t need = jobs_bytes + counters_bytes + scratch_bytes;
        if (need > ws->capacity)
            return fail(MMR_STATUS_INVALID_ARGUMENT, "the submission needs %zu bytes of device scratch, the allocation-free entry point "
                        "has %zu (MMR_RECON_NOALLOC_SCRATCH_BYTES); submit fewer or smaller pictures", need, ws->capacity);
        d_block = ws->base;
        rec.host_jobs.assign((const char*)jobs, (const char*)jobs + bytes);
        if (patch) patch(rec.host_jobs.data(), scratch_ptr());
        hipError_t e = hipSuccess;
        if (bytes) e = hipMemcpyAsync(d_block, rec.host_jobs.data(), bytes, hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = mmr::launch_fill_zero(counters_ptr(), (uint32_t)counter_words, stream);
        if (e != hipSuccess) return hip_fail(e, "job upload");
        return MMR_STATUS_OK;
    }
    // Waits for the submission and returns its status.
    mmr_status end_workspace() {
        uint32_t word = 0;
        hipError_t e = hipMemcpyAsync(&word, status_ptr(), sizeof word, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        d_block = nullptr;
        if (e != hipSuccess) return hip_fail(e, "allocation-free submission");
        return status_from_word(word);
    }

    mmr_status begin_capture(const void* jobs, size_t bytes, size_t scratch_bytes,
                             const std::function<void(char* host_jobs, char* scratch)>& patch) {
        rec.host_jobs.assign((const char*)jobs, (const char*)jobs + bytes);
        hipError_t e = hipMalloc((void**)&d_block, jobs_bytes + counters_bytes + scratch_bytes);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc (prepared batch)");
        if (patch) patch(rec.host_jobs.data(), scratch_ptr());
        if (bytes) e = hipMemcpy(d_block, rec.host_jobs.data(), bytes, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return hip_fail(e, "prepared batch upload / capture");
        capturing = true;
        e = mmr::launch_fill_zero(counters_ptr(), (uint32_t)counter_words, stream);
        if (e != hipSuccess) return hip_fail(e, "counter fill");
        return MMR_STATUS_OK;
    }
    mmr_status end_capture() {
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(stream, &g);
        capturing = false;
        if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
        cap->graph = g;
        e = hipGraphInstantiate(&cap->exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
        cap->d_block = d_block;
        cap->status = status_ptr();
        d_block = nullptr;
        return MMR_STATUS_OK;
    }

    // Device block: [jobs (uploaded)][counters + status word (zeroed)][scratch (uninitialised)].
    // `patch` (optional) edits the host copy of the job records once the device block exists, so
    // records can point into the block's scratch.
    mmr_status begin(const void* jobs, size_t bytes, size_t counters, size_t scratch_bytes = 0,
                     const std::function<void(char* host_jobs, char* scratch)>& patch = nullptr) {
        if (cap || ws) {
            jobs_bytes = (bytes + 255) & ~size_t(255);
            counter_words = counters + 1;
            counters_bytes = (counter_words * sizeof(uint32_t) + 255) & ~size_t(255);
            return cap ? begin_capture(jobs, bytes, scratch_bytes, patch) : begin_workspace(jobs, bytes, scratch_bytes, patch);
        }
        if (ctx->tickets.size() >= ctx->queue_depth) reap(ctx);
        if (ctx->tickets.size() >= ctx->queue_depth)
            return fail(MMR_STATUS_BUSY, "queue full: %u tickets in flight", ctx->queue_depth);
        jobs_bytes = (bytes + 255) & ~size_t(255);
        counter_words = counters + 1;  // + status word
        counters_bytes = (counter_words * sizeof(uint32_t) + 255) & ~size_t(255);
        hipError_t e = take_event(ctx, &rec.start);
        if (e == hipSuccess) e = take_event(ctx, &rec.kernels_done);
        if (e == hipSuccess) e = take_event(ctx, &rec.done);
        if (e == hipSuccess) e = take_pinned(ctx, kPinnedJobs + bytes, &rec.pinned, &rec.pinned_bytes);
        if (e != hipSuccess) return hip_fail(e, "ticket setup");
        rec.host_status = reinterpret_cast<uint32_t*>(rec.pinned);
        *rec.host_status = 0;
        char* staged = rec.pinned + kPinnedJobs;
        if (bytes) memcpy(staged, jobs, bytes);
        const size_t need = jobs_bytes + counters_bytes + scratch_bytes;
        if (stream == ctx->stream) {
            if (need > ctx->stream_block_cap) {
                if (ctx->stream_block) (void)hipFreeAsync(ctx->stream_block, stream);
                ctx->stream_block = nullptr;
                ctx->stream_block_cap = 0;
                const size_t grown = std::max(need, need + need / 4);
                e = hipMallocAsync((void**)&ctx->stream_block, grown, stream);
                if (e != hipSuccess) {
                    ctx->stream_block = nullptr;
                    return hip_fail(e, "hipMallocAsync");
                }
                ctx->stream_block_cap = grown;
            }
            d_block = ctx->stream_block;
            owns_block = false;
        } else {
            e = hipMallocAsync((void**)&d_block, need, stream);
            if (e != hipSuccess) return hip_fail(e, "hipMallocAsync");
            owns_block = true;
        }
        enqueued = true;
        if (patch) patch(staged, scratch_ptr());
        if (bytes) e = hipMemcpyAsync(d_block, staged, bytes, hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = mmr::launch_fill_zero(counters_ptr(), (uint32_t)counter_words, stream);
        if (e == hipSuccess) e = hipEventRecord(rec.start, stream);
        if (e != hipSuccess) return hip_fail(e, "job upload");
        return MMR_STATUS_OK;
    }
    // Ends a launch phase: an event on the stream between the kernels before and after it (timed
    // submissions only; prepared graphs and the allocation-free entry point keep one phase).
    mmr_status mark() {
        if (cap || ws || rec.mark_count >= MMR_RECON_MAX_PHASES - 1) return MMR_STATUS_OK;
        hipEvent_t ev = nullptr;
        hipError_t e = take_event(ctx, &ev);
        if (e != hipSuccess) return hip_fail(e, "hipEventCreate (phase)");
        rec.marks[rec.mark_count++] = ev;
        e = hipEventRecord(ev, stream);
        if (e != hipSuccess) return hip_fail(e, "hipEventRecord (phase)");
        return MMR_STATUS_OK;
    }
    void* jobs_ptr() const { return d_block; }
    uint32_t* counters_ptr() const { return reinterpret_cast<uint32_t*>(d_block + jobs_bytes); }
    uint32_t* status_ptr() const { return counters_ptr() + (counter_words - 1); }
    char* scratch_ptr() const { return d_block + jobs_bytes + counters_bytes; }

    mmr_status end(uint64_t* out_ticket) {
        if (cap) return end_capture();
        if (ws) return end_workspace();
        hipError_t e = hipEventRecord(rec.kernels_done, stream);
        if (e == hipSuccess) e = hipMemcpyAsync(rec.host_status, status_ptr(), sizeof(uint32_t), hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipEventRecord(rec.done, stream);
        if (e == hipSuccess && owns_block) e = hipFreeAsync(d_block, stream);
        d_block = nullptr;
        if (e != hipSuccess) return hip_fail(e, "submission end");
        const uint64_t id = ctx->next_ticket++;
        ctx->tickets.emplace(id, std::move(rec));
        rec = TicketRecord();
        enqueued = false;
        if (out_ticket) *out_ticket = id;
        return MMR_STATUS_OK;
    }
    ~Submission() {
        if (capturing) {   // a launch failed during capture: end it and drop the partial graph
            hipGraph_t g = nullptr;
            if (hipStreamEndCapture(stream, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
        }
        if (d_block && !ws) {
            if (cap) (void)hipFree(d_block);
            else if (owns_block) (void)hipFreeAsync(d_block, stream);
        } else if (d_block) {
            (void)hipStreamSynchronize(stream);   // a failed workspace submission: the next call may reuse the memory
        }
        // A submission that failed after enqueueing work: the upload may still read the pinned
        // staging and the events may still be recorded, so wait before they go back to the pools.
        if (enqueued && (rec.pinned || rec.start)) (void)hipStreamSynchronize(stream);
        release_ticket(ctx, rec);
    }
};

template <typename T>
bool aligned(const T* p, size_t a) { return (reinterpret_cast<uintptr_t>(p) & (a - 1)) == 0; }

uint32_t groups_for(mmr_recon_context* ctx, uint32_t per_cu) {
    return ctx->persistent_groups ? ctx->persistent_groups : ctx->cu_count * per_cu;
}

// Appends the tiles of one job (ceil(blocks / 32) entries holding `job`) to a tile -> job table
// and returns the index of its first tile.
// Appends the tiles of `blocks` work items (tiles of `per_tile`) of job `job` to a tile -> job
// table; returns the job's first tile.
uint32_t add_tiles(std::vector<uint32_t>& table, uint32_t job, uint32_t blocks, uint32_t per_tile = 32) {
    const uint32_t first = (uint32_t)table.size();
    table.insert(table.end(), (blocks + per_tile - 1) / per_tile, job);
    return first;
}

// Appends `bytes` at a 16-byte aligned offset of `blob` and returns that offset.
size_t blob_append(std::vector<char>& blob, const void* data, size_t bytes) {
    const size_t off = (blob.size() + 15) & ~size_t(15);
    blob.resize(off + bytes);
    if (bytes) memcpy(blob.data() + off, data, bytes);
    return off;
}

#define CHECK_CTX(ctx)                                                                 \
    do {                                                                               \
        if (!(ctx)) return fail(MMR_STATUS_INVALID_ARGUMENT, "context is NULL");       \
    } while (0)

#define LOCK_AND_SET_DEVICE(ctx)                                                       \
    std::lock_guard<std::mutex> lock_((ctx)->mu);                                      \
    do {                                                                               \
        hipError_t e_ = hipSetDevice((ctx)->device);                                   \
        if (e_ != hipSuccess) return hip_fail(e_, "hipSetDevice");                     \
    } while (0)

// ------------------------------------------------------------------ MJPEG helpers
mmr_status build_mjpeg(const mmr_recon_mjpeg_frame& f, mmr::MjpegJob& j, uint32_t& rgb_quads) {
    if (f.struct_size < sizeof(mmr_recon_mjpeg_frame)) return fail(MMR_STATUS_INVALID_ARGUMENT, "MJPEG frame struct_size too small");
    if (f.width == 0 || f.height == 0 || f.width > 65535 || f.height > 65535)
        return fail(MMR_STATUS_INVALID_ARGUMENT, "MJPEG frame size %ux%u is invalid", f.width, f.height);
    if (f.component_count < 1 || f.component_count > 4)
        return fail(MMR_STATUS_INVALID_ARGUMENT, "MJPEG component count %u is invalid", f.component_count);
    memset(&j, 0, sizeof j);
    uint32_t hmax = 0, vmax = 0, per_mcu = 0;
    for (uint32_t c = 0; c < f.component_count; ++c) {
        if (f.h_sampling[c] < 1 || f.h_sampling[c] > 4 || f.v_sampling[c] < 1 || f.v_sampling[c] > 4)
            return fail(MMR_STATUS_INVALID_ARGUMENT, "MJPEG sampling factors of component %u are invalid", c);
        if (f.quant_table[c] > 3) return fail(MMR_STATUS_INVALID_ARGUMENT, "MJPEG quantization table index > 3");
        if (!f.planes[c]) return fail(MMR_STATUS_INVALID_ARGUMENT, "MJPEG plane %u is NULL", c);
        hmax = std::max<uint32_t>(hmax, f.h_sampling[c]);
        vmax = std::max<uint32_t>(vmax, f.v_sampling[c]);
        per_mcu += f.h_sampling[c] * f.v_sampling[c];
    }
    const bool interleaved = f.component_count > 1;
    if (interleaved && per_mcu > 10) return fail(MMR_STATUS_INVALID_DATA, "MJPEG MCU holds %u blocks (max 10)", per_mcu);
    const uint32_t mcu_w = interleaved ? hmax * 8 : 8, mcu_h = interleaved ? vmax * 8 : 8;
    const uint32_t cols = (f.width + mcu_w - 1) / mcu_w, rows = (f.height + mcu_h - 1) / mcu_h;
    const uint64_t blocks = (uint64_t)cols * rows * (interleaved ? per_mcu : 1);
    if (blocks != f.block_count)
        return fail(MMR_STATUS_INVALID_DATA, "MJPEG block_count %u does not match the MCU walk (%llu)", f.block_count,
                    (unsigned long long)blocks);
    if (!f.levels || !f.quant_tables) return fail(MMR_STATUS_INVALID_ARGUMENT, "MJPEG levels / quant_tables is NULL");
    if (!aligned(f.levels, 16)) return fail(MMR_STATUS_INVALID_ARGUMENT, "MJPEG levels must be 16-byte aligned");
    j.levels = f.levels;
    j.quant = f.quant_tables;
    j.mcu_cols = cols;
    j.blocks_per_mcu = interleaved ? per_mcu : 1;
    j.block_count = f.block_count;
    uint32_t k = 0;
    for (uint32_t c = 0; c < f.component_count; ++c) {
        j.planes[c] = f.planes[c];
        j.plane_w[c] = (f.width * f.h_sampling[c] + hmax - 1) / hmax;
        j.plane_h[c] = (f.height * f.v_sampling[c] + vmax - 1) / vmax;
        j.comp_h[c] = interleaved ? f.h_sampling[c] : 1;
        j.comp_v[c] = interleaved ? f.v_sampling[c] : 1;
        j.qsel[c] = f.quant_table[c];
        const uint32_t bh = interleaved ? f.v_sampling[c] : 1, bw = interleaved ? f.h_sampling[c] : 1;
        for (uint32_t by = 0; by < bh; ++by)
            for (uint32_t bx = 0; bx < bw; ++bx, ++k) {
                j.blk_comp[k] = (uint8_t)c;
                j.blk_bx[k] = (uint8_t)bx;
                j.blk_by[k] = (uint8_t)by;
            }
    }
    rgb_quads = 0;
    if (f.rgb) {
        if (f.component_count != 3) return fail(MMR_STATUS_UNSUPPORTED, "RGB output needs 3 components");
        if (f.h_sampling[0] != hmax || f.v_sampling[0] != vmax)
            return fail(MMR_STATUS_UNSUPPORTED, "RGB output needs the luma component at the largest sampling factors");
        rgb_quads = ((f.width + 3) / 4) * f.height;
    }
    return MMR_STATUS_OK;
}

mmr::RgbJob rgb_job_of(const mmr_recon_mjpeg_frame& f, const mmr::MjpegJob& j) {
    mmr::RgbJob r;
    memset(&r, 0, sizeof r);
    r.planes[0] = j.planes[0];
    r.planes[1] = j.planes[1];
    r.planes[2] = j.planes[2];
    r.rgb = f.rgb;
    r.width = f.width;
    r.height = f.height;
    r.cw = j.plane_w[1];
    r.ch = j.plane_h[1];
    r.luma_stride = j.plane_w[0];
    r.chroma_stride = j.plane_w[1];
    r.range = f.rgb_range ? f.rgb_range : MMR_RANGE_FULL;
    r.matrix = MMR_MATRIX_BT601;
    r.inv_width = 1.0 / f.width;
    r.inv_height = 1.0 / f.height;
    mmr::set_rgb_layout(r);
    mmr::set_rgb_terms(r);
    return r;
}

mmr_status check_rgb_job(const mmr_recon_yuv_to_rgb_job& s) {
    if (s.struct_size < sizeof(mmr_recon_yuv_to_rgb_job)) return fail(MMR_STATUS_INVALID_ARGUMENT, "YUV->RGB struct_size too small");
    if (!s.width || !s.height || !s.chroma_width || !s.chroma_height)
        return fail(MMR_STATUS_INVALID_ARGUMENT, "YUV->RGB job has zero dimensions");
    if (s.width > 65535 || s.height > 65535 || s.chroma_width > s.width || s.chroma_height > s.height)
        return fail(MMR_STATUS_INVALID_ARGUMENT, "YUV->RGB job: sizes above 65535 or chroma larger than luma");
    if (s.luma_stride < s.width || s.chroma_stride < s.chroma_width)
        return fail(MMR_STATUS_INVALID_ARGUMENT, "YUV->RGB stride smaller than width");
    if (!s.planes[0] || !s.planes[1] || !s.planes[2] || (!s.rgb && !s.rgba))
        return fail(MMR_STATUS_INVALID_ARGUMENT, "YUV->RGB job has a NULL plane or no output");
    if (s.range != MMR_RANGE_FULL && s.range != MMR_RANGE_LIMITED) return fail(MMR_STATUS_INVALID_ARGUMENT, "YUV->RGB range invalid");
    if (s.matrix != MMR_MATRIX_BT601 && s.matrix != MMR_MATRIX_BT709) return fail(MMR_STATUS_INVALID_ARGUMENT, "YUV->RGB matrix invalid");
    return MMR_STATUS_OK;
}

}  // namespace

namespace mmr_host {

mmr_recon_tuning default_tuning() {
    mmr_recon_tuning t;
    memset(&t, 0, sizeof t);
    t.struct_size = sizeof t;
    t.mjpeg_group_blocks = 32;
    t.mpeg2_group_blocks = 32;
    t.dv_group_blocks = 64;
    t.h264_deblock_band_rows = 8;
    t.h264_deblock_groups_per_cu = 2;
    t.h264_deblock_publish = 2;
    return t;
}

mmr_status check_tuning(const mmr_recon_tuning& t) {
    for (uint32_t b : {t.mjpeg_group_blocks, t.mpeg2_group_blocks, t.dv_group_blocks})
        if (b != 16 && b != 32 && b != 64) return fail(MMR_STATUS_INVALID_ARGUMENT, "blocks per workgroup must be 16, 32 or 64 (got %u)", b);
    if (!mmr::h264_deblock_rows_supported(t.h264_deblock_band_rows))
        return fail(MMR_STATUS_INVALID_ARGUMENT, "deblocking band rows must be 4, 8 or 16 (got %u)", t.h264_deblock_band_rows);
    if (t.h264_deblock_groups_per_cu < 1 || t.h264_deblock_groups_per_cu > 4)
        return fail(MMR_STATUS_INVALID_ARGUMENT, "deblocking groups per compute unit must be 1..4");
    if (t.h264_deblock_publish < 1 || t.h264_deblock_publish > 32)
        return fail(MMR_STATUS_INVALID_ARGUMENT, "deblocking publish interval must be 1..32");
    return MMR_STATUS_OK;
}

mmr_status mjpeg_submit_on(mmr_recon_context* ctx, hipStream_t stream, const mmr_recon_mjpeg_frame* frames, uint32_t count, uint32_t flags,
                                  uint64_t* out_ticket, PreparedCapture* cap = nullptr) {
    CHECK_CTX(ctx);
    if (!frames || count == 0) return fail(MMR_STATUS_INVALID_ARGUMENT, "no frames");
    std::vector<mmr::MjpegJob> jobs(count);
    std::vector<mmr::RgbJob> rgb;
    std::vector<uint32_t> tiles, rgb_tiles;
    const uint32_t bpg = ctx->tuning.mjpeg_group_blocks;
    for (uint32_t i = 0; i < count; ++i) {
        uint32_t quads = 0;
        mmr_status s = build_mjpeg(frames[i], jobs[i], quads);
        if (s) return s;
        jobs[i].tile_begin = add_tiles(tiles, i, jobs[i].block_count, bpg);
        if (quads) {
            mmr::RgbJob r = rgb_job_of(frames[i], jobs[i]);
            r.tile_begin = add_tiles(rgb_tiles, (uint32_t)rgb.size(), mmr::rgb_job_tiles(r), 1);
            rgb.push_back(r);
        }
    }
    std::vector<char> blob;
    const size_t jobs_off = blob_append(blob, jobs.data(), jobs.size() * sizeof(mmr::MjpegJob));
    const size_t rgb_off = blob_append(blob, rgb.data(), rgb.size() * sizeof(mmr::RgbJob));
    const size_t tiles_off = blob_append(blob, tiles.data(), tiles.size() * sizeof(uint32_t));
    const size_t rgb_tiles_off = blob_append(blob, rgb_tiles.data(), rgb_tiles.size() * sizeof(uint32_t));
    Submission sub(ctx, stream, cap);
    mmr_status s = sub.begin(blob.data(), blob.size(), 0);
    if (s) return s;
    const char* base = (const char*)sub.jobs_ptr();
    auto* dj = reinterpret_cast<const mmr::MjpegJob*>(base + jobs_off);
    auto* dr = reinterpret_cast<const mmr::RgbJob*>(base + rgb_off);
    auto* dt = reinterpret_cast<const uint32_t*>(base + tiles_off);
    auto* drt = reinterpret_cast<const uint32_t*>(base + rgb_tiles_off);
    hipError_t e = hipSuccess;
    if (flags & MMR_SUBMIT_PER_FRAME_LAUNCH) {
        size_t r = 0;
        for (uint32_t i = 0; i < count && e == hipSuccess; ++i) {
            e = mmr::launch_mjpeg(dj, dt, jobs[i].tile_begin, (jobs[i].block_count + bpg - 1) / bpg, bpg, stream);
            if (e == hipSuccess && frames[i].rgb) {
                e = mmr::launch_yuv_to_rgb(dr, drt, rgb[r].tile_begin, mmr::rgb_job_tiles(rgb[r]), stream);
                ++r;
            }
        }
    } else {
        e = mmr::launch_mjpeg(dj, dt, 0, (uint32_t)tiles.size(), bpg, stream);
        if (e == hipSuccess && !rgb.empty()) {
            if ((s = sub.mark())) return s;
            e = mmr::launch_yuv_to_rgb(dr, drt, 0, (uint32_t)rgb_tiles.size(), stream);
        }
    }
    if (e != hipSuccess) return hip_fail(e, "MJPEG launch");
    return sub.end(out_ticket);
}

mmr_status mpeg2_submit_on(mmr_recon_context* ctx, hipStream_t stream, const mmr_recon_mpeg2_picture* pics, uint32_t count, uint32_t flags,
                                  uint64_t* out_ticket, PreparedCapture* cap = nullptr, const Workspace* ws = nullptr) {
    CHECK_CTX(ctx);
    if (!pics || count == 0) return fail(MMR_STATUS_INVALID_ARGUMENT, "no pictures");
    const bool per_frame = (flags & MMR_SUBMIT_PER_FRAME_LAUNCH) != 0;
    const uint32_t bpg = ctx->tuning.mpeg2_group_blocks;
    std::vector<mmr::Mpeg2Job> jobs(count);
    for (uint32_t i = 0; i < count; ++i) {
        const mmr_recon_mpeg2_picture& p = pics[i];
        // descriptors that end before field_motion (built without field prediction) stay valid
        constexpr size_t kMpeg2Base = offsetof(mmr_recon_mpeg2_picture, field_motion);
        if (p.struct_size < kMpeg2Base) return fail(MMR_STATUS_INVALID_ARGUMENT, "MPEG-2 struct_size too small");
        const mmr_recon_mpeg2_field_motion* field_motion =
            p.struct_size >= kMpeg2Base + sizeof(p.field_motion) ? p.field_motion : nullptr;
        if (p.width == 0 || p.height == 0 || p.width > 16384 || p.height > 16384)
            return fail(MMR_STATUS_INVALID_ARGUMENT, "MPEG-2 picture %u size %ux%u is invalid", i, p.width, p.height);
        if (p.picture_type < MMR_MPEG2_PICTURE_I || p.picture_type > MMR_MPEG2_PICTURE_B)
            return fail(MMR_STATUS_INVALID_ARGUMENT, "MPEG-2 picture %u type %u is invalid", i, p.picture_type);
        if (p.alternate_scan > 1 || p.intra_dc_precision > 3)
            return fail(MMR_STATUS_INVALID_ARGUMENT, "MPEG-2 picture %u scan / dc precision invalid", i);
        if (!p.matrices || !p.macroblocks || !p.levels || !p.planes[0] || !p.planes[1] || !p.planes[2])
            return fail(MMR_STATUS_INVALID_ARGUMENT, "MPEG-2 picture %u has a NULL buffer", i);
        if (!aligned(p.levels, 16) || !aligned(p.macroblocks, 4) || !aligned(field_motion, 2))
            return fail(MMR_STATUS_INVALID_ARGUMENT, "MPEG-2 picture %u: levels must be 16-byte aligned, macroblocks 4-byte, "
                        "field_motion 2-byte", i);
        const bool needs_f = p.picture_type != MMR_MPEG2_PICTURE_I, needs_b = p.picture_type == MMR_MPEG2_PICTURE_B;
        if ((needs_f && (!p.forward[0] || !p.forward[1] || !p.forward[2])) || (needs_b && (!p.backward[0] || !p.backward[1] || !p.backward[2])))
            return fail(MMR_STATUS_INVALID_DATA, "missing %s MPEG-2 reference for picture %u", needs_b ? "forward/backward" : "forward", i);
        if (!per_frame && (p.forward_source >= (int32_t)i || p.backward_source >= (int32_t)i || p.forward_source < -1 || p.backward_source < -1))
            return fail(MMR_STATUS_INVALID_ARGUMENT, "MPEG-2 picture %u depends on a picture that is not earlier in decode order", i);
        mmr::Mpeg2Job& j = jobs[i];
        memset(&j, 0, sizeof j);
        j.matrices = p.matrices;
        j.mbs = p.macroblocks;
        j.levels = p.levels;
        j.field_motion = field_motion;
        for (int k = 0; k < 3; ++k) {
            mmr::Mpeg2Plane& pl = j.plane[k];
            pl.dst = p.planes[k];
            pl.fwd = p.forward[k] ? p.forward[k] : p.planes[k];
            pl.bwd = p.backward[k] ? p.backward[k] : (p.forward[k] ? p.forward[k] : p.planes[k]);
        }
        j.fwd_src = per_frame ? -1 : p.forward_source;
        j.bwd_src = per_frame ? -1 : p.backward_source;
        j.mb_w = (p.width + 15) / 16;
        j.mb_h = (p.height + 15) / 16;
        for (int k = 0; k < 3; ++k) {
            j.plane[k].stride = k ? j.mb_w * 8 : j.mb_w * 16;
            j.plane[k].height = k ? j.mb_h * 8 : j.mb_h * 16;
        }
        j.alt_scan = p.alternate_scan;
        j.dc_shift = 3 - p.intra_dc_precision;
        j.tile_count = (j.mb_w * j.mb_h * 6 + bpg - 1) / bpg;
        j.refs_aligned = 1;
        for (int k = 0; k < 3; ++k)
            if (!aligned(j.plane[k].fwd, 4) || !aligned(j.plane[k].bwd, 4)) j.refs_aligned = 0;
    }
    // Dependency levels: a picture runs one launch after the latest picture it references.
    std::vector<uint32_t> level(count, 0);
    uint32_t levels = 1;
    if (!per_frame) {
        for (uint32_t i = 0; i < count; ++i) {
            if (pics[i].forward_source >= 0) level[i] = std::max(level[i], level[pics[i].forward_source] + 1);
            if (pics[i].backward_source >= 0) level[i] = std::max(level[i], level[pics[i].backward_source] + 1);
            levels = std::max(levels, level[i] + 1);
        }
    }
    // Jobs grouped by level (stable); the tiles of one level are contiguous in the tile table.
    std::vector<mmr::Mpeg2Job> ordered;
    std::vector<uint32_t> tile_table;
    std::vector<uint32_t> level_tile_first(levels + 1, 0);
    if (per_frame) {
        ordered = jobs;
        for (uint32_t i = 0; i < count; ++i) ordered[i].tile_begin = add_tiles(tile_table, i, ordered[i].tile_count * bpg, bpg);
    } else {
        for (uint32_t lv = 0; lv < levels; ++lv) {
            level_tile_first[lv] = (uint32_t)tile_table.size();
            for (uint32_t i = 0; i < count; ++i)
                if (level[i] == lv) {
                    mmr::Mpeg2Job j = jobs[i];
                    j.tile_begin = add_tiles(tile_table, (uint32_t)ordered.size(), j.tile_count * bpg, bpg);
                    ordered.push_back(j);
                }
        }
        level_tile_first[levels] = (uint32_t)tile_table.size();
    }
    std::vector<char> blob;
    const size_t jobs_off = blob_append(blob, ordered.data(), ordered.size() * sizeof(mmr::Mpeg2Job));
    const size_t tiles_off = blob_append(blob, tile_table.data(), tile_table.size() * sizeof(uint32_t));
    // Scratch: each picture's block order (6 entries of 16 bytes per macroblock).
    std::vector<size_t> order_off(ordered.size());
    size_t scratch_total = 0;
    uint32_t order_rows = 0;
    for (size_t k = 0; k < ordered.size(); ++k) {
        order_off[k] = scratch_total;
        scratch_total += ((size_t)ordered[k].mb_w * ordered[k].mb_h * 6 * 16 + 255) & ~size_t(255);   // 16-byte entries
        ordered[k].row_begin = per_frame ? 0 : order_rows;
        order_rows += ordered[k].mb_h;
    }
    Submission sub(ctx, stream, cap, ws);
    mmr_status s = sub.begin(blob.data(), blob.size(), 0, scratch_total, [&](char* host, char* d_scratch) {
        auto* hj = reinterpret_cast<mmr::Mpeg2Job*>(host + jobs_off);
        for (size_t k = 0; k < ordered.size(); ++k) {
            hj[k].order = reinterpret_cast<uint32_t*>(d_scratch + order_off[k]);
            hj[k].row_begin = ordered[k].row_begin;
        }
    });
    if (s) return s;
    auto* dj = reinterpret_cast<const mmr::Mpeg2Job*>((const char*)sub.jobs_ptr() + jobs_off);
    auto* dt = reinterpret_cast<const uint32_t*>((const char*)sub.jobs_ptr() + tiles_off);
    hipError_t e = hipSuccess;
    if (per_frame) {
        for (uint32_t i = 0; i < count && e == hipSuccess; ++i) {
            e = mmr::launch_mpeg2_order(dj + i, 1, ordered[i].mb_h, sub.status_ptr(), stream);
            if (e == hipSuccess)
                e = mmr::launch_mpeg2(dj, dt, ordered[i].tile_begin, ordered[i].tile_count, bpg, sub.status_ptr(), stream);
        }
    } else {
        e = mmr::launch_mpeg2_order(dj, (uint32_t)ordered.size(), order_rows, sub.status_ptr(), stream);
        if (e == hipSuccess && (s = sub.mark())) return s;
        for (uint32_t lv = 0; lv < levels && e == hipSuccess; ++lv)
            e = mmr::launch_mpeg2(dj, dt, level_tile_first[lv], level_tile_first[lv + 1] - level_tile_first[lv], bpg,
                                  sub.status_ptr(), stream);
    }
    if (e != hipSuccess) return hip_fail(e, "MPEG-2 launch");
    return sub.end(out_ticket);
}

mmr_status dv_submit_on(mmr_recon_context* ctx, hipStream_t stream, const mmr_recon_dv_frame* frames, uint32_t count, uint32_t flags,
                               uint64_t* out_ticket, PreparedCapture* cap = nullptr) {
    CHECK_CTX(ctx);
    if (!frames || count == 0) return fail(MMR_STATUS_INVALID_ARGUMENT, "no frames");
    const bool per_frame = (flags & MMR_SUBMIT_PER_FRAME_LAUNCH) != 0;
    const uint32_t bpg = ctx->tuning.dv_group_blocks;
    std::vector<mmr::DvJob> jobs(count);
    std::vector<uint32_t> tiles;
    for (uint32_t i = 0; i < count; ++i) {
        const mmr_recon_dv_frame& f = frames[i];
        if (f.struct_size < sizeof(mmr_recon_dv_frame)) return fail(MMR_STATUS_INVALID_ARGUMENT, "DV struct_size too small");
        if (f.system != MMR_DV_SYSTEM_525_60 && f.system != MMR_DV_SYSTEM_625_50)
            return fail(MMR_STATUS_INVALID_ARGUMENT, "DV frame %u system %u is invalid", i, f.system);
        if (!f.levels || !f.block_info || !f.planes[0] || !f.planes[1] || !f.planes[2])
            return fail(MMR_STATUS_INVALID_ARGUMENT, "DV frame %u has a NULL buffer", i);
        if (!aligned(f.levels, 16)) return fail(MMR_STATUS_INVALID_ARGUMENT, "DV levels must be 16-byte aligned");
        mmr::DvJob& j = jobs[i];
        memset(&j, 0, sizeof j);
        j.levels = f.levels;
        j.info = f.block_info;
        for (int k = 0; k < 3; ++k) j.planes[k] = f.planes[k];
        j.system = f.system;
        j.block_count = (f.system == MMR_DV_SYSTEM_625_50 ? 12u : 10u) * 27u * 5u * 6u;
        j.tile_begin = add_tiles(tiles, i, j.block_count, bpg);
    }
    std::vector<char> blob;
    const size_t jobs_off = blob_append(blob, jobs.data(), jobs.size() * sizeof(mmr::DvJob));
    const size_t tiles_off = blob_append(blob, tiles.data(), tiles.size() * sizeof(uint32_t));
    Submission sub(ctx, stream, cap);
    mmr_status s = sub.begin(blob.data(), blob.size(), 0);
    if (s) return s;
    auto* dj = reinterpret_cast<const mmr::DvJob*>((const char*)sub.jobs_ptr() + jobs_off);
    auto* dt = reinterpret_cast<const uint32_t*>((const char*)sub.jobs_ptr() + tiles_off);
    hipError_t e = hipSuccess;
    if (per_frame) {
        for (uint32_t i = 0; i < count && e == hipSuccess; ++i)
            e = mmr::launch_dv(dj, dt, jobs[i].tile_begin, (jobs[i].block_count + bpg - 1) / bpg, bpg, stream);
    } else {
        e = mmr::launch_dv(dj, dt, 0, (uint32_t)tiles.size(), bpg, stream);
    }
    if (e != hipSuccess) return hip_fail(e, "DV launch");
    return sub.end(out_ticket);
}

mmr_status h264_submit_on(mmr_recon_context* ctx, hipStream_t stream, const mmr_recon_h264_picture* pics, uint32_t count, uint32_t flags,
                                 uint64_t* out_ticket, PreparedCapture* cap = nullptr) {
    CHECK_CTX(ctx);
    if (!pics || count == 0) return fail(MMR_STATUS_INVALID_ARGUMENT, "no pictures");
    const bool per_frame = (flags & MMR_SUBMIT_PER_FRAME_LAUNCH) != 0;
    std::vector<mmr::H264Job> jobs(count);
    std::vector<uint32_t> tiles;
    uint32_t bands = 0, intra_bands = 0;
    const uint32_t intra_rows_per_band = mmr::h264_intra_band_rows();
    const uint32_t band_rows = ctx->tuning.h264_deblock_band_rows, publish = ctx->tuning.h264_deblock_publish;
    // Jobs are built longest wavefront first (width + 2 x height macroblocks, descending; ties in
    // submission order): deblocking and intra band tickets follow job order, so the pictures whose
    // dependency chains set the launch's length get workgroups first. The pictures of a batch are
    // independent, so their order changes no result. Job k is picture order[k].
    std::vector<uint32_t> order(count);
    for (uint32_t i = 0; i < count; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return pics[a].width_mbs + 2 * pics[a].height_mbs > pics[b].width_mbs + 2 * pics[b].height_mbs;
    });
    for (uint32_t k = 0; k < count; ++k) {
        const uint32_t i = order[k];
        const mmr_recon_h264_picture& p = pics[i];
        // descriptors that end before intra (built without intra records) stay valid
        constexpr size_t kH264Base = offsetof(mmr_recon_h264_picture, intra);
        if (p.struct_size < kH264Base) return fail(MMR_STATUS_INVALID_ARGUMENT, "H.264 struct_size too small");
        const mmr_recon_h264_intra* intra = p.struct_size >= kH264Base + sizeof(p.intra) ? p.intra : nullptr;
        if (!aligned(intra, 4)) return fail(MMR_STATUS_INVALID_ARGUMENT, "H.264 picture %u: intra records must be 4-byte aligned", i);
        if (p.width_mbs == 0 || p.height_mbs == 0 || p.width_mbs > 1024 || p.height_mbs > 1024)
            return fail(MMR_STATUS_INVALID_ARGUMENT, "H.264 picture %u size invalid", i);
        if (!p.macroblocks || !p.blocks || !p.levels || !p.planes[0] || !p.planes[1] || !p.planes[2])
            return fail(MMR_STATUS_INVALID_ARGUMENT, "H.264 picture %u has a NULL buffer", i);
        if (p.ref_count[0] > MMR_H264_MAX_REFS || p.ref_count[1] > MMR_H264_MAX_REFS)
            return fail(MMR_STATUS_INVALID_DATA, "H.264 picture %u reference counts invalid", i);
        // An all-intra picture (I or IDR) has no references. Every reference slot then points at the
        // picture's own planes, an in-bounds stand-in that only an inter block of invalid data reads
        // (and the device reports such a block as invalid data).
        const bool no_refs = p.ref_count[0] + p.ref_count[1] == 0;
        if (p.weighted_mode > MMR_H264_WEIGHTED_TABLE) return fail(MMR_STATUS_INVALID_ARGUMENT, "H.264 weighted_mode invalid");
        if (p.weighted_mode == MMR_H264_WEIGHTED_TABLE && (!p.weights || !p.bi_weights))
            return fail(MMR_STATUS_INVALID_ARGUMENT, "H.264 picture %u: weighted mode needs weight tables", i);
        if (!aligned(p.planes[0], 4) || !aligned(p.planes[1], 4) || !aligned(p.planes[2], 4))
            return fail(MMR_STATUS_INVALID_ARGUMENT, "H.264 planes must be 4-byte aligned");
        mmr::H264Job& j = jobs[k];
        memset(&j, 0, sizeof j);
        j.mbs = p.macroblocks;
        j.blocks = p.blocks;
        j.levels = p.levels;
        j.weights = p.weights;
        j.bi_weights = p.bi_weights;
        j.scaling = p.scaling_lists ? p.scaling_lists : ctx->flat_scaling;
        for (int k = 0; k < 3; ++k) j.planes[k] = p.planes[k];
        for (int list = 0; list < 2; ++list) {
            j.ref_count[list] = p.ref_count[list];
            for (uint32_t r = 0; r < MMR_H264_MAX_REFS; ++r) {
                const uint32_t src = r < p.ref_count[list] ? r : 0;
                const mmr_recon_h264_reference& ref = p.refs[list][src];
                for (int k = 0; k < 3; ++k) {
                    const uint8_t* plane = no_refs ? p.planes[k] : p.ref_count[list] ? ref.planes[k] : p.refs[1 - list][0].planes[k];
                    if (!plane) return fail(MMR_STATUS_INVALID_ARGUMENT, "H.264 picture %u: reference plane is NULL", i);
                    j.ref_planes[list][r][k] = plane;
                }
                j.ref_poc[list][r] = p.ref_count[list] ? ref.poc : 0;
            }
        }
        j.width_mbs = p.width_mbs;
        j.height_mbs = p.height_mbs;
        j.weighted = p.weighted_mode;
        j.deblock = p.deblock ? 1u : 0u;
        j.chroma_qp_offset[0] = p.chroma_qp_offset[0];
        j.chroma_qp_offset[1] = p.chroma_qp_offset[1];
        j.tile_begin = add_tiles(tiles, k, (p.width_mbs * p.height_mbs) * 8);   // 4 macroblocks per 32-entry tile unit
        j.refs_aligned = 1;
        for (int list = 0; list < 2; ++list)
            for (uint32_t r = 0; r < MMR_H264_MAX_REFS; ++r)
                for (int k = 0; k < 3; ++k)
                    if (!aligned(j.ref_planes[list][r][k], 4)) j.refs_aligned = 0;
        j.band_begin = per_frame ? 0 : bands;
        bands += j.deblock ? (p.height_mbs + band_rows - 1) / band_rows : 0;
        j.intra = intra;
        constexpr size_t kBypassEnd = offsetof(mmr_recon_h264_picture, transform_bypass) + sizeof(uint32_t);
        j.transform_bypass = p.struct_size >= kBypassEnd && p.transform_bypass ? 1u : 0u;
        j.intra_band_begin = per_frame ? 0 : intra_bands;
        intra_bands += intra ? (p.height_mbs + intra_rows_per_band - 1) / intra_rows_per_band : 0;
    }
    Submission sub(ctx, stream, cap);
    // Counters: deblocking band tickets and hand-over progress, then intra band tickets and
    // hand-over progress (per picture with one launch per picture).
    auto pic_intra_bands = [&](uint32_t i) { return (pics[order[i]].height_mbs + intra_rows_per_band - 1) / intra_rows_per_band; };
    size_t counters = 0;
    if (per_frame) {
        for (uint32_t i = 0; i < count; ++i)
            counters += 1 + (jobs[i].deblock ? pics[order[i]].height_mbs : 0) + (jobs[i].intra ? 1 + pic_intra_bands(i) : 0);
    } else {
        counters = 1 + bands + (intra_bands ? 1 + intra_bands : 0);
    }
    // Scratch: 16 bytes of boundary strengths per macroblock of every deblocked picture.
    std::vector<size_t> bs_off(count, 0);
    size_t scratch_total = 0;
    for (uint32_t i = 0; i < count; ++i) {
        bs_off[i] = scratch_total;
        if (jobs[i].deblock) scratch_total += ((size_t)pics[order[i]].width_mbs * pics[order[i]].height_mbs * 16 + 255) & ~size_t(255);
    }
    std::vector<char> blob;
    const size_t jobs_off = blob_append(blob, jobs.data(), jobs.size() * sizeof(mmr::H264Job));
    const size_t tiles_off = blob_append(blob, tiles.data(), tiles.size() * sizeof(uint32_t));
    mmr_status s = sub.begin(blob.data(), blob.size(), counters, scratch_total, [&](char* host, char* d_scratch) {
        auto* hj = reinterpret_cast<mmr::H264Job*>(host + jobs_off);
        for (uint32_t i = 0; i < count; ++i) hj[i].bs = jobs[i].deblock ? reinterpret_cast<uint32_t*>(d_scratch + bs_off[i]) : nullptr;
    });
    if (s) return s;
    auto* dj = reinterpret_cast<const mmr::H264Job*>((const char*)sub.jobs_ptr() + jobs_off);
    auto* dt = reinterpret_cast<const uint32_t*>((const char*)sub.jobs_ptr() + tiles_off);
    hipError_t e = hipSuccess;
    const uint32_t groups = ctx->persistent_groups ? ctx->persistent_groups : ctx->cu_count * ctx->tuning.h264_deblock_groups_per_cu;
    const uint32_t intra_groups = ctx->cu_count;   // one band (8 waves) per CU
    if (per_frame) {
        uint32_t* c = sub.counters_ptr();
        for (uint32_t i = 0; i < count && e == hipSuccess; ++i) {
            const uint32_t mc_tiles = (pics[order[i]].width_mbs * pics[order[i]].height_mbs + 3) / 4;
            if (jobs[i].deblock) e = mmr::launch_h264_bs(dj, dt, jobs[i].tile_begin, mc_tiles, stream);
            if (e == hipSuccess) e = mmr::launch_h264_mc(dj, dt, jobs[i].tile_begin, mc_tiles, sub.status_ptr(), stream);
            if (e == hipSuccess && jobs[i].intra) {
                e = mmr::launch_h264_intra(dj + i, 1, pic_intra_bands(i), c, intra_groups, publish, sub.status_ptr(), stream);
                c += 1 + pic_intra_bands(i);
            }
            if (e == hipSuccess && jobs[i].deblock) {
                const uint32_t nb = (pics[order[i]].height_mbs + band_rows - 1) / band_rows;
                e = mmr::launch_h264_deblock(dj + i, 1, nb, c, groups, band_rows, publish, sub.status_ptr(), stream);
                c += 1 + pics[order[i]].height_mbs;
            } else {
                c += 1;
            }
        }
    } else {
        if (bands) {
            e = mmr::launch_h264_bs(dj, dt, 0, (uint32_t)tiles.size(), stream);
            if (e == hipSuccess && (s = sub.mark())) return s;
        }
        if (e == hipSuccess) e = mmr::launch_h264_mc(dj, dt, 0, (uint32_t)tiles.size(), sub.status_ptr(), stream);
        if (e == hipSuccess && intra_bands) {
            if ((s = sub.mark())) return s;
            e = mmr::launch_h264_intra(dj, count, intra_bands, sub.counters_ptr() + 1 + bands, intra_groups, publish,
                                       sub.status_ptr(), stream);
        }
        if (e == hipSuccess && bands) {
            if ((s = sub.mark())) return s;
            e = mmr::launch_h264_deblock(dj, count, bands, sub.counters_ptr(), groups, band_rows, publish, sub.status_ptr(), stream);
        }
    }
    if (e != hipSuccess) return hip_fail(e, "H.264 launch");
    return sub.end(out_ticket);
}

// The device job of a checked conversion job (its tile range is the caller's).
mmr_status build_rgb_job(const mmr_recon_yuv_to_rgb_job& in, mmr::RgbJob& r) {
    mmr_status s = check_rgb_job(in);
    if (s) return s;
    memset(&r, 0, sizeof r);
    for (int k = 0; k < 3; ++k) r.planes[k] = in.planes[k];
    r.rgb = in.rgb;
    r.rgba = in.rgba;
    r.width = in.width;
    r.height = in.height;
    r.cw = in.chroma_width;
    r.ch = in.chroma_height;
    r.luma_stride = in.luma_stride;
    r.chroma_stride = in.chroma_stride;
    r.range = in.range;
    r.matrix = in.matrix;
    r.inv_width = 1.0 / in.width;
    r.inv_height = 1.0 / in.height;
    mmr::set_rgb_layout(r);
    mmr::set_rgb_terms(r);
    return MMR_STATUS_OK;
}

mmr_status yuv_to_rgb_submit_on(mmr_recon_context* ctx, hipStream_t stream, const mmr_recon_yuv_to_rgb_job* in, uint32_t count, uint32_t flags,
                                       uint64_t* out_ticket) {
    CHECK_CTX(ctx);
    if (!in || count == 0) return fail(MMR_STATUS_INVALID_ARGUMENT, "no jobs");
    const bool per_frame = (flags & MMR_SUBMIT_PER_FRAME_LAUNCH) != 0;
    std::vector<mmr::RgbJob> jobs(count);
    std::vector<uint32_t> tiles;
    for (uint32_t i = 0; i < count; ++i) {
        mmr_status s = build_rgb_job(in[i], jobs[i]);
        if (s) return s;
        mmr::RgbJob& r = jobs[i];
        r.tile_begin = add_tiles(tiles, i, mmr::rgb_job_tiles(r), 1);
    }
    std::vector<char> blob;
    const size_t jobs_off = blob_append(blob, jobs.data(), jobs.size() * sizeof(mmr::RgbJob));
    const size_t tiles_off = blob_append(blob, tiles.data(), tiles.size() * sizeof(uint32_t));
    Submission sub(ctx, stream);
    mmr_status s = sub.begin(blob.data(), blob.size(), 0);
    if (s) return s;
    auto* dj = reinterpret_cast<const mmr::RgbJob*>((const char*)sub.jobs_ptr() + jobs_off);
    auto* dt = reinterpret_cast<const uint32_t*>((const char*)sub.jobs_ptr() + tiles_off);
    hipError_t e = hipSuccess;
    if (per_frame) {
        for (uint32_t i = 0; i < count && e == hipSuccess; ++i)
            e = mmr::launch_yuv_to_rgb(dj, dt, jobs[i].tile_begin, mmr::rgb_job_tiles(jobs[i]), stream);
    } else {
        e = mmr::launch_yuv_to_rgb(dj, dt, 0, (uint32_t)tiles.size(), stream);
    }
    if (e != hipSuccess) return hip_fail(e, "YUV->RGB launch");
    return sub.end(out_ticket);
}

// The device job of composite job c (index i for messages), checked; its tile range is the caller's.
mmr_status build_composite_job(const mmr_recon_composite_job& c, uint32_t i, mmr::CompositeJobDev& j) {
    if (c.struct_size < sizeof(mmr_recon_composite_job)) return fail(MMR_STATUS_INVALID_ARGUMENT, "composite struct_size too small");
    if (!c.target || !c.width || !c.height || c.width > 65535 || c.height > 65535 || c.stride < c.width * 4 ||
        c.layer_count > MMR_RECON_MAX_LAYERS)
        return fail(MMR_STATUS_INVALID_ARGUMENT, "composite job %u is invalid", i);
    if (!aligned(c.target, 4) || c.stride % 4) return fail(MMR_STATUS_INVALID_ARGUMENT, "composite target must be 4-byte aligned");
    for (uint32_t l = 0; l < c.layer_count; ++l) {
        const mmr_recon_layer& L = c.layers[l];
        if (!L.rgba || L.stride < L.width * 4 || L.stride % 4 || !aligned(L.rgba, 4))
            return fail(MMR_STATUS_INVALID_ARGUMENT, "composite job %u layer %u storage is invalid", i, l);
        if ((uint64_t)L.x + L.width > c.width || (uint64_t)L.y + L.height > c.height)
            return fail(MMR_STATUS_INVALID_DATA, "RGBA overlay storage or placement is invalid (job %u layer %u)", i, l);
        if (L.op != MMR_BLEND_SOURCE_OVER && L.op != MMR_BLEND_PREMULTIPLIED_OVER)
            return fail(MMR_STATUS_INVALID_ARGUMENT, "composite operator %u invalid", L.op);
        if (L.opacity > 255) return fail(MMR_STATUS_INVALID_ARGUMENT, "opacity must be 0..255");
    }
    memset(&j, 0, sizeof j);
    j.target = c.target;
    j.width = c.width;
    j.height = c.height;
    j.stride = c.stride;
    j.layer_count = c.layer_count;
    for (uint32_t l = 0; l < c.layer_count; ++l) j.layers[l] = c.layers[l];
    j.waves_per_row = (c.width + 255) / 256;
    return MMR_STATUS_OK;
}

mmr_status composite_submit_on(mmr_recon_context* ctx, hipStream_t stream, const mmr_recon_composite_job* in, uint32_t count, uint32_t flags,
                                      uint64_t* out_ticket) {
    CHECK_CTX(ctx);
    if (!in || count == 0) return fail(MMR_STATUS_INVALID_ARGUMENT, "no jobs");
    const bool per_frame = (flags & MMR_SUBMIT_PER_FRAME_LAUNCH) != 0;
    std::vector<mmr::CompositeJobDev> jobs(count);
    std::vector<uint32_t> tiles;
    for (uint32_t i = 0; i < count; ++i) {
        mmr_status s = build_composite_job(in[i], i, jobs[i]);
        if (s) return s;
        mmr::CompositeJobDev& j = jobs[i];
        j.tile_begin = add_tiles(tiles, i, mmr::composite_job_tiles(j.waves_per_row, j.height), 1);
    }
    std::vector<char> blob;
    const size_t jobs_off = blob_append(blob, jobs.data(), jobs.size() * sizeof(mmr::CompositeJobDev));
    const size_t tiles_off = blob_append(blob, tiles.data(), tiles.size() * sizeof(uint32_t));
    Submission sub(ctx, stream);
    mmr_status s = sub.begin(blob.data(), blob.size(), 0);
    if (s) return s;
    auto* dj = reinterpret_cast<const mmr::CompositeJobDev*>((const char*)sub.jobs_ptr() + jobs_off);
    auto* dt = reinterpret_cast<const uint32_t*>((const char*)sub.jobs_ptr() + tiles_off);
    hipError_t e = hipSuccess;
    if (per_frame) {
        for (uint32_t i = 0; i < count && e == hipSuccess; ++i)
            e = mmr::launch_composite(dj, dt, jobs[i].tile_begin, mmr::composite_job_tiles(jobs[i].waves_per_row, jobs[i].height), stream);
    } else {
        e = mmr::launch_composite(dj, dt, 0, (uint32_t)tiles.size(), stream);
    }
    if (e != hipSuccess) return hip_fail(e, "composite launch");
    return sub.end(out_ticket);
}

// Conversion conv[i] followed by composite comp[i] in one pass: comp[i]'s target must be conv[i]'s
// RGBA output (rows packed: stride 4 x width, the same size, no RGB output).
mmr_status yuv_composite_submit_on(mmr_recon_context* ctx, hipStream_t stream, const mmr_recon_yuv_to_rgb_job* conv,
                                   const mmr_recon_composite_job* comp, uint32_t count, uint32_t flags, uint64_t* out_ticket) {
    CHECK_CTX(ctx);
    if (!conv || !comp || count == 0) return fail(MMR_STATUS_INVALID_ARGUMENT, "no jobs");
    const bool per_frame = (flags & MMR_SUBMIT_PER_FRAME_LAUNCH) != 0;
    std::vector<mmr::RgbJob> rjobs(count);
    std::vector<mmr::CompositeJobDev> cjobs(count);
    std::vector<uint32_t> tiles;
    for (uint32_t i = 0; i < count; ++i) {
        mmr_status s = build_rgb_job(conv[i], rjobs[i]);
        if (s) return s;
        if ((s = build_composite_job(comp[i], i, cjobs[i]))) return s;
        if (conv[i].rgb || conv[i].rgba != comp[i].target || conv[i].width != comp[i].width || conv[i].height != comp[i].height ||
            comp[i].stride != comp[i].width * 4)
            return fail(MMR_STATUS_INVALID_ARGUMENT,
                        "conversion + composite job %u: the composite's target must be the conversion's packed RGBA output", i);
        cjobs[i].tile_begin = add_tiles(tiles, i, mmr::composite_job_tiles(cjobs[i].waves_per_row, cjobs[i].height), 1);
    }
    std::vector<char> blob;
    const size_t rjobs_off = blob_append(blob, rjobs.data(), rjobs.size() * sizeof(mmr::RgbJob));
    const size_t cjobs_off = blob_append(blob, cjobs.data(), cjobs.size() * sizeof(mmr::CompositeJobDev));
    const size_t tiles_off = blob_append(blob, tiles.data(), tiles.size() * sizeof(uint32_t));
    Submission sub(ctx, stream);
    mmr_status s = sub.begin(blob.data(), blob.size(), 0);
    if (s) return s;
    auto* drj = reinterpret_cast<const mmr::RgbJob*>((const char*)sub.jobs_ptr() + rjobs_off);
    auto* dcj = reinterpret_cast<const mmr::CompositeJobDev*>((const char*)sub.jobs_ptr() + cjobs_off);
    auto* dt = reinterpret_cast<const uint32_t*>((const char*)sub.jobs_ptr() + tiles_off);
    hipError_t e = hipSuccess;
    if (per_frame) {
        for (uint32_t i = 0; i < count && e == hipSuccess; ++i)
            e = mmr::launch_yuv_composite(drj, dcj, dt, cjobs[i].tile_begin, mmr::composite_job_tiles(cjobs[i].waves_per_row, cjobs[i].height),
                                          stream);
    } else {
        e = mmr::launch_yuv_composite(drj, dcj, dt, 0, (uint32_t)tiles.size(), stream);
    }
    if (e != hipSuccess) return hip_fail(e, "conversion + composite launch");
    return sub.end(out_ticket);
}

// srgb_to_linear_u16 / linear_u16_to_srgb (crates/mmfx/src/render.rs:1856-1880) as device tables,
// built on the host with the C library's pow (what Rust's f64::powf calls on Linux) on first use.
mmr_status ensure_linear_tables(mmr_recon_context* ctx) {
    if (ctx->linear_tables) return MMR_STATUS_OK;
    std::vector<uint8_t> host(mmr::kLinearTableBytes);
    uint16_t* to_linear = reinterpret_cast<uint16_t*>(host.data());
    for (int v = 0; v < 256; ++v) {
        const double encoded = (double)v / 255.0;
        const double linear = encoded <= 0.04045 ? encoded / 12.92 : pow((encoded + 0.055) / 1.055, 2.4);
        to_linear[v] = (uint16_t)round(linear * 65535.0);
    }
    uint8_t* to_srgb = host.data() + 512;
    for (uint32_t v = 0; v < 65536; ++v) {
        const double linear = (double)v / 65535.0;
        const double encoded = linear <= 0.0031308 ? linear * 12.92 : 1.055 * pow(linear, 1.0 / 2.4) - 0.055;
        const double r = round(encoded * 255.0);
        to_srgb[v] = (uint8_t)(r < 0.0 ? 0.0 : (r > 255.0 ? 255.0 : r));
    }
    uint8_t* d = nullptr;
    hipError_t e = hipMalloc((void**)&d, host.size());
    if (e != hipSuccess) return hip_fail(e, "hipMalloc (linear-light tables)");
    e = hipMemcpy(d, host.data(), host.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return hip_fail(e, "linear-light table upload");
    }
    ctx->linear_tables = d;
    return MMR_STATUS_OK;
}

mmr_status linear_composite_submit_on(mmr_recon_context* ctx, hipStream_t stream, const mmr_recon_linear_composite_job* in,
                                      uint32_t count, uint32_t flags, uint64_t* out_ticket) {
    CHECK_CTX(ctx);
    if (!in || count == 0) return fail(MMR_STATUS_INVALID_ARGUMENT, "no jobs");
    const bool per_frame = (flags & MMR_SUBMIT_PER_FRAME_LAUNCH) != 0;
    std::vector<mmr::LinearCompositeDev> jobs(count);
    std::vector<uint32_t> tiles;
    for (uint32_t i = 0; i < count; ++i) {
        const mmr_recon_linear_composite_job& c = in[i];
        if (c.struct_size < sizeof(mmr_recon_linear_composite_job))
            return fail(MMR_STATUS_INVALID_ARGUMENT, "linear composite struct_size too small");
        if (!c.target || !c.width || !c.height || c.width > 65535 || c.height > 65535 || c.stride < c.width ||
            c.layer_count > MMR_RECON_MAX_LAYERS)
            return fail(MMR_STATUS_INVALID_ARGUMENT, "linear composite job %u is invalid", i);
        if (!aligned(c.target, 8)) return fail(MMR_STATUS_INVALID_ARGUMENT, "linear composite target must be 8-byte aligned");
        for (uint32_t l = 0; l < c.layer_count; ++l) {
            const mmr_recon_linear_layer& L = c.layers[l];
            if (L.format != MMR_LINEAR_SOURCE_SURFACE && L.format != MMR_LINEAR_SOURCE_SRGBA8)
                return fail(MMR_STATUS_INVALID_ARGUMENT, "linear composite job %u layer %u: format %u invalid", i, l, L.format);
            if (!L.pixels || L.width > 65535 || L.height > 65535 || L.stride < L.width ||
                !aligned(L.pixels, L.format == MMR_LINEAR_SOURCE_SURFACE ? 8 : 4))
                return fail(MMR_STATUS_INVALID_ARGUMENT, "linear composite job %u layer %u storage is invalid", i, l);
            if (L.opacity > 65535) return fail(MMR_STATUS_INVALID_ARGUMENT, "linear opacity must be 0..65535");
        }
        mmr::LinearCompositeDev& j = jobs[i];
        memset(&j, 0, sizeof j);
        j.target = c.target;
        j.width = c.width;
        j.height = c.height;
        j.stride = c.stride;
        j.layer_count = c.layer_count;
        for (uint32_t l = 0; l < c.layer_count; ++l) j.layers[l] = c.layers[l];
        j.waves_per_row = (c.width + 63) / 64;
        j.tile_begin = add_tiles(tiles, i, mmr::linear_job_tiles(j.waves_per_row, j.height), 1);
    }
    mmr_status s = ensure_linear_tables(ctx);
    if (s) return s;
    std::vector<char> blob;
    const size_t jobs_off = blob_append(blob, jobs.data(), jobs.size() * sizeof(mmr::LinearCompositeDev));
    const size_t tiles_off = blob_append(blob, tiles.data(), tiles.size() * sizeof(uint32_t));
    Submission sub(ctx, stream);
    s = sub.begin(blob.data(), blob.size(), 0);
    if (s) return s;
    auto* dj = reinterpret_cast<const mmr::LinearCompositeDev*>((const char*)sub.jobs_ptr() + jobs_off);
    auto* dt = reinterpret_cast<const uint32_t*>((const char*)sub.jobs_ptr() + tiles_off);
    const uint16_t* to_linear = reinterpret_cast<const uint16_t*>(ctx->linear_tables);
    hipError_t e = hipSuccess;
    if (per_frame) {
        for (uint32_t i = 0; i < count && e == hipSuccess; ++i)
            e = mmr::launch_linear_composite(dj, dt, jobs[i].tile_begin, mmr::linear_job_tiles(jobs[i].waves_per_row, jobs[i].height),
                                             to_linear, stream);
    } else {
        e = mmr::launch_linear_composite(dj, dt, 0, (uint32_t)tiles.size(), to_linear, stream);
    }
    if (e != hipSuccess) return hip_fail(e, "linear composite launch");
    return sub.end(out_ticket);
}

mmr_status linear_convert_submit_on(mmr_recon_context* ctx, hipStream_t stream, const mmr_recon_linear_convert_job* in,
                                    uint32_t count, uint32_t flags, uint64_t* out_ticket) {
    CHECK_CTX(ctx);
    if (!in || count == 0) return fail(MMR_STATUS_INVALID_ARGUMENT, "no jobs");
    const bool per_frame = (flags & MMR_SUBMIT_PER_FRAME_LAUNCH) != 0;
    std::vector<mmr::LinearConvertDev> jobs(count);
    std::vector<uint32_t> tiles;
    for (uint32_t i = 0; i < count; ++i) {
        const mmr_recon_linear_convert_job& c = in[i];
        if (c.struct_size < sizeof(mmr_recon_linear_convert_job))
            return fail(MMR_STATUS_INVALID_ARGUMENT, "linear convert struct_size too small");
        if (c.direction != MMR_LINEAR_TO_SRGBA8 && c.direction != MMR_LINEAR_FROM_SRGBA8)
            return fail(MMR_STATUS_INVALID_ARGUMENT, "linear convert job %u: direction %u invalid", i, c.direction);
        if (!c.linear || !c.rgba || !c.width || !c.height || c.width > 65535 || c.height > 65535 || c.linear_stride < c.width ||
            c.rgba_stride < (uint64_t)c.width * 4 || c.rgba_stride % 4 || !aligned(c.linear, 8) || !aligned(c.rgba, 4))
            return fail(MMR_STATUS_INVALID_ARGUMENT, "linear convert job %u storage is invalid", i);
        mmr::LinearConvertDev& j = jobs[i];
        memset(&j, 0, sizeof j);
        j.linear = c.linear;
        j.rgba = c.rgba;
        j.direction = c.direction;
        j.width = c.width;
        j.height = c.height;
        j.linear_stride = c.linear_stride;
        j.rgba_stride = c.rgba_stride;
        j.waves_per_row = (c.width + 63) / 64;
        j.tile_begin = add_tiles(tiles, i, mmr::linear_job_tiles(j.waves_per_row, j.height), 1);
    }
    mmr_status s = ensure_linear_tables(ctx);
    if (s) return s;
    std::vector<char> blob;
    const size_t jobs_off = blob_append(blob, jobs.data(), jobs.size() * sizeof(mmr::LinearConvertDev));
    const size_t tiles_off = blob_append(blob, tiles.data(), tiles.size() * sizeof(uint32_t));
    Submission sub(ctx, stream);
    s = sub.begin(blob.data(), blob.size(), 0);
    if (s) return s;
    auto* dj = reinterpret_cast<const mmr::LinearConvertDev*>((const char*)sub.jobs_ptr() + jobs_off);
    auto* dt = reinterpret_cast<const uint32_t*>((const char*)sub.jobs_ptr() + tiles_off);
    hipError_t e = hipSuccess;
    if (per_frame) {
        for (uint32_t i = 0; i < count && e == hipSuccess; ++i)
            e = mmr::launch_linear_convert(dj, dt, jobs[i].tile_begin, mmr::linear_job_tiles(jobs[i].waves_per_row, jobs[i].height),
                                           ctx->linear_tables, stream);
    } else {
        e = mmr::launch_linear_convert(dj, dt, 0, (uint32_t)tiles.size(), ctx->linear_tables, stream);
    }
    if (e != hipSuccess) return hip_fail(e, "linear convert launch");
    return sub.end(out_ticket);
}

mmr_status yuv_blend_submit_on(mmr_recon_context* ctx, hipStream_t stream, const mmr_recon_yuv_blend_job* in, uint32_t count, uint32_t flags,
                                      uint64_t* out_ticket) {
    CHECK_CTX(ctx);
    if (!in || count == 0) return fail(MMR_STATUS_INVALID_ARGUMENT, "no jobs");
    uint32_t max_px = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const mmr_recon_yuv_blend_job& j = in[i];
        if (j.struct_size < sizeof(mmr_recon_yuv_blend_job)) return fail(MMR_STATUS_INVALID_ARGUMENT, "YUV blend struct_size too small");
        if ((j.width | j.height | j.x | j.y) & 1 || (uint64_t)j.x + j.width > j.target_width || (uint64_t)j.y + j.height > j.target_height)
            return fail(MMR_STATUS_INVALID_DATA, "YUV overlay storage or placement is invalid (job %u)", i);
        for (int k = 0; k < 3; ++k)
            if (!j.target[k] || !j.overlay[k]) return fail(MMR_STATUS_INVALID_ARGUMENT, "YUV blend job %u has a NULL plane", i);
        if (!j.rgba || !j.chroma_alpha) return fail(MMR_STATUS_INVALID_ARGUMENT, "YUV blend job %u has no alpha", i);
        max_px = std::max(max_px, j.width * j.height);
    }
    Submission sub(ctx, stream);
    mmr_status s = sub.begin(in, sizeof(mmr_recon_yuv_blend_job) * count, 0);
    if (s) return s;
    hipError_t e = hipSuccess;
    auto* dj = reinterpret_cast<const mmr_recon_yuv_blend_job*>(sub.jobs_ptr());
    if (flags & MMR_SUBMIT_PER_FRAME_LAUNCH) {
        for (uint32_t i = 0; i < count && e == hipSuccess; ++i) e = mmr::launch_yuv_blend(dj + i, 1, in[i].width * in[i].height, stream);
    } else {
        e = mmr::launch_yuv_blend(dj, count, max_px, stream);
    }
    if (e != hipSuccess) return hip_fail(e, "YUV blend launch");
    return sub.end(out_ticket);
}

mmr_status overlay_prepare_submit_on(mmr_recon_context* ctx, hipStream_t stream, const mmr_recon_overlay_prepare_job* in, uint32_t count,
                                     uint32_t flags, uint64_t* out_ticket) {
    CHECK_CTX(ctx);
    if (!in || count == 0) return fail(MMR_STATUS_INVALID_ARGUMENT, "no jobs");
    uint32_t max_cells = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const mmr_recon_overlay_prepare_job& j = in[i];
        if (j.struct_size < sizeof(mmr_recon_overlay_prepare_job)) return fail(MMR_STATUS_INVALID_ARGUMENT, "overlay struct_size too small");
        if (!j.rgba || j.width == 0 || j.height == 0 || j.width > 16384 || j.height > 16384)
            return fail(MMR_STATUS_INVALID_ARGUMENT, "overlay job %u: no pixels or size out of range", i);
        if (!aligned(j.rgba, 4)) return fail(MMR_STATUS_INVALID_ARGUMENT, "overlay job %u: rgba must be 4-byte aligned", i);
        max_cells = std::max(max_cells, ((j.width + 1) / 2) * ((j.height + 1) / 2));
    }
    Submission sub(ctx, stream);
    mmr_status s = sub.begin(in, sizeof(mmr_recon_overlay_prepare_job) * count, 0);
    if (s) return s;
    auto* dj = reinterpret_cast<const mmr_recon_overlay_prepare_job*>(sub.jobs_ptr());
    hipError_t e = hipSuccess;
    if (flags & MMR_SUBMIT_PER_FRAME_LAUNCH) {
        for (uint32_t i = 0; i < count && e == hipSuccess; ++i)
            e = mmr::launch_overlay_prepare(dj + i, 1, ((in[i].width + 1) / 2) * ((in[i].height + 1) / 2), stream);
    } else {
        e = mmr::launch_overlay_prepare(dj, count, max_cells, stream);
    }
    if (e != hipSuccess) return hip_fail(e, "overlay prepare launch");
    return sub.end(out_ticket);
}

// Taps of output coordinate o for a pass from n to m samples: image::imageops::resize (the image
// crate 0.25.6, called by crates/render/src/compositor.rs:117-122) -- centre (o + 0.5) n/m, support
// s max(n/m, 1) with s = 3 for Lanczos3 and 1 for Triangle, taps clamped to the source, weights
// kernel((i - centre + 0.5) / max(n/m, 1)) normalised by their f32 sum, all in f32 in this order.
static float lanczos_sinc(float t) {
    const float a = t * 3.14159265358979323846f;
    return t == 0.0f ? 1.0f : sinf(a) / a;
}
static float lanczos3(float x) { return fabsf(x) < 3.0f ? lanczos_sinc(x) * lanczos_sinc(x / 3.0f) : 0.0f; }
static float triangle(float x) { return fabsf(x) < 1.0f ? 1.0f - fabsf(x) : 0.0f; }
static int32_t scale_taps(uint32_t filter, uint32_t n, uint32_t m, uint32_t o, std::vector<float>& w) {
    const float ratio = (float)n / (float)m;
    const float sratio = ratio < 1.0f ? 1.0f : ratio;
    const float support = (filter == MMR_SCALE_FILTER_TRIANGLE ? 1.0f : 3.0f) * sratio;
    float c = ((float)o + 0.5f) * ratio;
    const int64_t left = std::clamp<int64_t>((int64_t)floorf(c - support), 0, (int64_t)n - 1);
    const int64_t right = std::clamp<int64_t>((int64_t)ceilf(c + support), left + 1, (int64_t)n);
    c = c - 0.5f;
    w.clear();
    float sum = 0.0f;
    for (int64_t i = left; i < right; ++i) {
        const float x = ((float)i - c) / sratio;
        w.push_back(filter == MMR_SCALE_FILTER_TRIANGLE ? triangle(x) : lanczos3(x));
        sum += w.back();
    }
    for (float& v : w) v /= sum;
    return (int32_t)left;
}

// scaled_dimensions / scaled_even (crates/render/src/compositor.rs:211-254).
static void scaled_dimensions(uint32_t sw, uint32_t sh, uint32_t cw, uint32_t ch, uint32_t mode, uint32_t& ow, uint32_t& oh) {
    auto even = [](uint64_t source, uint64_t target, uint64_t den) -> uint32_t {
        const uint64_t units = (source * target + den) / (2 * den);
        return (uint32_t)std::max<uint64_t>(units * 2, 2);
    };
    if (mode == MMR_SCALE_STRETCH) {
        ow = cw, oh = ch;
    } else if (mode == MMR_SCALE_NATIVE) {
        ow = sw, oh = sh;
    } else {
        const bool width_limited = (uint64_t)cw * sh <= (uint64_t)ch * sw;
        if (mode == MMR_SCALE_FILL ? !width_limited : width_limited) ow = cw, oh = even(sh, cw, sw);
        else ow = even(sw, ch, sh), oh = ch;
    }
}

// The taps of (filter, n, m) for every output coordinate, from the context's cache.
static std::shared_ptr<const ScaleTaps> cached_taps(mmr_recon_context* ctx, uint32_t filter, uint32_t n, uint32_t m) {
    const auto key = std::make_tuple(filter, n, m);
    auto it = ctx->scale_taps.find(key);
    if (it != ctx->scale_taps.end()) return it->second;
    if (ctx->scale_taps.size() >= 64) ctx->scale_taps.clear();
    auto t = std::make_shared<ScaleTaps>();
    std::vector<std::vector<float>> per(m);
    t->left.resize(m);
    t->count.resize(m);
    for (uint32_t o = 0; o < m; ++o) {
        t->left[o] = scale_taps(filter, n, m, o, per[o]);
        t->count[o] = (uint32_t)per[o].size();
        t->max = std::max(t->max, t->count[o]);
    }
    t->w.assign((size_t)m * t->max, 0.0f);
    for (uint32_t o = 0; o < m; ++o) std::copy(per[o].begin(), per[o].end(), t->w.begin() + (size_t)o * t->max);
    ctx->scale_taps.emplace(key, t);
    return t;
}

mmr_status scale_submit_on(mmr_recon_context* ctx, hipStream_t stream, const mmr_recon_scale_job* in, uint32_t count, uint32_t flags,
                           uint64_t* out_ticket) {
    CHECK_CTX(ctx);
    if (!in || count == 0) return fail(MMR_STATUS_INVALID_ARGUMENT, "no jobs");
    // Per plane: the device record plus byte offsets of its tap arrays in the blob (one copy per
    // distinct geometry in the submission) and of its f32 intermediate in scratch.
    struct PlaneHost {
        mmr::ScalePlaneDev d;
        size_t vleft, vcount, vw, hleft, hcount, hw, tmp;
    };
    struct Placed {
        size_t left, count, w;
        uint32_t max;
    };
    std::vector<PlaneHost> planes;
    std::vector<char> taps;
    std::map<std::tuple<uint32_t, uint32_t, uint32_t>, Placed> placed;
    size_t scratch_total = 0;
    auto put = [&taps](const void* data, size_t bytes) {
        const size_t off = (taps.size() + 15) & ~size_t(15);
        taps.resize(off + bytes);
        if (bytes) memcpy(taps.data() + off, data, bytes);
        return off;
    };
    auto place = [&](uint32_t filter, uint32_t n, uint32_t m) -> Placed {
        const auto key = std::make_tuple(filter, n, m);
        auto it = placed.find(key);
        if (it != placed.end()) return it->second;
        auto t = cached_taps(ctx, filter, n, m);
        Placed p{put(t->left.data(), t->left.size() * 4), put(t->count.data(), t->count.size() * 4),
                 put(t->w.data(), t->w.size() * 4), t->max};
        placed.emplace(key, p);
        return p;
    };
    for (uint32_t i = 0; i < count; ++i) {
        const mmr_recon_scale_job& j = in[i];
        if (j.struct_size < sizeof(mmr_recon_scale_job)) return fail(MMR_STATUS_INVALID_ARGUMENT, "scale struct_size too small");
        if (j.width == 0 || j.height == 0 || j.canvas_width == 0 || j.canvas_height == 0 || ((j.width | j.height | j.canvas_width | j.canvas_height) & 1) ||
            j.width > 16384 || j.height > 16384 || j.canvas_width > 16384 || j.canvas_height > 16384)
            return fail(MMR_STATUS_INVALID_DATA, "scale job %u: frame and canvas sizes must be even, non-zero and at most 16384", i);
        if (j.mode > MMR_SCALE_NATIVE) return fail(MMR_STATUS_INVALID_ARGUMENT, "scale job %u: mode %u is invalid", i, j.mode);
        if (j.filter > MMR_SCALE_FILTER_TRIANGLE) return fail(MMR_STATUS_INVALID_ARGUMENT, "scale job %u: filter %u is invalid", i, j.filter);
        uint32_t sw, sh;
        scaled_dimensions(j.width, j.height, j.canvas_width, j.canvas_height, j.mode, sw, sh);
        if (sw > 65536 || sh > 65536) return fail(MMR_STATUS_INVALID_DATA, "scale job %u: scaled size out of range", i);
        for (int k = 0; k < 3; ++k) {
            const uint32_t div = k ? 2 : 1;
            const uint32_t pw = j.width / div, ph = j.height / div, nw = sw / div, nh = sh / div;
            const uint32_t W = j.canvas_width / div, H = j.canvas_height / div;
            if (!j.src[k] || !j.dst[k] || j.src_stride[k] < pw || j.dst_stride[k] < W)
                return fail(MMR_STATUS_INVALID_ARGUMENT, "scale job %u plane %d: NULL plane or stride below its width", i, k);
            PlaneHost plane{};
            mmr::ScalePlaneDev& d = plane.d;
            d.src = j.src[k];
            d.dst = j.dst[k];
            d.src_w = pw, d.src_h = ph, d.src_stride = j.src_stride[k];
            d.dst_w = W, d.dst_h = H, d.dst_stride = j.dst_stride[k];
            d.copy_w = std::min(nw, W), d.copy_h = std::min(nh, H);
            d.sx = (nw - d.copy_w) / 2, d.sy = (nh - d.copy_h) / 2;
            d.dst_x = (W - d.copy_w) / 2, d.dst_y = (H - d.copy_h) / 2;
            d.background = k ? 128u : (j.range == MMR_RANGE_LIMITED ? 16u : 0u);
            d.resize = !(nw == pw && nh == ph);
            d.tiled = 1;
            if (d.resize) {
                // The window's rows / columns of the whole-axis tables: row r of the window is
                // scaled row sy + r, column c is scaled column sx + c.
                const Placed v = place(j.filter, ph, nh), h = place(j.filter, pw, nw);
                // One-pass tiles when every run of kScaleTileW window columns reaches at most
                // kScaleTileSrc source columns (taps are monotonic in the output coordinate) and
                // a column has at most kScaleTileTaps taps.
                const auto ht = cached_taps(ctx, j.filter, pw, nw);
                if (ht->max > (uint32_t)mmr::kScaleTileTaps) d.tiled = 0;
                for (uint32_t c = 0; c < d.copy_w && d.tiled; ++c) {
                    const uint32_t e = d.sx + std::min(c + (uint32_t)mmr::kScaleTileW - 1, d.copy_w - 1);
                    if ((int64_t)ht->left[e] + ht->count[e] - ht->left[d.sx + c] > mmr::kScaleTileSrc) d.tiled = 0;
                }
                d.vmax = v.max, d.hmax = h.max;
                plane.vleft = v.left + (size_t)d.sy * 4;
                plane.vcount = v.count + (size_t)d.sy * 4;
                plane.vw = v.w + (size_t)d.sy * v.max * 4;
                plane.hleft = h.left + (size_t)d.sx * 4;
                plane.hcount = h.count + (size_t)d.sx * 4;
                plane.hw = h.w + (size_t)d.sx * h.max * 4;
                if (!d.tiled) {
                    plane.tmp = scratch_total;
                    scratch_total += (((size_t)pw * d.copy_h * 4) + 255) & ~size_t(255);
                }
            }
            planes.push_back(plane);
        }
    }
    std::vector<char> blob(planes.size() * sizeof(mmr::ScalePlaneDev));
    const size_t taps_off = (blob.size() + 255) & ~size_t(255);
    blob.resize(taps_off + taps.size());
    if (!taps.empty()) memcpy(blob.data() + taps_off, taps.data(), taps.size());
    for (size_t p = 0; p < planes.size(); ++p) memcpy(blob.data() + p * sizeof(mmr::ScalePlaneDev), &planes[p].d, sizeof(mmr::ScalePlaneDev));
    Submission sub(ctx, stream);
    mmr_status s = sub.begin(blob.data(), blob.size(), 0, scratch_total, [&](char* host, char* d_scratch) {
        char* base = static_cast<char*>(sub.jobs_ptr()) + taps_off;
        auto* hp = reinterpret_cast<mmr::ScalePlaneDev*>(host);
        for (size_t p = 0; p < planes.size(); ++p) {
            if (!planes[p].d.resize) continue;
            hp[p].vleft = reinterpret_cast<const int32_t*>(base + planes[p].vleft);
            hp[p].vcount = reinterpret_cast<const uint32_t*>(base + planes[p].vcount);
            hp[p].vw = reinterpret_cast<const float*>(base + planes[p].vw);
            hp[p].hleft = reinterpret_cast<const int32_t*>(base + planes[p].hleft);
            hp[p].hcount = reinterpret_cast<const uint32_t*>(base + planes[p].hcount);
            hp[p].hw = reinterpret_cast<const float*>(base + planes[p].hw);
            hp[p].tmp = planes[p].d.tiled ? nullptr : reinterpret_cast<float*>(d_scratch + planes[p].tmp);
        }
    });
    if (s) return s;
    auto* dp = reinterpret_cast<const mmr::ScalePlaneDev*>(sub.jobs_ptr());
    auto sizes = [&](size_t from, size_t to, uint32_t& mt, uint32_t& mv, uint32_t& mo) {
        mt = mv = mo = 0;
        for (size_t p = from; p < to; ++p) {
            const mmr::ScalePlaneDev& d = planes[p].d;
            if (d.tiled) {
                const uint32_t tiles = ((d.dst_w + mmr::kScaleTileW - 1) / mmr::kScaleTileW) *
                                       ((d.dst_h + mmr::kScaleTileH - 1) / mmr::kScaleTileH);
                mt = std::max(mt, tiles);
                continue;
            }
            if (d.resize) mv = std::max(mv, d.src_w * d.copy_h);
            mo = std::max(mo, d.dst_w * d.dst_h);
        }
    };
    hipError_t e = hipSuccess;
    uint32_t mt = 0, mv = 0, mo = 0;
    if (flags & MMR_SUBMIT_PER_FRAME_LAUNCH) {
        for (uint32_t i = 0; i < count && e == hipSuccess; ++i) {
            sizes(3 * i, 3 * i + 3, mt, mv, mo);
            e = mmr::launch_scale(dp + 3 * i, 3, mt, mv, mo, stream);
        }
    } else {
        sizes(0, planes.size(), mt, mv, mo);
        e = mmr::launch_scale(dp, (uint32_t)planes.size(), mt, mv, mo, stream);
    }
    if (e != hipSuccess) return hip_fail(e, "scale launch");
    return sub.end(out_ticket);
}

}  // namespace mmr_host

extern "C" {

uint32_t mmr_recon_abi_version(void) { return MMR_RECON_ABI_VERSION; }

size_t mmr_recon_last_error_message(char* buffer, size_t capacity) {
    const size_t len = strnlen(g_last_error, sizeof g_last_error - 1);
    if (buffer && capacity) {
        const size_t n = std::min(capacity - 1, len);
        memcpy(buffer, g_last_error, n);
        buffer[n] = 0;
    }
    return len + 1;
}

mmr_status mmr_recon_context_create(const mmr_recon_config* config, mmr_recon_context** out_context) {
    return guarded("mmr_recon_context_create", [&]() -> mmr_status {
        if (!out_context) return fail(MMR_STATUS_INVALID_ARGUMENT, "out_context is NULL");
        *out_context = nullptr;
        mmr_recon_config cfg;
        memset(&cfg, 0, sizeof cfg);
        if (config) {
            if (config->struct_size < sizeof(mmr_recon_config)) return fail(MMR_STATUS_INVALID_ARGUMENT, "config struct_size too small");
            cfg = *config;
        }
        int count = 0;
        hipError_t e = hipGetDeviceCount(&count);
        if (e != hipSuccess || count == 0) return fail(MMR_STATUS_UNSUPPORTED, "no HIP device is available");
        if (cfg.device_ordinal < 0 || cfg.device_ordinal >= count)
            return fail(MMR_STATUS_INVALID_ARGUMENT, "device ordinal %d out of range (%d devices)", cfg.device_ordinal, count);
        auto* ctx = new mmr_recon_context();
        ctx->device = cfg.device_ordinal;
        ctx->queue_depth = cfg.queue_depth ? cfg.queue_depth : 256;
        ctx->persistent_groups = cfg.persistent_groups;
        ctx->tuning = mmr_host::default_tuning();
        e = hipSetDevice(ctx->device);
        hipDeviceProp_t prop;
        if (e == hipSuccess) e = hipGetDeviceProperties(&prop, ctx->device);
        if (e == hipSuccess) ctx->cu_count = (uint32_t)prop.multiProcessorCount;
        if (e == hipSuccess) e = noalloc_scratch(ctx->device, &ctx->noalloc_scratch);   // with the device current
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
        if (e == hipSuccess) {
            // Submission blocks come from the device's stream-ordered pool; keeping freed memory in
            // the pool (instead of returning it at each synchronisation) lets a submission allocate
            // without waiting for the device while earlier ones are in flight.
            hipMemPool_t pool = nullptr;
            if (hipDeviceGetDefaultMemPool(&pool, ctx->device) == hipSuccess && pool) {
                uint64_t keep = UINT64_MAX;
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
            }
        }
        if (e == hipSuccess) e = mmr::dv_upload_tables(ctx->stream);
        if (e == hipSuccess) e = hipMalloc((void**)&ctx->flat_scaling, 96 + 128);
        if (e == hipSuccess) {
            uint8_t flat[224];
            memset(flat, 16, sizeof flat);
            e = hipMemcpy(ctx->flat_scaling, flat, sizeof flat, hipMemcpyHostToDevice);
        }
        if (e != hipSuccess) {
            mmr_status s = hip_fail(e, "context setup");
            if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
            if (ctx->flat_scaling) (void)hipFree(ctx->flat_scaling);
            if (ctx->linear_tables) (void)hipFree(ctx->linear_tables);
            delete ctx;
            return s;
        }
        *out_context = ctx;
        return MMR_STATUS_OK;
    });
}

void mmr_recon_context_destroy(mmr_recon_context* ctx) {
    guarded_void("mmr_recon_context_destroy", [&]() {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        for (auto& kv : ctx->tickets) release_ticket(ctx, kv.second);
        for (hipEvent_t ev : ctx->free_events) (void)hipEventDestroy(ev);
        for (auto& kv : ctx->free_pinned) (void)hipHostFree(kv.second);
        if (ctx->ring) (void)hipHostFree(ctx->ring);
        if (ctx->stream_block) {
            (void)hipFreeAsync(ctx->stream_block, ctx->stream);
            (void)hipStreamSynchronize(ctx->stream);
        }
        if (ctx->flat_scaling) (void)hipFree(ctx->flat_scaling);
        if (ctx->linear_tables) (void)hipFree(ctx->linear_tables);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

mmr_status mmr_recon_device_alloc(mmr_recon_context* ctx, size_t bytes, void** out) {
    return guarded("mmr_recon_device_alloc", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        if (!out) return fail(MMR_STATUS_INVALID_ARGUMENT, "out pointer is NULL");
        LOCK_AND_SET_DEVICE(ctx);
        hipError_t e = hipMalloc(out, bytes ? bytes : 1);
        return e == hipSuccess ? MMR_STATUS_OK : hip_fail(e, "hipMalloc");
    });
}

mmr_status mmr_recon_device_free(mmr_recon_context* ctx, void* p) {
    return guarded("mmr_recon_device_free", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        LOCK_AND_SET_DEVICE(ctx);
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess && p) e = hipFree(p);
        return e == hipSuccess ? MMR_STATUS_OK : hip_fail(e, "hipFree");
    });
}

mmr_status mmr_recon_copy_to_device(mmr_recon_context* ctx, void* dst, const void* src, size_t bytes) {
    return guarded("mmr_recon_copy_to_device", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        if ((!dst || !src) && bytes) return fail(MMR_STATUS_INVALID_ARGUMENT, "NULL copy pointer");
        LOCK_AND_SET_DEVICE(ctx);
        hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        return e == hipSuccess ? MMR_STATUS_OK : hip_fail(e, "copy to device");
    });
}

mmr_status mmr_recon_copy_to_host(mmr_recon_context* ctx, void* dst, const void* src, size_t bytes) {
    return guarded("mmr_recon_copy_to_host", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        if ((!dst || !src) && bytes) return fail(MMR_STATUS_INVALID_ARGUMENT, "NULL copy pointer");
        LOCK_AND_SET_DEVICE(ctx);
        hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        return e == hipSuccess ? MMR_STATUS_OK : hip_fail(e, "copy to host");
    });
}

}  // extern "C"

// ------------------------------------------------------------------ resident frame cache
// DeviceResourceCache (crates/render/src/device_cache.rs:53-287) holding device allocations.
// Keys order as FrameResourceKey's derived Ord does (namespace, owner, revision, local_frame,
// width, height, variant), which breaks ties between equally old eviction candidates.
struct mmr_recon_frame_cache {
    using Key = std::tuple<uint32_t, uint64_t, uint64_t, int64_t, uint32_t, uint32_t, uint32_t>;
    struct Entry {
        mmr_recon_frame_descriptor descriptor;
        void* storage;
        uint64_t bytes;
        uint64_t last_used;
    };
    mmr_recon_context* ctx = nullptr;
    uint64_t budget = 0, retained = 0, generation = 0, reuses = 0, insertions = 0, evictions = 0;
    std::map<Key, Entry> entries;
};

namespace {

mmr_recon_frame_cache::Key cache_key(const mmr_recon_frame_key& k) {
    return {k.namespace_id, k.owner, k.revision, k.local_frame, k.width, k.height, k.variant};
}

bool same_descriptor(const mmr_recon_frame_descriptor& a, const mmr_recon_frame_descriptor& b) {
    return a.width == b.width && a.height == b.height && a.format == b.format && a.color == b.color &&
           a.alpha_mode == b.alpha_mode;
}

// Stream-ordered release on the context's stream: submissions enqueued before it finish first.
void cache_release(mmr_recon_frame_cache* c, mmr_recon_frame_cache::Entry& e) {
    if (e.storage) (void)hipFreeAsync(e.storage, c->ctx->stream);
    e.storage = nullptr;
    c->retained -= std::min(c->retained, e.bytes);
}

// evict_for (device_cache.rs:263-287): least recently used first, never the current generation.
mmr_status cache_evict_for(mmr_recon_frame_cache* c, uint64_t additional) {
    while (c->retained + additional > c->budget) {
        auto victim = c->entries.end();
        for (auto it = c->entries.begin(); it != c->entries.end(); ++it)
            if (it->second.last_used < c->generation && (victim == c->entries.end() || it->second.last_used < victim->second.last_used))
                victim = it;   // map order: the smallest key wins among equally old entries
        if (victim == c->entries.end())
            return fail(MMR_STATUS_INVALID_STATE, "device resource budget cannot fit %llu more bytes without releasing a resource "
                        "used in generation %llu", (unsigned long long)additional, (unsigned long long)c->generation);
        cache_release(c, victim->second);
        c->entries.erase(victim);
        ++c->evictions;
    }
    return MMR_STATUS_OK;
}

}  // namespace

extern "C" {

mmr_status mmr_recon_frame_cache_create(mmr_recon_context* ctx, uint64_t byte_budget, mmr_recon_frame_cache** out) {
    return guarded("mmr_recon_frame_cache_create", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        if (!out) return fail(MMR_STATUS_INVALID_ARGUMENT, "out_cache is NULL");
        if (byte_budget == 0) return fail(MMR_STATUS_INVALID_DATA, "device resource cache byte budget must be positive");
        auto* c = new mmr_recon_frame_cache();
        c->ctx = ctx;
        c->budget = byte_budget;
        *out = c;
        return MMR_STATUS_OK;
    });
}

void mmr_recon_frame_cache_destroy(mmr_recon_frame_cache* c) {
    guarded_void("mmr_recon_frame_cache_destroy", [&]() {
        if (!c) return;
        {
            std::lock_guard<std::mutex> lock(c->ctx->mu);
            (void)hipSetDevice(c->ctx->device);
            for (auto& kv : c->entries) cache_release(c, kv.second);
            (void)hipStreamSynchronize(c->ctx->stream);
        }
        delete c;
    });
}

mmr_status mmr_recon_frame_cache_begin_frame(mmr_recon_frame_cache* c) {
    return guarded("mmr_recon_frame_cache_begin_frame", [&]() -> mmr_status {
        if (!c) return fail(MMR_STATUS_INVALID_ARGUMENT, "cache is NULL");
        std::lock_guard<std::mutex> lock(c->ctx->mu);
        if (c->generation != UINT64_MAX) ++c->generation;
        return MMR_STATUS_OK;
    });
}

mmr_status mmr_recon_frame_cache_retain(mmr_recon_frame_cache* c, const mmr_recon_frame_key* key, const mmr_recon_frame_descriptor* d,
                                        uint64_t bytes, void** out_ptr, uint32_t* out_status) {
    return guarded("mmr_recon_frame_cache_retain", [&]() -> mmr_status {
        if (!c || !key || !d || !out_ptr) return fail(MMR_STATUS_INVALID_ARGUMENT, "cache / key / descriptor / out pointer is NULL");
        LOCK_AND_SET_DEVICE(c->ctx);
        *out_ptr = nullptr;
        const auto k = cache_key(*key);
        auto it = c->entries.find(k);
        if (it != c->entries.end()) {
            if (!same_descriptor(it->second.descriptor, *d))
                return fail(MMR_STATUS_INVALID_STATE, "stable device resource key was reused with a different descriptor");
            it->second.last_used = c->generation;
            ++c->reuses;
            *out_ptr = it->second.storage;
            if (out_status) *out_status = MMR_CACHE_REUSED;
            return MMR_STATUS_OK;
        }
        if (bytes == 0 || bytes > c->budget)
            return fail(MMR_STATUS_INVALID_DATA, "device resource requires %llu bytes with a %llu byte budget", (unsigned long long)bytes,
                        (unsigned long long)c->budget);
        void* p = nullptr;
        hipError_t e = hipMallocAsync(&p, bytes, c->ctx->stream);
        if (e != hipSuccess) return hip_fail(e, "hipMallocAsync (frame cache)");
        mmr_status s = cache_evict_for(c, bytes);
        if (s) {
            (void)hipFreeAsync(p, c->ctx->stream);   // as the reference drops the created resource
            return s;
        }
        c->entries.emplace(k, mmr_recon_frame_cache::Entry{*d, p, bytes, c->generation});
        c->retained += bytes;
        ++c->insertions;
        *out_ptr = p;
        if (out_status) *out_status = MMR_CACHE_INSERTED;
        return MMR_STATUS_OK;
    });
}

mmr_status mmr_recon_frame_cache_get(mmr_recon_frame_cache* c, const mmr_recon_frame_key* key, const mmr_recon_frame_descriptor* d,
                                     void** out_ptr) {
    return guarded("mmr_recon_frame_cache_get", [&]() -> mmr_status {
        if (!c || !key || !d || !out_ptr) return fail(MMR_STATUS_INVALID_ARGUMENT, "cache / key / descriptor / out pointer is NULL");
        std::lock_guard<std::mutex> lock(c->ctx->mu);
        *out_ptr = nullptr;
        auto it = c->entries.find(cache_key(*key));
        if (it == c->entries.end()) return MMR_STATUS_OK;
        if (!same_descriptor(it->second.descriptor, *d))
            return fail(MMR_STATUS_INVALID_STATE, "stable device resource key was reused with a different descriptor");
        it->second.last_used = c->generation;
        ++c->reuses;
        *out_ptr = it->second.storage;
        return MMR_STATUS_OK;
    });
}

mmr_status mmr_recon_frame_cache_remove(mmr_recon_frame_cache* c, const mmr_recon_frame_key* key, int32_t* out_removed) {
    return guarded("mmr_recon_frame_cache_remove", [&]() -> mmr_status {
        if (!c || !key) return fail(MMR_STATUS_INVALID_ARGUMENT, "cache / key is NULL");
        LOCK_AND_SET_DEVICE(c->ctx);
        auto it = c->entries.find(cache_key(*key));
        const bool found = it != c->entries.end();
        if (found) {
            cache_release(c, it->second);
            c->entries.erase(it);
        }
        if (out_removed) *out_removed = found ? 1 : 0;
        return MMR_STATUS_OK;
    });
}

mmr_status mmr_recon_frame_cache_release_idle(mmr_recon_frame_cache* c, uint64_t max_idle, uint64_t* out_released) {
    return guarded("mmr_recon_frame_cache_release_idle", [&]() -> mmr_status {
        if (!c) return fail(MMR_STATUS_INVALID_ARGUMENT, "cache is NULL");
        LOCK_AND_SET_DEVICE(c->ctx);
        uint64_t released = 0;
        for (auto it = c->entries.begin(); it != c->entries.end();) {
            const uint64_t idle = c->generation - std::min(c->generation, it->second.last_used);
            if (idle == 0 || idle <= max_idle) {
                ++it;
                continue;
            }
            cache_release(c, it->second);
            it = c->entries.erase(it);
            ++released;
        }
        c->evictions += released;
        if (out_released) *out_released = released;
        return MMR_STATUS_OK;
    });
}

mmr_status mmr_recon_frame_cache_clear(mmr_recon_frame_cache* c) {
    return guarded("mmr_recon_frame_cache_clear", [&]() -> mmr_status {
        if (!c) return fail(MMR_STATUS_INVALID_ARGUMENT, "cache is NULL");
        LOCK_AND_SET_DEVICE(c->ctx);
        for (auto& kv : c->entries) cache_release(c, kv.second);
        c->entries.clear();
        c->retained = 0;
        return MMR_STATUS_OK;
    });
}

mmr_status mmr_recon_frame_cache_stats_get(mmr_recon_frame_cache* c, mmr_recon_frame_cache_stats* out) {
    return guarded("mmr_recon_frame_cache_stats_get", [&]() -> mmr_status {
        if (!c || !out) return fail(MMR_STATUS_INVALID_ARGUMENT, "cache / out_stats is NULL");
        if (out->struct_size < sizeof(mmr_recon_frame_cache_stats))
            return fail(MMR_STATUS_INVALID_ARGUMENT, "frame cache stats struct_size too small");
        std::lock_guard<std::mutex> lock(c->ctx->mu);
        out->resources = c->entries.size();
        out->retained_bytes = c->retained;
        out->byte_budget = c->budget;
        out->reuses = c->reuses;
        out->insertions = c->insertions;
        out->evictions = c->evictions;
        out->generation = c->generation;
        return MMR_STATUS_OK;
    });
}

mmr_status mmr_recon_wait(mmr_recon_context* ctx, uint64_t ticket) {
    return guarded("mmr_recon_wait", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        LOCK_AND_SET_DEVICE(ctx);
        auto it = ctx->tickets.find(ticket);
        if (it == ctx->tickets.end()) {
            if (ticket == 0 || ticket >= ctx->next_ticket) return fail(MMR_STATUS_INVALID_ARGUMENT, "unknown ticket %llu", (unsigned long long)ticket);
            return MMR_STATUS_OK;  // already retired
        }
        mmr_status s = finish(ctx, it->second);
        it->second.reported = true;
        return s;
    });
}

mmr_status mmr_recon_poll(mmr_recon_context* ctx, uint64_t ticket, int32_t* out_done) {
    return guarded("mmr_recon_poll", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        if (!out_done) return fail(MMR_STATUS_INVALID_ARGUMENT, "out_done is NULL");
        LOCK_AND_SET_DEVICE(ctx);
        auto it = ctx->tickets.find(ticket);
        if (it == ctx->tickets.end()) {
            if (ticket == 0 || ticket >= ctx->next_ticket) return fail(MMR_STATUS_INVALID_ARGUMENT, "unknown ticket %llu", (unsigned long long)ticket);
            *out_done = 1;
            return MMR_STATUS_OK;
        }
        const hipError_t q = it->second.finished ? hipSuccess : hipEventQuery(it->second.done);
        if (q == hipErrorNotReady) {
            *out_done = 0;
            return MMR_STATUS_OK;
        }
        *out_done = 1;
        it->second.reported = true;
        return finish(ctx, it->second);
    });
}

mmr_status mmr_recon_synchronize(mmr_recon_context* ctx) {
    return guarded("mmr_recon_synchronize", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        LOCK_AND_SET_DEVICE(ctx);
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
        // Errors already returned by wait / poll (or carried by server completions) are not repeated.
        mmr_status worst = MMR_STATUS_OK;
        for (auto& kv : ctx->tickets) {
            mmr_status s = finish(ctx, kv.second);
            if (s != MMR_STATUS_OK && worst == MMR_STATUS_OK && !kv.second.reported) worst = s;
            kv.second.reported = true;
        }
        reap(ctx);
        return worst;
    });
}

mmr_status mmr_recon_last_kernel_ms(mmr_recon_context* ctx, uint64_t ticket, double* out_ms) {
    return guarded("mmr_recon_last_kernel_ms", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        if (!out_ms) return fail(MMR_STATUS_INVALID_ARGUMENT, "out_ms is NULL");
        LOCK_AND_SET_DEVICE(ctx);
        auto it = ctx->tickets.find(ticket);
        if (it == ctx->tickets.end()) return fail(MMR_STATUS_INVALID_STATE, "ticket %llu is not tracked", (unsigned long long)ticket);
        mmr_status s = finish(ctx, it->second);
        *out_ms = it->second.kernel_ms;
        return s;
    });
}

mmr_status mmr_recon_kernel_phases_ms(mmr_recon_context* ctx, uint64_t ticket, double* out_ms, uint32_t capacity,
                                      uint32_t* out_count) {
    return guarded("mmr_recon_kernel_phases_ms", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        if (!out_count || (capacity && !out_ms)) return fail(MMR_STATUS_INVALID_ARGUMENT, "out_ms / out_count is NULL");
        LOCK_AND_SET_DEVICE(ctx);
        auto it = ctx->tickets.find(ticket);
        if (it == ctx->tickets.end()) return fail(MMR_STATUS_INVALID_STATE, "ticket %llu is not tracked", (unsigned long long)ticket);
        mmr_status s = finish(ctx, it->second);
        const std::vector<double>& ph = it->second.phase_ms;
        *out_count = (uint32_t)ph.size();
        for (uint32_t i = 0; i < capacity && i < ph.size(); ++i) out_ms[i] = ph[i];
        return s;
    });
}

// ------------------------------------------------------------------ MJPEG
mmr_status mmr_recon_mjpeg_submit(mmr_recon_context* ctx, const mmr_recon_mjpeg_frame* frames, uint32_t count, uint32_t flags,
                                  uint64_t* out_ticket) {
    return guarded("mmr_recon_mjpeg_submit", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        LOCK_AND_SET_DEVICE(ctx);
        return mmr_host::mjpeg_submit_on(ctx, ctx->stream, frames, count, flags, out_ticket);
    });
}

// ------------------------------------------------------------------ MPEG-2
mmr_status mmr_recon_mpeg2_submit(mmr_recon_context* ctx, const mmr_recon_mpeg2_picture* pics, uint32_t count, uint32_t flags,
                                  uint64_t* out_ticket) {
    return guarded("mmr_recon_mpeg2_submit", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        LOCK_AND_SET_DEVICE(ctx);
        return mmr_host::mpeg2_submit_on(ctx, ctx->stream, pics, count, flags, out_ticket);
    });
}

mmr_status mmr_recon_mpeg2_reconstruct_noalloc(mmr_recon_context* ctx, const mmr_recon_mpeg2_picture* pics, uint32_t count,
                                               uint32_t flags) {
    return guarded("mmr_recon_mpeg2_reconstruct_noalloc", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        LOCK_AND_SET_DEVICE(ctx);
        if (!ctx->noalloc_scratch) return fail(MMR_STATUS_INTERNAL_ERROR, "the device scratch of this context is not resolved");
        const Workspace ws{ctx->noalloc_scratch, MMR_RECON_NOALLOC_SCRATCH_BYTES};
        return mmr_host::mpeg2_submit_on(ctx, ctx->stream, pics, count, flags, nullptr, nullptr, &ws);
    });
}

// ------------------------------------------------------------------ DV
mmr_status mmr_recon_dv_submit(mmr_recon_context* ctx, const mmr_recon_dv_frame* frames, uint32_t count, uint32_t flags,
                               uint64_t* out_ticket) {
    return guarded("mmr_recon_dv_submit", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        LOCK_AND_SET_DEVICE(ctx);
        return mmr_host::dv_submit_on(ctx, ctx->stream, frames, count, flags, out_ticket);
    });
}

// ------------------------------------------------------------------ H.264
mmr_status mmr_recon_h264_implicit_weights(int32_t poc0, int32_t poc1, int32_t cur, int32_t lt0, int32_t lt1, int32_t* w0,
                                           int32_t* w1) {
    return guarded("mmr_recon_h264_implicit_weights", [&]() -> mmr_status {
        // implicit_b_prediction_weights, crates/codecs/h264/src/decoder.rs:3577-3620.
        if (!w0 || !w1) return fail(MMR_STATUS_INVALID_ARGUMENT, "NULL output");
        int64_t td = std::clamp<int64_t>((int64_t)poc1 - poc0, -128, 127);
        int32_t weight1 = 32;
        if (!(lt0 || lt1 || td == 0)) {
            const int64_t tb = std::clamp<int64_t>((int64_t)cur - poc0, -128, 127);
            const int64_t tx = (16384 + (td < 0 ? -td : td) / 2) / td;
            const int64_t scale = std::clamp<int64_t>((tb * tx + 32) >> 6, -1024, 1023);
            const int64_t candidate = scale >> 2;
            weight1 = (candidate >= -64 && candidate <= 128) ? (int32_t)candidate : 32;
        }
        *w0 = 64 - weight1;
        *w1 = weight1;
        return MMR_STATUS_OK;
    });
}

mmr_status mmr_recon_h264_submit(mmr_recon_context* ctx, const mmr_recon_h264_picture* pics, uint32_t count, uint32_t flags,
                                 uint64_t* out_ticket) {
    return guarded("mmr_recon_h264_submit", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        LOCK_AND_SET_DEVICE(ctx);
        return mmr_host::h264_submit_on(ctx, ctx->stream, pics, count, flags, out_ticket);
    });
}

// ------------------------------------------------------------------ colour / blend
mmr_status mmr_recon_yuv_to_rgb_submit(mmr_recon_context* ctx, const mmr_recon_yuv_to_rgb_job* in, uint32_t count, uint32_t flags,
                                       uint64_t* out_ticket) {
    return guarded("mmr_recon_yuv_to_rgb_submit", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        LOCK_AND_SET_DEVICE(ctx);
        return mmr_host::yuv_to_rgb_submit_on(ctx, ctx->stream, in, count, flags, out_ticket);
    });
}

mmr_status mmr_recon_composite_submit(mmr_recon_context* ctx, const mmr_recon_composite_job* in, uint32_t count, uint32_t flags,
                                      uint64_t* out_ticket) {
    return guarded("mmr_recon_composite_submit", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        LOCK_AND_SET_DEVICE(ctx);
        return mmr_host::composite_submit_on(ctx, ctx->stream, in, count, flags, out_ticket);
    });
}

mmr_status mmr_recon_yuv_composite_submit(mmr_recon_context* ctx, const mmr_recon_yuv_to_rgb_job* conversions,
                                          const mmr_recon_composite_job* composites, uint32_t count, uint32_t flags, uint64_t* out_ticket) {
    return guarded("mmr_recon_yuv_composite_submit", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        LOCK_AND_SET_DEVICE(ctx);
        return mmr_host::yuv_composite_submit_on(ctx, ctx->stream, conversions, composites, count, flags, out_ticket);
    });
}

mmr_status mmr_recon_linear_composite_submit(mmr_recon_context* ctx, const mmr_recon_linear_composite_job* in, uint32_t count,
                                             uint32_t flags, uint64_t* out_ticket) {
    return guarded("mmr_recon_linear_composite_submit", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        LOCK_AND_SET_DEVICE(ctx);
        return mmr_host::linear_composite_submit_on(ctx, ctx->stream, in, count, flags, out_ticket);
    });
}

mmr_status mmr_recon_linear_convert_submit(mmr_recon_context* ctx, const mmr_recon_linear_convert_job* in, uint32_t count,
                                           uint32_t flags, uint64_t* out_ticket) {
    return guarded("mmr_recon_linear_convert_submit", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        LOCK_AND_SET_DEVICE(ctx);
        return mmr_host::linear_convert_submit_on(ctx, ctx->stream, in, count, flags, out_ticket);
    });
}

mmr_status mmr_recon_yuv_blend_submit(mmr_recon_context* ctx, const mmr_recon_yuv_blend_job* in, uint32_t count, uint32_t flags,
                                      uint64_t* out_ticket) {
    return guarded("mmr_recon_yuv_blend_submit", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        LOCK_AND_SET_DEVICE(ctx);
        return mmr_host::yuv_blend_submit_on(ctx, ctx->stream, in, count, flags, out_ticket);
    });
}

mmr_status mmr_recon_overlay_prepare_submit(mmr_recon_context* ctx, const mmr_recon_overlay_prepare_job* jobs, uint32_t count, uint32_t flags,
                                            uint64_t* out_ticket) {
    return guarded("mmr_recon_overlay_prepare_submit", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        LOCK_AND_SET_DEVICE(ctx);
        return mmr_host::overlay_prepare_submit_on(ctx, ctx->stream, jobs, count, flags, out_ticket);
    });
}

mmr_status mmr_recon_scale_submit(mmr_recon_context* ctx, const mmr_recon_scale_job* jobs, uint32_t count, uint32_t flags, uint64_t* out_ticket) {
    return guarded("mmr_recon_scale_submit", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        LOCK_AND_SET_DEVICE(ctx);
        return mmr_host::scale_submit_on(ctx, ctx->stream, jobs, count, flags, out_ticket);
    });
}

}  // extern "C"

// ====================================================================== prepared batches
//
// A batch prepared once and replayed: the submission's job records go once into a persistent
// device block and its launches (counter fill, kernels) are captured into a HIP graph on a private
// stream; a launch replays the graph on the context's stream between the ticket's events. The
// graph holds no memset node (the counters are cleared by a fill kernel), so every replay starts
// from zeroed counters.

struct mmr_recon_prepared {
    PreparedCapture cap;
    uint32_t codec = 0;
};

extern "C" {

mmr_status mmr_recon_prepare(mmr_recon_context* ctx, uint32_t codec, const void* descriptors, uint32_t count, uint32_t flags,
                             mmr_recon_prepared** out) {
    return guarded("mmr_recon_prepare", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        if (!out) return fail(MMR_STATUS_INVALID_ARGUMENT, "out_prepared is NULL");
        *out = nullptr;
        if (!descriptors || count == 0) return fail(MMR_STATUS_INVALID_ARGUMENT, "no descriptors");
        LOCK_AND_SET_DEVICE(ctx);
        auto* p = new mmr_recon_prepared();
        p->codec = codec;
        hipError_t e = hipStreamCreateWithFlags(&p->cap.stream, hipStreamNonBlocking);   // default priority
        if (e != hipSuccess) {
            delete p;
            return hip_fail(e, "capture stream");
        }
        mmr_status s;
        switch (codec) {
        case MMR_CODEC_MJPEG:
            s = mmr_host::mjpeg_submit_on(ctx, p->cap.stream, static_cast<const mmr_recon_mjpeg_frame*>(descriptors), count, flags, nullptr, &p->cap);
            break;
        case MMR_CODEC_MPEG2:
            s = mmr_host::mpeg2_submit_on(ctx, p->cap.stream, static_cast<const mmr_recon_mpeg2_picture*>(descriptors), count, flags, nullptr, &p->cap);
            break;
        case MMR_CODEC_DV:
            s = mmr_host::dv_submit_on(ctx, p->cap.stream, static_cast<const mmr_recon_dv_frame*>(descriptors), count, flags, nullptr, &p->cap);
            break;
        case MMR_CODEC_H264:
            s = mmr_host::h264_submit_on(ctx, p->cap.stream, static_cast<const mmr_recon_h264_picture*>(descriptors), count, flags, nullptr, &p->cap);
            break;
        default:
            s = fail(MMR_STATUS_INVALID_ARGUMENT, "codec %u invalid", codec);
        }
        if (s != MMR_STATUS_OK || !p->cap.exec) {
            if (p->cap.exec) (void)hipGraphExecDestroy(p->cap.exec);
            if (p->cap.graph) (void)hipGraphDestroy(p->cap.graph);
            if (p->cap.d_block) (void)hipFree(p->cap.d_block);
            (void)hipStreamDestroy(p->cap.stream);
            delete p;
            return s != MMR_STATUS_OK ? s : fail(MMR_STATUS_INTERNAL_ERROR, "graph capture produced no graph");
        }
        *out = p;
        return MMR_STATUS_OK;
    });
}

mmr_status mmr_recon_prepared_launch(mmr_recon_context* ctx, mmr_recon_prepared* p, uint64_t* out_ticket) {
    return guarded("mmr_recon_prepared_launch", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        if (!p) return fail(MMR_STATUS_INVALID_ARGUMENT, "prepared batch is NULL");
        LOCK_AND_SET_DEVICE(ctx);
        if (ctx->tickets.size() >= ctx->queue_depth) reap(ctx);
        if (ctx->tickets.size() >= ctx->queue_depth) return fail(MMR_STATUS_BUSY, "queue full: %u tickets in flight", ctx->queue_depth);
        TicketRecord rec;
        hipError_t e = take_event(ctx, &rec.start);
        if (e == hipSuccess) e = take_event(ctx, &rec.kernels_done);
        if (e == hipSuccess) e = take_event(ctx, &rec.done);
        if (e == hipSuccess) e = take_pinned(ctx, kPinnedJobs, &rec.pinned, &rec.pinned_bytes);
        if (e == hipSuccess) {
            rec.host_status = reinterpret_cast<uint32_t*>(rec.pinned);
            *rec.host_status = 0;
            e = hipEventRecord(rec.start, ctx->stream);
        }
        if (e == hipSuccess) e = hipGraphLaunch(p->cap.exec, ctx->stream);
        if (e == hipSuccess) e = hipEventRecord(rec.kernels_done, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(rec.host_status, p->cap.status, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipEventRecord(rec.done, ctx->stream);
        if (e != hipSuccess) {
            // the events may be in flight: wait for the stream before releasing them
            (void)hipStreamSynchronize(ctx->stream);
            release_ticket(ctx, rec);
            return hip_fail(e, "prepared batch launch");
        }
        const uint64_t id = ctx->next_ticket++;
        ctx->tickets.emplace(id, std::move(rec));
        if (out_ticket) *out_ticket = id;
        return MMR_STATUS_OK;
    });
}

void mmr_recon_prepared_destroy(mmr_recon_context* ctx, mmr_recon_prepared* p) {
    guarded_void("mmr_recon_prepared_destroy", [&]() {
        if (!p) return;
        if (ctx) {
            std::lock_guard<std::mutex> lock(ctx->mu);
            (void)hipSetDevice(ctx->device);
            (void)hipStreamSynchronize(ctx->stream);   // no launch of this batch may still be running
        }
        if (p->cap.exec) (void)hipGraphExecDestroy(p->cap.exec);
        if (p->cap.graph) (void)hipGraphDestroy(p->cap.graph);
        if (p->cap.d_block) (void)hipFree(p->cap.d_block);
        if (p->cap.stream) (void)hipStreamDestroy(p->cap.stream);
        delete p;
    });
}

}  // extern "C"

// ====================================================================== tuning

extern "C" {

mmr_status mmr_recon_get_tuning(mmr_recon_context* ctx, mmr_recon_tuning* out) {
    return guarded("mmr_recon_get_tuning", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        if (!out || out->struct_size < sizeof(mmr_recon_tuning)) return fail(MMR_STATUS_INVALID_ARGUMENT, "tuning struct_size too small");
        std::lock_guard<std::mutex> lock(ctx->mu);
        *out = ctx->tuning;
        return MMR_STATUS_OK;
    });
}

mmr_status mmr_recon_set_tuning(mmr_recon_context* ctx, const mmr_recon_tuning* in) {
    return guarded("mmr_recon_set_tuning", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        if (!in || in->struct_size < sizeof(mmr_recon_tuning)) return fail(MMR_STATUS_INVALID_ARGUMENT, "tuning struct_size too small");
        mmr_status s = mmr_host::check_tuning(*in);
        if (s) return s;
        std::lock_guard<std::mutex> lock(ctx->mu);
        ctx->tuning = *in;
        ctx->tuning.struct_size = sizeof(mmr_recon_tuning);
        return MMR_STATUS_OK;
    });
}

mmr_status mmr_recon_autotune(mmr_recon_context* ctx, uint32_t codec, const void* descriptors, uint32_t count, uint32_t flags,
                              uint32_t repeats, mmr_recon_tuning* out_best, mmr_recon_tuning_trial* trials, uint32_t trial_capacity,
                              uint32_t* out_trials) {
    return guarded("mmr_recon_autotune", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        if (!descriptors || count == 0) return fail(MMR_STATUS_INVALID_ARGUMENT, "no descriptors");
        if (repeats == 0) repeats = 3;
        mmr_recon_tuning base;
        {
            std::lock_guard<std::mutex> lock(ctx->mu);
            base = ctx->tuning;
        }
        // Candidates: the codec's own fields varied, the others as they are.
        std::vector<mmr_recon_tuning> cands;
        const uint32_t groups[3] = {16, 32, 64};
        switch (codec) {
        case MMR_CODEC_MJPEG:
        case MMR_CODEC_MPEG2:
        case MMR_CODEC_DV:
            for (uint32_t g : groups) {
                mmr_recon_tuning t = base;
                (codec == MMR_CODEC_MJPEG ? t.mjpeg_group_blocks : codec == MMR_CODEC_MPEG2 ? t.mpeg2_group_blocks : t.dv_group_blocks) = g;
                cands.push_back(t);
            }
            break;
        case MMR_CODEC_H264:
            for (uint32_t rows : {4u, 8u, 16u})
                for (uint32_t per_cu : {1u, 2u})
                    for (uint32_t publish : {1u, 2u, 4u, 8u}) {
                        mmr_recon_tuning t = base;
                        t.h264_deblock_band_rows = rows;
                        t.h264_deblock_groups_per_cu = per_cu;
                        t.h264_deblock_publish = publish;
                        cands.push_back(t);
                    }
            break;
        default:
            return fail(MMR_STATUS_INVALID_ARGUMENT, "codec %u invalid", codec);
        }
        auto submit = [&](uint64_t* ticket) -> mmr_status {
            switch (codec) {
            case MMR_CODEC_MJPEG: return mmr_recon_mjpeg_submit(ctx, static_cast<const mmr_recon_mjpeg_frame*>(descriptors), count, flags, ticket);
            case MMR_CODEC_MPEG2: return mmr_recon_mpeg2_submit(ctx, static_cast<const mmr_recon_mpeg2_picture*>(descriptors), count, flags, ticket);
            case MMR_CODEC_DV: return mmr_recon_dv_submit(ctx, static_cast<const mmr_recon_dv_frame*>(descriptors), count, flags, ticket);
            default: return mmr_recon_h264_submit(ctx, static_cast<const mmr_recon_h264_picture*>(descriptors), count, flags, ticket);
            }
        };
        auto restore = [&]() {
            std::lock_guard<std::mutex> lock(ctx->mu);
            ctx->tuning = base;
        };
        size_t best = 0;
        std::vector<double> medians(cands.size());
        for (size_t c = 0; c < cands.size(); ++c) {
            {
                std::lock_guard<std::mutex> lock(ctx->mu);
                ctx->tuning = cands[c];
            }
            std::vector<double> ms;
            for (uint32_t r = 0; r <= repeats; ++r) {
                uint64_t t = 0;
                mmr_status s = submit(&t);
                if (s == MMR_STATUS_OK) s = mmr_recon_wait(ctx, t);
                double k = 0.0;
                if (s == MMR_STATUS_OK) s = mmr_recon_last_kernel_ms(ctx, t, &k);
                if (s != MMR_STATUS_OK) {
                    restore();
                    return s;
                }
                if (r > 0) ms.push_back(k);   // the first run is untimed
            }
            std::sort(ms.begin(), ms.end());
            medians[c] = ms[ms.size() / 2];
            if (medians[c] < medians[best]) best = c;
            if (trials && c < trial_capacity) {
                trials[c].tuning = cands[c];
                trials[c].kernel_ms = medians[c];
            }
        }
        {
            std::lock_guard<std::mutex> lock(ctx->mu);
            ctx->tuning = cands[best];
        }
        if (out_best) *out_best = cands[best];
        if (out_trials) *out_trials = (uint32_t)cands.size();
        return MMR_STATUS_OK;
    });
}

}  // extern "C"

// ====================================================================== serving
//
// Continuous batching. A scheduler thread owns `slots` HIP streams. Each iteration it
//   1. retires batches whose completion event has fired (frame completions with latency
//      enqueue -> observed completion; requests that finish release their admitted bytes),
//   2. admits pending requests in FIFO order while admitted bytes + request bytes <= budget
//      (a request larger than the whole budget is admitted alone),
//   3. while a slot is free and admitted requests have frames left, launches one batch: up to
//      max_frames_per_request_per_batch next frames of every admitted request (round robin),
//      at most max_batch_frames per codec; per codec one batched submit, then the batch's
//      YUV->RGB and composite jobs, all on the slot's stream, then the batch event.
// Frames of one request, and requests of one stream, are ordered: a batch waits (device-side,
// hipStreamWaitEvent) on the previous batch that carried frames of the same request or stream.

#include <chrono>
#include <condition_variable>
#include <list>
#include <memory>
#include <thread>
#include <unordered_map>

namespace mmr_host {

using Clock = std::chrono::steady_clock;

struct BatchEvent {
    hipEvent_t ev = nullptr;
    ~BatchEvent() {
        if (ev) (void)hipEventDestroy(ev);
    }
};

size_t codec_struct_size(uint32_t codec) {
    switch (codec) {
    case MMR_CODEC_MJPEG: return sizeof(mmr_recon_mjpeg_frame);
    case MMR_CODEC_MPEG2: return sizeof(mmr_recon_mpeg2_picture);
    case MMR_CODEC_DV: return sizeof(mmr_recon_dv_frame);
    case MMR_CODEC_H264: return sizeof(mmr_recon_h264_picture);
    default: return 0;
    }
}

struct ServerRequest {
    uint64_t id = 0, tag = 0, hbm = 0;
    uint32_t codec = 0, stream_id = 0, frame_count = 0;
    std::vector<char> frames;
    std::vector<mmr_recon_yuv_to_rgb_job> convert;
    std::vector<mmr_recon_composite_job> composite;
    uint32_t next = 0;   // next frame to place in a batch
    uint32_t done = 0;   // frames completed
    mmr_status status = MMR_STATUS_OK;
    Clock::time_point t0;
    std::shared_ptr<BatchEvent> last;   // last launched batch carrying frames of this request
    const char* frame_ptr(uint32_t i) const { return frames.data() + (size_t)i * codec_struct_size(codec); }
};

struct InflightBatch {
    std::shared_ptr<BatchEvent> ev;
    int slot = -1;
    std::vector<std::pair<ServerRequest*, uint32_t>> frames;
    std::vector<uint64_t> tickets;
};

}  // namespace mmr_host

struct mmr_recon_server {
    mmr_recon_context* ctx = nullptr;
    mmr_recon_server_config cfg{};
    std::vector<hipStream_t> streams;
    std::vector<bool> slot_busy;
    std::mutex mu;
    std::condition_variable cv;
    std::condition_variable idle_cv;
    std::deque<std::unique_ptr<mmr_host::ServerRequest>> pending;
    std::list<std::unique_ptr<mmr_host::ServerRequest>> active;
    std::deque<mmr_host::InflightBatch> inflight;
    std::deque<mmr_recon_frame_completion> completions;
    std::unordered_map<uint32_t, std::shared_ptr<mmr_host::BatchEvent>> stream_last;   // per stream id
    mmr_recon_server_stats stats{};
    uint64_t next_id = 1;
    bool stop = false;
    std::thread thread;
};

namespace mmr_host {

mmr_status launch_batch(mmr_recon_server* s, int slot, InflightBatch& b, std::vector<ServerRequest*>& reqs_in_batch) {
    mmr_recon_context* ctx = s->ctx;
    hipStream_t st = s->streams[slot];
    // Device-side ordering after earlier batches of the same requests.
    std::vector<hipEvent_t> waits;
    for (ServerRequest* r : reqs_in_batch)
        if (r->last && r->last->ev) waits.push_back(r->last->ev);
    std::sort(waits.begin(), waits.end());
    waits.erase(std::unique(waits.begin(), waits.end()), waits.end());
    for (hipEvent_t e : waits) {
        hipError_t err = hipStreamWaitEvent(st, e, 0);
        if (err != hipSuccess) return hip_fail(err, "hipStreamWaitEvent");
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const uint32_t codecs[4] = {MMR_CODEC_MJPEG, MMR_CODEC_MPEG2, MMR_CODEC_DV, MMR_CODEC_H264};
    std::vector<mmr_recon_yuv_to_rgb_job> conv;
    std::vector<mmr_recon_composite_job> comp;
    for (uint32_t codec : codecs) {
        std::vector<std::pair<ServerRequest*, uint32_t>> items;
        for (auto& f : b.frames)
            if (f.first->codec == codec) items.push_back(f);
        if (items.empty()) continue;
        const size_t sz = codec_struct_size(codec);
        std::vector<char> blob(items.size() * sz);
        for (size_t i = 0; i < items.size(); ++i) memcpy(blob.data() + i * sz, items[i].first->frame_ptr(items[i].second), sz);
        if (codec == MMR_CODEC_MPEG2) {
            // Translate in-request reference indices to positions in this batch (-1 when the
            // producing picture ran in an earlier batch, which this stream already waited for).
            auto* pics = reinterpret_cast<mmr_recon_mpeg2_picture*>(blob.data());
            for (size_t i = 0; i < items.size(); ++i) {
                for (int32_t* src : {&pics[i].forward_source, &pics[i].backward_source}) {
                    int32_t pos = -1;
                    if (*src >= 0)
                        for (size_t k = 0; k < i; ++k)
                            if (items[k].first == items[i].first && items[k].second == (uint32_t)*src) pos = (int32_t)k;
                    *src = pos;
                }
            }
        }
        uint64_t t = 0;
        mmr_status rc;
        switch (codec) {
        case MMR_CODEC_MJPEG: rc = mjpeg_submit_on(ctx, st, reinterpret_cast<mmr_recon_mjpeg_frame*>(blob.data()), (uint32_t)items.size(), 0, &t); break;
        case MMR_CODEC_MPEG2: rc = mpeg2_submit_on(ctx, st, reinterpret_cast<mmr_recon_mpeg2_picture*>(blob.data()), (uint32_t)items.size(), 0, &t); break;
        case MMR_CODEC_DV: rc = dv_submit_on(ctx, st, reinterpret_cast<mmr_recon_dv_frame*>(blob.data()), (uint32_t)items.size(), 0, &t); break;
        default: rc = h264_submit_on(ctx, st, reinterpret_cast<mmr_recon_h264_picture*>(blob.data()), (uint32_t)items.size(), 0, &t); break;
        }
        if (rc) return rc;
        b.tickets.push_back(t);
    }
    for (auto& f : b.frames) {
        if (!f.first->convert.empty()) conv.push_back(f.first->convert[f.second]);
        if (!f.first->composite.empty()) comp.push_back(f.first->composite[f.second]);
    }
    uint64_t t = 0;
    if (!conv.empty()) {
        mmr_status rc = yuv_to_rgb_submit_on(ctx, st, conv.data(), (uint32_t)conv.size(), 0, &t);
        if (rc) return rc;
        b.tickets.push_back(t);
    }
    if (!comp.empty()) {
        mmr_status rc = composite_submit_on(ctx, st, comp.data(), (uint32_t)comp.size(), 0, &t);
        if (rc) return rc;
        b.tickets.push_back(t);
    }
    b.ev = std::make_shared<BatchEvent>();
    e = hipEventCreateWithFlags(&b.ev->ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(b.ev->ev, st);
    if (e != hipSuccess) return hip_fail(e, "batch event");
    return MMR_STATUS_OK;
}

// Fails every frame of the active and pending requests that no batch carries yet (after an
// exception inside the loop); frames already in flight complete normally.
void fail_unplaced(mmr_recon_server* s, mmr_status st) noexcept {
    try {
        for (auto& r : s->active) {
            while (r->next < r->frame_count) {
                s->completions.push_back(mmr_recon_frame_completion{r->id, r->tag, r->next, st, 0.0});
                r->next++;
                r->done++;
                if (!r->status) r->status = st;
            }
        }
        for (auto& r : s->pending) {
            while (r->next < r->frame_count) {
                s->completions.push_back(mmr_recon_frame_completion{r->id, r->tag, r->next, st, 0.0});
                r->next++;
                r->done++;
            }
        }
    } catch (...) {
        // no memory even for the completions: the requests stay and the next step retries
    }
}

// One scheduler step (retire, admit, launch, wait); returns true when the server stops.
bool server_step(mmr_recon_server* s, std::unique_lock<std::mutex>& lk, size_t& rr) {
    {
        bool progressed = false;
        // 1. retire completed batches (in launch order per slot; any order across slots)
        for (auto it = s->inflight.begin(); it != s->inflight.end();) {
            if (hipEventQuery(it->ev->ev) != hipSuccess) {
                ++it;
                continue;
            }
            const auto now = Clock::now();
            mmr_status st = MMR_STATUS_OK;
            {
                std::lock_guard<std::mutex> clk(s->ctx->mu);
                for (uint64_t t : it->tickets) {
                    auto tr = s->ctx->tickets.find(t);
                    if (tr != s->ctx->tickets.end()) {
                        tr->second.reported = true;   // carried by the frame completions
                        mmr_status x = finish(s->ctx, tr->second);
                        if (x && !st) st = x;
                    }
                }
                reap(s->ctx);
            }
            for (auto& f : it->frames) {
                ServerRequest* r = f.first;
                if (st && !r->status) r->status = st;
                mmr_recon_frame_completion c;
                c.request_id = r->id;
                c.user_tag = r->tag;
                c.frame_index = f.second;
                c.status = st;
                c.latency_ms = std::chrono::duration<double, std::milli>(now - r->t0).count();
                s->completions.push_back(c);
                s->stats.frames_completed++;
                r->done++;
            }
            s->slot_busy[it->slot] = false;
            it = s->inflight.erase(it);
            progressed = true;
        }
        for (auto it = s->active.begin(); it != s->active.end();) {
            if ((*it)->done == (*it)->frame_count) {
                s->stats.admitted_bytes -= (*it)->hbm;
                s->stats.requests_completed++;
                it = s->active.erase(it);
                progressed = true;
            } else {
                ++it;
            }
        }
        // 2. admission by HBM budget (FIFO)
        while (!s->pending.empty()) {
            auto& r = s->pending.front();
            const uint64_t budget = s->cfg.hbm_budget_bytes;
            if (budget && !s->active.empty() && s->stats.admitted_bytes + r->hbm > budget) break;
            s->stats.admitted_bytes += r->hbm;
            s->stats.peak_admitted_bytes = std::max(s->stats.peak_admitted_bytes, s->stats.admitted_bytes);
            s->active.push_back(std::move(r));
            s->pending.pop_front();
            progressed = true;
        }
        // 3. launch batches while slots are free
        for (;;) {
            int slot = -1;
            for (size_t i = 0; i < s->slot_busy.size(); ++i)
                if (!s->slot_busy[i]) { slot = (int)i; break; }
            if (slot < 0) break;
            InflightBatch b;
            b.slot = slot;
            std::vector<ServerRequest*> reqs;
            uint32_t per_codec[5] = {0, 0, 0, 0, 0};
            const size_t n = s->active.size();
            // A stream's requests are batched in enqueue order: only its earliest request with
            // frames left is eligible.
            std::unordered_map<uint32_t, uint64_t> first_open;
            for (auto& ar : s->active)
                if (ar->stream_id && ar->next < ar->frame_count) {
                    auto f = first_open.find(ar->stream_id);
                    if (f == first_open.end() || ar->id < f->second) first_open[ar->stream_id] = ar->id;
                }
            auto start = s->active.begin();
            for (size_t k = 0; k < (rr % (n ? n : 1)); ++k) ++start;
            auto it = start;
            for (size_t visited = 0; visited < n; ++visited) {
                if (it == s->active.end()) it = s->active.begin();
                ServerRequest* r = it->get();
                ++it;
                if (r->next >= r->frame_count) continue;
                if (r->stream_id && first_open[r->stream_id] != r->id) continue;
                uint32_t take = 0;
                while (r->next < r->frame_count && take < s->cfg.max_frames_per_request_per_batch &&
                       per_codec[r->codec] < s->cfg.max_batch_frames) {
                    b.frames.push_back({r, r->next});
                    r->next++;
                    take++;
                    per_codec[r->codec]++;
                }
                if (take) reqs.push_back(r);
            }
            if (b.frames.empty()) break;
            rr++;
            // A request's first batch also waits for the last batch of its stream's previous request.
            for (ServerRequest* r : reqs)
                if (r->stream_id && !r->last) {
                    auto sl = s->stream_last.find(r->stream_id);
                    if (sl != s->stream_last.end()) r->last = sl->second;
                }
            mmr_status rc = guarded("server batch", [&]() -> mmr_status { return launch_batch(s, slot, b, reqs); });
            if (rc) {
                for (auto& f : b.frames) {
                    mmr_recon_frame_completion c{f.first->id, f.first->tag, f.second, rc, 0.0};
                    s->completions.push_back(c);
                    f.first->done++;
                    if (!f.first->status) f.first->status = rc;
                }
                continue;
            }
            for (ServerRequest* r : reqs) {
                r->last = b.ev;
                if (r->stream_id) s->stream_last[r->stream_id] = b.ev;
            }
            s->slot_busy[slot] = true;
            s->stats.batches_launched++;
            s->inflight.push_back(std::move(b));
            progressed = true;
        }
        s->stats.pending = (uint32_t)s->pending.size();
        s->stats.active = (uint32_t)s->active.size();
        if (s->pending.empty() && s->active.empty() && s->inflight.empty()) {
            s->idle_cv.notify_all();
            if (s->stop) return true;
            s->cv.wait(lk);
            return false;
        }
        if (!progressed) {
            if (s->inflight.empty()) s->cv.wait_for(lk, std::chrono::milliseconds(1));
            else s->cv.wait_for(lk, std::chrono::microseconds(20));
        }
    }
    return false;
}

// The scheduler thread. An exception inside a step (host memory exhausted) fails the frames no
// batch carries yet instead of ending the process; batches in flight still complete.
void server_loop(mmr_recon_server* s) {
    std::unique_lock<std::mutex> lk(s->mu);
    size_t rr = 0;
    for (;;) {
        try {
            if (server_step(s, lk, rr)) break;
        } catch (...) {
            fail_unplaced(s, MMR_STATUS_INTERNAL_ERROR);
            s->cv.wait_for(lk, std::chrono::milliseconds(1));
        }
    }
}

}  // namespace mmr_host

extern "C" {

mmr_status mmr_recon_server_create(mmr_recon_context* ctx, const mmr_recon_server_config* config, mmr_recon_server** out) {
    return guarded("mmr_recon_server_create", [&]() -> mmr_status {
        CHECK_CTX(ctx);
        if (!out) return fail(MMR_STATUS_INVALID_ARGUMENT, "out_server is NULL");
        *out = nullptr;
        mmr_recon_server_config cfg{};
        if (config) {
            if (config->struct_size < sizeof(mmr_recon_server_config)) return fail(MMR_STATUS_INVALID_ARGUMENT, "server config struct_size too small");
            cfg = *config;
        }
        if (!cfg.slots) cfg.slots = 4;
        if (cfg.slots > 32) return fail(MMR_STATUS_INVALID_ARGUMENT, "at most 32 slots");
        if (!cfg.max_batch_frames) cfg.max_batch_frames = 64;
        if (!cfg.max_frames_per_request_per_batch) cfg.max_frames_per_request_per_batch = 8;
        if (!cfg.max_pending) cfg.max_pending = 4096;
        auto* s = new mmr_recon_server();
        s->ctx = ctx;
        s->cfg = cfg;
        hipError_t e = hipSetDevice(ctx->device);
        for (uint32_t i = 0; i < cfg.slots && e == hipSuccess; ++i) {
            hipStream_t st = nullptr;
            e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);   // default priority
            if (e == hipSuccess) s->streams.push_back(st);
        }
        if (e != hipSuccess) {
            for (auto st : s->streams) (void)hipStreamDestroy(st);
            delete s;
            return hip_fail(e, "server streams");
        }
        s->slot_busy.assign(cfg.slots, false);
        s->stats.struct_size = sizeof(mmr_recon_server_stats);
        try {
            s->thread = std::thread(mmr_host::server_loop, s);
        } catch (...) {
            for (auto st : s->streams) (void)hipStreamDestroy(st);
            delete s;
            throw;
        }
        *out = s;
        return MMR_STATUS_OK;
    });
}

}  // extern "C"

namespace {

// Device bytes a request's descriptors write: every output plane, RGB / RGBA image and composite
// target its frames name, from their sizes. Admission counts at least this much for a request,
// whatever hbm_bytes it declares.
uint64_t request_footprint(const mmr_recon_request& req) {
    uint64_t bytes = 0;
    for (uint32_t i = 0; i < req.frame_count; ++i) {
        switch (req.codec) {
        case MMR_CODEC_MJPEG: {
            const auto& f = reinterpret_cast<const mmr_recon_mjpeg_frame*>(req.frames)[i];
            uint32_t hmax = 1, vmax = 1;
            for (uint32_t c = 0; c < f.component_count && c < 4; ++c) {
                hmax = std::max<uint32_t>(hmax, f.h_sampling[c]);
                vmax = std::max<uint32_t>(vmax, f.v_sampling[c]);
            }
            for (uint32_t c = 0; c < f.component_count && c < 4; ++c)
                bytes += (uint64_t)((f.width * f.h_sampling[c] + hmax - 1) / hmax) * ((f.height * f.v_sampling[c] + vmax - 1) / vmax);
            if (f.rgb) bytes += (uint64_t)f.width * f.height * 3;
            break;
        }
        case MMR_CODEC_MPEG2: {
            const auto& p = reinterpret_cast<const mmr_recon_mpeg2_picture*>(req.frames)[i];
            const uint64_t mbs = (uint64_t)((p.width + 15) / 16) * ((p.height + 15) / 16);
            bytes += mbs * 384;
            break;
        }
        case MMR_CODEC_DV: {
            const auto& f = reinterpret_cast<const mmr_recon_dv_frame*>(req.frames)[i];
            bytes += f.system == MMR_DV_SYSTEM_625_50 ? 720 * 576 + 2 * 360 * 288 : 720 * 480 + 2 * 180 * 480;
            break;
        }
        case MMR_CODEC_H264: {
            const auto& p = reinterpret_cast<const mmr_recon_h264_picture*>(req.frames)[i];
            bytes += (uint64_t)p.width_mbs * p.height_mbs * 384;
            break;
        }
        default:
            break;
        }
        if (req.convert) {
            const auto& c = req.convert[i];
            bytes += (uint64_t)c.width * c.height * ((c.rgb ? 3 : 0) + (c.rgba ? 4 : 0));
        }
        if (req.composite) bytes += (uint64_t)req.composite[i].stride * req.composite[i].height;
    }
    return bytes;
}

}  // namespace

extern "C" {

mmr_status mmr_recon_server_enqueue(mmr_recon_server* s, const mmr_recon_request* req, uint64_t* out_id) {
    return guarded("mmr_recon_server_enqueue", [&]() -> mmr_status {
        if (!s) return fail(MMR_STATUS_INVALID_ARGUMENT, "server is NULL");
        if (!req || req->struct_size < sizeof(mmr_recon_request)) return fail(MMR_STATUS_INVALID_ARGUMENT, "request invalid");
        const size_t sz = mmr_host::codec_struct_size(req->codec);
        if (!sz) return fail(MMR_STATUS_INVALID_ARGUMENT, "request codec %u invalid", req->codec);
        if (!req->frame_count || !req->frames) return fail(MMR_STATUS_INVALID_ARGUMENT, "request has no frames");
        auto r = std::make_unique<mmr_host::ServerRequest>();
        r->codec = req->codec;
        r->stream_id = req->stream_id;
        r->frame_count = req->frame_count;
        r->hbm = std::max<uint64_t>(req->hbm_bytes, request_footprint(*req));
        r->tag = req->user_tag;
        r->frames.assign((const char*)req->frames, (const char*)req->frames + sz * req->frame_count);
        if (req->codec == MMR_CODEC_MPEG2) {
            auto* pics = reinterpret_cast<const mmr_recon_mpeg2_picture*>(req->frames);
            for (uint32_t i = 0; i < req->frame_count; ++i)
                if (pics[i].forward_source >= (int32_t)i || pics[i].backward_source >= (int32_t)i)
                    return fail(MMR_STATUS_INVALID_ARGUMENT, "request picture %u references a later picture", i);
        }
        if (req->convert) r->convert.assign(req->convert, req->convert + req->frame_count);
        if (req->composite) r->composite.assign(req->composite, req->composite + req->frame_count);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->stop) return fail(MMR_STATUS_INVALID_STATE, "server is shutting down");
            if (s->pending.size() >= s->cfg.max_pending) return fail(MMR_STATUS_BUSY, "server queue full (%u pending)", s->cfg.max_pending);
            r->id = s->next_id++;
            r->t0 = mmr_host::Clock::now();
            if (out_id) *out_id = r->id;
            s->pending.push_back(std::move(r));
            s->stats.requests_enqueued++;
        }
        s->cv.notify_one();
        return MMR_STATUS_OK;
    });
}

mmr_status mmr_recon_server_poll(mmr_recon_server* s, mmr_recon_frame_completion* out, uint32_t capacity, uint32_t* out_count) {
    return guarded("mmr_recon_server_poll", [&]() -> mmr_status {
        if (!s || !out_count || (capacity && !out)) return fail(MMR_STATUS_INVALID_ARGUMENT, "invalid poll arguments");
        std::lock_guard<std::mutex> lk(s->mu);
        uint32_t n = 0;
        while (n < capacity && !s->completions.empty()) {
            out[n++] = s->completions.front();
            s->completions.pop_front();
        }
        *out_count = n;
        return MMR_STATUS_OK;
    });
}

mmr_status mmr_recon_server_drain(mmr_recon_server* s) {
    return guarded("mmr_recon_server_drain", [&]() -> mmr_status {
        if (!s) return fail(MMR_STATUS_INVALID_ARGUMENT, "server is NULL");
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv.notify_all();
        s->idle_cv.wait(lk, [s] { return s->pending.empty() && s->active.empty() && s->inflight.empty(); });
        return MMR_STATUS_OK;
    });
}

mmr_status mmr_recon_server_stats_get(mmr_recon_server* s, mmr_recon_server_stats* out) {
    return guarded("mmr_recon_server_stats_get", [&]() -> mmr_status {
        if (!s || !out || out->struct_size < sizeof(mmr_recon_server_stats)) return fail(MMR_STATUS_INVALID_ARGUMENT, "invalid stats arguments");
        std::lock_guard<std::mutex> lk(s->mu);
        *out = s->stats;
        out->struct_size = sizeof(mmr_recon_server_stats);
        return MMR_STATUS_OK;
    });
}

void mmr_recon_server_destroy(mmr_recon_server* s) {
    guarded_void("mmr_recon_server_destroy", [&]() {
        if (!s) return;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->pending.clear();
            s->stop = true;
        }
        s->cv.notify_all();
        s->thread.join();
        for (auto st : s->streams) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
        delete s;
    });
}

}  // extern "C"
