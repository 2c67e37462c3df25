// Interface between the HIP runtime model (hip_model.cpp) and the kernel address models
// (kernel_model.cpp). Test infrastructure only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

namespace model {

struct Stats {
    uint64_t kernels = 0, checks = 0, ops_run = 0, jobs = 0, allocations = 0, events = 0;
};

std::recursive_mutex& mutex();
// Records a violation (printf-style message).
void violation(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
// True when [p, p + bytes) lies inside one live allocation; otherwise records a violation.
bool check(const void* p, size_t bytes, const char* what);
// Identity of the bytes of the allocation holding p: changes whenever a copy or fill writes it
// (so per-buffer scans can be cached).
uint64_t content_key(const void* p);
// Queues a kernel on stream s; body runs as the stream operation.
void enqueue(hipStream_t s, const char* name, std::function<void()> body);
Stats& stats();
const std::vector<std::string>& violations();
size_t violation_count();

}  // namespace model
