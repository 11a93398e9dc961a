// device_common.cpp — HIP error reporting, device probing and the scratch cache.
#include "device_common.hpp"

#include <mutex>
#include <vector>

namespace cmg::hip {

void throw_hip(hipError_t e, const char* what, const char* file, int line) {
    std::string msg = std::string("HIP error ") + hipGetErrorName(e) + " (" +
                      hipGetErrorString(e) + ") in " + what + " at " + file + ":" +
                      std::to_string(line);
    throw HipError(msg);
}

bool device_available() noexcept {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return false;
    }
    return true;
}

int cu_count() {
    static int cached[64] = {0};
    int dev = 0;
    CMG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) dev = 0;
    if (!cached[dev]) {
        int cus = 0;
        CMG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        cached[dev] = cus > 0 ? cus : 1;
    }
    return cached[dev];
}

void* Scratch::get(size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    if (bytes <= cap_) return ptr_;
    if (ptr_) {
        CMG_HIP(hipDeviceSynchronize());
        CMG_HIP(hipFree(ptr_));
        ptr_ = nullptr;
        cap_ = 0;
    }
    size_t grow = bytes + bytes / 4;
    CMG_HIP(hipMalloc(&ptr_, grow));
    cap_ = grow;
    return ptr_;
}

// The runtime may already be torn down when static destructors run, so the buffer is
// left to process exit.
Scratch::~Scratch() = default;

namespace {
std::mutex g_mu[64];
Scratch g_scratch[64];
}  // namespace

ScratchLease::ScratchLease() {
    CMG_HIP(hipGetDevice(&device));
    if (device < 0 || device >= 64) throw HipError("device ordinal out of range");
    g_mu[device].lock();
    s = &g_scratch[device];
}

ScratchLease::~ScratchLease() { g_mu[device].unlock(); }

}  // namespace cmg::hip

namespace cmg::hip {

namespace {
struct TimerState {
    bool enabled = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
};
std::mutex g_timer_mu;
std::vector<std::pair<std::string, TimerState>> g_timers;

TimerState* find_timer(const char* name, bool create) {
    for (auto& t : g_timers)
        if (t.first == name) return &t.second;
    if (!create) return nullptr;
    g_timers.emplace_back(name, TimerState{});
    return &g_timers.back().second;
}
}  // namespace

void timer_enable(const char* name, bool on) {
    std::lock_guard<std::mutex> lk(g_timer_mu);
    find_timer(name, true)->enabled = on;
}

void timer_read(const char* name, double* total_ms, uint64_t* count) {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    {
        std::lock_guard<std::mutex> lk(g_timer_mu);
        TimerState* t = find_timer(name, false);
        if (t) ev.swap(t->events);
    }
    double sum = 0;
    for (auto& e : ev) {
        CMG_HIP(hipEventSynchronize(e.second));
        float ms = 0;
        CMG_HIP(hipEventElapsedTime(&ms, e.first, e.second));
        sum += ms;
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (total_ms) *total_ms = sum;
    if (count) *count = ev.size();
}

TimedLaunch::TimedLaunch(const char* n, hipStream_t s) : name(n), stream(s), on(false) {
    // launches recorded into a graph are not timed (events in a capture would become
    // graph nodes and be read before the graph ever runs)
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    std::lock_guard<std::mutex> lk(g_timer_mu);
    TimerState* t = find_timer(name, false);
    if (!t || !t->enabled) return;
    hipEvent_t a, b;
    CMG_HIP(hipEventCreate(&a));
    CMG_HIP(hipEventCreate(&b));
    CMG_HIP(hipEventRecord(a, stream));
    t->events.emplace_back(a, b);
    on = true;
}

TimedLaunch::~TimedLaunch() {
    if (!on) return;
    std::lock_guard<std::mutex> lk(g_timer_mu);
    TimerState* t = find_timer(name, false);
    if (t && !t->events.empty()) (void)hipEventRecord(t->events.back().second, stream);
}

}  // namespace cmg::hip
