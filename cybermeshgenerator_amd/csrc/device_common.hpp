// device_common.hpp — HIP error handling, launch helpers and the per-device scratch
// cache shared by every kernel family of libcmg_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace cmg::hip {

// A HIP runtime failure; the C ABI maps it to CMG_ERR_INTERNAL with what() in errbuf.
struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Input rejected before any device work (maps to CMG_ERR_INVALID_INPUT).
struct InvalidInput : std::runtime_error {
    using std::runtime_error::runtime_error;
};

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line);

#define CMG_HIP(call)                                                            \
    do {                                                                         \
        hipError_t _e = (call);                                                  \
        if (_e != hipSuccess) ::cmg::hip::throw_hip(_e, #call, __FILE__, __LINE__); \
    } while (0)

// Check the last launch (launch-configuration errors are reported here).
#define CMG_HIP_LAUNCH() CMG_HIP(hipGetLastError())

// True once a usable gfx9xx device is visible to this process.
bool device_available() noexcept;

// Number of compute units of the current device (cached).
int cu_count();

// Grid size for a grid-stride kernel over `work` items with `block` threads per
// block: enough blocks to fill every CU `waves_per_cu` times, never more than needed.
inline unsigned grid_for(size_t work, unsigned block, unsigned waves_per_cu = 8) {
    size_t need = (work + block - 1) / block;
    size_t fill = (size_t)cu_count() * waves_per_cu * 64 / block;
    if (fill < 1) fill = 1;
    size_t g = need < fill ? need : fill;
    return g ? (unsigned)g : 1u;
}

// Workgroup index for grid-stride loops whose neighbouring items share data (queries
// sorted by cell): workgroups are dispatched round-robin over the 8 XCDs (workgroup b to
// XCD b % 8), each XCD with its own L2, so b is renumbered to (b % 8) * (G / 8) + b / 8
// when the grid G is a multiple of 8. In every round of the stride the workgroups of one
// XCD then cover one contiguous range of items instead of every eighth block of 256. A
// bijection of [0, G), so every item is still visited exactly once.
__device__ inline long long xcd_block() {
    const unsigned G = gridDim.x, b = blockIdx.x;
    return (G % 8u) ? (long long)b : (long long)(b % 8u) * (G / 8u) + b / 8u;
}

// Growable device scratch owned by the library, one per device. Calls that use it
// serialise on its mutex and synchronise their stream before releasing it.
class Scratch {
  public:
    // Device pointer to at least `bytes` bytes (256-byte aligned); grows as needed.
    void* get(size_t bytes);
    ~Scratch();

  private:
    void* ptr_ = nullptr;
    size_t cap_ = 0;
};

// Locks and returns the current device's scratch.
struct ScratchLease {
    ScratchLease();
    ~ScratchLease();
    Scratch* s;
    int device;
};

}  // namespace cmg::hip

namespace cmg::hip {

// Optional per-kernel timing with HIP events recorded on the launch stream around named
// kernels (enabled per name through cmg_hip_kernel_timer). The benchmark uses it to time
// the dominant kernel of a call on the stream that kernel runs on.
void timer_enable(const char* name, bool on);
// Total milliseconds and launch count of the timed launches of `name` so far
// (synchronises on the recorded events); resets the record.
void timer_read(const char* name, double* total_ms, uint64_t* count);

// RAII helper: times the launches made during its lifetime if `name` is enabled.
struct TimedLaunch {
    TimedLaunch(const char* name, hipStream_t s);
    ~TimedLaunch();
    const char* name;
    hipStream_t stream;
    bool on;
};

}  // namespace cmg::hip
