// voxelize_internal.hpp — pieces shared by the occupancy, SDF and batch TUs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "cmg_hip_internal.hpp"
#include "device_common.hpp"

namespace cmg::hip {

// Stream-ordered bump allocator over device blocks it owns. Blocks are released by the
// destructor after synchronising the stream, so every pointer stays valid until then.
class DeviceArena {
  public:
    explicit DeviceArena(hipStream_t s, size_t block = size_t(64) << 20)
        : stream_(s), block_(block) {}
    DeviceArena(const DeviceArena&) = delete;
    DeviceArena& operator=(const DeviceArena&) = delete;
    ~DeviceArena() {
        if (!blocks_.empty()) (void)hipStreamSynchronize(stream_);
        for (void* p : blocks_) (void)hipFree(p);
    }

    template <class T>
    T* alloc(size_t n) {
        size_t bytes = (n * sizeof(T) + 255) & ~size_t(255);
        if (bytes == 0) bytes = 256;
        if (used_ + bytes > cap_) {
            size_t sz = bytes > block_ ? bytes : block_;
            void* p = nullptr;
            CMG_HIP(hipMalloc(&p, sz));
            blocks_.push_back(p);
            cur_ = static_cast<char*>(p);
            cap_ = sz;
            used_ = 0;
        }
        T* out = reinterpret_cast<T*>(cur_ + used_);
        used_ += bytes;
        return out;
    }

    hipStream_t stream() const { return stream_; }

  private:
    hipStream_t stream_;
    size_t block_;
    std::vector<void*> blocks_;
    char* cur_ = nullptr;
    size_t cap_ = 0, used_ = 0;
};

struct OccPlan {
    int ntx = 0, nty = 0, words = 0;
};
OccPlan occupancy_plan(const GridSpec& g);

}  // namespace cmg::hip
