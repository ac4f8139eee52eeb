// GPU plumbing shared by every handle: status mapping for HIP errors, a growable device
// buffer, and a per-handle context holding the HIP stream and the optional per-kernel
// event timer.
#pragma once
#include <utility>
#include "aoclda_error.h"
#include "errors.hpp"
#include <hip/hip_runtime_api.h>
#include <map>
#include <string>
#include <vector>

namespace dagpu {

// Maps a HIP error onto a da_status and records a message on the error stack.
da_status hip_status(ErrorStack *err, hipError_t e, const char *what);

#define DA_HIP(err, call)                                                                 \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return ::dagpu::hip_status((err), e_, #call);                                 \
    } while (0)

// Returns da_status_success when a usable GPU is present, an error otherwise.
da_status require_gpu(ErrorStack *err);

class DeviceBuffer {
  public:
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer() { release(); }
    // Ensures at least `bytes` bytes are allocated (contents are not preserved).
    hipError_t reserve(size_t bytes);
    void release();
    void *ptr() const { return ptr_; }
    template <typename T> T *as() const { return static_cast<T *>(ptr_); }
    size_t size() const { return bytes_; }
    void swap(DeviceBuffer &o) {
        std::swap(ptr_, o.ptr_);
        std::swap(bytes_, o.bytes_);
    }

  private:
    void *ptr_ = nullptr;
    size_t bytes_ = 0;
};

// Strided copy with the arguments of hipMemcpy2DAsync. Host sides are handled with one
// contiguous transfer: rows of a strided host array are packed (or unpacked) on the host
// around a single copy, because a pitched copy from pageable memory moves one row at a time
// (seconds for millions of short rows). Strided device sides use the device's own 2-D copy.
// Packed transfers are complete when the call returns; the others follow stream order.
hipError_t copy_2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width,
                   size_t rows, hipMemcpyKind kind, hipStream_t st);

class GpuContext {
  public:
    GpuContext() = default;
    ~GpuContext();
    da_status init(ErrorStack *err);
    hipStream_t stream() const { return stream_; }

    // Per-kernel timing with HIP events recorded on this context's stream.
    void enable_profile(bool on) { profile_ = on; }
    bool profiling() const { return profile_; }
    void begin(const char *name);
    void end(const char *name);
    void reset_profile();
    // Waits for the stream and returns the accumulated milliseconds of a kernel name.
    bool profile_get(const std::string &name, double &ms, int &count);
    std::vector<std::string> profile_names();

  private:
    struct Pending {
        std::string name;
        hipEvent_t start, stop;
    };
    void collect();
    hipStream_t stream_ = nullptr;
    bool profile_ = false;
    std::vector<Pending> pending_;
    Pending current_{};
    std::map<std::string, std::pair<double, int>> totals_;
    std::vector<hipEvent_t> free_events_;
    hipEvent_t get_event();
};

// Brackets a kernel launch with profile events when profiling is on.
struct KernelScope {
    GpuContext &ctx;
    const char *name;
    KernelScope(GpuContext &c, const char *n) : ctx(c), name(n) {
        if (ctx.profiling())
            ctx.begin(name);
    }
    ~KernelScope() {
        if (ctx.profiling())
            ctx.end(name);
    }
};

} // namespace dagpu
