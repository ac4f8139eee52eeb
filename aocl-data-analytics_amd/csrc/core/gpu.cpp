#include "gpu.hpp"
#include <cstring>
#include <cstdio>

namespace dagpu {

da_status hip_status(ErrorStack *err, hipError_t e, const char *what) {
    da_status s = (e == hipErrorOutOfMemory) ? da_status_memory_error : da_status_internal_error;
    std::string msg = std::string("HIP error in ") + what + ": " + hipGetErrorString(e);
    if (err)
        err->error(s, msg);
    return s;
}

da_status require_gpu(ErrorStack *err) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) {
        std::string msg = "No AMD GPU is available: this library computes on the GPU only.";
        if (err)
            err->error(da_status_arch_not_supported, msg);
        return da_status_arch_not_supported;
    }
    return da_status_success;
}

hipError_t copy_2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width,
                   size_t rows, hipMemcpyKind kind, hipStream_t st) {
    if (rows == 0 || width == 0)
        return hipSuccess;
    if (kind == hipMemcpyDeviceToDevice)
        return hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, kind, st);
    const bool dc = dpitch == width || rows == 1, sc = spitch == width || rows == 1;
    const size_t bytes = width * rows;
    if (dc && sc)
        return hipMemcpyAsync(dst, src, bytes, kind, st);
    hipError_t e = hipSuccess;
    if (kind == hipMemcpyHostToDevice) {
        const char *h = static_cast<const char *>(src);
        std::vector<char> pack;
        if (!sc) {
            pack.resize(bytes);
            for (size_t r = 0; r < rows; r++)
                std::memcpy(pack.data() + r * width, h + r * spitch, width);
            h = pack.data();
        }
        if (dc) {
            e = hipMemcpyAsync(dst, h, bytes, kind, st);
        } else {
            DeviceBuffer stage;
            e = stage.reserve(bytes);
            if (e == hipSuccess)
                e = hipMemcpyAsync(stage.ptr(), h, bytes, kind, st);
            if (e == hipSuccess)
                e = hipMemcpy2DAsync(dst, dpitch, stage.ptr(), width, width, rows,
                                     hipMemcpyDeviceToDevice, st);
            if (e == hipSuccess)
                e = hipStreamSynchronize(st);
            return e;
        }
        if (e == hipSuccess && !sc)
            e = hipStreamSynchronize(st);
        return e;
    }
    // Device to host.
    DeviceBuffer stage;
    const void *d = src;
    if (!sc) {
        e = stage.reserve(bytes);
        if (e == hipSuccess)
            e = hipMemcpy2DAsync(stage.ptr(), width, src, spitch, width, rows,
                                 hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess)
            return e;
        d = stage.ptr();
    }
    if (dc) {
        e = hipMemcpyAsync(dst, d, bytes, kind, st);
        if (e == hipSuccess && !sc)
            e = hipStreamSynchronize(st);
        return e;
    }
    std::vector<char> buf(bytes);
    e = hipMemcpyAsync(buf.data(), d, bytes, kind, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return e;
    char *h = static_cast<char *>(dst);
    for (size_t r = 0; r < rows; r++)
        std::memcpy(h + r * dpitch, buf.data() + r * width, width);
    return hipSuccess;
}

hipError_t DeviceBuffer::reserve(size_t bytes) {
    if (bytes <= bytes_ && ptr_)
        return hipSuccess;
    release();
    if (bytes == 0)
        bytes = 256;
    hipError_t e = hipMalloc(&ptr_, bytes);
    if (e != hipSuccess) {
        ptr_ = nullptr;
        bytes_ = 0;
        (void)hipGetLastError();
        return e;
    }
    bytes_ = bytes;
    return hipSuccess;
}

void DeviceBuffer::release() {
    if (ptr_)
        (void)hipFree(ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
}

GpuContext::~GpuContext() {
    if (stream_) {
        (void)hipStreamSynchronize(stream_);
        for (auto &p : pending_) {
            (void)hipEventDestroy(p.start);
            (void)hipEventDestroy(p.stop);
        }
        for (auto ev : free_events_)
            (void)hipEventDestroy(ev);
        (void)hipStreamDestroy(stream_);
    }
}

da_status GpuContext::init(ErrorStack *err) {
    if (stream_)
        return da_status_success;
    da_status s = require_gpu(err);
    if (s != da_status_success)
        return s;
    DA_HIP(err, hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    return da_status_success;
}

hipEvent_t GpuContext::get_event() {
    if (!free_events_.empty()) {
        hipEvent_t e = free_events_.back();
        free_events_.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

void GpuContext::begin(const char *name) {
    current_.name = name;
    current_.start = get_event();
    current_.stop = get_event();
    (void)hipEventRecord(current_.start, stream_);
}

void GpuContext::end(const char *name) {
    (void)name;
    (void)hipEventRecord(current_.stop, stream_);
    pending_.push_back(current_);
    if (pending_.size() > 4096)
        collect();
}

void GpuContext::collect() {
    if (pending_.empty())
        return;
    (void)hipStreamSynchronize(stream_);
    for (auto &p : pending_) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
            auto &t = totals_[p.name];
            t.first += ms;
            t.second += 1;
        }
        free_events_.push_back(p.start);
        free_events_.push_back(p.stop);
    }
    pending_.clear();
}

void GpuContext::reset_profile() {
    collect();
    totals_.clear();
}

bool GpuContext::profile_get(const std::string &name, double &ms, int &count) {
    collect();
    auto it = totals_.find(name);
    if (it == totals_.end()) {
        ms = 0;
        count = 0;
        return false;
    }
    ms = it->second.first;
    count = it->second.second;
    return true;
}

std::vector<std::string> GpuContext::profile_names() {
    collect();
    std::vector<std::string> out;
    for (auto &kv : totals_)
        out.push_back(kv.first);
    return out;
}

} // namespace dagpu
