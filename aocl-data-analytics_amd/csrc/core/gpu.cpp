#include "gpu.hpp"
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
