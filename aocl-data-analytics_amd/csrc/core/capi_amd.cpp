// Device-memory helpers and per-handle timing of the MI355X extension API (aoclda_amd.h).
#include "aoclda_amd.h"
#include "handle.hpp"
#include <cstring>

using namespace dagpu;

static da_status map(hipError_t e) {
    if (e == hipSuccess)
        return da_status_success;
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? da_status_memory_error : da_status_internal_error;
}

extern "C" {

da_status da_amd_malloc(size_t bytes, void **d_ptr) {
    if (!d_ptr)
        return da_status_invalid_pointer;
    if (require_gpu(nullptr) != da_status_success)
        return da_status_arch_not_supported;
    return map(hipMalloc(d_ptr, bytes ? bytes : 1));
}
da_status da_amd_free(void *d_ptr) { return map(hipFree(d_ptr)); }
da_status da_amd_memcpy_htod(void *d_dst, const void *h_src, size_t bytes) {
    return map(hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice));
}
da_status da_amd_memcpy_dtoh(void *h_dst, const void *d_src, size_t bytes) {
    return map(hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost));
}
da_status da_amd_memcpy_dtod(void *d_dst, const void *d_src, size_t bytes) {
    return map(hipMemcpy(d_dst, d_src, bytes, hipMemcpyDeviceToDevice));
}
da_status da_amd_device_synchronize(void) { return map(hipDeviceSynchronize()); }
da_status da_amd_device_name(char *name, da_int len) {
    if (!name || len < 1)
        return da_status_invalid_pointer;
    if (require_gpu(nullptr) != da_status_success)
        return da_status_arch_not_supported;
    hipDeviceProp_t prop;
    da_status s = map(hipGetDeviceProperties(&prop, 0));
    if (s != da_status_success)
        return s;
    std::strncpy(name, prop.gcnArchName, (size_t)len - 1);
    name[len - 1] = '\0';
    return da_status_success;
}

da_status da_amd_synchronize(da_handle handle) {
    if (!handle)
        return da_status_handle_not_initialized;
    if (!handle->gpu.stream())
        return da_status_success;
    return map(hipStreamSynchronize(handle->gpu.stream()));
}

da_status da_amd_profile_enable(da_handle handle, da_int on) {
    if (!handle)
        return da_status_handle_not_initialized;
    handle->gpu.enable_profile(on != 0);
    return da_status_success;
}
da_status da_amd_profile_reset(da_handle handle) {
    if (!handle)
        return da_status_handle_not_initialized;
    handle->gpu.reset_profile();
    return da_status_success;
}
da_status da_amd_profile_get(da_handle handle, const char *kernel, double *total_ms,
                             da_int *count) {
    if (!handle || !kernel || !total_ms || !count)
        return da_status_invalid_pointer;
    double ms = 0;
    int c = 0;
    handle->gpu.profile_get(kernel, ms, c);
    *total_ms = ms;
    *count = c;
    return da_status_success;
}
da_status da_amd_profile_names(da_handle handle, char *names, da_int *lnames) {
    if (!handle || !lnames)
        return da_status_invalid_pointer;
    std::string all;
    for (auto &n : handle->gpu.profile_names()) {
        if (!all.empty())
            all += ",";
        all += n;
    }
    if (!names || (size_t)*lnames <= all.size()) {
        *lnames = (da_int)all.size() + 1;
        return da_status_invalid_array_dimension;
    }
    std::memcpy(names, all.c_str(), all.size() + 1);
    return da_status_success;
}

} // extern "C"
