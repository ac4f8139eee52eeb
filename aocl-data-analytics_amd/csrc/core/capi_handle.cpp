// C entry points for handles, options and results (aoclda_handle.h, aoclda_options.h,
// aoclda_result.h). Behaviour follows source/core/utilities/da_handle_public.cpp:30-275
// and options_public.cpp:30-140 of the reference.
#include "aoclda_handle.h"
#include "aoclda_options.h"
#include "aoclda_result.h"
#include "factories.hpp"
#include <cstdio>
#include <cstring>
#include <new>

using namespace dagpu;

template <typename T> static da_status handle_init(da_handle *handle, da_handle_type type) {
    if (handle == nullptr)
        return da_status_invalid_pointer;
    try {
        *handle = new _da_handle;
    } catch (std::bad_alloc &) {
        return da_status_memory_error;
    }
    _da_handle *h = *handle;
    h->handle_type = type;
    h->precision = std::is_same_v<T, double> ? da_double : da_single;
    try {
        switch (type) {
        case da_handle_uninitialized:
            return da_status_success;
        case da_handle_nn:
            h->alg = make_nn_handle(h->precision, h->err, h->gpu);
            break;
        case da_handle_kmeans:
            h->alg = make_kmeans_handle(h->precision, h->err, h->gpu);
            break;
        case da_handle_dbscan:
            h->alg = make_dbscan_handle(h->precision, h->err, h->gpu);
            break;
        case da_handle_approx_nn:
            h->alg = make_approx_nn_handle(h->precision, h->err, h->gpu);
            break;
        default:
            return h->err.error(da_status_not_implemented,
                                "This handle type is not implemented by the MI355X build: "
                                "only da_handle_nn, da_handle_kmeans, da_handle_dbscan and "
                                "da_handle_approx_nn are available.");
        }
    } catch (std::bad_alloc &) {
        return h->err.error(da_status_memory_error, "Memory allocation error");
    }
    return da_status_success;
}

extern "C" {

da_status da_handle_init_d(da_handle *handle, da_handle_type handle_type) {
    return handle_init<double>(handle, handle_type);
}
da_status da_handle_init_s(da_handle *handle, da_handle_type handle_type) {
    return handle_init<float>(handle, handle_type);
}

da_status da_handle_print_error_message(da_handle handle) {
    if (!handle)
        return da_status_invalid_input;
    std::string t = handle->err.text();
    if (!t.empty())
        std::printf("%s\n", t.c_str());
    return da_status_success;
}

void da_handle_destroy(da_handle *handle) {
    if (handle && *handle) {
        delete *handle;
        *handle = nullptr;
    }
}

da_status da_handle_get_error_message(da_handle handle, char **message) {
    if (!handle || !message)
        return da_status_invalid_input;
    std::string t = handle->err.text();
    *message = (char *)std::malloc(t.size() + 1);
    if (!*message)
        return da_status_memory_error;
    std::memcpy(*message, t.c_str(), t.size() + 1);
    return da_status_success;
}

da_status da_handle_get_error_severity(da_handle handle, da_severity *severity) {
    if (!handle || !severity)
        return da_status_invalid_input;
    *severity = handle->err.severity();
    return da_status_success;
}

void da_handle_refresh(da_handle handle) {
    if (handle && handle->alg)
        handle->alg->refresh();
}

} // extern "C"

static da_status get_opts(da_handle handle, OptionRegistry **opts) {
    if (!handle->alg)
        return handle->err.error(da_status_handle_not_initialized,
                                 "The handle has not been initialized with a valid type.");
    *opts = &handle->alg->opts;
    return da_status_success;
}

extern "C" {

da_status da_options_set_int(da_handle handle, const char *option, da_int value) {
    if (!handle)
        return da_status_handle_not_initialized;
    handle->err.clear();
    OptionRegistry *o;
    da_status s = get_opts(handle, &o);
    if (s != da_status_success)
        return s;
    if (!option)
        return handle->err.error(da_status_invalid_pointer, "The option name is null.");
    s = o->set_int(option, value);
    if (s != da_status_success)
        return handle->err.error(s, o->errmsg());
    handle->alg->refresh();
    return da_status_success;
}

da_status da_options_set_string(da_handle handle, const char *option, const char *value) {
    if (!handle)
        return da_status_handle_not_initialized;
    handle->err.clear();
    OptionRegistry *o;
    da_status s = get_opts(handle, &o);
    if (s != da_status_success)
        return s;
    if (!option || !value)
        return handle->err.error(da_status_invalid_pointer, "The option name or value is null.");
    s = o->set_string(option, value);
    if (s != da_status_success)
        return handle->err.error(s, o->errmsg());
    handle->alg->refresh();
    return da_status_success;
}

} // extern "C"

template <typename T> static da_status set_real(da_handle handle, const char *option, T value) {
    if (!handle)
        return da_status_handle_not_initialized;
    handle->err.clear();
    da_status s = handle->check_precision<T>();
    if (s != da_status_success)
        return s;
    OptionRegistry *o;
    s = get_opts(handle, &o);
    if (s != da_status_success)
        return s;
    if (!option)
        return handle->err.error(da_status_invalid_pointer, "The option name is null.");
    s = o->set_real(option, (double)value);
    if (s != da_status_success)
        return handle->err.error(s, o->errmsg());
    handle->alg->refresh();
    return da_status_success;
}

extern "C" {

da_status da_options_set_real_s(da_handle handle, const char *option, float value) {
    return set_real<float>(handle, option, value);
}
da_status da_options_set_real_d(da_handle handle, const char *option, double value) {
    return set_real<double>(handle, option, value);
}

da_status da_options_get_int(da_handle handle, const char *option, da_int *value) {
    if (!handle)
        return da_status_handle_not_initialized;
    handle->err.clear();
    OptionRegistry *o;
    da_status s = get_opts(handle, &o);
    if (s != da_status_success)
        return s;
    if (!option || !value)
        return handle->err.error(da_status_invalid_pointer, "Null argument.");
    s = o->get_int(option, *value);
    if (s != da_status_success)
        return handle->err.error(s, o->errmsg());
    return da_status_success;
}

} // extern "C"

template <typename T> static da_status get_real(da_handle handle, const char *option, T *value) {
    if (!handle)
        return da_status_handle_not_initialized;
    handle->err.clear();
    da_status s = handle->check_precision<T>();
    if (s != da_status_success)
        return s;
    OptionRegistry *o;
    s = get_opts(handle, &o);
    if (s != da_status_success)
        return s;
    if (!option || !value)
        return handle->err.error(da_status_invalid_pointer, "Null argument.");
    double v = 0;
    s = o->get_real(option, v);
    if (s != da_status_success)
        return handle->err.error(s, o->errmsg());
    *value = (T)v;
    return da_status_success;
}

extern "C" {

da_status da_options_get_real_s(da_handle handle, const char *option, float *value) {
    return get_real<float>(handle, option, value);
}
da_status da_options_get_real_d(da_handle handle, const char *option, double *value) {
    return get_real<double>(handle, option, value);
}

da_status da_options_get_string_key(da_handle handle, const char *option, char *value,
                                    da_int *lvalue, da_int *key) {
    if (!handle)
        return da_status_handle_not_initialized;
    handle->err.clear();
    OptionRegistry *o;
    da_status s = get_opts(handle, &o);
    if (s != da_status_success)
        return s;
    if (!option || !value || !lvalue)
        return handle->err.error(da_status_invalid_pointer, "Null argument.");
    std::string sv;
    s = o->get_string(option, sv, key);
    if (s != da_status_success)
        return handle->err.error(s, o->errmsg());
    size_t n = sv.size();
    if (n >= (size_t)(*lvalue)) {
        *lvalue = (da_int)(n + 1);
        return handle->err.error(da_status_invalid_input,
                                 "target storage where to store option string value is too "
                                 "small, make it at least " +
                                     std::to_string(n + 1) + " characters long");
    }
    std::memcpy(value, sv.c_str(), n + 1);
    return da_status_success;
}

da_status da_options_get_string(da_handle handle, const char *option, char *value,
                                da_int *lvalue) {
    return da_options_get_string_key(handle, option, value, lvalue, nullptr);
}

da_status da_options_print(da_handle handle) {
    if (!handle)
        return da_status_handle_not_initialized;
    OptionRegistry *o;
    da_status s = get_opts(handle, &o);
    if (s != da_status_success)
        return s;
    o->print();
    handle->err.clear();
    return da_status_success;
}

} // extern "C"

template <typename T>
static da_status get_result(da_handle handle, da_result query, da_int *dim, T *result) {
    if (!handle)
        return da_status_handle_not_initialized;
    handle->err.clear();
    if (dim == nullptr)
        return handle->err.error(da_status_invalid_input, "dim has not been defined");
    if (result == nullptr)
        return handle->err.error(da_status_invalid_input,
                                 "The result array has not been allocated");
    if constexpr (!std::is_same_v<T, da_int>) {
        da_status s = handle->check_precision<T>();
        if (s != da_status_success)
            return s;
    }
    if (!handle->alg)
        return handle->err.error(da_status_handle_not_initialized,
                                 "The handle does not have any results to export. Have you "
                                 "initialized the handle and performed any calculation?");
    return handle->alg->get_result(query, dim, result);
}

extern "C" {

da_status da_handle_get_result_d(const da_handle handle, da_result query, da_int *dim,
                                 double *result) {
    return get_result(handle, query, dim, result);
}
da_status da_handle_get_result_s(const da_handle handle, da_result query, da_int *dim,
                                 float *result) {
    return get_result(handle, query, dim, result);
}
da_status da_handle_get_result_int(const da_handle handle, da_result query, da_int *dim,
                                   da_int *result) {
    return get_result(handle, query, dim, result);
}

} // extern "C"
