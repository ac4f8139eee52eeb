// Stored models: da_handle_save_model, da_handle_load_model, da_handle_print_model_versions
// (aoclda_handle.h:177-217) and da_print_model_metadata (aoclda_utils.h:213). Behaviour
// follows source/core/utilities/da_handle_public.cpp:294-357, da_handle.cpp:63-175,
// basic_handle.cpp:276-346 and model_persistence.cpp:184-258 of the reference: argument
// checks, the metadata block, then the option registry and the algorithm's fields; a load
// that fails after its metadata were accepted leaves the new handle for the caller to
// destroy.
#include "aoclda_handle.h"
#include "aoclda_utils.h"
#include "factories.hpp"
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <new>

using namespace dagpu;

namespace {

da_status read_file(const char *name, std::vector<char> &data) {
    std::ifstream f(name, std::ios::binary | std::ios::ate);
    if (!f.is_open())
        return da_status_io_error;
    const std::streamsize size = f.tellg();
    if (size <= 0)
        return da_status_io_error;
    f.seekg(0, std::ios::beg);
    try {
        data.resize((size_t)size);
    } catch (std::bad_alloc &) {
        return da_status_memory_error;
    }
    if (!f.read(data.data(), size))
        return da_status_io_error;
    return da_status_success;
}

} // namespace

extern "C" {

da_status da_handle_save_model(da_handle handle, const char *file_name) {
    if (handle == nullptr)
        return da_status_invalid_pointer;
    handle->err.clear();
    if (file_name == nullptr)
        return handle->err.error(da_status_invalid_pointer, "file_name cannot be null.");
    if (!handle->alg)
        return handle->err.error(da_status_handle_not_initialized,
                                 "The handle has not been initialized with a valid type.");
    AlgHandle &alg = *handle->alg;
    da_status s = alg.save_check();
    if (s != da_status_success)
        return s;
    const bool single = handle->precision == da_single;
    ModelBuffer b;
    s = b.put_metadata(handle->handle_type, single ? (da_int)sizeof(float) : (da_int)sizeof(double),
                       alg.serialization_version);
    if (s == da_status_success)
        s = alg.opts.save(b, single);
    if (s == da_status_success)
        s = alg.serialize(b);
    if (s != da_status_success)
        return handle->err.error(s, "Failure serializing model.");
    std::ofstream f(file_name, std::ios::binary | std::ios::trunc);
    if (!f.is_open())
        return handle->err.error(da_status_io_error, "Error while opening/creating save file.");
    f.write(b.bytes().data(), (std::streamsize)b.bytes().size());
    if (!f)
        return handle->err.error(da_status_io_error, "Failed to write to file.");
    return da_status_success;
}

da_status da_handle_load_model(da_handle *handle, const char *file_name) {
    if (handle == nullptr || *handle != nullptr || file_name == nullptr)
        return da_status_invalid_pointer;
    std::vector<char> data;
    da_status s = read_file(file_name, data);
    if (s != da_status_success)
        return s;
    ModelBuffer b(data.data(), data.size());
    da_handle_type type = da_handle_uninitialized;
    da_int precision = 0, version = 0;
    std::string lib_version;
    s = b.get_metadata(type, precision, version, lib_version);
    if (s != da_status_success)
        return s;
    const bool single = precision == (da_int)sizeof(float);
    s = single ? da_handle_init_s(handle, type) : da_handle_init_d(handle, type);
    if (s != da_status_success)
        return s;
    _da_handle *h = *handle;
    if (!h->alg)
        return h->err.error(da_status_invalid_file_data,
                            "The stored model does not name a handle type with a model.");
    AlgHandle &alg = *h->alg;
    if (version != alg.serialization_version)
        return h->err.error(da_status_version_mismatch,
                            "Saved model's version is incompatible with the current "
                            "library version.");
    alg.saved_version = lib_version;
    s = alg.opts.load(b, single);
    if (s != da_status_success)
        return h->err.error(s, "Failure deserializing registry.");
    s = alg.serialize(b);
    if (s != da_status_success)
        return h->err.error(s, "Failure deserializing model.");
    s = alg.restore();
    if (s != da_status_success)
        return s;
    return da_status_success;
}

da_status da_handle_print_model_versions(da_handle handle) {
    if (handle == nullptr)
        return da_status_invalid_pointer;
    if (!handle->alg || handle->alg->saved_version.empty())
        return handle->err.error(da_status_invalid_input,
                                 "The handle doesn't contain a valid loaded model.");
    std::cout << "===== Loaded Model Versions =====\n"
              << std::setw(30) << "Serialization version:" << handle->alg->serialization_version
              << "\n"
              << std::setw(30) << "Saved AOCL-DA build version:" << handle->alg->saved_version
              << "\n=================================\n"
              << std::flush;
    return da_status_success;
}

da_status da_print_model_metadata(const char *filename) {
    if (filename == nullptr)
        return da_status_invalid_pointer;
    std::vector<char> data;
    {
        std::ifstream f(filename, std::ios::binary | std::ios::ate);
        if (!f.is_open())
            return da_status_io_error;
        const std::streamsize size = f.tellg();
        if (size <= 0)
            return da_status_invalid_file_data;
        f.seekg(0, std::ios::beg);
        data.resize((size_t)size);
        if (!f.read(data.data(), size))
            return da_status_io_error;
    }
    ModelBuffer b(data.data(), data.size());
    std::string header, lib_version;
    da_int int_size = 0, version = 0, precision = 0;
    da_handle_type type = da_handle_uninitialized;
    da_status s = b.get(header);
    if (s == da_status_success)
        s = b.get(int_size);
    if (s == da_status_success)
        s = b.get(version);
    if (s == da_status_success)
        s = b.get(type);
    if (s == da_status_success)
        s = b.get(precision);
    if (s == da_status_success)
        s = b.get(lib_version);
    if (s != da_status_success)
        return s;
    std::cout << "===== Serialized Model Metadata =====\n"
              << std::setw(30) << "Header keyword:" << header << "\n"
              << std::setw(30) << "Integer size:" << (int_size == 4 ? "LP64" : "ILP64") << "\n"
              << std::setw(30) << "Serialization version:" << version << "\n"
              << std::setw(30) << "Handle type:" << (int)type << "\n"
              << std::setw(30) << "Precision: Float" << precision * 8 << "\n"
              << std::setw(30) << "Saved AOCL-DA build version:" << lib_version << "\n"
              << "=====================================\n"
              << std::flush;
    return da_status_success;
}

} // extern "C"
