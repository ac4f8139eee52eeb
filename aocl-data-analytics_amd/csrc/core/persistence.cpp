// Stored-model metadata (source/core/utilities/model_persistence.cpp:108-182).
#include "persistence.hpp"

namespace dagpu {

const char *library_version() { return "AOCL-DA 5.3.2 (MI355X)"; }

da_status ModelBuffer::put_metadata(da_handle_type type, da_int precision, da_int version) {
    da_status st = put(std::string(MODEL_HEADER));
    if (st == da_status_success)
        st = put((da_int)sizeof(da_int));
    if (st == da_status_success)
        st = put(version);
    if (st == da_status_success)
        st = put(type);
    if (st == da_status_success)
        st = put(precision);
    if (st == da_status_success)
        st = put(std::string(library_version()));
    return st;
}

da_status ModelBuffer::get_metadata(da_handle_type &type, da_int &precision, da_int &version,
                                    std::string &lib_version) {
    std::string header;
    da_int int_size = 0;
    da_status st = get(header);
    if (st != da_status_success || header != MODEL_HEADER)
        return da_status_invalid_file_data;
    st = get(int_size);
    if (st == da_status_success)
        st = get(version);
    if (st == da_status_success)
        st = get(type);
    if (st == da_status_success)
        st = get(precision);
    if (st == da_status_success)
        st = get(lib_version);
    if (st != da_status_success)
        return st;
    if (version < MODEL_MIN_VERSION)
        return da_status_version_mismatch;
    // A 32-bit da_int build reads only models stored with 32-bit integers.
    if (int_size != (da_int)sizeof(da_int))
        return da_status_invalid_file_data;
    if (precision != (da_int)sizeof(float) && precision != (da_int)sizeof(double))
        return da_status_invalid_file_data;
    return da_status_success;
}

} // namespace dagpu
