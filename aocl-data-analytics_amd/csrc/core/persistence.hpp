// Model persistence: the stored-model byte format of the reference
// (source/core/utilities/model_persistence.hpp:47-209, model_persistence.cpp:52-560).
//
// A stored model is a flat little-endian byte stream:
//   metadata: "AOCLDA_STORED_MODEL", sizeof(da_int), the serialization version, the handle
//             type, the precision (sizeof(float) or sizeof(double)), the library version;
//   the option registry: the number of options, then per option its name and value;
//   the algorithm's own fields, in the order its serialize() visits them.
// Scalars are stored in fixed-width types: bool as uint8, da_int and enums as int64,
// float and double as themselves. A container (std::vector, std::string) is its element
// count as int64 followed by its elements, each stored by the same rules.
#pragma once
#include "aoclda_error.h"
#include "aoclda_handle.h"
#include "aoclda_types.h"
#include <cstdint>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

namespace dagpu {

// Oldest serialization version a stored model may carry (model_persistence.hpp:52) and the
// version this library writes.
constexpr da_int MODEL_MIN_VERSION = 50301;
constexpr const char *MODEL_HEADER = "AOCLDA_STORED_MODEL";
// Library version recorded in stored models.
const char *library_version();

class ModelBuffer {
  public:
    enum class Mode { Save, Load };

    // A buffer that collects a model being saved.
    ModelBuffer() : mode_(Mode::Save) {}
    // A buffer over the bytes of a stored model being loaded.
    ModelBuffer(const char *data, size_t size) : mode_(Mode::Load), in_(data), size_(size) {}

    Mode mode() const { return mode_; }
    bool loading() const { return mode_ == Mode::Load; }
    const std::vector<char> &bytes() const { return out_; }
    size_t offset() const { return off_; }

    // Stores (Save) or reads back (Load) one value, scalar or container.
    template <typename V> da_status io(V &v) {
        if (mode_ == Mode::Save)
            return put(v);
        return get(v);
    }
    template <typename V> da_status put(const V &v);
    template <typename V> da_status get(V &v);

    // Metadata block (model_persistence.cpp:108-182). On load, checks the header keyword,
    // the integer size (int64 files are refused by a 32-bit da_int build), the version and
    // the precision; the handle type and versions are returned.
    da_status put_metadata(da_handle_type type, da_int precision, da_int version);
    da_status get_metadata(da_handle_type &type, da_int &precision, da_int &version,
                           std::string &lib_version);

  private:
    template <typename V> struct is_vector : std::false_type {};
    template <typename U> struct is_vector<std::vector<U>> : std::true_type {};
    template <typename V>
    static constexpr bool is_scalar_v = std::is_arithmetic_v<V> || std::is_enum_v<V>;
    template <typename V>
    using stored_t = std::conditional_t<
        std::is_same_v<V, bool>, uint8_t,
        std::conditional_t<std::is_integral_v<V> && !std::is_same_v<V, char>, int64_t,
                           std::conditional_t<std::is_enum_v<V>, int64_t, V>>>;

    da_status raw_put(const void *p, size_t n) {
        const char *c = static_cast<const char *>(p);
        out_.insert(out_.end(), c, c + n);
        return da_status_success;
    }
    da_status raw_get(void *p, size_t n) {
        if (off_ > size_ || size_ - off_ < n)
            return da_status_invalid_file_data;
        std::memcpy(p, in_ + off_, n);
        off_ += n;
        return da_status_success;
    }

    Mode mode_;
    std::vector<char> out_;
    const char *in_ = nullptr;
    size_t size_ = 0, off_ = 0;
};

template <typename V> da_status ModelBuffer::put(const V &v) {
    if constexpr (is_scalar_v<V>) {
        const stored_t<V> s = static_cast<stored_t<V>>(v);
        return raw_put(&s, sizeof(s));
    } else if constexpr (std::is_same_v<V, std::string> || is_vector<V>::value) {
        const int64_t n = (int64_t)v.size();
        da_status st = raw_put(&n, sizeof(n));
        for (size_t i = 0; i < v.size() && st == da_status_success; i++)
            st = put(v[i]);
        return st;
    } else {
        static_assert(is_scalar_v<V>, "unsupported stored type");
    }
}

template <typename V> da_status ModelBuffer::get(V &v) {
    if constexpr (is_scalar_v<V>) {
        stored_t<V> s;
        da_status st = raw_get(&s, sizeof(s));
        if (st == da_status_success)
            v = static_cast<V>(s);
        return st;
    } else if constexpr (std::is_same_v<V, std::string> || is_vector<V>::value) {
        int64_t n = 0;
        da_status st = raw_get(&n, sizeof(n));
        if (st != da_status_success)
            return st;
        // Every element takes at least one byte: a count beyond the remaining bytes is corrupt.
        if (n < 0 || (uint64_t)n > size_ - off_)
            return da_status_invalid_file_data;
        try {
            v.resize((size_t)n);
        } catch (...) {
            return da_status_memory_error;
        }
        for (int64_t i = 0; i < n && st == da_status_success; i++)
            st = get(v[(size_t)i]);
        return st;
    } else {
        static_assert(is_scalar_v<V>, "unsupported stored type");
    }
}

} // namespace dagpu
