// Error stack carried by every handle. A call clears it on entry; errors and warnings
// are appended with their message, following the record/trace behaviour of
// source/core/utilities/da_error.hpp:150-360 in the reference (action DA_RECORD).
#pragma once
#include "aoclda_error.h"
#include <string>
#include <vector>

namespace dagpu {

class ErrorStack {
  public:
    void clear() {
        entries_.clear();
        severity_ = DA_NOTSET;
        status_ = da_status_success;
    }
    da_status error(da_status s, const std::string &msg) {
        entries_.push_back({s, msg});
        severity_ = DA_ERROR;
        status_ = s;
        return s;
    }
    da_status warn(da_status s, const std::string &msg) {
        entries_.push_back({s, msg});
        if (severity_ != DA_ERROR)
            severity_ = DA_WARNING;
        status_ = s;
        return s;
    }
    da_status status() const { return status_; }
    da_severity severity() const { return severity_; }
    std::string text() const {
        std::string out;
        for (const auto &e : entries_) {
            if (!out.empty())
                out += "\n";
            out += e.msg;
        }
        return out;
    }

  private:
    struct Entry {
        da_status status;
        std::string msg;
    };
    std::vector<Entry> entries_;
    da_severity severity_ = DA_NOTSET;
    da_status status_ = da_status_success;
};

} // namespace dagpu
