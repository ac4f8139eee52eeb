// Typed, string-keyed option registry with bounds, mirroring the behaviour of the
// reference's OptionRegistry (source/core/utilities/options.hpp): names and string values
// are trimmed, blank-squeezed and lower-cased before lookup (options.hpp:192-201); an
// unknown name returns da_status_option_not_found, a value of the wrong kind
// da_status_option_wrong_type, a numeric value outside its bounds
// da_status_option_invalid_bounds and an unknown string value
// da_status_option_invalid_value.
#pragma once
#include "aoclda_error.h"
#include "aoclda_types.h"
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace dagpu {

class ModelBuffer;

enum class LowerBound { m_inf, greater_than, greater_equal };
enum class UpperBound { p_inf, less_than, less_equal };

std::string normalise_option_string(const std::string &s);

class OptionRegistry {
  public:
    void add_int(const std::string &name, const std::string &desc, da_int lo,
                 LowerBound lbt, da_int hi, UpperBound ubt, da_int def);
    void add_real(const std::string &name, const std::string &desc, double lo,
                  LowerBound lbt, double hi, UpperBound ubt, double def);
    void add_string(const std::string &name, const std::string &desc,
                    const std::vector<std::pair<std::string, da_int>> &choices,
                    const std::string &def);

    da_status set_int(const std::string &name, da_int v);
    da_status set_real(const std::string &name, double v);
    da_status set_string(const std::string &name, const std::string &v);

    da_status get_int(const std::string &name, da_int &v) const;
    da_status get_real(const std::string &name, double &v) const;
    da_status get_string(const std::string &name, std::string &v, da_int *key = nullptr) const;

    // Convenience for internal reads of options that are known to exist.
    da_int ival(const std::string &name) const;
    double rval(const std::string &name) const;
    da_int skey(const std::string &name) const;

    // Changes the bounds (and optionally the value) of an integer option, as the
    // reference does for "n_clusters" once the data size is known (kmeans_options.hpp:151-172).
    void rebound_int(const std::string &name, da_int hi, UpperBound ubt);

    void print() const;
    const std::string &errmsg() const { return errmsg_; }

    // Stored-model form of the registry (the reference's save_registry / load_registry,
    // source/core/utilities/options.hpp:956-1009): the number of options, then per option
    // its name and value: integers as int64, reals in the handle's precision (float when
    // `single`), strings as their selected label. Loading sets the stored values without the
    // bounds checks of a user setting; an unknown name or label, or more options than this
    // registry holds, is da_status_invalid_file_data.
    da_status save(ModelBuffer &b, bool single) const;
    da_status load(ModelBuffer &b, bool single);

  private:
    enum class Kind { Int, Real, String };
    struct Opt {
        std::string name, desc;
        Kind kind;
        da_int ival = 0, idef = 0, ilo = 0, ihi = 0;
        double rval = 0, rdef = 0, rlo = 0, rhi = 0;
        LowerBound lbt = LowerBound::m_inf;
        UpperBound ubt = UpperBound::p_inf;
        std::vector<std::pair<std::string, da_int>> choices;
        std::string sval, sdef;
    };
    const Opt *find(const std::string &name) const;
    Opt *find(const std::string &name);
    std::map<std::string, Opt> opts_;
    mutable std::string errmsg_;
};

} // namespace dagpu
