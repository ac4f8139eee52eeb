#include "options.hpp"
#include "persistence.hpp"
#include <cctype>
#include <cstdio>
#include <sstream>

namespace dagpu {

std::string normalise_option_string(const std::string &s) {
    std::string out;
    bool pending_space = false;
    for (char c : s) {
        if (std::isspace((unsigned char)c)) {
            pending_space = !out.empty();
            continue;
        }
        if (pending_space)
            out.push_back(' ');
        pending_space = false;
        out.push_back((char)std::tolower((unsigned char)c));
    }
    return out;
}

void OptionRegistry::add_int(const std::string &name, const std::string &desc, da_int lo,
                             LowerBound lbt, da_int hi, UpperBound ubt, da_int def) {
    Opt o;
    o.name = normalise_option_string(name);
    o.desc = desc;
    o.kind = Kind::Int;
    o.ilo = lo;
    o.ihi = hi;
    o.lbt = lbt;
    o.ubt = ubt;
    o.idef = o.ival = def;
    opts_[o.name] = o;
}

void OptionRegistry::add_real(const std::string &name, const std::string &desc, double lo,
                              LowerBound lbt, double hi, UpperBound ubt, double def) {
    Opt o;
    o.name = normalise_option_string(name);
    o.desc = desc;
    o.kind = Kind::Real;
    o.rlo = lo;
    o.rhi = hi;
    o.lbt = lbt;
    o.ubt = ubt;
    o.rdef = o.rval = def;
    opts_[o.name] = o;
}

void OptionRegistry::add_string(const std::string &name, const std::string &desc,
                                const std::vector<std::pair<std::string, da_int>> &choices,
                                const std::string &def) {
    Opt o;
    o.name = normalise_option_string(name);
    o.desc = desc;
    o.kind = Kind::String;
    for (auto &c : choices)
        o.choices.push_back({normalise_option_string(c.first), c.second});
    o.sdef = o.sval = normalise_option_string(def);
    opts_[o.name] = o;
}

const OptionRegistry::Opt *OptionRegistry::find(const std::string &name) const {
    auto it = opts_.find(normalise_option_string(name));
    return it == opts_.end() ? nullptr : &it->second;
}
OptionRegistry::Opt *OptionRegistry::find(const std::string &name) {
    auto it = opts_.find(normalise_option_string(name));
    return it == opts_.end() ? nullptr : &it->second;
}

template <typename V>
static bool in_bounds(V v, V lo, LowerBound lbt, V hi, UpperBound ubt) {
    if (lbt == LowerBound::greater_than && !(v > lo))
        return false;
    if (lbt == LowerBound::greater_equal && !(v >= lo))
        return false;
    if (ubt == UpperBound::less_than && !(v < hi))
        return false;
    if (ubt == UpperBound::less_equal && !(v <= hi))
        return false;
    return true;
}

da_status OptionRegistry::set_int(const std::string &name, da_int v) {
    Opt *o = find(name);
    if (!o) {
        errmsg_ = "Option '" + name + "' not found.";
        return da_status_option_not_found;
    }
    if (o->kind != Kind::Int) {
        errmsg_ = "Option '" + name + "' is not an integer option.";
        return da_status_option_wrong_type;
    }
    if (!in_bounds<da_int>(v, o->ilo, o->lbt, o->ihi, o->ubt)) {
        errmsg_ = "Option '" + name + "': value " + std::to_string(v) + " is out of bounds.";
        return da_status_option_invalid_bounds;
    }
    o->ival = v;
    return da_status_success;
}

da_status OptionRegistry::set_real(const std::string &name, double v) {
    Opt *o = find(name);
    if (!o) {
        errmsg_ = "Option '" + name + "' not found.";
        return da_status_option_not_found;
    }
    if (o->kind != Kind::Real) {
        errmsg_ = "Option '" + name + "' is not a real option.";
        return da_status_option_wrong_type;
    }
    if (!in_bounds<double>(v, o->rlo, o->lbt, o->rhi, o->ubt)) {
        errmsg_ = "Option '" + name + "': value is out of bounds.";
        return da_status_option_invalid_bounds;
    }
    o->rval = v;
    return da_status_success;
}

da_status OptionRegistry::set_string(const std::string &name, const std::string &v) {
    Opt *o = find(name);
    if (!o) {
        errmsg_ = "Option '" + name + "' not found.";
        return da_status_option_not_found;
    }
    if (o->kind != Kind::String) {
        errmsg_ = "Option '" + name + "' is not a string option.";
        return da_status_option_wrong_type;
    }
    std::string nv = normalise_option_string(v);
    for (auto &c : o->choices) {
        if (c.first == nv) {
            o->sval = nv;
            return da_status_success;
        }
    }
    errmsg_ = "Option '" + name + "': value '" + v + "' is not valid.";
    return da_status_option_invalid_value;
}

da_status OptionRegistry::get_int(const std::string &name, da_int &v) const {
    const Opt *o = find(name);
    if (!o) {
        errmsg_ = "Option '" + name + "' not found.";
        return da_status_option_not_found;
    }
    if (o->kind != Kind::Int) {
        errmsg_ = "Option '" + name + "' is not an integer option.";
        return da_status_option_wrong_type;
    }
    v = o->ival;
    return da_status_success;
}

da_status OptionRegistry::get_real(const std::string &name, double &v) const {
    const Opt *o = find(name);
    if (!o) {
        errmsg_ = "Option '" + name + "' not found.";
        return da_status_option_not_found;
    }
    if (o->kind != Kind::Real) {
        errmsg_ = "Option '" + name + "' is not a real option.";
        return da_status_option_wrong_type;
    }
    v = o->rval;
    return da_status_success;
}

da_status OptionRegistry::get_string(const std::string &name, std::string &v,
                                     da_int *key) const {
    const Opt *o = find(name);
    if (!o) {
        errmsg_ = "Option '" + name + "' not found.";
        return da_status_option_not_found;
    }
    if (o->kind != Kind::String) {
        errmsg_ = "Option '" + name + "' is not a string option.";
        return da_status_option_wrong_type;
    }
    v = o->sval;
    if (key) {
        for (auto &c : o->choices)
            if (c.first == o->sval)
                *key = c.second;
    }
    return da_status_success;
}

da_int OptionRegistry::ival(const std::string &name) const {
    da_int v = 0;
    get_int(name, v);
    return v;
}
double OptionRegistry::rval(const std::string &name) const {
    double v = 0;
    get_real(name, v);
    return v;
}
da_int OptionRegistry::skey(const std::string &name) const {
    std::string s;
    da_int k = -1;
    get_string(name, s, &k);
    return k;
}

void OptionRegistry::rebound_int(const std::string &name, da_int hi, UpperBound ubt) {
    Opt *o = find(name);
    if (!o || o->kind != Kind::Int)
        return;
    o->ihi = hi;
    o->ubt = ubt;
}

void OptionRegistry::print() const {
    std::printf("     Option                                    Value\n");
    for (auto &kv : opts_) {
        const Opt &o = kv.second;
        std::ostringstream val;
        if (o.kind == Kind::Int)
            val << o.ival;
        else if (o.kind == Kind::Real)
            val << o.rval;
        else
            val << o.sval;
        std::printf("     %-40s  %s\n", o.name.c_str(), val.str().c_str());
    }
}

da_status OptionRegistry::save(ModelBuffer &b, bool single) const {
    da_status st = b.put((da_int)opts_.size());
    for (const auto &kv : opts_) {
        if (st != da_status_success)
            break;
        const Opt &o = kv.second;
        st = b.put(o.name);
        if (st != da_status_success)
            break;
        switch (o.kind) {
        case Kind::Int:
            st = b.put(o.ival);
            break;
        case Kind::Real:
            st = single ? b.put((float)o.rval) : b.put(o.rval);
            break;
        case Kind::String:
            st = b.put(o.sval);
            break;
        }
    }
    return st;
}

da_status OptionRegistry::load(ModelBuffer &b, bool single) {
    da_int count = 0;
    da_status st = b.get(count);
    if (st != da_status_success)
        return st;
    if (count < 0 || count > (da_int)opts_.size())
        return da_status_invalid_file_data;
    for (da_int i = 0; i < count; i++) {
        std::string name;
        st = b.get(name);
        if (st != da_status_success)
            return st;
        Opt *o = find(name);
        if (o == nullptr)
            return da_status_invalid_file_data;
        switch (o->kind) {
        case Kind::Int:
            st = b.get(o->ival);
            break;
        case Kind::Real:
            if (single) {
                float v = 0;
                st = b.get(v);
                o->rval = v;
            } else {
                st = b.get(o->rval);
            }
            break;
        case Kind::String: {
            std::string v;
            st = b.get(v);
            if (st != da_status_success)
                return st;
            bool known = false;
            for (auto &c : o->choices)
                known |= c.first == v;
            if (!known)
                return da_status_invalid_file_data;
            o->sval = v;
            break;
        }
        }
        if (st != da_status_success)
            return st;
    }
    return da_status_success;
}

} // namespace dagpu
