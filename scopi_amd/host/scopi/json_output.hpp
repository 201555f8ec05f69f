// json_output.hpp -- the output files of ScopiSolver::write_output_files
// (include/scopi/solver.hpp:369-438, objects/methods/write_objects.hpp:45-195,
// objects/neighbor.hpp:23-70, contact/property.hpp) for the host mirror.
//
// The reference builds an nlohmann::json document and prints it with `file << std::setw(4)
// << json` (or writes json::to_bson).  The layout those produce is reproduced here without
// that library:
//   * objects keep their keys sorted (nlohmann's default std::map), members and array
//     elements one per line, indent 4, "key": value, empty containers as {} / [], a key
//     assigned nothing but `{}` stays null (the reference's "contacts" with no contact);
//   * integers print as integers, doubles as the shortest decimal string that reads back to
//     the same double, in fixed notation when the decimal exponent n (value = 0.d1d2.. x 10^n)
//     satisfies -4 < n <= 15, else as d.ddde+XX (at least two exponent digits); 0 as 0.0,
//     -0 as -0.0, non-finite values as null;
//   * BSON (binary_output) per the BSON 1.1 specification as nlohmann's to_bson maps the
//     types: document 0x03, array 0x04 (keys "0", "1", ...), double 0x01, string 0x02,
//     boolean 0x08, null 0x0A, integers as int32 0x10 when they fit, else int64 0x12.
#pragma once

#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace scopi {
namespace output {

class value {
  public:
    enum kind_t { null_v, bool_v, uint_v, int_v, float_v, string_v, array_v, object_v };

    value() = default;
    value(bool b) : kind_(bool_v), b_(b) {}
    value(double d) : kind_(float_v), d_(d) {}
    value(int i) : kind_(i < 0 ? int_v : uint_v), i_(i), u_(i < 0 ? 0 : (uint64_t)i) {}
    value(long i) : kind_(i < 0 ? int_v : uint_v), i_(i), u_(i < 0 ? 0 : (uint64_t)i) {}
    value(long long i) : kind_(i < 0 ? int_v : uint_v), i_(i), u_(i < 0 ? 0 : (uint64_t)i) {}
    value(unsigned long u) : kind_(uint_v), u_(u) {}
    value(unsigned long long u) : kind_(uint_v), u_(u) {}
    value(unsigned u) : kind_(uint_v), u_(u) {}
    value(const char* s) : kind_(string_v), s_(s) {}
    value(std::string s) : kind_(string_v), s_(std::move(s)) {}
    template <class T, std::size_t N>
    value(const std::array<T, N>& a) : kind_(array_v) {
        for (const auto& x : a) arr_.emplace_back(x);
    }

    static value array() {
        value v;
        v.kind_ = array_v;
        return v;
    }
    static value object() {
        value v;
        v.kind_ = object_v;
        return v;
    }

    kind_t kind() const { return kind_; }
    // a null value becomes an object / array on first use, as nlohmann's does
    value& operator[](const std::string& key) {
        if (kind_ == null_v) kind_ = object_v;
        if (kind_ != object_v) throw std::runtime_error("json value is not an object");
        return obj_[key];
    }
    void push_back(value v) {
        if (kind_ == null_v) kind_ = array_v;
        if (kind_ != array_v) throw std::runtime_error("json value is not an array");
        arr_.push_back(std::move(v));
    }
    const std::vector<value>& elements() const { return arr_; }
    const std::map<std::string, value>& members() const { return obj_; }
    bool boolean() const { return b_; }
    uint64_t uint() const { return u_; }
    int64_t integer() const { return i_; }
    double number() const { return d_; }
    const std::string& string() const { return s_; }

  private:
    kind_t kind_ = null_v;
    bool b_ = false;
    int64_t i_ = 0;
    uint64_t u_ = 0;
    double d_ = 0.0;
    std::string s_;
    std::vector<value> arr_;
    std::map<std::string, value> obj_;
};

// The shortest round-trip decimal digits of d > 0 (std::to_chars) laid out as described above.
inline std::string format_double(double d) {
    if (!std::isfinite(d)) return "null";
    std::string out;
    if (std::signbit(d)) {
        out.push_back('-');
        d = -d;
    }
    if (d == 0.0) return out + "0.0";
    char buf[64];
    const auto res = std::to_chars(buf, buf + sizeof(buf), d, std::chars_format::scientific);
    const std::string sci(buf, res.ptr);  // d.ddde[+-]XX
    const std::size_t epos = sci.find('e');
    std::string digits;
    for (std::size_t k = 0; k < epos; ++k)
        if (sci[k] != '.') digits.push_back(sci[k]);
    const int e10 = std::atoi(sci.c_str() + epos + 1);  // value = d.ddd x 10^e10
    const int k = (int)digits.size();
    const int n = e10 + 1;  // value = 0.ddd x 10^n
    constexpr int kMinExp = -4, kMaxExp = 15;
    if (k <= n && n <= kMaxExp) {
        out += digits;
        out.append((std::size_t)(n - k), '0');
        out += ".0";
    } else if (0 < n && n <= kMaxExp) {
        out += digits.substr(0, (std::size_t)n);
        out.push_back('.');
        out += digits.substr((std::size_t)n);
    } else if (kMinExp < n && n <= 0) {
        out += "0.";
        out.append((std::size_t)(-n), '0');
        out += digits;
    } else {
        out.push_back(digits[0]);
        if (k > 1) {
            out.push_back('.');
            out += digits.substr(1);
        }
        out.push_back('e');
        int e = n - 1;
        out.push_back(e < 0 ? '-' : '+');
        if (e < 0) e = -e;
        if (e < 10) out.push_back('0');
        out += std::to_string(e);
    }
    return out;
}

inline void escape_string(const std::string& s, std::string& out) {
    out.push_back('"');
    for (unsigned char c : s) {
        switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\b': out += "\\b"; break;
            case '\f': out += "\\f"; break;
            case '\n': out += "\\n"; break;
            case '\r': out += "\\r"; break;
            case '\t': out += "\\t"; break;
            default:
                if (c < 0x20) {
                    char u[8];
                    std::snprintf(u, sizeof(u), "\\u%04x", (unsigned)c);
                    out += u;
                } else {
                    out.push_back((char)c);
                }
        }
    }
    out.push_back('"');
}

inline void dump(const value& v, std::string& out, int indent = 4, int level = 0) {
    switch (v.kind()) {
        case value::null_v: out += "null"; return;
        case value::bool_v: out += v.boolean() ? "true" : "false"; return;
        case value::uint_v: out += std::to_string(v.uint()); return;
        case value::int_v: out += std::to_string(v.integer()); return;
        case value::float_v: out += format_double(v.number()); return;
        case value::string_v: escape_string(v.string(), out); return;
        case value::array_v: {
            if (v.elements().empty()) {
                out += "[]";
                return;
            }
            out += "[\n";
            const std::string pad((std::size_t)(indent * (level + 1)), ' ');
            bool first = true;
            for (const auto& e : v.elements()) {
                if (!first) out += ",\n";
                first = false;
                out += pad;
                dump(e, out, indent, level + 1);
            }
            out.push_back('\n');
            out.append((std::size_t)(indent * level), ' ');
            out.push_back(']');
            return;
        }
        case value::object_v: {
            if (v.members().empty()) {
                out += "{}";
                return;
            }
            out += "{\n";
            const std::string pad((std::size_t)(indent * (level + 1)), ' ');
            bool first = true;
            for (const auto& kv : v.members()) {
                if (!first) out += ",\n";
                first = false;
                out += pad;
                escape_string(kv.first, out);
                out += ": ";
                dump(kv.second, out, indent, level + 1);
            }
            out.push_back('\n');
            out.append((std::size_t)(indent * level), ' ');
            out.push_back('}');
            return;
        }
    }
}

inline std::string dump(const value& v, int indent = 4) {
    std::string out;
    dump(v, out, indent, 0);
    return out;
}

namespace detail {
inline void put_le(std::vector<uint8_t>& b, uint64_t x, int bytes) {
    for (int k = 0; k < bytes; ++k) b.push_back((uint8_t)(x >> (8 * k)));
}
inline void bson_element(std::vector<uint8_t>& b, const std::string& key, const value& v);
inline void bson_document(std::vector<uint8_t>& b, const value& v) {
    const std::size_t start = b.size();
    put_le(b, 0, 4);
    if (v.kind() == value::object_v) {
        for (const auto& kv : v.members()) bson_element(b, kv.first, kv.second);
    } else {
        std::size_t k = 0;
        for (const auto& e : v.elements()) bson_element(b, std::to_string(k++), e);
    }
    b.push_back(0);
    const uint32_t size = (uint32_t)(b.size() - start);
    for (int k = 0; k < 4; ++k) b[start + k] = (uint8_t)(size >> (8 * k));
}
inline void bson_element(std::vector<uint8_t>& b, const std::string& key, const value& v) {
    auto head = [&](uint8_t type) {
        b.push_back(type);
        b.insert(b.end(), key.begin(), key.end());
        b.push_back(0);
    };
    switch (v.kind()) {
        case value::null_v: head(0x0A); return;
        case value::bool_v: head(0x08); b.push_back(v.boolean() ? 1 : 0); return;
        case value::float_v: {
            head(0x01);
            uint64_t bits;
            const double d = v.number();
            std::memcpy(&bits, &d, 8);
            put_le(b, bits, 8);
            return;
        }
        case value::uint_v:
            if (v.uint() <= (uint64_t)INT32_MAX) {
                head(0x10);
                put_le(b, v.uint(), 4);
            } else {
                if (v.uint() > (uint64_t)INT64_MAX) throw std::runtime_error("bson: integer out of range");
                head(0x12);
                put_le(b, v.uint(), 8);
            }
            return;
        case value::int_v:
            if (v.integer() >= INT32_MIN && v.integer() <= INT32_MAX) {
                head(0x10);
                put_le(b, (uint64_t)(uint32_t)(int32_t)v.integer(), 4);
            } else {
                head(0x12);
                put_le(b, (uint64_t)v.integer(), 8);
            }
            return;
        case value::string_v: {
            head(0x02);
            put_le(b, v.string().size() + 1, 4);
            b.insert(b.end(), v.string().begin(), v.string().end());
            b.push_back(0);
            return;
        }
        case value::array_v: head(0x04); bson_document(b, v); return;
        case value::object_v: head(0x03); bson_document(b, v); return;
    }
}
}  // namespace detail

// json::to_bson: the top level must be an object
inline std::vector<uint8_t> to_bson(const value& v) {
    if (v.kind() != value::object_v) throw std::runtime_error("bson: the top level must be an object");
    std::vector<uint8_t> b;
    detail::bson_document(b, v);
    return b;
}

}  // namespace output
}  // namespace scopi
