// toml_lite.cpp -- see toml_lite.hpp.
#include "toml_lite.hpp"

#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace ycn {

static std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

// strip a trailing comment that is not inside a string
static std::string strip_comment(const std::string &s) {
    bool in_str = false;
    char q = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        const char c = s[i];
        if (in_str) {
            if (c == '\\' && q == '"') { ++i; continue; }
            if (c == q) in_str = false;
        } else if (c == '"' || c == '\'') {
            in_str = true;
            q = c;
        } else if (c == '#') {
            return s.substr(0, i);
        }
    }
    return s;
}

static TomlValue parse_value(const std::string &v, int line) {
    TomlValue out;
    out.raw = v;
    if (v.empty()) throw std::runtime_error("toml: empty value on line " + std::to_string(line));
    if (v[0] == '"' || v[0] == '\'') {
        const char q = v[0];
        if (v.size() < 2 || v.back() != q) throw std::runtime_error("toml: bad string on line " + std::to_string(line));
        std::string s;
        for (size_t i = 1; i + 1 < v.size(); ++i) {
            if (q == '"' && v[i] == '\\' && i + 2 < v.size()) {
                const char e = v[++i];
                s += (e == 'n') ? '\n' : (e == 't') ? '\t' : e;
            } else {
                s += v[i];
            }
        }
        out.kind = TomlValue::String;
        out.s = s;
        return out;
    }
    if (v == "true" || v == "false") {
        out.kind = TomlValue::Bool;
        out.b = (v == "true");
        return out;
    }
    std::string num;
    for (char c : v)
        if (c != '_') num += c;
    const bool is_float = num.find_first_of(".eE") != std::string::npos || num == "inf" || num == "+inf" ||
                          num == "-inf" || num == "nan";
    errno = 0;
    char *end = nullptr;
    if (is_float) {
        out.kind = TomlValue::Float;
        out.f = std::strtod(num.c_str(), &end);
    } else {
        out.kind = TomlValue::Integer;
        out.i = std::strtoll(num.c_str(), &end, 0);
        out.f = (double)out.i;
    }
    if (!end || *end != '\0' || errno)
        throw std::runtime_error("toml: cannot parse value '" + v + "' on line " + std::to_string(line));
    return out;
}

Toml Toml::parse_string(const std::string &text) {
    Toml t;
    std::istringstream in(text);
    std::string line, table;
    int ln = 0;
    while (std::getline(in, line)) {
        ++ln;
        const std::string s = trim(strip_comment(line));
        if (s.empty()) continue;
        if (s[0] == '[') {
            if (s.back() != ']' || s.size() < 3) throw std::runtime_error("toml: bad table on line " + std::to_string(ln));
            table = trim(s.substr(1, s.size() - 2));
            t.tables_[table];
            continue;
        }
        const size_t eq = s.find('=');
        if (eq == std::string::npos) throw std::runtime_error("toml: expected key = value on line " + std::to_string(ln));
        std::string key = trim(s.substr(0, eq));
        if (key.size() >= 2 && (key[0] == '"' || key[0] == '\'')) key = key.substr(1, key.size() - 2);
        t.tables_[table][key] = parse_value(trim(s.substr(eq + 1)), ln);
    }
    return t;
}

Toml Toml::parse_file(const std::string &path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("toml: cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse_string(ss.str());
}

bool Toml::has(const std::string &table, const std::string &key) const {
    auto it = tables_.find(table);
    return it != tables_.end() && it->second.count(key);
}

const TomlValue &Toml::at(const std::string &table, const std::string &key) const {
    auto it = tables_.find(table);
    if (it == tables_.end()) throw std::runtime_error("toml: no table [" + table + "]");
    auto jt = it->second.find(key);
    if (jt == it->second.end()) throw std::runtime_error("toml: no key " + table + "." + key);
    return jt->second;
}

std::string Toml::get_string(const std::string &t, const std::string &k) const {
    const TomlValue &v = at(t, k);
    if (v.kind != TomlValue::String) throw std::runtime_error("toml: " + t + "." + k + " is not a string");
    return v.s;
}
long long Toml::get_int(const std::string &t, const std::string &k) const {
    const TomlValue &v = at(t, k);
    if (v.kind != TomlValue::Integer) throw std::runtime_error("toml: " + t + "." + k + " is not an integer");
    return v.i;
}
double Toml::get_double(const std::string &t, const std::string &k) const {
    const TomlValue &v = at(t, k);
    if (v.kind != TomlValue::Float && v.kind != TomlValue::Integer)
        throw std::runtime_error("toml: " + t + "." + k + " is not a number");
    return v.f;
}
bool Toml::get_bool(const std::string &t, const std::string &k) const {
    const TomlValue &v = at(t, k);
    if (v.kind != TomlValue::Bool) throw std::runtime_error("toml: " + t + "." + k + " is not a boolean");
    return v.b;
}
std::string Toml::opt_string(const std::string &t, const std::string &k, const std::string &d) const {
    try { return get_string(t, k); } catch (const std::exception &) { return d; }
}
long long Toml::opt_int(const std::string &t, const std::string &k, long long d) const {
    try { return get_int(t, k); } catch (const std::exception &) { return d; }
}
double Toml::opt_double(const std::string &t, const std::string &k, double d) const {
    try { return get_double(t, k); } catch (const std::exception &) { return d; }
}
bool Toml::opt_bool(const std::string &t, const std::string &k, bool d) const {
    try { return get_bool(t, k); } catch (const std::exception &) { return d; }
}

}  // namespace ycn
