// toml_lite.hpp -- the TOML subset that code/input.toml and runs.py's toml.dump
// output use: [tables], key = value with strings, integers, floats and booleans,
// '#' comments.  Replaces toml11 (code/src/main.cu:113-170), which the reference
// links from an external checkout.
#pragma once
#include <map>
#include <stdexcept>
#include <string>

namespace ycn {

struct TomlValue {
    enum Kind { String, Integer, Float, Bool } kind = String;
    std::string s;
    long long i = 0;
    double f = 0.0;
    bool b = false;
    std::string raw;
};

class Toml {
public:
    static Toml parse_file(const std::string &path);
    static Toml parse_string(const std::string &text);

    bool has(const std::string &table, const std::string &key) const;
    const TomlValue &at(const std::string &table, const std::string &key) const;

    // toml::find<T> semantics: the value must have the requested type; a float key
    // also accepts an integer literal.  Throws std::runtime_error otherwise.
    std::string get_string(const std::string &t, const std::string &k) const;
    long long get_int(const std::string &t, const std::string &k) const;
    double get_double(const std::string &t, const std::string &k) const;
    bool get_bool(const std::string &t, const std::string &k) const;

    // optional variants: the default when the key is absent or has another type
    // (the reference's try { find<...> } catch blocks, main.cu:146-157)
    std::string opt_string(const std::string &t, const std::string &k, const std::string &d) const;
    long long opt_int(const std::string &t, const std::string &k, long long d) const;
    double opt_double(const std::string &t, const std::string &k, double d) const;
    bool opt_bool(const std::string &t, const std::string &k, bool d) const;

    const std::map<std::string, std::map<std::string, TomlValue>> &tables() const { return tables_; }

private:
    std::map<std::string, std::map<std::string, TomlValue>> tables_;
};

}  // namespace ycn
