// Minimal command-line parser for the host tools: positional arguments,
// --long / -s options with a value, fixed or open-ended value counts, flags
// and defaults.  Parse errors throw std::runtime_error with a usage text.
#pragma once

#include <cstdlib>
#include <deque>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

class ArgParser {
public:
    struct Arg {
        std::vector<std::string> names;
        std::string help;
        int nargs = 1;          // values per use; -1 = one or more
        bool flag = false;      // no value; true when present
        bool required = false;
        bool positional = false;
        std::vector<std::string> def;
        std::vector<std::string> val;
        bool used = false;
    };

    explicit ArgParser(std::string prog) : prog_(std::move(prog)) {}

    Arg& add(std::vector<std::string> names, std::string help) {
        args_.push_back(Arg{});
        Arg& a = args_.back();
        a.names = std::move(names);
        a.help = std::move(help);
        a.positional = a.names[0][0] != '-';
        return a;
    }

    void parse(int argc, char** argv) {
        std::vector<std::string> t(argv + 1, argv + argc);
        size_t pos_i = 0;
        for (size_t i = 0; i < t.size(); ++i) {
            if (t[i].size() > 1 && t[i][0] == '-' && !is_number(t[i])) {
                Arg& a = find(t[i]);
                a.used = true;
                if (a.flag) continue;
                a.val.clear();
                if (a.nargs < 0) {
                    while (i + 1 < t.size() && !(t[i + 1][0] == '-' && !is_number(t[i + 1]))) a.val.push_back(t[++i]);
                    if (a.val.empty()) fail(t[i] + " needs at least one value");
                } else {
                    for (int k = 0; k < a.nargs; ++k) {
                        if (i + 1 >= t.size()) fail(a.names[0] + " needs " + std::to_string(a.nargs) + " value(s)");
                        a.val.push_back(t[++i]);
                    }
                }
            } else {
                Arg* a = nullptr;
                for (size_t k = 0, seen = 0; k < args_.size(); ++k)
                    if (args_[k].positional && seen++ == pos_i) a = &args_[k];
                if (!a) fail("unexpected argument " + t[i]);
                ++pos_i;
                a->used = true;
                a->val = {t[i]};
                for (int k = 1; k < a->nargs; ++k) {
                    if (i + 1 >= t.size()) fail(a->names[0] + " needs " + std::to_string(a->nargs) + " value(s)");
                    a->val.push_back(t[++i]);
                }
            }
        }
        for (auto& a : args_)
            if (a.required && !a.used) fail("missing required argument " + a.names[0]);
    }

    bool is_used(const std::string& name) const { return cfind(name).used; }
    std::vector<std::string> raw(const std::string& name) const {
        const Arg& a = cfind(name);
        if (a.flag) return {a.used ? "1" : "0"};
        if (a.used) return a.val;
        if (a.def.empty()) throw std::runtime_error("argument " + name + " has no value");
        return a.def;
    }
    std::string str(const std::string& name) const { return raw(name).at(0); }
    double dbl(const std::string& name) const { return std::stod(str(name)); }
    long long integer(const std::string& name) const { return std::stoll(str(name)); }
    bool flag(const std::string& name) const { return cfind(name).used; }
    std::vector<double> dbls(const std::string& name) const {
        std::vector<double> out;
        for (const auto& v : raw(name)) out.push_back(std::stod(v));
        return out;
    }

    std::string usage() const {
        std::ostringstream o;
        o << "usage: " << prog_ << "\n";
        for (const auto& a : args_) {
            o << "  ";
            for (size_t k = 0; k < a.names.size(); ++k) o << (k ? ", " : "") << a.names[k];
            o << "\t" << a.help;
            if (!a.def.empty()) o << " (default " << a.def[0] << (a.def.size() > 1 ? " ..." : "") << ")";
            if (a.required) o << " [required]";
            o << "\n";
        }
        return o.str();
    }

private:
    std::string prog_;
    std::deque<Arg> args_;  // deque: add() references stay valid

    static bool is_number(const std::string& s) {
        char* end = nullptr;
        std::strtod(s.c_str(), &end);
        return end && *end == '\0';
    }
    [[noreturn]] void fail(const std::string& m) const { throw std::runtime_error(m + "\n" + usage()); }
    Arg& find(const std::string& n) {
        for (auto& a : args_)
            for (const auto& x : a.names)
                if (x == n) return a;
        fail("unknown option " + n);
    }
    const Arg& cfind(const std::string& n) const {
        for (const auto& a : args_)
            for (const auto& x : a.names)
                if (x == n) return a;
        throw std::logic_error("undeclared argument " + n);
    }
};
