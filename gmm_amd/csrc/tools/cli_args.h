// Minimal command-line parser for the trainGMM / getLogLike tools: long
// options (--name VALUE, --name=VALUE), short aliases (-x VALUE), boolean
// switches with an implicit true value (--mean, --mean=0), repeatable
// options (-i a -i b), as the reference's boost::program_options setup.
#pragma once

#include <cstdlib>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace gmm {

class CliArgs {
public:
    void option(const std::string &name, char alias, bool is_switch, const std::string &def = "") {
        specs_[name] = Spec{alias, is_switch};
        if (alias) alias_[alias] = name;
        if (!def.empty()) defaults_[name] = def;
    }
    void parse(int argc, char **argv) {
        for (int i = 1; i < argc; i++) {
            std::string a = argv[i], name, val;
            bool has_val = false;
            if (a.rfind("--", 0) == 0) {
                name = a.substr(2);
                const size_t eq = name.find('=');
                if (eq != std::string::npos) {
                    val = name.substr(eq + 1);
                    name = name.substr(0, eq);
                    has_val = true;
                }
            } else if (a.size() == 2 && a[0] == '-') {
                auto it = alias_.find(a[1]);
                if (it == alias_.end()) throw std::runtime_error("unknown option " + a);
                name = it->second;
            } else {
                throw std::runtime_error("unexpected argument " + a);
            }
            auto sp = specs_.find(name);
            if (sp == specs_.end()) throw std::runtime_error("unknown option --" + name);
            if (!has_val) {
                if (sp->second.is_switch) {
                    val = "1";
                    if (i + 1 < argc && (std::string(argv[i + 1]) == "0" || std::string(argv[i + 1]) == "1"))
                        val = argv[++i];
                } else {
                    if (i + 1 >= argc) throw std::runtime_error("missing value for --" + name);
                    val = argv[++i];
                }
            }
            values_[name].push_back(val);
        }
    }
    bool has(const std::string &n) const { return values_.count(n) || defaults_.count(n); }
    // given on the command line (defaults aside)
    bool given(const std::string &n) const { return values_.count(n) > 0; }
    std::string str(const std::string &n) const {
        auto it = values_.find(n);
        if (it != values_.end()) return it->second.back();
        auto d = defaults_.find(n);
        return d != defaults_.end() ? d->second : std::string();
    }
    std::vector<std::string> all(const std::string &n) const {
        auto it = values_.find(n);
        return it != values_.end() ? it->second : std::vector<std::string>();
    }
    int i(const std::string &n) const { return std::atoi(str(n).c_str()); }
    double f(const std::string &n) const { return std::atof(str(n).c_str()); }
    bool b(const std::string &n) const {
        const std::string v = str(n);
        return !v.empty() && v != "0" && v != "false";
    }

private:
    struct Spec {
        char alias;
        bool is_switch;
    };
    std::map<std::string, Spec> specs_;
    std::map<char, std::string> alias_;
    std::map<std::string, std::string> defaults_;
    std::map<std::string, std::vector<std::string>> values_;
};

}  // namespace gmm
