// Minimal command-line flags in the gflags syntax the reference's tools accept
// (src/define_args.h): --name=value, --name value, -name=value, and for
// booleans --name, --noname, --name=true|false.
#pragma once
#include <cstdlib>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>

namespace VerifyCurves {

class Flags {
 public:
  void add(const std::string &name, const std::string &def, const std::string &help,
           bool is_bool = false) {
    values_[name] = def;
    help_[name] = help;
    bool_[name] = is_bool;
  }
  void parse(int argc, char **argv) {
    for (int k = 1; k < argc; ++k) {
      std::string a = argv[k];
      if (a == "--help" || a == "-help") {
        for (auto &kv : help_)
          std::cout << "  --" << kv.first << " (" << kv.second << ") default: " << values_[kv.first]
                    << "\n";
        std::exit(0);
      }
      if (a.size() < 2 || a[0] != '-') throw std::runtime_error("unexpected argument: " + a);
      a = a.substr(a[1] == '-' ? 2 : 1);
      std::string key = a, val;
      bool has = false;
      const size_t eq = a.find('=');
      if (eq != std::string::npos) {
        key = a.substr(0, eq);
        val = a.substr(eq + 1);
        has = true;
      }
      if (!values_.count(key) && key.rfind("no", 0) == 0 && bool_.count(key.substr(2)) &&
          bool_[key.substr(2)] && !has) {
        values_[key.substr(2)] = "false";
        continue;
      }
      if (!values_.count(key)) throw std::runtime_error("unknown command line flag '" + key + "'");
      if (bool_[key]) {
        values_[key] = has ? val : "true";
      } else {
        if (!has) {
          if (k + 1 >= argc) throw std::runtime_error("flag '" + key + "' is missing its argument");
          val = argv[++k];
        }
        values_[key] = val;
      }
    }
  }
  std::string str(const std::string &k) const { return values_.at(k); }
  double num(const std::string &k) const { return std::atof(values_.at(k).c_str()); }
  bool flag(const std::string &k) const {
    const std::string &v = values_.at(k);
    return v == "true" || v == "1" || v == "yes" || v == "t" || v == "y";
  }

 private:
  std::map<std::string, std::string> values_, help_;
  std::map<std::string, bool> bool_;
};

}  // namespace VerifyCurves
