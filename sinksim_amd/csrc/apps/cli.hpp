// Command-line parsing shared by the tools: positional arguments, --key value options and --flag switches, with
// the reference's behaviour (apps/cli.hpp:15-51: an option without a value exits with status 2).
#pragma once

#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <string>
#include <vector>

namespace sinksim::cli {

struct Args {
  std::vector<std::string> positional;
  std::map<std::string, std::string> options;
  std::set<std::string> flags;

  static Args parse(int argc, char** argv, const std::set<std::string>& flagNames) {
    Args a;
    for (int i = 1; i < argc; ++i) {
      const std::string s = argv[i];
      if (s.size() > 2 && s[0] == '-' && s[1] == '-') {
        const std::string key = s.substr(2);
        if (flagNames.count(key)) {
          a.flags.insert(key);
        } else if (i + 1 < argc) {
          a.options[key] = argv[++i];
        } else {
          std::fprintf(stderr, "option --%s needs a value\n", key.c_str());
          std::exit(2);
        }
      } else {
        a.positional.push_back(s);
      }
    }
    return a;
  }
  bool has(const std::string& k) const { return options.count(k) != 0; }
  bool flag(const std::string& k) const { return flags.count(k) != 0; }
  std::string get(const std::string& k, const std::string& def) const {
    const auto it = options.find(k);
    return it == options.end() ? def : it->second;
  }
  double num(const std::string& k, double def) const {
    const auto it = options.find(k);
    return it == options.end() ? def : std::atof(it->second.c_str());
  }
};

inline std::string minutes(double seconds) {
  char buf[32];
  std::snprintf(buf, sizeof buf, "%.2f min", seconds / 60.0);
  return buf;
}

// The numerics a GPU run computes: --numerics portable|portable32, or the reference GPU runner's --precision
// double|mixed (apps/sinksim_cuda_run.cpp:94-98). The oracle and std numerics exist on the CPU only.
inline bool gpu_numerics(const Args& args, const char* tool, const std::string& fallback, bool& mixed) {
  std::string n = fallback;
  if (args.has("precision")) {
    const std::string p = args.get("precision", "mixed");
    if (p != "mixed" && p != "double") {
      std::fprintf(stderr, "unknown --precision %s (use mixed or double)\n", p.c_str());
      return false;
    }
    n = p == "mixed" ? "portable32" : "portable";
  }
  if (args.has("numerics")) n = args.get("numerics", n);
  if (n == "portable32") { mixed = true; return true; }
  if (n == "portable") { mixed = false; return true; }
  if (n == "oracle" || n == "std")
    std::fprintf(stderr, "%s: the %s numerics run on the CPU only (the platform pow and serial sums); the GPU computes portable and portable32\n",
                 tool, n.c_str());
  else
    std::fprintf(stderr, "unknown --numerics %s (use portable or portable32)\n", n.c_str());
  return false;
}

}  // namespace sinksim::cli
