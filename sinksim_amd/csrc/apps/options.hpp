// The tools' command lines: positional arguments, `--name value` options and `--name` switches. A `--name` that is not
// one of the tool's switches takes the next argument as its value, so options a tool does not use are accepted and
// ignored; one given last without a value is a usage error.
#pragma once

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <initializer_list>
#include <string>
#include <utility>
#include <vector>

namespace sinksim::tools {

class CommandLine {
 public:
  CommandLine(int argc, char** argv, std::initializer_list<const char*> switches) {
    for (int i = 1; i < argc; ++i) {
      const std::string word = argv[i];
      if (word.size() <= 2 || word.compare(0, 2, "--") != 0) {
        positional_.push_back(word);
        continue;
      }
      const std::string name = word.substr(2);
      bool isSwitch = false;
      for (const char* s : switches) isSwitch = isSwitch || name == s;
      if (isSwitch) {
        given_.emplace_back(name, "");
      } else if (i + 1 < argc) {
        given_.emplace_back(name, argv[++i]);
      } else if (error_.empty()) {
        error_ = "option --" + name + " needs a value";
      }
    }
  }

  // the first usage error ("" when none)
  const std::string& error() const { return error_; }
  const std::vector<std::string>& positional() const { return positional_; }
  bool has(const std::string& name) const { return find(name) != nullptr; }
  // the last value given for the option, or the fallback
  std::string text(const std::string& name, const std::string& fallback) const {
    const std::string* v = find(name);
    return v ? *v : fallback;
  }
  // a number; a value that is not one is a usage error reported through error()
  double number(const std::string& name, double fallback) {
    const std::string* v = find(name);
    if (!v) return fallback;
    char* end = nullptr;
    errno = 0;
    const double x = std::strtod(v->c_str(), &end);
    if (end == v->c_str() || *end != '\0' || errno == ERANGE) {
      if (error_.empty()) error_ = "option --" + name + " needs a number, not \"" + *v + "\"";
      return fallback;
    }
    return x;
  }

 private:
  const std::string* find(const std::string& name) const {
    for (auto it = given_.rbegin(); it != given_.rend(); ++it)
      if (it->first == name) return &it->second;
    return nullptr;
  }
  std::vector<std::string> positional_;
  std::vector<std::pair<std::string, std::string>> given_;
  std::string error_;
};

// --numerics: the reference's four names (engine/src/simulation.cpp numerics_name / parse_numerics)
inline bool numerics_code(const std::string& name, int& code) {
  const char* const names[4] = {"oracle", "portable", "portable32", "std"};
  for (int k = 0; k < 4; ++k)
    if (name == names[k]) { code = k; return true; }
  return false;
}
inline const char* numerics_name(int code) {
  const char* const names[4] = {"oracle", "portable", "portable32", "std"};
  return code >= 0 && code < 4 ? names[code] : "?";
}

// The GPU tools' numerics: --precision mixed|double (the reference GPU runner's option; mixed by default) or
// --numerics portable32|portable, which wins when both are given; `fallback` without either. False after printing the
// reason to stderr.
inline bool device_numerics(const CommandLine& cl, const char* tool, bool& mixed, const char* fallback = "portable32") {
  std::string name = fallback;
  if (cl.has("precision")) {
    const std::string p = cl.text("precision", "mixed");
    if (p != "mixed" && p != "double") {
      std::fprintf(stderr, "unknown --precision %s (use mixed or double)\n", p.c_str());
      return false;
    }
    name = p == "mixed" ? "portable32" : "portable";
  }
  name = cl.text("numerics", name);
  int code = -1;
  if (!numerics_code(name, code)) {
    std::fprintf(stderr, "unknown --numerics %s (use portable or portable32)\n", name.c_str());
    return false;
  }
  if (code != 1 && code != 2) {
    std::fprintf(stderr, "%s: the %s numerics run on the CPU only (sinksim_sweep computes them); the GPU computes portable and portable32\n", tool,
                 name.c_str());
    return false;
  }
  mixed = code == 2;
  return true;
}

// "%dh%02dm", the validation table's way of printing a time
inline std::string hours_minutes(double seconds) {
  const int m = static_cast<int>(seconds / 60);
  char buf[32];
  std::snprintf(buf, sizeof buf, "%dh%02dm", m / 60, m % 60);
  return buf;
}

}  // namespace sinksim::tools
