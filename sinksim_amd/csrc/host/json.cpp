#include "json.hpp"

#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

namespace sinksim::json {

Value::Value(const std::vector<double>& v) : kind_(Kind::Array) {
  items_.reserve(v.size());
  for (double d : v) items_.emplace_back(d);
}

bool Value::contains(const std::string& k) const { return find(k) != nullptr; }

const Value* Value::find(const std::string& k) const {
  if (kind_ != Kind::Object) return nullptr;
  for (const auto& m : members_)
    if (m.first == k) return &m.second;
  return nullptr;
}

Value& Value::operator[](const std::string& k) {
  if (kind_ == Kind::Null) kind_ = Kind::Object;
  for (auto& m : members_)
    if (m.first == k) return m.second;
  members_.emplace_back(k, Value());
  return members_.back().second;
}

// ------------------------------------------------------------------------------------------- parser

namespace {

struct Parser {
  const char* p;
  const char* end;
  std::string where;

  [[noreturn]] void fail(const std::string& what) const {
    throw std::runtime_error(where + ": JSON parse error: " + what);
  }
  void ws() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
  }
  bool eat(char c) {
    ws();
    if (p < end && *p == c) { ++p; return true; }
    return false;
  }
  void expect(char c) {
    if (!eat(c)) fail(std::string("expected '") + c + "'");
  }
  static void utf8(std::string& out, uint32_t cp) {
    if (cp < 0x80) out.push_back(static_cast<char>(cp));
    else if (cp < 0x800) { out.push_back(static_cast<char>(0xC0 | (cp >> 6))); out.push_back(static_cast<char>(0x80 | (cp & 0x3F))); }
    else if (cp < 0x10000) {
      out.push_back(static_cast<char>(0xE0 | (cp >> 12)));
      out.push_back(static_cast<char>(0x80 | ((cp >> 6) & 0x3F)));
      out.push_back(static_cast<char>(0x80 | (cp & 0x3F)));
    } else {
      out.push_back(static_cast<char>(0xF0 | (cp >> 18)));
      out.push_back(static_cast<char>(0x80 | ((cp >> 12) & 0x3F)));
      out.push_back(static_cast<char>(0x80 | ((cp >> 6) & 0x3F)));
      out.push_back(static_cast<char>(0x80 | (cp & 0x3F)));
    }
  }
  uint32_t hex4() {
    if (end - p < 4) fail("truncated \\u escape");
    uint32_t v = 0;
    for (int k = 0; k < 4; ++k) {
      const char c = *p++;
      v <<= 4;
      if (c >= '0' && c <= '9') v |= static_cast<uint32_t>(c - '0');
      else if (c >= 'a' && c <= 'f') v |= static_cast<uint32_t>(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= static_cast<uint32_t>(c - 'A' + 10);
      else fail("bad \\u escape");
    }
    return v;
  }
  std::string str() {
    expect('"');
    std::string out;
    while (true) {
      if (p >= end) fail("unterminated string");
      const char c = *p++;
      if (c == '"') break;
      if (c != '\\') { out.push_back(c); continue; }
      if (p >= end) fail("unterminated escape");
      const char e = *p++;
      switch (e) {
        case '"': out.push_back('"'); break;
        case '\\': out.push_back('\\'); break;
        case '/': out.push_back('/'); break;
        case 'b': out.push_back('\b'); break;
        case 'f': out.push_back('\f'); break;
        case 'n': out.push_back('\n'); break;
        case 'r': out.push_back('\r'); break;
        case 't': out.push_back('\t'); break;
        case 'u': {
          uint32_t cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00 && end - p >= 6 && p[0] == '\\' && p[1] == 'u') {
            p += 2;
            const uint32_t lo = hex4();
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          utf8(out, cp);
          break;
        }
        default: fail("bad escape");
      }
    }
    return out;
  }
  Value num() {
    const char* s = p;
    if (p < end && *p == '-') ++p;
    bool isFloat = false;
    while (p < end && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' || *p == '+' || *p == '-')) {
      if (*p == '.' || *p == 'e' || *p == 'E') isFloat = true;
      ++p;
    }
    if (!isFloat && p - s <= 18) {
      long long v = 0;
      const auto r = std::from_chars(s, p, v);
      if (r.ec != std::errc() || r.ptr != p) fail("bad number");
      return Value(v);
    }
    double d = 0;
    const auto r = std::from_chars(s, p, d);  // correctly rounded
    if (r.ec != std::errc() || r.ptr != p) fail("bad number");
    return Value(d);
  }
  Value value(int depth) {
    if (depth > 256) fail("nesting too deep");
    ws();
    if (p >= end) fail("unexpected end");
    const char c = *p;
    if (c == '{') {
      ++p;
      Value o = Value::object();
      if (eat('}')) return o;
      do {
        ws();
        std::string k = str();
        expect(':');
        o[k] = value(depth + 1);
      } while (eat(','));
      expect('}');
      return o;
    }
    if (c == '[') {
      ++p;
      Value a = Value::array();
      if (eat(']')) return a;
      do a.push_back(value(depth + 1));
      while (eat(','));
      expect(']');
      return a;
    }
    if (c == '"') return Value(str());
    if (end - p >= 4 && std::strncmp(p, "true", 4) == 0) { p += 4; return Value(true); }
    if (end - p >= 5 && std::strncmp(p, "false", 5) == 0) { p += 5; return Value(false); }
    if (end - p >= 4 && std::strncmp(p, "null", 4) == 0) { p += 4; return Value(); }
    if (c == '-' || (c >= '0' && c <= '9')) return num();
    fail(std::string("unexpected character '") + c + "'");
  }
};

void dump_string(const std::string& s, std::string& out) {
  out.push_back('"');
  for (unsigned char c : s) {
    switch (c) {
      case '"': out += "\\\""; break;
      case '\\': out += "\\\\"; break;
      case '\n': out += "\\n"; break;
      case '\r': out += "\\r"; break;
      case '\t': out += "\\t"; break;
      case '\b': out += "\\b"; break;
      case '\f': out += "\\f"; break;
      default:
        if (c < 0x20) {
          char buf[8];
          std::snprintf(buf, sizeof buf, "\\u%04x", c);
          out += buf;
        } else {
          out.push_back(static_cast<char>(c));
        }
    }
  }
  out.push_back('"');
}

void dump_impl(const Value& v, std::string& out, int indent, int level) {
  switch (v.kind()) {
    case Value::Kind::Null: out += "null"; break;
    case Value::Kind::Bool: out += v.as_bool() ? "true" : "false"; break;
    case Value::Kind::Int: out += std::to_string(static_cast<long long>(v.as_double())); break;
    case Value::Kind::Float: out += format_number(v.as_double()); break;
    case Value::Kind::String: dump_string(v.as_string(), out); break;
    case Value::Kind::Array: {
      out.push_back('[');
      bool first = true;
      for (const Value& x : v.items()) {
        if (!first) out.push_back(',');
        first = false;
        dump_impl(x, out, -1, level + 1);
      }
      out.push_back(']');
      break;
    }
    case Value::Kind::Object: {
      if (v.members().empty()) { out += "{}"; break; }
      out.push_back('{');
      const std::string pad = indent >= 0 ? "\n" + std::string(static_cast<size_t>((level + 1) * indent), ' ') : "";
      const std::string close = indent >= 0 ? "\n" + std::string(static_cast<size_t>(level * indent), ' ') : "";
      bool first = true;
      for (const auto& m : v.members()) {
        if (!first) out.push_back(',');
        first = false;
        out += pad;
        dump_string(m.first, out);
        out += indent >= 0 ? ": " : ":";
        dump_impl(m.second, out, indent, level + 1);
      }
      out += close;
      out.push_back('}');
      break;
    }
  }
}

}  // namespace

Value parse(const std::string& text, const std::string& where) {
  Parser ps{text.data(), text.data() + text.size(), where};
  Value v = ps.value(0);
  ps.ws();
  if (ps.p != ps.end) ps.fail("trailing characters");
  return v;
}

Value read_file(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open " + path);
  std::stringstream ss;
  ss << in.rdbuf();
  return parse(ss.str(), path);
}

std::string format_number(double d) {
  if (!std::isfinite(d)) return "null";
  if (d == 0) d = 0.0;
  char buf[64];
  const auto r = std::to_chars(buf, buf + sizeof buf, d);
  return std::string(buf, r.ptr);
}

std::string dump(const Value& v, int indent) {
  std::string out;
  dump_impl(v, out, indent, 0);
  return out;
}

void write_file(const std::string& path, const Value& v, int indent) {
  std::ofstream out(path, std::ios::binary);
  if (!out) throw std::runtime_error("cannot write " + path);
  out << dump(v, indent) << '\n';
}

const Value& member(const Value& obj, const char* key, const std::string& where) {
  if (!obj.is_object()) throw std::runtime_error(where + ": expected an object holding \"" + key + "\"");
  const Value* v = obj.find(key);
  if (!v) throw std::runtime_error(where + ": missing member \"" + key + "\"");
  return *v;
}

double number(const Value& obj, const char* key, const std::string& where) {
  const Value& v = member(obj, key, where);
  if (!v.is_number()) throw std::runtime_error(where + ": member \"" + key + "\" is not a number");
  return v.as_double();
}

long long integer(const Value& obj, const char* key, const std::string& where) {
  const double d = number(obj, key, where);
  if (d != std::floor(d)) throw std::runtime_error(where + ": member \"" + key + "\" is not an integer");
  return static_cast<long long>(d);
}

bool boolean(const Value& obj, const char* key, const std::string& where) {
  const Value& v = member(obj, key, where);
  if (!v.is_bool()) throw std::runtime_error(where + ": member \"" + key + "\" is not a boolean");
  return v.as_bool();
}

std::string string(const Value& obj, const char* key, const std::string& where) {
  const Value& v = member(obj, key, where);
  if (!v.is_string()) throw std::runtime_error(where + ": member \"" + key + "\" is not a string");
  return v.as_string();
}

std::string string_or(const Value& obj, const char* key, const std::string& fallback) {
  const Value* v = obj.find(key);
  return (v && v->is_string()) ? v->as_string() : fallback;
}

double number_or(const Value& obj, const char* key, double fallback) {
  const Value* v = obj.find(key);
  return (v && v->is_number()) ? v->as_double() : fallback;
}

}  // namespace sinksim::json
