// A small JSON document model for the engine's files: objects keep member order, numbers keep whether they were
// written as integers, and numbers are read correctly rounded and written as the shortest decimal that round-trips
// (the reference's format_number, engine/src/json.cpp:56-62). Layout of written files: objects one member per
// line, arrays on one line (engine/src/json.cpp:13-53).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace sinksim::json {

class Value {
 public:
  enum class Kind { Null, Bool, Int, Float, String, Array, Object };

  Value() = default;
  Value(std::nullptr_t) {}
  Value(bool b) : kind_(Kind::Bool), b_(b) {}
  Value(int v) : kind_(Kind::Int), i_(v) {}
  Value(long v) : kind_(Kind::Int), i_(v) {}
  Value(long long v) : kind_(Kind::Int), i_(v) {}
  Value(double d) : kind_(Kind::Float), d_(d) {}
  Value(const char* s) : kind_(Kind::String), s_(s) {}
  Value(std::string s) : kind_(Kind::String), s_(std::move(s)) {}
  Value(const std::vector<double>& v);

  static Value array() { Value v; v.kind_ = Kind::Array; return v; }
  static Value object() { Value v; v.kind_ = Kind::Object; return v; }

  Kind kind() const { return kind_; }
  bool is_null() const { return kind_ == Kind::Null; }
  bool is_bool() const { return kind_ == Kind::Bool; }
  bool is_number() const { return kind_ == Kind::Int || kind_ == Kind::Float; }
  bool is_string() const { return kind_ == Kind::String; }
  bool is_array() const { return kind_ == Kind::Array; }
  bool is_object() const { return kind_ == Kind::Object; }

  bool as_bool() const { return b_; }
  double as_double() const { return kind_ == Kind::Int ? static_cast<double>(i_) : d_; }
  const std::string& as_string() const { return s_; }

  // arrays
  size_t size() const { return kind_ == Kind::Array ? items_.size() : members_.size(); }
  const Value& operator[](size_t i) const { return items_.at(i); }
  void push_back(Value v) { items_.push_back(std::move(v)); }
  const std::vector<Value>& items() const { return items_; }

  // objects (member order is insertion order)
  bool contains(const std::string& k) const;
  const Value* find(const std::string& k) const;
  Value& operator[](const std::string& k);  // inserts null when absent
  const std::vector<std::pair<std::string, Value>>& members() const { return members_; }

 private:
  Kind kind_ = Kind::Null;
  bool b_ = false;
  long long i_ = 0;
  double d_ = 0;
  std::string s_;
  std::vector<Value> items_;
  std::vector<std::pair<std::string, Value>> members_;
};

Value parse(const std::string& text, const std::string& where);
Value read_file(const std::string& path);
std::string dump(const Value& v, int indent = 2);
void write_file(const std::string& path, const Value& v, int indent = 2);

// shortest round-trip decimal; -0 as 0; non-finite as null
std::string format_number(double d);

// member access with the reference's error texts (engine/src/json.cpp:86-115)
const Value& member(const Value& obj, const char* key, const std::string& where);
double number(const Value& obj, const char* key, const std::string& where);
long long integer(const Value& obj, const char* key, const std::string& where);
bool boolean(const Value& obj, const char* key, const std::string& where);
std::string string(const Value& obj, const char* key, const std::string& where);
std::string string_or(const Value& obj, const char* key, const std::string& fallback);
double number_or(const Value& obj, const char* key, double fallback);

}  // namespace sinksim::json
