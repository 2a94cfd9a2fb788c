// Minimal JSON reader for scene configs (objects, arrays, numbers, strings,
// booleans, null). The reference parses scene files with nlohmann::json
// (scene.cpp:119-131); only reading is needed here.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace hm {

struct JsonError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

class Json {
 public:
  enum Type { Null, Bool, Number, String, Array, Object };

  Json() = default;
  static Json parse(const std::string& text);

  Type type() const { return type_; }
  bool is_null() const { return type_ == Null; }
  bool is_number() const { return type_ == Number; }
  bool is_string() const { return type_ == String; }
  bool is_bool() const { return type_ == Bool; }
  bool is_array() const { return type_ == Array; }
  bool is_object() const { return type_ == Object; }

  // Typed accessors throw JsonError on a type mismatch, like nlohmann's
  // implicit conversions throw json::exception (scene.cpp relies on that to
  // fall back to defaults).
  double num() const;
  float f() const { return (float)num(); }
  int i() const { return (int)num(); }
  bool b() const;
  const std::string& str() const;
  size_t size() const;
  const Json& at(size_t idx) const;
  const Json& at(const std::string& key) const;  // throws if missing
  bool has(const std::string& key) const;
  const std::vector<Json>& items() const;
  const std::vector<std::pair<std::string, Json>>& members() const;

 private:
  Type type_ = Null;
  double num_ = 0.0;
  bool bool_ = false;
  std::string str_;
  std::vector<Json> arr_;
  std::vector<std::pair<std::string, Json>> obj_;
  friend struct JsonParser;
};

}  // namespace hm
