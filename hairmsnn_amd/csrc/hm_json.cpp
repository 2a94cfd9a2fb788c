#include "hm_json.h"

#include <cstdlib>
#include <cstring>

namespace hm {

struct JsonParser {
  const char* p;
  const char* end;

  [[noreturn]] void fail(const char* what) {
    throw JsonError(std::string("json parse error: ") + what);
  }
  void ws() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
  }
  bool lit(const char* s) {
    size_t n = std::strlen(s);
    if ((size_t)(end - p) >= n && std::strncmp(p, s, n) == 0) {
      p += n;
      return true;
    }
    return false;
  }
  static void put_utf8(std::string& out, unsigned cp) {
    if (cp < 0x80) {
      out += (char)cp;
    } else if (cp < 0x800) {
      out += (char)(0xC0 | (cp >> 6));
      out += (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      out += (char)(0xE0 | (cp >> 12));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    } else {
      out += (char)(0xF0 | (cp >> 18));
      out += (char)(0x80 | ((cp >> 12) & 0x3F));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    }
  }
  unsigned hex4() {
    if (end - p < 4) fail("short \\u escape");
    unsigned v = 0;
    for (int k = 0; k < 4; ++k) {
      char c = *p++;
      v <<= 4;
      if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
      else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
      else fail("bad \\u escape");
    }
    return v;
  }
  std::string string_body() {
    // p points just past the opening quote
    std::string out;
    while (true) {
      if (p >= end) fail("unterminated string");
      char c = *p++;
      if (c == '"') break;
      if (c == '\\') {
        if (p >= end) fail("bad escape");
        char e = *p++;
        switch (e) {
          case '"': out += '"'; break;
          case '\\': out += '\\'; break;
          case '/': out += '/'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'n': out += '\n'; break;
          case 'r': out += '\r'; break;
          case 't': out += '\t'; break;
          case 'u': {
            unsigned cp = hex4();
            if (cp >= 0xD800 && cp < 0xDC00 && end - p >= 6 && p[0] == '\\' && p[1] == 'u') {
              p += 2;
              unsigned lo = hex4();
              cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
            }
            put_utf8(out, cp);
            break;
          }
          default: fail("bad escape");
        }
      } else {
        out += c;
      }
    }
    return out;
  }
  Json value(int depth) {
    if (depth > 256) fail("nesting too deep");
    ws();
    if (p >= end) fail("unexpected end");
    Json j;
    char c = *p;
    if (c == '{') {
      ++p;
      j.type_ = Json::Object;
      ws();
      if (p < end && *p == '}') {
        ++p;
        return j;
      }
      while (true) {
        ws();
        if (p >= end || *p != '"') fail("expected key");
        ++p;
        std::string key = string_body();
        ws();
        if (p >= end || *p != ':') fail("expected ':'");
        ++p;
        Json v = value(depth + 1);
        bool replaced = false;
        for (auto& kv : j.obj_) {
          if (kv.first == key) {  // last duplicate wins, as in nlohmann
            kv.second = std::move(v);
            replaced = true;
            break;
          }
        }
        if (!replaced) j.obj_.emplace_back(std::move(key), std::move(v));
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == '}') {
          ++p;
          break;
        }
        fail("expected ',' or '}'");
      }
      return j;
    }
    if (c == '[') {
      ++p;
      j.type_ = Json::Array;
      ws();
      if (p < end && *p == ']') {
        ++p;
        return j;
      }
      while (true) {
        j.arr_.push_back(value(depth + 1));
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == ']') {
          ++p;
          break;
        }
        fail("expected ',' or ']'");
      }
      return j;
    }
    if (c == '"') {
      ++p;
      j.type_ = Json::String;
      j.str_ = string_body();
      return j;
    }
    if (lit("true")) {
      j.type_ = Json::Bool;
      j.bool_ = true;
      return j;
    }
    if (lit("false")) {
      j.type_ = Json::Bool;
      j.bool_ = false;
      return j;
    }
    if (lit("null")) return j;
    if (c == '-' || (c >= '0' && c <= '9')) {
      std::string tok;
      while (p < end && (std::strchr("+-0123456789.eE", *p) != nullptr)) tok += *p++;
      char* e = nullptr;
      double v = std::strtod(tok.c_str(), &e);
      if (e == tok.c_str() || *e != '\0') fail("bad number");
      j.type_ = Json::Number;
      j.num_ = v;
      return j;
    }
    fail("unexpected character");
  }
};

Json Json::parse(const std::string& text) {
  JsonParser ps{text.data(), text.data() + text.size()};
  Json j = ps.value(0);
  ps.ws();
  if (ps.p != ps.end) ps.fail("trailing characters");
  return j;
}

double Json::num() const {
  if (type_ == Number) return num_;
  if (type_ == Bool) return bool_ ? 1.0 : 0.0;  // nlohmann converts booleans to numbers
  throw JsonError("json: not a number");
}
bool Json::b() const {
  if (type_ == Bool) return bool_;
  if (type_ == Number) return num_ != 0.0;
  throw JsonError("json: not a boolean");
}
const std::string& Json::str() const {
  if (type_ != String) throw JsonError("json: not a string");
  return str_;
}
size_t Json::size() const {
  if (type_ == Array) return arr_.size();
  if (type_ == Object) return obj_.size();
  return 0;
}
const Json& Json::at(size_t idx) const {
  if (type_ != Array || idx >= arr_.size()) throw JsonError("json: index out of range");
  return arr_[idx];
}
const Json& Json::at(const std::string& key) const {
  if (type_ == Object)
    for (const auto& kv : obj_)
      if (kv.first == key) return kv.second;
  throw JsonError("json: missing key '" + key + "'");
}
bool Json::has(const std::string& key) const {
  if (type_ != Object) return false;
  for (const auto& kv : obj_)
    if (kv.first == key) return true;
  return false;
}
const std::vector<Json>& Json::items() const {
  if (type_ != Array) throw JsonError("json: not an array");
  return arr_;
}
const std::vector<std::pair<std::string, Json>>& Json::members() const {
  if (type_ != Object) throw JsonError("json: not an object");
  return obj_;
}

}  // namespace hm
