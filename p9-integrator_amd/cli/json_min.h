// json_min.h — a small JSON reader for mint configuration files.
//
// The reference parses its config with the vendored nlohmann/json
// (src/simutils.cuh:14,257).  The schema is flat (arrays of objects of
// numbers and strings), so a recursive-descent reader of RFC 8259 values is
// enough: objects, arrays, strings (with \uXXXX escapes kept as UTF-8),
// numbers, true/false/null.
#pragma once
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace jsonmin {

struct Value {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Value> arr;
  std::map<std::string, Value> obj;

  bool has(const std::string &k) const { return kind == Object && obj.count(k) != 0; }
  const Value &at(const std::string &k) const {
    auto it = obj.find(k);
    if (kind != Object || it == obj.end()) throw std::runtime_error("missing key \"" + k + "\"");
    return it->second;
  }
  double number() const {
    if (kind != Number) throw std::runtime_error("expected a number");
    return num;
  }
  double number_or(const std::string &k, double d) const { return has(k) ? at(k).number() : d; }
  std::string string_or(const std::string &k, const std::string &d) const {
    if (!has(k)) return d;
    const Value &v = at(k);
    if (v.kind != String) throw std::runtime_error("expected a string for \"" + k + "\"");
    return v.str;
  }
};

class Parser {
 public:
  explicit Parser(const std::string &t) : s_(t) {}
  Value parse() {
    Value v = value();
    ws();
    if (i_ != s_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string &s_;
  size_t i_ = 0;

  [[noreturn]] void fail(const std::string &m) {
    throw std::runtime_error("JSON parse error at byte " + std::to_string(i_) + ": " + m);
  }
  void ws() {
    while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\n' || s_[i_] == '\r' || s_[i_] == '\t')) ++i_;
  }
  char peek() {
    ws();
    if (i_ >= s_.size()) fail("unexpected end of input");
    return s_[i_];
  }
  void expect(char c) {
    if (peek() != c) fail(std::string("expected '") + c + "'");
    ++i_;
  }
  bool lit(const char *w) {
    size_t n = 0;
    while (w[n]) ++n;
    if (s_.compare(i_, n, w) == 0) {
      i_ += n;
      return true;
    }
    return false;
  }
  Value value() {
    const char c = peek();
    Value v;
    if (c == '{') {
      v.kind = Value::Object;
      ++i_;
      if (peek() == '}') {
        ++i_;
        return v;
      }
      for (;;) {
        if (peek() != '"') fail("expected a key");
        std::string k = string();
        expect(':');
        v.obj[k] = value();
        const char d = peek();
        ++i_;
        if (d == '}') break;
        if (d != ',') fail("expected ',' or '}'");
      }
    } else if (c == '[') {
      v.kind = Value::Array;
      ++i_;
      if (peek() == ']') {
        ++i_;
        return v;
      }
      for (;;) {
        v.arr.push_back(value());
        const char d = peek();
        ++i_;
        if (d == ']') break;
        if (d != ',') fail("expected ',' or ']'");
      }
    } else if (c == '"') {
      v.kind = Value::String;
      v.str = string();
    } else if (lit("true")) {
      v.kind = Value::Bool;
      v.b = true;
    } else if (lit("false")) {
      v.kind = Value::Bool;
    } else if (lit("null")) {
      v.kind = Value::Null;
    } else {
      v.kind = Value::Number;
      const char *b = s_.c_str() + i_;
      char *e = nullptr;
      v.num = std::strtod(b, &e);
      if (e == b) fail("bad value");
      i_ += (size_t)(e - b);
    }
    return v;
  }
  static void utf8(std::string &o, unsigned cp) {
    if (cp < 0x80) {
      o += (char)cp;
    } else if (cp < 0x800) {
      o += (char)(0xC0 | (cp >> 6));
      o += (char)(0x80 | (cp & 0x3F));
    } else {
      o += (char)(0xE0 | (cp >> 12));
      o += (char)(0x80 | ((cp >> 6) & 0x3F));
      o += (char)(0x80 | (cp & 0x3F));
    }
  }
  std::string string() {
    expect('"');
    std::string o;
    while (true) {
      if (i_ >= s_.size()) fail("unterminated string");
      char c = s_[i_++];
      if (c == '"') break;
      if (c != '\\') {
        o += c;
        continue;
      }
      if (i_ >= s_.size()) fail("bad escape");
      c = s_[i_++];
      switch (c) {
        case 'n': o += '\n'; break;
        case 't': o += '\t'; break;
        case 'r': o += '\r'; break;
        case 'b': o += '\b'; break;
        case 'f': o += '\f'; break;
        case 'u': {
          if (i_ + 4 > s_.size()) fail("bad \\u escape");
          utf8(o, (unsigned)std::strtoul(s_.substr(i_, 4).c_str(), nullptr, 16));
          i_ += 4;
          break;
        }
        default: o += c;
      }
    }
    return o;
  }
};

inline Value parse(const std::string &text) { return Parser(text).parse(); }

}  // namespace jsonmin
