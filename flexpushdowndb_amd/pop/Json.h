// A small JSON reader for the metadata objects operators exchange (bloom
// filter descriptions, hash functions).  Integers without a fraction or
// exponent are kept exactly as int64; every number is also kept as a double.
#pragma once

#include <cctype>
#include <cerrno>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "Expected.h"

namespace fpdb::amd {

struct JsonValue {
  enum Kind { NUL, BOOL, NUMBER, STRING, ARRAY, OBJECT } kind = NUL;
  bool b = false;
  bool integral = false;
  int64_t i = 0;
  double d = 0;
  std::string s;
  std::vector<JsonValue> items;
  std::map<std::string, JsonValue> fields;

  bool contains(const std::string& key) const { return kind == OBJECT && fields.count(key); }
  const JsonValue& operator[](const std::string& key) const { return fields.at(key); }
};

namespace json_detail {

struct Reader {
  const std::string& t;
  size_t p = 0;
  std::string err;

  void ws() {
    while (p < t.size() && (t[p] == ' ' || t[p] == '\n' || t[p] == '\r' || t[p] == '\t')) ++p;
  }
  bool fail(const std::string& what) {
    if (err.empty()) err = what + " at offset " + std::to_string(p);
    return false;
  }
  bool lit(const char* w) {
    size_t n = std::char_traits<char>::length(w);
    if (t.compare(p, n, w) != 0) return fail(std::string("expected '") + w + "'");
    p += n;
    return true;
  }
  bool str(std::string& out) {
    if (p >= t.size() || t[p] != '"') return fail("expected a string");
    ++p;
    while (p < t.size() && t[p] != '"') {
      char c = t[p++];
      if (c == '\\') {
        if (p >= t.size()) return fail("unterminated escape");
        char e = t[p++];
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {
            if (p + 4 > t.size()) return fail("short \\u escape");
            unsigned cp = (unsigned)std::strtoul(t.substr(p, 4).c_str(), nullptr, 16);
            p += 4;
            if (cp < 0x80) {
              out += (char)cp;
            } else if (cp < 0x800) {
              out += (char)(0xC0 | (cp >> 6));
              out += (char)(0x80 | (cp & 0x3F));
            } else {
              out += (char)(0xE0 | (cp >> 12));
              out += (char)(0x80 | ((cp >> 6) & 0x3F));
              out += (char)(0x80 | (cp & 0x3F));
            }
            break;
          }
          default: out += e;
        }
      } else {
        out += c;
      }
    }
    if (p >= t.size()) return fail("unterminated string");
    ++p;
    return true;
  }
  bool value(JsonValue& v, int depth) {
    if (depth > 64) return fail("nesting too deep");
    ws();
    if (p >= t.size()) return fail("unexpected end");
    char c = t[p];
    if (c == '{') {
      v.kind = JsonValue::OBJECT;
      ++p;
      ws();
      if (p < t.size() && t[p] == '}') return ++p, true;
      while (true) {
        ws();
        std::string key;
        if (!str(key)) return false;
        ws();
        if (p >= t.size() || t[p] != ':') return fail("expected ':'");
        ++p;
        if (!value(v.fields[key], depth + 1)) return false;
        ws();
        if (p < t.size() && t[p] == ',') { ++p; continue; }
        if (p < t.size() && t[p] == '}') return ++p, true;
        return fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      v.kind = JsonValue::ARRAY;
      ++p;
      ws();
      if (p < t.size() && t[p] == ']') return ++p, true;
      while (true) {
        v.items.emplace_back();
        if (!value(v.items.back(), depth + 1)) return false;
        ws();
        if (p < t.size() && t[p] == ',') { ++p; continue; }
        if (p < t.size() && t[p] == ']') return ++p, true;
        return fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.kind = JsonValue::STRING;
      return str(v.s);
    }
    if (c == 't') { v.kind = JsonValue::BOOL; v.b = true; return lit("true"); }
    if (c == 'f') { v.kind = JsonValue::BOOL; v.b = false; return lit("false"); }
    if (c == 'n') { v.kind = JsonValue::NUL; return lit("null"); }
    size_t start = p;
    if (t[p] == '-') ++p;
    bool frac = false;
    while (p < t.size() && (std::isdigit((unsigned char)t[p]) || t[p] == '.' || t[p] == 'e' || t[p] == 'E' ||
                            t[p] == '+' || t[p] == '-')) {
      if (t[p] == '.' || t[p] == 'e' || t[p] == 'E') frac = true;
      ++p;
    }
    if (p == start || (p == start + 1 && t[start] == '-')) return fail("unexpected character");
    std::string num = t.substr(start, p - start);
    v.kind = JsonValue::NUMBER;
    char* end = nullptr;
    v.d = std::strtod(num.c_str(), &end);
    if (end != num.c_str() + num.size()) return fail("malformed number");
    if (!frac) {
      errno = 0;
      long long x = std::strtoll(num.c_str(), &end, 10);
      if (errno == 0 && end == num.c_str() + num.size()) {
        v.integral = true;
        v.i = x;
      }
    }
    return true;
  }
};

}  // namespace json_detail

inline Expected<JsonValue> parseJson(const std::string& text) {
  json_detail::Reader r{text};
  JsonValue v;
  if (!r.value(v, 0)) return make_unexpected("JSON parse error: " + r.err);
  r.ws();
  if (r.p != text.size()) return make_unexpected("JSON parse error: trailing characters at offset " + std::to_string(r.p));
  return v;
}

// Field accessors that name the missing or mistyped key.
inline Expected<int64_t> jsonInt(const JsonValue& o, const std::string& key, const std::string& what) {
  if (!o.contains(key)) return make_unexpected(key + " not specified in " + what + " JSON");
  const JsonValue& v = o[key];
  if (v.kind != JsonValue::NUMBER || !v.integral) return make_unexpected(key + " in " + what + " JSON is not an integer");
  return v.i;
}

inline Expected<double> jsonDouble(const JsonValue& o, const std::string& key, const std::string& what) {
  if (!o.contains(key)) return make_unexpected(key + " not specified in " + what + " JSON");
  const JsonValue& v = o[key];
  if (v.kind != JsonValue::NUMBER) return make_unexpected(key + " in " + what + " JSON is not a number");
  return v.d;
}

inline Expected<bool> jsonBool(const JsonValue& o, const std::string& key, const std::string& what) {
  if (!o.contains(key)) return make_unexpected(key + " not specified in " + what + " JSON");
  const JsonValue& v = o[key];
  if (v.kind != JsonValue::BOOL) return make_unexpected(key + " in " + what + " JSON is not a boolean");
  return v.b;
}

}  // namespace fpdb::amd
