// Minimal JSON reader for the flat config objects of the RHLLC module
// (hllc_cfg.json, solve_cfg.json, directories_cfg.json): one object whose
// values are numbers, strings, booleans or null.  Nested objects/arrays are
// skipped.  Header-only.
#pragma once

#include <cctype>
#include <cstdlib>
#include <map>
#include <string>

namespace jsonmin {

struct Value {
  enum Kind { Null, Number, String, Bool, Other } kind = Null;
  double num = 0;
  std::string str;
  bool b = false;
};

inline bool get(const Value &v, double &out) {
  if (v.kind != Value::Number) return false;
  out = v.num;
  return true;
}
inline bool get(const Value &v, int &out) {
  if (v.kind != Value::Number) return false;
  out = (int)v.num;
  return true;
}
inline bool get(const Value &v, std::string &out) {
  if (v.kind != Value::String) return false;
  out = v.str;
  return true;
}

namespace detail {
struct P {
  const std::string &s;
  size_t i = 0;
  std::string err;
  void ws() {
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
  }
  bool fail(const char *m) {
    err = std::string(m) + " at offset " + std::to_string(i);
    return false;
  }
  bool str(std::string &out) {
    if (s[i] != '"') return fail("expected string");
    ++i;
    out.clear();
    while (i < s.size() && s[i] != '"') {
      char c = s[i++];
      if (c == '\\' && i < s.size()) {
        const char e = s[i++];
        switch (e) {
          case 'n': c = '\n'; break;
          case 't': c = '\t'; break;
          case 'r': c = '\r'; break;
          case 'b': c = '\b'; break;
          case 'f': c = '\f'; break;
          case 'u': {
            if (i + 4 > s.size()) return fail("bad escape");
            const long cp = std::strtol(s.substr(i, 4).c_str(), nullptr, 16);
            i += 4;
            c = cp < 128 ? (char)cp : '?';
            break;
          }
          default: c = e;
        }
      }
      out.push_back(c);
    }
    if (i >= s.size()) return fail("unterminated string");
    ++i;
    return true;
  }
  bool skip_nested() {
    int depth = 0;
    do {
      if (i >= s.size()) return fail("unterminated value");
      const char c = s[i];
      if (c == '"') {
        std::string t;
        if (!str(t)) return false;
        continue;
      }
      if (c == '{' || c == '[') ++depth;
      if (c == '}' || c == ']') --depth;
      ++i;
    } while (depth > 0);
    return true;
  }
  bool value(Value &v) {
    ws();
    if (i >= s.size()) return fail("expected value");
    const char c = s[i];
    if (c == '"') {
      v.kind = Value::String;
      return str(v.str);
    }
    if (c == '{' || c == '[') {
      v.kind = Value::Other;
      return skip_nested();
    }
    if (s.compare(i, 4, "true") == 0) return v.kind = Value::Bool, v.b = true, i += 4, true;
    if (s.compare(i, 5, "false") == 0) return v.kind = Value::Bool, v.b = false, i += 5, true;
    if (s.compare(i, 4, "null") == 0) return v.kind = Value::Null, i += 4, true;
    char *end = nullptr;
    v.num = std::strtod(s.c_str() + i, &end);
    if (end == s.c_str() + i) return fail("bad number");
    v.kind = Value::Number;
    i = (size_t)(end - s.c_str());
    return true;
  }
};
}  // namespace detail

inline bool parse_object(const std::string &text, std::map<std::string, Value> &out, std::string &err) {
  detail::P p{text};
  p.ws();
  if (p.i >= text.size() || text[p.i] != '{') return err = "expected '{'", false;
  ++p.i;
  p.ws();
  if (p.i < text.size() && text[p.i] == '}') return true;
  for (;;) {
    p.ws();
    std::string key;
    if (!p.str(key)) return err = p.err, false;
    p.ws();
    if (p.i >= text.size() || text[p.i] != ':') return err = "expected ':'", false;
    ++p.i;
    Value v;
    if (!p.value(v)) return err = p.err, false;
    out[key] = v;
    p.ws();
    if (p.i < text.size() && text[p.i] == ',') {
      ++p.i;
      continue;
    }
    if (p.i < text.size() && text[p.i] == '}') return true;
    return err = "expected ',' or '}'", false;
  }
}

}  // namespace jsonmin
