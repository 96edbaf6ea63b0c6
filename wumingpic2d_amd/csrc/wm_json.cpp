// wm_json.cpp -- the JSON document model of wm_json.h: a recursive-descent
// parser, a printer, and path lookup.
#include "wm_json.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <cctype>
#include <fstream>
#include <sstream>

namespace wmjson {

Value::~Value() {
  for (auto &m : obj) delete m.second;
  for (auto *a : arr) delete a;
}

Value *Value::member(const std::string &key) const {
  for (auto &m : obj)
    if (m.first == key) return m.second;
  return nullptr;
}

Value *Value::add(const std::string &key, Value *v) {
  v->parent = this;
  v->name = key;
  if (kind == Arr)
    arr.push_back(v);
  else
    obj.emplace_back(key, v);
  return v;
}

void Value::set_member(const std::string &key, Value *v) {
  for (auto &m : obj)
    if (m.first == key) {
      delete m.second;
      m.second = v;
      v->parent = this;
      v->name = key;
      return;
    }
  add(key, v);
}

Value *make_obj() {
  Value *v = new Value();
  v->kind = Value::Obj;
  return v;
}
Value *make_arr() {
  Value *v = new Value();
  v->kind = Value::Arr;
  return v;
}
Value *make_int(long long x) {
  Value *v = new Value();
  v->kind = Value::Int;
  v->i = x;
  return v;
}
Value *make_real(double x) {
  Value *v = new Value();
  v->kind = Value::Real;
  v->r = x;
  return v;
}
Value *make_str(const std::string &x) {
  Value *v = new Value();
  v->kind = Value::Str;
  v->s = x;
  return v;
}
Value *make_bool(bool x) {
  Value *v = new Value();
  v->kind = Value::Bool;
  v->b = x;
  return v;
}

namespace {

struct Parser {
  const std::string &t;
  size_t p = 0;
  std::string err;
  explicit Parser(const std::string &text) : t(text) {}

  void ws() {
    while (p < t.size() && isspace((unsigned char)t[p])) ++p;
  }
  bool fail(const char *what) {
    if (err.empty()) {
      char b[128];
      snprintf(b, sizeof b, "%s at offset %zu", what, p);
      err = b;
    }
    return false;
  }
  bool lit(const char *w) {
    const size_t n = strlen(w);
    if (t.compare(p, n, w) != 0) return false;
    p += n;
    return true;
  }
  bool string(std::string &out) {
    if (p >= t.size() || t[p] != '"') return fail("expected a string");
    ++p;
    out.clear();
    while (p < t.size() && t[p] != '"') {
      char c = t[p++];
      if (c == '\\') {
        if (p >= t.size()) return fail("bad escape");
        c = t[p++];
        switch (c) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {
            if (p + 4 > t.size()) return fail("bad \\u escape");
            const unsigned cp = (unsigned)strtoul(t.substr(p, 4).c_str(), nullptr, 16);
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
          default: out += c;
        }
      } else {
        out += c;
      }
    }
    if (p >= t.size()) return fail("unterminated string");
    ++p;
    return true;
  }
  Value *value() {
    ws();
    if (p >= t.size()) {
      fail("unexpected end");
      return nullptr;
    }
    const char c = t[p];
    if (c == '{') {
      ++p;
      Value *o = make_obj();
      ws();
      if (p < t.size() && t[p] == '}') {
        ++p;
        return o;
      }
      for (;;) {
        ws();
        std::string k;
        if (!string(k)) break;
        ws();
        if (p >= t.size() || t[p] != ':') {
          fail("expected ':'");
          break;
        }
        ++p;
        Value *v = value();
        if (!v) break;
        o->add(k, v);
        ws();
        if (p < t.size() && t[p] == ',') {
          ++p;
          continue;
        }
        if (p < t.size() && t[p] == '}') {
          ++p;
          return o;
        }
        fail("expected ',' or '}'");
        break;
      }
      delete o;
      return nullptr;
    }
    if (c == '[') {
      ++p;
      Value *a = make_arr();
      ws();
      if (p < t.size() && t[p] == ']') {
        ++p;
        return a;
      }
      for (;;) {
        Value *v = value();
        if (!v) break;
        a->add("", v);
        ws();
        if (p < t.size() && t[p] == ',') {
          ++p;
          continue;
        }
        if (p < t.size() && t[p] == ']') {
          ++p;
          return a;
        }
        fail("expected ',' or ']'");
        break;
      }
      delete a;
      return nullptr;
    }
    if (c == '"') {
      std::string s;
      if (!string(s)) return nullptr;
      return make_str(s);
    }
    if (lit("true")) return make_bool(true);
    if (lit("false")) return make_bool(false);
    if (lit("null")) return new Value();
    // a number: integer unless it has a fraction or an exponent
    const size_t s0 = p;
    if (p < t.size() && (t[p] == '-' || t[p] == '+')) ++p;
    bool real = false;
    while (p < t.size()) {
      const char d = t[p];
      if (isdigit((unsigned char)d)) {
        ++p;
      } else if (d == '.' || d == 'e' || d == 'E' || d == 'd' || d == 'D') {
        real = true;
        ++p;
        if ((d == 'e' || d == 'E' || d == 'd' || d == 'D') && p < t.size() &&
            (t[p] == '-' || t[p] == '+'))
          ++p;
      } else {
        break;
      }
    }
    if (p == s0) {
      fail("unexpected character");
      return nullptr;
    }
    std::string num = t.substr(s0, p - s0);
    for (char &ch : num)
      if (ch == 'd' || ch == 'D') ch = 'e';
    return real ? make_real(strtod(num.c_str(), nullptr)) : make_int(strtoll(num.c_str(), nullptr, 10));
  }
};

void put_str(std::string &o, const std::string &s) {
  o += '"';
  for (char c : s) {
    switch (c) {
      case '"': o += "\\\""; break;
      case '\\': o += "\\\\"; break;
      case '\n': o += "\\n"; break;
      case '\t': o += "\\t"; break;
      default: o += c;
    }
  }
  o += '"';
}

void put(std::string &o, const Value *v, bool pretty, int depth) {
  const std::string pad = pretty ? std::string(2 * (depth + 1), ' ') : "";
  const std::string end = pretty ? std::string(2 * depth, ' ') : "";
  char b[64];
  switch (v->kind) {
    case Value::Null: o += "null"; break;
    case Value::Bool: o += v->b ? "true" : "false"; break;
    case Value::Int:
      snprintf(b, sizeof b, "%lld", v->i);
      o += b;
      break;
    case Value::Real:
      snprintf(b, sizeof b, "%.17E", v->r);
      o += b;
      break;
    case Value::Str: put_str(o, v->s); break;
    case Value::Arr:
      o += '[';
      for (size_t k = 0; k < v->arr.size(); ++k) {
        if (k) o += ',';
        if (pretty) o += "\n" + pad;
        put(o, v->arr[k], pretty, depth + 1);
      }
      if (pretty && !v->arr.empty()) o += "\n" + end;
      o += ']';
      break;
    case Value::Obj:
      o += '{';
      for (size_t k = 0; k < v->obj.size(); ++k) {
        if (k) o += ',';
        if (pretty) o += "\n" + pad;
        put_str(o, v->obj[k].first);
        o += pretty ? ": " : ":";
        put(o, v->obj[k].second, pretty, depth + 1);
      }
      if (pretty && !v->obj.empty()) o += "\n" + end;
      o += '}';
      break;
  }
}

}  // namespace

Value *parse(const std::string &text, std::string *err) {
  Parser P(text);
  Value *v = P.value();
  if (v) {
    P.ws();
    if (P.p != text.size()) {
      delete v;
      v = nullptr;
      P.fail("trailing characters");
    }
  }
  if (!v && err) *err = P.err;
  return v;
}

Value *load(const std::string &path, std::string *err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    if (err) *err = "cannot open " + path;
    return nullptr;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  return parse(ss.str(), err);
}

std::string dump(const Value *v, bool pretty) {
  std::string o;
  put(o, v, pretty, 0);
  if (pretty) o += '\n';
  return o;
}

bool save(const Value *v, const std::string &path) {
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) return false;
  f << dump(v, true);
  return (bool)f;
}

Value *get_path(Value *root, const std::string &path) {
  Value *cur = root;
  size_t p = 0;
  while (cur && p <= path.size()) {
    size_t q = path.find('.', p);
    if (q == std::string::npos) q = path.size();
    std::string part = path.substr(p, q - p);
    int index = -1;
    const size_t lp = part.find('(');
    if (lp != std::string::npos) {
      const size_t rp = part.find(')', lp);
      index = atoi(part.substr(lp + 1, rp - lp - 1).c_str());
      part = part.substr(0, lp);
    }
    if (!part.empty()) cur = cur->kind == Value::Obj ? cur->member(part) : nullptr;
    if (cur && index >= 0)
      cur = cur->kind == Value::Arr && index >= 1 && (size_t)index <= cur->arr.size()
                ? cur->arr[index - 1]
                : nullptr;
    p = q + 1;
    if (q == path.size()) break;
  }
  return cur;
}

}  // namespace wmjson
