// wm_json.h -- a small JSON document model for the restart files (the
// JSON + raw pairs of common/paraio.f90) and for the Fortran json_module the
// reference's apps use.  Objects keep their keys in insertion order, as the
// files the reference writes list them.
#pragma once

#include <string>
#include <utility>
#include <vector>

namespace wmjson {

struct Value {
  enum Kind { Null, Bool, Int, Real, Str, Arr, Obj } kind = Null;
  bool b = false;
  long long i = 0;
  double r = 0.0;
  std::string s;
  std::vector<std::pair<std::string, Value *>> obj;   // Obj: members in order
  std::vector<Value *> arr;                            // Arr
  Value *parent = nullptr;
  std::string name;                                    // key in the parent object

  ~Value();
  Value *member(const std::string &key) const;         // null if absent
  Value *add(const std::string &key, Value *v);        // appends (Obj) or pushes (Arr)
  void set_member(const std::string &key, Value *v);   // replaces or appends
  double as_real() const { return kind == Int ? (double)i : r; }
  long long as_int() const { return kind == Real ? (long long)r : i; }
};

Value *make_obj();
Value *make_arr();
Value *make_int(long long v);
Value *make_real(double v);
Value *make_str(const std::string &v);
Value *make_bool(bool v);

// Parses text; returns null and sets err on a syntax error.
Value *parse(const std::string &text, std::string *err);
Value *load(const std::string &path, std::string *err);
// Two-space indented (pretty) or compact text.
std::string dump(const Value *v, bool pretty);
bool save(const Value *v, const std::string &path);

// Path lookup: members separated by '.', array elements as name(k) with a
// 1-based k (spaces inside the parentheses allowed), e.g. "up.shape( 1)".
Value *get_path(Value *root, const std::string &path);

}  // namespace wmjson
