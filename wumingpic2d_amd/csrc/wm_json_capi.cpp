// wm_json_capi.cpp -- the C ABI of include/wmpic_json.h over wm_json.h.
#include <string.h>

#include "../../include/wmpic_json.h"
#include "wm_json.h"

using wmjson::Value;

namespace {
Value *V(void *p) { return (Value *)p; }
Value *at(void *v, const char *path) {
  if (!v) return nullptr;
  return path && *path ? wmjson::get_path(V(v), path) : V(v);
}
// parent object of path and the last member's name
Value *owner(void *v, const char *path, std::string *leaf) {
  std::string p(path ? path : "");
  const size_t dot = p.rfind('.');
  *leaf = dot == std::string::npos ? p : p.substr(dot + 1);
  return dot == std::string::npos ? V(v) : wmjson::get_path(V(v), p.substr(0, dot));
}
int update(void *v, const char *path, Value *x) {
  std::string leaf;
  Value *o = v ? owner(v, path, &leaf) : nullptr;
  if (!o || o->kind != Value::Obj) {
    delete x;
    return 0;
  }
  const bool found = o->member(leaf) != nullptr;
  o->set_member(leaf, x);
  return found ? 1 : 0;
}
int add(void *v, const char *name, Value *x) {
  if (!v || (V(v)->kind != Value::Obj && V(v)->kind != Value::Arr)) {
    delete x;
    return 0;
  }
  V(v)->add(name ? name : "", x);
  return 1;
}
}  // namespace

extern "C" {

void *wm_json_new_object(void) { return wmjson::make_obj(); }
void *wm_json_load(const char *path) { return wmjson::load(path, nullptr); }
void *wm_json_parse(const char *text, int len) {
  return wmjson::parse(std::string(text, len < 0 ? strlen(text) : (size_t)len), nullptr);
}
void wm_json_free(void *root) { delete V(root); }
void *wm_json_child(void *v, const char *path) { return at(v, path); }

int wm_json_get_int(void *v, const char *path, long long *out) {
  Value *x = at(v, path);
  if (!x || (x->kind != Value::Int && x->kind != Value::Real)) return 0;
  *out = x->as_int();
  return 1;
}
int wm_json_get_real(void *v, const char *path, double *out) {
  Value *x = at(v, path);
  if (!x || (x->kind != Value::Int && x->kind != Value::Real)) return 0;
  *out = x->as_real();
  return 1;
}
int wm_json_get_bool(void *v, const char *path, int *out) {
  Value *x = at(v, path);
  if (!x || x->kind != Value::Bool) return 0;
  *out = x->b ? 1 : 0;
  return 1;
}
int wm_json_get_str(void *v, const char *path, char *buf, int cap) {
  Value *x = at(v, path);
  if (!x || x->kind != Value::Str) return -1;
  const int n = (int)x->s.size();
  if (buf && cap > 0) memcpy(buf, x->s.data(), (size_t)(n < cap ? n : cap));
  return n;
}
int wm_json_array_size(void *v, const char *path) {
  Value *x = at(v, path);
  return x && x->kind == Value::Arr ? (int)x->arr.size() : -1;
}

int wm_json_add_child(void *v, const char *name, void *child) {
  if (!v || !child) return 0;
  return add(v, name, V(child));
}
int wm_json_add_int(void *v, const char *name, long long x) { return add(v, name, wmjson::make_int(x)); }
int wm_json_add_real(void *v, const char *name, double x) { return add(v, name, wmjson::make_real(x)); }
int wm_json_add_bool(void *v, const char *name, int x) { return add(v, name, wmjson::make_bool(x != 0)); }
int wm_json_add_str(void *v, const char *name, const char *x) { return add(v, name, wmjson::make_str(x)); }
int wm_json_add_ints(void *v, const char *name, const long long *x, int n) {
  Value *a = wmjson::make_arr();
  for (int k = 0; k < n; ++k) a->add("", wmjson::make_int(x[k]));
  return add(v, name, a);
}
int wm_json_add_reals(void *v, const char *name, const double *x, int n) {
  Value *a = wmjson::make_arr();
  for (int k = 0; k < n; ++k) a->add("", wmjson::make_real(x[k]));
  return add(v, name, a);
}

int wm_json_update_str(void *v, const char *path, const char *x) { return update(v, path, wmjson::make_str(x)); }
int wm_json_update_int(void *v, const char *path, long long x) { return update(v, path, wmjson::make_int(x)); }
int wm_json_update_real(void *v, const char *path, double x) { return update(v, path, wmjson::make_real(x)); }

int wm_json_print(void *v, const char *path) { return v && wmjson::save(V(v), path) ? 1 : 0; }
int wm_json_serialize(void *v, char *buf, int cap) {
  if (!v) return -1;
  const std::string s = wmjson::dump(V(v), false);
  if (buf && cap > 0) memcpy(buf, s.data(), s.size() < (size_t)cap ? s.size() : (size_t)cap);
  return (int)s.size();
}

}  // extern "C"
