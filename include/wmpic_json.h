/*
 * wmpic_json.h -- C ABI of the JSON document model in libwmpic, for the
 * Fortran module json_module that the reference's apps use (a subset of the
 * json-fortran 8.2.1 API vendored in the reference's utils/json: json_core,
 * json_file, json_value).  No counterpart in the reference's hot path: the
 * apps read their configuration and write their outputs' metadata through
 * it (proj/weibel/app.f90:162-258, 550-627).
 *
 * Values are handles; a document is freed through its root.  Paths name
 * members separated by '.', array elements as name(k), 1-based.  Getters
 * return 1 when the path exists and holds a value of a convertible kind,
 * else 0 and leave *out alone.
 */
#ifndef WMPIC_JSON_H
#define WMPIC_JSON_H

#ifdef __cplusplus
extern "C" {
#endif

void *wm_json_new_object(void);
void *wm_json_load(const char *path);            /* null if unreadable or malformed */
void *wm_json_parse(const char *text, int len);  /* null if malformed */
void wm_json_free(void *root);
void *wm_json_child(void *v, const char *path);  /* null if absent */
int wm_json_get_int(void *v, const char *path, long long *out);
int wm_json_get_real(void *v, const char *path, double *out);
int wm_json_get_bool(void *v, const char *path, int *out);
/* copies at most cap bytes of the string; returns its full length, -1 if
 * absent or not a string */
int wm_json_get_str(void *v, const char *path, char *buf, int cap);
int wm_json_array_size(void *v, const char *path);
/* adds a member (obj) or element (arr); child is reparented to v */
int wm_json_add_child(void *v, const char *name, void *child);
int wm_json_add_int(void *v, const char *name, long long x);
int wm_json_add_real(void *v, const char *name, double x);
int wm_json_add_bool(void *v, const char *name, int x);
int wm_json_add_str(void *v, const char *name, const char *x);
int wm_json_add_ints(void *v, const char *name, const long long *x, int n);
int wm_json_add_reals(void *v, const char *name, const double *x, int n);
/* replaces the value at path (member of an object) by x; returns 1 if the
 * member existed, 0 if it was added */
int wm_json_update_str(void *v, const char *path, const char *x);
int wm_json_update_int(void *v, const char *path, long long x);
int wm_json_update_real(void *v, const char *path, double x);
int wm_json_print(void *v, const char *path);    /* pretty text to a file; 0 on error */
/* compact text of v: copies at most cap bytes, returns the full length */
int wm_json_serialize(void *v, char *buf, int cap);

#ifdef __cplusplus
}
#endif
#endif
