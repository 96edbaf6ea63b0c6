// wm_paraio.cpp -- the restart files of common/paraio.f90 (paraio__init,
// paraio__output, paraio__input; paraio.f90:34-426): a JSON file naming
// every attribute and dataset with its datatype, byte offset, size and shape
// (jsonio.f90 put_metadata), beside a raw file holding them at those offsets
// in the machine's byte order.  One process writes what the reference's
// ranks write collectively for nproc = 1.
//
// The raw file holds, in order: the attributes dummy_attribute (i4, the
// offset kind 8), it, nxs, nxe, ndim, np, nxgs, nxge, nygs, nyge, nsp, nproc
// (i4), delx, delt, c (f8), r(nsp), q(nsp) (f8); then the datasets up
// (f8, ndim x np x ny x nsp x nproc: the whole array, every slot of every
// row), np2 (i4, ny x nsp x nproc) and uf (f8, 6 x (nx+4) x (ny+4) x nproc,
// ghost cells included).
//
// Besides the reference's host-array calls, the device-resident state is
// written and read row by row: the rows' active records go to their slots of
// up, the rest of up is left a hole of the file (reading it back gives
// zeros where the reference's files hold whatever the slots held; np2 says
// which slots count either way), and no host copy of the whole state is made.
#include <errno.h>
#include <fcntl.h>
#include <string.h>
#include <unistd.h>

#include <string>
#include <vector>

#include "wm_internal.h"
#include "wm_json.h"

namespace {

using wmjson::Value;

int io_error(wm_ctx *ctx, const char *fmt, const char *a, const char *b = "") {
  wm_set_error(ctx, WM_ERR_ARG, fmt, a, b);
  return WM_ERR_ARG;
}

struct Raw {
  int fd = -1;
  ~Raw() {
    if (fd >= 0) close(fd);
  }
  bool write_at(long long off, const void *p, size_t n) const {
    const char *c = (const char *)p;
    while (n > 0) {
      const ssize_t k = pwrite(fd, c, n, off);
      if (k <= 0) return false;
      c += k;
      n -= (size_t)k;
      off += k;
    }
    return true;
  }
  bool read_at(long long off, void *p, size_t n) const {
    char *c = (char *)p;
    while (n > 0) {
      const ssize_t k = pread(fd, c, n, off);
      if (k <= 0) return false;
      c += k;
      n -= (size_t)k;
      off += k;
    }
    return true;
  }
};

// jsonio.f90 put_metadata: datatype, offset, size, ndim, shape, description.
Value *metadata(Value *dst, const char *name, const char *type, long long disp, long long size,
                const std::vector<long long> &shape, const char *desc) {
  Value *o = wmjson::make_obj();
  o->add("datatype", wmjson::make_str(type));
  o->add("offset", wmjson::make_int(disp));
  o->add("size", wmjson::make_int(size));
  o->add("ndim", wmjson::make_int((long long)shape.size()));
  Value *sh = wmjson::make_arr();
  for (long long v : shape) sh->add("", wmjson::make_int(v));
  o->add("shape", sh);
  o->add("description", wmjson::make_str(desc));
  return dst->add(name, o);
}

// The writer of one pair of files: attributes go to the JSON (with their
// value under "data", jsonio.f90 put_scalar_*) and to the raw file at disp.
struct Writer {
  wm_ctx *ctx;
  Raw raw;
  long long disp = 0;
  Value *root, *attr, *data;
  std::string json_path;
  bool ok = true;
  Writer(wm_ctx *c, const std::string &base, const std::string &file) : ctx(c) {
    json_path = base + ".json";
    raw.fd = open((base + ".raw").c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
    ok = raw.fd >= 0;
    root = wmjson::make_obj();
    Value *meta = root->add("meta", wmjson::make_obj());
    const int one = 1;
    meta->add("endian", wmjson::make_int(*(const unsigned char *)&one));   // 1: little endian
    meta->add("rawfile", wmjson::make_str(file + ".raw"));
    attr = root->add("attribute", wmjson::make_obj());
    data = root->add("dataset", wmjson::make_obj());
  }
  ~Writer() { delete root; }
  void i4(const char *name, int v) {
    metadata(attr, name, "i4", disp, 4, {1}, "")->add("data", wmjson::make_int(v));
    ok = ok && raw.write_at(disp, &v, 4);
    disp += 4;
  }
  void f8(const char *name, double v) {
    metadata(attr, name, "f8", disp, 8, {1}, "")->add("data", wmjson::make_real(v));
    ok = ok && raw.write_at(disp, &v, 8);
    disp += 8;
  }
  void f8s(const char *name, const double *v, int n) {
    Value *m = metadata(attr, name, "f8", disp, 8LL * n, {n}, "");
    Value *a = m->add("data", wmjson::make_arr());
    for (int k = 0; k < n; ++k) a->add("", wmjson::make_real(v[k]));
    ok = ok && raw.write_at(disp, v, 8 * (size_t)n);
    disp += 8LL * n;
  }
  // a dataset's metadata; its bytes are written by the caller at the
  // returned offset
  long long dataset(const char *name, const char *type, long long size,
                    const std::vector<long long> &shape, const char *desc) {
    metadata(data, name, type, disp, size, shape, desc);
    const long long at = disp;
    disp += size;
    return at;
  }
  int finish() {
    if (ok) ok = ftruncate(raw.fd, disp) == 0;     // the holes of up read back as zeros
    if (ok) ok = wmjson::save(root, json_path);
    if (!ok) {
      wm_set_error(ctx, WM_ERR_ARG, "paraio__output: cannot write %s (%s)", json_path.c_str(),
                   strerror(errno));
      return WM_ERR_ARG;
    }
    return WM_OK;
  }
};

// paraio.f90:143-156 and put_metadata (:956-1002): the attributes in order.
void attributes(Writer &W, const wm_ctx *ctx, int it, int nxs, int nxe) {
  const wm_grid &g = ctx->g;
  W.i4("dummy_attribute", 8);                    // MPI_OFFSET_KIND
  W.i4("it", it);
  W.i4("nxs", nxs);
  W.i4("nxe", nxe);
  W.i4("ndim", g.ndim);
  W.i4("np", g.np);
  W.i4("nxgs", g.nxgs);
  W.i4("nxge", g.nxge);
  W.i4("nygs", g.nygs);
  W.i4("nyge", g.nyge);
  W.i4("nsp", g.nsp);
  W.i4("nproc", ctx->io_nproc);
  W.f8("delx", g.delx);
  W.f8("delt", g.delt);
  W.f8("c", g.c);
  W.f8s("r", g.r, g.nsp);
  W.f8s("q", g.q, g.nsp);
}

struct Dims {
  long long up, np2, uf;        // elements
  long long nxg, nyl;
};
Dims dims_of(const wm_ctx *ctx) {
  const wm_grid &g = ctx->g;
  Dims d;
  d.nyl = g.ny + 4;
  d.nxg = g.nx + 4;
  d.up = (long long)g.ndim * g.np * g.ny * g.nsp;
  d.np2 = (long long)g.ny * g.nsp;
  d.uf = 6LL * d.nxg * d.nyl;
  return d;
}

int require_io(wm_ctx *ctx) {
  if (!ctx->ioinit) {
    wm_set_error(ctx, WM_ERR_NOT_INIT, "Initialize first by calling paraio__init()");
    return WM_ERR_NOT_INIT;
  }
  return WM_OK;
}

// The reader: the JSON, the raw file, and the checks of paraio.f90:262-330.
struct Reader {
  wm_ctx *ctx;
  Value *root = nullptr;
  Raw raw;
  ~Reader() { delete root; }
  int open_files(const std::string &base) {
    std::string err;
    root = wmjson::load(base + ".json", &err);
    if (!root) return io_error(ctx, "paraio__input: %s (%s)", (base + ".json").c_str(), err.c_str());
    raw.fd = ::open((base + ".raw").c_str(), O_RDONLY);
    if (raw.fd < 0) return io_error(ctx, "paraio__input: cannot open %s%s", (base + ".raw").c_str());
    return WM_OK;
  }
  Value *meta(const char *group, const char *name) {
    Value *g = root->member(group);
    return g ? g->member(name) : nullptr;
  }
  template <class T>
  int attr(const char *name, T *out, int n = 1) {
    Value *m = meta("attribute", name);
    Value *off = m ? m->member("offset") : nullptr;
    if (!off || !raw.read_at(off->as_int(), out, sizeof(T) * n))
      return io_error(ctx, "paraio__input: attribute %s missing or unreadable%s", name);
    return WM_OK;
  }
  // dataset offset after its shape check (nd entries against want)
  int dataset(const char *name, const std::vector<long long> &want, long long *off) {
    Value *m = meta("dataset", name);
    Value *nd = m ? m->member("ndim") : nullptr, *sh = m ? m->member("shape") : nullptr;
    bool ok = nd && sh && nd->as_int() == (long long)want.size() && sh->kind == Value::Arr &&
              sh->arr.size() >= want.size();
    for (size_t k = 0; ok && k < want.size(); ++k) ok = sh->arr[k]->as_int() == want[k];
    if (!ok) {
      wm_set_error(ctx, WM_ERR_ARG, "Fatal error in reading particle data");
      return WM_ERR_ARG;
    }
    *off = m->member("offset")->as_int();
    return WM_OK;
  }
  // paraio.f90:262-322: the attributes, checked against the context's grid
  int header(int *nxs, int *nxe, int *it) {
    const wm_grid &g = ctx->g;
    int v[8];
    WM_TRY(attr("it", it));
    WM_TRY(attr("nxs", nxs));
    WM_TRY(attr("nxe", nxe));
    const char *names[8] = {"ndim", "np", "nxgs", "nxge", "nygs", "nyge", "nsp", "nproc"};
    for (int k = 0; k < 8; ++k) WM_TRY(attr(names[k], &v[k]));
    if (!(v[2] == g.nxgs && v[3] == g.nxge && v[4] == g.nygs && v[5] == g.nyge && v[1] == g.np &&
          v[6] == g.nsp && v[7] == ctx->io_nproc)) {
      wm_set_error(ctx, WM_ERR_ARG, "*** Error: invalid input parameters ***");
      return WM_ERR_ARG;
    }
    return WM_OK;
  }
};

std::string base_of(const wm_ctx *ctx, const char *filename) { return ctx->iodir + filename; }

// put_metadata (paraio.f90:956-1002) after dummy_attribute and it: the grid
// and the physical constants.
void common_attributes(Writer &W, const wm_ctx *ctx, int it) {
  const wm_grid &g = ctx->g;
  W.i4("dummy_attribute", 8);
  W.i4("it", it);
  W.i4("ndim", g.ndim);
  W.i4("np", g.np);
  W.i4("nxgs", g.nxgs);
  W.i4("nxge", g.nxge);
  W.i4("nygs", g.nygs);
  W.i4("nyge", g.nyge);
  W.i4("nsp", g.nsp);
  W.i4("nproc", ctx->io_nproc);
  W.f8("delx", g.delx);
  W.f8("delt", g.delt);
  W.f8("c", g.c);
  W.f8s("r", g.r, g.nsp);
  W.f8s("q", g.q, g.nsp);
}

std::string numbered(int it, const char *suffix) {
  char b[64];
  snprintf(b, sizeof b, "%07d%s", it, suffix);
  return b;
}

// write_particle (paraio.f90:738-863): the interior fields, then per species
// the active particles (mode 0) or those with a positive ID (mode 1, the
// tracers), rows in order, as (ndim, count) records.
int write_particles(wm_ctx *ctx, const double *up, const double *uf, const int *np2, int it,
                    int mode, const char *suffix) {
  WM_TRY(require_io(ctx));
  const wm_grid &g = ctx->g;
  const std::string file = numbered(it, suffix);
  Writer W(ctx, base_of(ctx, file.c_str()), file);
  common_attributes(W, ctx, it);
  const long long gx = g.nx + 4;
  std::vector<double> f(6 * (size_t)g.nx * g.ny);
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i)
      for (int k = 0; k < 6; ++k)
        f[((size_t)j * g.nx + i) * 6 + k] = uf[((size_t)(j + 2) * gx + (i + 2)) * 6 + k];
  long long at = W.dataset("uf", "f8", 8LL * (long long)f.size(), {6, g.nx, g.ny},
                           "electromagnetic field");
  W.ok = W.ok && W.raw.write_at(at, f.data(), 8 * f.size());
  for (int s = 0; s < g.nsp; ++s) {
    std::vector<double> buf;
    for (int j = 0; j < g.ny; ++j)
      for (int k = 0; k < np2[(size_t)s * g.ny + j]; ++k) {
        const double *r = up + (((size_t)s * g.ny + j) * g.np + k) * g.ndim;
        long long id;
        memcpy(&id, r + g.ndim - 1, 8);
        if (mode == 0 || id > 0) buf.insert(buf.end(), r, r + g.ndim);
      }
    char name[16], desc[32];
    snprintf(name, sizeof name, "up%02d", s + 1);
    snprintf(desc, sizeof desc, "particle species #%02d", s + 1);
    const long long n = (long long)buf.size() / g.ndim;
    at = W.dataset(name, "f8", 8LL * (long long)buf.size(), {g.ndim, n}, desc);
    W.ok = W.ok && (buf.empty() || W.raw.write_at(at, buf.data(), 8 * buf.size()));
  }
  return W.finish();
}

}  // namespace

extern "C" {

int wm_paraio_init(wm_ctx *ctx, int ndim, int np, int nsp, int nxgs, int nxge, int nygs, int nyge,
                   int nys, int nye, int nproc, int nrank, double delx, double delt, double c,
                   const double *q, const double *r, const char *dir) {
  if (!ctx || !dir || !q || !r) return WM_ERR_ARG;
  const wm_grid &g = ctx->g;
  if (!ctx->have_grid || g.ndim != ndim || g.np != np || g.nsp != nsp || g.nxgs != nxgs ||
      g.nxge != nxge || g.nygs != nygs || g.nyge != nyge || g.nys != nys || g.nye != nye) {
    wm_set_error(ctx, WM_ERR_ARG, "paraio__init: the grid differs from the modules' (or none yet)");
    return WM_ERR_ARG;
  }
  if (nproc != 1 || nrank != 0) {
    wm_set_error(ctx, WM_ERR_ARG, "paraio__init: one process holds the whole domain (nproc=%d)",
                 nproc);
    return WM_ERR_ARG;
  }
  (void)delx;
  (void)delt;
  (void)c;
  ctx->iodir = dir;
  ctx->io_nproc = nproc;
  ctx->io_nrank = nrank;
  ctx->ioinit = true;
  return WM_OK;
}

int wm_paraio_output(wm_ctx *ctx, const double *up, const double *uf, const int *np2, int nxs,
                     int nxe, int it, const char *filename) {
  if (!ctx || !up || !uf || !np2 || !filename) return WM_ERR_ARG;
  WM_TRY(require_io(ctx));
  const Dims d = dims_of(ctx);
  Writer W(ctx, base_of(ctx, filename), filename);
  attributes(W, ctx, it, nxs, nxe);
  const wm_grid &g = ctx->g;
  long long at = W.dataset("up", "f8", 8 * d.up, {g.ndim, g.np, g.ny, g.nsp, 1}, "particles");
  W.ok = W.ok && W.raw.write_at(at, up, 8 * (size_t)d.up);
  at = W.dataset("np2", "i4", 4 * d.np2, {g.ny, g.nsp, 1}, "number of active particles");
  W.ok = W.ok && W.raw.write_at(at, np2, 4 * (size_t)d.np2);
  at = W.dataset("uf", "f8", 8 * d.uf, {6, d.nxg, d.nyl, 1},
                 "electromagnetic fields including ghost cells");
  W.ok = W.ok && W.raw.write_at(at, uf, 8 * (size_t)d.uf);
  return W.finish();
}

int wm_paraio_input(wm_ctx *ctx, double *up, double *uf, int *np2, int *nxs, int *nxe, int *it,
                    const char *filename) {
  if (!ctx || !up || !uf || !np2 || !nxs || !nxe || !it || !filename) return WM_ERR_ARG;
  WM_TRY(require_io(ctx));
  const Dims d = dims_of(ctx);
  const wm_grid &g = ctx->g;
  Reader R{ctx};
  WM_TRY(R.open_files(base_of(ctx, filename)));
  WM_TRY(R.header(nxs, nxe, it));
  long long off;
  WM_TRY(R.dataset("up", {g.ndim, g.np, g.ny, g.nsp, 1}, &off));
  if (!R.raw.read_at(off, up, 8 * (size_t)d.up)) return io_error(ctx, "paraio__input: short %s%s", "up");
  WM_TRY(R.dataset("np2", {g.ny, g.nsp, 1}, &off));
  if (!R.raw.read_at(off, np2, 4 * (size_t)d.np2)) return io_error(ctx, "paraio__input: short %s%s", "np2");
  WM_TRY(R.dataset("uf", {6, d.nxg, d.nyl, 1}, &off));
  if (!R.raw.read_at(off, uf, 8 * (size_t)d.uf)) return io_error(ctx, "paraio__input: short %s%s", "uf");
  return WM_OK;
}

int wm_paraio_param(wm_ctx *ctx, int n0, double wpe, double wpi, double wge, double wgi,
                    double vti, double vte, const char *filename) {
  if (!ctx || !filename) return WM_ERR_ARG;
  WM_TRY(require_io(ctx));
  const wm_grid &g = ctx->g;
  Writer W(ctx, base_of(ctx, filename), filename);
  // paraio.f90:478-533, in that order
  W.i4("dummy_attribute", 8);
  W.i4("nx", g.nx);
  W.i4("ny", g.nyge - g.nygs + 1);
  W.i4("np", g.np);
  W.f8("delx", g.delx);
  W.f8("delt", g.delt);
  W.f8("c", g.c);
  W.f8s("r", g.r, g.nsp);
  W.f8s("q", g.q, g.nsp);
  W.f8("wpe", wpe);
  W.f8("wge", wge);
  W.f8("wpi", wpi);
  W.f8("wgi", wgi);
  W.f8("vti", vti);
  W.f8("vte", vte);
  W.f8("vai", g.c * wgi / wpi);
  W.f8("vae", g.c * wge / wpe);
  W.i4("n0", n0);
  return W.finish();
}

int wm_paraio_mom(wm_ctx *ctx, double *mom, const double *uf, int it) {
  if (!ctx || !mom || !uf) return WM_ERR_ARG;
  WM_TRY(require_io(ctx));
  const wm_grid &g = ctx->g;
  const std::string file = numbered(it, "_mom");
  Writer W(ctx, base_of(ctx, file.c_str()), file);
  common_attributes(W, ctx, it);
  const long long mx = g.nx + 2, my = g.ny + 2, gx = g.nx + 4;
  auto M = [&](int k, int i, int j, int s) -> double & {     // mom(k+1, nxgs-1+i, nys-1+j, s+1)
    return mom[(((size_t)s * my + j) * mx + i) * 7 + k];
  };
  // paraio.f90:625-680: velocities and temperatures divided by the density in
  // place over the whole array, as the reference does
  for (int s = 0; s < g.nsp; ++s)
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i)
        for (int k = 1; k < 7; ++k) M(k, i, j, s) = M(k, i, j, s) / M(0, i, j, s);
  std::vector<double> den, vel, temp;
  for (int s = 0; s < g.nsp; ++s)
    for (int j = 1; j <= g.ny; ++j)
      for (int i = 1; i <= g.nx; ++i) {
        den.push_back(M(0, i, j, s));
        for (int k = 1; k <= 3; ++k) vel.push_back(M(k, i, j, s));
        for (int k = 4; k <= 6; ++k) temp.push_back(M(k, i, j, s));
      }
  long long at = W.dataset("den", "f8", 8LL * (long long)den.size(), {g.nx, g.ny, g.nsp}, "density");
  W.ok = W.ok && W.raw.write_at(at, den.data(), 8 * den.size());
  at = W.dataset("vel", "f8", 8LL * (long long)vel.size(), {3, g.nx, g.ny, g.nsp}, "velocity");
  W.ok = W.ok && W.raw.write_at(at, vel.data(), 8 * vel.size());
  at = W.dataset("temp", "f8", 8LL * (long long)temp.size(), {3, g.nx, g.ny, g.nsp}, "temperature");
  W.ok = W.ok && W.raw.write_at(at, temp.data(), 8 * temp.size());
  // cell-centred fields, paraio.f90:682-692 (Bz takes (i+1, j) twice there)
  auto F = [&](int k, int i, int j) { return uf[((size_t)(j + 2) * gx + (i + 2)) * 6 + k]; };
  std::vector<double> t;
  for (int j = 0; j < g.ny; ++j)
    for (int i = 0; i < g.nx; ++i) {
      t.push_back((F(0, i, j) + F(0, i, j + 1)) / 2);
      t.push_back((F(1, i, j) + F(1, i + 1, j)) / 2);
      t.push_back((F(2, i, j) + F(2, i + 1, j) + F(2, i + 1, j) + F(2, i + 1, j + 1)) / 4);
      t.push_back((F(3, i, j) + F(3, i + 1, j)) / 2);
      t.push_back((F(4, i, j) + F(4, i, j + 1)) / 2);
      t.push_back(F(5, i, j));
    }
  at = W.dataset("uf", "f8", 8LL * (long long)t.size(), {6, g.nx, g.ny}, "electromagnetic field");
  W.ok = W.ok && W.raw.write_at(at, t.data(), 8 * t.size());
  return W.finish();
}

int wm_paraio_ptcl(wm_ctx *ctx, const double *up, const double *uf, const int *np2, int it) {
  if (!ctx || !up || !uf || !np2) return WM_ERR_ARG;
  return write_particles(ctx, up, uf, np2, it, 0, "_ptcl");
}

int wm_paraio_orb(wm_ctx *ctx, const double *up, const double *uf, const int *np2, int it) {
  if (!ctx || !up || !uf || !np2) return WM_ERR_ARG;
  return write_particles(ctx, up, uf, np2, it, 1, "_orb");
}

int wm_paraio_output_state(wm_ctx *ctx, int it, const char *filename) {
  if (!ctx || !filename) return WM_ERR_ARG;
  WM_TRY(require_io(ctx));
  if (!ctx->resident) {
    wm_set_error(ctx, WM_ERR_NOT_INIT, "paraio output: no device-resident state");
    return WM_ERR_NOT_INIT;
  }
  const wm_grid &g = ctx->g;
  const Dims d = dims_of(ctx);
  const int nseg = g.nsp * g.ny;
  std::vector<int> n2(nseg);
  WM_TRY(wm_state_np2(ctx, n2.data()));
  Writer W(ctx, base_of(ctx, filename), filename);
  attributes(W, ctx, it, g.nxgs, g.nxge);
  const long long up_at = W.dataset("up", "f8", 8 * d.up, {g.ndim, g.np, g.ny, g.nsp, 1}, "particles");
  for (int s0 = 0; s0 < nseg && W.ok;) {
    int w = 1;
    const int s1 = wm_batch_end(ctx, n2.data(), s0, &w);
    std::vector<double> h((size_t)(s1 - s0) * w * g.ndim);
    WM_TRY(wm_rows_out(ctx, s0, s1, w, h.data()));
    for (int s = s0; s < s1 && W.ok; ++s)
      W.ok = W.raw.write_at(up_at + 8LL * s * g.np * g.ndim, &h[(size_t)(s - s0) * w * g.ndim],
                            8 * (size_t)n2[s] * g.ndim);
    s0 = s1;
  }
  long long at = W.dataset("np2", "i4", 4 * d.np2, {g.ny, g.nsp, 1}, "number of active particles");
  W.ok = W.ok && W.raw.write_at(at, n2.data(), 4 * (size_t)d.np2);
  std::vector<double> f((size_t)d.uf);
  WM_TRY(wm_field_out(ctx, f.data()));
  at = W.dataset("uf", "f8", 8 * d.uf, {6, d.nxg, d.nyl, 1},
                 "electromagnetic fields including ghost cells");
  W.ok = W.ok && W.raw.write_at(at, f.data(), 8 * (size_t)d.uf);
  return W.finish();
}

int wm_paraio_input_state(wm_ctx *ctx, int *it, const char *filename) {
  if (!ctx || !it || !filename) return WM_ERR_ARG;
  WM_TRY(require_io(ctx));
  const wm_grid &g = ctx->g;
  const Dims d = dims_of(ctx);
  const int nseg = g.nsp * g.ny;
  Reader R{ctx};
  WM_TRY(R.open_files(base_of(ctx, filename)));
  int nxs = 0, nxe = 0;
  WM_TRY(R.header(&nxs, &nxe, it));
  long long up_at, n2_at, uf_at;
  WM_TRY(R.dataset("up", {g.ndim, g.np, g.ny, g.nsp, 1}, &up_at));
  WM_TRY(R.dataset("np2", {g.ny, g.nsp, 1}, &n2_at));
  WM_TRY(R.dataset("uf", {6, d.nxg, d.nyl, 1}, &uf_at));
  std::vector<int> n2(nseg);
  if (!R.raw.read_at(n2_at, n2.data(), 4 * (size_t)nseg))
    return io_error(ctx, "paraio__input: short %s%s", "np2");
  for (int s = 0; s < nseg; ++s)
    if (n2[s] < 0 || n2[s] > g.np) {
      wm_set_error(ctx, WM_ERR_RANGE, "paraio__input: np2 = %d of row segment %d outside 0..np", n2[s], s);
      return WM_ERR_RANGE;
    }
  WM_TRY(wm_rows_begin(ctx));
  for (int s0 = 0; s0 < nseg;) {
    int w = 1;
    const int s1 = wm_batch_end(ctx, n2.data(), s0, &w);
    std::vector<double> h((size_t)(s1 - s0) * w * g.ndim);
    for (int s = s0; s < s1; ++s)
      if (!R.raw.read_at(up_at + 8LL * s * g.np * g.ndim, &h[(size_t)(s - s0) * w * g.ndim],
                         8 * (size_t)n2[s] * g.ndim))
        return io_error(ctx, "paraio__input: short %s%s", "up");
    WM_TRY(wm_rows_in(ctx, s0, s1, w, n2.data(), h.data()));
    s0 = s1;
  }
  WM_TRY(wm_rows_done(ctx, n2.data()));
  std::vector<double> f((size_t)d.uf);
  if (!R.raw.read_at(uf_at, f.data(), 8 * (size_t)d.uf))
    return io_error(ctx, "paraio__input: short %s%s", "uf");
  return wm_field_in(ctx, f.data());
}

}  // extern "C"
