// extern "C" boundary of libx3d: context management and the Fortran-mirror
// entries of derive.f90 / poisson.f90 / 2DECOMP transposes.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>

#include "x3d_internal.h"

namespace x3d {

namespace {
thread_local std::string g_err;
std::unique_ptr<Context> g_ctx;
std::mutex g_ctx_mu;

// coefficient sets handed in by a caller that are not one of the library's
// own (uploaded on first use and kept)
std::vector<TriSet> g_foreign;
// metric arrays passed by callers that differ from x3d_init's (e.g. computed
// by another libm): host values in the precision given, and their device copy
struct ForeignMetric {
  std::vector<double> v64;
  std::vector<float> v32;
  void *dev = nullptr;
};
std::vector<ForeignMetric> g_foreign_metric;
}  // namespace

void set_error(const std::string &m) { g_err = m; }
int fail(int code, const std::string &m) {
  g_err = m;
  return code;
}
Context *ctx() { return g_ctx.get(); }
int require_ctx(Context **c) {
  *c = g_ctx.get();
  if (!*c) return fail(X3D_ERR_STATE, "x3d_init has not been called");
  return X3D_OK;
}

template <typename T>
int launch_reorder_ji(const T *src, T *dst, int n1, int n2, int n3, hipStream_t st);

namespace {

void sm_vector(const std::vector<double> &ff, const std::vector<double> &fs, const std::vector<double> &fw,
               double alfa, std::vector<double> &r, double &den) {
  const int n = (int)ff.size();
  r.assign(n, 0.0);
  r[0] = -1.0;
  r[n - 1] = alfa;
  for (int i = 1; i < n; ++i) r[i] = r[i] - r[i - 1] * fs[i];
  r[n - 1] = r[n - 1] * fw[n - 1];
  for (int i = n - 2; i >= 0; --i) r[i] = (r[i] - ff[i] * r[i + 1]) * fw[i];
  den = 1.0 + r[0] - alfa * r[n - 1];
}

// Device tridiagonal set for the arrays (ff,fs,fw) a caller passed.
template <typename T>
int resolve_set(Context &c, int dir, int n, const T *ff, const T *fs, const T *fw, bool periodic, double alfa,
                TriDev<T> *out) {
  const TriSet *s = nullptr;
  if constexpr (sizeof(T) == 8) s = match_set(c, dir, n, ff, fs, fw);
  else s = match_set32(c, dir, n, ff, fs, fw);
  if (s && (!periodic || (s->periodic && s->alfa == alfa))) {
    if constexpr (sizeof(T) == 8) *out = s->dev64();
    else *out = s->dev32();
    out->periodic = periodic ? 1 : 0;
    return X3D_OK;
  }
  // foreign set: look it up by content, else upload it
  for (const TriSet &f : g_foreign) {
    if ((int)f.ff.size() != n || (f.periodic != (periodic ? 1 : 0)) || (periodic && f.alfa != alfa)) continue;
    bool eq = true;
    for (int i = 0; i < n && eq; ++i) eq = (T)f.ff[i] == ff[i] && (T)f.fs[i] == fs[i] && (T)f.fw[i] == fw[i];
    if (eq) {
      if constexpr (sizeof(T) == 8) *out = f.dev64();
      else *out = f.dev32();
      return X3D_OK;
    }
  }
  if (g_foreign.size() >= 64) return fail(X3D_ERR_ARG, "too many distinct coefficient arrays");
  TriSet f;
  f.name = "foreign";
  f.ff.assign(ff, ff + n);
  f.fs.assign(fs, fs + n);
  f.fw.assign(fw, fw + n);
  f.periodic = periodic ? 1 : 0;
  f.alfa = alfa;
  if (periodic) sm_vector(f.ff, f.fs, f.fw, alfa, f.sm, f.sm_den);
  else f.sm.assign(n, 0.0);
  X3D_TRY(upload_set(f));
  g_foreign.push_back(std::move(f));
  if constexpr (sizeof(T) == 8) *out = g_foreign.back().dev64();
  else *out = g_foreign.back().dev32();
  return X3D_OK;
}

int check_dims(const int *a, const int *b, const int *c) {
  if (!a || !b || !c) return fail(X3D_ERR_ARG, "null size argument");
  if (*a <= 0 || *b <= 0 || *c <= 0) return fail(X3D_ERR_ARG, "non-positive size");
  return X3D_OK;
}

// derx/dery/derz/derxx/deryy/derzz_<bc1><bcn> (derive.f90:7-3795)
// The metric a y derivative multiplies by on a stretched mesh (derive.f90:
// 409-417 ppy for dery, :4572-4578 ppyi for deryvp, :4905-4911 ppy for
// derypv), as the reference does with the array its caller passes: the
// context's device copy when the values are x3d_init's (or none is passed),
// else the caller's values, uploaded once and kept by content.
template <typename T>
int y_metric(const Context &c, bool pressure, const T *caller, int n, const T **dev) {
  *dev = nullptr;
  if (c.p.istret == 0) return X3D_OK;
  const std::vector<double> &m = pressure ? c.ppyi : c.ppy;
  if (n > (int)m.size()) return fail(X3D_ERR_ARG, "stretched y: line longer than the metric");
  bool same = true;
  if (caller)
    for (int j = 0; j < n && same; ++j) same = caller[j] == (T)m[j];
  if (same) {
    if constexpr (sizeof(T) == 8) *dev = (const T *)(pressure ? c.ppyi64 : c.ppy64);
    else *dev = (const T *)(pressure ? c.ppyi32 : c.ppy32);
    return X3D_OK;
  }
  for (const ForeignMetric &f : g_foreign_metric) {
    bool hit;
    if constexpr (sizeof(T) == 8) hit = f.v64.size() == (size_t)n && std::equal(f.v64.begin(), f.v64.end(), caller);
    else hit = f.v32.size() == (size_t)n && std::equal(f.v32.begin(), f.v32.end(), caller);
    if (hit) {
      *dev = (const T *)f.dev;
      return X3D_OK;
    }
  }
  if (g_foreign_metric.size() >= 64) return fail(X3D_ERR_ARG, "too many distinct metric arrays");
  ForeignMetric f;
  if constexpr (sizeof(T) == 8) f.v64.assign(caller, caller + n);
  else f.v32.assign(caller, caller + n);
  X3D_HIP(hipMalloc(&f.dev, n * sizeof(T)));
  X3D_HIP(hipMemcpy(f.dev, caller, n * sizeof(T), hipMemcpyHostToDevice));
  g_foreign_metric.push_back(std::move(f));
  *dev = (const T *)g_foreign_metric.back().dev;
  return X3D_OK;
}

template <typename T>
int der_entry(int dir, int op, int bc1, int bcn, T *t, const T *u, const T *ff, const T *fs, const T *fw,
              const int *nx, const int *ny, const int *nz, const int *npaire, const T *pp = nullptr) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  X3D_TRY(check_dims(nx, ny, nz));
  if (!t || !u || !ff || !fs || !fw || !npaire) return fail(X3D_ERR_ARG, "null pointer argument");
  if (*npaire != 0 && *npaire != 1) return fail(X3D_ERR_ARG, "npaire must be 0 or 1");
  const int n = dir == 0 ? *nx : (dir == 1 ? *ny : *nz);
  const DirCoef &D = c->dir[dir];
  const bool periodic = bc1 == 0;
  const double alfa = op == OP_DER1 ? D.s.alfai : D.s.alsai;
  LineCall<T> L;
  L.dir = dir;
  L.op = op;
  L.nx = *nx;
  L.ny = *ny;
  L.nz = *nz;
  L.n_out = n;
  L.u = u;
  L.t = t;
  X3D_TRY(resolve_set<T>(*c, dir, n, ff, fs, fw, periodic, alfa, &L.tri));
  L.s = cast_sten<T>(D.s);
  L.bc1 = bc1;
  L.bcn = bcn;
  L.npaire = *npaire;
  // dery multiplies by ppy; deryy does not (its caller combines pp2y and
  // pp4y, transeq.f90:337-346)
  if (dir == 1 && op == OP_DER1 && c->p.istret != 0) X3D_TRY(y_metric<T>(*c, false, pp, n, &L.pp));
  X3D_TRY(launch_line<T>(L, c->stream));
  X3D_HIP(hipStreamSynchronize(c->stream));
  return X3D_OK;
}

// staggered operators (derive.f90:3796-5615); periodicity follows the
// module flag of the direction (nclx/ncly/nclz), as in the reference
template <typename T>
int stag_entry(int dir, int op, T *t, const T *u, const T *cf1, const T *cs1, const T *cw1, const T *cf2,
               const T *cs2, const T *cw2, int n1, int n2, int n3, int n_out, int npaire,
               const T *pp = nullptr) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!t || !u || !cf1 || !cs1 || !cw1) return fail(X3D_ERR_ARG, "null pointer argument");
  if (npaire != 0 && npaire != 1) return fail(X3D_ERR_ARG, "npaire must be 0 or 1");
  const DirCoef &D = c->dir[dir];
  const bool periodic = D.bc1 == 0;
  if (!periodic && npaire == 0 && op != OP_DERVP)
    return fail(X3D_ERR_UNSUPPORTED, "the reference has no odd branch for this non-periodic operator");
  const T *ff = cf1, *fs = cs1, *fw = cw1;
  if ((op == OP_DERPV || op == OP_INTERPV) && periodic) {
    if (!cf2 || !cs2 || !cw2) return fail(X3D_ERR_ARG, "null pointer argument");
    ff = cf2; fs = cs2; fw = cw2;
  }
  const double alfa = (op == OP_DERVP || op == OP_DERPV) ? D.s.alcai : D.s.ailcai;
  LineCall<T> L;
  L.dir = dir;
  L.op = op;
  L.nx = n1;
  L.ny = n2;
  L.nz = n3;
  L.n_out = n_out;
  L.u = u;
  L.t = t;
  X3D_TRY(resolve_set<T>(*c, dir, n_out, ff, fs, fw, periodic, alfa, &L.tri));
  L.s = cast_sten<T>(D.s);
  L.bc1 = D.bc1;
  L.bcn = D.bcn;
  L.npaire = npaire;
  if (dir == 1 && (op == OP_DERVP || op == OP_DERPV))
    X3D_TRY(y_metric<T>(*c, op == OP_DERVP, pp, n_out, &L.pp));
  X3D_TRY(launch_line<T>(L, c->stream));
  X3D_HIP(hipStreamSynchronize(c->stream));
  return X3D_OK;
}

int find_array(const Context &c, const char *name, const std::vector<double> **out) {
  if (c.p.istret != 0) {
    // stretching.f90 outputs (the reference's module arrays yp, ypi, ppy,
    // ppyi, pp2y, pp4y)
    const std::string s = name;
    const std::vector<double> *m = s == "yp" ? &c.yp : s == "ypi" ? &c.ypi : s == "ppy" ? &c.ppy
                                 : s == "ppyi" ? &c.ppyi : s == "pp2y" ? &c.pp2y : s == "pp4y" ? &c.pp4y
                                 : nullptr;
    if (m) {
      *out = m;
      return X3D_OK;
    }
  }
  for (int d = 0; d < 3; ++d)
    for (const auto &kv : c.dir[d].arrays)
      if (kv.first == name) {
        *out = &kv.second;
        return X3D_OK;
      }
  return fail(X3D_ERR_ARG, std::string("unknown coefficient array ") + name);
}

}  // namespace
}  // namespace x3d

using namespace x3d;

extern "C" {

const char *x3d_last_error(void) { return g_err.c_str(); }

int x3d_init(const x3d_params *p) {
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  if (!p) return fail(X3D_ERR_ARG, "null params");
  if (g_ctx) return fail(X3D_ERR_STATE, "x3d_init called twice (call x3d_finalize first)");
  auto c = std::make_unique<Context>();
  c->p = *p;
  if (c->p.ipinter == 0) c->p.ipinter = 3;   // parameters.f90:708 default
  if (c->p.ifirstder == 0) c->p.ifirstder = 4;
  if (c->p.isecondder == 0) c->p.isecondder = 4;
  X3D_HIP(hipGetDevice(&c->device));
  X3D_TRY(build_schemes(*c));
  X3D_TRY(upload_sets(*c));
  {
    int r = build_stretching(*c);
    if (r != X3D_OK) {
      free_stretching(*c);
      free_sets(*c);
      return r;
    }
  }
  // a grid without a Poisson solver (a boundary combination the reference
  // does not solve, sizes the FFT passes do not cover) still gets the
  // derivative entries
  {
    int r = poisson_setup(*c);
    if (r != X3D_OK && r != X3D_ERR_UNSUPPORTED) {
      free_stretching(*c);
      free_sets(*c);
      return r;
    }
  }
  g_ctx = std::move(c);
  return X3D_OK;
}

int x3d_finalize(void) {
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  if (!g_ctx) return X3D_OK;
  (void)hipDeviceSynchronize();
  poisson_free(*g_ctx);
  free_stretching(*g_ctx);
  free_sets(*g_ctx);
  for (TriSet &f : g_foreign) {
    if (f.d64) (void)hipFree(f.d64);
    if (f.d32) (void)hipFree(f.d32);
  }
  g_foreign.clear();
  for (ForeignMetric &f : g_foreign_metric)
    if (f.dev) (void)hipFree(f.dev);
  g_foreign_metric.clear();
  if (g_ctx->host_stage) (void)hipFree(g_ctx->host_stage);
  g_ctx.reset();
  return X3D_OK;
}

int x3d_set_stream(void *s) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  c->stream = static_cast<hipStream_t>(s);
  return X3D_OK;
}

int x3d_synchronize(void) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  X3D_HIP(hipStreamSynchronize(c->stream));
  return X3D_OK;
}

int x3d_device_name(char *buf, int len) {
  if (!buf || len <= 0) return fail(X3D_ERR_ARG, "bad buffer");
  int dev = 0;
  X3D_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  X3D_HIP(hipGetDeviceProperties(&prop, dev));
  std::snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  return X3D_OK;
}

int x3d_get_coefficients(const char *name, double *out, int cap, int *n) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!name || !n) return fail(X3D_ERR_ARG, "null argument");
  const std::vector<double> *v = nullptr;
  X3D_TRY(find_array(*c, name, &v));
  *n = (int)v->size();
  if (out) {
    if (cap < (int)v->size()) return fail(X3D_ERR_ARG, "output buffer too small");
    std::memcpy(out, v->data(), v->size() * sizeof(double));
  }
  return X3D_OK;
}

int x3d_schemes_host(const x3d_params *p, const char *name, double *out, int cap, int *n) {
  if (!p || !name || !n) return fail(X3D_ERR_ARG, "null argument");
  Context c;
  c.p = *p;
  if (c.p.ipinter == 0) c.p.ipinter = 3;
  if (c.p.ifirstder == 0) c.p.ifirstder = 4;
  if (c.p.isecondder == 0) c.p.isecondder = 4;
  X3D_TRY(build_schemes(c));
  X3D_TRY(compute_stretching(c));
  const std::string s = name;
  for (int d = 0; d < 3; ++d)
    for (const auto &kv : c.dir[d].scalars)
      if (kv.first == s) {
        *n = 1;
        if (out && cap >= 1) out[0] = kv.second;
        return X3D_OK;
      }
  const std::vector<double> *v = nullptr;
  X3D_TRY(find_array(c, name, &v));
  *n = (int)v->size();
  if (out) {
    if (cap < (int)v->size()) return fail(X3D_ERR_ARG, "output buffer too small");
    std::memcpy(out, v->data(), v->size() * sizeof(double));
  }
  return X3D_OK;
}

int x3d_get_scalar(const char *name, double *out) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!name || !out) return fail(X3D_ERR_ARG, "null argument");
  const std::string s = name;
  if (s == "xnu") { *out = c->xnu; return X3D_OK; }
  for (int d = 0; d < 3; ++d)
    for (const auto &kv : c->dir[d].scalars)
      if (kv.first == s) {
        *out = kv.second;
        return X3D_OK;
      }
  return fail(X3D_ERR_ARG, "unknown scalar " + s);
}

// ---------------------------------------------------------------- derivatives
#define X3D_DEF_DERX(fname, T, dir, op, b1, bn)                                                   \
  int fname(T *t, const T *u, T *r, T *s, const T *ff, const T *fs, const T *fw, const int *nx, \
            const int *ny, const int *nz, const int *npaire, const T *lind) {                    \
    (void)r; (void)s; (void)lind;                                                                \
    return der_entry<T>(dir, op, b1, bn, t, u, ff, fs, fw, nx, ny, nz, npaire);                  \
  }
#define X3D_DEF_DERY(fname, T, b1, bn)                                                             \
  int fname(T *t, const T *u, T *r, T *s, const T *ff, const T *fs, const T *fw, const T *pp,    \
            const int *nx, const int *ny, const int *nz, const int *npaire, const T *lind) {     \
    (void)r; (void)s; (void)lind; /* pp: ppy, applied when istret != 0 */                        \
    return der_entry<T>(1, OP_DER1, b1, bn, t, u, ff, fs, fw, nx, ny, nz, npaire, pp);           \
  }
#define X3D_DEF_FAMILY(T, SFX)                                                  \
  X3D_DEF_DERX(x3d_derx_00##SFX, T, 0, OP_DER1, 0, 0)                           \
  X3D_DEF_DERX(x3d_derx_11##SFX, T, 0, OP_DER1, 1, 1)                           \
  X3D_DEF_DERX(x3d_derx_12##SFX, T, 0, OP_DER1, 1, 2)                           \
  X3D_DEF_DERX(x3d_derx_21##SFX, T, 0, OP_DER1, 2, 1)                           \
  X3D_DEF_DERX(x3d_derx_22##SFX, T, 0, OP_DER1, 2, 2)                           \
  X3D_DEF_DERY(x3d_dery_00##SFX, T, 0, 0)                                       \
  X3D_DEF_DERY(x3d_dery_11##SFX, T, 1, 1)                                       \
  X3D_DEF_DERY(x3d_dery_12##SFX, T, 1, 2)                                       \
  X3D_DEF_DERY(x3d_dery_21##SFX, T, 2, 1)                                       \
  X3D_DEF_DERY(x3d_dery_22##SFX, T, 2, 2)                                       \
  X3D_DEF_DERX(x3d_derz_00##SFX, T, 2, OP_DER1, 0, 0)                           \
  X3D_DEF_DERX(x3d_derz_11##SFX, T, 2, OP_DER1, 1, 1)                           \
  X3D_DEF_DERX(x3d_derz_12##SFX, T, 2, OP_DER1, 1, 2)                           \
  X3D_DEF_DERX(x3d_derz_21##SFX, T, 2, OP_DER1, 2, 1)                           \
  X3D_DEF_DERX(x3d_derz_22##SFX, T, 2, OP_DER1, 2, 2)                           \
  X3D_DEF_DERX(x3d_derxx_00##SFX, T, 0, OP_DER2, 0, 0)                          \
  X3D_DEF_DERX(x3d_derxx_11##SFX, T, 0, OP_DER2, 1, 1)                          \
  X3D_DEF_DERX(x3d_derxx_12##SFX, T, 0, OP_DER2, 1, 2)                          \
  X3D_DEF_DERX(x3d_derxx_21##SFX, T, 0, OP_DER2, 2, 1)                          \
  X3D_DEF_DERX(x3d_derxx_22##SFX, T, 0, OP_DER2, 2, 2)                          \
  X3D_DEF_DERX(x3d_deryy_00##SFX, T, 1, OP_DER2, 0, 0)                          \
  X3D_DEF_DERX(x3d_deryy_11##SFX, T, 1, OP_DER2, 1, 1)                          \
  X3D_DEF_DERX(x3d_deryy_12##SFX, T, 1, OP_DER2, 1, 2)                          \
  X3D_DEF_DERX(x3d_deryy_21##SFX, T, 1, OP_DER2, 2, 1)                          \
  X3D_DEF_DERX(x3d_deryy_22##SFX, T, 1, OP_DER2, 2, 2)                          \
  X3D_DEF_DERX(x3d_derzz_00##SFX, T, 2, OP_DER2, 0, 0)                          \
  X3D_DEF_DERX(x3d_derzz_11##SFX, T, 2, OP_DER2, 1, 1)                          \
  X3D_DEF_DERX(x3d_derzz_12##SFX, T, 2, OP_DER2, 1, 2)                          \
  X3D_DEF_DERX(x3d_derzz_21##SFX, T, 2, OP_DER2, 2, 1)                          \
  X3D_DEF_DERX(x3d_derzz_22##SFX, T, 2, OP_DER2, 2, 2)

X3D_DEF_FAMILY(double, )
X3D_DEF_FAMILY(float, _f32)

// ---------------------------------------------------------------- staggered
#define X3D_DEF_STAG(T, SFX)                                                                                     \
  int x3d_derxvp##SFX(T *t, const T *u, T *r, T *s, const T *cf, const T *cs, const T *cw, const int *nx,        \
                      const int *nxm, const int *ny, const int *nz, const int *npaire) {                         \
    (void)r; (void)s;                                                                                            \
    X3D_TRY(check_dims(nx, ny, nz)); X3D_TRY(check_dims(nxm, nxm, npaire ? nxm : nullptr));                      \
    return stag_entry<T>(0, OP_DERVP, t, u, cf, cs, cw, nullptr, nullptr, nullptr, *nx, *ny, *nz, *nxm, *npaire); \
  }                                                                                                              \
  int x3d_interxvp##SFX(T *t, const T *u, T *r, T *s, const T *cf, const T *cs, const T *cw, const int *nx,      \
                        const int *nxm, const int *ny, const int *nz, const int *npaire) {                       \
    (void)r; (void)s;                                                                                            \
    X3D_TRY(check_dims(nx, ny, nz)); X3D_TRY(check_dims(nxm, nxm, npaire ? nxm : nullptr));                      \
    return stag_entry<T>(0, OP_INTERVP, t, u, cf, cs, cw, nullptr, nullptr, nullptr, *nx, *ny, *nz, *nxm,        \
                         *npaire);                                                                               \
  }                                                                                                              \
  int x3d_derxpv##SFX(T *t, const T *u, T *r, T *s, const T *cf1, const T *cs1, const T *cw1, const T *cf2,      \
                      const T *cs2, const T *cw2, const int *nxm, const int *nx, const int *ny, const int *nz,   \
                      const int *npaire) {                                                                       \
    (void)r; (void)s;                                                                                            \
    X3D_TRY(check_dims(nxm, ny, nz)); X3D_TRY(check_dims(nx, nx, npaire ? nx : nullptr));                        \
    return stag_entry<T>(0, OP_DERPV, t, u, cf1, cs1, cw1, cf2, cs2, cw2, *nxm, *ny, *nz, *nx, *npaire);          \
  }                                                                                                              \
  int x3d_interxpv##SFX(T *t, const T *u, T *r, T *s, const T *cf1, const T *cs1, const T *cw1, const T *cf2,    \
                        const T *cs2, const T *cw2, const int *nxm, const int *nx, const int *ny, const int *nz, \
                        const int *npaire) {                                                                     \
    (void)r; (void)s;                                                                                            \
    X3D_TRY(check_dims(nxm, ny, nz)); X3D_TRY(check_dims(nx, nx, npaire ? nx : nullptr));                        \
    return stag_entry<T>(0, OP_INTERPV, t, u, cf1, cs1, cw1, cf2, cs2, cw2, *nxm, *ny, *nz, *nx, *npaire);       \
  }                                                                                                              \
  int x3d_interyvp##SFX(T *t, const T *u, T *r, T *s, const T *cf, const T *cs, const T *cw, const int *nx,      \
                        const int *ny, const int *nym, const int *nz, const int *npaire) {                       \
    (void)r; (void)s;                                                                                            \
    X3D_TRY(check_dims(nx, ny, nz)); X3D_TRY(check_dims(nym, nym, npaire ? nym : nullptr));                      \
    return stag_entry<T>(1, OP_INTERVP, t, u, cf, cs, cw, nullptr, nullptr, nullptr, *nx, *ny, *nz, *nym,        \
                         *npaire);                                                                               \
  }                                                                                                              \
  int x3d_deryvp##SFX(T *t, const T *u, T *r, T *s, const T *cf, const T *cs, const T *cw, const T *ppyi,        \
                      const int *nx, const int *ny, const int *nym, const int *nz, const int *npaire) {          \
    (void)r; (void)s;                                                                                            \
    X3D_TRY(check_dims(nx, ny, nz)); X3D_TRY(check_dims(nym, nym, npaire ? nym : nullptr));                      \
    return stag_entry<T>(1, OP_DERVP, t, u, cf, cs, cw, nullptr, nullptr, nullptr, *nx, *ny, *nz, *nym, *npaire, \
                         ppyi);                                                                                  \
  }                                                                                                              \
  int x3d_interypv##SFX(T *t, const T *u, T *r, T *s, const T *cf1, const T *cs1, const T *cw1, const T *cf2,    \
                        const T *cs2, const T *cw2, const int *nx, const int *nym, const int *ny, const int *nz, \
                        const int *npaire) {                                                                     \
    (void)r; (void)s;                                                                                            \
    X3D_TRY(check_dims(nx, nym, nz)); X3D_TRY(check_dims(ny, ny, npaire ? ny : nullptr));                        \
    return stag_entry<T>(1, OP_INTERPV, t, u, cf1, cs1, cw1, cf2, cs2, cw2, *nx, *nym, *nz, *ny, *npaire);       \
  }                                                                                                              \
  int x3d_derypv##SFX(T *t, const T *u, T *r, T *s, const T *cf1, const T *cs1, const T *cw1, const T *cf2,      \
                      const T *cs2, const T *cw2, const T *ppy, const int *nx, const int *nym, const int *ny,    \
                      const int *nz, const int *npaire) {                                                        \
    (void)r; (void)s;                                                                                            \
    X3D_TRY(check_dims(nx, nym, nz)); X3D_TRY(check_dims(ny, ny, npaire ? ny : nullptr));                        \
    return stag_entry<T>(1, OP_DERPV, t, u, cf1, cs1, cw1, cf2, cs2, cw2, *nx, *nym, *nz, *ny, *npaire, ppy);    \
  }                                                                                                              \
  int x3d_derzvp##SFX(T *t, const T *u, T *r, T *s, const T *cf, const T *cs, const T *cw, const int *nx,        \
                      const int *ny, const int *nz, const int *nzm, const int *npaire) {                         \
    (void)r; (void)s;                                                                                            \
    X3D_TRY(check_dims(nx, ny, nz)); X3D_TRY(check_dims(nzm, nzm, npaire ? nzm : nullptr));                      \
    return stag_entry<T>(2, OP_DERVP, t, u, cf, cs, cw, nullptr, nullptr, nullptr, *nx, *ny, *nz, *nzm, *npaire); \
  }                                                                                                              \
  int x3d_interzvp##SFX(T *t, const T *u, T *r, T *s, const T *cf, const T *cs, const T *cw, const int *nx,      \
                        const int *ny, const int *nz, const int *nzm, const int *npaire) {                       \
    (void)r; (void)s;                                                                                            \
    X3D_TRY(check_dims(nx, ny, nz)); X3D_TRY(check_dims(nzm, nzm, npaire ? nzm : nullptr));                      \
    return stag_entry<T>(2, OP_INTERVP, t, u, cf, cs, cw, nullptr, nullptr, nullptr, *nx, *ny, *nz, *nzm,        \
                         *npaire);                                                                               \
  }                                                                                                              \
  int x3d_derzpv##SFX(T *t, const T *u, T *r, T *s, const T *cf1, const T *cs1, const T *cw1, const T *cf2,      \
                      const T *cs2, const T *cw2, const int *nx, const int *ny, const int *nzm, const int *nz,   \
                      const int *npaire) {                                                                       \
    (void)r; (void)s;                                                                                            \
    X3D_TRY(check_dims(nx, ny, nzm)); X3D_TRY(check_dims(nz, nz, npaire ? nz : nullptr));                        \
    return stag_entry<T>(2, OP_DERPV, t, u, cf1, cs1, cw1, cf2, cs2, cw2, *nx, *ny, *nzm, *nz, *npaire);         \
  }                                                                                                              \
  int x3d_interzpv##SFX(T *t, const T *u, T *r, T *s, const T *cf1, const T *cs1, const T *cw1, const T *cf2,    \
                        const T *cs2, const T *cw2, const int *nx, const int *ny, const int *nzm, const int *nz, \
                        const int *npaire) {                                                                     \
    (void)r; (void)s;                                                                                            \
    X3D_TRY(check_dims(nx, ny, nzm)); X3D_TRY(check_dims(nz, nz, npaire ? nz : nullptr));                        \
    return stag_entry<T>(2, OP_INTERPV, t, u, cf1, cs1, cw1, cf2, cs2, cw2, *nx, *ny, *nzm, *nz, *npaire);       \
  }

X3D_DEF_STAG(double, )
X3D_DEF_STAG(float, _f32)

int x3d_line_op_async(int dir, int op, int npaire, int precision, const void *u, void *t, int n1, int n2,
                      int n3) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (dir < 0 || dir > 2 || op < 0 || op > 5 || (npaire != 0 && npaire != 1) || !u || !t || n1 <= 0 ||
      n2 <= 0 || n3 <= 0)
    return fail(X3D_ERR_ARG, "bad argument");
  // on a stretched mesh the y operators that the named entries multiply by
  // a metric do so here too: der1 by ppy, dervp by ppyi, derpv by ppy
  const bool metric = dir == 1 && c->p.istret != 0 && (op == OP_DER1 || op == OP_DERVP || op == OP_DERPV);
  if (precision == 0) {
    LineCall<double> L;
    X3D_TRY(make_line_call<double>(*c, dir, op, npaire, &L));
    L.nx = n1; L.ny = n2; L.nz = n3;
    L.u = static_cast<const double *>(u);
    L.t = static_cast<double *>(t);
    if (metric) X3D_TRY(y_metric<double>(*c, op == OP_DERVP, nullptr, L.n_out, &L.pp));
    return launch_line<double>(L, c->stream);
  }
  if (precision == 1) {
    LineCall<float> L;
    X3D_TRY(make_line_call<float>(*c, dir, op, npaire, &L));
    L.nx = n1; L.ny = n2; L.nz = n3;
    L.u = static_cast<const float *>(u);
    L.t = static_cast<float *>(t);
    if (metric) X3D_TRY(y_metric<float>(*c, op == OP_DERVP, nullptr, L.n_out, &L.pp));
    return launch_line<float>(L, c->stream);
  }
  return fail(X3D_ERR_ARG, "precision must be 0 or 1");
}

}  // extern "C"

// ---------------------------------------------------------------- Poisson
template <typename T>
static int poisson_enqueue(Context *c, T *rhs, int n1, int n2, int n3) {
  if (n1 <= 0 || n2 <= 0 || n3 <= 0) return fail(X3D_ERR_ARG, "non-positive size");
  if (!rhs) return fail(X3D_ERR_ARG, "null rhs");
  if (!c->poisson) return fail(X3D_ERR_UNSUPPORTED, "no Poisson solver for these boundary conditions/sizes");
  if (n1 != c->nxm || n2 != c->nym || n3 != c->nzm)
    return fail(X3D_ERR_ARG, "rhs extents differ from the pressure grid (nxm, nym, nzm)");
  return poisson_solve<T>(*c, rhs, c->stream);
}

template <typename T>
static int poisson_entry(T *rhs, const int *n1, const int *n2, const int *n3) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  X3D_TRY(check_dims(n1, n2, n3));
  X3D_TRY(poisson_enqueue<T>(c, rhs, *n1, *n2, *n3));
  X3D_HIP(hipStreamSynchronize(c->stream));
  return X3D_OK;
}

extern "C" {

int x3d_poisson(double *rhs, const int *n1, const int *n2, const int *n3) {
  return poisson_entry<double>(rhs, n1, n2, n3);
}
int x3d_poisson_f32(float *rhs, const int *n1, const int *n2, const int *n3) {
  return poisson_entry<float>(rhs, n1, n2, n3);
}
int x3d_poisson_async(int precision, void *rhs, int n1, int n2, int n3) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (precision == 0) return poisson_enqueue<double>(c, static_cast<double *>(rhs), n1, n2, n3);
  if (precision == 1) return poisson_enqueue<float>(c, static_cast<float *>(rhs), n1, n2, n3);
  return fail(X3D_ERR_ARG, "precision must be 0 or 1");
}

}  // extern "C"

// decomp_2d_fft_3d (2DECOMP&FFT v2.0.4, physical space in z-pencils)
template <typename T>
static int fft3d_entry(bool forward, const void *in, void *out, const int *n1, const int *n2, const int *n3) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  X3D_TRY(check_dims(n1, n2, n3));
  if (!in || !out) return fail(X3D_ERR_ARG, "null pointer");
  if (in == out) return fail(X3D_ERR_ARG, "the 3-D FFT is out of place");
  if (*n1 != c->nxm || *n2 != c->nym || *n3 != c->nzm)
    return fail(X3D_ERR_ARG, "real extents differ from the pressure grid (nxm, nym, nzm)");
  X3D_TRY(fft3d<T>(*c, forward, in, out, c->stream));
  X3D_HIP(hipStreamSynchronize(c->stream));
  return X3D_OK;
}

extern "C" {

int x3d_fft3d_r2c(const double *in, void *out, const int *n1, const int *n2, const int *n3) {
  return fft3d_entry<double>(true, in, out, n1, n2, n3);
}
int x3d_fft3d_c2r(const void *in, double *out, const int *n1, const int *n2, const int *n3) {
  return fft3d_entry<double>(false, in, out, n1, n2, n3);
}
int x3d_fft3d_r2c_f32(const float *in, void *out, const int *n1, const int *n2, const int *n3) {
  return fft3d_entry<float>(true, in, out, n1, n2, n3);
}
int x3d_fft3d_c2r_f32(const void *in, float *out, const int *n1, const int *n2, const int *n3) {
  return fft3d_entry<float>(false, in, out, n1, n2, n3);
}

int x3d_poisson_async_noalloc(int precision, void *rhs, int n1, int n2, int n3) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (precision != 0 && precision != 1) return fail(X3D_ERR_ARG, "precision must be 0 or 1");
  if (n1 <= 0 || n2 <= 0 || n3 <= 0) return fail(X3D_ERR_ARG, "non-positive size");
  if (!rhs) return fail(X3D_ERR_ARG, "null rhs");
  if (!c->poisson) return fail(X3D_ERR_UNSUPPORTED, "no Poisson solver for these boundary conditions/sizes");
  if (n1 != c->nxm || n2 != c->nym || n3 != c->nzm)
    return fail(X3D_ERR_ARG, "rhs extents differ from the pressure grid (nxm, nym, nzm)");
  const size_t need = precision == 0 ? poisson_work_bytes<double>(*c) : poisson_work_bytes<float>(*c);
  if (need > POISSON_SCRATCH_BYTES)
    return fail(X3D_ERR_UNSUPPORTED, "x3d_poisson_async_noalloc: the spectral work array needs " +
                                         std::to_string(need) + " bytes, beyond the " +
                                         std::to_string(POISSON_SCRATCH_BYTES) + "-byte scratch");
  void *ptr = poisson_scratch(c->device);
  if (!ptr) return fail(X3D_ERR_STATE, "no scratch address for this device (x3d_init not run on it)");
  if (precision == 0) return poisson_run<double>(*c, static_cast<double *>(rhs), ptr, c->stream);
  return poisson_run<float>(*c, static_cast<float *>(rhs), ptr, c->stream);
}

}  // extern "C"

// ---------------------------------------------------------------- transposes
template <typename T>
static int transpose_enqueue(Context *c, const T *src, T *dst, int n1, int n2, int n3) {
  if (n1 <= 0 || n2 <= 0 || n3 <= 0) return fail(X3D_ERR_ARG, "non-positive size");
  if (!src || !dst) return fail(X3D_ERR_ARG, "null pointer");
  const size_t n = (size_t)n1 * n2 * n3;
  if (src != dst) X3D_TRY(launch_copy<T>(src, dst, n, c->stream));
  return X3D_OK;
}

template <typename T>
static int transpose_entry(const T *src, T *dst, const int *n1, const int *n2, const int *n3) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  X3D_TRY(check_dims(n1, n2, n3));
  X3D_TRY(transpose_enqueue<T>(c, src, dst, *n1, *n2, *n3));
  X3D_HIP(hipStreamSynchronize(c->stream));
  return X3D_OK;
}

extern "C" {

#define X3D_DEF_TRANSPOSE(name, T)                                                            \
  int name(const T *s, T *d, const int *a, const int *b, const int *c) {                     \
    return transpose_entry<T>(s, d, a, b, c);                                                 \
  }
X3D_DEF_TRANSPOSE(x3d_transpose_x_to_y, double)
X3D_DEF_TRANSPOSE(x3d_transpose_y_to_z, double)
X3D_DEF_TRANSPOSE(x3d_transpose_z_to_y, double)
X3D_DEF_TRANSPOSE(x3d_transpose_y_to_x, double)
X3D_DEF_TRANSPOSE(x3d_transpose_x_to_y_f32, float)
X3D_DEF_TRANSPOSE(x3d_transpose_y_to_z_f32, float)
X3D_DEF_TRANSPOSE(x3d_transpose_z_to_y_f32, float)
X3D_DEF_TRANSPOSE(x3d_transpose_y_to_x_f32, float)

int x3d_transpose_async(int precision, const void *src, void *dst, int n1, int n2, int n3) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (precision == 0)
    return transpose_enqueue<double>(c, static_cast<const double *>(src), static_cast<double *>(dst), n1, n2, n3);
  if (precision == 1)
    return transpose_enqueue<float>(c, static_cast<const float *>(src), static_cast<float *>(dst), n1, n2, n3);
  return fail(X3D_ERR_ARG, "precision must be 0 or 1");
}
int x3d_reorder_ji(const double *src, double *dst, const int *n1, const int *n2, const int *n3) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  X3D_TRY(check_dims(n1, n2, n3));
  if (!src || !dst || src == dst) return fail(X3D_ERR_ARG, "bad pointers (reorder is out of place)");
  X3D_TRY(launch_reorder_ji<double>(src, dst, *n1, *n2, *n3, c->stream));
  X3D_HIP(hipStreamSynchronize(c->stream));
  return X3D_OK;
}

}  // extern "C"
