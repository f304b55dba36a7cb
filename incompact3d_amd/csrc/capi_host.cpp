// Host-array twins of the device entries: x3d_host_<name> takes the same
// arguments as x3d_<name>, with the field arrays in host memory, as the
// reference's allocatables are (variables.f90:127-187).  The input is copied
// to HBM, the device entry runs on it, and the output is copied back; the
// coefficient arrays are passed through unchanged (the device entries read
// them on the host).  For callers that keep their fields on the host, never
// on the measured path.  On one GPU a pencil transpose is an identity copy
// (variables.f90:121-123), so the host transposes copy host memory directly.
#include <cstring>

#include "x3d_internal.h"

using namespace x3d;

namespace {

constexpr size_t STAGE_ALIGN = 256;

size_t round_up(size_t b) { return (b + STAGE_ALIGN - 1) / STAGE_ALIGN * STAGE_ALIGN; }

int stage(Context *c, size_t bytes, void **p) {
  if (c->host_stage_bytes < bytes) {
    if (c->host_stage) X3D_HIP(hipFree(c->host_stage));
    c->host_stage = nullptr;
    c->host_stage_bytes = 0;
    X3D_HIP(hipMalloc(&c->host_stage, bytes));
    c->host_stage_bytes = bytes;
  }
  *p = c->host_stage;
  return X3D_OK;
}

int dims_ok(const int *a, const int *b, const int *c) {
  if (!a || !b || !c) return fail(X3D_ERR_ARG, "null size argument");
  if (*a <= 0 || *b <= 0 || *c <= 0) return fail(X3D_ERR_ARG, "non-positive size");
  return X3D_OK;
}

size_t words(const int *a, const int *b, const int *c) { return (size_t)*a * (size_t)*b * (size_t)*c; }

// stage `in_bytes` of u into HBM, run f(device in, device out), bring
// `out_bytes` back into t; in place (t == u) when `inplace`
template <typename F>
int host_call(const void *u, void *t, size_t in_bytes, size_t out_bytes, bool inplace, F &&f) {
  Context *c;
  X3D_TRY(require_ctx(&c));
  if (!u || !t) return fail(X3D_ERR_ARG, "null pointer");
  void *p;
  X3D_TRY(stage(c, inplace ? round_up(in_bytes) : round_up(in_bytes) + round_up(out_bytes), &p));
  char *din = static_cast<char *>(p);
  char *dout = inplace ? din : din + round_up(in_bytes);
  X3D_HIP(hipMemcpy(din, u, in_bytes, hipMemcpyHostToDevice));
  X3D_TRY(f(din, dout));
  X3D_HIP(hipMemcpy(t, dout, out_bytes, hipMemcpyDeviceToHost));
  return X3D_OK;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- derivatives
#define X3D_HOST_DERX(name, T)                                                                              \
  int x3d_host_##name(T *t, const T *u, T *r, T *s, const T *ff, const T *fs, const T *fw, const int *nx, \
                      const int *ny, const int *nz, const int *npaire, const T *lind) {                    \
    X3D_TRY(dims_ok(nx, ny, nz));                                                                          \
    const size_t b = words(nx, ny, nz) * sizeof(T);                                                        \
    return host_call(u, t, b, b, false, [&](char *du, char *dt) {                                          \
      return x3d_##name((T *)dt, (const T *)du, r, s, ff, fs, fw, nx, ny, nz, npaire, lind);               \
    });                                                                                                    \
  }
#define X3D_HOST_DERY(name, T)                                                                              \
  int x3d_host_##name(T *t, const T *u, T *r, T *s, const T *ff, const T *fs, const T *fw, const T *pp,   \
                      const int *nx, const int *ny, const int *nz, const int *npaire, const T *lind) {     \
    X3D_TRY(dims_ok(nx, ny, nz));                                                                          \
    const size_t b = words(nx, ny, nz) * sizeof(T);                                                        \
    return host_call(u, t, b, b, false, [&](char *du, char *dt) {                                          \
      return x3d_##name((T *)dt, (const T *)du, r, s, ff, fs, fw, pp, nx, ny, nz, npaire, lind);           \
    });                                                                                                    \
  }
#define X3D_HOST_DER_BC(op, T, SFX, MAC) \
  MAC(op##_00##SFX, T) MAC(op##_11##SFX, T) MAC(op##_12##SFX, T) MAC(op##_21##SFX, T) MAC(op##_22##SFX, T)
#define X3D_HOST_DER_FAMILY(T, SFX)               \
  X3D_HOST_DER_BC(derx, T, SFX, X3D_HOST_DERX)    \
  X3D_HOST_DER_BC(dery, T, SFX, X3D_HOST_DERY)    \
  X3D_HOST_DER_BC(derz, T, SFX, X3D_HOST_DERX)    \
  X3D_HOST_DER_BC(derxx, T, SFX, X3D_HOST_DERX)   \
  X3D_HOST_DER_BC(deryy, T, SFX, X3D_HOST_DERX)   \
  X3D_HOST_DER_BC(derzz, T, SFX, X3D_HOST_DERX)

X3D_HOST_DER_FAMILY(double, )
X3D_HOST_DER_FAMILY(float, _f32)

// ---------------------------------------------------------------- staggered operators
// velocity -> pressure grid: (tx,ux,rx,sx,cf,cs,cw,<extents>,npaire), input
// extents (a,b,c), output with the staggered direction at m
#define X3D_HOST_VP(name, T, A1, A2, A3, A4, IN1, IN2, IN3, OUT1, OUT2, OUT3)                               \
  int x3d_host_##name(T *t, const T *u, T *r, T *s, const T *cf, const T *cs, const T *cw, const int *A1,  \
                      const int *A2, const int *A3, const int *A4, const int *npaire) {                    \
    X3D_TRY(dims_ok(IN1, IN2, IN3));                                                                       \
    X3D_TRY(dims_ok(OUT1, OUT2, OUT3));                                                                    \
    return host_call(u, t, words(IN1, IN2, IN3) * sizeof(T), words(OUT1, OUT2, OUT3) * sizeof(T), false,  \
                     [&](char *du, char *dt) {                                                             \
                       return x3d_##name((T *)dt, (const T *)du, r, s, cf, cs, cw, A1, A2, A3, A4, npaire); \
                     });                                                                                   \
  }
#define X3D_HOST_PV(name, T, A1, A2, A3, A4, IN1, IN2, IN3, OUT1, OUT2, OUT3)                                 \
  int x3d_host_##name(T *t, const T *u, T *r, T *s, const T *cf1, const T *cs1, const T *cw1, const T *cf2, \
                      const T *cs2, const T *cw2, const int *A1, const int *A2, const int *A3, const int *A4, \
                      const int *npaire) {                                                                   \
    X3D_TRY(dims_ok(IN1, IN2, IN3));                                                                         \
    X3D_TRY(dims_ok(OUT1, OUT2, OUT3));                                                                      \
    return host_call(u, t, words(IN1, IN2, IN3) * sizeof(T), words(OUT1, OUT2, OUT3) * sizeof(T), false,    \
                     [&](char *du, char *dt) {                                                               \
                       return x3d_##name((T *)dt, (const T *)du, r, s, cf1, cs1, cw1, cf2, cs2, cw2, A1, A2, \
                                         A3, A4, npaire);                                                    \
                     });                                                                                     \
  }

#define X3D_HOST_STAG(T, SFX)                                                                                  \
  X3D_HOST_VP(derxvp##SFX, T, nx, nxm, ny, nz, nx, ny, nz, nxm, ny, nz)                                        \
  X3D_HOST_VP(interxvp##SFX, T, nx, nxm, ny, nz, nx, ny, nz, nxm, ny, nz)                                      \
  X3D_HOST_PV(derxpv##SFX, T, nxm, nx, ny, nz, nxm, ny, nz, nx, ny, nz)                                        \
  X3D_HOST_PV(interxpv##SFX, T, nxm, nx, ny, nz, nxm, ny, nz, nx, ny, nz)                                      \
  X3D_HOST_VP(interyvp##SFX, T, nx, ny, nym, nz, nx, ny, nz, nx, nym, nz)                                      \
  X3D_HOST_PV(interypv##SFX, T, nx, nym, ny, nz, nx, nym, nz, nx, ny, nz)                                      \
  X3D_HOST_VP(derzvp##SFX, T, nx, ny, nz, nzm, nx, ny, nz, nx, ny, nzm)                                        \
  X3D_HOST_VP(interzvp##SFX, T, nx, ny, nz, nzm, nx, ny, nz, nx, ny, nzm)                                      \
  X3D_HOST_PV(derzpv##SFX, T, nx, ny, nzm, nz, nx, ny, nzm, nx, ny, nz)                                        \
  X3D_HOST_PV(interzpv##SFX, T, nx, ny, nzm, nz, nx, ny, nzm, nx, ny, nz)                                      \
  int x3d_host_deryvp##SFX(T *t, const T *u, T *r, T *s, const T *cf, const T *cs, const T *cw, const T *ppyi, \
                           const int *nx, const int *ny, const int *nym, const int *nz, const int *npaire) {   \
    X3D_TRY(dims_ok(nx, ny, nz));                                                                              \
    X3D_TRY(dims_ok(nx, nym, nz));                                                                             \
    return host_call(u, t, words(nx, ny, nz) * sizeof(T), words(nx, nym, nz) * sizeof(T), false,              \
                     [&](char *du, char *dt) {                                                                 \
                       return x3d_deryvp##SFX((T *)dt, (const T *)du, r, s, cf, cs, cw, ppyi, nx, ny, nym, nz, \
                                              npaire);                                                         \
                     });                                                                                       \
  }                                                                                                            \
  int x3d_host_derypv##SFX(T *t, const T *u, T *r, T *s, const T *cf1, const T *cs1, const T *cw1,            \
                           const T *cf2, const T *cs2, const T *cw2, const T *ppy, const int *nx,             \
                           const int *nym, const int *ny, const int *nz, const int *npaire) {                 \
    X3D_TRY(dims_ok(nx, nym, nz));                                                                             \
    X3D_TRY(dims_ok(nx, ny, nz));                                                                              \
    return host_call(u, t, words(nx, nym, nz) * sizeof(T), words(nx, ny, nz) * sizeof(T), false,              \
                     [&](char *du, char *dt) {                                                                 \
                       return x3d_derypv##SFX((T *)dt, (const T *)du, r, s, cf1, cs1, cw1, cf2, cs2, cw2, ppy, \
                                              nx, nym, ny, nz, npaire);                                        \
                     });                                                                                       \
  }

X3D_HOST_STAG(double, )
X3D_HOST_STAG(float, _f32)

// ---------------------------------------------------------------- Poisson and 3-D FFT
int x3d_host_poisson(double *rhs, const int *n1, const int *n2, const int *n3) {
  X3D_TRY(dims_ok(n1, n2, n3));
  const size_t b = words(n1, n2, n3) * sizeof(double);
  return host_call(rhs, rhs, b, b, true, [&](char *d, char *) { return x3d_poisson((double *)d, n1, n2, n3); });
}
int x3d_host_poisson_f32(float *rhs, const int *n1, const int *n2, const int *n3) {
  X3D_TRY(dims_ok(n1, n2, n3));
  const size_t b = words(n1, n2, n3) * sizeof(float);
  return host_call(rhs, rhs, b, b, true, [&](char *d, char *) { return x3d_poisson_f32((float *)d, n1, n2, n3); });
}

#define X3D_HOST_FFT(SFX, T)                                                                                     \
  int x3d_host_fft3d_r2c##SFX(const T *in, void *out, const int *n1, const int *n2, const int *n3) {            \
    X3D_TRY(dims_ok(n1, n2, n3));                                                                                \
    const size_t bi = words(n1, n2, n3) * sizeof(T), bo = (size_t)*n1 * *n2 * (*n3 / 2 + 1) * 2 * sizeof(T);     \
    return host_call(in, out, bi, bo, false,                                                                     \
                     [&](char *d, char *o) { return x3d_fft3d_r2c##SFX((const T *)d, o, n1, n2, n3); });        \
  }                                                                                                              \
  int x3d_host_fft3d_c2r##SFX(const void *in, T *out, const int *n1, const int *n2, const int *n3) {            \
    X3D_TRY(dims_ok(n1, n2, n3));                                                                                \
    const size_t bo = words(n1, n2, n3) * sizeof(T), bi = (size_t)*n1 * *n2 * (*n3 / 2 + 1) * 2 * sizeof(T);     \
    return host_call(in, out, bi, bo, false,                                                                     \
                     [&](char *d, char *o) { return x3d_fft3d_c2r##SFX(d, (T *)o, n1, n2, n3); });              \
  }
X3D_HOST_FFT(, double)
X3D_HOST_FFT(_f32, float)

// ---------------------------------------------------------------- transposes
#define X3D_HOST_TRANSPOSE(name, T)                                                                \
  int x3d_host_##name(const T *src, T *dst, const int *n1, const int *n2, const int *n3) {       \
    Context *c;                                                                                   \
    X3D_TRY(require_ctx(&c));                                                                     \
    X3D_TRY(dims_ok(n1, n2, n3));                                                                 \
    if (!src || !dst) return fail(X3D_ERR_ARG, "null pointer");                                   \
    if (src != dst) std::memmove(dst, src, words(n1, n2, n3) * sizeof(T));                        \
    return X3D_OK;                                                                                \
  }
X3D_HOST_TRANSPOSE(transpose_x_to_y, double)
X3D_HOST_TRANSPOSE(transpose_y_to_z, double)
X3D_HOST_TRANSPOSE(transpose_z_to_y, double)
X3D_HOST_TRANSPOSE(transpose_y_to_x, double)
X3D_HOST_TRANSPOSE(transpose_x_to_y_f32, float)
X3D_HOST_TRANSPOSE(transpose_y_to_z_f32, float)
X3D_HOST_TRANSPOSE(transpose_z_to_y_f32, float)
X3D_HOST_TRANSPOSE(transpose_y_to_x_f32, float)

}  // extern "C"
