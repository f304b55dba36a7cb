// C-ABI of the closure network and the box filter (include/x3d.h, "closure
// network" and "x3d_box_filter").
#include "x3d_internal.h"

using namespace x3d;

extern "C" {

int x3d_mlp_nparams(void) { return MLP_PARAMS; }

int x3d_mlp_create(x3d_mlp **out) {
  if (!out) return fail(X3D_ERR_ARG, "null output");
  MlpDev *M = nullptr;
  X3D_TRY(mlp_alloc(M));
  *out = reinterpret_cast<x3d_mlp *>(M);
  return X3D_OK;
}

int x3d_mlp_destroy(x3d_mlp *m) {
  if (!m) return X3D_OK;
  (void)hipDeviceSynchronize();
  mlp_free(reinterpret_cast<MlpDev *>(m));
  return X3D_OK;
}

static hipStream_t lib_stream() {
  Context *c = ctx();
  return c ? c->stream : nullptr;
}

int x3d_mlp_set_params(x3d_mlp *m, const float *host) {
  if (!m || !host) return fail(X3D_ERR_ARG, "null pointer");
  return mlp_set_params(reinterpret_cast<MlpDev *>(m), host, lib_stream());
}

int x3d_mlp_get_params(x3d_mlp *m, float *host) {
  if (!m || !host) return fail(X3D_ERR_ARG, "null pointer");
  return mlp_get_params(reinterpret_cast<MlpDev *>(m), host, lib_stream());
}

int x3d_mlp_forward(x3d_mlp *m, const float *x, float *y, int n) {
  if (!m || !x || !y) return fail(X3D_ERR_ARG, "null pointer");
  hipStream_t st = lib_stream();
  X3D_TRY(mlp_forward(reinterpret_cast<MlpDev *>(m), x, y, n, st));
  X3D_HIP(hipStreamSynchronize(st));
  return X3D_OK;
}

int x3d_mlp_train_step(x3d_mlp *m, const float *x, const float *t, int n, float lr, float *loss) {
  if (!m || !x || !t) return fail(X3D_ERR_ARG, "null pointer");
  if (!(lr > 0.0f)) return fail(X3D_ERR_ARG, "learning rate must be positive");
  return mlp_train_batch(reinterpret_cast<MlpDev *>(m), x, t, n, 1, lr, loss, lib_stream());
}

int x3d_mlp_gradient(x3d_mlp *m, const float *x, const float *t, int n, float *host_grad, float *loss) {
  if (!m || !x || !t || !host_grad) return fail(X3D_ERR_ARG, "null pointer");
  hipStream_t st = lib_stream();
  X3D_TRY(mlp_train_batch(reinterpret_cast<MlpDev *>(m), x, t, n, 0, 0.0f, loss, st));
  return mlp_get_grad(reinterpret_cast<MlpDev *>(m), host_grad, st);
}

int x3d_mlp_closure(x3d_mlp *m, int precision, const void *const *g, void *const *t, long long n, double gscale,
                    double delta, double delta0, double cs, double beta, int project, long long *nfallback) {
  if (!m || !g || !t || n <= 0) return fail(X3D_ERR_ARG, "bad argument");
  if (project != 0 && project != 1) return fail(X3D_ERR_ARG, "project must be 0 or 1");
  for (int i = 0; i < 9; ++i)
    if (!g[i]) return fail(X3D_ERR_ARG, "null gradient field");
  for (int i = 0; i < 6; ++i)
    if (!t[i]) return fail(X3D_ERR_ARG, "null stress field");
  if (!(gscale > 0) || !(delta > 0) || !(delta0 > 0)) return fail(X3D_ERR_ARG, "scales must be positive");
  hipStream_t st = lib_stream();
  int *cnt = nullptr;
  if (nfallback) {
    X3D_HIP(hipMalloc(&cnt, sizeof(int)));
    X3D_HIP(hipMemsetAsync(cnt, 0, sizeof(int), st));
  }
  int r;
  if (precision == 0) {
    const double *gg[9];
    double *tt[6];
    for (int i = 0; i < 9; ++i) gg[i] = static_cast<const double *>(g[i]);
    for (int i = 0; i < 6; ++i) tt[i] = static_cast<double *>(t[i]);
    r = launch_closure<double>(reinterpret_cast<MlpDev *>(m), gg, tt, (size_t)n, gscale, delta / delta0, delta, cs,
                               beta, project, cnt, st);
  } else if (precision == 1) {
    const float *gg[9];
    float *tt[6];
    for (int i = 0; i < 9; ++i) gg[i] = static_cast<const float *>(g[i]);
    for (int i = 0; i < 6; ++i) tt[i] = static_cast<float *>(t[i]);
    r = launch_closure<float>(reinterpret_cast<MlpDev *>(m), gg, tt, (size_t)n, gscale, delta / delta0, delta, cs,
                              beta, project, cnt, st);
  } else {
    r = fail(X3D_ERR_ARG, "precision must be 0 or 1");
  }
  if (r == X3D_OK) {
    int h = 0;
    if (cnt) {
      X3D_HIP(hipMemcpyAsync(&h, cnt, sizeof(int), hipMemcpyDeviceToHost, st));
    }
    X3D_HIP(hipStreamSynchronize(st));
    if (nfallback) *nfallback = h;
  }
  if (cnt) (void)hipFree(cnt);
  return r;
}

int x3d_box_filter(const double *ux, const double *uy, const double *uz, int nx, int ny, int nz, int f,
                   double *const *ub, double *const *tau) {
  if (!ux || !uy || !uz || !ub || !tau || nx <= 0 || ny <= 0 || nz <= 0) return fail(X3D_ERR_ARG, "bad argument");
  for (int i = 0; i < 3; ++i)
    if (!ub[i]) return fail(X3D_ERR_ARG, "null output");
  for (int i = 0; i < 6; ++i)
    if (!tau[i]) return fail(X3D_ERR_ARG, "null output");
  const double *u[3] = {ux, uy, uz};
  hipStream_t st = lib_stream();
  X3D_TRY(launch_box_filter(u, nx, ny, nz, f, ub, tau, st));
  X3D_HIP(hipStreamSynchronize(st));
  return X3D_OK;
}

}  // extern "C"
