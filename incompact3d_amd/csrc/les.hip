// Elementwise kernels of the LES subgrid-stress path (les_models.f90) and the
// box filter that turns a fine velocity field into training data for the
// closure network.  The derivatives themselves are line-solve launches.
#include "x3d_internal.h"

namespace x3d {

constexpr int LES_THREADS = 256;

static inline int ew_blocks(size_t n) {
  const size_t b = (n + LES_THREADS - 1) / LES_THREADS;
  return (int)(b < 65536 ? b : 65536);
}

// smag (les_models.f90:214-300): S_ij from the nine gradients
// (gxx gyx gzx gxy gyy gzy gxz gyz gzz), srt = S:S with off-diagonals twice,
// nut = (Cs del)^2 sqrt(2 srt)
template <typename T>
__global__ void __launch_bounds__(LES_THREADS) strain_nut_kernel(const T *__restrict__ g0, const T *__restrict__ g1,
                                                                 const T *__restrict__ g2, const T *__restrict__ g3,
                                                                 const T *__restrict__ g4, const T *__restrict__ g5,
                                                                 const T *__restrict__ g6, const T *__restrict__ g7,
                                                                 const T *__restrict__ g8, T *__restrict__ sxx,
                                                                 T *__restrict__ syy, T *__restrict__ szz,
                                                                 T *__restrict__ sxy, T *__restrict__ sxz,
                                                                 T *__restrict__ syz, T *__restrict__ nut,
                                                                 double cs_d2, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const T a = g0[i], e = g4[i], z = g8[i];
    const T xy = T(0.5) * (g3[i] + g1[i]);
    const T xz = T(0.5) * (g2[i] + g6[i]);
    const T yz = T(0.5) * (g7[i] + g5[i]);
    sxx[i] = a;
    syy[i] = e;
    szz[i] = z;
    sxy[i] = xy;
    sxz[i] = xz;
    syz[i] = yz;
    const T srt = a * a + e * e + z * z + T(2) * xy * xy + T(2) * xz * xz + T(2) * yz * yz;
    nut[i] = (T)cs_d2 * sqrt(T(2) * srt);
  }
}

template <typename T>
int launch_strain_nut(const T *const g[9], T *const s[6], T *nut, double cs_d2, size_t n, hipStream_t st) {
  hipLaunchKernelGGL((strain_nut_kernel<T>), dim3(ew_blocks(n)), dim3(LES_THREADS), 0, st, g[0], g[1], g[2], g[3],
                     g[4], g[5], g[6], g[7], g[8], s[0], s[1], s[2], s[3], s[4], s[5], nut, cs_d2, n);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}
template int launch_strain_nut<double>(const double *const[9], double *const[6], double *, double, size_t,
                                       hipStream_t);
template int launch_strain_nut<float>(const float *const[9], float *const[6], float *, double, size_t, hipStream_t);

// acc += a * (x * y), as "sgs = sgs + two * s * ta" of les_models.f90:1129,1192
template <typename T>
__global__ void __launch_bounds__(LES_THREADS) fma3_kernel(T *__restrict__ acc, const T *__restrict__ x,
                                                           const T *__restrict__ y, T a, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) acc[i] = acc[i] + a * x[i] * y[i];
}

template <typename T>
int launch_fma3(T *acc, const T *x, const T *y, double a, size_t n, hipStream_t st) {
  hipLaunchKernelGGL((fma3_kernel<T>), dim3(ew_blocks(n)), dim3(LES_THREADS), 0, st, acc, x, y, (T)a, n);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}
template int launch_fma3<double>(double *, const double *, const double *, double, size_t, hipStream_t);
template int launch_fma3<float>(float *, const float *, const float *, double, size_t, hipStream_t);

// y += x (the SGS force added to the RHS, transeq.f90:500-502)
template <typename T>
__global__ void __launch_bounds__(LES_THREADS) add_kernel(T *__restrict__ y, const T *__restrict__ x, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) y[i] = y[i] + x[i];
}

template <typename T>
int launch_add(T *y, const T *x, size_t n, hipStream_t st) {
  hipLaunchKernelGGL((add_kernel<T>), dim3(ew_blocks(n)), dim3(LES_THREADS), 0, st, y, x, n);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}
template int launch_add<double>(double *, const double *, size_t, hipStream_t);
template int launch_add<float>(float *, const float *, size_t, hipStream_t);

// T_ij = 2 nut S_ij (sgs_mom_conservative, les_models.f90:1358-1363)
template <typename T>
__global__ void __launch_bounds__(LES_THREADS) smag_stress_kernel(const T *__restrict__ nut, const T *s0,
                                                                  const T *s1, const T *s2, const T *s3, const T *s4,
                                                                  const T *s5, T *t0, T *t1, T *t2,
                                                                  T *t3, T *t4, T *t5, size_t n) {
  // t may be s (in place): each point reads its s_ij before writing t_ij
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const T v = T(2) * nut[i];
    t0[i] = v * s0[i];
    t1[i] = v * s1[i];
    t2[i] = v * s2[i];
    t3[i] = v * s3[i];
    t4[i] = v * s4[i];
    t5[i] = v * s5[i];
  }
}

template <typename T>
int launch_smag_stress(const T *nut, const T *const s[6], T *const t[6], size_t n, hipStream_t st) {
  hipLaunchKernelGGL((smag_stress_kernel<T>), dim3(ew_blocks(n)), dim3(LES_THREADS), 0, st, nut, s[0], s[1], s[2],
                     s[3], s[4], s[5], t[0], t[1], t[2], t[3], t[4], t[5], n);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}
template int launch_smag_stress<double>(const double *, const double *const[6], double *const[6], size_t,
                                        hipStream_t);
template int launch_smag_stress<float>(const float *, const float *const[6], float *const[6], size_t, hipStream_t);

// ---------------------------------------------------------------- box filter
// One thread per coarse point (I,J,K): averages over the fine points
// f*I .. f*I+f-1 (likewise J, K) of u_i and u_i u_j, then
// tau_ij = bar(u_i u_j) - bar(u_i) bar(u_j).  Order: xx yy zz xy xz yz.
__global__ void __launch_bounds__(LES_THREADS) box_filter_kernel(const double *__restrict__ ux,
                                                                 const double *__restrict__ uy,
                                                                 const double *__restrict__ uz, int nx, int ny,
                                                                 int nz, int f, double *__restrict__ bx,
                                                                 double *__restrict__ by, double *__restrict__ bz,
                                                                 double *__restrict__ t0, double *__restrict__ t1,
                                                                 double *__restrict__ t2, double *__restrict__ t3,
                                                                 double *__restrict__ t4, double *__restrict__ t5) {
  const int cx = nx / f, cy = ny / f, cz = nz / f;
  const size_t nc = (size_t)cx * cy * cz;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += stride) {
    const int I = (int)(c % cx), J = (int)((c / cx) % cy), K = (int)(c / ((size_t)cx * cy));
    double su[3] = {0, 0, 0}, sp[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < f; ++k)
      for (int j = 0; j < f; ++j)
        for (int i = 0; i < f; ++i) {
          const size_t o = ((size_t)(K * f + k) * ny + (size_t)(J * f + j)) * nx + (size_t)(I * f + i);
          const double a = ux[o], b = uy[o], d = uz[o];
          su[0] += a; su[1] += b; su[2] += d;
          sp[0] += a * a; sp[1] += b * b; sp[2] += d * d;
          sp[3] += a * b; sp[4] += a * d; sp[5] += b * d;
        }
    const double w = 1.0 / ((double)f * f * f);
    const double ax = su[0] * w, ay = su[1] * w, az = su[2] * w;
    bx[c] = ax;
    by[c] = ay;
    bz[c] = az;
    t0[c] = sp[0] * w - ax * ax;
    t1[c] = sp[1] * w - ay * ay;
    t2[c] = sp[2] * w - az * az;
    t3[c] = sp[3] * w - ax * ay;
    t4[c] = sp[4] * w - ax * az;
    t5[c] = sp[5] * w - ay * az;
  }
}

int launch_box_filter(const double *const u[3], int nx, int ny, int nz, int f, double *const ub[3],
                      double *const tau[6], hipStream_t st) {
  if (f <= 0 || nx % f || ny % f || nz % f) return fail(X3D_ERR_ARG, "box filter: sizes must be multiples of f");
  const size_t nc = (size_t)(nx / f) * (ny / f) * (nz / f);
  hipLaunchKernelGGL(box_filter_kernel, dim3(ew_blocks(nc)), dim3(LES_THREADS), 0, st, u[0], u[1], u[2], nx, ny, nz,
                     f, ub[0], ub[1], ub[2], tau[0], tau[1], tau[2], tau[3], tau[4], tau[5]);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}

}  // namespace x3d
