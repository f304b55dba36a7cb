// Streaming field kernels: pencil copies, AB3 update, precision casts, the
// Taylor-Green initial condition and deterministic FP64 reductions.
#include <cmath>

#include "state64.h"
#include "x3d_internal.h"

namespace x3d {

constexpr int EW_THREADS = 256;

static inline unsigned ew_blocks(size_t n, int per_thread = 4) {
  size_t b = (n + (size_t)EW_THREADS * per_thread - 1) / ((size_t)EW_THREADS * per_thread);
  if (b < 1) b = 1;
  if (b > 65535u * 8u) b = 65535u * 8u;
  return (unsigned)b;
}

// ---------------------------------------------------------------- copy
template <typename T>
__global__ void __launch_bounds__(EW_THREADS) copy_kernel(const T *__restrict__ src, T *__restrict__ dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// 16-byte chunks, four per thread with every load issued before the first
// store, one pass over the data (no grid-stride loop).  Loads and stores are
// non-temporal: the copied field is not read again by this kernel, and
// streaming it past the caches took a 512^3 copy from 70 to 75 % of HBM peak
// (FP64) and from 68 to 77 % (FP32).
constexpr int COPY_CHUNKS = 4;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(EW_THREADS) copy16_kernel(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst,
                                                            size_t n16) {
  const size_t base = (size_t)blockIdx.x * (EW_THREADS * COPY_CHUNKS) + threadIdx.x;
  u32x4 v[COPY_CHUNKS];
#pragma unroll
  for (int k = 0; k < COPY_CHUNKS; ++k) {
    const size_t i = base + (size_t)k * EW_THREADS;
    if (i < n16) v[k] = __builtin_nontemporal_load(src + i);
  }
#pragma unroll
  for (int k = 0; k < COPY_CHUNKS; ++k) {
    const size_t i = base + (size_t)k * EW_THREADS;
    if (i < n16) __builtin_nontemporal_store(v[k], dst + i);
  }
}

template <typename T>
int launch_copy(const T *src, T *dst, size_t n, hipStream_t st) {
  if (n == 0) return X3D_OK;
  const size_t bytes = n * sizeof(T);
  if ((bytes % 16) == 0 && (reinterpret_cast<uintptr_t>(src) % 16) == 0 &&
      (reinterpret_cast<uintptr_t>(dst) % 16) == 0) {
    const size_t n16 = bytes / 16;
    const size_t per_block = (size_t)EW_THREADS * COPY_CHUNKS;
    const size_t blocks = (n16 + per_block - 1) / per_block;
    if (blocks > 0x7fffffffu) return fail(X3D_ERR_UNSUPPORTED, "copy: field too large");
    hipLaunchKernelGGL(copy16_kernel, dim3((unsigned)blocks), dim3(EW_THREADS), 0, st,
                       reinterpret_cast<const u32x4 *>(src), reinterpret_cast<u32x4 *>(dst), n16);
  } else {
    hipLaunchKernelGGL((copy_kernel<T>), dim3(ew_blocks(n)), dim3(EW_THREADS), 0, st, src, dst, n);
  }
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}
template int launch_copy<double>(const double *, double *, size_t, hipStream_t);
template int launch_copy<float>(const float *, float *, size_t, hipStream_t);

// ---------------------------------------------------------------- LDS-tiled (i,j) -> (j,i) reorder
// dst(j,i,k) = src(i,j,k): the layout change of a y-pencil stored y-fastest.
template <typename T>
__global__ void __launch_bounds__(256) reorder_ji_kernel(const T *__restrict__ src, T *__restrict__ dst, int n1,
                                                         int n2) {
  __shared__ T tile[32][33];
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const long long k = blockIdx.z;
  const T *s = src + k * (long long)n1 * n2;
  T *d = dst + k * (long long)n1 * n2;
  for (int r = threadIdx.y; r < 32; r += 8) {
    const int i = i0 + threadIdx.x, j = j0 + r;
    if (i < n1 && j < n2) tile[r][threadIdx.x] = s[(long long)j * n1 + i];
  }
  __syncthreads();
  for (int r = threadIdx.y; r < 32; r += 8) {
    const int j = j0 + threadIdx.x, i = i0 + r;
    if (i < n1 && j < n2) d[(long long)i * n2 + j] = tile[threadIdx.x][r];
  }
}

template <typename T>
int launch_reorder_ji(const T *src, T *dst, int n1, int n2, int n3, hipStream_t st) {
  dim3 g((n1 + 31) / 32, (n2 + 31) / 32, n3);
  hipLaunchKernelGGL((reorder_ji_kernel<T>), g, dim3(32, 8), 0, st, src, dst, n1, n2);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}
template int launch_reorder_ji<double>(const double *, double *, int, int, int, hipStream_t);

// ---------------------------------------------------------------- AB3 (time_integrators.f90:85-100)
// u <- a*f1 + b*f2 + c*f3 + u, in FP64; the history is read at its stored
// precision H; optionally writes an FP32 shadow of the new velocity.
template <typename H>
__global__ void __launch_bounds__(EW_THREADS) ab_kernel(double *__restrict__ u, const H *__restrict__ f1,
                                                        const H *__restrict__ f2, const H *__restrict__ f3,
                                                        double a, double b, double c, int nterms,
                                                        float *__restrict__ shadow, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double r;
    if (nterms == 1) r = a * (double)f1[i] + u[i];
    else if (nterms == 2) r = a * (double)f1[i] + b * (double)f2[i] + u[i];
    else r = a * (double)f1[i] + b * (double)f2[i] + c * (double)f3[i] + u[i];
    u[i] = r;
    if (shadow) shadow[i] = (float)r;
  }
}

template <typename H>
int launch_ab(double *u, const H *f1, const H *f2, const H *f3, double a, double b, double c, int nterms,
              float *shadow, size_t n, hipStream_t st) {
  hipLaunchKernelGGL((ab_kernel<H>), dim3(ew_blocks(n)), dim3(EW_THREADS), 0, st, u, f1, f2, f3, a, b, c, nterms,
                     shadow, n);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}
template int launch_ab<double>(double *, const double *, const double *, const double *, double, double, double, int,
                               float *, size_t, hipStream_t);
template int launch_ab<float>(double *, const float *, const float *, const float *, double, double, double, int,
                              float *, size_t, hipStream_t);

// u64 <- u64 - p (p at precision P) and shadow <- (float)u64: cor_vel
// (navier.f90:243-245) for the FP32-bulk path, which keeps the velocity in FP64.
template <typename P>
__global__ void __launch_bounds__(EW_THREADS) corvel_kernel(double *__restrict__ u, const P *__restrict__ p,
                                                            float *__restrict__ shadow, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double r = u[i] - (double)p[i];
    u[i] = r;
    if (shadow) shadow[i] = (float)r;
  }
}
template <typename P>
int launch_corvel(double *u, const P *p, float *shadow, size_t n, hipStream_t st) {
  hipLaunchKernelGGL((corvel_kernel<P>), dim3(ew_blocks(n)), dim3(EW_THREADS), 0, st, u, p, shadow, n);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}
template int launch_corvel<float>(double *, const float *, float *, size_t, hipStream_t);
template int launch_corvel<double>(double *, const double *, float *, size_t, hipStream_t);

__global__ void __launch_bounds__(EW_THREADS) cast_kernel(const double *__restrict__ a, float *__restrict__ b,
                                                          size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = (float)a[i];
}
int launch_cast(const double *a, float *b, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(cast_kernel, dim3(ew_blocks(n)), dim3(EW_THREADS), 0, st, a, b, n);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}
// FP64 field <-> (FP32 rounding, integer residual) pair, exact both ways
// for finite values of magnitude at least 2^-127 (state64.h)
__global__ void __launch_bounds__(EW_THREADS) split_kernel(const double *__restrict__ a, float *__restrict__ hi,
                                                           int32_t *__restrict__ r, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double u = a[i];
    const float h = (float)u;
    hi[i] = h;
    r[i] = split_residual(u, h);
  }
}
int launch_split(const double *a, float *hi, int32_t *r, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(split_kernel, dim3(ew_blocks(n)), dim3(EW_THREADS), 0, st, a, hi, r, n);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}
__global__ void __launch_bounds__(EW_THREADS) join_kernel(const float *__restrict__ hi, const int32_t *__restrict__ r,
                                                          double *__restrict__ a, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) a[i] = join_split(hi[i], r[i]);
}
int launch_join(const float *hi, const int32_t *r, double *a, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(join_kernel, dim3(ew_blocks(n)), dim3(EW_THREADS), 0, st, hi, r, a, n);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}
__global__ void __launch_bounds__(EW_THREADS) widen_kernel(const float *__restrict__ a, double *__restrict__ b,
                                                           size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = (double)a[i];
}
int launch_widen(const float *a, double *b, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(widen_kernel, dim3(ew_blocks(n)), dim3(EW_THREADS), 0, st, a, b, n);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}

// ---------------------------------------------------------------- TGV initial condition (Case-TGV.f90:90-95)
__global__ void __launch_bounds__(EW_THREADS) tgv_kernel(double *__restrict__ ux, double *__restrict__ uy,
                                                         double *__restrict__ uz, int nx, int ny, int nz, double dx,
                                                         double dy, double dz) {
  const size_t n = (size_t)nx * ny * nz;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += stride) {
    const int i = (int)(idx % nx);
    const int j = (int)((idx / nx) % ny);
    const int k = (int)(idx / ((size_t)nx * ny));
    const double x = (double)i * dx, y = (double)j * dy, z = (double)k * dz;
    ux[idx] = sin(x) * cos(y) * cos(z);
    uy[idx] = -cos(x) * sin(y) * cos(z);
    uz[idx] = 0.0;
  }
}
int launch_tgv(double *ux, double *uy, double *uz, int nx, int ny, int nz, double dx, double dy, double dz,
               hipStream_t st) {
  const size_t n = (size_t)nx * ny * nz;
  hipLaunchKernelGGL(tgv_kernel, dim3(ew_blocks(n)), dim3(EW_THREADS), 0, st, ux, uy, uz, nx, ny, nz, dx, dy, dz);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}

// ---------------------------------------------------------------- FP64 reductions
// Deterministic two-stage sums: each block reduces a fixed slice in a fixed
// order; the partials are summed by one block.
constexpr int RED_BLOCKS = 1024;

template <typename F>
__device__ void block_sum_store(double v, double *out) {
  __shared__ double red[EW_THREADS];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = EW_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

struct Dummy {};

// Per point: 0 eek, 1 eps, 2 enst, 3 u.lap(u)   (Case-TGV.f90:287-370)
// g[9] = (dux/dx, duy/dx, duz/dx, dux/dy, duy/dy, duz/dy, dux/dz, duy/dz, duz/dz)
template <typename T>
__global__ void __launch_bounds__(EW_THREADS) tgv_partials(const T *__restrict__ ux, const T *__restrict__ uy,
                                                           const T *__restrict__ uz, const T *g0, const T *g1,
                                                           const T *g2, const T *g3, const T *g4, const T *g5,
                                                           const T *g6, const T *g7, const T *g8, const T *lx,
                                                           const T *ly, const T *lz, double xnu, size_t n,
                                                           int nx, int ny, int nxc, int nyc, int nzc,
                                                           double *__restrict__ partial) {
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  const size_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const size_t beg = (size_t)blockIdx.x * chunk;
  const size_t end = beg + chunk < n ? beg + chunk : n;
  const bool trim = nxc < nx || nyc < ny || (size_t)nzc * nx * ny < n;
  for (size_t i = beg + threadIdx.x; i < end; i += blockDim.x) {
    if (trim) {
      // free-slip directions drop their last plane (Case-TGV.f90:213-242)
      const size_t ij = i % ((size_t)nx * ny);
      if ((int)(ij % nx) >= nxc || (int)(ij / nx) >= nyc || (int)(i / ((size_t)nx * ny)) >= nzc) continue;
    }
    const double a = ux[i], b = uy[i], c = uz[i];
    s0 += 0.5 * (a * a + b * b + c * c);
    const double ta1 = g0[i], tb1 = g1[i], tc1 = g2[i], td1 = g3[i], te1 = g4[i], tf1 = g5[i], tg1 = g6[i],
                 th1 = g7[i], ti1 = g8[i];
    s1 += 0.5 * xnu *
          ((2.0 * ta1) * (2.0 * ta1) + (2.0 * te1) * (2.0 * te1) + (2.0 * ti1) * (2.0 * ti1) +
           2.0 * (td1 + tb1) * (td1 + tb1) + 2.0 * (tg1 + tc1) * (tg1 + tc1) + 2.0 * (th1 + tf1) * (th1 + tf1));
    s2 += 0.5 * ((tf1 - th1) * (tf1 - th1) + (tg1 - tc1) * (tg1 - tc1) + (tb1 - td1) * (tb1 - td1));
    if (lx) s3 += (-xnu) * (a * (double)lx[i] + b * (double)ly[i] + c * (double)lz[i]);
  }
  __shared__ double red[4][EW_THREADS];
  red[0][threadIdx.x] = s0;
  red[1][threadIdx.x] = s1;
  red[2][threadIdx.x] = s2;
  red[3][threadIdx.x] = s3;
  __syncthreads();
  for (int s = EW_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 4) partial[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = red[threadIdx.x][0];
}

__global__ void __launch_bounds__(EW_THREADS) sum_partials(const double *__restrict__ partial, int nb, int nq,
                                                           double *__restrict__ out) {
  __shared__ double red[EW_THREADS];
  for (int q = 0; q < nq; ++q) {
    double s = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partial[(size_t)q * nb + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = EW_THREADS / 2; k > 0; k >>= 1) {
      if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = red[0];
    __syncthreads();
  }
}

template <typename T>
int launch_tgv_sums(const T *ux, const T *uy, const T *uz, const T *const g[9], const T *lx, const T *ly,
                    const T *lz, double xnu, int nx, int ny, int nz, int nxc, int nyc, int nzc, double *partial,
                    double *out4, hipStream_t st) {
  const size_t n = (size_t)nx * ny * nz;
  hipLaunchKernelGGL((tgv_partials<T>), dim3(RED_BLOCKS), dim3(EW_THREADS), 0, st, ux, uy, uz, g[0], g[1], g[2], g[3],
                     g[4], g[5], g[6], g[7], g[8], lx, ly, lz, xnu, n, nx, ny, nxc, nyc, nzc, partial);
  hipLaunchKernelGGL(sum_partials, dim3(1), dim3(EW_THREADS), 0, st, partial, RED_BLOCKS, 4, out4);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}
template int launch_tgv_sums<double>(const double *, const double *, const double *, const double *const[9],
                                     const double *, const double *, const double *, double, int, int, int, int, int,
                                     int, double *, double *, hipStream_t);
template int launch_tgv_sums<float>(const float *, const float *, const float *, const float *const[9], const float *,
                                    const float *, const float *, double, int, int, int, int, int, int, double *,
                                    double *, hipStream_t);

// pre_correc on free-slip faces (navier.f90:600-614, 686-704, 752-770): the
// normal velocity component on plane `idx` of direction `dir` set to zero,
// in the FP64 velocity and its FP32 copy when there is one.
__global__ void __launch_bounds__(EW_THREADS) zero_face_kernel(double *__restrict__ u, float *__restrict__ u32,
                                                               int nx, int ny, int nz, int dir, int idx) {
  const int na = dir == 0 ? ny : nx, nb = dir == 2 ? ny : nz;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= na * nb) return;
  const int a = q % na, b = q / na;
  size_t o;
  if (dir == 0) o = (size_t)idx + (size_t)nx * (a + (size_t)ny * b);
  else if (dir == 1) o = (size_t)a + (size_t)nx * (idx + (size_t)ny * b);
  else o = (size_t)a + (size_t)nx * (b + (size_t)ny * idx);
  u[o] = 0.0;
  if (u32) u32[o] = 0.0f;
}

int launch_zero_face(double *u, float *u32, int nx, int ny, int nz, int dir, int idx, hipStream_t st) {
  const int cnt = dir == 0 ? ny * nz : (dir == 1 ? nx * nz : nx * ny);
  hipLaunchKernelGGL(zero_face_kernel, dim3((cnt + EW_THREADS - 1) / EW_THREADS), dim3(EW_THREADS), 0, st, u, u32, nx,
                     ny, nz, dir, idx);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}

// max and mean |.| of a field (navier.f90:346-361)
template <typename T>
__global__ void __launch_bounds__(EW_THREADS) absmax_partials(const T *__restrict__ a, size_t n,
                                                              double *__restrict__ partial) {
  double mx = -1.0e300, sm = 0;
  const size_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const size_t beg = (size_t)blockIdx.x * chunk;
  const size_t end = beg + chunk < n ? beg + chunk : n;
  for (size_t i = beg + threadIdx.x; i < end; i += blockDim.x) {
    const double v = a[i];
    mx = v > mx ? v : mx;
    sm += fabs(v);
  }
  __shared__ double r0[EW_THREADS], r1[EW_THREADS];
  r0[threadIdx.x] = mx;
  r1[threadIdx.x] = sm;
  __syncthreads();
  for (int s = EW_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      r0[threadIdx.x] = r0[threadIdx.x] > r0[threadIdx.x + s] ? r0[threadIdx.x] : r0[threadIdx.x + s];
      r1[threadIdx.x] += r1[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = r0[0];
    partial[gridDim.x + blockIdx.x] = r1[0];
  }
}
__global__ void __launch_bounds__(EW_THREADS) finish_absmax(const double *__restrict__ partial, int nb, size_t n,
                                                            double *__restrict__ out) {
  __shared__ double r0[EW_THREADS], r1[EW_THREADS];
  double mx = -1.0e300, sm = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    mx = partial[i] > mx ? partial[i] : mx;
    sm += partial[nb + i];
  }
  r0[threadIdx.x] = mx;
  r1[threadIdx.x] = sm;
  __syncthreads();
  for (int s = EW_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      r0[threadIdx.x] = r0[threadIdx.x] > r0[threadIdx.x + s] ? r0[threadIdx.x] : r0[threadIdx.x + s];
      r1[threadIdx.x] += r1[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = r0[0];
    out[1] = r1[0] / (double)n;
  }
}
// the two stages apart: partials of a field (RED_BLOCKS pairs), and the
// finish over nb pairs of any origin
template <typename T>
int launch_absmax_partials(const T *a, size_t n, double *partial, int *nb, hipStream_t st) {
  hipLaunchKernelGGL((absmax_partials<T>), dim3(RED_BLOCKS), dim3(EW_THREADS), 0, st, a, n, partial);
  X3D_HIP(hipGetLastError());
  *nb = RED_BLOCKS;
  return X3D_OK;
}
template int launch_absmax_partials<double>(const double *, size_t, double *, int *, hipStream_t);
template int launch_absmax_partials<float>(const float *, size_t, double *, int *, hipStream_t);

int launch_finish_absmax(const double *partial, int nb, size_t n, double *out2, hipStream_t st) {
  hipLaunchKernelGGL(finish_absmax, dim3(1), dim3(EW_THREADS), 0, st, partial, nb, n, out2);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}

template <typename T>
int launch_max_mean(const T *a, size_t n, double *partial, double *out2, hipStream_t st) {
  hipLaunchKernelGGL((absmax_partials<T>), dim3(RED_BLOCKS), dim3(EW_THREADS), 0, st, a, n, partial);
  hipLaunchKernelGGL(finish_absmax, dim3(1), dim3(EW_THREADS), 0, st, partial, RED_BLOCKS, n, out2);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}
template int launch_max_mean<double>(const double *, size_t, double *, double *, hipStream_t);
template int launch_max_mean<float>(const float *, size_t, double *, double *, hipStream_t);

}  // namespace x3d
