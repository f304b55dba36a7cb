// Wall-bounded (channel) pieces of the device time step:
//   init_channel        Case-Channel.f90:96-128  (laminar profile + noise)
//   channel_cfr         Case-Channel.f90:220-261 (constant flow rate)
//   momentum_forcing    Case-Channel.f90:396-410 (constant pressure gradient)
//   pre_correc, ncly=2  navier.f90:616-690       (wall velocity)
//   gradp, ncly=2       navier.f90:456-479       (wall pressure gradient / gdt)
// The fields are the reference's (nx,ny,nz) arrays, i fastest; the FP64
// velocity and, in the FP32-bulk step, its FP32 copy are updated together.
#include <cmath>

#include "x3d_internal.h"

namespace x3d {

namespace {

constexpr int CH_THREADS = 256;
constexpr int CH_RED_BLOCKS = 1024;

unsigned ch_blocks(size_t n) {
  size_t b = (n + CH_THREADS * 4 - 1) / (CH_THREADS * 4);
  if (b < 1) b = 1;
  if (b > 65535u * 8u) b = 65535u * 8u;
  return (unsigned)b;
}

// Uniform [0,1) of the counter seed*2^40 + comp*2^36 + idx: the SplitMix64
// output function (Steele, Lea & Flood, 2014) applied to counter * golden
// gamma, top 53 bits.  A counter-based stream gives every point its own
// number independently of the launch shape, on the host and the device.
__device__ __forceinline__ double hash_uniform(unsigned long long seed, int comp, unsigned long long idx) {
  unsigned long long z = ((seed << 40) + ((unsigned long long)comp << 36) + idx) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// init_channel, iin=1 with idir_stream=1 (Case-Channel.f90:96-128):
// u = init_noise um (2 r - 1) + 1 - y^2, v and w = init_noise um (2 r - 1),
// um = exp(-0.2 y^2), y = yc[j] (the wall-normal coordinate minus yly/2).
__global__ void __launch_bounds__(CH_THREADS)
    channel_init_kernel(double *__restrict__ ux, double *__restrict__ uy, double *__restrict__ uz,
                        const double *__restrict__ yc, int nx, int ny, size_t n, double noise,
                        unsigned long long seed) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    const int j = (int)((q / (size_t)nx) % (size_t)ny);
    const double y = yc[j];
    const double um = exp(-0.2 * y * y);
    const double r0 = hash_uniform(seed, 0, q), r1 = hash_uniform(seed, 1, q), r2 = hash_uniform(seed, 2, q);
    ux[q] = noise * um * (2.0 * r0 - 1.0) + 1.0 - y * y;
    uy[q] = noise * um * (2.0 * r1 - 1.0);
    uz[q] = noise * um * (2.0 * r2 - 1.0);
  }
}

// Partial sums of u / ppy(j) over contiguous chunks (channel_cfr's ub before
// its coefficient); ppy == nullptr stands for the uniform mesh's ppy = 1.
__global__ void __launch_bounds__(CH_THREADS)
    bulk_partials(const double *__restrict__ u, const double *__restrict__ ppy, int nx, int ny, size_t n,
                  double *__restrict__ partial) {
  const size_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const size_t beg = (size_t)blockIdx.x * chunk;
  const size_t end = beg + chunk < n ? beg + chunk : n;
  double sm = 0.0;
  for (size_t q = beg + threadIdx.x; q < end; q += blockDim.x) {
    const double v = u[q];
    sm += ppy ? v / ppy[(q / (size_t)nx) % (size_t)ny] : v;
  }
  __shared__ double r[CH_THREADS];
  r[threadIdx.x] = sm;
  __syncthreads();
  for (int s = CH_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) r[threadIdx.x] += r[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = r[0];
}

// out[0] = ub = coeff * sum, out[1] = can = -(constant - ub)
__global__ void __launch_bounds__(CH_THREADS)
    bulk_finish(const double *__restrict__ partial, int nb, double coeff, double constant, double *__restrict__ out) {
  __shared__ double r[CH_THREADS];
  double sm = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) sm += partial[i];
  r[threadIdx.x] = sm;
  __syncthreads();
  for (int s = CH_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) r[threadIdx.x] += r[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double ub = r[0] * coeff;
    out[0] = ub;
    out[1] = -(constant - ub);
  }
}

// u <- u - can (Case-Channel.f90:250-256), can read from the device
__global__ void __launch_bounds__(CH_THREADS)
    cfr_shift_kernel(double *__restrict__ u, float *__restrict__ u32, const double *__restrict__ can_p, size_t n) {
  const double can = can_p[1];
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    const double r = u[q] - can;
    u[q] = r;
    if (u32) u32[q] = (float)r;
  }
}

// f += c (momentum_forcing_channel with cpg, Case-Channel.f90:403-410)
template <typename T>
__global__ void __launch_bounds__(CH_THREADS) add_const_kernel(T *__restrict__ f, double c, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) f[q] = (T)((double)f[q] + c);
}

// the spin-up rotation of momentum_forcing_channel (Case-Channel.f90:413-418):
// f += a v, with a = -wrotation for dux (v = uy) and +wrotation for duy (v = ux)
template <typename T>
__global__ void __launch_bounds__(CH_THREADS) axpy_kernel(T *__restrict__ f, const T *__restrict__ v, T a, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) f[q] = f[q] + a * v[q];
}

// pre_correc on the y walls (navier.f90:616-690, no MHD, byx = byy = byz = 0):
// ux = 0 + dpdx gdt, uy = 0, uz = 0 + dpdz gdt on plane j of each wall; the
// stored planes are dpdx(i,k), dpdz(i,k) at [k*nx + i].
__global__ void __launch_bounds__(CH_THREADS)
    wall_velocity_kernel(double *__restrict__ ux, double *__restrict__ uy, double *__restrict__ uz,
                         float *__restrict__ vx, float *__restrict__ vy, float *__restrict__ vz,
                         const double *__restrict__ dpdx, const double *__restrict__ dpdz, double gdt, int nx,
                         int ny, int nz, int j) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nx * nz) return;
  const int i = q % nx, k = q / nx;
  const size_t o = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
  const double a = 0.0 + dpdx[q] * gdt, b = 0.0, c = 0.0 + dpdz[q] * gdt;
  ux[o] = a;
  uy[o] = b;
  uz[o] = c;
  if (vx) {
    vx[o] = (float)a;
    vy[o] = (float)b;
    vz[o] = (float)c;
  }
}

// gradp on the y walls (navier.f90:456-479): plane j of a gradient component
// over gdt into the stored plane dst[k*nx + i]
template <typename T>
__global__ void __launch_bounds__(CH_THREADS)
    wall_store_kernel(const T *__restrict__ g, double *__restrict__ dst, double gdt, int nx, int ny, int nz, int j) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nx * nz) return;
  const int i = q % nx, k = q / nx;
  dst[q] = (double)g[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)] / gdt;
}

}  // namespace

int launch_channel_init(double *ux, double *uy, double *uz, const double *yc, int nx, int ny, int nz, double noise,
                        unsigned long long seed, hipStream_t st) {
  const size_t n = (size_t)nx * ny * nz;
  if (n >= ((size_t)1 << 36)) return fail(X3D_ERR_UNSUPPORTED, "channel noise: more than 2^36 points");
  hipLaunchKernelGGL(channel_init_kernel, dim3(ch_blocks(n)), dim3(CH_THREADS), 0, st, ux, uy, uz, yc, nx, ny, n,
                     noise, seed);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}

// out2[0] = ub = coeff * sum u / ppy, out2[1] = -(constant - ub); partial
// holds CH_RED_BLOCKS doubles
int launch_bulk(const double *u, const double *ppy, int nx, int ny, int nz, double coeff, double constant,
                double *partial, double *out2, hipStream_t st) {
  const size_t n = (size_t)nx * ny * nz;
  hipLaunchKernelGGL(bulk_partials, dim3(CH_RED_BLOCKS), dim3(CH_THREADS), 0, st, u, ppy, nx, ny, n, partial);
  hipLaunchKernelGGL(bulk_finish, dim3(1), dim3(CH_THREADS), 0, st, partial, CH_RED_BLOCKS, coeff, constant, out2);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}

int launch_cfr_shift(double *u, float *u32, const double *out2, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(cfr_shift_kernel, dim3(ch_blocks(n)), dim3(CH_THREADS), 0, st, u, u32, out2, n);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}

template <typename T>
int launch_add_const(T *f, double c, size_t n, hipStream_t st) {
  hipLaunchKernelGGL((add_const_kernel<T>), dim3(ch_blocks(n)), dim3(CH_THREADS), 0, st, f, c, n);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}
template int launch_add_const<double>(double *, double, size_t, hipStream_t);
template int launch_add_const<float>(float *, double, size_t, hipStream_t);

template <typename T>
int launch_axpy(T *f, const T *v, double a, size_t n, hipStream_t st) {
  hipLaunchKernelGGL((axpy_kernel<T>), dim3(ch_blocks(n)), dim3(CH_THREADS), 0, st, f, v, (T)a, n);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}
template int launch_axpy<double>(double *, const double *, double, size_t, hipStream_t);
template int launch_axpy<float>(float *, const float *, double, size_t, hipStream_t);

int launch_wall_velocity(double *const u[3], float *const v[3], const double *dpdx, const double *dpdz, double gdt,
                         int nx, int ny, int nz, int j, hipStream_t st) {
  const int cnt = nx * nz;
  hipLaunchKernelGGL(wall_velocity_kernel, dim3((cnt + CH_THREADS - 1) / CH_THREADS), dim3(CH_THREADS), 0, st, u[0],
                     u[1], u[2], v[0], v[1], v[2], dpdx, dpdz, gdt, nx, ny, nz, j);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}

template <typename T>
int launch_wall_store(const T *g, double *dst, double gdt, int nx, int ny, int nz, int j, hipStream_t st) {
  const int cnt = nx * nz;
  hipLaunchKernelGGL((wall_store_kernel<T>), dim3((cnt + CH_THREADS - 1) / CH_THREADS), dim3(CH_THREADS), 0, st, g,
                     dst, gdt, nx, ny, nz, j);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}
template int launch_wall_store<double>(const double *, double *, double, int, int, int, int, hipStream_t);
template int launch_wall_store<float>(const float *, double *, double, int, int, int, int, hipStream_t);

}  // namespace x3d
