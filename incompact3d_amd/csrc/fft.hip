// Spectral pressure Poisson solver for gfx950 (no FFT library).
//
// poisson_000 (poisson.f90:298-410): its shift post-processing cancels
// exactly (forward and backward factors are complex conjugates around the
// conjugated division), so the solve is  p_hat = -f_hat / K  with modes of
// K < 1e-16 set to zero, K being the modified wavenumber of waves()
// (poisson.f90:1706-1783).  poisson_010 (poisson.f90:665-1013) is the same
// with a DCT-II along y computed, as in the reference, from an even/odd
// reordering plus a complex FFT (Makhoul's algorithm), and with the
// (nx/2+1, :, nz/2+1) line zeroed.  poisson_11x (poisson.f90:1019-1465,
// istret=0) adds the same even/odd reorder along x and, for bcz=1, along z;
// along x, the real-to-complex direction here, the half spectrum is shifted by
// e^{-i pi k/(2 nx)}, after which its real part holds cosine mode k and its
// imaginary part mode nx-k, so the two are divided by K at those two modes
// (the real/imaginary split of waves(), poisson.f90:1740-1800).
//
// Passes over the half spectrum cw (nxh = nx/2+1, ny, nz), kx fastest:
//   1 r2c along x (real line packed as nx/2 complex points, FFT, unpack)
//   2 c2c forward along y            (DCT-II for a non-periodic y)
//   3 c2c forward along z, divide by -K (FP64), c2c inverse along z
//   4 c2c inverse along y            (DCT-III for a non-periodic y)
//   5 c2r along x
// Line FFTs run in registers: each thread holds E = 16 points of a line in
// cyclic order (point t + m*NT for thread t of NT), does radix-16 (then
// radix-2/4/8) Stockham butterflies on them and exchanges through LDS only
// between stages -- one exchange for lines of up to 256 points.  The
// first stage reads its inputs straight from HBM and the last stage leaves
// the outputs in the same cyclic order, so they are stored straight back;
// the DCT reorders, the Poisson divide and the real-to-complex packing are
// folded into those loads and stores.
#include <algorithm>
#include <cmath>

#include "x3d_internal.h"

namespace x3d {

template <typename T>
struct alignas(2 * sizeof(T)) cplx {
  T x, y;
};
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> cconj(cplx<T> a) { return {a.x, -a.y}; }

template <typename T>
struct PoissonTables {
  const cplx<T> *twx, *twy, *twz;   // exp(-2 pi i m / N) for N = nx/2, ny, nz
  const cplx<T> *rtx;               // exp(-2 pi i k / nx), k = 0..nx/2 (r2c unpack)
  const cplx<T> *dcty;              // exp(-i pi k / (2 ny)), k = 0..ny-1 (DCT along y)
  const cplx<T> *dctx;              // exp(-i pi k / (2 nx)), k = 0..nx/2 (DCT along x)
  const cplx<T> *dctz;              // exp(-i pi k / (2 nz)), k = 0..nz-1 (DCT along z)
  const double *kx2, *tx;           // modified wavenumber^2 and transfer fn, per kx (nxh; nx+1 modes with xdct)
  const double *ky2, *ty;           // per ky (ny)
  const double *kz2, *tz;           // per kz (nz)
};

struct PoissonPlan {
  int nx, ny, nz, nxh;              // pressure-grid sizes (ny = nym when y is not periodic)
  int ydct;                         // 1 when y is non-periodic (poisson_010, poisson_11x)
  int xdct, zdct;                   // 1 when x / z are non-periodic (poisson_11x; x alone: poisson_100)
  double scale;                     // 1/(nx*ny*nz)
  void *tab64 = nullptr, *tab32 = nullptr;
  PoissonTables<double> t64;
  PoissonTables<float> t32;
  void *cw64 = nullptr, *cw32 = nullptr;   // spectral work arrays
  // stretched y (poisson_010 with istret != 0): factors of the pentadiagonal
  // systems, five FP64 arrays [kz][system][row][kx] (see build_penta_factors);
  // ystret = the number of systems per (kx, kz): 2 (istret 1, 2), 1 (istret 3)
  int ystret = 0;
  double *pf = nullptr;
  size_t pf_stride = 0;
};

static int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return (1 << l) == n ? l : -1;
}

// ---------------------------------------------------------------- radix-R DFTs in registers
// Natural order in and out; forward uses exp(-2 pi i/R), inverse exp(+2 pi i/R).
template <typename T, bool INV>
__device__ __forceinline__ cplx<T> rot_mi(cplx<T> a) {   // a*(-i) forward, a*(+i) inverse
  return INV ? cplx<T>{-a.y, a.x} : cplx<T>{a.y, -a.x};
}

template <typename T, bool INV>
__device__ __forceinline__ void dft4(cplx<T> &x0, cplx<T> &x1, cplx<T> &x2, cplx<T> &x3) {
  const cplx<T> a = cadd(x0, x2), b = csub(x0, x2), c = cadd(x1, x3), d = rot_mi<T, INV>(csub(x1, x3));
  x0 = cadd(a, c);
  x1 = cadd(b, d);
  x2 = csub(a, c);
  x3 = csub(b, d);
}

// exp(-+2 pi i m/16) for the constant twiddles of the radix-8/16 kernels
template <typename T, bool INV, int M>
__device__ __forceinline__ cplx<T> w16() {
  constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173, h = 0.70710678118654752440;
  constexpr int m = M & 15;
  constexpr double re = m == 0 ? 1.0 : m == 1 ? c1 : m == 2 ? h : m == 3 ? s1 : m == 4 ? 0.0 : m == 5 ? -s1
                      : m == 6 ? -h : m == 7 ? -c1 : m == 8 ? -1.0 : m == 9 ? -c1 : m == 10 ? -h : m == 11 ? -s1
                      : m == 12 ? 0.0 : m == 13 ? s1 : m == 14 ? h : c1;
  constexpr double im = m == 0 ? 0.0 : m == 1 ? -s1 : m == 2 ? -h : m == 3 ? -c1 : m == 4 ? -1.0 : m == 5 ? -c1
                      : m == 6 ? -h : m == 7 ? -s1 : m == 8 ? 0.0 : m == 9 ? s1 : m == 10 ? h : m == 11 ? c1
                      : m == 12 ? 1.0 : m == 13 ? c1 : m == 14 ? h : s1;
  return cplx<T>{(T)re, (T)(INV ? -im : im)};
}

template <typename T, int R, bool INV>
struct Dft;

template <typename T, bool INV>
struct Dft<T, 1, INV> {
  static __device__ __forceinline__ void run(cplx<T> (&)[1]) {}
};
template <typename T, bool INV>
struct Dft<T, 2, INV> {
  static __device__ __forceinline__ void run(cplx<T> (&v)[2]) {
    const cplx<T> a = cadd(v[0], v[1]), b = csub(v[0], v[1]);
    v[0] = a;
    v[1] = b;
  }
};
template <typename T, bool INV>
struct Dft<T, 4, INV> {
  static __device__ __forceinline__ void run(cplx<T> (&v)[4]) { dft4<T, INV>(v[0], v[1], v[2], v[3]); }
};
// 8 = 4 x 2: r = 2a + b, k = k1 + 4 k2
template <typename T, bool INV>
struct Dft<T, 8, INV> {
  static __device__ __forceinline__ void run(cplx<T> (&v)[8]) {
    dft4<T, INV>(v[0], v[2], v[4], v[6]);
    dft4<T, INV>(v[1], v[3], v[5], v[7]);
    v[3] = cmul(v[3], w16<T, INV, 2>());
    v[5] = rot_mi<T, INV>(v[5]);
    v[7] = cmul(v[7], w16<T, INV, 6>());
    cplx<T> o[8];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      o[k1] = cadd(v[2 * k1], v[2 * k1 + 1]);
      o[k1 + 4] = csub(v[2 * k1], v[2 * k1 + 1]);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = o[k];
  }
};
// 16 = 4 x 4: r = 4a + b, k = k1 + 4 k2
template <typename T, bool INV>
struct Dft<T, 16, INV> {
  static __device__ __forceinline__ void run(cplx<T> (&v)[16]) {
#pragma unroll
    for (int b = 0; b < 4; ++b) dft4<T, INV>(v[b], v[4 + b], v[8 + b], v[12 + b]);
    // Y_b[k1] sits in v[4 k1 + b]; twiddle by w16^(b k1)
    v[5] = cmul(v[5], w16<T, INV, 1>());
    v[6] = cmul(v[6], w16<T, INV, 2>());
    v[7] = cmul(v[7], w16<T, INV, 3>());
    v[9] = cmul(v[9], w16<T, INV, 2>());
    v[10] = rot_mi<T, INV>(v[10]);
    v[11] = cmul(v[11], w16<T, INV, 6>());
    v[13] = cmul(v[13], w16<T, INV, 3>());
    v[14] = cmul(v[14], w16<T, INV, 6>());
    v[15] = cmul(v[15], w16<T, INV, 9>());
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) dft4<T, INV>(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
    // X[k1 + 4 k2] sits in v[4 k1 + k2]
    cplx<T> o[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) o[k1 + 4 * k2] = v[4 * k1 + k2];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = o[k];
  }
};

// ---------------------------------------------------------------- line FFT
// Points per thread and threads per line for a line of N points.
template <int N>
struct LineShape {
  static constexpr int E = N >= 16 ? 16 : N;
  static constexpr int NT = N / E;
};

// Stockham stages on a line held in cyclic order: thread t owns points
// t + m*NT (m < E).  A stage of radix R does E/R butterflies per thread,
// j = t + q*NT, whose inputs are the thread's points q + r*E/R.  Its outputs
// go to (j/Ns)*Ns*R + j%Ns + r*Ns; for the last stage (Ns*R = N) that is the
// input slot again, otherwise they go through LDS and are read back in
// cyclic order.  tw[m] = exp(-2 pi i m/N) in LDS.  LD maps a point index to
// the thread's LDS slot.  Every thread of the block calls this.
template <typename T, int N, bool INV, int NS, typename LD>
__device__ __forceinline__ void fft_stages(cplx<T> (&x)[LineShape<N>::E], const int t, const LD &ld,
                                           const cplx<T> *tw) {
  constexpr int E = LineShape<N>::E, NT = LineShape<N>::NT;
  constexpr int R = (N / NS) < E ? (N / NS) : E;
  constexpr int Q = E / R;
  constexpr bool LAST = NS * R == N;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = t + q * NT;
    const int k = j & (NS - 1);
    cplx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = x[q + r * Q];
    if (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cplx<T> w = tw[r * k * (N / (NS * R))];
        if (INV) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    }
    Dft<T, R, INV>::run(v);
#pragma unroll
    for (int r = 0; r < R; ++r) x[q + r * Q] = v[r];
  }
  if constexpr (!LAST) {
    __syncthreads();   // the previous use of the exchange buffer is over
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int j = t + q * NT;
      const int k = j & (NS - 1);
      const int o = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) ld.store(o + r * NS, x[q + r * Q]);
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < E; ++m) x[m] = ld.load(t + m * NT);
    fft_stages<T, N, INV, NS * R, LD>(x, t, ld, tw);
  }
}

// LDS slot maps.  Strided (y/z) blocks: point-major rows of TL lines, lanes
// run over lines.  x blocks: line-major, lanes run over points; one pad slot
// per 16 points keeps the radix-16 scatter off a single bank.
template <typename T>
struct LdRows {
  cplx<T> *buf;
  int TL, l;
  __device__ __forceinline__ void store(int p, cplx<T> v) const { buf[p * TL + l] = v; }
  __device__ __forceinline__ cplx<T> load(int p) const { return buf[p * TL + l]; }
};
template <typename T>
struct LdLine {
  cplx<T> *buf;
  int base;
  __device__ __forceinline__ void store(int p, cplx<T> v) const { buf[base + p + (p >> 4)] = v; }
  __device__ __forceinline__ cplx<T> load(int p) const { return buf[base + p + (p >> 4)]; }
};

template <typename T>
__device__ __forceinline__ void stage_table(cplx<T> *dst, const cplx<T> *__restrict__ src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

constexpr int FFT_THREADS = 256;

// x blocks: XT lines of M = nx/2 complex points, NT threads per line.
template <int M>
struct XShape {
  static constexpr int NT = LineShape<M>::NT;
  static constexpr int TL = FFT_THREADS / NT;
  static constexpr int LP = M + M / 16;             // padded line length in LDS
  template <typename T>
  static constexpr size_t lds() { return sizeof(cplx<T>) * ((size_t)TL * LP + 2 * M + 1); }
};
// strided blocks: TL kx x N points, NT threads per line.  TL = 256/NT lines
// for FP64 and twice that for FP32, so that each point row a block reads is
// 256 bytes either way (at most 64 lines, 512 threads).  The nx/2+1 kx lines
// do not fill whole blocks: either the last column of blocks runs partly
// empty, or (EXTRA) each block takes one of the remaining (kx, plane) lines
// as well, on NT more threads.
template <typename T, int N>
struct SShape {
  static constexpr int NT = LineShape<N>::NT;
  static constexpr int TL0 = (FFT_THREADS / NT) * (sizeof(T) == 4 ? 2 : 1);
  static constexpr int TL = TL0 < 64 ? TL0 : 64;
  static constexpr int threads(bool extra) { return (TL + (extra ? 1 : 0)) * NT; }
  // ztab: the z pass's tz and kz2 tables in FP64 (lines up to 256 points,
  // where the 4 KB keep two blocks per CU)
  static constexpr bool ZTAB = N <= 256;
  static constexpr size_t lds(bool extra, bool ztab) {
    return sizeof(cplx<T>) * ((size_t)(TL + (extra ? 1 : 0)) * N + 2 * N) + (ztab ? 2 * N * sizeof(double) : 0);
  }
};

// row of point p under the even/odd reorder of poisson.f90:688-697 / :997-1008
__device__ __forceinline__ int dct_row(int p, int N) { return p < (N >> 1) ? 2 * p : 2 * (N - 1 - p) + 1; }

// ---------------------------------------------------------------- pass 1: r2c along x
// X_k = (Z_k + conj Z_{M-k})/2 - i/2 e^{-2 pi i k/nx} (Z_k - conj Z_{M-k}), k = 0..M
template <typename T, int M>
__global__ void __launch_bounds__(FFT_THREADS) r2c_x_kernel(const T *__restrict__ in, cplx<T> *__restrict__ out,
                                                             int nlines, int xdct, PoissonTables<T> tb) {
  using S = XShape<M>;
  constexpr int E = LineShape<M>::E, NT = S::NT, TL = S::TL;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
  cplx<T> *tw = buf + TL * S::LP, *rt = tw + M;
  const int t = threadIdx.x % NT, l = threadIdx.x / NT;
  const int line = blockIdx.x * TL + l;
  const bool valid = line < nlines;
  stage_table(tw, tb.twx, M);
  stage_table(rt, tb.rtx, M + 1);
  cplx<T> x[E];
  if (xdct) {
    // even/odd reorder along x (poisson.f90:1086-1095): point p of the FFT
    // input is row dct_row(p) of the line
    const T *src = in + (size_t)line * 2 * M;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int q = t + m * NT;
      x[m] = valid ? cplx<T>{src[dct_row(2 * q, 2 * M)], src[dct_row(2 * q + 1, 2 * M)]} : cplx<T>{T(0), T(0)};
    }
  } else {
    const cplx<T> *src = reinterpret_cast<const cplx<T> *>(in + (size_t)line * 2 * M);
#pragma unroll
    for (int m = 0; m < E; ++m) x[m] = valid ? src[t + m * NT] : cplx<T>{T(0), T(0)};
  }
  __syncthreads();
  const LdLine<T> ld{buf, l * S::LP};
  fft_stages<T, M, false, 1>(x, t, ld, tw);
  // Z_{M-k} belongs to another thread: exchange
  __syncthreads();
#pragma unroll
  for (int m = 0; m < E; ++m) ld.store(t + m * NT, x[m]);
  __syncthreads();
  if (valid) {
    cplx<T> *dst = out + (size_t)line * (M + 1);
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int k = t + m * NT;
      const cplx<T> zk = x[m];
      const cplx<T> zm = cconj(ld.load(k == 0 ? 0 : M - k));
      const cplx<T> e = {T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y + zm.y)};
      const cplx<T> o = {T(0.5) * (zk.y - zm.y), -T(0.5) * (zk.x - zm.x)};   // (Z_k - conj Z_{M-k}) / (2i)
      cplx<T> v = cadd(e, cmul(rt[k], o));
      // cosine transform along x: shift by e^{-i pi k/(2 nx)} (poisson.f90:1116-1125)
      if (xdct) v = cmul(tb.dctx[k], v);
      dst[k] = v;
      if (k == 0) {   // k = M uses Z_0 on both sides
        cplx<T> vm = cadd(e, cmul(rt[M], o));
        if (xdct) vm = cmul(tb.dctx[M], vm);
        dst[M] = vm;
      }
    }
  }
}

// ---------------------------------------------------------------- pass 5: c2r along x
// Z_k = (X_k + conj X_{M-k}) + i e^{+2 pi i k/nx} (X_k - conj X_{M-k}),  k < M
template <typename T, int M>
__global__ void __launch_bounds__(FFT_THREADS) c2r_x_kernel(const cplx<T> *__restrict__ in, T *__restrict__ out,
                                                             int nlines, int xdct, PoissonTables<T> tb) {
  using S = XShape<M>;
  constexpr int E = LineShape<M>::E, NT = S::NT, TL = S::TL;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
  cplx<T> *tw = buf + TL * S::LP, *rt = tw + M;
  const int t = threadIdx.x % NT, l = threadIdx.x / NT;
  const int line = blockIdx.x * TL + l;
  const bool valid = line < nlines;
  stage_table(tw, tb.twx, M);
  stage_table(rt, tb.rtx, M + 1);
  const cplx<T> *src = in + (size_t)line * (M + 1);
  cplx<T> xk[E], xm[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int k = t + m * NT;
    xk[m] = valid ? src[k] : cplx<T>{T(0), T(0)};
    xm[m] = valid ? src[M - k] : cplx<T>{T(0), T(0)};
    if (xdct) {
      // undo the cosine-transform shift: e^{+i pi k/(2 nx)} (poisson.f90:1428-1437)
      xk[m] = cmul(cconj(tb.dctx[k]), xk[m]);
      xm[m] = cmul(cconj(tb.dctx[M - k]), xm[m]);
    }
  }
  __syncthreads();
  cplx<T> x[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int k = t + m * NT;
    const cplx<T> c = cconj(xm[m]);
    const cplx<T> e = cadd(xk[m], c);
    const cplx<T> od = cmul(cconj(rt[k]), csub(xk[m], c));
    x[m] = cplx<T>{e.x - od.y, e.y + od.x};
  }
  const LdLine<T> ld{buf, l * S::LP};
  fft_stages<T, M, true, 1>(x, t, ld, tw);
  if (valid && xdct) {
    // inverse reorder along x (poisson.f90:1452-1460)
    T *dst = out + (size_t)line * 2 * M;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int q = t + m * NT;
      dst[dct_row(2 * q, 2 * M)] = x[m].x;
      dst[dct_row(2 * q + 1, 2 * M)] = x[m].y;
    }
  } else if (valid) {
    cplx<T> *dst = reinterpret_cast<cplx<T> *>(out + (size_t)line * 2 * M);
#pragma unroll
    for (int m = 0; m < E; ++m) dst[t + m * NT] = x[m];
  }
}

// ---------------------------------------------------------------- passes 2-4: strided lines
// Z_FWD / Z_INV: the z transform alone, around the stretched-y solve
enum { FFT_Y_FWD = 0, FFT_Y_INV = 1, FFT_Z_SOLVE = 2, FFT_Z_FWD = 3, FFT_Z_INV = 4 };
template <int MODE>
struct ZPass {
  static constexpr bool value = MODE == FFT_Z_SOLVE || MODE == FFT_Z_FWD || MODE == FFT_Z_INV;
};



// COS: the poisson_11x z pass (cosine transform along x, and along z when
// zdct), compiled separately so the periodic passes keep their registers.
template <typename T, int N, int MODE, bool COS, bool EXTRA>
__global__ void __launch_bounds__((SShape<T, N>::threads(EXTRA)))
    strided_kernel(cplx<T> *__restrict__ cw, int nmain, int nxh, int ny, int nz, int ydct, double scale,
                   int zero_line, int xdct, int zdct, PoissonTables<T> tb) {
  using S = SShape<T, N>;
  constexpr int E = LineShape<N>::E, NT = S::NT, TL = S::TL, TLB = TL + (EXTRA ? 1 : 0);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
  cplx<T> *tw = buf + TLB * N, *dct = tw + N;
  const int tid = threadIdx.x;
  // line l of the block: a = kx, b = kz (y passes) or ky (z pass)
  int l, t, a, b;
  bool valid;
  if (!EXTRA || tid < TL * NT) {
    l = tid % TL;
    t = tid / TL;
    a = blockIdx.x * TL + l;
    b = blockIdx.y;
    valid = a < nxh;
  } else {
    // the extra line: g-th of the (nxh - nmain) x nplanes lines past the
    // whole blocks (the host launches at least that many blocks)
    const int nplanes = gridDim.y, ntail = (nxh - nmain) * nplanes;
    const int g = blockIdx.y * gridDim.x + blockIdx.x;
    l = TL;
    t = tid - TL * NT;
    valid = g < ntail;
    const int gg = valid ? g : 0;
    a = nmain + gg / nplanes;
    b = gg % nplanes;
  }
  constexpr bool ZL = ZPass<MODE>::value;
  const size_t se = ZL ? (size_t)nxh * ny : (size_t)nxh;
  const size_t base = ZL ? (size_t)b * nxh + a : (size_t)b * nxh * ny + a;
  stage_table(tw, ZL ? tb.twz : tb.twy, N);
  // the cosine transform of this pass: y in the y passes, z in the z pass
  const bool dct_here = ZL ? (MODE == FFT_Z_SOLVE && COS && zdct != 0) : ydct != 0;
  if (dct_here) stage_table(dct, ZL ? tb.dctz : tb.dcty, N);
  // the z pass reads tz and kz2 at its 16 points per thread: staged in LDS
  // with the twiddles instead of read from global memory after the FFT
  constexpr bool ZT = MODE == FFT_Z_SOLVE && !COS && SShape<T, N>::ZTAB;
  double *ztab = reinterpret_cast<double *>(dct + N);
  if constexpr (ZT) {
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
      ztab[i] = tb.tz[i];
      ztab[N + i] = tb.kz2[i];
    }
  }
  const bool reorder_in = (MODE == FFT_Y_FWD || MODE == FFT_Z_SOLVE) && dct_here;
  cplx<T> x[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int p = t + m * NT;
    const int row = reorder_in ? dct_row(p, N) : p;
    x[m] = valid ? cw[base + (size_t)row * se] : cplx<T>{T(0), T(0)};
  }
  if constexpr (MODE == FFT_Y_INV) {
    if (dct_here) {
      // DCT-III input side: V_k = e^{+i pi k/2N} (C_k - i C_{N-k}), V_0 = C_0
      // (poisson.f90:924-943); C_{N-k} is read from its own row
      cplx<T> cm[E];
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int k = t + m * NT;
        cm[m] = (valid && k > 0) ? cw[base + (size_t)(N - k) * se] : cplx<T>{T(0), T(0)};
      }
      __syncthreads();   // dct table
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int k = t + m * NT;
        x[m] = cmul(cconj(dct[k]), cplx<T>{x[m].x + cm[m].y, x[m].y - cm[m].x});
      }
    }
  }
  __syncthreads();   // tables staged
  const LdRows<T> ld{buf, TLB, l};
  fft_stages<T, N, MODE == FFT_Y_INV || MODE == FFT_Z_INV, 1>(x, t, ld, tw);

  if ((MODE == FFT_Y_FWD || MODE == FFT_Z_SOLVE) && dct_here) {
    // DCT-II output side: C_k = (e^{-i pi k/2N} V_k + e^{+i pi k/2N} V_{N-k}) / 2,
    // C_0 = V_0 (poisson.f90:761-781, :1139-1160); V_{N-k} belongs to another thread
    __syncthreads();
#pragma unroll
    for (int m = 0; m < E; ++m) ld.store(t + m * NT, x[m]);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int k = t + m * NT;
      if (k > 0) {
        const cplx<T> vm = ld.load(N - k);
        const cplx<T> p = cmul(dct[k], x[m]);
        const cplx<T> q = cmul(cconj(dct[k]), vm);
        x[m] = cplx<T>{T(0.5) * (p.x + q.x), T(0.5) * (p.y + q.y)};
      }
    }
  }

  if (MODE == FFT_Z_SOLVE) {
    // divide by -K in FP64 (poisson.f90:362-371 / :795-819 / :1191-1207).
    // With a cosine transform along x the real part is mode kx and the
    // imaginary part mode nx-kx, each divided by its own K.
    cplx<T> part[E];
    if (COS && !dct_here) {
      // periodic z after the x cosine transform: the z FFT has mixed the
      // real (mode kx) and imaginary (mode nx-kx) parts; they are the
      // Hermitian and anti-Hermitian halves in kz, A_k = (D_k + conj D_{N-k})/2
      // and iB_k = (D_k - conj D_{N-k})/2, each divided by its own K below
      __syncthreads();
#pragma unroll
      for (int m = 0; m < E; ++m) ld.store(t + m * NT, x[m]);
      __syncthreads();
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int k = t + m * NT;
        part[m] = cconj(ld.load(k == 0 ? 0 : N - k));
      }
      __syncthreads();
    }
    if (valid && COS) {
      const double ky2 = tb.ky2[b], tyv = tb.ty[b];
      const int am = 2 * (nxh - 1) - a;
      const double txr = tb.tx[a], kx2r = tb.kx2[a], txi = tb.tx[am], kx2i = tb.kx2[am];
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int kz = t + m * NT;
        const double tzv = tb.tz[kz], kz2 = tb.kz2[kz];
        const double yz2 = (tyv * tzv) * (tyv * tzv);
        const double Kr = kx2r * yz2 + ky2 * (txr * tzv) * (txr * tzv) + kz2 * (txr * tyv) * (txr * tyv);
        const double Ki = kx2i * yz2 + ky2 * (txi * tzv) * (txi * tzv) + kz2 * (txi * tyv) * (txi * tyv);
        const double fr = Kr < 1.0e-16 ? 0.0 : -scale / Kr;
        const double fi = Ki < 1.0e-16 ? 0.0 : -scale / Ki;
        if (dct_here) {
          x[m] = cplx<T>{(T)((double)x[m].x * fr), (T)((double)x[m].y * fi)};
        } else {
          const double ar = 0.5 * ((double)x[m].x + (double)part[m].x), ai = 0.5 * ((double)x[m].y + (double)part[m].y);
          const double br = 0.5 * ((double)x[m].x - (double)part[m].x), bi = 0.5 * ((double)x[m].y - (double)part[m].y);
          x[m] = cplx<T>{(T)(ar * fr + br * fi), (T)(ai * fr + bi * fi)};
        }
      }
    } else if (valid) {
      const double ky2 = tb.ky2[b], tyv = tb.ty[b];
      const double txv = tb.tx[a], kx2 = tb.kx2[a];
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int kz = t + m * NT;
        const double tzv = ZT ? ztab[kz] : tb.tz[kz];
        const double yz = tyv * tzv, xz = txv * tzv, xy = txv * tyv;
        const double K = kx2 * (yz * yz) + ky2 * (xz * xz) + (ZT ? ztab[N + kz] : tb.kz2[kz]) * (xy * xy);
        const bool zero = (K < 1.0e-16) || (zero_line && a == (nxh - 1) && kz == (N >> 1));
        const double f = zero ? 0.0 : -scale / K;
        x[m] = cplx<T>{(T)((double)x[m].x * f), (T)((double)x[m].y * f)};
      }
    }
    if (COS && dct_here) {
      // DCT-III input side along z: V_k = e^{+i pi k/2N} (C_k - i C_{N-k}),
      // V_0 = C_0 (poisson.f90:1330-1351); C_{N-k} belongs to another thread
      __syncthreads();
#pragma unroll
      for (int m = 0; m < E; ++m) ld.store(t + m * NT, x[m]);
      __syncthreads();
      cplx<T> cm[E];
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int k = t + m * NT;
        cm[m] = k > 0 ? ld.load(N - k) : cplx<T>{T(0), T(0)};
      }
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int k = t + m * NT;
        x[m] = cmul(cconj(dct[k]), cplx<T>{x[m].x + cm[m].y, x[m].y - cm[m].x});
      }
      __syncthreads();
    }
    fft_stages<T, N, true, 1>(x, t, ld, tw);
  }

  if (valid) {
    const bool reorder_out = (MODE == FFT_Y_INV || MODE == FFT_Z_SOLVE) && dct_here;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int p = t + m * NT;
      const int row = reorder_out ? dct_row(p, N) : p;
      cw[base + (size_t)row * se] = x[m];
    }
  }
}

// ---------------------------------------------------------------- setup
static void wave_1d(const DirCoef &D, int bc, std::vector<double> &k2, std::vector<double> &tf) {
  // poisson.f90:1566-1596 (bc=0) and :1597-1612 (bc=1); transfer function of
  // the interpolation at w = es*d (poisson.f90:1716-1731)
  const double twopi = 2.0 * std::acos(-1.0);
  const Sten<double> &c = D.s;
  const double d = D.d, L = D.len;
  int n;
  std::vector<double> es;
  if (bc == 0) {
    n = D.n;
    k2.assign(n, 0.0);
    es.assign(n, 0.0);
    for (int i = 0; i <= n / 2; ++i) {
      const double w = twopi * i / n;
      double wp = c.aci * 2.0 * d * std::sin(w * 0.5) + c.bci * 2.0 * d * std::sin(1.5 * w);
      wp = wp / (1.0 + 2.0 * c.alcai * std::cos(w));
      es[i] = n * w / L;
      const double kk = n * wp / L;
      k2[i] = kk * kk;
    }
    for (int i = n / 2 + 1; i < n; ++i) {
      k2[i] = k2[n - i];
      es[i] = es[n - i];
    }
  } else {
    n = D.nm;
    k2.assign(n, 0.0);
    es.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
      const double w = twopi * 0.5 * i / D.nm;
      double wp = c.aci * 2.0 * d * std::sin(w * 0.5) + (c.bci * 2.0 * d) * std::sin(1.5 * w);
      wp = wp / (1.0 + 2.0 * c.alcai * std::cos(w));
      es[i] = D.nm * w / L;
      const double kk = D.nm * wp / L;
      k2[i] = kk * kk;
    }
    k2[0] = 0.0;
    es[0] = 0.0;
  }
  tf.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    const double r = es[i] * d;
    const double tt = 2.0 * (c.bici * std::cos(r * 1.5) + c.cici * std::cos(r * 2.5) + c.dici * std::cos(r * 3.5));
    const double tt1 = 2.0 * c.aici * std::cos(r * 0.5);
    const double t1 = 1.0 + 2.0 * c.ailcai * std::cos(r);
    tf[i] = (tt1 + tt) / t1;
  }
}

// k^2 and the transfer function at mode m of a non-periodic direction
// (w = pi m / nm), the formulas of wave_1d's bc=1 branch, for m = nm.
static void wave_1d_mode(const DirCoef &D, int m, std::vector<double> &k2, std::vector<double> &tf) {
  const double pi = std::acos(-1.0);
  const Sten<double> &c = D.s;
  const double d = D.d, L = D.len;
  const double w = pi * m / D.nm;
  double wp = c.aci * 2.0 * d * std::sin(w * 0.5) + (c.bci * 2.0 * d) * std::sin(1.5 * w);
  wp = wp / (1.0 + 2.0 * c.alcai * std::cos(w));
  const double kk = D.nm * wp / L, es = D.nm * w / L;
  const double r = es * d;
  const double tt = 2.0 * (c.bici * std::cos(r * 1.5) + c.cici * std::cos(r * 2.5) + c.dici * std::cos(r * 3.5));
  const double tt1 = 2.0 * c.aici * std::cos(r * 0.5);
  const double t1 = 1.0 + 2.0 * c.ailcai * std::cos(r);
  k2.assign(1, kk * kk);
  tf.assign(1, (tt1 + tt) / t1);
}

template <typename T>
static void append_cplx(std::vector<unsigned char> &blob, const std::vector<double> &re,
                        const std::vector<double> &im, size_t *off) {
  while (blob.size() % 16) blob.push_back(0);
  *off = blob.size();
  for (size_t i = 0; i < re.size(); ++i) {
    T v[2] = {(T)re[i], (T)im[i]};
    const unsigned char *p = reinterpret_cast<const unsigned char *>(v);
    blob.insert(blob.end(), p, p + sizeof(v));
  }
}
static void append_real(std::vector<unsigned char> &blob, const std::vector<double> &v, size_t *off) {
  while (blob.size() % 16) blob.push_back(0);
  *off = blob.size();
  const unsigned char *p = reinterpret_cast<const unsigned char *>(v.data());
  blob.insert(blob.end(), p, p + v.size() * sizeof(double));
}

template <typename T>
static int build_tables(const PoissonPlan &P, const std::vector<double> *tabs, void **dev,
                        PoissonTables<T> *out) {
  // tabs: kx2, tx, ky2, ty, kz2, tz
  const double pi = std::acos(-1.0);
  auto twiddle = [&](int N, std::vector<double> &re, std::vector<double> &im) {
    re.resize(N);
    im.resize(N);
    for (int m = 0; m < N; ++m) {
      re[m] = std::cos(2.0 * pi * m / N);
      im[m] = -std::sin(2.0 * pi * m / N);
    }
  };
  std::vector<unsigned char> blob;
  size_t o_twx, o_twy, o_twz, o_rtx, o_dct, o_r[6];
  std::vector<double> re, im;
  twiddle(P.nx / 2, re, im);
  append_cplx<T>(blob, re, im, &o_twx);
  twiddle(P.ny, re, im);
  append_cplx<T>(blob, re, im, &o_twy);
  twiddle(P.nz, re, im);
  append_cplx<T>(blob, re, im, &o_twz);
  re.assign(P.nxh, 0.0);
  im.assign(P.nxh, 0.0);
  for (int k = 0; k < P.nxh; ++k) {
    re[k] = std::cos(2.0 * pi * k / P.nx);
    im[k] = -std::sin(2.0 * pi * k / P.nx);
  }
  append_cplx<T>(blob, re, im, &o_rtx);
  // DCT twiddles e^{-i pi k/(2N)}: the cos/sin pair of abxyz bc=1 (poisson.f90:1502-1508)
  re.assign(P.ny, 0.0);
  im.assign(P.ny, 0.0);
  for (int k = 0; k < P.ny; ++k) {
    re[k] = std::cos((double)k * pi * 0.5 / (double)P.ny);
    im[k] = -std::sin((double)k * pi * 0.5 / (double)P.ny);
  }
  append_cplx<T>(blob, re, im, &o_dct);
  size_t o_dctx, o_dctz;
  re.assign(P.nx / 2 + 1, 0.0);
  im.assign(P.nx / 2 + 1, 0.0);
  for (int k = 0; k <= P.nx / 2; ++k) {
    re[k] = std::cos((double)k * pi * 0.5 / (double)P.nx);
    im[k] = -std::sin((double)k * pi * 0.5 / (double)P.nx);
  }
  append_cplx<T>(blob, re, im, &o_dctx);
  re.assign(P.nz, 0.0);
  im.assign(P.nz, 0.0);
  for (int k = 0; k < P.nz; ++k) {
    re[k] = std::cos((double)k * pi * 0.5 / (double)P.nz);
    im[k] = -std::sin((double)k * pi * 0.5 / (double)P.nz);
  }
  append_cplx<T>(blob, re, im, &o_dctz);
  for (int i = 0; i < 6; ++i) append_real(blob, tabs[i], &o_r[i]);
  void *d = nullptr;
  X3D_HIP(hipMalloc(&d, blob.size()));
  X3D_HIP(hipMemcpy(d, blob.data(), blob.size(), hipMemcpyHostToDevice));
  *dev = d;
  unsigned char *b = static_cast<unsigned char *>(d);
  out->twx = reinterpret_cast<const cplx<T> *>(b + o_twx);
  out->twy = reinterpret_cast<const cplx<T> *>(b + o_twy);
  out->twz = reinterpret_cast<const cplx<T> *>(b + o_twz);
  out->rtx = reinterpret_cast<const cplx<T> *>(b + o_rtx);
  out->dcty = reinterpret_cast<const cplx<T> *>(b + o_dct);
  out->dctx = reinterpret_cast<const cplx<T> *>(b + o_dctx);
  out->dctz = reinterpret_cast<const cplx<T> *>(b + o_dctz);
  out->kx2 = reinterpret_cast<const double *>(b + o_r[0]);
  out->tx = reinterpret_cast<const double *>(b + o_r[1]);
  out->ky2 = reinterpret_cast<const double *>(b + o_r[2]);
  out->ty = reinterpret_cast<const double *>(b + o_r[3]);
  out->kz2 = reinterpret_cast<const double *>(b + o_r[4]);
  out->tz = reinterpret_cast<const double *>(b + o_r[5]);
  return X3D_OK;
}

// ---------------------------------------------------------------- stretched y
// poisson_010 with istret != 0 (poisson.f90:846-895): after the x, y
// (cosine) and z transforms, the even cosine modes 0, 2, 4, ... and the odd
// modes 1, 3, 5, ... of each (kx, kz) line solve the pentadiagonal systems of
// matrice_refinement (poisson.f90:1814-2038) with inversion5_v1
// (tools.f90:1225-1362); with istret = 3 all modes solve one system
// (poisson.f90:2126-2247, inversion5_v2 at tools.f90:1365-1497).  Their matrices are real (the reference's real and
// imaginary parts are equal for periodic x and z) and do not depend on the
// right-hand side, so the elimination runs once here, on the host in FP64, in
// the reference's order, and the kernel applies the stored factors:
//   F1[m], F2[m]  multipliers of rows m+1, m+2 by row m (m < nh-2)
//   RP[m], A1[m], B1[m]  back substitution e_m = e_m RP - A1 e_{m+1} - B1 e_{m+2}
//   last row n = nh-1: RP[n] = b1 (0 when |b1| <= 1e-16), A1[n] = sr/b1,
//   e_n = e_n/b1 - A1 e_{n-1}
// A pivot of magnitude <= 1e-16 gives a zero factor, as the reference's
// guarded divisions; an exactly zero pivot during the elimination (a line
// whose matrix is zero) gives a zero multiplier.
static void build_penta_factors(const Context &c, const PoissonPlan &P, const std::vector<double> &kx2,
                                const std::vector<double> &tx, const std::vector<double> &ty,
                                const std::vector<double> &kz2, const std::vector<double> &tz,
                                std::vector<double> &out) {
  const DirCoef &Y = c.dir[1];
  const double pi = std::acos(-1.0);
  // istret = 3: one system over all nym modes (poisson.f90:2126-2247,
  // inversion5_v2), else the even and odd modes apart
  const bool one = c.p.istret == 3;
  const int nsys = one ? 1 : 2;
  const int nym = P.ny, nh = nym / nsys, nxh = P.nxh, nz = P.nz;
  const double eps = 1.0e-16;
  // yky = nym wp without 1/yly on a stretched mesh (poisson.f90:1614)
  std::vector<double> yky(nym, 0.0);
  for (int j = 1; j < nym; ++j) {
    const double w = 2.0 * pi * 0.5 * j / nym;
    double wp = Y.s.aci * 2.0 * Y.d * std::sin(w * 0.5) + (Y.s.bci * 2.0 * Y.d) * std::sin(1.5 * w);
    wp = wp / (1.0 + 2.0 * Y.s.alcai * std::cos(w));
    yky[j] = nym * wp;
  }
  const double alpha = c.alpha, beta = c.p.beta;
  const double xa0 = alpha / pi + 0.5 / beta / pi;
  const double xa1 = c.p.istret == 1 ? 1.0 / 4.0 / beta / pi : -(1.0 / 4.0 / beta / pi);   // 2 and 3: -
  const double xa0_2 = xa0 * xa0, xa1_2 = xa1 * xa1, xa01 = xa0 * xa1, xa0p1_2 = (xa0 + xa1) * (xa0 + xa1);
  // the z tables are symmetric (wave_1d copies mode nz-kz from mode kz), so
  // the matrices of kz and nz-kz are equal: planes kz = 0 .. nz/2 are built
  // and the kernel reads plane min(kz, nz-kz)
  const int nzp = nz / 2 + 1;
  const size_t stride = (size_t)nzp * nsys * nh * nxh;
  out.assign(5 * stride, 0.0);
  std::vector<double> cw(nh), a(5 * nh);
  for (int kz = 0; kz < nzp; ++kz)
    for (int par = 0; par < nsys; ++par)
      for (int kx = 0; kx < nxh; ++kx) {
        const double TX = tx[kx], TZ = tz[kz], TX2 = TX * TX, TZ2 = TZ * TZ, XK2 = kx2[kx], ZK2 = kz2[kz];
        for (int j = 0; j < nh; ++j) cw[j] = TX * (yky[nsys * j + par] * TZ);
        auto A = [&](int j, int d) -> double & { return a[5 * j + d]; };
        std::fill(a.begin(), a.end(), 0.0);
        auto kxz = [&](int mode) { const double t2 = ty[mode] * ty[mode]; return XK2 * t2 * TZ2 + ZK2 * t2 * TX2; };
        const int n = nh - 1;
        if (one) {
          // istret = 3 (poisson.f90:2148-2233): the interior formulas on
          // every row, the first row's off-diagonals without the factor 2
          // of the even system of istret 1, 2
          for (int j = 1; j < n; ++j)
            A(j, 2) = -(kxz(j) + xa0_2 * (cw[j] * cw[j]) + xa1_2 * cw[j] * (cw[j - 1] + cw[j + 1]));
          A(0, 2) = -(kxz(0) + xa0_2 * (cw[0] * cw[0]) + xa1_2 * cw[0] * cw[1]);
          A(n, 2) = -(kxz(n) + xa0_2 * (cw[n] * cw[n]) + xa1_2 * cw[n] * cw[n - 1]);
          for (int j = 0; j < n; ++j) A(j, 3) = xa01 * (cw[j + 1] * (cw[j] + cw[j + 1]));
          for (int j = 0; j + 2 < nh; ++j) A(j, 4) = -xa1_2 * (cw[j + 1] * cw[j + 2]);
          for (int j = 1; j < nh; ++j) A(j, 1) = xa01 * (cw[j - 1] * (cw[j] + cw[j - 1]));
          for (int j = 2; j < nh; ++j) A(j, 0) = -xa1_2 * (cw[j - 1] * cw[j - 2]);
          if (XK2 == 0.0 && ZK2 == 0.0) {
            A(0, 2) = 1.0;
            A(0, 3) = 0.0;
            A(0, 4) = 0.0;
          }
        } else {
          // main diagonal
          for (int j = 1; j < n; ++j)
            A(j, 2) = -(kxz(2 * j + par) + xa0_2 * (cw[j] * cw[j]) + xa1_2 * cw[j] * (cw[j - 1] + cw[j + 1]));
          if (par == 0) {
            A(0, 2) = -(kxz(0) + xa0_2 * (cw[0] * cw[0]) + xa1_2 * cw[0] * cw[1]);
            A(n, 2) = -(kxz(2 * n) + xa0_2 * (cw[n] * cw[n]) + xa1_2 * cw[n] * cw[n - 1]);
          } else {
            A(0, 2) = -(kxz(1) + (xa0_2 - xa1_2) * (cw[0] * cw[0]) + xa1_2 * cw[0] * cw[1]);
            A(n, 2) = -(kxz(2 * n + 1) + xa0p1_2 * (cw[n] * cw[n]) + xa1_2 * cw[n] * cw[n - 1]);
          }
          // +1
          for (int j = 1; j < n; ++j) A(j, 3) = xa01 * (cw[j + 1] * (cw[j] + cw[j + 1]));
          if (par == 0) {
            A(0, 3) = 2.0 * xa01 * (cw[0] * cw[1] + cw[1] * cw[1]);
          } else {
            A(0, 3) = (xa0 - xa1) * xa1 * (cw[0] * cw[1]) + xa0 * xa1 * (cw[1] * cw[1]);
            A(n - 1, 3) = xa0 * xa1 * cw[n - 1] * cw[n] + (xa0 + xa1) * xa1 * (cw[n] * cw[n]);
            A(n, 3) = 0.0;
          }
          // +2
          for (int j = 0; j < nh - 2; ++j) A(j, 4) = xa1_2 * (-cw[j + 1] * cw[j + 2]);
          if (par == 0) A(0, 4) = 2.0 * A(0, 4);
          A(n - 1, 4) = 0.0;
          A(n, 4) = 0.0;
          // -1
          for (int j = 1; j < nh; ++j) A(j, 1) = xa01 * (cw[j - 1] * (cw[j] + cw[j - 1]));
          A(0, 1) = 0.0;
          if (par == 1) {
            A(1, 1) = xa0 * xa1 * (cw[1] * cw[0]) + (xa0 + xa1) * xa1 * (cw[0] * cw[0]);
            A(n, 1) = (xa0 + xa1) * xa1 * (cw[n] * cw[n - 1]) + xa0 * xa1 * (cw[n - 1] * cw[n - 1]);
          }
          // -2
          for (int j = 2; j < nh; ++j) A(j, 0) = xa1_2 * (-cw[j - 1] * cw[j - 2]);
          // the mean mode: unit first row of the even system
          if (par == 0 && XK2 == 0.0 && ZK2 == 0.0) {
            A(0, 2) = 1.0;
            A(0, 3) = 0.0;
            A(0, 4) = 0.0;
          }
        }
        // inversion5_v1's elimination, recording the factors
        const size_t line = ((size_t)(kz * nsys + par) * nh) * nxh + kx;
        auto F = [&](int q, int m) -> double & { return out[q * stride + line + (size_t)m * nxh]; };
        for (int m = 0; m + 2 < nh; ++m) {
          const double piv = A(m, 2);
          const double s1 = piv != 0.0 ? A(m + 1, 1) / piv : 0.0;
          A(m + 1, 2) = A(m + 1, 2) - s1 * A(m, 3);
          A(m + 1, 3) = A(m + 1, 3) - s1 * A(m, 4);
          const double s2 = piv != 0.0 ? A(m + 2, 0) / piv : 0.0;
          A(m + 2, 1) = A(m + 2, 1) - s2 * A(m, 3);
          A(m + 2, 2) = A(m + 2, 2) - s2 * A(m, 4);
          F(0, m) = s1;
          F(1, m) = s2;
        }
        {
          const double sr = std::fabs(A(n - 1, 2)) > eps ? A(n, 1) / A(n - 1, 2) : 0.0;
          const double b1 = A(n, 2) - sr * A(n - 1, 3);
          const bool ok = std::fabs(b1) > eps;
          F(2, n) = ok ? b1 : 0.0;
          F(3, n) = ok ? sr / b1 : 0.0;
          const double rp = std::fabs(A(n - 1, 2)) > eps ? 1.0 / A(n - 1, 2) : 0.0;
          F(2, n - 1) = rp;
          F(3, n - 1) = A(n - 1, 3) * rp;
        }
        for (int i = nh - 3; i >= 0; --i) {
          const double rp = std::fabs(A(i, 2)) > eps ? 1.0 / A(i, 2) : 0.0;
          F(2, i) = rp;
          F(3, i) = A(i, 3) * rp;
          F(4, i) = A(i, 4) * rp;
        }
      }
}

// One thread per (kx, kz, system): forward elimination with the stored
// multipliers (the scaled right-hand side read once, the eliminated values
// written back), then the last 2x2 block and the back substitution.  The
// (nx/2+1, :, nz/2+1) line is zeroed (poisson.f90:902-908).
template <typename T>
__global__ void __launch_bounds__(64) penta_y_kernel(cplx<T> *__restrict__ cw, const double *__restrict__ pf,
                                                     size_t stride, int nxh, int ny, int nz, int nsys,
                                                     double scale) {
  const int kx = blockIdx.x * blockDim.x + threadIdx.x;
  const int kz = blockIdx.y / nsys, par = blockIdx.y - kz * nsys;
  if (kx >= nxh) return;
  const int nh = ny / nsys, n = nh - 1;
  const int kzp = kz <= nz - kz ? kz : nz - kz;   // factor plane: kz and nz-kz share one
  const size_t line = ((size_t)(kzp * nsys + par) * nh) * nxh + kx;
  const double *F1 = pf + line, *F2 = pf + stride + line, *RP = pf + 2 * stride + line;
  const double *A1 = pf + 3 * stride + line, *B1 = pf + 4 * stride + line;
  auto at = [&](int m) -> cplx<T> & { return cw[((size_t)kz * ny + (nsys * m + par)) * nxh + kx]; };
  if (kx == nxh - 1 && kz == (nz >> 1)) {
    for (int m = 0; m < nh; ++m) at(m) = cplx<T>{T(0), T(0)};
    return;
  }
  auto ld = [&](int m) {
    const cplx<T> v = at(m);
    return cplx<double>{(double)v.x * scale, (double)v.y * scale};
  };
  auto st = [&](int m, cplx<double> v) { at(m) = cplx<T>{(T)v.x, (T)v.y}; };
  // forward: window (e0, e1, e2) = rows m, m+1, m+2
  cplx<double> e0 = ld(0), e1 = ld(1), e2;
  // unrolled so that the loads of later rows issue ahead of the chain
#pragma unroll 4
  for (int m = 0; m + 2 < nh; ++m) {
    e2 = ld(m + 2);
    const double s1 = F1[(size_t)m * nxh], s2 = F2[(size_t)m * nxh];
    e1 = cplx<double>{e1.x - s1 * e0.x, e1.y - s1 * e0.y};
    e2 = cplx<double>{e2.x - s2 * e0.x, e2.y - s2 * e0.y};
    st(m, e0);
    e0 = e1;
    e1 = e2;
  }
  // e0, e1 = rows n-1, n
  const double b1 = RP[(size_t)n * nxh], t1 = A1[(size_t)n * nxh];
  cplx<double> xn = b1 != 0.0 ? cplx<double>{e1.x / b1 - t1 * e0.x, e1.y / b1 - t1 * e0.y} : cplx<double>{0.0, 0.0};
  const double rp = RP[(size_t)(n - 1) * nxh], a1 = A1[(size_t)(n - 1) * nxh];
  cplx<double> xm = {e0.x * rp - a1 * xn.x, e0.y * rp - a1 * xn.y};
  st(n, xn);
  st(n - 1, xm);
  // back substitution: (xm, xn) = rows i+1, i+2
#pragma unroll 4
  for (int i = nh - 3; i >= 0; --i) {
    const cplx<T> v = at(i);
    const double r = RP[(size_t)i * nxh], a = A1[(size_t)i * nxh], bb = B1[(size_t)i * nxh];
    const cplx<double> x = {((double)v.x * r - a * xm.x) - bb * xn.x, ((double)v.y * r - a * xm.y) - bb * xn.y};
    st(i, x);
    xn = xm;
    xm = x;
  }
}

// ---------------------------------------------------------------- poisson_100
// x non-periodic, y and z periodic (poisson.f90:413-661).  After the x
// cosine transform the real part of entry kx is cosine mode kx and the
// imaginary part mode nx-kx; the complex FFTs along y and z then mix the two.
// Both parts are transforms of real fields, so they are the Hermitian and
// anti-Hermitian halves over (ky, kz) -> (-ky, -kz): A = (D + conj D')/2,
// iB = (D - conj D')/2, divided by -K of mode kx and of mode nx-kx (FP64,
// zero where K < 1e-16).  One thread per pair of points, in place; the
// partner's results are conj(A) and -conj(iB) with the same K (the tables
// are symmetric in ky and kz).
template <typename T>
__global__ void __launch_bounds__(256) split_divide_kernel(cplx<T> *__restrict__ cw, int nxh, int ny, int nz,
                                                           double scale, PoissonTables<T> tb) {
  const size_t total = (size_t)nxh * ny * nz;
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= total) return;
  const int a = (int)(q % nxh);
  const size_t r = q / nxh;
  const int ky = (int)(r % ny), kz = (int)(r / ny);
  const int kyp = ky ? ny - ky : 0, kzp = kz ? nz - kz : 0;
  const size_t qp = ((size_t)kzp * ny + kyp) * nxh + a;
  if (qp < q) return;   // the partner's thread writes both
  const cplx<T> d = cw[q], dp = cw[qp];
  const int am = 2 * (nxh - 1) - a;
  const double tyv = tb.ty[ky], tzv = tb.tz[kz], ky2 = tb.ky2[ky], kz2 = tb.kz2[kz];
  const double txr = tb.tx[a], kx2r = tb.kx2[a], txi = tb.tx[am], kx2i = tb.kx2[am];
  const double yz2 = (tyv * tzv) * (tyv * tzv);
  const double Kr = kx2r * yz2 + ky2 * (txr * tzv) * (txr * tzv) + kz2 * (txr * tyv) * (txr * tyv);
  const double Ki = kx2i * yz2 + ky2 * (txi * tzv) * (txi * tzv) + kz2 * (txi * tyv) * (txi * tyv);
  const double fr = Kr < 1.0e-16 ? 0.0 : -scale / Kr;
  const double fi = Ki < 1.0e-16 ? 0.0 : -scale / Ki;
  const double ar = 0.5 * ((double)d.x + (double)dp.x), ai = 0.5 * ((double)d.y - (double)dp.y);
  const double br = 0.5 * ((double)d.x - (double)dp.x), bi = 0.5 * ((double)d.y + (double)dp.y);
  cw[q] = cplx<T>{(T)(ar * fr + br * fi), (T)(ai * fr + bi * fi)};
  if (qp != q) cw[qp] = cplx<T>{(T)(ar * fr - br * fi), (T)(-ai * fr + bi * fi)};
}

// ---------------------------------------------------------------- scratch without allocation
// The spectral work array of x3d_poisson_async_noalloc: a file-scope device
// array of POISSON_SCRATCH_BYTES that every GPU the library runs on holds
// until the process ends.  Its device address is taken once per device at
// x3d_init (with that device current) and kept in g_scratch; host code uses
// only that address, never the array's name.  Every call on one GPU shares
// the array: calls that overlap on one GPU race.
constexpr int MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[POISSON_SCRATCH_BYTES];
static void *g_scratch[MAX_DEVICES] = {};

static int scratch_init(int device) {
  if (device < 0 || device >= MAX_DEVICES) return fail(X3D_ERR_UNSUPPORTED, "device ordinal beyond the scratch table");
  if (g_scratch[device]) return X3D_OK;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(X3D_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[device] = ptr;
  return X3D_OK;
}

void *poisson_scratch(int device) {
  return device >= 0 && device < MAX_DEVICES ? g_scratch[device] : nullptr;
}

int poisson_setup(Context &c) {
  X3D_TRY(scratch_init(c.device));
  const DirCoef &X = c.dir[0], &Y = c.dir[1], &Z = c.dir[2];
  // dispatch of decomp_2d_poisson_init (poisson.f90:98-108): 000, 100 (x
  // non-periodic, y and z periodic), 010 and 11x (x and y non-periodic, z
  // either)
  if (X.bc1 != 0 && Y.bc1 == 0 && Z.bc1 != 0)
    return fail(X3D_ERR_UNSUPPORTED, "Poisson: x and z non-periodic with y periodic is not a reference case");
  if (X.bc1 == 0 && Z.bc1 != 0)
    return fail(X3D_ERR_UNSUPPORTED, "Poisson: z non-periodic needs x and y non-periodic (poisson_11x)");
  auto *P = new PoissonPlan();
  P->xdct = X.bc1 != 0;
  P->ydct = Y.bc1 != 0;
  P->zdct = Z.bc1 != 0;
  P->nx = P->xdct ? X.nm : X.n;
  P->ny = Y.nm;
  P->nz = P->zdct ? Z.nm : Z.n;
  P->nxh = P->nx / 2 + 1;
  P->scale = 1.0 / ((double)P->nx * (double)P->ny * (double)P->nz);
  const int lx = ilog2(P->nx), ly = ilog2(P->ny), lz = ilog2(P->nz);
  if (lx < 4 || ly < 3 || lz < 3 || P->nx > 2048 || P->ny > 1024 || P->nz > 1024) {
    delete P;
    return fail(X3D_ERR_UNSUPPORTED, "Poisson: pressure-grid sizes must be powers of two (nx 16..2048, ny/nz 8..1024)");
  }
  std::vector<double> tabs[6];
  wave_1d(X, P->xdct ? 1 : 0, tabs[0], tabs[1]);
  wave_1d(Y, P->ydct ? 1 : 0, tabs[2], tabs[3]);
  wave_1d(Z, P->zdct ? 1 : 0, tabs[4], tabs[5]);
  if (P->xdct) {
    // the imaginary part of kx carries mode nx-kx: modes 0..nx (poisson.f90:1646-1655)
    std::vector<double> k2e, tfe;
    wave_1d_mode(X, P->nx, k2e, tfe);
    tabs[0].push_back(k2e[0]);
    tabs[1].push_back(tfe[0]);
  } else {
    tabs[0].resize(P->nxh);
    tabs[1].resize(P->nxh);
  }
  if (c.p.istret != 0) {
    // stretched y: poisson_010's pentadiagonal branch; istret = 3 solves one
    // system over all nym modes (inversion5_v2), istret = 1, 2 the even and
    // odd modes apart (inversion5_v1)
    if (P->xdct || P->zdct || !P->ydct) {
      delete P;
      return fail(X3D_ERR_UNSUPPORTED, "Poisson on a stretched y mesh: poisson_010 (x and z periodic) only");
    }
    std::vector<double> f;
    build_penta_factors(c, *P, tabs[0], tabs[1], tabs[3], tabs[4], tabs[5], f);
    P->ystret = c.p.istret == 3 ? 1 : 2;   // number of systems per (kx, kz)
    P->pf_stride = f.size() / 5;
    hipError_t e = hipMalloc(&P->pf, f.size() * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(P->pf, f.data(), f.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      if (P->pf) (void)hipFree(P->pf);
      delete P;
      return fail(X3D_ERR_HIP, std::string("Poisson factors: ") + hipGetErrorString(e));
    }
  }
  int r = build_tables<double>(*P, tabs, &P->tab64, &P->t64);
  if (r == X3D_OK) r = build_tables<float>(*P, tabs, &P->tab32, &P->t32);
  if (r != X3D_OK) {
    if (P->pf) (void)hipFree(P->pf);
    delete P;
    return r;
  }
  c.poisson = P;
  return X3D_OK;
}

void poisson_free(Context &c) {
  auto *P = static_cast<PoissonPlan *>(c.poisson);
  if (!P) return;
  if (P->tab64) (void)hipFree(P->tab64);
  if (P->tab32) (void)hipFree(P->tab32);
  if (P->cw64) (void)hipFree(P->cw64);
  if (P->cw32) (void)hipFree(P->cw32);
  if (P->pf) (void)hipFree(P->pf);
  delete P;
  c.poisson = nullptr;
}

template <typename T, int M>
static void launch_x(bool fwd, T *rhs, cplx<T> *cw, int nlines, int xdct, const PoissonTables<T> &tb,
                     hipStream_t st) {
  using S = XShape<M>;
  const dim3 grid((nlines + S::TL - 1) / S::TL), block(S::TL * S::NT);
  const size_t sh = S::template lds<T>();
  if (fwd) hipLaunchKernelGGL((r2c_x_kernel<T, M>), grid, block, sh, st, rhs, cw, nlines, xdct, tb);
  else hipLaunchKernelGGL((c2r_x_kernel<T, M>), grid, block, sh, st, cw, rhs, nlines, xdct, tb);
}

template <typename T, int N, int MODE, bool COS, bool EXTRA>
static void launch_s1(const PoissonPlan &P, cplx<T> *cw, const PoissonTables<T> &tb, int nmain, int ncols,
                      hipStream_t st) {
  using S = SShape<T, N>;
  const dim3 grid(ncols, ZPass<MODE>::value ? P.ny : P.nz), block(S::threads(EXTRA));
  // poisson_010 zeroes the (nx/2+1, :, nz/2+1) line (poisson.f90:902-908); poisson_11x does not
  const int zero_line = MODE == FFT_Z_SOLVE && P.ydct && !P.xdct;
  const bool zt = MODE == FFT_Z_SOLVE && !COS && S::ZTAB;
  hipLaunchKernelGGL((strided_kernel<T, N, MODE, COS, EXTRA>), grid, block, S::lds(EXTRA, zt), st, cw, nmain, P.nxh,
                     P.ny, P.nz, P.ydct, P.scale, zero_line, P.xdct, P.zdct, tb);
}

template <typename T, int N, int MODE>
static void launch_s(const PoissonPlan &P, cplx<T> *cw, const PoissonTables<T> &tb, hipStream_t st) {
  using S = SShape<T, N>;
  const int ncols = P.nxh / S::TL, nmain = ncols * S::TL, rem = P.nxh - nmain;
  if constexpr (MODE == FFT_Z_SOLVE) {
    // the cosine-x variant keeps whole blocks plus a partial last column
    if (P.xdct) return launch_s1<T, N, MODE, true, false>(P, cw, tb, nmain, (P.nxh + S::TL - 1) / S::TL, st);
  }
  // FP32: one extra line per block when the blocks outnumber the remaining
  // lines (measured at 256^3: 34/34/52 us against 43/43/63 for the y, y and z
  // passes with a partial last column; FP64 measured slower with it, 65/66/89
  // against 64/60/70 us)
  if (sizeof(T) == 4 && rem > 0 && ncols >= rem) launch_s1<T, N, MODE, false, true>(P, cw, tb, nmain, ncols, st);
  else launch_s1<T, N, MODE, false, false>(P, cw, tb, nmain, (P.nxh + S::TL - 1) / S::TL, st);
}

template <typename T, int MODE>
static bool dispatch_s(int N, const PoissonPlan &P, cplx<T> *cw, const PoissonTables<T> &tb, hipStream_t st) {
  switch (N) {
    case 8: launch_s<T, 8, MODE>(P, cw, tb, st); return true;
    case 16: launch_s<T, 16, MODE>(P, cw, tb, st); return true;
    case 32: launch_s<T, 32, MODE>(P, cw, tb, st); return true;
    case 64: launch_s<T, 64, MODE>(P, cw, tb, st); return true;
    case 128: launch_s<T, 128, MODE>(P, cw, tb, st); return true;
    case 256: launch_s<T, 256, MODE>(P, cw, tb, st); return true;
    case 512: launch_s<T, 512, MODE>(P, cw, tb, st); return true;
    case 1024: launch_s<T, 1024, MODE>(P, cw, tb, st); return true;
    default: return false;
  }
}

template <typename T>
static bool dispatch_x(bool fwd, int M, T *rhs, cplx<T> *cw, int nlines, int xdct, const PoissonTables<T> &tb,
                       hipStream_t st) {
  switch (M) {
    case 8: launch_x<T, 8>(fwd, rhs, cw, nlines, xdct, tb, st); return true;
    case 16: launch_x<T, 16>(fwd, rhs, cw, nlines, xdct, tb, st); return true;
    case 32: launch_x<T, 32>(fwd, rhs, cw, nlines, xdct, tb, st); return true;
    case 64: launch_x<T, 64>(fwd, rhs, cw, nlines, xdct, tb, st); return true;
    case 128: launch_x<T, 128>(fwd, rhs, cw, nlines, xdct, tb, st); return true;
    case 256: launch_x<T, 256>(fwd, rhs, cw, nlines, xdct, tb, st); return true;
    case 512: launch_x<T, 512>(fwd, rhs, cw, nlines, xdct, tb, st); return true;
    case 1024: launch_x<T, 1024>(fwd, rhs, cw, nlines, xdct, tb, st); return true;
    default: return false;
  }
}

template <typename T>
size_t poisson_work_bytes(const Context &c) {
  const auto *P = static_cast<const PoissonPlan *>(c.poisson);
  return P ? (size_t)P->nxh * P->ny * P->nz * sizeof(cplx<T>) : 0;
}
template size_t poisson_work_bytes<double>(const Context &);
template size_t poisson_work_bytes<float>(const Context &);

// The five (stretched: seven) passes of the solve on the spectral work array
// cw of poisson_work_bytes<T>(c) bytes; enqueues kernels only.
template <typename T>
int poisson_run(const Context &c, T *rhs, void *work, hipStream_t st) {
  const auto *P = static_cast<const PoissonPlan *>(c.poisson);
  if (!P) return fail(X3D_ERR_STATE, "Poisson solver not initialised");
  cplx<T> *cw = static_cast<cplx<T> *>(work);
  PoissonTables<T> tb;
  if constexpr (sizeof(T) == 8) tb = P->t64;
  else tb = P->t32;
  const int nlines = P->ny * P->nz, M = P->nx / 2;
  // sizes were checked by poisson_setup: every dispatch below finds its case
  bool ok = dispatch_x<T>(true, M, rhs, cw, nlines, P->xdct, tb, st);
  ok = ok && dispatch_s<T, FFT_Y_FWD>(P->ny, *P, cw, tb, st);
  if (P->xdct && !P->ydct) {
    // poisson_100: forward z, the pairwise split and divide, inverse z
    ok = ok && dispatch_s<T, FFT_Z_FWD>(P->nz, *P, cw, tb, st);
    if (ok) {
      const size_t total = (size_t)P->nxh * P->ny * P->nz;
      hipLaunchKernelGGL((split_divide_kernel<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, cw,
                         P->nxh, P->ny, P->nz, P->scale, tb);
    }
    ok = ok && dispatch_s<T, FFT_Z_INV>(P->nz, *P, cw, tb, st);
  } else if (P->ystret) {
    // stretched y: z transform, pentadiagonal solves along y, inverse z
    ok = ok && dispatch_s<T, FFT_Z_FWD>(P->nz, *P, cw, tb, st);
    if (ok) {
      const dim3 grid((P->nxh + 63) / 64, P->ystret * P->nz), block(64);
      hipLaunchKernelGGL((penta_y_kernel<T>), grid, block, 0, st, cw, P->pf, P->pf_stride, P->nxh, P->ny, P->nz,
                         P->ystret, P->scale);
    }
    ok = ok && dispatch_s<T, FFT_Z_INV>(P->nz, *P, cw, tb, st);
  } else {
    ok = ok && dispatch_s<T, FFT_Z_SOLVE>(P->nz, *P, cw, tb, st);
  }
  ok = ok && dispatch_s<T, FFT_Y_INV>(P->ny, *P, cw, tb, st);
  ok = ok && dispatch_x<T>(false, M, rhs, cw, nlines, P->xdct, tb, st);
  if (!ok) return fail(X3D_ERR_UNSUPPORTED, "Poisson: no FFT kernel for this size");
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(X3D_ERR_HIP, std::string("poisson launch: ") + hipGetErrorString(e));
  return X3D_OK;
}
template int poisson_run<double>(const Context &, double *, void *, hipStream_t);
template int poisson_run<float>(const Context &, float *, void *, hipStream_t);

// The library's own work array, allocated at the first solve of each precision.
template <typename T>
int poisson_solve(Context &c, T *rhs, hipStream_t st) {
  auto *P = static_cast<PoissonPlan *>(c.poisson);
  if (!P) return fail(X3D_ERR_STATE, "Poisson solver not initialised");
  void **cwp = sizeof(T) == 8 ? &P->cw64 : &P->cw32;
  if (!*cwp) X3D_HIP(hipMalloc(cwp, poisson_work_bytes<T>(c)));
  return poisson_run<T>(c, rhs, *cwp, st);
}

template int poisson_solve<double>(Context &, double *, hipStream_t);
template int poisson_solve<float>(Context &, float *, hipStream_t);

// ---------------------------------------------------------------- 3-D FFT (decomp_2d_fft_3d)
// The unnormalised transform 2DECOMP's decomp_2d_fft_3d performs with
// physical space in z-pencils (poisson.f90:325-330, 405): real (nx,ny,nz) to
// complex (nx, ny, nz/2+1), i fastest, exp(-i k.x) forward and exp(+i k.x)
// backward, on the pressure grid.  The passes are the Poisson solve's own
// (r2c along x, c2c along y and z) on its half-x work array (nx/2+1, ny, nz);
// one more pass maps between that layout and the half-z layout of the
// reference through the symmetry of a real field's spectrum,
// X(kx, ky, kz) = conj X(-kx, -ky, -kz).
template <typename T>
__global__ void __launch_bounds__(256) halfx_to_halfz(const cplx<T> *__restrict__ cw, cplx<T> *__restrict__ out,
                                                      int nx, int ny, int nz, int nxh) {
  const size_t total = (size_t)nx * ny * (nz / 2 + 1);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += stride) {
    const int kx = (int)(q % nx);
    const size_t r = q / nx;
    const int ky = (int)(r % ny), kz = (int)(r / ny);
    out[q] = kx < nxh ? cw[kx + (size_t)nxh * (ky + (size_t)ny * kz)]
                      : cconj(cw[(nx - kx) + (size_t)nxh * ((ny - ky) % ny + (size_t)ny * ((nz - kz) % nz))]);
  }
}

template <typename T>
__global__ void __launch_bounds__(256) halfz_to_halfx(const cplx<T> *__restrict__ in, cplx<T> *__restrict__ cw,
                                                      int nx, int ny, int nz, int nxh) {
  const int nzh = nz / 2 + 1;
  const size_t total = (size_t)nxh * ny * nz;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += stride) {
    const int kx = (int)(q % nxh);
    const size_t r = q / nxh;
    const int ky = (int)(r % ny), kz = (int)(r / ny);
    cw[q] = kz < nzh ? in[kx + (size_t)nx * (ky + (size_t)ny * kz)]
                     : cconj(in[(nx - kx) % nx + (size_t)nx * ((ny - ky) % ny + (size_t)ny * (nz - kz))]);
  }
}

template <typename T>
int fft3d(Context &c, bool forward, const void *in, void *out, hipStream_t st) {
  auto *P = static_cast<PoissonPlan *>(c.poisson);
  if (!P) return fail(X3D_ERR_STATE, "no FFT plan for this grid (see x3d_init)");
  void **cwp = sizeof(T) == 8 ? &P->cw64 : &P->cw32;
  if (!*cwp) X3D_HIP(hipMalloc(cwp, poisson_work_bytes<T>(c)));
  cplx<T> *cw = static_cast<cplx<T> *>(*cwp);
  // the plain transforms: no cosine pre/post-processing, whatever the
  // boundaries the Poisson plan was built for
  PoissonPlan Q = *P;
  Q.xdct = Q.ydct = Q.zdct = 0;
  Q.ystret = 0;
  PoissonTables<T> tb;
  if constexpr (sizeof(T) == 8) tb = P->t64;
  else tb = P->t32;
  const int nlines = Q.ny * Q.nz, M = Q.nx / 2;
  const size_t nout = forward ? (size_t)Q.nx * Q.ny * (Q.nz / 2 + 1) : (size_t)Q.nxh * Q.ny * Q.nz;
  const unsigned nb = (unsigned)std::min<size_t>((nout + 255) / 256, 65535u * 8u);
  bool ok;
  if (forward) {
    ok = dispatch_x<T>(true, M, const_cast<T *>(static_cast<const T *>(in)), cw, nlines, 0, tb, st);
    ok = ok && dispatch_s<T, FFT_Y_FWD>(Q.ny, Q, cw, tb, st);
    ok = ok && dispatch_s<T, FFT_Z_FWD>(Q.nz, Q, cw, tb, st);
    if (ok)
      hipLaunchKernelGGL((halfx_to_halfz<T>), dim3(nb), dim3(256), 0, st, cw, static_cast<cplx<T> *>(out), Q.nx,
                         Q.ny, Q.nz, Q.nxh);
  } else {
    hipLaunchKernelGGL((halfz_to_halfx<T>), dim3(nb), dim3(256), 0, st, static_cast<const cplx<T> *>(in), cw, Q.nx,
                       Q.ny, Q.nz, Q.nxh);
    ok = dispatch_s<T, FFT_Z_INV>(Q.nz, Q, cw, tb, st);
    ok = ok && dispatch_s<T, FFT_Y_INV>(Q.ny, Q, cw, tb, st);
    ok = ok && dispatch_x<T>(false, M, static_cast<T *>(out), cw, nlines, 0, tb, st);
  }
  if (!ok) return fail(X3D_ERR_UNSUPPORTED, "FFT: no kernel for this size");
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(X3D_ERR_HIP, std::string("fft3d launch: ") + hipGetErrorString(e));
  return X3D_OK;
}
template int fft3d<double>(Context &, bool, const void *, void *, hipStream_t);
template int fft3d<float>(Context &, bool, const void *, void *, hipStream_t);

}  // namespace x3d
