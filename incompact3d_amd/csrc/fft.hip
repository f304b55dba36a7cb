// Spectral pressure Poisson solver for gfx950 (no FFT library).
//
// poisson_000 (poisson.f90:298-410): its shift post-processing cancels
// exactly (forward and backward factors are complex conjugates around the
// conjugated division), so the solve is  p_hat = -f_hat / K  with modes of
// K < 1e-16 set to zero, K being the modified wavenumber of waves()
// (poisson.f90:1706-1783).  poisson_010 (poisson.f90:665-1013) is the same
// with a DCT-II along y computed, as in the reference, from an even/odd
// reordering plus a complex FFT (Makhoul's algorithm), and with the
// (nx/2+1, :, nz/2+1) line zeroed.
//
// Passes over the half spectrum cw (nxh = nx/2+1, ny, nz), kx fastest:
//   1 r2c along x (real line packed as nx/2 complex points, LDS FFT, unpack)
//   2 c2c forward along y            (DCT-II for a non-periodic y)
//   3 c2c forward along z, divide by -K (FP64), c2c inverse along z
//   4 c2c inverse along y            (DCT-III for a non-periodic y)
//   5 c2r along x
// Every line FFT is a radix-4 (+radix-2) Stockham transform held in LDS;
// blocks hold N*TL = 4096 complex points so that strided y/z lines are read
// and written as rows of TL consecutive kx.
#include <cmath>

#include "x3d_internal.h"

namespace x3d {

template <typename T>
struct cplx {
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

constexpr int FFT_BLOCK_POINTS = 4096;

__device__ __forceinline__ int ilog2_dev(int n) { return 31 - __clz(n); }
constexpr int FFT_THREADS = 256;

// In-place Stockham FFT of TL lines of length N held in LDS as buf[e*TLp + l].
// tw[m] = exp(-2 pi i m / N); inverse uses the conjugate.
template <typename T>
__device__ void lds_fft(cplx<T> *buf, int N, int lg, int TL, int TLp, const cplx<T> *__restrict__ tw,
                        bool inverse) {
  const int tid = threadIdx.x, nthr = blockDim.x;
  int Ns = 1;
  const int nr4 = lg >> 1;
  for (int st = 0; st < nr4; ++st) {
    const int quarter = N >> 2;
    const int nb = quarter * TL;
    cplx<T> v[4][4];
    int jj[4], ll[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int bidx = tid + q * nthr;
      jj[q] = -1;
      if (bidx < nb) {
        const int l = bidx % TL, j = bidx / TL;
        jj[q] = j;
        ll[q] = l;
        const int k = j % Ns;
        const int step = N / (Ns * 4);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          cplx<T> x = buf[(j + r * quarter) * TLp + l];
          if (r > 0) {
            cplx<T> w = tw[r * k * step];
            if (inverse) w.y = -w.y;
            x = cmul(x, w);
          }
          v[q][r] = x;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (jj[q] >= 0) {
        const cplx<T> a = cadd(v[q][0], v[q][2]);
        const cplx<T> b = csub(v[q][0], v[q][2]);
        const cplx<T> c = cadd(v[q][1], v[q][3]);
        const cplx<T> dd = csub(v[q][1], v[q][3]);
        // forward: d*(-i); inverse: d*(+i)
        const cplx<T> d = inverse ? cplx<T>{-dd.y, dd.x} : cplx<T>{dd.y, -dd.x};
        const int j = jj[q], l = ll[q];
        const int k = j % Ns;
        const int o = (j / Ns) * Ns * 4 + k;
        buf[(o)*TLp + l] = cadd(a, c);
        buf[(o + Ns) * TLp + l] = cadd(b, d);
        buf[(o + 2 * Ns) * TLp + l] = csub(a, c);
        buf[(o + 3 * Ns) * TLp + l] = csub(b, d);
      }
    }
    __syncthreads();
    Ns *= 4;
  }
  if (lg & 1) {
    const int half = N >> 1;
    const int nb = half * TL;
    cplx<T> v[8][2];
    int jj[8], ll[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int bidx = tid + q * nthr;
      jj[q] = -1;
      if (bidx < nb) {
        const int l = bidx % TL, j = bidx / TL;
        jj[q] = j;
        ll[q] = l;
        const int k = j % Ns;
        const int step = N / (Ns * 2);
        v[q][0] = buf[j * TLp + l];
        cplx<T> x = buf[(j + half) * TLp + l];
        cplx<T> w = tw[k * step];
        if (inverse) w.y = -w.y;
        v[q][1] = cmul(x, w);
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (jj[q] >= 0) {
        const int j = jj[q], l = ll[q];
        const int k = j % Ns;
        const int o = (j / Ns) * Ns * 2 + k;
        buf[o * TLp + l] = cadd(v[q][0], v[q][1]);
        buf[(o + Ns) * TLp + l] = csub(v[q][0], v[q][1]);
      }
    }
    __syncthreads();
  }
}

template <typename T>
struct PoissonTables {
  const cplx<T> *twx, *twy, *twz;   // exp(-2 pi i m / N) for N = nx/2, ny, nz
  const cplx<T> *rtx;               // exp(-2 pi i k / nx), k = 0..nx/2 (r2c unpack)
  const cplx<T> *dcty;              // exp(-i pi k / (2 ny)), k = 0..ny-1 (DCT along y)
  const double *kx2, *tx;           // modified wavenumber^2 and transfer fn, per kx (nxh)
  const double *ky2, *ty;           // per ky (ny)
  const double *kz2, *tz;           // per kz (nz)
};

struct PoissonPlan {
  int nx, ny, nz, nxh;              // pressure-grid sizes (ny = nym when y is not periodic)
  int ydct;                         // 1 when y is non-periodic (poisson_010)
  double scale;                     // 1/(nx*ny*nz)
  void *tab64 = nullptr, *tab32 = nullptr;
  PoissonTables<double> t64;
  PoissonTables<float> t32;
  void *cw64 = nullptr, *cw32 = nullptr;   // spectral work arrays
};

static int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return (1 << l) == n ? l : -1;
}

// ---------------------------------------------------------------- pass 1: r2c along x
template <typename T>
__global__ void __launch_bounds__(FFT_THREADS) r2c_x_kernel(const T *__restrict__ in, cplx<T> *__restrict__ out,
                                                             int nx, int nlines, PoissonTables<T> tb) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
  const int M = nx >> 1, lg = ilog2_dev(M);
  const int TL = FFT_BLOCK_POINTS / M, TLp = TL + 1;
  const int l0 = blockIdx.x * TL;
  const int nv = min(TL, nlines - l0);
  const int tid = threadIdx.x, nthr = blockDim.x;
  const T *src = in + (long long)l0 * nx;
  for (int idx = tid; idx < nv * M; idx += nthr) {
    const int l = idx / M, m = idx - l * M;
    const T *p = src + (long long)l * nx + 2 * m;
    buf[m * TLp + l] = cplx<T>{p[0], p[1]};
  }
  for (int idx = nv * M + tid; idx < TL * M; idx += nthr) {
    const int l = idx / M, m = idx - l * M;
    buf[m * TLp + l] = cplx<T>{T(0), T(0)};
  }
  __syncthreads();
  lds_fft<T>(buf, M, lg, TL, TLp, tb.twx, false);
  // unpack: X_k = (Z_k + conj Z_{M-k})/2 - i/2 e^{-2 pi i k/nx} (Z_k - conj Z_{M-k})
  const int nxh = M + 1;
  cplx<T> *dst = out + (long long)l0 * nxh;
  for (int idx = tid; idx < nv * nxh; idx += nthr) {
    const int l = idx / nxh, k = idx - l * nxh;
    const cplx<T> zk = buf[(k == M ? 0 : k) * TLp + l];
    const cplx<T> zm = cconj(buf[(k == 0 ? 0 : M - k) * TLp + l]);
    const cplx<T> e = {T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y + zm.y)};
    const cplx<T> o0 = {T(0.5) * (zk.x - zm.x), T(0.5) * (zk.y - zm.y)};
    const cplx<T> o = {o0.y, -o0.x};  // (Z_k - conj Z_{M-k}) / (2i)
    dst[(long long)l * nxh + k] = cadd(e, cmul(tb.rtx[k], o));
  }
}

// ---------------------------------------------------------------- pass 5: c2r along x
template <typename T>
__global__ void __launch_bounds__(FFT_THREADS) c2r_x_kernel(const cplx<T> *__restrict__ in, T *__restrict__ out,
                                                             int nx, int nlines, PoissonTables<T> tb) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
  const int M = nx >> 1, lg = ilog2_dev(M), nxh = M + 1;
  const int TL = FFT_BLOCK_POINTS / M, TLp = TL + 1;
  const int l0 = blockIdx.x * TL;
  const int nv = min(TL, nlines - l0);
  const int tid = threadIdx.x, nthr = blockDim.x;
  const cplx<T> *src = in + (long long)l0 * nxh;
  // Z_k = (X_k + conj X_{M-k}) + i e^{+2 pi i k/nx} (X_k - conj X_{M-k}),  k < M
  for (int idx = tid; idx < TL * M; idx += nthr) {
    const int l = idx / M, k = idx - l * M;
    cplx<T> z = {T(0), T(0)};
    if (l < nv) {
      const cplx<T> xk = src[(long long)l * nxh + k];
      const cplx<T> xm = cconj(src[(long long)l * nxh + (M - k)]);
      const cplx<T> e = cadd(xk, xm);
      const cplx<T> d = csub(xk, xm);
      const cplx<T> w = cconj(tb.rtx[k]);
      const cplx<T> od = cmul(w, d);
      z = cplx<T>{e.x - od.y, e.y + od.x};
    }
    buf[k * TLp + l] = z;
  }
  __syncthreads();
  lds_fft<T>(buf, M, lg, TL, TLp, tb.twx, true);
  T *dst = out + (long long)l0 * nx;
  for (int idx = tid; idx < nv * M; idx += nthr) {
    const int l = idx / M, m = idx - l * M;
    const cplx<T> z = buf[m * TLp + l];
    T *p = dst + (long long)l * nx + 2 * m;
    p[0] = z.x;
    p[1] = z.y;
  }
}

// ---------------------------------------------------------------- passes 2-4: strided lines
enum { FFT_Y_FWD = 0, FFT_Y_INV = 1, FFT_Z_SOLVE = 2 };

template <typename T, int MODE>
__global__ void __launch_bounds__(FFT_THREADS) strided_kernel(cplx<T> *__restrict__ cw, int nxh, int ny, int nz,
                                                               int ydct, double scale, int zero_line,
                                                               PoissonTables<T> tb) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
  const int N = MODE == FFT_Z_SOLVE ? nz : ny;
  const int lg = ilog2_dev(N);
  const int TL = FFT_BLOCK_POINTS / N, TLp = TL + 1;
  const int a0 = blockIdx.x * TL;            // kx tile
  const int b = blockIdx.y;                  // kz (y passes) or ky (z pass)
  const int nv = min(TL, nxh - a0);
  const int tid = threadIdx.x, nthr = blockDim.x;
  const long long se = MODE == FFT_Z_SOLVE ? (long long)nxh * ny : (long long)nxh;
  const long long base = MODE == FFT_Z_SOLVE ? (long long)b * nxh + a0 : (long long)b * nxh * ny + a0;
  const cplx<T> *tw = MODE == FFT_Z_SOLVE ? tb.twz : tb.twy;

  // ---- load (DCT forward: even/odd reorder, poisson.f90:688-697)
  for (int idx = tid; idx < TL * N; idx += nthr) {
    const int e = idx / TL, l = idx - e * TL;
    cplx<T> x = {T(0), T(0)};
    if (l < nv) x = cw[base + (long long)e * se + l];
    int pos = e;
    if (MODE == FFT_Y_FWD && ydct) pos = (e & 1) ? N - 1 - (e >> 1) : (e >> 1);
    buf[pos * TLp + l] = x;
  }
  __syncthreads();

  if (MODE == FFT_Y_INV && ydct) {
    // DCT-III input side: V_k = e^{+i pi k/2N} (C_k - i C_{N-k}), V_0 = C_0
    // (poisson.f90:924-943).  Each thread owns the pair (k, N-k).
    const int half = N >> 1;
    for (int bidx = tid; bidx < (half + 1) * TL; bidx += nthr) {
      const int l = bidx % TL, k = bidx / TL;
      const cplx<T> ck = buf[k * TLp + l];
      const cplx<T> cm = (k == 0) ? cplx<T>{T(0), T(0)} : buf[(N - k) * TLp + l];
      const cplx<T> vk = cmul(cconj(tb.dcty[k]), cplx<T>{ck.x + cm.y, ck.y - cm.x});
      if (k > 0 && k < half) {
        const cplx<T> vm = cmul(cconj(tb.dcty[N - k]), cplx<T>{cm.x + ck.y, cm.y - ck.x});
        buf[(N - k) * TLp + l] = vm;
      }
      buf[k * TLp + l] = vk;
    }
    __syncthreads();
  }

  lds_fft<T>(buf, N, lg, TL, TLp, tw, MODE == FFT_Y_INV);

  if (MODE == FFT_Y_FWD && ydct) {
    // DCT-II output side: C_k = (e^{-i pi k/2N} V_k + e^{+i pi k/2N} V_{N-k}) / 2
    // (poisson.f90:761-781).  Each thread owns the pair (k, N-k).
    const int half = N >> 1;
    for (int bidx = tid; bidx < (half + 1) * TL; bidx += nthr) {
      const int l = bidx % TL, k = bidx / TL;
      const cplx<T> vk = buf[k * TLp + l];
      if (k == 0) continue;
      const cplx<T> vm = buf[(N - k) * TLp + l];
      const cplx<T> a = cmul(tb.dcty[k], vk);
      const cplx<T> bb = cmul(cconj(tb.dcty[k]), vm);
      if (k < half) {
        const cplx<T> a2 = cmul(tb.dcty[N - k], vm);
        const cplx<T> b2 = cmul(cconj(tb.dcty[N - k]), vk);
        buf[(N - k) * TLp + l] = cplx<T>{T(0.5) * (a2.x + b2.x), T(0.5) * (a2.y + b2.y)};
      }
      buf[k * TLp + l] = cplx<T>{T(0.5) * (a.x + bb.x), T(0.5) * (a.y + bb.y)};
    }
    __syncthreads();
  }

  if (MODE == FFT_Z_SOLVE) {
    // divide by -K in FP64 (poisson.f90:362-371 / :795-819)
    const int ky = b;
    const double ky2 = tb.ky2[ky], tyv = tb.ty[ky];
    for (int idx = tid; idx < TL * N; idx += nthr) {
      const int kz = idx / TL, l = idx - kz * TL;
      const int kx = a0 + l;
      cplx<T> x = buf[kz * TLp + l];
      if (kx < nxh) {
        const double txv = tb.tx[kx], tzv = tb.tz[kz];
        const double yz = tyv * tzv, xz = txv * tzv, xy = txv * tyv;
        const double K = tb.kx2[kx] * (yz * yz) + ky2 * (xz * xz) + tb.kz2[kz] * (xy * xy);
        const bool zero = (K < 1.0e-16) || (zero_line && kx == (nxh - 1) && kz == (N >> 1));
        const double f = zero ? 0.0 : -scale / K;
        x = cplx<T>{(T)((double)x.x * f), (T)((double)x.y * f)};
      }
      buf[kz * TLp + l] = x;
    }
    __syncthreads();
    lds_fft<T>(buf, N, lg, TL, TLp, tw, true);
  }

  // ---- store (DCT inverse: undo the even/odd reorder, poisson.f90:997-1008)
  for (int idx = tid; idx < TL * N; idx += nthr) {
    const int e = idx / TL, l = idx - e * TL;
    if (l < nv) {
      int pos = e;
      if (MODE == FFT_Y_INV && ydct) pos = (e & 1) ? N - 1 - (e >> 1) : (e >> 1);
      cw[base + (long long)e * se + l] = buf[pos * TLp + l];
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
  out->kx2 = reinterpret_cast<const double *>(b + o_r[0]);
  out->tx = reinterpret_cast<const double *>(b + o_r[1]);
  out->ky2 = reinterpret_cast<const double *>(b + o_r[2]);
  out->ty = reinterpret_cast<const double *>(b + o_r[3]);
  out->kz2 = reinterpret_cast<const double *>(b + o_r[4]);
  out->tz = reinterpret_cast<const double *>(b + o_r[5]);
  return X3D_OK;
}

int poisson_setup(Context &c) {
  const DirCoef &X = c.dir[0], &Y = c.dir[1], &Z = c.dir[2];
  if (X.bc1 != 0 || Z.bc1 != 0) return fail(X3D_ERR_UNSUPPORTED, "Poisson: x and z must be periodic (poisson_100/11x are not provided)");
  auto *P = new PoissonPlan();
  P->nx = X.n;
  P->ny = Y.nm;
  P->nz = Z.n;
  P->nxh = P->nx / 2 + 1;
  P->ydct = Y.bc1 != 0;
  P->scale = 1.0 / ((double)P->nx * (double)P->ny * (double)P->nz);
  const int lx = ilog2(P->nx), ly = ilog2(P->ny), lz = ilog2(P->nz);
  if (lx < 4 || ly < 3 || lz < 3 || P->nx > 2048 || P->ny > 1024 || P->nz > 1024) {
    delete P;
    return fail(X3D_ERR_UNSUPPORTED, "Poisson: pressure-grid sizes must be powers of two (nx 16..2048, ny/nz 8..1024)");
  }
  std::vector<double> tabs[6];
  wave_1d(X, 0, tabs[0], tabs[1]);
  wave_1d(Y, P->ydct ? 1 : 0, tabs[2], tabs[3]);
  wave_1d(Z, 0, tabs[4], tabs[5]);
  tabs[0].resize(P->nxh);
  tabs[1].resize(P->nxh);
  int r = build_tables<double>(*P, tabs, &P->tab64, &P->t64);
  if (r == X3D_OK) r = build_tables<float>(*P, tabs, &P->tab32, &P->t32);
  if (r != X3D_OK) {
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
  delete P;
  c.poisson = nullptr;
}

template <typename T>
int poisson_solve(Context &c, T *rhs, hipStream_t st) {
  auto *P = static_cast<PoissonPlan *>(c.poisson);
  if (!P) return fail(X3D_ERR_STATE, "Poisson solver not initialised");
  void **cwp = sizeof(T) == 8 ? &P->cw64 : &P->cw32;
  const size_t ncw = (size_t)P->nxh * P->ny * P->nz;
  if (!*cwp) X3D_HIP(hipMalloc(cwp, ncw * sizeof(cplx<T>)));
  cplx<T> *cw = static_cast<cplx<T> *>(*cwp);
  PoissonTables<T> tb;
  if constexpr (sizeof(T) == 8) tb = P->t64;
  else tb = P->t32;
  const int nlines = P->ny * P->nz;
  // pass 1
  {
    const int M = P->nx / 2, TL = FFT_BLOCK_POINTS / M;
    const size_t sh = sizeof(cplx<T>) * (size_t)M * (TL + 1);
    hipLaunchKernelGGL((r2c_x_kernel<T>), dim3((nlines + TL - 1) / TL), dim3(FFT_THREADS), sh, st, rhs, cw, P->nx,
                       nlines, tb);
  }
  // pass 2
  {
    const int N = P->ny, TL = FFT_BLOCK_POINTS / N;
    const size_t sh = sizeof(cplx<T>) * (size_t)N * (TL + 1);
    hipLaunchKernelGGL((strided_kernel<T, FFT_Y_FWD>), dim3((P->nxh + TL - 1) / TL, P->nz), dim3(FFT_THREADS), sh, st,
                       cw, P->nxh, P->ny, P->nz, P->ydct, P->scale, 0, tb);
  }
  // pass 3
  {
    const int N = P->nz, TL = FFT_BLOCK_POINTS / N;
    const size_t sh = sizeof(cplx<T>) * (size_t)N * (TL + 1);
    hipLaunchKernelGGL((strided_kernel<T, FFT_Z_SOLVE>), dim3((P->nxh + TL - 1) / TL, P->ny), dim3(FFT_THREADS), sh,
                       st, cw, P->nxh, P->ny, P->nz, P->ydct, P->scale, P->ydct, tb);
  }
  // pass 4
  {
    const int N = P->ny, TL = FFT_BLOCK_POINTS / N;
    const size_t sh = sizeof(cplx<T>) * (size_t)N * (TL + 1);
    hipLaunchKernelGGL((strided_kernel<T, FFT_Y_INV>), dim3((P->nxh + TL - 1) / TL, P->nz), dim3(FFT_THREADS), sh, st,
                       cw, P->nxh, P->ny, P->nz, P->ydct, P->scale, 0, tb);
  }
  // pass 5
  {
    const int M = P->nx / 2, TL = FFT_BLOCK_POINTS / M;
    const size_t sh = sizeof(cplx<T>) * (size_t)M * (TL + 1);
    hipLaunchKernelGGL((c2r_x_kernel<T>), dim3((nlines + TL - 1) / TL), dim3(FFT_THREADS), sh, st, cw, rhs, P->nx,
                       nlines, tb);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(X3D_ERR_HIP, std::string("poisson launch: ") + hipGetErrorString(e));
  return X3D_OK;
}

template int poisson_solve<double>(Context &, double *, hipStream_t);
template int poisson_solve<float>(Context &, float *, hipStream_t);

}  // namespace x3d
