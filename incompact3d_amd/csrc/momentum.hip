// Fused momentum right-hand side of transeq.f90:73-470 for a periodic box.
//
// One launch per (sweep direction d, velocity component c) computes the
// three compact solves the reference does for that pair
//     D1_d(u_d u_c)   (transeq.f90:125/198/259)
//     D1_d(u_c)       (transeq.f90:128/201/262)
//     D2_d(u_c)       (transeq.f90:301/336/442)
// and accumulates  -1/2 D1_d(u_d u_c) - 1/2 u_d D1_d(u_c) + nu D2_d(u_c)
// into the running right-hand side, so each (d, c) pair moves four fields
// through HBM (u_d, u_c, the accumulator in, the accumulator out) instead of
// ten.  Each input window is read once; the three solves (two on the
// first-derivative matrix, one on the second-derivative matrix) run between
// one set of block barriers.  The x sweep, the last one in the
// reference's order, also applies AB3 (time_integrators.f90:85-100): the new
// velocity goes to a second buffer, because the other components still read
// the old one.  With FP32 compute the FP64 velocity is not touched here: the
// sweep writes the step's increment in FP32 and the new FP32 velocity, and the
// gradient correction adds the increment to the FP64 velocity, so that the
// FP64 state is read and written once per step.
#include <algorithm>
#include <cmath>
#include <vector>

#include "lines_impl.h"

namespace x3d {

template <typename T>
struct MomArgs {
  const T *ud, *uc, *acc;
  T *out;
  const T *pk1, *pk2;           // packed first/second-derivative sets
  T alfa1, den1, alfa2, den2;
  T rden1, rden2;               // 1 / den
  Sten<T> s;
  T xnu;
  int n, na, nb;
  long long sa, sb, se;         // line strides (y/z sweeps)
  int kdepth;                   // carry-chain depth in segments (y/z sweeps)
  // AB3 epilogue (x sweep only), du = ca*f1 + cb*f2 + cc*f3 in FP64:
  //   FP64: u_new = du + u_c
  //   FP32: delta = du, shadow = du + u_c (the FP64 velocity adds delta later)
  const T *f2, *f3;
  double *u_new;
  float *delta;
  float *shadow;
  double ca, cb, cc;
  int nterms;                   // 0: no AB3
};

// ---------------------------------------------------------------- y / z sweeps
template <typename T, int L>
__global__ void __launch_bounds__(1024) mom_yz_kernel(const MomArgs<T> A) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *smem = reinterpret_cast<T *>(smem_raw);
  const int TL = blockDim.x, P = blockDim.y;
  const int line = threadIdx.x, seg = threadIdx.y;
  const int tid = seg * TL + line, nthr = TL * P;
  const int n = A.n;
  // LDS: two coefficient tables, carries of three right-hand sides
  T *cs1 = smem, *cs2 = cs1 + 4 * n, *cf = cs2 + 4 * n, *cb = cf + 3 * P * TL;
  T *gf = cb + 3 * P * TL, *gb = gf + 2 * P, *x0s = gb + 2 * P, *xns = x0s + 3 * TL;
  int bx, b;
  block_coords<sizeof(T) == 4>(bx, b);
  const int a = bx * TL + line;
  const bool active = a < A.na;
  // lines past the end compute on a copy of the last line (no load branches)
  const int al = active ? a : A.na - 1;
  const int e0 = seg * L;        // n = P*L: every segment is full
  // byte offsets in 32 bits (the host checks the field size), so each load is
  // a uniform base plus one VGPR offset
  const unsigned base = ((unsigned)al * (unsigned)A.sa + (unsigned)b * (unsigned)A.sb) * (unsigned)sizeof(T);
  const unsigned step = (unsigned)A.se * (unsigned)sizeof(T);
  // windows reach at most 4 points past either end of the line (n >= 8)
  auto wrap = [&](int j) { return (unsigned)(j < 0 ? j + n : (j >= n ? j - n : j)); };
  // right-hand sides: w[0] of D1(u_d u_c), w[1] of D1(u_c) (first-derivative
  // matrix), w[2] of D2(u_c) (second-derivative matrix); u_c is read over
  // [e0-4, e0+L+4), u_d over [e0-2, e0+L+2), each once, before the
  // coefficient copy and its barrier
  __builtin_amdgcn_s_setprio(3);   // issue this block's loads ahead of other blocks' arithmetic
  TableCopy<T, 2> tc;
  tc.load(A.pk1, A.pk2, tid, nthr, 4 * n);
  T w[3][L], ud[L];
  {
    T wc[L + 8];
#pragma unroll
    for (int k = 0; k < L + 8; ++k) wc[k] = at(A.uc, base + wrap(e0 - 4 + k) * step);
    {
      T p[L + 4];
#pragma unroll
      for (int k = 0; k < L + 4; ++k) {
        const T d = at(A.ud, base + wrap(e0 - 2 + k) * step);
        p[k] = d * wc[k + 2];
        if (k >= 2 && k < L + 2) ud[k - 2] = d;
      }
#pragma unroll
      for (int m = 0; m < L; ++m) w[0][m] = interior_row<OP_DER1, T, L + 4>(p, m + 2, A.s);
    }
#pragma unroll
    for (int m = 0; m < L; ++m) {
      w[1][m] = interior_row<OP_DER1, T, L + 8>(wc, m + 4, A.s);
      w[2][m] = interior_row<OP_DER2, T, L + 8>(wc, m + 4, A.s);
    }
  }
  __builtin_amdgcn_s_setprio(0);   // back to normal priority once the loads are out
  tc.store(cs1, cs2, A.pk1, A.pk2, tid, nthr, 4 * n);
  __syncthreads();
  seg_solve<T, L, 3, 2>(w, e0, cs1, A.alfa1, A.rden1, cs2, A.alfa2, A.rden2, cf, cb, gf, gb, x0s, xns, line, seg,
                        TL, P, A.kdepth);
  if (!active) return;
  // -1/2 D1(u_d u_c) - 1/2 u_d D1(u_c) + nu D2(u_c)
#pragma unroll
  for (int m = 0; m < L; ++m) w[0][m] = T(-0.5) * w[0][m] + T(-0.5) * (ud[m] * w[1][m]) + A.xnu * w[2][m];
  if (A.acc) {
    // acc may be out (in-place accumulation): all of this thread's acc
    // points are read before any is written
    T ac[L];
#pragma unroll
    for (int m = 0; m < L; ++m) ac[m] = at(A.acc, base + (unsigned)(e0 + m) * step);
#pragma unroll
    for (int m = 0; m < L; ++m) at(A.out, base + (unsigned)(e0 + m) * step) = ac[m] + w[0][m];
  } else {
#pragma unroll
    for (int m = 0; m < L; ++m) at(A.out, base + (unsigned)(e0 + m) * step) = w[0][m];
  }
}

// ---------------------------------------------------------------- one line on one wavefront
// The two coefficient sets of a line, V points per lane (lane k owns points
// kV .. kV+V-1), held in registers.
template <typename T, int V>
struct LineCoef {
  T fs1[V], fw1[V], ffw1[V], sm1[V], fs2[V], fw2[V], ffw2[V], sm2[V];
  __device__ __forceinline__ void load(const MomArgs<T> &A, const int lane) {
#pragma unroll
    for (int m = 0; m < V; ++m) {
      const int e = lane * V + m;
      fs1[m] = A.pk1[4 * e]; fw1[m] = A.pk1[4 * e + 1]; ffw1[m] = A.pk1[4 * e + 2]; sm1[m] = A.pk1[4 * e + 3];
      fs2[m] = A.pk2[4 * e]; fw2[m] = A.pk2[4 * e + 1]; ffw2[m] = A.pk2[4 * e + 2]; sm2[m] = A.pk2[4 * e + 3];
    }
  }
};

// -1/2 D1(u_d u_c) - 1/2 u_d D1(u_c) + nu D2(u_c) of one periodic line held
// by a wavefront: the halos come from the neighbour lanes (DPP), the three
// solves are wave_solve's lane-parallel Thomas sweeps.
template <typename T, int V>
__device__ __forceinline__ void mom_line(const MomArgs<T> &A, const LineCoef<T, V> &C, const T (&ud)[V],
                                         const T (&uc)[V], T (&r)[V], const int lane) {
  {
    T p[V];
#pragma unroll
    for (int m = 0; m < V; ++m) p[m] = ud[m] * uc[m];
    T v[V + 4];
    halo_window<T, V, 2, 2>(p, v, lane);
    T w[V];
#pragma unroll
    for (int m = 0; m < V; ++m) w[m] = A.s.afi * (v[m + 3] - v[m + 1]) + A.s.bfi * (v[m + 4] - v[m]);
    wave_solve<T, V>(w, C.fs1, C.fw1, C.ffw1, C.sm1, A.alfa1, A.rden1, lane);
#pragma unroll
    for (int m = 0; m < V; ++m) r[m] = T(-0.5) * w[m];
  }
  {
    T v[V + 4];
    halo_window<T, V, 2, 2>(uc, v, lane);
    T w[V];
#pragma unroll
    for (int m = 0; m < V; ++m) w[m] = A.s.afi * (v[m + 3] - v[m + 1]) + A.s.bfi * (v[m + 4] - v[m]);
    wave_solve<T, V>(w, C.fs1, C.fw1, C.ffw1, C.sm1, A.alfa1, A.rden1, lane);
#pragma unroll
    for (int m = 0; m < V; ++m) r[m] += T(-0.5) * (ud[m] * w[m]);
  }
  {
    T v[V + 8];
    halo_window<T, V, 4, 4>(uc, v, lane);
    T w[V];
#pragma unroll
    for (int m = 0; m < V; ++m) w[m] = interior_row<OP_DER2, T, V + 8>(v, m + 4, A.s);
    wave_solve<T, V>(w, C.fs2, C.fw2, C.ffw2, C.sm2, A.alfa2, A.rden2, lane);
#pragma unroll
    for (int m = 0; m < V; ++m) r[m] += A.xnu * w[m];
  }
}

// ---------------------------------------------------------------- x sweep (one wavefront per line)
template <typename T, int V>
__global__ void __launch_bounds__(256) mom_x_kernel(const MomArgs<T> A, int nlines) {
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  const int n = A.n;
  const int e0 = lane * V;
  LineCoef<T, V> C;
  C.load(A, lane);
  // the next line's velocities are loaded before this line is solved, so
  // that their latency overlaps the solves (each wave loops over several lines)
  T pud[V], puc[V];
  if (wave < nlines) {
    load_run<T, V>(A.ud + (long long)wave * n + e0, pud);
    load_run<T, V>(A.uc + (long long)wave * n + e0, puc);
  }
  for (int ln = wave; ln < nlines; ln += nwaves) {
    const long long base = (long long)ln * n + e0;
    T ud[V], uc[V];
#pragma unroll
    for (int m = 0; m < V; ++m) { ud[m] = pud[m]; uc[m] = puc[m]; }
    if (ln + nwaves < nlines) {
      load_run<T, V>(A.ud + (long long)(ln + nwaves) * n + e0, pud);
      load_run<T, V>(A.uc + (long long)(ln + nwaves) * n + e0, puc);
    }
    T r[V];
    mom_line<T, V>(A, C, ud, uc, r, lane);
    // every epilogue input is read before the first store (the compiler
    // cannot move loads above stores to possibly aliasing arrays)
    T acc[V], f2[V], f3[V];
    if (A.acc) load_run<T, V>(A.acc + base, acc);
    if (A.nterms >= 2) load_run<T, V>(A.f2 + base, f2);
    if (A.nterms >= 3) load_run<T, V>(A.f3 + base, f3);
    if (A.acc) {
#pragma unroll
      for (int m = 0; m < V; ++m) r[m] = acc[m] + r[m];
    }
    store_run<T, V>(A.out + base, r);
    if (A.nterms) {
      // AB3 with FP64 accumulation: du = a f1 + b f2 + c f3, u_new = du + u_old
      double du[V];
#pragma unroll
      for (int m = 0; m < V; ++m) {
        if (A.nterms == 1) du[m] = A.ca * (double)r[m];
        else if (A.nterms == 2) du[m] = A.ca * (double)r[m] + A.cb * (double)f2[m];
        else du[m] = A.ca * (double)r[m] + A.cb * (double)f2[m] + A.cc * (double)f3[m];
      }
      if constexpr (sizeof(T) == 8) {
        // u_c is the FP64 velocity itself
        double un[V];
#pragma unroll
        for (int m = 0; m < V; ++m) un[m] = du[m] + (double)uc[m];
        store_run<double, V>(A.u_new + base, un);
      } else {
        float d[V], sh[V];
#pragma unroll
        for (int m = 0; m < V; ++m) {
          d[m] = (float)du[m];
          sh[m] = (float)(du[m] + (double)uc[m]);
        }
        store_run<float, V>(A.delta + base, d);
        store_run<float, V>(A.shadow + base, sh);
      }
    }
  }
}

// ---------------------------------------------------------------- host side

// The fused sweeps cover periodic lines with nx = 64, 128, 256 or 512 (one
// wavefront per x line, 1..8 points per lane) and ny, nz that YzShape can
// plan (a multiple of 16 whose block fits; 16..512 for both precisions).  The product derivative
// uses the set of parity d==c (transeq.f90:125) and the velocity derivative
// the opposite one; one set serves both, so they must hold the same
// coefficients, which is the case for periodic lines (schemes.f90:583-596
// builds ffp from ff and prepares both from the same diagonals when
// ncl1 = ncln = 0).
bool momentum_fusable(const Context &c) {
  for (int d = 0; d < 3; ++d) {
    const DirCoef &D = c.dir[d];
    if (D.bc1 != 0 || D.bcn != 0) return false;
    if (!same_set(c.sets[D.set_d1_even], c.sets[D.set_d1_odd])) return false;
    if (d == 0) {
      const int V = D.n / 64;
      if (D.n % 64 != 0 || !(V == 1 || V == 2 || V == 4 || V == 8)) return false;
    } else if (!YzShape<double>().plan(D.n, 2, 3) || !YzShape<float>().plan(D.n, 2, 3)) {
      return false;
    }
  }
  return true;
}

template <typename T>
int launch_momentum(const Context &c, int dir, int comp, const T *const u[3], const T *acc, T *out,
                    const T *f2, const T *f3, double *u_new, float *delta, float *shadow, double ca,
                    double cb, double cc, int nterms, int nx, int ny, int nz, hipStream_t st) {
  const DirCoef &D = c.dir[dir];
  if (D.bc1 != 0) return fail(X3D_ERR_UNSUPPORTED, "fused momentum sweep: periodic lines only");
  if (!same_set(c.sets[D.set_d1_even], c.sets[D.set_d1_odd]))
    return fail(X3D_ERR_UNSUPPORTED, "fused momentum sweep: product and velocity sets differ");
  const int pv = dir == comp ? 0 : 1;
  const TriSet &s1 = c.sets[pv == 1 ? D.set_d1_even : D.set_d1_odd];
  const TriSet &s2 = c.sets[pv == 1 ? D.set_d2_even : D.set_d2_odd];
  MomArgs<T> A{};
  A.ud = u[dir];
  A.uc = u[comp];
  A.acc = acc;
  A.out = out;
  TriDev<T> t1, t2;
  if constexpr (sizeof(T) == 8) { t1 = s1.dev64(); t2 = s2.dev64(); }
  else { t1 = s1.dev32(); t2 = s2.dev32(); }
  A.pk1 = t1.pk; A.pk2 = t2.pk;
  A.alfa1 = t1.alfa; A.den1 = t1.sm_den; A.alfa2 = t2.alfa; A.den2 = t2.sm_den;
  A.rden1 = t1.rden; A.rden2 = t2.rden;
  A.s = cast_sten<T>(D.s);
  A.xnu = (T)c.xnu;
  A.n = D.n;
  A.f2 = f2; A.f3 = f3; A.u_new = u_new; A.delta = delta; A.shadow = shadow;
  A.ca = ca; A.cb = cb; A.cc = cc; A.nterms = nterms;
  const int n = D.n;
  if (dir == 0) {
    const int V = n / 64;
    if (n % 64 != 0 || !(V == 1 || V == 2 || V == 4 || V == 8))
      return fail(X3D_ERR_UNSUPPORTED, "fused x sweep needs nx = 64, 128, 256 or 512");
    if (nterms && (sizeof(T) == 8 ? !u_new : (!delta || !shadow)))
      return fail(X3D_ERR_ARG, "momentum: AB3 output buffers missing");
    const int nlines = ny * nz;
    const int waves = nlines < 8192 ? nlines : 8192;
    const dim3 grid((waves + 3) / 4), block(256);
    switch (V) {
      case 1: hipLaunchKernelGGL((mom_x_kernel<T, 1>), grid, block, 0, st, A, nlines); break;
      case 2: hipLaunchKernelGGL((mom_x_kernel<T, 2>), grid, block, 0, st, A, nlines); break;
      case 4: hipLaunchKernelGGL((mom_x_kernel<T, 4>), grid, block, 0, st, A, nlines); break;
      default: hipLaunchKernelGGL((mom_x_kernel<T, 8>), grid, block, 0, st, A, nlines); break;
    }
  } else {
    if (nterms) return fail(X3D_ERR_ARG, "AB3 is fused into the x sweep only");
    if (dir == 1) {
      A.na = nx; A.nb = nz; A.sa = 1; A.sb = (long long)nx * ny; A.se = nx;
    } else {
      A.na = nx; A.nb = ny; A.sa = 1; A.sb = nx; A.se = (long long)nx * ny;
    }
    constexpr int L = YzShape<T>::L;
    YzShape<T> Y;
    if (!Y.plan(n, 2, 3)) return fail(X3D_ERR_UNSUPPORTED, "fused y/z sweep: line length not supported");
    if ((size_t)nx * ny * nz * sizeof(T) >= ((size_t)1 << 32))
      return fail(X3D_ERR_UNSUPPORTED, "fused y/z sweep: field larger than 4 GiB");
    A.kdepth = std::max(t1.kdepth, t2.kdepth);
    const dim3 grid((A.na + Y.TL - 1) / Y.TL, A.nb), block(Y.TL, Y.P);
    const size_t shmem = Y.shmem;
    hipLaunchKernelGGL((mom_yz_kernel<T, L>), grid, block, shmem, st, A);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(X3D_ERR_HIP, std::string("momentum launch: ") + hipGetErrorString(e));
  return X3D_OK;
}

template int launch_momentum<double>(const Context &, int, int, const double *const[3], const double *, double *,
                                     const double *, const double *, double *, float *, float *, double, double,
                                     double, int, int, int, int, hipStream_t);
template int launch_momentum<float>(const Context &, int, int, const float *const[3], const float *, float *,
                                    const float *, const float *, double *, float *, float *, double, double,
                                    double, int, int, int, int, hipStream_t);

}  // namespace x3d
