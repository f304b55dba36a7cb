// Fused pressure projection sweeps for a periodic box.
//
//   divergence        navier.f90:257-372   x, y, z sweeps (velocity grid ->
//                                          pressure grid, staggered "vp" ops)
//   gradp + cor_vel   navier.f90:386-431,  z, y, x sweeps (pressure grid ->
//                     206-249              velocity grid, "pv" ops)
//
// One launch per sweep applies every staggered derivative and interpolation
// the reference applies along that direction: the interpolations (up to two
// right-hand sides sharing their matrix) and the derivative are solved in the
// same block, and the outputs are the reference's sums, so the eight
// divergence/gradient launches per direction pair become three.  Outputs may
// overwrite inputs: a line never leaves its block (y/z) or its wavefront (x),
// and every window is loaded before the first barrier, stores come after
// the last one.
#include <algorithm>
#include <cmath>
#include <vector>

#include "lines_impl.h"
#include "state64.h"

namespace x3d {

// Inputs of each sweep and the solve they feed: I0/I1 interpolations, D0 the
// derivative (-1: absent).
//   DIV_X   out0 = Dvp(u), out1 = Ivp(v), out2 = Ivp(w)        navier.f90:297,313-314
//   DIV_Y   out0 = Ivp(in0) + Dvp(in1), out1 = Ivp(in2)         navier.f90:321-327
//   DIV_Z   out0 = Ivp(in0) + Dvp(in1)                          navier.f90:333-339
//   GRAD_Z  out0 = Ipv(in0), out1 = Dpv(in0)                    navier.f90:400-403
//   GRAD_Y  out0 = Ipv(in0), out1 = Dpv(in0), out2 = Ipv(in1)   navier.f90:414-419
//   GRAD_X  u -= Dpv(in0), v -= Ipv(in1), w -= Ipv(in2)         navier.f90:426-431, cor_vel 206-249
template <int SW>
struct Recipe;
template <>
struct Recipe<SW_DIV_X> {
  static constexpr bool PV = false;
  static constexpr int NIN = 3, NI = 2, I0 = 1, I1 = 2, D0 = 0;
};
template <>
struct Recipe<SW_DIV_Y> {
  static constexpr bool PV = false;
  static constexpr int NIN = 3, NI = 2, I0 = 0, I1 = 2, D0 = 1;
};
template <>
struct Recipe<SW_DIV_Z> {
  static constexpr bool PV = false;
  static constexpr int NIN = 2, NI = 1, I0 = 0, I1 = -1, D0 = 1;
};
template <>
struct Recipe<SW_GRAD_Z> {
  static constexpr bool PV = true;
  static constexpr int NIN = 1, NI = 1, I0 = 0, I1 = -1, D0 = 0;
};
template <>
struct Recipe<SW_GRAD_Y> {
  static constexpr bool PV = true;
  static constexpr int NIN = 2, NI = 2, I0 = 0, I1 = 1, D0 = 0;
};
template <>
struct Recipe<SW_GRAD_X> {
  static constexpr bool PV = true;
  static constexpr int NIN = 3, NI = 2, I0 = 1, I1 = 2, D0 = 0;
};

template <typename T>
struct ProjArgs {
  const T *in[3];
  T *out[3];
  const T *pkI, *pkD;           // packed interpolation and derivative sets
  T alfaI, denI, alfaD, denD;
  T rdenI, rdenD;               // 1 / den
  Sten<T> s;
  int n, na, nb;
  long long sa, sb, se;         // line strides (y/z sweeps)
  int kdepth;                   // carry-chain depth in segments (y/z sweeps)
  double *u64[3];               // GRAD_X: the FP64 velocity, corrected in place
  float *u32[3];                // GRAD_X, FP32 bulk: its FP32 copy (else null)
  const float *du[3];           // GRAD_X, FP32 bulk: the AB3 increment (else null)
  // GRAD_X, FP32 bulk with the FP64 velocity held as (FP32 rounding, integer
  // residual) pairs (state64.h): hold is the rounding before the step, res
  // the residual (read, then rewritten), u32 receives the new rounding; u64
  // is then unused (null)
  const float *hold[3];
  int32_t *res[3];
  double *divp;                 // DIV_Z: per-block max and sum |div u*| (null: none)
};

// ---------------------------------------------------------------- y / z sweeps
template <typename T, int L, int SW>
__global__ void __launch_bounds__(1024) proj_yz_kernel(const ProjArgs<T> A) {
  using R = Recipe<SW>;
  constexpr int OPI = R::PV ? OP_INTERPV : OP_INTERVP;
  constexpr int OPD = R::PV ? OP_DERPV : OP_DERVP;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *smem = reinterpret_cast<T *>(smem_raw);
  const int TL = blockDim.x, P = blockDim.y;
  const int line = threadIdx.x, seg = threadIdx.y;
  const int tid = seg * TL + line, nthr = TL * P;
  const int n = A.n;
  constexpr int N = R::NI + 1;   // interpolations, then the derivative
  T *csI = smem, *csD = csI + 4 * n, *cf = csD + 4 * n, *cb = cf + N * P * TL;
  T *gf = cb + N * P * TL, *gb = gf + 2 * P, *x0s = gb + 2 * P, *xns = x0s + N * TL;
  int bx, b;
  block_coords<sizeof(T) == 4>(bx, b);
  const int a = bx * TL + line;
  const bool active = a < A.na;
  // lines past the end compute on a copy of the last line (no load branches)
  const int al = active ? a : A.na - 1;
  const int e0 = seg * L;        // n = P*L: every segment is full
  const unsigned base = ((unsigned)al * (unsigned)A.sa + (unsigned)b * (unsigned)A.sb) * (unsigned)sizeof(T);
  const unsigned step = (unsigned)A.se * (unsigned)sizeof(T);
  auto wrap = [&](int j) { return (unsigned)(j < 0 ? j + n : (j >= n ? j - n : j)); };
  __builtin_amdgcn_s_setprio(3);   // issue this block's loads ahead of other blocks' arithmetic
  TableCopy<T, 2> tc;
  tc.load(A.pkI, A.pkD, tid, nthr, 4 * n);
  // w[0..NI-1]: interpolations, w[NI]: derivative
  T w[N][L];
#pragma unroll
  for (int i = 0; i < R::NIN; ++i) {
    // window U(e0-4 .. e0+L+3) of input i; stencils reach at most 4 points
    T win[L + 8];
#pragma unroll
    for (int k = 0; k < L + 8; ++k) win[k] = at(A.in[i], base + wrap(e0 - 4 + k) * step);
    if (i == R::I0) {
#pragma unroll
      for (int m = 0; m < L; ++m) w[0][m] = interior_row<OPI, T, L + 8>(win, m + 4, A.s);
    }
    if constexpr (R::NI == 2) {
      if (i == R::I1) {
#pragma unroll
        for (int m = 0; m < L; ++m) w[1][m] = interior_row<OPI, T, L + 8>(win, m + 4, A.s);
      }
    }
    if (i == R::D0) {
#pragma unroll
      for (int m = 0; m < L; ++m) w[N - 1][m] = interior_row<OPD, T, L + 8>(win, m + 4, A.s);
    }
  }
  __builtin_amdgcn_s_setprio(0);   // back to normal priority once the loads are out
  tc.store(csI, csD, A.pkI, A.pkD, tid, nthr, 4 * n);
  __syncthreads();
  seg_solve<T, L, N, R::NI>(w, e0, csI, A.alfaI, A.rdenI, csD, A.alfaD, A.rdenD, cf, cb, gf, gb, x0s, xns, line,
                            seg, TL, P, A.kdepth);
  if (active) {
#pragma unroll
    for (int m = 0; m < L; ++m) {
      const unsigned o = base + (unsigned)(e0 + m) * step;
      if constexpr (SW == SW_DIV_Y) {
        at(A.out[0], o) = w[0][m] + w[N - 1][m];
        at(A.out[1], o) = w[1][m];
      } else if constexpr (SW == SW_DIV_Z) {
        at(A.out[0], o) = w[0][m] + w[N - 1][m];
      } else if constexpr (SW == SW_GRAD_Z) {   // streaming stores (see st_nt)
        st_nt(A.out[0], o, w[0][m]);
        st_nt(A.out[1], o, w[N - 1][m]);
      } else {   // SW_GRAD_Y, streaming stores
        st_nt(A.out[0], o, w[0][m]);
        st_nt(A.out[1], o, w[N - 1][m]);
        st_nt(A.out[2], o, w[1][m]);
      }
    }
  }
  if constexpr (SW == SW_DIV_Z) {
    // the divergence statistics of navier.f90:348-361 (max and sum of |.|
    // over the pressure grid) as per-block partials, from the values stored
    if (A.divp) {
      double mx = -1.0e300, sm = 0.0;
      if (active) {
#pragma unroll
        for (int m = 0; m < L; ++m) {
          const double v = (double)(w[0][m] + w[N - 1][m]);
          mx = v > mx ? v : mx;
          sm += fabs(v);
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double om = __shfl_xor(mx, off), os = __shfl_xor(sm, off);
        mx = om > mx ? om : mx;
        sm += os;
      }
      __shared__ double rmx[16], rsm[16];
      if ((tid & 63) == 0) {
        rmx[tid >> 6] = mx;
        rsm[tid >> 6] = sm;
      }
      __syncthreads();
      if (tid == 0) {
        double bm = rmx[0], bs = rsm[0];
        for (int i = 1; i < (nthr + 63) / 64; ++i) {
          bm = rmx[i] > bm ? rmx[i] : bm;
          bs += rsm[i];
        }
        const unsigned nblk = gridDim.x * gridDim.y, blk = blockIdx.y * gridDim.x + blockIdx.x;
        A.divp[blk] = bm;
        A.divp[nblk + blk] = bs;
      }
    }
  }
}

// ---------------------------------------------------------------- x sweeps (one wavefront per line)
// cor_vel (navier.f90:206-249) on one run of V points: u -= g in FP64, and
// for FP32-bulk compute first u += the step's AB3 increment, then the FP32
// copy refreshed.  The FP64 velocity is read from and written to either a
// plain FP64 array or its split form (hold/res in, u32/res out).
template <int V, typename T>
__device__ __forceinline__ void correct(const ProjArgs<T> &A, const int c, const long long base, const T (&g)[V]) {
  double u[V];
  const bool split = sizeof(T) == 4 && A.res[c];
  if (split) {
    float h[V];
    int32_t r[V];
    load_run<float, V>(A.hold[c] + base, h);
    load_run<int32_t, V>(A.res[c] + base, r);
#pragma unroll
    for (int m = 0; m < V; ++m) u[m] = join_split(h[m], r[m]);
  } else {
    load_run<double, V>(A.u64[c] + base, u);
  }
  if (sizeof(T) == 4 && A.du[c]) {
    float d[V];
    load_run<float, V>(A.du[c] + base, d);
#pragma unroll
    for (int m = 0; m < V; ++m) u[m] = u[m] + (double)d[m];
  }
#pragma unroll
  for (int m = 0; m < V; ++m) u[m] = u[m] - (double)g[m];
  if (split) {
    float f[V];
    int32_t r[V];
#pragma unroll
    for (int m = 0; m < V; ++m) {
      f[m] = (float)u[m];
      r[m] = split_residual(u[m], f[m]);
    }
    store_run<float, V>(A.u32[c] + base, f);
    store_run<int32_t, V>(A.res[c] + base, r);
    return;
  }
  store_run<double, V>(A.u64[c] + base, u);
  if (sizeof(T) == 4 && A.u32[c]) {
    float f[V];
#pragma unroll
    for (int m = 0; m < V; ++m) f[m] = (float)u[m];
    store_run<float, V>(A.u32[c] + base, f);
  }
}

// FP32 gradient sweep over 256-point lines: held to 128 VGPRs (4 waves per
// SIMD instead of 3), which it fits without spilling
template <typename T, int V, int SW>
constexpr int projx_wpe() { return (sizeof(T) == 4 && V == 4 && SW == SW_GRAD_X) ? 4 : 1; }

template <typename T, int V, int SW>
__global__ void __launch_bounds__(256, (projx_wpe<T, V, SW>())) proj_x_kernel(const ProjArgs<T> A, int nlines) {
  using R = Recipe<SW>;
  constexpr int OPI = R::PV ? OP_INTERPV : OP_INTERVP;
  constexpr int OPD = R::PV ? OP_DERPV : OP_DERVP;
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  const int n = A.n;
  const int e0 = lane * V;
  T fsI[V], fwI[V], ffwI[V], smI[V], fsD[V], fwD[V], ffwD[V], smD[V];
#pragma unroll
  for (int m = 0; m < V; ++m) {
    const int e = e0 + m;
    fsI[m] = A.pkI[4 * e]; fwI[m] = A.pkI[4 * e + 1]; ffwI[m] = A.pkI[4 * e + 2]; smI[m] = A.pkI[4 * e + 3];
    fsD[m] = A.pkD[4 * e]; fwD[m] = A.pkD[4 * e + 1]; ffwD[m] = A.pkD[4 * e + 2]; smD[m] = A.pkD[4 * e + 3];
  }
  // the next line's inputs are loaded before this line is solved, so that
  // their latency overlaps the solves (each wave loops over several lines)
  T pre[R::NIN][V];
  if (wave < nlines) {
#pragma unroll
    for (int i = 0; i < R::NIN; ++i) load_run<T, V>(A.in[i] + (long long)wave * n + e0, pre[i]);
  }
  for (int ln = wave; ln < nlines; ln += nwaves) {
    const long long base = (long long)ln * n + e0;
    T cur[R::NIN][V];
#pragma unroll
    for (int i = 0; i < R::NIN; ++i)
#pragma unroll
      for (int m = 0; m < V; ++m) cur[i][m] = pre[i][m];
    if (ln + nwaves < nlines) {
#pragma unroll
      for (int i = 0; i < R::NIN; ++i) load_run<T, V>(A.in[i] + (long long)(ln + nwaves) * n + e0, pre[i]);
    }
    T wI[2][V], wD[V];
#pragma unroll
    for (int i = 0; i < R::NIN; ++i) {
      T(&own)[V] = cur[i];
      T v[V + 8];
      halo_window<T, V, 4, 4>(own, v, lane);
      if (i == R::I0) {
#pragma unroll
        for (int m = 0; m < V; ++m) wI[0][m] = interior_row<OPI, T, V + 8>(v, m + 4, A.s);
      }
      if (i == R::I1) {
#pragma unroll
        for (int m = 0; m < V; ++m) wI[1][m] = interior_row<OPI, T, V + 8>(v, m + 4, A.s);
      }
      if (i == R::D0) {
#pragma unroll
        for (int m = 0; m < V; ++m) wD[m] = interior_row<OPD, T, V + 8>(v, m + 4, A.s);
      }
    }
    wave_solve<T, V>(wI[0], fsI, fwI, ffwI, smI, A.alfaI, A.rdenI, lane);
    wave_solve<T, V>(wI[1], fsI, fwI, ffwI, smI, A.alfaI, A.rdenI, lane);
    wave_solve<T, V>(wD, fsD, fwD, ffwD, smD, A.alfaD, A.rdenD, lane);
    if constexpr (SW == SW_DIV_X) {
      store_run<T, V>(A.out[0] + base, wD);
      store_run<T, V>(A.out[1] + base, wI[0]);
      store_run<T, V>(A.out[2] + base, wI[1]);
    } else {   // SW_GRAD_X: u_c -= grad_c p with the FP64 velocity (cor_vel)
      correct<V>(A, 0, base, wD);
      correct<V>(A, 1, base, wI[0]);
      correct<V>(A, 2, base, wI[1]);
    }
  }
}

// ---------------------------------------------------------------- host side
// The fused projection covers the grids of the fused momentum sweeps: periodic
// lines, nx = 64..512 (x wavefronts) and ny, nz planned by YzShape.
bool projection_fusable(const Context &c) {
  for (int d = 0; d < 3; ++d) {
    const DirCoef &D = c.dir[d];
    if (D.bc1 != 0 || D.bcn != 0 || D.nm != D.n) return false;
    if (d == 0) {
      const int V = D.n / 64;
      if (D.n % 64 != 0 || !(V == 1 || V == 2 || V == 4 || V == 8)) return false;
    } else if (!YzShape<double>().plan(D.n, 2, 3) || !YzShape<float>().plan(D.n, 2, 3)) {
      return false;
    }
  }
  return true;
}

template <typename T, int SW>
static void launch_yz(const ProjArgs<T> &A, const YzShape<T> &Y, hipStream_t st) {
  const dim3 grid((A.na + Y.TL - 1) / Y.TL, A.nb), block(Y.TL, Y.P);
  hipLaunchKernelGGL((proj_yz_kernel<T, YzShape<T>::L, SW>), grid, block, Y.shmem, st, A);
}

template <typename T, int SW>
static void launch_x(const ProjArgs<T> &A, int V, int nlines, hipStream_t st) {
  const int waves = nlines < 8192 ? nlines : 8192;
  const dim3 grid((waves + 3) / 4), block(256);
  switch (V) {
    case 1: hipLaunchKernelGGL((proj_x_kernel<T, 1, SW>), grid, block, 0, st, A, nlines); break;
    case 2: hipLaunchKernelGGL((proj_x_kernel<T, 2, SW>), grid, block, 0, st, A, nlines); break;
    case 4: hipLaunchKernelGGL((proj_x_kernel<T, 4, SW>), grid, block, 0, st, A, nlines); break;
    default: hipLaunchKernelGGL((proj_x_kernel<T, 8, SW>), grid, block, 0, st, A, nlines); break;
  }
}

// One projection sweep.  in/out follow the recipes above; u64/u32 are used
// by SW_GRAD_X only (u32 null for FP64 compute).
template <typename T>
int launch_projection(const Context &c, int sw, const T *const in[3], T *const out[3], double *const u64[3],
                      float *const u32[3], const float *const du[3], int nx, int ny, int nz, hipStream_t st,
                      double *divp, int *divp_blocks, const float *const hold[3], int32_t *const res[3]) {
  if (sw < SW_DIV_X || sw > SW_GRAD_X) return fail(X3D_ERR_ARG, "projection: unknown sweep");
  const int dir = (sw == SW_DIV_X || sw == SW_GRAD_X) ? 0 : (sw == SW_DIV_Y || sw == SW_GRAD_Y) ? 1 : 2;
  const bool pv = sw >= SW_GRAD_Z;
  const DirCoef &D = c.dir[dir];
  if (D.bc1 != 0 || D.nm != D.n) return fail(X3D_ERR_UNSUPPORTED, "fused projection: periodic lines only");
  const TriSet &sI = c.sets[pv ? D.set_ci6 : D.set_ci6p];
  const TriSet &sD = c.sets[D.set_c6];
  ProjArgs<T> A{};
  for (int i = 0; i < 3; ++i) {
    A.in[i] = in[i];
    A.out[i] = out ? out[i] : nullptr;
    A.u64[i] = u64 ? u64[i] : nullptr;
    A.u32[i] = u32 ? u32[i] : nullptr;
    A.du[i] = du ? du[i] : nullptr;
    A.hold[i] = hold ? hold[i] : nullptr;
    A.res[i] = res ? res[i] : nullptr;
  }
  TriDev<T> tI, tD;
  if constexpr (sizeof(T) == 8) { tI = sI.dev64(); tD = sD.dev64(); }
  else { tI = sI.dev32(); tD = sD.dev32(); }
  A.pkI = tI.pk; A.pkD = tD.pk;
  A.alfaI = tI.alfa; A.denI = tI.sm_den; A.alfaD = tD.alfa; A.denD = tD.sm_den;
  A.rdenI = tI.rden; A.rdenD = tD.rden;
  A.s = cast_sten<T>(D.s);
  A.n = D.n;
  const int n = D.n;
  if (dir == 0) {
    const int V = n / 64;
    if (n % 64 != 0 || !(V == 1 || V == 2 || V == 4 || V == 8))
      return fail(X3D_ERR_UNSUPPORTED, "fused x sweep needs nx = 64, 128, 256 or 512");
    if (sw == SW_GRAD_X) {
      bool ok = true;
      for (int i = 0; i < 3; ++i) {
        if (res && sizeof(T) == 4) ok = ok && hold && hold[i] && res[i] && u32 && u32[i] && du && du[i];
        else ok = ok && u64 && u64[i];
      }
      if (!ok) return fail(X3D_ERR_ARG, "projection: GRAD_X needs the FP64 velocity (plain or split)");
    }
    const int nlines = ny * nz;
    if (sw == SW_DIV_X) launch_x<T, SW_DIV_X>(A, V, nlines, st);
    else launch_x<T, SW_GRAD_X>(A, V, nlines, st);
  } else {
    if (dir == 1) {
      A.na = nx; A.nb = nz; A.sa = 1; A.sb = (long long)nx * ny; A.se = nx;
    } else {
      A.na = nx; A.nb = ny; A.sa = 1; A.sb = nx; A.se = (long long)nx * ny;
    }
    YzShape<T> Y;
    if (!Y.plan(n, 2, 3)) return fail(X3D_ERR_UNSUPPORTED, "fused y/z sweep: line length not supported");
    if ((size_t)nx * ny * nz * sizeof(T) >= ((size_t)1 << 32))
      return fail(X3D_ERR_UNSUPPORTED, "fused y/z sweep: field larger than 4 GiB");
    A.kdepth = std::max(tI.kdepth, tD.kdepth);
    if (sw == SW_DIV_Z && divp) {
      A.divp = divp;
      if (divp_blocks) *divp_blocks = ((A.na + Y.TL - 1) / Y.TL) * A.nb;
    }
    switch (sw) {
      case SW_DIV_Y: launch_yz<T, SW_DIV_Y>(A, Y, st); break;
      case SW_DIV_Z: launch_yz<T, SW_DIV_Z>(A, Y, st); break;
      case SW_GRAD_Z: launch_yz<T, SW_GRAD_Z>(A, Y, st); break;
      default: launch_yz<T, SW_GRAD_Y>(A, Y, st); break;
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(X3D_ERR_HIP, std::string("projection launch: ") + hipGetErrorString(e));
  return X3D_OK;
}

template int launch_projection<double>(const Context &, int, const double *const[3], double *const[3],
                                       double *const[3], float *const[3], const float *const[3], int, int, int,
                                       hipStream_t, double *, int *, const float *const[3], int32_t *const[3]);
template int launch_projection<float>(const Context &, int, const float *const[3], float *const[3],
                                      double *const[3], float *const[3], const float *const[3], int, int, int,
                                      hipStream_t, double *, int *, const float *const[3], int32_t *const[3]);

}  // namespace x3d
