// Line-batched compact-scheme solves for gfx950.
//
// Every operator of derive.f90 has the same shape along one direction:
//   t = A^{-1} (B u)
// with B an explicit stencil (interior rows + boundary closures) and A a
// tridiagonal matrix whose Thomas factors (ff, fs, fw) come from prepare()
// (schemes.f90:413-439); periodic lines add the Sherman-Morrison correction
// of derive.f90:55-59.
//
// GPU mapping (one kernel template for all operators and directions):
//  * a block owns TL lines x P segments; thread (line, seg) owns L
//    consecutive points of one line in registers;
//  * y and z lines are strided in memory, and the TL lines of a block are TL
//    consecutive i: every load/store of a point row is TL contiguous words
//    (512 B in FP64), read straight from HBM into registers;
//  * x lines are contiguous: the block stages TL whole lines through LDS with
//    fully contiguous loads/stores, and threads read their segment from LDS;
//  * the two first-order recurrences of the Thomas algorithm
//      y_i = r_i - fs_i y_{i-1},   x_i = (y_i - ff_i x_{i+1}) fw_i
//    are solved per segment with zero carry-in, then the carries are chained
//    across the P segments of a line through LDS and applied as
//    y_i += G_i c (G_i = prod(-fs)), x_i += H_i c (H_i = prod(-ff fw)).
//    |fs|,|ff fw| < 1 for these diagonally dominant systems, so the result
//    equals the sequential sweep to rounding.
#include "x3d_internal.h"

namespace x3d {

template <typename T>
struct LineArgs {
  const T *u, *u2;
  T *t;
  const T *pp;
  const T *acc, *mul;                 // epilogue: t = ac*acc + rc*mul*result
  T ac, rc;
  const T *ff, *fs, *fw, *sm;
  T alfa, sm_den;
  Sten<T> s;
  int periodic, bc1, bcn, npaire;
  int n_in, n_out;
  int na, nb;                         // line grid extents (a fastest)
  long long sa_in, sb_in, se_in;      // input strides (elements)
  long long sa_out, sb_out, se_out;   // output strides
  int row;                            // LDS row stride (x mode)
};

// ---------------------------------------------------------------- stencils
// Interior rows: v[k] holds U(i) of output row i; offsets are compile-time.
template <int OP, typename T, int W>
__device__ __forceinline__ T interior_row(const T (&v)[W], const int k, const Sten<T> &s) {
  if constexpr (OP == OP_DER1) {
    // derive.f90:34-36
    return s.afi * (v[k + 1] - v[k - 1]) + s.bfi * (v[k + 2] - v[k - 2]);
  } else if constexpr (OP == OP_DER2) {
    // derive.f90:1414-1421
    return s.asi * (v[k + 1] - v[k] - v[k] + v[k - 1]) + s.bsi * (v[k + 2] - v[k] - v[k] + v[k - 2]) +
           s.csi * (v[k + 3] - v[k] - v[k] + v[k - 3]) + s.dsi * (v[k + 4] - v[k] - v[k] + v[k - 4]);
  } else if constexpr (OP == OP_DERVP) {
    // derive.f90:3826-3827
    return s.aci * (v[k + 1] - v[k]) + s.bci * (v[k + 2] - v[k - 1]);
  } else if constexpr (OP == OP_INTERVP) {
    // derive.f90:3947-3950
    return s.aici * (v[k + 1] + v[k]) + s.bici * (v[k + 2] + v[k - 1]) + s.cici * (v[k + 3] + v[k - 2]) +
           s.dici * (v[k + 4] + v[k - 3]);
  } else if constexpr (OP == OP_DERPV) {
    // derive.f90:4068-4069
    return s.aci * (v[k] - v[k - 1]) + s.bci * (v[k + 1] - v[k - 2]);
  } else {
    // derive.f90:4170-4173
    return s.aici * (v[k] + v[k - 1]) + s.bici * (v[k + 1] + v[k - 2]) + s.cici * (v[k + 2] + v[k - 3]) +
           s.dici * (v[k + 3] + v[k - 4]);
  }
}

// Rows of a non-periodic line that use a boundary closure.
template <int OP>
__device__ __forceinline__ bool is_boundary_row(int i, int n_out) {
  if constexpr (OP == OP_DER1) return i < 2 || i >= n_out - 2;
  else if constexpr (OP == OP_DER2) return i < 4 || i >= n_out - 4;
  else if constexpr (OP == OP_DERVP) return i == 0 || i == n_out - 1;
  else if constexpr (OP == OP_INTERVP) return i < 3 || i >= n_out - 3;
  else if constexpr (OP == OP_DERPV) return i < 2 || i >= n_out - 2;
  else return i < 4 || i >= n_out - 4;
}

template <typename T>
__device__ __forceinline__ T d2t(T a, T up, T u0, T um) {
  return a * (up - u0 - u0 + um);
}

// Boundary closures, written row by row after the reference.
template <int OP, typename T, typename LD>
__device__ T boundary_row(int i, int n_in, int n_out, const Sten<T> &s, int bc1, int bcn, int npaire,
                          LD U) {
  if constexpr (OP == OP_DER1) {
    // derx_11 derive.f90:88-130, derx_12 :161-200, derx_21 :231-270, derx_22 :299-306
    const int n = n_out;
    const T a = s.afi, b = s.bfi;
    if (i < 2) {
      if (bc1 == 1) {
        if (npaire == 1) return i == 0 ? T(0) : a * (U(2) - U(0)) + b * (U(3) - U(1));
        return i == 0 ? a * (U(1) + U(1)) + b * (U(2) + U(2)) : a * (U(2) - U(0)) + b * (U(3) + U(1));
      }
      return i == 0 ? s.af1 * U(0) + s.bf1 * U(1) + s.cf1 * U(2) : s.af2 * (U(2) - U(0));
    }
    if (bcn == 1) {
      if (npaire == 1)
        return i == n - 2 ? a * (U(n - 1) - U(n - 3)) + b * (U(n - 2) - U(n - 4)) : T(0);
      return i == n - 2 ? a * (U(n - 1) - U(n - 3)) + b * ((-U(n - 2)) - U(n - 4))
                        : a * ((-U(n - 2)) - U(n - 2)) + b * ((-U(n - 3)) - U(n - 3));
    }
    return i == n - 2 ? s.afm * (U(n - 1) - U(n - 3))
                      : (-s.afn * U(n - 1)) - s.bfn * U(n - 2) - s.cfn * U(n - 3);
  } else if constexpr (OP == OP_DER2) {
    // derxx_11 derive.f90:1481-1662, derxx_22 :1978-2048 (12/21 combine both)
    const int n = n_out;
    const T A = s.asi, B = s.bsi, C = s.csi, D = s.dsi;
    if (i < 4) {
      if (bc1 == 1 && npaire == 1) {
        auto g = [&](int j) { return j < 0 ? U(-j) : U(j); };
        return d2t(A, g(i + 1), U(i), g(i - 1)) + d2t(B, g(i + 2), U(i), g(i - 2)) +
               d2t(C, g(i + 3), U(i), g(i - 3)) + d2t(D, g(i + 4), U(i), g(i - 4));
      }
      if (bc1 == 1) {
        if (i == 0) return T(0);
        if (i == 1)
          return d2t(A, U(2), U(1), U(0)) + B * (U(3) - U(1) - U(1) - U(1)) + C * (U(4) - U(1) - U(1) - U(2)) +
                 D * (U(5) - U(1) - U(1) - U(3));
        if (i == 2)
          return d2t(A, U(3), U(2), U(1)) + d2t(B, U(4), U(2), U(0)) + C * (U(5) - U(2) - U(2) - U(1)) +
                 D * (U(6) - U(2) - U(2) - U(2));
        return d2t(A, U(4), U(3), U(2)) + d2t(B, U(5), U(3), U(1)) + C * (U(6) - U(3) - U(3) - U(0)) +
               D * (U(7) - U(3) - U(3) - U(1));
      }
      if (i == 0) return s.as1 * U(0) + s.bs1 * U(1) + s.cs1 * U(2) + s.ds1 * U(3);
      if (i == 1) return d2t(s.as2, U(2), U(1), U(0));
      if (i == 2) return d2t(s.as3, U(3), U(2), U(1)) + d2t(s.bs3, U(4), U(2), U(0));
      return d2t(s.as4, U(4), U(3), U(2)) + d2t(s.bs4, U(5), U(3), U(1)) + d2t(s.cs4, U(6), U(3), U(0));
    }
    if (bcn == 1 && npaire == 1) {
      auto g = [&](int j) { return j > n - 1 ? U(2 * (n - 1) - j) : U(j); };
      return d2t(A, g(i + 1), U(i), g(i - 1)) + d2t(B, g(i + 2), U(i), g(i - 2)) +
             d2t(C, g(i + 3), U(i), g(i - 3)) + d2t(D, g(i + 4), U(i), g(i - 4));
    }
    if (bcn == 1) {
      if (i == n - 1) return T(0);
      if (i == n - 4)
        return d2t(A, U(n - 3), U(n - 4), U(n - 5)) + d2t(B, U(n - 2), U(n - 4), U(n - 6)) +
               C * (-U(n - 1) - U(n - 4) - U(n - 4) + U(n - 7)) + D * (-U(n - 2) - U(n - 4) - U(n - 4) + U(n - 8));
      if (i == n - 3)
        return d2t(A, U(n - 2), U(n - 3), U(n - 4)) + d2t(B, U(n - 1), U(n - 3), U(n - 5)) +
               C * (-U(n - 2) - U(n - 3) - U(n - 3) + U(n - 6)) + D * (-U(n - 3) - U(n - 3) - U(n - 3) + U(n - 7));
      return d2t(A, U(n - 1), U(n - 2), U(n - 3)) + B * (-U(n - 2) - U(n - 2) - U(n - 2) + U(n - 4)) +
             C * (-U(n - 3) - U(n - 2) - U(n - 2) + U(n - 5)) + D * (-U(n - 4) - U(n - 2) - U(n - 2) + U(n - 6));
    }
    if (i == n - 4)
      return d2t(s.astt, U(n - 3), U(n - 4), U(n - 5)) + d2t(s.bstt, U(n - 2), U(n - 4), U(n - 6)) +
             d2t(s.cstt, U(n - 1), U(n - 4), U(n - 7));
    if (i == n - 3) return d2t(s.ast, U(n - 2), U(n - 3), U(n - 4)) + d2t(s.bst, U(n - 1), U(n - 3), U(n - 5));
    if (i == n - 2) return d2t(s.asm_, U(n - 1), U(n - 2), U(n - 3));
    return s.asn * U(n - 1) + s.bsn * U(n - 2) + s.csn * U(n - 3) + s.dsn * U(n - 4);
  } else if constexpr (OP == OP_DERVP) {
    // derxvp non-periodic, derive.f90:3851-3899 (nm = n_out = n_in - 1)
    const int n = n_in, nm = n_out;
    const T a = s.aci, b = s.bci;
    if (npaire == 1) {
      if (i == 0) return a * (U(1) - U(0)) + b * (U(2) - U(1));
      return a * (U(n - 1) - U(nm - 1)) + b * (U(nm - 1) - U(nm - 2));
    }
    if (i == 0) return a * (U(1) - U(0)) + b * (U(2) - T(2) * U(0) + U(1));
    return a * (U(n - 1) - U(nm - 1)) + b * (T(2) * U(n - 1) - U(nm - 1) - U(nm - 2));
  } else if constexpr (OP == OP_INTERVP) {
    // interxvp non-periodic npaire=1, derive.f90:3989-4031: the rows are the
    // interior formula with u mirrored about the end points.
    const int n = n_in;
    auto w = [&](int j) { return j < 0 ? U(-j) : (j > n - 1 ? U(2 * (n - 1) - j) : U(j)); };
    return s.aici * (w(i + 1) + w(i)) + s.bici * (w(i + 2) + w(i - 1)) + s.cici * (w(i + 3) + w(i - 2)) +
           s.dici * (w(i + 4) + w(i - 3));
  } else if constexpr (OP == OP_DERPV) {
    // derxpv non-periodic npaire=1, derive.f90:4096-4116
    const int n = n_out;
    const T a = s.aci, b = s.bci;
    if (i == 0 || i == n - 1) return T(0);
    if (i == 1) return a * (U(1) - U(0)) + b * (U(2) - U(0));
    return a * (U(n - 2) - U(n - 3)) + b * (U(n - 2) - U(n - 4));
  } else {
    // interxpv non-periodic npaire=1, derive.f90:4209-4255: interior formula
    // with u mirrored about the pressure half-points -1/2 and nm-1/2.
    const int nm = n_in;
    auto w = [&](int j) { return j < 0 ? U(-1 - j) : (j > nm - 1 ? U(2 * nm - 1 - j) : U(j)); };
    return s.aici * (w(i) + w(i - 1)) + s.bici * (w(i + 1) + w(i - 2)) + s.cici * (w(i + 2) + w(i - 3)) +
           s.dici * (w(i + 3) + w(i - 4));
  }
}

// ---------------------------------------------------------------- kernel
template <typename T, int OP, int L, bool XMODE>
__global__ void __launch_bounds__(1024) line_kernel(const LineArgs<T> A) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *smem = reinterpret_cast<T *>(smem_raw);
  const int TL = blockDim.x, P = blockDim.y;
  const int line = threadIdx.x, seg = threadIdx.y;
  const int tid = seg * TL + line, nthr = TL * P;
  const int a = blockIdx.x * TL + line, b = blockIdx.y;
  const bool lv = a < A.na;
  const int n_in = A.n_in, n_out = A.n_out;
  const int e0 = seg * L;

  T *tile = smem;                                     // XMODE only: TL x row
  T *carry = smem + (XMODE ? (size_t)TL * A.row : 0); // P*TL
  T *gend = carry + P * TL;                           // P
  T *x0s = gend + P;                                  // TL
  T *xns = x0s + TL;                                  // TL

  // ---- stage x lines through LDS (contiguous tile of TL lines)
  if constexpr (XMODE) {
    const long long base = (long long)blockIdx.x * TL * A.sa_in + (long long)b * A.sb_in;
    const int nv = min(TL, A.na - (int)blockIdx.x * TL);
    const int tot = nv * n_in;
    for (int idx = tid; idx < tot; idx += nthr) {
      const int l = idx / n_in;
      const int e = idx - l * n_in;
      T x = A.u[base + idx];
      if (A.u2) x *= A.u2[base + idx];
      tile[l * A.row + e] = x;
    }
    __syncthreads();
  }

  const long long lbase_in = (long long)a * A.sa_in + (long long)b * A.sb_in;
  auto ldu = [&](int j) -> T {
    if constexpr (XMODE) {
      return tile[line * A.row + j];
    } else {
      const long long o = lbase_in + (long long)j * A.se_in;
      T x = A.u[o];
      if (A.u2) x *= A.u2[o];
      return x;
    }
  };

  // ---- window U(e0-4 .. e0+L+3) into registers
  T v[L + 8];
#pragma unroll
  for (int k = 0; k < L + 8; ++k) {
    int j = e0 - 4 + k;
    if (A.periodic) {
      if (j < 0) j += n_in;            // j >= -4 and n_in >= 8
      while (j >= n_in) j -= n_in;     // j < n_in + L + 4
    } else {
      j = j < 0 ? 0 : (j >= n_in ? n_in - 1 : j);
    }
    v[k] = (lv && e0 < n_out) ? ldu(j) : T(0);
  }

  // ---- explicit right-hand side, in place (row m only needs v[m..m+8])
#pragma unroll
  for (int m = 0; m < L; ++m) {
    const int e = e0 + m;
    T r = T(0);
    if (lv && e < n_out) {
      if (!A.periodic && is_boundary_row<OP>(e, n_out))
        r = boundary_row<OP, T>(e, n_in, n_out, A.s, A.bc1, A.bcn, A.npaire, ldu);
      else
        r = interior_row<OP, T, L + 8>(v, m + 4, A.s);
    }
    v[m] = r;
  }

  // ---- forward elimination, local then chained (derive.f90:45-48)
  T prev = T(0), g = T(1);
#pragma unroll
  for (int m = 0; m < L; ++m) {
    const int e = e0 + m;
    if (e < n_out) {
      const T f = A.fs[e];
      prev = v[m] - prev * f;
      v[m] = prev;
      g = -f * g;
    }
  }
  carry[seg * TL + line] = prev;
  if (line == 0) gend[seg] = g;
  __syncthreads();
  if (seg > 0) {
    T cin = T(0);
    for (int s = 0; s < seg; ++s) cin = carry[s * TL + line] + gend[s] * cin;
    T gg = T(1);
#pragma unroll
    for (int m = 0; m < L; ++m) {
      const int e = e0 + m;
      if (e < n_out) {
        gg = -A.fs[e] * gg;
        v[m] += gg * cin;
      }
    }
  }
  __syncthreads();

  // ---- back substitution, local then chained (derive.f90:49-54)
  T nxt = T(0), h = T(1);
#pragma unroll
  for (int m = L - 1; m >= 0; --m) {
    const int e = e0 + m;
    if (e < n_out) {
      const T w = A.fw[e], f = A.ff[e];
      nxt = (v[m] - f * nxt) * w;
      v[m] = nxt;
      h = -(f * w) * h;
    }
  }
  carry[seg * TL + line] = nxt;
  if (line == 0) gend[seg] = h;
  __syncthreads();
  if (seg < P - 1) {
    T c = T(0);
    for (int s = P - 1; s > seg; --s) c = carry[s * TL + line] + gend[s] * c;
    T hh = T(1);
#pragma unroll
    for (int m = L - 1; m >= 0; --m) {
      const int e = e0 + m;
      if (e < n_out) {
        hh = -(A.ff[e] * A.fw[e]) * hh;
        v[m] += hh * c;
      }
    }
  }

  // ---- Sherman-Morrison correction of periodic lines (derive.f90:55-59)
  if (A.periodic) {
    if (seg == 0) x0s[line] = v[0];
#pragma unroll
    for (int m = 0; m < L; ++m)
      if (e0 + m == n_out - 1) xns[line] = v[m];
    __syncthreads();
    const T sx = (x0s[line] - A.alfa * xns[line]) / A.sm_den;
#pragma unroll
    for (int m = 0; m < L; ++m) {
      const int e = e0 + m;
      if (e < n_out) v[m] = v[m] - sx * A.sm[e];
    }
  }
  if (A.pp) {
#pragma unroll
    for (int m = 0; m < L; ++m) {
      const int e = e0 + m;
      if (e < n_out) v[m] *= A.pp[e];
    }
  }

  // ---- store
  if constexpr (XMODE) {
    // every read of the tile happened before the first barrier above
#pragma unroll
    for (int m = 0; m < L; ++m) {
      const int e = e0 + m;
      if (e < n_out) tile[line * A.row + e] = v[m];
    }
    __syncthreads();
    const long long base = (long long)blockIdx.x * TL * A.sa_out + (long long)b * A.sb_out;
    const int nv = min(TL, A.na - (int)blockIdx.x * TL);
    const int tot = nv * n_out;
    for (int idx = tid; idx < tot; idx += nthr) {
      const int l = idx / n_out;
      const int e = idx - l * n_out;
      T r = tile[l * A.row + e];
      if (A.mul) r *= A.mul[base + idx];
      r = A.rc * r;
      if (A.acc) r = A.ac * A.acc[base + idx] + r;
      A.t[base + idx] = r;
    }
  } else {
    if (lv) {
      const long long lbase_out = (long long)a * A.sa_out + (long long)b * A.sb_out;
#pragma unroll
      for (int m = 0; m < L; ++m) {
        const int e = e0 + m;
        if (e < n_out) {
          const long long o = lbase_out + (long long)e * A.se_out;
          T r = v[m];
          if (A.mul) r *= A.mul[o];
          r = A.rc * r;
          if (A.acc) r = A.ac * A.acc[o] + r;
          A.t[o] = r;
        }
      }
    }
  }
}

// ---------------------------------------------------------------- launcher
template <typename T, int OP, int L>
static int launch_l(const LineArgs<T> &a, bool xmode, int TL, int P, dim3 grid, size_t shmem,
                    hipStream_t st) {
  dim3 block(TL, P);
  if (xmode)
    hipLaunchKernelGGL((line_kernel<T, OP, L, true>), grid, block, shmem, st, a);
  else
    hipLaunchKernelGGL((line_kernel<T, OP, L, false>), grid, block, shmem, st, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(X3D_ERR_HIP, std::string("line kernel launch: ") + hipGetErrorString(e));
  return X3D_OK;
}

template <typename T, int OP>
static int launch_op(const LineArgs<T> &a, bool xmode, int nmax, dim3 grid_base, hipStream_t st) {
  int L = nmax <= 128 ? 16 : (nmax <= 512 ? 32 : 64);
  int P = (nmax + L - 1) / L;
  if (nmax > 1024) return fail(X3D_ERR_UNSUPPORTED, "line length > 1024");
  int TL = xmode ? 16 : 64;
  if (TL * P > 1024) return fail(X3D_ERR_UNSUPPORTED, "line kernel block too large");
  dim3 grid((a.na + TL - 1) / TL, grid_base.y);
  size_t shmem = sizeof(T) * ((size_t)P * TL + P + 2 * TL);
  if (xmode) shmem += sizeof(T) * (size_t)TL * a.row;
  if (shmem > 160 * 1024) return fail(X3D_ERR_UNSUPPORTED, "line kernel LDS > 160 KB");
  switch (L) {
    case 16: return launch_l<T, OP, 16>(a, xmode, TL, P, grid, shmem, st);
    case 32: return launch_l<T, OP, 32>(a, xmode, TL, P, grid, shmem, st);
    default: return launch_l<T, OP, 64>(a, xmode, TL, P, grid, shmem, st);
  }
}

template <typename T>
int launch_line(const LineCall<T> &c, hipStream_t st) {
  if (!c.u || !c.t || !c.tri.ff || !c.tri.fs || !c.tri.fw) return fail(X3D_ERR_ARG, "null pointer");
  if (c.tri.periodic && !c.tri.sm) return fail(X3D_ERR_ARG, "periodic set without SM vector");
  LineArgs<T> a{};
  a.u = c.u; a.u2 = c.u2; a.t = c.t; a.pp = c.pp;
  a.acc = c.acc; a.mul = c.mul; a.ac = c.ac; a.rc = c.rc;
  a.ff = c.tri.ff; a.fs = c.tri.fs; a.fw = c.tri.fw; a.sm = c.tri.sm;
  a.alfa = c.tri.alfa; a.sm_den = c.tri.sm_den;
  a.s = c.s;
  a.periodic = c.tri.periodic; a.bc1 = c.bc1; a.bcn = c.bcn; a.npaire = c.npaire;
  const int nx = c.nx, ny = c.ny, nz = c.nz;
  const int n_in = c.dir == 0 ? nx : (c.dir == 1 ? ny : nz);
  const int n_out = c.n_out;
  if (c.tri.n != n_out) return fail(X3D_ERR_ARG, "coefficient length != output line length");
  if (n_in < 8 || n_out < 8) return fail(X3D_ERR_UNSUPPORTED, "lines shorter than 8 points");
  if (c.tri.periodic && n_in != n_out) return fail(X3D_ERR_ARG, "periodic line with n_in != n_out");
  a.n_in = n_in; a.n_out = n_out;
  dim3 g;
  bool xmode = false;
  if (c.dir == 0) {
    xmode = true;
    a.na = ny; a.nb = nz;
    a.se_in = 1; a.sa_in = n_in; a.sb_in = (long long)n_in * ny;
    a.se_out = 1; a.sa_out = n_out; a.sb_out = (long long)n_out * ny;
    int row = (n_in > n_out ? n_in : n_out);
    if (sizeof(T) == 8) row |= 1;          // odd stride in 8-byte words: no bank conflicts
    else row = (row % 2 == 0) ? row + 1 : row;
    a.row = row;
    g = dim3(1, nz);
  } else if (c.dir == 1) {
    a.na = nx; a.nb = nz;
    a.se_in = nx; a.sa_in = 1; a.sb_in = (long long)nx * n_in;
    a.se_out = nx; a.sa_out = 1; a.sb_out = (long long)nx * n_out;
    g = dim3(1, nz);
  } else {
    a.na = nx; a.nb = ny;
    a.se_in = (long long)nx * ny; a.sa_in = 1; a.sb_in = nx;
    a.se_out = (long long)nx * ny; a.sa_out = 1; a.sb_out = nx;
    g = dim3(1, ny);
  }
  const int nmax = n_in > n_out ? n_in : n_out;
  switch (c.op) {
    case OP_DER1: return launch_op<T, OP_DER1>(a, xmode, nmax, g, st);
    case OP_DER2: return launch_op<T, OP_DER2>(a, xmode, nmax, g, st);
    case OP_DERVP: return launch_op<T, OP_DERVP>(a, xmode, nmax, g, st);
    case OP_INTERVP: return launch_op<T, OP_INTERVP>(a, xmode, nmax, g, st);
    case OP_DERPV: return launch_op<T, OP_DERPV>(a, xmode, nmax, g, st);
    case OP_INTERPV: return launch_op<T, OP_INTERPV>(a, xmode, nmax, g, st);
  }
  return fail(X3D_ERR_ARG, "unknown line operator");
}

template int launch_line<double>(const LineCall<double> &, hipStream_t);
template int launch_line<float>(const LineCall<float> &, hipStream_t);

}  // namespace x3d
