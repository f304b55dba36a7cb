#pragma once
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
//  * y and z lines are strided in memory: the TL = 64 lines of a block are 64
//    consecutive i, so every load/store of a point row is 64 contiguous words
//    (512 B in FP64) and one wavefront holds one segment of 64 lines -- the
//    per-point coefficients are then wave-uniform and come in through scalar
//    loads;
//  * x lines are contiguous: the block stages TL whole lines (and the
//    coefficients) through LDS with fully contiguous loads/stores;
//  * the two first-order recurrences of the Thomas algorithm
//      y_i = r_i - fs_i y_{i-1},   x_i = fw_i y_i - (ff_i fw_i) x_{i+1}
//    are solved per segment with zero carry-in, then the carries are chained
//    across the P segments of a line through LDS and applied as
//    y_i += G_i c (G_i = prod(-fs)), x_i += H_i c (H_i = prod(-ff fw)).
//    |fs|,|ff fw| < 1 for these diagonally dominant systems, so the result
//    equals the sequential sweep to rounding;
//  * non-periodic boundary closures (at most 4 rows per end) are evaluated
//    once per line into LDS before the sweep, so the unrolled interior loop
//    carries no boundary code.
#include "x3d_internal.h"

namespace x3d {

template <typename T>
struct LineArgs {
  const T *u, *u2;
  T *t;
  const T *pp;
  const T *acc, *mul;                 // epilogue: t = ac*acc + rc*(mul*result)
  T ac, rc;
  const T *pk;                        // packed {fs, fw, ff*fw, r} per point
  T alfa, sm_den, rden;                // rden = 1 / sm_den
  Sten<T> s;
  int bc1, bcn, npaire;
  int n_in, n_out;
  int na, nb;                         // line grid extents (a fastest)
  long long sa_in, sb_in, se_in;      // input strides (elements)
  long long sa_out, sb_out, se_out;   // output strides
  int row;                            // LDS row stride (x mode)
};

// Window offsets an interior row needs: U(i-HL) .. U(i+HR).
template <int OP>
struct Halo;
template <> struct Halo<OP_DER1> { static constexpr int L = 2, R = 2; };
template <> struct Halo<OP_DER2> { static constexpr int L = 4, R = 4; };
template <> struct Halo<OP_DERVP> { static constexpr int L = 1, R = 2; };
template <> struct Halo<OP_INTERVP> { static constexpr int L = 3, R = 4; };
template <> struct Halo<OP_DERPV> { static constexpr int L = 2, R = 1; };
template <> struct Halo<OP_INTERPV> { static constexpr int L = 4, R = 3; };

// ---------------------------------------------------------------- stencils
// Interior rows: v[k] holds U(i) of output row i; offsets are compile-time.
// D2S selects the shared-second-difference form of the second derivative
// (the direct four-term form needs fewer registers where a kernel is at its
// register budget).
template <int OP, typename T, int W, bool D2S = true>
__device__ __forceinline__ T interior_row(const T (&v)[W], const int k, const Sten<T> &s) {
  if constexpr (OP == OP_DER1) {
    // derive.f90:34-36
    return s.afi * (v[k + 1] - v[k - 1]) + s.bfi * (v[k + 2] - v[k - 2]);
  } else if constexpr (OP == OP_DER2 && !D2S) {
    // derive.f90:1414-1421
    return s.asi * (v[k + 1] - v[k] - v[k] + v[k - 1]) + s.bsi * (v[k + 2] - v[k] - v[k] + v[k - 2]) +
           s.csi * (v[k + 3] - v[k] - v[k] + v[k - 3]) + s.dsi * (v[k + 4] - v[k] - v[k] + v[k - 4]);
  } else if constexpr (OP == OP_DER2) {
    // derive.f90:1414-1421, sum_j c_j (U(i+j) - 2U(i) + U(i-j)), written on
    // the second differences D(i) = (U(i+1) - U(i)) - (U(i) - U(i-1)):
    // U(i+j) - 2U(i) + U(i-j) = sum_{|q|<j} (j - |q|) D(i+q), so the row is
    // W0 D(i) + W1 (D(i-1) + D(i+1)) + W2 (..) + W3 (..) with W0 = c1 + 2c2 +
    // 3c3 + 4c4, W1 = c2 + 2c3 + 3c4, W2 = c3 + 2c4, W3 = c4.  The first and
    // second differences are shared by neighbouring rows of a window (9
    // operations a row instead of 16), and the sums are of small numbers.
    auto d1 = [&](int q) { return v[k + q + 1] - v[k + q]; };
    auto d2 = [&](int q) { return d1(q) - d1(q - 1); };
    const T w3 = s.dsi, w2 = s.csi + T(2) * s.dsi, w1 = s.bsi + T(2) * s.csi + T(3) * s.dsi,
            w0 = s.asi + T(2) * s.bsi + T(3) * s.csi + T(4) * s.dsi;
    return w0 * d2(0) + w1 * (d2(-1) + d2(1)) + w2 * (d2(-2) + d2(2)) + w3 * (d2(-3) + d2(3));
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

// Rows of a non-periodic line that use a boundary closure (<= 4 per end).
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

// ---------------------------------------------------------------- y / z segment solve
// Block: TL consecutive i (one row of TL words per line point) x P segments
// of L points; a thread owns one segment of one line.  The two recurrences
// of each solve run locally per segment with a zero carry-in; the carries
// are then chained through LDS.  The chain stops K segments back: the factor
// a carry picks up across a segment is the product of that segment's
// elimination factors, and K is chosen on the host so that the product over
// any K consecutive segments is below the precision's resolution (the terms
// dropped are smaller than the rounding of the full chain).
//
// N right-hand sides: the first N1 belong to matrix 1 (table c1, periodic
// closure alfa1 and 1/den1), the rest to matrix 2.  Every phase runs on all
// of them between the same three barriers, so the operators of one sweep
// direction share one set of barriers.  A segment's carry-in reads the
// carries of the K segments before it (after it, backward) four at a time,
// the LDS reads ahead of the multiply-add chain.
//
// LDS: cf and cb hold N*P*TL carries each, gf and gb 2*P segment factors,
// x0s and xns N*TL line ends.
template <typename T, int L, int N, int N1 = N, bool PER = true>
__device__ __forceinline__ void seg_solve(T (&v)[N][L], const int e0, const T *__restrict__ c1, const T alfa1,
                                          const T rden1, const T *__restrict__ c2, const T alfa2, const T rden2,
                                          T *cf, T *cb, T *gf, T *gb, T *x0s, T *xns, const int line,
                                          const int seg, const int TL, const int P, const int K) {
  static_assert(N1 >= 1 && N1 <= N, "matrix 1 needs a right-hand side");
  constexpr bool TWO = N1 < N;
  const int PT = P * TL;
  // forward elimination, zero carry-in
  T prev[N], g1 = T(1), g2 = T(1);
#pragma unroll
  for (int q = 0; q < N; ++q) prev[q] = T(0);
#pragma unroll
  for (int m = 0; m < L; ++m) {
    const T f1 = c1[4 * (e0 + m)];
    const T f2 = TWO ? c2[4 * (e0 + m)] : T(0);
#pragma unroll
    for (int q = 0; q < N; ++q) {
      prev[q] = v[q][m] - prev[q] * (q < N1 ? f1 : f2);
      v[q][m] = prev[q];
    }
    g1 = -f1 * g1;
    if (TWO) g2 = -f2 * g2;
  }
#pragma unroll
  for (int q = 0; q < N; ++q) cf[q * PT + seg * TL + line] = prev[q];
  if (line == 0) {
    gf[seg] = g1;
    if (TWO) gf[P + seg] = g2;
  }
  __syncthreads();
  if (seg > 0) {
    T cin[N];
#pragma unroll
    for (int q = 0; q < N; ++q) cin[q] = T(0);
    for (int s0 = seg - K > 0 ? seg - K : 0; s0 < seg; s0 += 4) {
      T gs1[4], gs2[4], cv[4][N];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int s = s0 + j < seg ? s0 + j : seg - 1;
        gs1[j] = gf[s];
        gs2[j] = TWO ? gf[P + s] : T(0);
#pragma unroll
        for (int q = 0; q < N; ++q) cv[j][q] = cf[q * PT + s * TL + line];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (s0 + j < seg) {
#pragma unroll
          for (int q = 0; q < N; ++q) cin[q] = cv[j][q] + (q < N1 ? gs1[j] : gs2[j]) * cin[q];
        }
      }
    }
    T gg1 = T(1), gg2 = T(1);
#pragma unroll
    for (int m = 0; m < L; ++m) {
      gg1 = -c1[4 * (e0 + m)] * gg1;
      if (TWO) gg2 = -c2[4 * (e0 + m)] * gg2;
#pragma unroll
      for (int q = 0; q < N; ++q) v[q][m] += (q < N1 ? gg1 : gg2) * cin[q];
    }
  }
  // back substitution, zero carry-in
  T nxt[N], h1 = T(1), h2 = T(1);
#pragma unroll
  for (int q = 0; q < N; ++q) nxt[q] = T(0);
#pragma unroll
  for (int m = L - 1; m >= 0; --m) {
    const T w1 = c1[4 * (e0 + m) + 1], ffw1 = c1[4 * (e0 + m) + 2];
    const T w2 = TWO ? c2[4 * (e0 + m) + 1] : T(0), ffw2 = TWO ? c2[4 * (e0 + m) + 2] : T(0);
#pragma unroll
    for (int q = 0; q < N; ++q) {
      nxt[q] = (q < N1 ? w1 : w2) * v[q][m] - (q < N1 ? ffw1 : ffw2) * nxt[q];
      v[q][m] = nxt[q];
    }
    h1 = -ffw1 * h1;
    if (TWO) h2 = -ffw2 * h2;
  }
#pragma unroll
  for (int q = 0; q < N; ++q) cb[q * PT + seg * TL + line] = nxt[q];
  if (line == 0) {
    gb[seg] = h1;
    if (TWO) gb[P + seg] = h2;
  }
  __syncthreads();
  if (seg < P - 1) {
    T c[N];
#pragma unroll
    for (int q = 0; q < N; ++q) c[q] = T(0);
    const int last = seg + K < P - 1 ? seg + K : P - 1;
    for (int s0 = last; s0 > seg; s0 -= 4) {
      T gs1[4], gs2[4], cv[4][N];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int s = s0 - j > seg ? s0 - j : seg + 1;
        gs1[j] = gb[s];
        gs2[j] = TWO ? gb[P + s] : T(0);
#pragma unroll
        for (int q = 0; q < N; ++q) cv[j][q] = cb[q * PT + s * TL + line];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (s0 - j > seg) {
#pragma unroll
          for (int q = 0; q < N; ++q) c[q] = cv[j][q] + (q < N1 ? gs1[j] : gs2[j]) * c[q];
        }
      }
    }
    T hh1 = T(1), hh2 = T(1);
#pragma unroll
    for (int m = L - 1; m >= 0; --m) {
      hh1 = -c1[4 * (e0 + m) + 2] * hh1;
      if (TWO) hh2 = -c2[4 * (e0 + m) + 2] * hh2;
#pragma unroll
      for (int q = 0; q < N; ++q) v[q][m] += (q < N1 ? hh1 : hh2) * c[q];
    }
  }
  // periodic closure (Sherman-Morrison with the precomputed vector sm);
  // a non-periodic line (PER false) is solved at this point
  if constexpr (!PER) return;
#pragma unroll
  for (int q = 0; q < N; ++q) {
    if (seg == 0) x0s[q * TL + line] = v[q][0];
    if (seg == P - 1) xns[q * TL + line] = v[q][L - 1];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < N; ++q) {
    const T sx = (x0s[q * TL + line] - (q < N1 ? alfa1 : alfa2) * xns[q * TL + line]) * (q < N1 ? rden1 : rden2);
    const T *cq = q < N1 ? c1 : c2;
#pragma unroll
    for (int m = 0; m < L; ++m) v[q][m] = v[q][m] - sx * cq[4 * (e0 + m) + 3];
  }
  // No barrier needed before the next solve: it writes cf/gf before its
  // first barrier (last read here before the second), cb/gb after its first
  // barrier (last read here before the third) and x0s/xns after its second.
}

// Copy of two coefficient tables (cnt words each) into LDS split around the
// caller's window loads: load() issues the first R*nthr words of each into
// registers, store() writes them and copies any rest, so the table and window
// latencies overlap before the barrier that must follow store().
template <typename T, int R>
struct TableCopy {
  T v1[R], v2[R];
  __device__ __forceinline__ void load(const T *a, const T *b, int tid, int nthr, int cnt) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = tid + r * nthr;
      if (i < cnt) {
        v1[r] = a[i];
        v2[r] = b[i];
      }
    }
  }
  __device__ __forceinline__ void store(T *da, T *db, const T *a, const T *b, int tid, int nthr, int cnt) const {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = tid + r * nthr;
      if (i < cnt) {
        da[i] = v1[r];
        db[i] = v2[r];
      }
    }
    for (int i = tid + R * nthr; i < cnt; i += nthr) {
      da[i] = a[i];
      db[i] = b[i];
    }
  }
};

// LDS words of a segmented sweep: nmat coefficient tables, the carries and
// line ends of nrhs right-hand sides (the layout of seg_solve).
inline size_t seg_lds_words(int n, int P, int TL, int nmat, int nrhs) {
  return 4 * (size_t)nmat * n + 2 * (size_t)nrhs * P * TL + 2 * (size_t)nmat * P + 2 * (size_t)nrhs * TL;
}

template <typename T>
__device__ __forceinline__ T &at(T *p, unsigned byte_off) {
  return *reinterpret_cast<T *>(reinterpret_cast<char *>(p) + byte_off);
}
template <typename T>
__device__ __forceinline__ const T &at(const T *p, unsigned byte_off) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + byte_off);
}

// Non-temporal (streaming) store at a byte offset, for output fields that
// the kernel does not read again; which kernels use it was settled by
// measurement (the segmented y/z derivative and gradient sweeps gained,
// the momentum and divergence sweeps and the x lines lost).
template <typename T>
__device__ __forceinline__ void st_nt(T *p, unsigned byte_off, T v) {
  __builtin_nontemporal_store(v, reinterpret_cast<T *>(reinterpret_cast<char *>(p) + byte_off));
}

// ---------------------------------------------------------------- x wavefront solve
// floor(a / b) for b > 0 (compile-time after unrolling)
constexpr int floor_div(int a, int b) { return (a >= 0) ? a / b : -((-a + b - 1) / b); }

// ---------------------------------------------------------------- cross-lane moves
// DPP moves (VALU, no LDS round trip) of 32- and 64-bit values.  CTRL is the
// DPP control: 0x101-0x10F row_shl (lane l reads lane l+k of its row of 16),
// 0x111-0x11F row_shr (lane l reads l-k), 0x130 wave_shl:1, 0x134 wave_rol:1,
// 0x138 wave_shr:1, 0x13C wave_ror:1.  Lanes without a source keep `old`.
template <int CTRL>
__device__ __forceinline__ float dpp(float x, float old) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(x), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double x, double old) {
  const long long xi = __double_as_longlong(x), oi = __double_as_longlong(old);
  const int lo = __builtin_amdgcn_update_dpp((int)oi, (int)xi, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(oi >> 32), (int)(xi >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float readlane(float x, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double readlane(double x, int l) {
  const long long xi = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)xi, l), hi = __builtin_amdgcn_readlane((int)(xi >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <typename T, int V, int HL, int HR>
__device__ __forceinline__ void halo_window(const T (&own)[V], T (&v)[V + HL + HR], const int lane) {
#pragma unroll
  for (int k = 0; k < V + HL + HR; ++k) {
    const int d = k - HL;
    if (d >= 0 && d < V) {
      v[k] = own[d];
    } else {
      const int lo = floor_div(d, V);
      const int el = d - lo * V;
      if (lo == -1) v[k] = dpp<0x13C>(own[el], T(0));        // wave_ror:1, lane l reads l-1 (0 reads 63)
      else if (lo == 1) v[k] = dpp<0x134>(own[el], T(0));    // wave_rol:1, lane l reads l+1 (63 reads 0)
      else v[k] = __shfl(own[el], (lane + lo) & 63);
    }
  }
}

// Composition of the affine maps x -> a x + b of the 64 lanes of a
// wavefront.  scan_up leaves in lane l the maps of lanes 0..l applied in lane
// order, scan_down the maps of lanes l..63 applied from lane 63 down: four
// DPP steps within each row of 16 lanes, then the row totals (read with
// readlane) carried across the four rows.
template <typename T>
__device__ __forceinline__ void scan_up(T &a, T &b, const int lane) {
#define X3D_UP(S)                                                                                                      \
  {                                                                                                                    \
    const T ap = dpp<0x110 + S>(a, T(1)), bp = dpp<0x110 + S>(b, T(0));                                               \
    b = a * bp + b;                                                                                                    \
    a = a * ap;                                                                                                        \
  }
  X3D_UP(1) X3D_UP(2) X3D_UP(4) X3D_UP(8)
#undef X3D_UP
  const T a0 = readlane(a, 15), b0 = readlane(b, 15);
  const T a1 = readlane(a, 31), b1 = readlane(b, 31);
  const T a2 = readlane(a, 47), b2 = readlane(b, 47);
  // carry into rows 1, 2, 3: rows 0..r-1 composed
  const T c2a = a1 * a0, c2b = a1 * b0 + b1;
  const T c3a = a2 * c2a, c3b = a2 * c2b + b2;
  const int row = lane >> 4;
  const T ca = row == 0 ? T(1) : (row == 1 ? a0 : (row == 2 ? c2a : c3a));
  const T cb = row == 0 ? T(0) : (row == 1 ? b0 : (row == 2 ? c2b : c3b));
  b = a * cb + b;
  a = a * ca;
}
template <typename T>
__device__ __forceinline__ void scan_down(T &a, T &b, const int lane) {
#define X3D_DN(S)                                                                                                      \
  {                                                                                                                    \
    const T an = dpp<0x100 + S>(a, T(1)), bn = dpp<0x100 + S>(b, T(0));                                               \
    b = a * bn + b;                                                                                                    \
    a = a * an;                                                                                                        \
  }
  X3D_DN(1) X3D_DN(2) X3D_DN(4) X3D_DN(8)
#undef X3D_DN
  const T a1 = readlane(a, 16), b1 = readlane(b, 16);
  const T a2 = readlane(a, 32), b2 = readlane(b, 32);
  const T a3 = readlane(a, 48), b3 = readlane(b, 48);
  // carry into rows 2, 1, 0: rows r+1..3 composed
  const T c1a = a2 * a3, c1b = a2 * b3 + b2;
  const T c0a = a1 * c1a, c0b = a1 * c1b + b1;
  const int row = lane >> 4;
  const T ca = row == 3 ? T(1) : (row == 2 ? a3 : (row == 1 ? c1a : c0a));
  const T cb = row == 3 ? T(0) : (row == 2 ? b3 : (row == 1 ? c1b : c0b));
  b = a * cb + b;
  a = a * ca;
}

// Periodic compact solve of one line held by a wavefront, V points per lane
// (lane k owns points kV .. kV+V-1): local Thomas sweeps, carries across the
// lanes by the affine-map scans, Sherman-Morrison closure (derive.f90:55-59)
// with rden = 1 / den.
template <typename T, int V>
__device__ __forceinline__ void wave_solve(T (&v)[V], const T (&fs)[V], const T (&fw)[V], const T (&ffw)[V],
                                           const T (&sm)[V], const T alfa, const T rden, const int lane) {
  T prev = T(0), g = T(1);
#pragma unroll
  for (int m = 0; m < V; ++m) {
    prev = v[m] - prev * fs[m];
    v[m] = prev;
    g = -fs[m] * g;
  }
  {
    T a = g, bb = prev;
    scan_up(a, bb, lane);
    const T cin = dpp<0x138>(bb, T(0));   // wave_shr:1, lane 0 reads 0
    T gg = T(1);
#pragma unroll
    for (int m = 0; m < V; ++m) {
      gg = -fs[m] * gg;
      v[m] += gg * cin;
    }
  }
  T nxt = T(0), h = T(1);
#pragma unroll
  for (int m = V - 1; m >= 0; --m) {
    nxt = fw[m] * v[m] - ffw[m] * nxt;
    v[m] = nxt;
    h = -ffw[m] * h;
  }
  {
    T a = h, bb = nxt;
    scan_down(a, bb, lane);
    const T c = dpp<0x130>(bb, T(0));     // wave_shl:1, lane 63 reads 0
    T hh = T(1);
#pragma unroll
    for (int m = V - 1; m >= 0; --m) {
      hh = -ffw[m] * hh;
      v[m] += hh * c;
    }
  }
  const T x0 = readlane(v[0], 0), xn = readlane(v[V - 1], 63);
  const T sx = (x0 - alfa * xn) * rden;
#pragma unroll
  for (int m = 0; m < V; ++m) v[m] = v[m] - sx * sm[m];
}

// Block coordinates (line tile, plane) of a y/z sweep.  Workgroups are handed
// to the eight XCDs in turn (id mod 8), so with a grid eight tiles wide each
// XCD would read one 128-byte column of every row.  With XCD set (and a
// block count divisible by 8) the ids are renumbered so that each XCD works
// through one contiguous eighth of the (tile, plane) space: whole rows of
// the field pass through one L2.
template <bool XCD>
__device__ __forceinline__ void block_coords(int &bx, int &by) {
  const int gx = (int)gridDim.x, nblk = gx * (int)gridDim.y;
  if (!XCD || (nblk & 7) != 0) {
    bx = (int)blockIdx.x;
    by = (int)blockIdx.y;
    return;
  }
  int l = (int)blockIdx.x + gx * (int)blockIdx.y;
  l = (l & 7) * (nblk >> 3) + (l >> 3);
  by = l / gx;
  bx = l - by * gx;
}

// Launch shape of a segmented y/z sweep for a line of n points: segments of
// L = 8 (FP64) or 16 (FP32) points, so that the fused sweeps' three solves
// keep their windows, right-hand sides and results within 128 VGPRs; TL
// lines per block with TL*P <= 1024 threads and the LDS (seg_lds_words)
// within 64 KiB.
template <typename T>
struct YzShape {
  static constexpr int L = sizeof(T) == 8 ? SEG_L64 : SEG_L32;
  int P = 0, TL = 0;
  size_t shmem = 0;
  bool plan(int n, int nmat, int nrhs) {
    if (n < 8 || n % L != 0) return false;
    P = n / L;
    // 128-byte rows (16 FP64 / 32 FP32 lines): full cache lines per point,
    // and at 256-point lines 512-thread blocks, two of which share a CU
    TL = 128 / (int)sizeof(T);
    while (TL >= 8 && (TL * P > 1024 || sizeof(T) * seg_lds_words(n, P, TL, nmat, nrhs) > 65536)) TL >>= 1;
    if (TL < 8) return false;
    shmem = sizeof(T) * seg_lds_words(n, P, TL, nmat, nrhs);
    return true;
  }
};

// ---------------------------------------------------------------- kernel
// AX: 0 x (LDS-staged), 1 y, 2 z (direct); EPI: fused product input and
// epilogue present (a plain derivative compiles without them).
// LB: the largest block the instantiation is launched with (the register
// budget follows from it).
template <typename T, int OP, int L, int AX, bool PERIODIC, bool EPI, int LB = 1024>
__global__ void __launch_bounds__(LB) line_kernel(const LineArgs<T> A) {
  constexpr bool XMODE = AX == 0;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *smem = reinterpret_cast<T *>(smem_raw);
  constexpr int HL = Halo<OP>::L, HR = Halo<OP>::R, W = L + HL + HR;
  const int TL = blockDim.x, P = blockDim.y;
  const int line = threadIdx.x, seg = threadIdx.y;
  const int tid = seg * TL + line, nthr = TL * P;
  const int a = blockIdx.x * TL + line, b = blockIdx.y;
  const bool lv = a < A.na;
  const int n_in = A.n_in, n_out = A.n_out;
  // seg is wave-uniform when TL == 64 (y/z mode): make the compiler see it
  const int e0 = (XMODE ? seg : __builtin_amdgcn_readfirstlane(seg)) * L;

  T *tile = smem;                                          // XMODE: TL x row
  T *cs = tile + (XMODE ? (size_t)TL * A.row : 0);         // XMODE: 4*n_out coefficients
  T *bnd = cs + (XMODE ? 4 * n_out : 0);                   // !PERIODIC: TL x 8 boundary rows
  T *carry = bnd + (PERIODIC ? 0 : 8 * TL);                // P*TL
  T *gend = carry + P * TL;                                // P
  T *x0s = gend + P;                                       // TL
  T *xns = x0s + TL;                                       // TL

  // ---- x lines (and coefficients) through LDS
  if constexpr (XMODE) {
    const long long base = (long long)blockIdx.x * TL * A.sa_in + (long long)b * A.sb_in;
    const int nv = min(TL, A.na - (int)blockIdx.x * TL);
    const int tot = nv * n_in;
    for (int idx = tid; idx < tot; idx += nthr) {
      const int l = idx / n_in;
      const int e = idx - l * n_in;
      T x = A.u[base + idx];
      if (EPI && A.u2) x *= A.u2[base + idx];
      tile[l * A.row + e] = x;
    }
    for (int idx = tid; idx < 4 * n_out; idx += nthr) cs[idx] = A.pk[idx];
    __syncthreads();
  }
  auto cf = [&](int e, int j) -> T {
    if constexpr (XMODE) return cs[4 * e + j];
    else return A.pk[4 * e + j];
  };

  const long long lbase_in = (long long)a * A.sa_in + (long long)b * A.sb_in;
  auto ldu = [&](int j) -> T {
    if constexpr (XMODE) {
      return tile[line * A.row + j];
    } else {
      const long long o = lbase_in + (long long)j * A.se_in;
      T x = A.u[o];
      if (EPI && A.u2) x *= A.u2[o];
      return x;
    }
  };

  // ---- boundary closures, once per line, into LDS
  if constexpr (!PERIODIC) {
    for (int q = seg; q < 8; q += P) {
      const int r = q < 4 ? q : n_out - 8 + q;
      if (lv && r >= 0 && r < n_out && is_boundary_row<OP>(r, n_out))
        bnd[line * 8 + q] = boundary_row<OP, T>(r, n_in, n_out, A.s, A.bc1, A.bcn, A.npaire, ldu);
    }
    __syncthreads();
  }

  // ---- window U(e0-HL .. e0+L+HR-1) into registers
  T v[W];
  const bool active = lv && e0 < n_out;
#pragma unroll
  for (int k = 0; k < W; ++k) {
    int j = e0 - HL + k;
    if constexpr (PERIODIC) {
      if (j < 0) j += n_in;            // j >= -4 and n_in >= 8
      while (j >= n_in) j -= n_in;     // j < n_in + L + 4
    } else {
      j = j < 0 ? 0 : (j >= n_in ? n_in - 1 : j);
    }
    v[k] = active ? ldu(j) : T(0);
  }

  // ---- explicit right-hand side, in place (row m only needs v[m..m+HL+HR])
#pragma unroll
  for (int m = 0; m < L; ++m) {
    const int e = e0 + m;
    T r = T(0);
    if (lv && e < n_out) {
      if constexpr (PERIODIC) {
        r = interior_row<OP, T, W>(v, m + HL, A.s);
      } else {
        if (is_boundary_row<OP>(e, n_out))
          r = bnd[line * 8 + (e < 4 ? e : e - (n_out - 8))];
        else
          r = interior_row<OP, T, W>(v, m + HL, A.s);
      }
    }
    v[m] = r;
  }

  // ---- forward elimination, local then chained (derive.f90:45-48)
  T prev = T(0), g = T(1);
#pragma unroll
  for (int m = 0; m < L; ++m) {
    const int e = e0 + m;
    if (e < n_out) {
      const T f = cf(e, 0);
      prev = v[m] - prev * f;
      v[m] = prev;
      g = -f * g;
    }
  }
  if (P > 1) {
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
          gg = -cf(e, 0) * gg;
          v[m] += gg * cin;
        }
      }
    }
    __syncthreads();
  }

  // ---- back substitution, local then chained (derive.f90:49-54)
  T nxt = T(0), h = T(1);
#pragma unroll
  for (int m = L - 1; m >= 0; --m) {
    const int e = e0 + m;
    if (e < n_out) {
      const T w = cf(e, 1), fw = cf(e, 2);
      nxt = w * v[m] - fw * nxt;
      v[m] = nxt;
      h = -fw * h;
    }
  }
  if (P > 1) {
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
          hh = -cf(e, 2) * hh;
          v[m] += hh * c;
        }
      }
    }
  }

  // ---- Sherman-Morrison correction of periodic lines (derive.f90:55-59)
  if constexpr (PERIODIC) {
    if (seg == 0) x0s[line] = v[0];
#pragma unroll
    for (int m = 0; m < L; ++m)
      if (e0 + m == n_out - 1) xns[line] = v[m];
    __syncthreads();
    const T sx = (x0s[line] - A.alfa * xns[line]) / A.sm_den;
#pragma unroll
    for (int m = 0; m < L; ++m) {
      const int e = e0 + m;
      if (e < n_out) v[m] = v[m] - sx * cf(e, 3);
    }
  }
  if (A.pp) {
#pragma unroll
    for (int m = 0; m < L; ++m) {
      const int e = e0 + m;
      if (e < n_out) v[m] *= A.pp[e];
    }
  }

  // ---- store (with the epilogue t = ac*acc + rc*(mul*result))
  if constexpr (XMODE) {
    __syncthreads();   // every read of the tile is done
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
      if (EPI && A.mul) r *= A.mul[base + idx];
      if (EPI) r = A.rc * r;
      if (EPI && A.acc) r = A.ac * A.acc[base + idx] + r;
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
          if (EPI && A.mul) r *= A.mul[o];
          if (EPI) r = A.rc * r;
          if (EPI && A.acc) r = A.ac * A.acc[o] + r;
          A.t[o] = r;
        }
      }
    }
  }
}

// ---------------------------------------------------------------- periodic y/z lines, segmented
// One operator on periodic y or z lines (n_in == n_out == P*L): block of TL
// lines (128-byte rows) x P segments of L points, every input read once into
// a window, the solve of seg_solve with the set's truncated carry chain, and
// the epilogue's inputs read before the first store (acc may be t).  Byte
// offsets are 32-bit (the host checks the field size).
template <typename T, int OP, int L, bool EPI>
__global__ void __launch_bounds__(1024) seg_line_kernel(const LineArgs<T> A, const T rden, const int kdepth) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *smem = reinterpret_cast<T *>(smem_raw);
  const int TL = blockDim.x, P = blockDim.y;
  const int line = threadIdx.x, seg = threadIdx.y;
  const int tid = seg * TL + line, nthr = TL * P;
  const int n = A.n_out;
  T *cs = smem, *cf = cs + 4 * n, *cb = cf + P * TL, *gf = cb + P * TL, *gb = gf + P, *x0s = gb + P,
    *xns = x0s + TL;
  const int a = blockIdx.x * TL + line, b = blockIdx.y;
  const bool active = a < A.na;
  // lines past the end compute on a copy of the last line (no load branches)
  const int al = active ? a : A.na - 1;
  const int e0 = seg * L;
  const unsigned base = ((unsigned)al * (unsigned)A.sa_in + (unsigned)b * (unsigned)A.sb_in) * (unsigned)sizeof(T);
  const unsigned step = (unsigned)A.se_in * (unsigned)sizeof(T);
  auto wrap = [&](int j) { return (unsigned)(j < 0 ? j + n : (j >= n ? j - n : j)); };
  // the window loads go out before the coefficient copy and its barrier; in
  // FP32 at raised wave priority (512^3 y/z derivatives +2-5 % of peak; the
  // FP64 z lines lost 6 % with it)
  if constexpr (sizeof(T) == 4) __builtin_amdgcn_s_setprio(3);
  T win[L + 8];
#pragma unroll
  for (int k = 0; k < L + 8; ++k) {
    const unsigned o = base + wrap(e0 - 4 + k) * step;
    T x = at(A.u, o);
    if (EPI && A.u2) x *= at(A.u2, o);
    win[k] = x;
  }
  for (int i = tid; i < 4 * n; i += nthr) cs[i] = A.pk[i];
  if constexpr (sizeof(T) == 4) __builtin_amdgcn_s_setprio(0);
  __syncthreads();
  T w[1][L];
#pragma unroll
  for (int m = 0; m < L; ++m) w[0][m] = interior_row<OP, T, L + 8>(win, m + 4, A.s);
  seg_solve<T, L, 1>(w, e0, cs, A.alfa, rden, cs, A.alfa, rden, cf, cb, gf, gb, x0s, xns, line, seg, TL, P,
                     kdepth);
  if (!active) return;
  if (A.pp) {
#pragma unroll
    for (int m = 0; m < L; ++m) w[0][m] *= A.pp[e0 + m];
  }
  if constexpr (EPI) {
    T mm[L], aa[L];
#pragma unroll
    for (int m = 0; m < L; ++m) {
      const unsigned o = base + (unsigned)(e0 + m) * step;
      mm[m] = A.mul ? at(A.mul, o) : T(1);
      aa[m] = A.acc ? at(A.acc, o) : T(0);
    }
#pragma unroll
    for (int m = 0; m < L; ++m) {
      T r = w[0][m];
      if (A.mul) r *= mm[m];
      r = A.rc * r;
      if (A.acc) r = A.ac * aa[m] + r;
      st_nt(A.t, base + (unsigned)(e0 + m) * step, r);
    }
  } else {
#pragma unroll
    for (int m = 0; m < L; ++m) st_nt(A.t, base + (unsigned)(e0 + m) * step, w[0][m]);
  }
}

template <typename T, int OP>
static int launch_seg_line(const LineArgs<T> &a, T rden, const YzShape<T> &Y, bool epi, int kdepth,
                           hipStream_t st) {
  // the segment solve needs one line per block row and every segment full
  const dim3 grid((a.na + Y.TL - 1) / Y.TL, a.nb), block(Y.TL, Y.P);
  const size_t shmem = Y.shmem;
  constexpr int L = YzShape<T>::L;
  if (epi) hipLaunchKernelGGL((seg_line_kernel<T, OP, L, true>), grid, block, shmem, st, a, rden, kdepth);
  else hipLaunchKernelGGL((seg_line_kernel<T, OP, L, false>), grid, block, shmem, st, a, rden, kdepth);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(X3D_ERR_HIP, std::string("segmented line launch: ") + hipGetErrorString(e));
  return X3D_OK;
}

// ---------------------------------------------------------------- non-periodic y/z lines, segmented
// One operator on y or z lines with free-slip or wall faces (the _11 _12 _21
// _22 variants of derive.f90): the layout and solve of seg_line_kernel (TL
// lines of 128-byte rows x P segments of L points, the set's truncated carry
// chain), with the window clamped at the line ends, the boundary-closure
// rows (at most four per end) evaluated once per line into LDS before the
// barrier, and no periodic closure.  The last segment may be partial: the
// coefficient table is padded in LDS up to P*L with decoupled rows (zero
// multipliers and a zero right-hand side), which carry nothing across.
// Input and output strides differ for the staggered operators (n_in !=
// n_out); byte offsets are 32-bit (the host checks both field sizes).
template <typename T, int OP, int L, bool EPI, int LB, int WPE = 1>
__global__ void __launch_bounds__(LB, WPE) seg_wall_kernel(const LineArgs<T> A, const int kdepth) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *smem = reinterpret_cast<T *>(smem_raw);
  const int TL = blockDim.x, P = blockDim.y;
  const int line = threadIdx.x, seg = threadIdx.y;
  const int tid = seg * TL + line, nthr = TL * P;
  const int n_in = A.n_in, n_out = A.n_out, np = P * L;
  T *cs = smem, *cf = cs + 4 * np, *cb = cf + P * TL, *gf = cb + P * TL, *gb = gf + P, *bnd = gb + P;
  const int a = blockIdx.x * TL + line, b = blockIdx.y;
  const bool active = a < A.na;
  const int al = active ? a : A.na - 1;
  const int e0 = seg * L;
  const unsigned ibase = ((unsigned)al * (unsigned)A.sa_in + (unsigned)b * (unsigned)A.sb_in) * (unsigned)sizeof(T);
  const unsigned istep = (unsigned)A.se_in * (unsigned)sizeof(T);
  auto ld = [&](int j) -> T {
    const unsigned o = ibase + (unsigned)j * istep;
    T x = at(A.u, o);
    if (EPI && A.u2) x *= at(A.u2, o);
    return x;
  };
  // boundary rows 0..3 and n_out-4..n_out-1 of this line (the threads of
  // the first eight segments), before the window is live
  for (int q = seg; q < 8; q += P) {
    const int r = q < 4 ? q : n_out - 8 + q;
    if (is_boundary_row<OP>(r, n_out))
      bnd[line * 8 + q] = boundary_row<OP, T>(r, n_in, n_out, A.s, A.bc1, A.bcn, A.npaire, ld);
  }
  T win[L + 8];
#pragma unroll
  for (int k = 0; k < L + 8; ++k) {
    const int j = e0 - 4 + k;
    win[k] = ld(j < 0 ? 0 : (j >= n_in ? n_in - 1 : j));
  }
  for (int i = tid; i < 4 * np; i += nthr) cs[i] = i < 4 * n_out ? A.pk[i] : T(0);
  __syncthreads();
  // the FP64 9-point segments with the 64-register budget: direct form
  constexpr bool D2S = !(sizeof(T) == 8 && L == SEG_L64 + 1 && WPE > 1);
  T w[1][L];
#pragma unroll
  for (int m = 0; m < L; ++m) {
    const int e = e0 + m;
    T r = T(0);
    if (e < n_out)
      r = is_boundary_row<OP>(e, n_out) ? bnd[line * 8 + (e < 4 ? e : e - (n_out - 8))]
                                        : interior_row<OP, T, L + 8, D2S>(win, m + 4, A.s);
    w[0][m] = r;
  }
  seg_solve<T, L, 1, 1, false>(w, e0, cs, A.alfa, T(0), cs, A.alfa, T(0), cf, cb, gf, gb, nullptr, nullptr, line,
                               seg, TL, P, kdepth);
  if (!active) return;
  const unsigned obase =
      ((unsigned)al * (unsigned)A.sa_out + (unsigned)b * (unsigned)A.sb_out) * (unsigned)sizeof(T);
  const unsigned ostep = (unsigned)A.se_out * (unsigned)sizeof(T);
  if (A.pp) {
#pragma unroll
    for (int m = 0; m < L; ++m)
      if (e0 + m < n_out) w[0][m] *= A.pp[e0 + m];
  }
  if constexpr (EPI) {
    T mm[L], aa[L];
#pragma unroll
    for (int m = 0; m < L; ++m) {
      const unsigned o = obase + (unsigned)(e0 + m) * ostep;
      const bool in = e0 + m < n_out;
      mm[m] = (A.mul && in) ? at(A.mul, o) : T(1);
      aa[m] = (A.acc && in) ? at(A.acc, o) : T(0);
    }
#pragma unroll
    for (int m = 0; m < L; ++m) {
      if (e0 + m >= n_out) continue;
      T r = w[0][m];
      if (A.mul) r *= mm[m];
      r = A.rc * r;
      if (A.acc) r = A.ac * aa[m] + r;
      st_nt(A.t, obase + (unsigned)(e0 + m) * ostep, r);
    }
  } else {
#pragma unroll
    for (int m = 0; m < L; ++m)
      if (e0 + m < n_out) st_nt(A.t, obase + (unsigned)(e0 + m) * ostep, w[0][m]);
  }
}

// Shape of seg_wall_kernel: 128-byte rows, segments of the precision's
// segment length; one point longer when the line needs more than 1024
// threads with it (the 513-point wall-normal lines of a 512-cell channel:
// 31 segments of 17 FP32 points, 57 of 9 FP64 points, with the short
// windows' registers); doubled beyond that (up to 640 threads then, so
// that the longer windows keep a register budget of 640 threads and stay
// out of scratch); longer lines go to line_kernel.
template <typename T>
struct WallShape {
  static constexpr int L1 = sizeof(T) == 8 ? SEG_L64 : SEG_L32;
  int L = 0, P = 0, TL = 0;
  size_t shmem = 0;
  bool plan(int n_out) {
    if (n_out < 8) return false;
    TL = 128 / (int)sizeof(T);
    L = L1;
    P = (n_out + L - 1) / L;
    if (TL * P > 1024) {
      L = L1 + 1;
      P = (n_out + L - 1) / L;
    }
    if (TL * P > 1024) {
      L = 2 * L1;
      P = (n_out + L - 1) / L;
      if (TL * P > 640) return false;
    }
    shmem = sizeof(T) * ((size_t)4 * P * L + 2 * (size_t)P * TL + 2 * (size_t)P + 8 * (size_t)TL);
    return shmem <= 65536;
  }
};

template <typename T, int OP>
static int launch_seg_wall(const LineArgs<T> &a, const WallShape<T> &Y, bool epi, int kdepth, hipStream_t st) {
  const dim3 grid((a.na + Y.TL - 1) / Y.TL, a.nb), block(Y.TL, Y.P);
  constexpr int L1 = WallShape<T>::L1;
  // the set's chain depth counts segments of L1 points (a segment of
  // L1 + 1 points attenuates at least as much as one of L1)
  const int K = std::min(Y.P, Y.L <= L1 + 1 ? kdepth : (kdepth + 1) / 2);
  if (Y.L == L1) {
    if (epi) hipLaunchKernelGGL((seg_wall_kernel<T, OP, L1, true, 1024>), grid, block, Y.shmem, st, a, K);
    else hipLaunchKernelGGL((seg_wall_kernel<T, OP, L1, false, 1024>), grid, block, Y.shmem, st, a, K);
  } else if (Y.L == L1 + 1) {
    // the plain operator fits two blocks of up to 16 waves per CU (at most
    // 64 VGPRs); the fused-operand form would spill under that budget
    if (epi) hipLaunchKernelGGL((seg_wall_kernel<T, OP, L1 + 1, true, 1024>), grid, block, Y.shmem, st, a, K);
    else hipLaunchKernelGGL((seg_wall_kernel<T, OP, L1 + 1, false, 1024, 8>), grid, block, Y.shmem, st, a, K);
  } else {
    if (epi) hipLaunchKernelGGL((seg_wall_kernel<T, OP, 2 * L1, true, 640>), grid, block, Y.shmem, st, a, K);
    else hipLaunchKernelGGL((seg_wall_kernel<T, OP, 2 * L1, false, 640>), grid, block, Y.shmem, st, a, K);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(X3D_ERR_HIP, std::string("segmented wall-line launch: ") + hipGetErrorString(e));
  return X3D_OK;
}

// ---------------------------------------------------------------- x lines, one wavefront per line
// Periodic x lines of n = 64*V points: lane k owns points [kV, kV+V) of the
// line (contiguous, so a wavefront reads/writes the line as one 64*V-word
// burst), the stencil halo comes from the neighbour lanes through
// ds_bpermute, and the segment carries of the two recurrences are chained by
// a Hillis-Steele scan of affine maps across the 64 lanes.  No LDS and no
// barriers; every wavefront keeps its coefficients in registers and walks a
// strided sequence of lines.
template <typename T, int V>
__device__ __forceinline__ void load_run(const T *__restrict__ p, T (&out)[V]) {
  if constexpr ((V * sizeof(T)) % 16 == 0) {
    using V4 = __attribute__((ext_vector_type(16 / sizeof(T)))) T;
    constexpr int PER = 16 / sizeof(T);
#pragma unroll
    for (int q = 0; q < V / PER; ++q) {
      const V4 x = *reinterpret_cast<const V4 *>(p + q * PER);
#pragma unroll
      for (int r = 0; r < PER; ++r) out[q * PER + r] = x[r];
    }
  } else {
#pragma unroll
    for (int m = 0; m < V; ++m) out[m] = p[m];
  }
}

template <typename T, int V>
__device__ __forceinline__ void store_run(T *__restrict__ p, const T (&in)[V]) {
  if constexpr ((V * sizeof(T)) % 16 == 0) {
    using V4 = __attribute__((ext_vector_type(16 / sizeof(T)))) T;
    constexpr int PER = 16 / sizeof(T);
#pragma unroll
    for (int q = 0; q < V / PER; ++q) {
      V4 x;
#pragma unroll
      for (int r = 0; r < PER; ++r) x[r] = in[q * PER + r];
      *reinterpret_cast<V4 *>(p + q * PER) = x;
    }
  } else {
#pragma unroll
    for (int m = 0; m < V; ++m) p[m] = in[m];
  }
}


template <typename T, int OP, int V, bool EPI>
__global__ void __launch_bounds__(256) xwave_kernel(const LineArgs<T> A, int nlines) {
  constexpr int HL = Halo<OP>::L, HR = Halo<OP>::R, W = V + HL + HR;
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  const int n = A.n_out;   // == n_in == 64*V
  const int e0 = lane * V;
  // 512-point lines read their coefficients from LDS for every line (laid
  // out [coefficient][point of the lane][lane], conflict-free) instead of
  // holding 4V of them in registers
  constexpr bool CL = V == 8;
  __shared__ T ctab[CL ? 4 * 64 * V : 1];
  T fs[V], fw[V], ffw[V], sm[V];
  if constexpr (CL) {
    for (int i = threadIdx.x; i < 4 * n; i += blockDim.x) {
      const int e = i >> 2, k = i & 3, ln = e / V, m = e - ln * V;
      ctab[(k * V + m) * 64 + ln] = A.pk[i];
    }
    __syncthreads();
  } else {
#pragma unroll
    for (int m = 0; m < V; ++m) {
      const int e = e0 + m;
      fs[m] = A.pk[4 * e + 0];
      fw[m] = A.pk[4 * e + 1];
      ffw[m] = A.pk[4 * e + 2];
      sm[m] = A.pk[4 * e + 3];
    }
  }
  // the next line's input is loaded while the current one is solved
  T nxt_u[V], nxt_u2[V];
  if (wave < nlines) {
    load_run<T, V>(A.u + (long long)wave * n + e0, nxt_u);
    if (EPI && A.u2) load_run<T, V>(A.u2 + (long long)wave * n + e0, nxt_u2);
  }
  for (int ln = wave; ln < nlines; ln += nwaves) {
    const long long base = (long long)ln * n + e0;
    T own[V];
#pragma unroll
    for (int m = 0; m < V; ++m) own[m] = nxt_u[m];
    if (EPI && A.u2) {
#pragma unroll
      for (int m = 0; m < V; ++m) own[m] *= nxt_u2[m];
    }
    if (ln + nwaves < nlines) {
      load_run<T, V>(A.u + base + (long long)nwaves * n, nxt_u);
      if (EPI && A.u2) load_run<T, V>(A.u2 + base + (long long)nwaves * n, nxt_u2);
    }
    // window U(e0-HL .. e0+V+HR-1): own points plus neighbour lanes' points
    T v[W];
    halo_window<T, V, HL, HR>(own, v, lane);
    // right-hand side in place, then the solve
#pragma unroll
    for (int m = 0; m < V; ++m) v[m] = interior_row<OP, T, W>(v, m + HL, A.s);
    T r[V];
#pragma unroll
    for (int m = 0; m < V; ++m) r[m] = v[m];
    if constexpr (CL) {
      asm volatile("" ::: "memory");   // keep the table reads inside the loop
#pragma unroll
      for (int m = 0; m < V; ++m) {
        fs[m] = ctab[(0 * V + m) * 64 + lane];
        fw[m] = ctab[(1 * V + m) * 64 + lane];
        ffw[m] = ctab[(2 * V + m) * 64 + lane];
        sm[m] = ctab[(3 * V + m) * 64 + lane];
      }
    }
    wave_solve<T, V>(r, fs, fw, ffw, sm, A.alfa, A.rden, lane);
#pragma unroll
    for (int m = 0; m < V; ++m) v[m] = r[m];
    // epilogue and store
    T res[V];
#pragma unroll
    for (int m = 0; m < V; ++m) res[m] = v[m];
    if (EPI && A.mul) {
      T mm[V];
      load_run<T, V>(A.mul + base, mm);
#pragma unroll
      for (int m = 0; m < V; ++m) res[m] *= mm[m];
    }
    if (EPI) {
#pragma unroll
      for (int m = 0; m < V; ++m) res[m] = A.rc * res[m];
    }
    if (EPI && A.acc) {
      T aa[V];
      load_run<T, V>(A.acc + base, aa);
#pragma unroll
      for (int m = 0; m < V; ++m) res[m] = A.ac * aa[m] + res[m];
    }
    store_run<T, V>(A.t + base, res);
  }
}

template <typename T, int OP, int V>
static void launch_xv(const LineArgs<T> &a, bool epi, int nlines, dim3 grid, dim3 block, hipStream_t st) {
  if (epi) hipLaunchKernelGGL((xwave_kernel<T, OP, V, true>), grid, block, 0, st, a, nlines);
  else hipLaunchKernelGGL((xwave_kernel<T, OP, V, false>), grid, block, 0, st, a, nlines);
}

template <typename T, int OP>
static int launch_xwave(const LineArgs<T> &a, bool epi, int nlines, hipStream_t st) {
  const int V = a.n_out / 64;
  const int waves = nlines < 16384 ? nlines : 16384;
  const dim3 grid((waves + 3) / 4), block(256);
  switch (V) {
    case 1: launch_xv<T, OP, 1>(a, epi, nlines, grid, block, st); break;
    case 2: launch_xv<T, OP, 2>(a, epi, nlines, grid, block, st); break;
    case 4: launch_xv<T, OP, 4>(a, epi, nlines, grid, block, st); break;
    case 8: launch_xv<T, OP, 8>(a, epi, nlines, grid, block, st); break;
    default: return fail(X3D_ERR_ARG, "xwave: unsupported line length");
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(X3D_ERR_HIP, std::string("xwave launch: ") + hipGetErrorString(e));
  return X3D_OK;
}

static inline bool aligned16(const void *p) { return p == nullptr || (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// ---------------------------------------------------------------- launcher
template <typename T, int OP, int L, int AX>
static void launch_ax(const LineArgs<T> &a, bool periodic, bool epi, dim3 grid, dim3 block, size_t shmem,
                      hipStream_t st) {
  if (periodic) {
    if (epi) hipLaunchKernelGGL((line_kernel<T, OP, L, AX, true, true>), grid, block, shmem, st, a);
    else hipLaunchKernelGGL((line_kernel<T, OP, L, AX, true, false>), grid, block, shmem, st, a);
  } else {
    if (epi) hipLaunchKernelGGL((line_kernel<T, OP, L, AX, false, true>), grid, block, shmem, st, a);
    else hipLaunchKernelGGL((line_kernel<T, OP, L, AX, false, false>), grid, block, shmem, st, a);
  }
}

template <typename T, int OP, int L>
static int launch_l(const LineArgs<T> &a, int ax, bool periodic, bool epi, int TL, int P, dim3 grid,
                    size_t shmem, hipStream_t st) {
  dim3 block(TL, P);
  // FP64 64-point segments of y/z wall lines up to 640 points (e.g. 513):
  // a 640-thread register budget (170 VGPRs instead of 128) keeps the
  // 72-point windows out of scratch (FP32 fits either way)
  if constexpr (L == 64 && sizeof(T) == 8) {
    if (ax != 0 && !periodic && TL * P <= 640) {
      if (ax == 1) {
        if (epi) hipLaunchKernelGGL((line_kernel<T, OP, L, 1, false, true, 640>), grid, block, shmem, st, a);
        else hipLaunchKernelGGL((line_kernel<T, OP, L, 1, false, false, 640>), grid, block, shmem, st, a);
      } else {
        if (epi) hipLaunchKernelGGL((line_kernel<T, OP, L, 2, false, true, 640>), grid, block, shmem, st, a);
        else hipLaunchKernelGGL((line_kernel<T, OP, L, 2, false, false, 640>), grid, block, shmem, st, a);
      }
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return fail(X3D_ERR_HIP, std::string("line kernel launch: ") + hipGetErrorString(e));
      return X3D_OK;
    }
  }
  if (ax == 0) launch_ax<T, OP, L, 0>(a, periodic, epi, grid, block, shmem, st);
  else if (ax == 1) launch_ax<T, OP, L, 1>(a, periodic, epi, grid, block, shmem, st);
  else launch_ax<T, OP, L, 2>(a, periodic, epi, grid, block, shmem, st);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(X3D_ERR_HIP, std::string("line kernel launch: ") + hipGetErrorString(e));
  return X3D_OK;
}

template <typename T, int OP>
static int launch_op(const LineArgs<T> &a, int ax, bool periodic, bool epi, int nmax, int nb, hipStream_t st) {
  const bool xmode = ax == 0;
  if (nmax > 1024) return fail(X3D_ERR_UNSUPPORTED, "line length > 1024");
  const int L = nmax <= 128 ? 16 : (nmax <= 512 ? 32 : 64);
  const int P = (nmax + L - 1) / L;
  int TL = 64;
  size_t fixed = sizeof(T) * ((size_t)(periodic ? 0 : 8 * TL) + (size_t)P * TL + P + 2 * TL);
  if (xmode) {
    // largest power-of-two line count whose tile fits in 64 KB
    TL = 64;
    while (TL > 8 && sizeof(T) * (size_t)TL * a.row > 64 * 1024) TL >>= 1;
    fixed = sizeof(T) * ((size_t)TL * a.row + 4 * (size_t)a.n_out + (periodic ? 0 : 8 * TL) + (size_t)P * TL +
                         P + 2 * TL);
  }
  if (TL * P > 1024) return fail(X3D_ERR_UNSUPPORTED, "line kernel block too large");
  if (fixed > 160 * 1024) return fail(X3D_ERR_UNSUPPORTED, "line kernel LDS > 160 KB");
  dim3 grid((a.na + TL - 1) / TL, nb);
  switch (L) {
    case 16: return launch_l<T, OP, 16>(a, ax, periodic, epi, TL, P, grid, fixed, st);
    case 32: return launch_l<T, OP, 32>(a, ax, periodic, epi, TL, P, grid, fixed, st);
    default: return launch_l<T, OP, 64>(a, ax, periodic, epi, TL, P, grid, fixed, st);
  }
}

template <typename T>
int launch_line(const LineCall<T> &c, hipStream_t st) {
  if (!c.u || !c.t || !c.tri.pk) return fail(X3D_ERR_ARG, "null pointer");
  LineArgs<T> a{};
  a.u = c.u; a.u2 = c.u2; a.t = c.t; a.pp = c.pp;
  a.acc = c.acc; a.mul = c.mul; a.ac = c.ac; a.rc = c.rc;
  a.pk = c.tri.pk;
  a.alfa = c.tri.alfa; a.sm_den = c.tri.sm_den; a.rden = c.tri.rden;
  a.s = c.s;
  a.bc1 = c.bc1; a.bcn = c.bcn; a.npaire = c.npaire;
  const bool periodic = c.tri.periodic != 0;
  const int nx = c.nx, ny = c.ny, nz = c.nz;
  const int n_in = c.dir == 0 ? nx : (c.dir == 1 ? ny : nz);
  const int n_out = c.n_out;
  if (c.tri.n != n_out) return fail(X3D_ERR_ARG, "coefficient length != output line length");
  if (n_in < 8 || n_out < 8) return fail(X3D_ERR_UNSUPPORTED, "lines shorter than 8 points");
  if (periodic && n_in != n_out) return fail(X3D_ERR_ARG, "periodic line with n_in != n_out");
  a.n_in = n_in; a.n_out = n_out;
  bool xmode = false;
  int nb;
  if (c.dir == 0) {
    xmode = true;
    a.na = ny; a.nb = nz; nb = nz;
    a.se_in = 1; a.sa_in = n_in; a.sb_in = (long long)n_in * ny;
    a.se_out = 1; a.sa_out = n_out; a.sb_out = (long long)n_out * ny;
    int row = (n_in > n_out ? n_in : n_out);
    if (row % 2 == 0) row += 1;   // odd stride: consecutive lines fall in different LDS banks
    a.row = row;
  } else if (c.dir == 1) {
    a.na = nx; a.nb = nz; nb = nz;
    a.se_in = nx; a.sa_in = 1; a.sb_in = (long long)nx * n_in;
    a.se_out = nx; a.sa_out = 1; a.sb_out = (long long)nx * n_out;
  } else {
    a.na = nx; a.nb = ny; nb = ny;
    a.se_in = (long long)nx * ny; a.sa_in = 1; a.sb_in = nx;
    a.se_out = (long long)nx * ny; a.sa_out = 1; a.sb_out = nx;
  }
  const int nmax = n_in > n_out ? n_in : n_out;
  const bool epi = c.u2 || c.acc || c.mul || c.rc != T(1);
  const int V = n_out / 64;
  if (xmode && periodic && n_in == n_out && n_out % 64 == 0 && (V == 1 || V == 2 || V == 4 || V == 8) &&
      !c.pp && aligned16(c.u) && aligned16(c.u2) && aligned16(c.t) && aligned16(c.acc) && aligned16(c.mul)) {
    const int nlines = ny * nz;
    switch (c.op) {
      case OP_DER1: return launch_xwave<T, OP_DER1>(a, epi, nlines, st);
      case OP_DER2: return launch_xwave<T, OP_DER2>(a, epi, nlines, st);
      case OP_DERVP: return launch_xwave<T, OP_DERVP>(a, epi, nlines, st);
      case OP_INTERVP: return launch_xwave<T, OP_INTERVP>(a, epi, nlines, st);
      case OP_DERPV: return launch_xwave<T, OP_DERPV>(a, epi, nlines, st);
      case OP_INTERPV: return launch_xwave<T, OP_INTERPV>(a, epi, nlines, st);
    }
  }
  YzShape<T> Y;
  if (!xmode && periodic && Y.plan(n_out, 1, 1) && (size_t)nx * ny * nz * sizeof(T) < ((size_t)1 << 32)) {
    const int K = c.tri.kdepth > 0 ? c.tri.kdepth : Y.P;
    const T rden = c.tri.rden;
    switch (c.op) {
      case OP_DER1: return launch_seg_line<T, OP_DER1>(a, rden, Y, epi, K, st);
      case OP_DER2: return launch_seg_line<T, OP_DER2>(a, rden, Y, epi, K, st);
      case OP_DERVP: return launch_seg_line<T, OP_DERVP>(a, rden, Y, epi, K, st);
      case OP_INTERVP: return launch_seg_line<T, OP_INTERVP>(a, rden, Y, epi, K, st);
      case OP_DERPV: return launch_seg_line<T, OP_DERPV>(a, rden, Y, epi, K, st);
      case OP_INTERPV: return launch_seg_line<T, OP_INTERPV>(a, rden, Y, epi, K, st);
    }
  }
  WallShape<T> Wsh;
  const size_t fin = (size_t)nx * ny * nz / (size_t)(c.dir == 1 ? ny : nz) * (size_t)n_in * sizeof(T);
  const size_t fout = (size_t)nx * ny * nz / (size_t)(c.dir == 1 ? ny : nz) * (size_t)n_out * sizeof(T);
  if (!xmode && !periodic && Wsh.plan(n_out) && fin < ((size_t)1 << 32) && fout < ((size_t)1 << 32)) {
    const int K = c.tri.kdepth > 0 ? c.tri.kdepth : Wsh.P;
    switch (c.op) {
      case OP_DER1: return launch_seg_wall<T, OP_DER1>(a, Wsh, epi, K, st);
      case OP_DER2: return launch_seg_wall<T, OP_DER2>(a, Wsh, epi, K, st);
      case OP_DERVP: return launch_seg_wall<T, OP_DERVP>(a, Wsh, epi, K, st);
      case OP_INTERVP: return launch_seg_wall<T, OP_INTERVP>(a, Wsh, epi, K, st);
      case OP_DERPV: return launch_seg_wall<T, OP_DERPV>(a, Wsh, epi, K, st);
      case OP_INTERPV: return launch_seg_wall<T, OP_INTERPV>(a, Wsh, epi, K, st);
    }
  }
  const int ax = c.dir;
  switch (c.op) {
    case OP_DER1: return launch_op<T, OP_DER1>(a, ax, periodic, epi, nmax, nb, st);
    case OP_DER2: return launch_op<T, OP_DER2>(a, ax, periodic, epi, nmax, nb, st);
    case OP_DERVP: return launch_op<T, OP_DERVP>(a, ax, periodic, epi, nmax, nb, st);
    case OP_INTERVP: return launch_op<T, OP_INTERVP>(a, ax, periodic, epi, nmax, nb, st);
    case OP_DERPV: return launch_op<T, OP_DERPV>(a, ax, periodic, epi, nmax, nb, st);
    case OP_INTERPV: return launch_op<T, OP_INTERPV>(a, ax, periodic, epi, nmax, nb, st);
  }
  return fail(X3D_ERR_ARG, "unknown line operator");
}


}  // namespace x3d
